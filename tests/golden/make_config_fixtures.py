"""Writes tests/golden/config_fixtures.json: oracle A, B, F for the two
large exact-Delhommeau cases, so the GPU tests compare against them without
running the oracle's per-pair quadrature at 800 and 4,000 panels.

  hemisphere40: floating hemisphere R = 1, 40 x 40 latitude-longitude mesh
    (800 immersed panels), heave, heading 0, rho = 1025, g = 9.81,
    Rankine + image + exact Delhommeau, direct and indirect, at 4 of the 64
    frequencies omega_i = 0.3 * 10^(i/63) (i = 0, 21, 42, 63).
  cylinder4000: truncated vertical cylinder R = 5, draft 10, 80 x 30 side
    wall + 20 bottom rings (4,000 panels), six rigid dofs about the origin,
    heading 0, rho = 1025, g = 9.81, Rankine + image + exact Delhommeau,
    direct, at omega = 0.2 + 2.3 * 15/31 (one of 32 equally spaced
    frequencies over 0.2 - 2.5).

Each entry records the mesh parameters and a checksum of the panel centres;
the tests regenerate the mesh and check the checksum before comparing.

Run:  python tests/golden/make_config_fixtures.py   (about 20 minutes on 8
cores; the oracle's exact Delhommeau term dominates).
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "marinehydro.jl_amd"), os.path.join(ROOT, "oracle")]

import marinehydro_amd as mh  # noqa: E402  (mesh generators only; no GPU)
import oracle as O  # noqa: E402

DELH = O.RANKINE | O.RANKINE_REFLECTED | O.DELHOMMEAU
RHO, G = 1025.0, 9.81


def checksum(mesh):
    return float(np.sum(mesh.centers * np.array([1.0, 2.0, 3.0])) + np.sum(mesh.areas))


def cplx(a):
    a = np.asarray(a)
    return {"re": a.real.tolist(), "im": a.imag.tolist()}


def main():
    out = {}
    t0 = time.time()
    m = mh.hemisphere(1.0, ntheta=40, nphi=40).mesh
    omegas = 0.3 * 10.0 ** (np.arange(64) / 63.0)
    idx = [0, 21, 42, 63]
    entry = {"mesh": {"kind": "hemisphere", "radius": 1.0, "ntheta": 40, "nphi": 40},
             "checksum": checksum(m), "nfaces": m.nfaces, "omega_index": idx,
             "omegas": omegas[idx].tolist()}
    for direct in (True, False):
        A_, B_, F_ = [], [], []
        for w in omegas[idx]:
            A, B, F = O.freq_solve(DELH, m, w, [[0.0, 0.0, 1.0]], betas=[0.0], direct=direct, g=G, rho=RHO)
            A_.append(A[0, 0])
            B_.append(B[0, 0])
            F_.append(F[0, 0])
            print(f"hemisphere40 direct={direct} omega={w:.4f} ({time.time() - t0:.0f} s)", flush=True)
        entry["direct" if direct else "indirect"] = {"A": A_, "B": B_, "F": cplx(F_)}
    out["hemisphere40"] = entry

    c = mh.vertical_cylinder(5.0, 10.0, ntheta=80, nz=30, nr=20).mesh
    w = 0.2 + 2.3 * 15 / 31
    dofs = mh.rigid_dofs(c, ["Surge", "Sway", "Heave", "Roll", "Pitch", "Yaw"])
    A, B, F = O.freq_solve(DELH, c, w, dofs, betas=[0.0], direct=True, g=G, rho=RHO)
    print(f"cylinder4000 omega={w:.4f} ({time.time() - t0:.0f} s)", flush=True)
    out["cylinder4000"] = {"mesh": {"kind": "vertical_cylinder", "radius": 5.0, "draft": 10.0, "ntheta": 80,
                                    "nz": 30, "nr": 20},
                           "checksum": checksum(c), "nfaces": c.nfaces, "omega": w,
                           "dofs": ["Surge", "Sway", "Heave", "Roll", "Pitch", "Yaw"],
                           "A": A.tolist(), "B": B.tolist(), "F": cplx(F[:, 0])}
    with open(os.path.join(HERE, "config_fixtures.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote config_fixtures.json ({time.time() - t0:.0f} s)")


if __name__ == "__main__":
    main()
