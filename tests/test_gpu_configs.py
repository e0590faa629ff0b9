"""The workload configurations at their stated sizes, on the GPU.

  config 2: hemisphere 40 x 40 (800 panels), 64 frequencies log-spaced over
            0.3 - 3.0 rad/s, {Wu, exact Delhommeau} x {direct, indirect};
            mixed precision against FP64 for every frequency, FP64 against
            the oracle at sampled frequencies (Wu live, Delhommeau from
            tests/golden/config_fixtures.json).
  config 3: two floating heave hemispheres, 20 x 20 each (400 panels),
            seeded (r1, r2, d, omega); per-body heave fields (A, B 2 x 2,
            F 2); gradients w.r.t. theta = (r1, r2, d, omega) against the
            oracle's central differences with h = 1e-6.
  config 4: truncated vertical cylinder R = 5, draft 10 (4,000 panels), six
            rigid dofs, 32 frequencies over 0.2 - 2.5 rad/s, exact
            Delhommeau, direct; mixed (FP32 LU + FP64 refinement) against
            FP64 for all 32, both against the oracle at one frequency;
            forward-mode gradients w.r.t. radius and draft at two
            frequencies against central differences of FP64 sweeps.

Errors are per quantity: |x - x_ref| <= tol * max|x_ref| over that
quantity's array (the off-diagonal coefficients change sign, so a per-entry
relative error is not defined everywhere).
"""
import json
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
WU = 1 | 2 | 8
DELH = 1 | 2 | 16
RHO, G = 1025.0, 9.81


def relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _fixtures():
    with open(os.path.join(HERE, "golden", "config_fixtures.json")) as f:
        return json.load(f)


def _checksum(mesh):
    return float(np.sum(mesh.centers * np.array([1.0, 2.0, 3.0])) + np.sum(mesh.areas))


def _cplx(d):
    return np.asarray(d["re"]) + 1j * np.asarray(d["im"])


def _gfs(mh, mask):
    return tuple(c() for c in (mh.Rankine, mh.RankineReflected, mh.GFWu, mh.ExactGuevelDelhommeau)
                 if mask & c.mask)


def _per_frequency(a, b):
    """largest |a - b| per frequency over the last axes, relative to max|b|"""
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b).reshape(a.shape[0], -1), axis=1) / np.max(np.abs(b))


# --------------------------------------------------------------------- config 2
CONFIG2_OMEGAS = 0.3 * 10.0 ** (np.arange(64) / 63.0)


@pytest.mark.parametrize("mask", [WU, DELH], ids=["wu", "delhommeau"])
@pytest.mark.parametrize("direct", [True, False], ids=["direct", "indirect"])
def test_config2_hemisphere800_64_frequencies(mh, oracle, gpu, mask, direct):
    m = mh.hemisphere(1.0, ntheta=40, nphi=40).mesh
    assert m.nfaces == 800
    gfs = _gfs(mh, mask)
    t0 = time.time()
    A64, B64, F64 = mh.freq_sweep(m, CONFIG2_OMEGAS, [[0, 0, 1.0]], betas=[0.0], green_functions=gfs,
                                  direct=direct, precision="fp64", rho=RHO, g=G)
    t1 = time.time()
    A32, B32, F32 = mh.freq_sweep(m, CONFIG2_OMEGAS, [[0, 0, 1.0]], betas=[0.0], green_functions=gfs,
                                  direct=direct, precision="mixed", rho=RHO, g=G)
    t2 = time.time()
    eA, eB, eF = _per_frequency(A32, A64), _per_frequency(B32, B64), _per_frequency(F32, F64)
    print(f"config2 mask={mask} direct={direct}: fp64 {t1 - t0:.2f} s, mixed {t2 - t1:.2f} s; "
          f"mixed vs fp64 per frequency max A {eA.max():.2e} B {eB.max():.2e} F {eF.max():.2e}", flush=True)
    assert eA.max() < 1e-6 and eB.max() < 1e-6 and eF.max() < 1e-6
    if mask == WU:
        idx = [0, 21, 42, 63]
        ref = [oracle.freq_solve(mask, m, CONFIG2_OMEGAS[i], [[0, 0, 1.0]], betas=[0.0], direct=direct, g=G,
                                 rho=RHO) for i in idx]
        A_o = np.array([r[0][0, 0] for r in ref])
        B_o = np.array([r[1][0, 0] for r in ref])
        F_o = np.array([r[2][0, 0] for r in ref])
    else:
        fx = _fixtures()["hemisphere40"]
        assert abs(fx["checksum"] - _checksum(m)) <= 1e-12 * abs(fx["checksum"])
        idx = fx["omega_index"]
        e = fx["direct" if direct else "indirect"]
        A_o, B_o, F_o = np.array(e["A"]), np.array(e["B"]), _cplx(e["F"])
    for X, Xo, name in ((A64[idx, 0, 0], A_o, "A"), (B64[idx, 0, 0], B_o, "B"), (F64[idx, 0, 0], F_o, "F"),
                        (A32[idx, 0, 0], A_o, "A mixed"), (B32[idx, 0, 0], B_o, "B mixed"),
                        (F32[idx, 0, 0], F_o, "F mixed")):
        err = relmax(X, Xo)
        print(f"   vs oracle {name}: {err:.2e}", flush=True)
        assert err < 1e-9, (name, err)


# --------------------------------------------------------------------- config 3
def config3_theta(seed):
    """(r1, r2, d, omega): r ~ U[0.5, 2], d ~ U[3, 10] (redrawn until the
    hulls are 0.5 m apart), omega ~ U[0.5, 2]."""
    rng = np.random.default_rng(seed)
    while True:
        r1, r2 = rng.uniform(0.5, 2.0, 2)
        d = rng.uniform(3.0, 10.0)
        w = rng.uniform(0.5, 2.0)
        if d > r1 + r2 + 0.5:
            return [float(r1), float(r2), float(d), float(w)]


def _pair(mh, t, res=20):
    b1 = mh.hemisphere(t[0], center=(0.0, 0.0, 0.0), ntheta=res, nphi=res)
    b2 = mh.hemisphere(t[1], center=(t[2], 0.0, 0.0), ntheta=res, nphi=res)
    mesh = mh.combine_meshes([b1.mesh, b2.mesh])
    n1, n2 = b1.mesh.nfaces, b2.mesh.nfaces
    h1 = np.zeros((n1 + n2, 3))
    h1[:n1, 2] = 1.0
    h2 = np.zeros((n1 + n2, 3))
    h2[n1:, 2] = 1.0
    z1 = np.zeros((b1.mesh.nvertices, 3))
    z2 = np.zeros((b2.mesh.nvertices, 3))
    dV = [np.vstack([b1.dvertices["radius"], z2]), np.vstack([z1, b2.dvertices["radius"]]),
          np.vstack([z1, b2.dvertices["x"]])]
    return mesh, [h1, h2], dV


@pytest.mark.parametrize("seed", [0, 1])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_config3_pair_gradients(mh, oracle, gpu, seed, precision):
    t = config3_theta(seed)
    mesh, dofs, dV = _pair(mh, t)
    assert mesh.nfaces == 400
    T = mh.shape_tangents(mesh, dV + [None], nomega=1, omega_param=3)
    A, B, F, dA, dB, dF = mh.freq_sweep_grad(mesh, [t[3]], dofs, betas=[0.0], tangents=T, precision=precision,
                                             rho=RHO, g=G)
    assert A.shape == (1, 2, 2) and F.shape == (1, 1, 2) and dA.shape == (1, 4, 2, 2)
    A_o, B_o, F_o = oracle.freq_solve(WU, mesh, t[3], dofs, betas=[0.0], g=G, rho=RHO)
    assert relmax(A[0], A_o) < 1e-9 and relmax(B[0], B_o) < 1e-9 and relmax(F[0, 0], F_o[:, 0]) < 1e-9
    h = 1e-6
    for p in range(4):
        outs = []
        for s in (+1, -1):
            tt = list(t)
            tt[p] += s * h
            mm, dd, _ = _pair(mh, tt)
            outs.append(oracle.freq_solve(WU, mm, tt[3], dd, betas=[0.0], g=G, rho=RHO))
        fA = (outs[0][0] - outs[1][0]) / (2 * h)
        fB = (outs[0][1] - outs[1][1]) / (2 * h)
        fF = (outs[0][2][:, 0] - outs[1][2][:, 0]) / (2 * h)
        eA, eB, eF = relmax(dA[0, p], fA), relmax(dB[0, p], fB), relmax(dF[0, p, 0], fF)
        print(f"config3 seed={seed} theta={np.round(t, 3)} {precision} param {p}: dA {eA:.1e} dB {eB:.1e} "
              f"dF {eF:.1e}", flush=True)
        assert max(eA, eB, eF) < 1e-6, (p, eA, eB, eF)


# --------------------------------------------------------------------- config 4
def test_config4_cylinder4000_six_dofs(mh, gpu):
    c = mh.vertical_cylinder(5.0, 10.0, ntheta=80, nz=30, nr=20).mesh
    assert c.nfaces == 4000
    fx = _fixtures()["cylinder4000"]
    assert abs(fx["checksum"] - _checksum(c)) <= 1e-12 * abs(fx["checksum"])
    dofs = mh.rigid_dofs(c, fx["dofs"])
    omegas = np.linspace(0.2, 2.5, 32)
    assert abs(omegas[15] - fx["omega"]) < 1e-14
    gfs = _gfs(mh, DELH)
    dm = mh.DeviceMesh(c)
    t0 = time.time()
    A64, B64, F64 = mh.freq_sweep(dm, omegas, dofs, betas=[0.0], green_functions=gfs, precision="fp64",
                                  rho=RHO, g=G)
    t1 = time.time()
    A32, B32, F32 = mh.freq_sweep(dm, omegas, dofs, betas=[0.0], green_functions=gfs, precision="mixed",
                                  rho=RHO, g=G)
    t2 = time.time()
    eA, eB, eF = _per_frequency(A32, A64), _per_frequency(B32, B64), _per_frequency(F32, F64)
    print(f"config4: fp64 sweep {t1 - t0:.2f} s, mixed sweep {t2 - t1:.2f} s (32 frequencies); "
          f"mixed vs fp64 max A {eA.max():.2e} B {eB.max():.2e} F {eF.max():.2e}", flush=True)
    assert eA.max() < 1e-6 and eB.max() < 1e-6 and eF.max() < 1e-6
    A_o, B_o, F_o = np.array(fx["A"]), np.array(fx["B"]), _cplx(fx["F"])
    for X, Xo, name in ((A64[15], A_o, "A"), (B64[15], B_o, "B"), (F64[15, 0], F_o, "F"),
                        (A32[15], A_o, "A mixed"), (B32[15], B_o, "B mixed"), (F32[15, 0], F_o, "F mixed")):
        err = relmax(X, Xo)
        print(f"   vs oracle {name}: {err:.2e}", flush=True)
        assert err < 1e-9, (name, err)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_config4_cylinder4000_six_dof_gradients(mh, gpu, precision):
    """forward-mode gradients of the 4,000-panel cylinder's six-dof A, B and
    F (heading 0) w.r.t. radius and draft at two of the 32 frequencies,
    against central differences (h = 1e-6 relative) of FP64 sweeps on
    regenerated meshes; the FP64 sweep itself is pinned to the oracle above.
    Seven right-hand sides: the pair kernels run in 2 x 2 column blocks."""
    fx = _fixtures()["cylinder4000"]
    R, H = 5.0, 10.0
    make = lambda r, h: mh.vertical_cylinder(r, h, ntheta=80, nz=30, nr=20)
    body = make(R, H)
    c = body.mesh
    names = fx["dofs"]
    omegas = np.linspace(0.2, 2.5, 32)[[5, 15]]
    gfs = _gfs(mh, DELH)
    axes = [None, None, None, [1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0]]
    T = mh.shape_tangents(c, [body.dvertices["radius"], body.dvertices["draft"]], dof_axes=axes)
    t0 = time.time()
    A, B, F, dA, dB, dF = mh.freq_sweep_grad(c, omegas, mh.rigid_dofs(c, names), [0.0], T, green_functions=gfs,
                                             precision=precision, rho=RHO, g=G)
    t1 = time.time()
    theta = [R, H]
    for m in range(2):
        step = 1e-6 * theta[m]
        outs = []
        for sgn in (1, -1):
            t = list(theta)
            t[m] += sgn * step
            mm = make(*t).mesh
            outs.append(mh.freq_sweep(mm, omegas, mh.rigid_dofs(mm, names), betas=[0.0], green_functions=gfs,
                                      precision="fp64", rho=RHO, g=G))
        for q, (got, name) in enumerate(((dA, "dA"), (dB, "dB"), (dF, "dF"))):
            ref = (np.asarray(outs[0][q]) - np.asarray(outs[1][q])) / (2 * step)
            err = relmax(got[:, m], ref)
            print(f"config4 {precision} d/d{('radius', 'draft')[m]} {name}: {err:.1e}", flush=True)
            assert err < 1e-6, (precision, m, name, err)
    print(f"config4 {precision} gradient sweep (2 parameters, 2 frequencies): {t1 - t0:.2f} s", flush=True)
