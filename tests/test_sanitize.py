"""The oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
(`make -C oracle sanitize`), run over every entry point on small inputs in a
child process with the ASan runtime loaded first: an out-of-bounds access, a
use after free or undefined behaviour in the C restatement aborts the child
and fails the test.  Host code only (no GPU)."""
import os
import shutil
import subprocess
import sys
import textwrap

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

SCRIPT = textwrap.dedent(r"""
    import sys
    import numpy as np
    sys.path[:0] = [%r, %r]
    import oracle as O
    import marinehydro_amd as mh

    rng = np.random.default_rng(0)
    mesh = mh.hemisphere(1.0, ntheta=6, nphi=8).mesh
    panels = [O.make_panel(mesh.centers[j], mesh.vertices[mesh.faces[j]], mesh.normals[j], mesh.areas[j],
                           mesh.radii[j]) for j in range(mesh.nfaces)]
    pts = [mesh.centers[3], mesh.centers[3] + [0.0, 0.0, -1e-7], np.array([5.0, 1.0, -0.4]),
           np.array([0.0, 0.0, -1e-9])]
    for mask in (1, 2, 3, 4, 8, 11, 16, 19):
        for x in pts:
            for p in panels[:6]:
                O.greens(mask, x, p.center, 0.7)
                O.gradient_greens(mask, x, p.center, 0.7, True)
                O.integral(mask, x, p, 0.7)
                O.integral_gradient(mask, x, p, 0.7, False)
    for x in pts:
        O.rankine_integral(x, panels[0])
        O.rankine_integral_gradient(x, panels[0], True)
    for h in (0.0, 1e-12, 0.3, 2.9, 3.1, 20.0, 300.0):
        for v in (-1e-12, -0.2, -3.0, -40.0):
            O.wu_dimless(h, v)
    for r, z in ((0.1, -0.1), (2.0, -0.5), (15.0, -3.0)):
        O.delhommeau_dimless(r, z)
    for z in (1e-3 + 0j, 3 + 2j, -20 + 1e-3j, 50 + 40j):
        O.exp_e1(z)
    for x in (0.0, 0.5, 2.9, 3.1, 12.0):
        O.struve_h0(x), O.struve_h1(x)
    for mask in (11, 19):
        for direct in (True, False):
            O.assemble(mask, mesh, 0.5, direct=direct, nthreads=1)
    A = rng.normal(size=(20, 20)) + 1j * rng.normal(size=(20, 20))
    LU, piv, info = O.zgetrf(A)
    O.zgetrs(LU, piv, rng.normal(size=(20, 2)) + 0j)
    O.zgetrs(LU, piv, rng.normal(size=(20, 2)) + 0j, trans=True)
    S, D = O.assemble(11, mesh, 0.5, nthreads=1)
    O.solve(D, S, np.ones(mesh.nfaces, dtype=complex))
    O.freq_solve(11, mesh, 1.2, [[0, 0, 1.0], [1.0, 0, 0]], betas=[0.0, 0.7])
    O.froude_krylov(mesh, 1.2, [[0, 0, 1.0], [1.0, 0, 0]], betas=(0.0, 0.7))
    print("sanitized oracle: ok")
""")


def test_oracle_under_asan_and_ubsan():
    if shutil.which("gcc") is None or shutil.which("make") is None:
        pytest.skip("no gcc/make")
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True)
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(asan) or not os.path.exists(asan):
        pytest.skip("no ASan runtime")
    env = dict(os.environ)
    # the ASan runtime first; whatever the environment already preloads stays
    env["LD_PRELOAD"] = asan + (":" + env["LD_PRELOAD"] if env.get("LD_PRELOAD") else "")
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["MH_ORACLE_LIB"] = os.path.join(ORACLE, "build", "liborc_san.so")
    script = SCRIPT % (ORACLE, os.path.join(ROOT, "marinehydro.jl_amd"))
    r = subprocess.run([sys.executable, "-c", script], env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "sanitized oracle: ok" in r.stdout
