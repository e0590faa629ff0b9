"""GPU parity of the exact Guevel-Delhommeau wave term against the oracle.

The reference integrates the theta-integral adaptively
(src/green_functions/exact_Guevel_Delhommeau.jl:8-26); the oracle does so in
long double and is pinned to mpmath at 30 digits (tests/test_oracle.py); the
device uses a fixed graded Gauss-Legendre rule in FP64.  Tolerances: 1e-11
relative on the dimensionless term and the matrices, 1e-10 on A, B, F.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GF_DEL = 1 | 2 | 16


def relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_wave_term_grid_delhommeau(mh, oracle, gpu):
    torch = gpu
    r = np.concatenate([[0.0, 1e-7, 1e-3], np.geomspace(0.01, 60.0, 30)])
    z = -np.geomspace(1e-4, 30.0, 20)
    R, Z = np.meshgrid(r, z, indexing="ij")
    R, Z = R.ravel(), Z.ravel()
    out = torch.empty((3, R.size), dtype=torch.complex128, device="cuda")
    dr, dz = torch.from_numpy(R).cuda(), torch.from_numpy(Z).cuda()
    mh._lib.check(mh._lib.lib().mh_wave_term_eval(mh._lib.context(0), 16, R.size, dr.data_ptr(),
                                                  dz.data_ptr(), out.data_ptr()))
    got = out.cpu().numpy()
    ref = np.array([oracle.delhommeau_dimless(a, b) for a, b in zip(R, Z)]).T
    for c in range(3):
        err = np.abs(got[c] - ref[c]) / np.maximum(1.0, np.abs(ref[c]))
        assert np.max(err) < 1e-11, (c, np.max(err))


def test_reference_constants_on_device(mh, gpu):
    # test/greens_function.jl:81-92 (Capytaine values, good to ~1e-5)
    e1 = {"center": [0.0, 0.0, -1.0]}
    c = np.array([1.0, 1.0, -2.0])
    e2 = {"center": c, "vertices": np.array([[-.5, -.5, 0], [.5, -.5, 0], [.5, .5, 0], [-.5, .5, 0]]) + c,
          "normal": [0, 0, 1.0], "area": 1.0, "radius": math.sqrt(2) / 2}
    g = mh.greens(mh.ExactGuevelDelhommeau(), e1, e2, 1.0)
    assert abs(g - (-0.9289959399878619 + 0.17490912138602516j)) < 1e-4
    gg = mh.gradient_greens(mh.ExactGuevelDelhommeau(), e1, e2, 1.0, with_respect_to_first_variable=True)
    ref = np.array([-0.06819068 + 0.12043414j, -0.06819068 + 0.12043414j, -0.32597325 + 0.17490912j])
    assert np.max(np.abs(gg - ref)) < 1e-4
    for k in [0.1, 1.0, 10.0]:  # Wu vs exact at atol 2e-2 (test/greens_function.jl:96-99)
        assert abs(mh.greens(mh.GFWu(), e1, e2, k) - mh.greens(mh.ExactGuevelDelhommeau(), e1, e2, k)) < 2e-2


@pytest.mark.parametrize("op", [0, 1, 2, 3])
def test_gf_eval_delhommeau(mh, oracle, gpu, op):
    rng = np.random.default_rng(11)
    m = mh.hemisphere(1.0, ntheta=8, nphi=8).mesh
    nf = m.nfaces
    i = rng.integers(0, nf, 80)
    j = rng.integers(0, nf, 80)
    pts = m.centers[i]
    recs = np.zeros((80, 20))
    recs[:, 0:3] = m.centers[j]
    recs[:, 3:15] = m.vertices[m.faces[j]].reshape(80, 12)
    recs[:, 15:18] = m.normals[j]
    recs[:, 18] = m.areas[j]
    recs[:, 19] = m.radii[j]
    gfs = (mh.Rankine(), mh.RankineReflected(), mh.ExactGuevelDelhommeau())
    for wrt in ([False] if op in (0, 2) else [False, True]):
        got = mh.gf_eval(gfs, pts, recs, 1.3, op, wrt)
        ref = []
        for p, r in zip(pts, recs):
            pan = oracle.make_panel(r[0:3], r[3:15].reshape(4, 3), r[15:18], r[18], r[19])
            ref.append([oracle.greens(GF_DEL, p, r[0:3], 1.3), oracle.gradient_greens(GF_DEL, p, r[0:3], 1.3, wrt),
                        oracle.integral(GF_DEL, p, pan, 1.3),
                        oracle.integral_gradient(GF_DEL, p, pan, 1.3, wrt)][op])
        ref = np.array(ref)
        ok = np.isfinite(ref)
        assert relmax(got[ok], ref[ok]) < 1e-11, (op, wrt)


def test_assembly_delhommeau_config1_mesh(mh, oracle, gpu):
    m = mh.hemisphere(1.0, ntheta=10, nphi=10).mesh
    k = 1.03 ** 2 / 9.81
    gfs = (mh.Rankine(), mh.RankineReflected(), mh.ExactGuevelDelhommeau())
    S, D, K = mh.assemble_matrices_batched(gfs, m, [k, 2.0], formulation=3)
    S, D, K = S.cpu().numpy(), D.cpu().numpy(), K.cpu().numpy()
    for q, kk in enumerate([k, 2.0]):
        So, Do = oracle.assemble(GF_DEL, m, kk, True)
        _, Ko = oracle.assemble(GF_DEL, m, kk, False)
        assert relmax(S[q], So) < 1e-11
        assert relmax(D[q], Do) < 1e-11
        assert relmax(K[q], Ko) < 1e-11


def test_config1_end_to_end(mh, oracle, gpu):
    """hemisphere R = 1, 10 x 10 (50 panels), heave, omega = 1.03, exact
    Delhommeau, indirect formulation, rho = 1025, heading 0"""
    m = mh.hemisphere(1.0, ntheta=10, nphi=10).mesh
    gfs = (mh.Rankine(), mh.RankineReflected(), mh.ExactGuevelDelhommeau())
    A, B, F = mh.freq_sweep(m, [1.03], [[0, 0, 1.0]], [0.0], green_functions=gfs, direct=False,
                            rho=1025.0)
    Ao, Bo, Fo = oracle.freq_solve(GF_DEL, m, 1.03, [[0, 0, 1.0]], [0.0], direct=False, rho=1025.0)
    assert abs(A[0, 0, 0] - Ao[0, 0]) < 1e-10 * abs(Ao[0, 0])
    assert abs(B[0, 0, 0] - Bo[0, 0]) < 1e-10 * abs(Bo[0, 0])
    assert abs(F[0, 0, 0] - Fo[0, 0]) < 1e-10 * abs(Fo[0, 0])


def test_hulme_with_exact_delhommeau(mh, gpu):
    # test/consistency_with_analytical_solutions.jl:149-166 (rho 1000)
    K = np.array([0.05, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 1.0, 1.2, 1.4, 1.6, 1.8, 2.0,
                  2.5, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0, 9.0, 10.0])
    A_h = np.array([0.8764, 0.8627, 0.7938, 0.7157, 0.6452, 0.5861, 0.5381, 0.4999, 0.4698, 0.4464,
                    0.4284, 0.4047, 0.3924, 0.3871, 0.3864, 0.3884, 0.3988, 0.4111, 0.4322, 0.4471,
                    0.4574, 0.4647, 0.4700, 0.4740, 0.4771])
    B_h = np.array([0.1036, 0.1816, 0.2793, 0.3254, 0.3410, 0.3391, 0.3271, 0.3098, 0.2899, 0.2691,
                    0.2484, 0.2096, 0.1756, 0.1469, 0.1229, 0.1031, 0.0674, 0.0452, 0.0219, 0.0116,
                    0.0066, 0.0040, 0.0026, 0.0017, 0.0012])
    m = mh.hemisphere(1.0, ntheta=10, nphi=10).mesh
    w = np.sqrt(K * 9.8)
    gfs = (mh.Rankine(), mh.RankineReflected(), mh.ExactGuevelDelhommeau())
    A, B, _ = mh.freq_sweep(m, w, [[0, 0, 1.0]], green_functions=gfs, rho=1000.0)
    An = A[:, 0, 0] / ((2 / 3) * math.pi * 1000.0)
    Bn = B[:, 0, 0] / ((2 / 3) * math.pi * 1000.0) / w
    nrm = lambda a, b: np.linalg.norm(a - b) <= 0.1 * max(np.linalg.norm(a), np.linalg.norm(b))
    assert nrm(An, A_h) and nrm(Bn, B_h)


def test_delhommeau_gradients_vs_oracle_fd(mh, oracle, gpu):
    res, omega = 6, 1.1
    body = mh.hemisphere(1.0, ntheta=res, nphi=res)
    mesh = body.mesh
    gfs = (mh.Rankine(), mh.RankineReflected(), mh.ExactGuevelDelhommeau())
    T = mh.shape_tangents(mesh, [body.dvertices["radius"], None], nomega=1, omega_param=1)
    A, B, F, dA, dB, dF = mh.freq_sweep_grad(mesh, [omega], [[0, 0, 1.0]], [0.0], T, green_functions=gfs,
                                             rho=1025.0)
    h = 1e-5
    for m, (mk, om_of) in enumerate([(lambda s: mh.hemisphere(1.0 + s * h, ntheta=res, nphi=res).mesh,
                                      lambda s: omega),
                                     (lambda s: mesh, lambda s: omega + s * h)]):
        vals = [oracle.freq_solve(GF_DEL, mk(s), om_of(s), [[0, 0, 1.0]], [0.0], rho=1025.0)
                for s in (1, -1)]
        gA, gB, gF = [(vals[0][c] - vals[1][c]) / (2 * h) for c in range(3)]
        assert abs(dA[0, m, 0, 0] - gA[0, 0]) < 1e-6 * abs(gA[0, 0]), (m, dA[0, m], gA)
        assert abs(dB[0, m, 0, 0] - gB[0, 0]) < 1e-6 * abs(gB[0, 0]), (m, dB[0, m], gB)
        assert abs(dF[0, m, 0, 0] - gF[0, 0]) < 1e-6 * abs(gF[0, 0]), (m, dF[0, m], gF)
