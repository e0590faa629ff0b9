"""GPU parity: libmhip.so (through the C ABI) against the FP64 CPU oracle.

Tolerances: FP64 kernels are compared entry by entry with
|gpu - oracle| <= tol * max|oracle| over the array (tol written per test);
they differ from the oracle only by floating-point association and the
device's libm.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GF_WU = 1 | 2 | 8


def relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _pairs(mh, m, rng, npairs):
    """Random (collocation point, source panel) pairs from a mesh, including
    the self pairs, near neighbours and far pairs."""
    nf = m.nfaces
    i = rng.integers(0, nf, npairs)
    j = rng.integers(0, nf, npairs)
    j[: nf if npairs >= nf else npairs] = i[: nf if npairs >= nf else npairs]  # self pairs
    pts = m.centers[i]
    recs = np.zeros((npairs, 20))
    recs[:, 0:3] = m.centers[j]
    recs[:, 3:15] = m.vertices[m.faces[j]].reshape(npairs, 12)
    recs[:, 15:18] = m.normals[j]
    recs[:, 18] = m.areas[j]
    recs[:, 19] = m.radii[j]
    return pts, recs


@pytest.mark.parametrize("mask", [1, 2, 4, 1 | 2, 8, GF_WU])
@pytest.mark.parametrize("op", [0, 1, 2, 3])
def test_gf_eval_matches_oracle(mh, oracle, gpu, mask, op):
    rng = np.random.default_rng(7)
    m = mh.hemisphere(1.0, ntheta=12, nphi=12).mesh
    pts, recs = _pairs(mh, m, rng, 300)
    k = 0.7
    for wrt in ([False] if op in (0, 2) else [False, True]):
        got = mh.gf_eval(_gfs(mh, mask), pts, recs, k, op, wrt)
        ref = []
        for p, r in zip(pts, recs):
            pan = oracle.make_panel(r[0:3], r[3:15].reshape(4, 3), r[15:18], r[18], r[19])
            if op == 0:
                ref.append(oracle.greens(mask, p, r[0:3], k))
            elif op == 1:
                ref.append(oracle.gradient_greens(mask, p, r[0:3], k, wrt))
            elif op == 2:
                ref.append(oracle.integral(mask, p, pan, k))
            else:
                ref.append(oracle.integral_gradient(mask, p, pan, k, wrt))
        ref = np.array(ref)
        ok = np.isfinite(ref)
        assert np.array_equal(ok, np.isfinite(got))
        assert relmax(got[ok], ref[ok]) < 1e-12, (mask, op, wrt)


def _gfs(mh, mask):
    out = []
    for cls in (mh.Rankine, mh.RankineReflected, mh.RankineReflectedNegative, mh.GFWu,
                mh.ExactGuevelDelhommeau):
        if mask & cls.mask:
            out.append(cls())
    return tuple(out)


def test_reference_element_pair(mh, gpu):
    # test/greens_function.jl:38-73 through the reference-named API
    e1 = {"center": [0.0, 0.0, -1.0]}
    c = np.array([1.0, 1.0, -2.0])
    e2 = {"center": c, "vertices": np.array([[-.5, -.5, 0], [.5, -.5, 0], [.5, .5, 0], [-.5, .5, 0]]) + c,
          "normal": [0, 0, 1.0], "area": 1.0, "radius": math.sqrt(2) / 2}
    assert abs(mh.greens(mh.Rankine(), e1, e2) - 1 / math.sqrt(3)) < 1e-8
    g = mh.gradient_greens(mh.Rankine(), e1, e2, with_respect_to_first_variable=True)
    assert np.allclose(g, (c - np.array([0, 0, -1.0])) / math.sqrt(3) ** 3, atol=1e-8)
    assert abs(mh.greens(mh.RankineReflected(), e1, e2) - 1 / math.sqrt(11)) < 1e-8
    assert abs(mh.integral(mh.Rankine(), e1, e2) - 1 / math.sqrt(3)) < 1e-1
    for k in [0.1, 1.0, 10.0]:  # Wu vs exact Delhommeau at atol 2e-2 (:96-99)
        w = mh.greens(mh.GFWu(), e1, e2, k)
        d = mh.greens(mh.ExactGuevelDelhommeau(), e1, e2, k)
        assert abs(w - d) < 2e-2, (k, w, d)


def test_wave_term_grid(mh, oracle, gpu):
    import ctypes as C

    torch = gpu
    r = np.concatenate([[0.0, 1e-9, 1e-6], np.linspace(0.01, 40.0, 97)])
    z = -np.concatenate([[1e-4, 0.01], np.linspace(0.05, 30.0, 49)])
    R, Z = np.meshgrid(r, z, indexing="ij")
    R, Z = R.ravel(), Z.ravel()
    dr = torch.from_numpy(R).cuda()
    dz = torch.from_numpy(Z).cuda()
    out = torch.empty((3, R.size), dtype=torch.complex128, device="cuda")
    ctx = mh._lib.context(0)
    mh._lib.check(mh._lib.lib().mh_wave_term_eval(ctx, 8, R.size, dr.data_ptr(), dz.data_ptr(),
                                                  out.data_ptr()))
    got = out.cpu().numpy()
    ref = np.array([oracle.wu_dimless(a, b) for a, b in zip(R, Z)]).T
    for c in range(3):
        assert relmax(got[c], ref[c]) < 1e-12, c


@pytest.mark.parametrize("mask", [1, 1 | 2, 1 | 4, GF_WU])
def test_assembly_matches_oracle(mh, oracle, gpu, mask):
    m = mh.hemisphere(1.0, ntheta=10, nphi=10).mesh  # ~50 panels (the CPU-runnable config)
    ks = [1.03 ** 2 / 9.81, 0.4, 2.2, 7.5, 0.01]
    S, D, K = mh.assemble_matrices_batched(_gfs(mh, mask), m, ks, formulation=3)
    S, D, K = S.cpu().numpy(), D.cpu().numpy(), K.cpu().numpy()
    for q, k in enumerate(ks):
        So, Do = oracle.assemble(mask, m, k, True)
        _, Ko = oracle.assemble(mask, m, k, False)
        assert relmax(S[q], So) < 1e-12, ("S", k)
        assert relmax(D[q], Do) < 1e-12, ("D", k)
        assert relmax(K[q], Ko) < 1e-12, ("K", k)


def test_assembly_ragged_sizes_and_matrix_api(mh, oracle, gpu):
    # sizes that are not multiples of the 64 x 64 tile, through the
    # reference-named single-wavenumber API
    for nt, nph in [(2, 3), (4, 7), (10, 13), (18, 25)]:
        m = mh.hemisphere(1.3, center=(0.2, -0.1, 0.0), ntheta=nt, nphi=nph).mesh
        S, D = mh.assemble_matrix_wu(m, 0.9)
        So, Do = oracle.assemble(GF_WU, m, 0.9, True)
        assert relmax(S.cpu().numpy(), So) < 1e-12
        assert relmax(D.cpu().numpy(), Do) < 1e-12
        S2, Kt = mh.assemble_matrices((mh.Rankine(), mh.RankineReflected(), mh.GFWu()), m, 0.9,
                                      direct=False)
        _, Ko = oracle.assemble(GF_WU, m, 0.9, False)
        assert relmax(Kt.cpu().numpy(), Ko) < 1e-12


@pytest.mark.parametrize("n,nrhs", [(1, 2), (5, 2), (31, 2), (32, 1), (33, 2), (64, 2), (100, 2),
                                     (127, 3), (128, 4), (129, 5), (257, 2), (700, 6)])
def test_lu_factor_and_solve(mh, gpu, n, nrhs):
    # slabs of 128 rows and 32-column blocks: sizes around both, and more
    # right-hand sides than one launch takes (4)
    torch = gpu
    rng = np.random.default_rng(n)
    batch = 3
    A = rng.standard_normal((batch, n, n)) + 1j * rng.standard_normal((batch, n, n))
    Bm = rng.standard_normal((batch, n, nrhs)) + 1j * rng.standard_normal((batch, n, nrhs))
    dA = torch.from_numpy(np.ascontiguousarray(A.transpose(0, 2, 1))).cuda()  # column-major
    piv = torch.empty((batch, n), dtype=torch.int64, device="cuda")
    ctx = mh._lib.context(0)
    L = mh._lib.lib()
    mh._lib.check(L.mh_lu_factor(ctx, n, batch, dA.data_ptr(), n, n * n, piv.data_ptr()))
    for trans in (0, 1):
        dB = torch.from_numpy(np.ascontiguousarray(Bm.transpose(0, 2, 1))).cuda()
        mh._lib.check(L.mh_lu_solve(ctx, n, batch, dA.data_ptr(), n, n * n, piv.data_ptr(), trans,
                                    nrhs, dB.data_ptr(), n, n * nrhs))
        X = dB.cpu().numpy().transpose(0, 2, 1)
        for b in range(batch):
            M = A[b].T if trans else A[b]
            ref = np.linalg.solve(M, Bm[b])
            assert relmax(X[b], ref) < 1e-10, (n, trans, b)


def test_lu_matches_oracle_pivots(mh, oracle, gpu):
    torch = gpu
    m = mh.hemisphere(1.0, ntheta=16, nphi=16).mesh
    _, D = oracle.assemble(GF_WU, m, 0.8, True)
    LUo, pivo, info = oracle.zgetrf(D)
    n = m.nfaces
    dA = torch.from_numpy(np.ascontiguousarray(D.T)).cuda()
    piv = torch.empty(n, dtype=torch.int64, device="cuda")
    mh._lib.check(mh._lib.lib().mh_lu_factor(mh._lib.context(0), n, 1, dA.data_ptr(), n, n * n,
                                             piv.data_ptr()))
    assert np.array_equal(piv.cpu().numpy(), pivo)
    assert relmax(dA.cpu().numpy().T, LUo) < 1e-12


@pytest.mark.parametrize("n", [700, 1100, 2100])
def test_lu_exact_factors_with_pivot_ties(mh, gpu, n):
    # A = L0 U0 with L0 unit lower in {-1, 0, 1} and U0 upper with entries in
    # {0, +-1, +-i} and diagonal in {+-1, +-i, +-1+-i}: every step of the
    # elimination is exact in FP64 (Gaussian integers, reciprocals of the
    # pivots are exact), and column c of the active matrix is
    # L0[c:, c] U0[c, c], so every row with L0[i, c] != 0 ties with row c in
    # |re| + |im|.  Partial pivoting that takes the first row of a tie (LAPACK's
    # izamax) keeps the diagonal: ipiv = 0..n-1 and LU = L0 + U0 - I exactly.
    # The sizes take the three leaf shapes (1, 2 and 4 rows per thread).
    torch = gpu
    rng = np.random.default_rng(n)
    L0 = np.tril(rng.integers(-1, 2, (n, n)), -1).astype(np.complex128) + np.eye(n)
    units = np.array([1, -1, 1j, -1j, 1 + 1j, 1 - 1j, -1 + 1j, -1 - 1j])
    U0 = np.triu(np.array([0, 1, -1, 1j, -1j])[rng.integers(0, 5, (n, n))], 1) + np.diag(units[rng.integers(0, 8, n)])
    A = L0 @ U0
    assert np.abs(A.real).max() < 2 ** 20  # products and sums stay exact
    dA = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
    piv = torch.empty(n, dtype=torch.int64, device="cuda")
    mh._lib.check(mh._lib.lib().mh_lu_factor(mh._lib.context(0), n, 1, dA.data_ptr(), n, n * n,
                                             piv.data_ptr()))
    assert np.array_equal(piv.cpu().numpy(), np.arange(n))
    LU = dA.cpu().numpy().T
    assert np.array_equal(LU, np.tril(L0, -1) + U0)


def test_lu_pivot_tie_across_waves_and_slots(mh, gpu):
    # equal |re| + |im| = 3 in four rows of the first column of a 1100 x 1100
    # matrix (panel rows 1030 and 1090 lie in the second register slot of
    # threads 6 and 66, rows 300 and 700 in the first slot of threads in
    # later waves); the first row, 300, is the pivot
    torch = gpu
    n = 1100
    rng = np.random.default_rng(11)
    A = 0.05 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    A[:, 0] *= 0.1
    A[1030, 0], A[300, 0], A[700, 0], A[1090, 0] = 2.0 + 1.0j, -1.0 - 2.0j, 3.0j, -3.0
    dA = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
    piv = torch.empty(n, dtype=torch.int64, device="cuda")
    ctx = mh._lib.context(0)
    L = mh._lib.lib()
    mh._lib.check(L.mh_lu_factor(ctx, n, 1, dA.data_ptr(), n, n * n, piv.data_ptr()))
    assert int(piv[0]) == 300
    b = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    dB = torch.from_numpy(np.ascontiguousarray(b.T)).cuda()
    mh._lib.check(L.mh_lu_solve(ctx, n, 1, dA.data_ptr(), n, n * n, piv.data_ptr(), 0, 2,
                                dB.data_ptr(), n, 2 * n))
    assert relmax(dB.cpu().numpy().T, np.linalg.solve(A, b)) < 1e-10


def test_singular_matrix_reports_error(mh, gpu):
    torch = gpu
    n = 40
    A = np.random.default_rng(1).standard_normal((n, n)).astype(np.complex128)
    A[:, 7] = 0.0
    dA = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
    piv = torch.empty(n, dtype=torch.int64, device="cuda")
    with pytest.raises(ArithmeticError):
        mh._lib.check(mh._lib.lib().mh_lu_factor(mh._lib.context(0), n, 1, dA.data_ptr(), n, n * n,
                                                 piv.data_ptr()))


@pytest.mark.parametrize("direct", [True, False])
def test_solve_matches_oracle(mh, oracle, gpu, direct):
    m = mh.hemisphere(1.0, ntheta=14, nphi=14).mesh
    S, DK = oracle.assemble(GF_WU, m, 0.6, direct)
    rng = np.random.default_rng(2)
    bc = rng.standard_normal((m.nfaces, 3)) + 1j * rng.standard_normal((m.nfaces, 3))
    phi = mh.solve(DK, S, bc, direct=direct).cpu().numpy()
    ref = oracle.solve(DK, S, bc, direct)
    assert relmax(phi, ref) < 1e-11
    phi1 = mh.solve(DK, S, bc[:, 0], direct=direct).cpu().numpy()
    assert relmax(phi1, ref[:, 0]) < 1e-11


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("direct", [True, False])
def test_freq_sweep_matches_oracle(mh, oracle, gpu, direct, precision):
    # mixed: FP32 LU + FP64 iterative refinement to a 1e-10 relative residual
    # (bound 1e-8 on the outputs; the mode's specification is 1e-6)
    m = mh.hemisphere(1.0, ntheta=12, nphi=12).mesh
    omegas = [0.3, 0.9, 1.03, 2.0, 3.0]
    dofs = mh.rigid_dofs(m, ["Surge", "Heave", "Pitch"])
    betas = [0.0, 0.7]
    A, B, F = mh.freq_sweep(m, omegas, dofs, betas, direct=direct, rho=1025.0, precision=precision)
    tol = 1e-10 if precision == "fp64" else 1e-8
    for w, om in enumerate(omegas):
        Ao, Bo, Fo = oracle.freq_solve(GF_WU, m, om, dofs, betas, direct=direct, rho=1025.0)
        assert relmax(A[w], Ao) < tol, om
        assert relmax(B[w], Bo) < tol, om
        assert relmax(F[w].T, Fo) < tol, om


def test_mixed_precision_2048_panels(mh, gpu):
    """2048 panels: mixed-precision sweep against the FP64 sweep; the mode is
    specified to 1e-6 relative, refinement to a 1e-10 residual gives ~1e-10."""
    m = mh.hemisphere(1.0, ntheta=64, nphi=64).mesh
    omegas = [0.3, 1.2, 3.0]
    A, B, F = mh.freq_sweep(m, omegas, [[0, 0, 1.0]], [0.0], rho=1025.0)
    Am, Bm, Fm = mh.freq_sweep(m, omegas, [[0, 0, 1.0]], [0.0], rho=1025.0, precision="mixed")
    assert relmax(Am, A) < 1e-8 and relmax(Bm, B) < 1e-8 and relmax(Fm, F) < 1e-8


def test_mixed_precision_above_2048_panels(mh, gpu):
    """2100 panels: FP32 leaf panels taller than 2048 rows (8 rows per thread)
    and a last block column of 52; mixed against FP64, with gradients."""
    body = mh.hemisphere(1.0, ntheta=70, nphi=60)
    m = body.mesh
    assert m.nfaces == 2100
    omegas = [0.5, 2.0]
    T = mh.shape_tangents(m, [body.dvertices["radius"], None], nomega=2, omega_param=1)
    r64 = mh.freq_sweep_grad(m, omegas, [[0, 0, 1.0]], [0.0], T, rho=1025.0)
    r32 = mh.freq_sweep_grad(m, omegas, [[0, 0, 1.0]], [0.0], T, rho=1025.0, precision="mixed")
    for a, b in zip(r32, r64):
        assert relmax(a, b) < 1e-8


@pytest.mark.parametrize("shape", [(10, 26), (12, 43), (10, 77)])
def test_mixed_precision_narrow_last_block_column(mh, gpu, shape):
    """130, 258 and 385 panels: the last block column of the FP32 LU is 2, 2
    and 1 wide (one-column leaves, triangular solves of width 1-2 and of the
    full 128); mixed against FP64, with gradients."""
    body = mh.hemisphere(1.0, ntheta=shape[0], nphi=shape[1])
    m = body.mesh
    assert m.nfaces % 128 in (1, 2)
    omegas = [0.4, 1.7]
    T = mh.shape_tangents(m, [body.dvertices["radius"], None], nomega=2, omega_param=1)
    r64 = mh.freq_sweep_grad(m, omegas, [[0, 0, 1.0]], [0.0], T, rho=1025.0)
    r32 = mh.freq_sweep_grad(m, omegas, [[0, 0, 1.0]], [0.0], T, rho=1025.0, precision="mixed")
    for a, b in zip(r32, r64):
        assert relmax(a, b) < 1e-8


@pytest.mark.parametrize("direct", [True, False])
def test_twelve_dofs_four_headings(mh, oracle, gpu, direct):
    """Two bodies with six rigid dofs each and four headings: 16 right-hand
    sides and 12 force weights, which take the widest matrix-vector variants
    (16 columns, S b and S^T c); FP64 sweep against the oracle."""
    b1 = mh.hemisphere(1.0, center=(0.0, 0.0, 0.0), ntheta=8, nphi=8)
    b2 = mh.hemisphere(0.8, center=(3.0, 0.0, 0.0), ntheta=8, nphi=8)
    m = mh.combine_meshes([b1.mesh, b2.mesh])
    names = ["Surge", "Sway", "Heave", "Roll", "Pitch", "Yaw"]
    n1, n2 = b1.mesh.nfaces, b2.mesh.nfaces
    dofs = [np.vstack([d, np.zeros((n2, 3))]) for d in mh.rigid_dofs(b1.mesh, names)]
    dofs += [np.vstack([np.zeros((n1, 3)), d])
             for d in mh.rigid_dofs(b2.mesh, names, rotation_center=(3.0, 0.0, 0.0))]
    betas = [0.0, 0.5, 1.0, 2.0]
    omegas = [0.8, 1.6]
    A, B, F = mh.freq_sweep(m, omegas, dofs, betas, direct=direct, rho=1025.0)
    for w, om in enumerate(omegas):
        Ao, Bo, Fo = oracle.freq_solve(GF_WU, m, om, dofs, betas, direct=direct, rho=1025.0)
        assert relmax(A[w], Ao) < 1e-10, om
        assert relmax(B[w], Bo) < 1e-10, om
        assert relmax(F[w].T, Fo) < 1e-10, om


def test_reference_entry_points(mh, oracle, gpu):
    # calculate_radiation_forces / DiffractionForce (src/waves.jl:29-52)
    m = mh.hemisphere(1.0, ntheta=10, nphi=10).mesh
    mh.set_rho(1025.0)
    try:
        AB = mh.calculate_radiation_forces(m, [0, 0, 1.0], 1.03)
        Fd = mh.DiffractionForce(m, 1.03, [0, 0, 1.0])
    finally:
        mh.set_rho(1000.0)
    Ao, Bo, Fo = oracle.freq_solve(GF_WU, m, 1.03, [[0, 0, 1.0]], [0.0], rho=1025.0)
    assert abs(AB[0] - Ao[0, 0]) < 1e-10 * abs(Ao[0, 0])
    assert abs(AB[1] - Bo[0, 0]) < 1e-10 * abs(Bo[0, 0])
    assert abs(Fd - Fo[0, 0]) < 1e-10 * abs(Fo[0, 0])


def test_host_entry_points(mh, oracle, gpu):
    import ctypes as C

    m = mh.hemisphere(1.0, ntheta=10, nphi=10).mesh
    L = mh._lib.lib()
    out = np.zeros(2)
    args = [m.nfaces, m.nvertices] + [a.ctypes.data for a in (m.vertices, m.faces, m.centers,
                                                             m.normals, m.areas, m.radii)]
    dof = np.array([0.0, 0.0, 1.0])
    mh._lib.check(L.mh_calculate_radiation_forces_host(*args, dof.ctypes.data, 1.03, 9.81, 1025.0,
                                                       out.ctypes.data))
    Ao, Bo, Fo = oracle.freq_solve(GF_WU, m, 1.03, [[0, 0, 1.0]], [0.0], rho=1025.0)
    assert abs(out[0] - Ao[0, 0]) < 1e-10 * abs(Ao[0, 0])
    assert abs(out[1] - Bo[0, 0]) < 1e-10 * abs(Bo[0, 0])
    mh._lib.check(L.mh_diffraction_force_host(*args, dof.ctypes.data, 1.03, 9.81, 1025.0,
                                              out.ctypes.data))
    assert abs(complex(out[0], out[1]) - Fo[0, 0]) < 1e-10 * abs(Fo[0, 0])


def test_hulme_hemisphere_heave_gpu(mh, gpu):
    # test/consistency_with_analytical_solutions.jl:169-189, all 25 wavenumbers in one sweep
    K = np.array([0.05, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 1.0, 1.2, 1.4, 1.6, 1.8, 2.0,
                  2.5, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0, 9.0, 10.0])
    A_h = [0.8764, 0.8627, 0.7938, 0.7157, 0.6452, 0.5861, 0.5381, 0.4999, 0.4698, 0.4464, 0.4284,
           0.4047, 0.3924, 0.3871, 0.3864, 0.3884, 0.3988, 0.4111, 0.4322, 0.4471, 0.4574, 0.4647,
           0.4700, 0.4740, 0.4771]
    B_h = [0.1036, 0.1816, 0.2793, 0.3254, 0.3410, 0.3391, 0.3271, 0.3098, 0.2899, 0.2691, 0.2484,
           0.2096, 0.1756, 0.1469, 0.1229, 0.1031, 0.0674, 0.0452, 0.0219, 0.0116, 0.0066, 0.0040,
           0.0026, 0.0017, 0.0012]
    rho = 1025.0
    m = mh.hemisphere(1.0, ntheta=10, nphi=10).mesh
    w = np.sqrt(K * 9.8)
    A, B, _ = mh.freq_sweep(m, w, [[0, 0, 1.0]], rho=rho)
    An = A[:, 0, 0] / ((2 / 3) * math.pi * rho)
    Bn = B[:, 0, 0] / ((2 / 3) * math.pi * rho) / w
    nrm = lambda a, b: np.linalg.norm(a - b) <= 0.1 * max(np.linalg.norm(a), np.linalg.norm(b))
    assert nrm(An, np.array(A_h)) and nrm(Bn, np.array(B_h))


def test_bad_mesh_index_is_rejected(mh, gpu):
    m = mh.hemisphere(1.0, ntheta=4, nphi=4).mesh
    dm = mh.DeviceMesh(m)
    dm.faces[0, 0] = 10_000  # out of range, caught by the device check
    with pytest.raises(ValueError):
        mh.assemble_matrix_wu(dm, 1.0)


def test_2048_panel_assembly_and_solve_properties(mh, oracle, gpu):
    """N = 2048 (64 x 64 hemisphere): assembly of sampled columns against the
    oracle, and the batched LU solve through its residual."""
    torch = gpu
    m = mh.hemisphere(1.0, ntheta=64, nphi=64).mesh
    n = m.nfaces
    ks = [0.3 ** 2 / 9.81, 3.0 ** 2 / 9.81]
    S, D = mh.assemble_matrices_batched(_gfs(mh, GF_WU), m, ks)
    So, Do = oracle.assemble(GF_WU, m, ks[1], True)
    assert relmax(S[1].cpu().numpy(), So) < 1e-12
    assert relmax(D[1].cpu().numpy(), Do) < 1e-12
    # residual of the batched LU on both matrices
    Dm = D.transpose(1, 2).contiguous()  # column-major copy
    D0 = Dm.clone()
    piv = torch.empty((2, n), dtype=torch.int64, device="cuda")
    L = mh._lib.lib()
    ctx = mh._lib.context(0)
    mh._lib.check(L.mh_lu_factor(ctx, n, 2, Dm.data_ptr(), n, n * n, piv.data_ptr()))
    rng = np.random.default_rng(0)
    b = torch.from_numpy(rng.standard_normal((2, 2, n)) + 1j * rng.standard_normal((2, 2, n))).cuda()
    x = b.clone()
    mh._lib.check(L.mh_lu_solve(ctx, n, 2, Dm.data_ptr(), n, n * n, piv.data_ptr(), 0, 2, x.data_ptr(),
                                n, 2 * n))
    for q in range(2):
        Aq = D0[q].transpose(0, 1)  # logical [i, j]
        res = Aq @ x[q].transpose(0, 1) - b[q].transpose(0, 1)
        rel = (torch.linalg.norm(res) / torch.linalg.norm(b[q])).item()
        assert rel < 1e-12, rel


def test_any_green_function_tuple_sums_like_the_reference(mh, oracle, gpu):
    """The reference sums any tuple of Green functions
    (abstract_greens_function.jl:18-83), repeats and several wave terms
    included.  Two wave terms against the oracle's sum of the same terms;
    a repeated term against twice the oracle's term (the free term 1/2 of D
    and K counted once); the pullback of a two-wave-term tuple against the
    sum of the pullbacks of its parts; the sweep refuses such a tuple."""
    m = mh.hemisphere(1.0, ntheta=10, nphi=10).mesh
    k = 0.8
    both = (mh.Rankine(), mh.RankineReflected(), mh.GFWu(), mh.ExactGuevelDelhommeau())
    for direct in (True, False):
        S, DK = mh.assemble_matrices(both, m, k, direct=direct)
        So, DKo = oracle.assemble(oracle.RANKINE | oracle.RANKINE_REFLECTED | oracle.WU | oracle.DELHOMMEAU,
                                  m, k, direct)
        assert relmax(S.cpu().numpy(), So) < 1e-11 and relmax(DK.cpu().numpy(), DKo) < 1e-11, direct
        twice = (mh.Rankine(), mh.RankineReflected(), mh.Rankine())
        S2, DK2 = mh.assemble_matrices(twice, m, k, direct=direct)
        Sa, DKa = oracle.assemble(oracle.RANKINE | oracle.RANKINE_REFLECTED, m, k, direct)
        Sb, DKb = oracle.assemble(oracle.RANKINE, m, k, direct)
        eye = np.eye(m.nfaces)
        assert relmax(S2.cpu().numpy(), Sa + Sb) < 1e-12
        assert relmax(DK2.cpu().numpy(), DKa + DKb - 0.5 * eye) < 1e-12
    # pair evaluations: a repeated term is twice the term
    e1, e2 = mh.element(m, 3), mh.element(m, 17)
    one = mh.integral(mh.Rankine(), e1, e2)
    two = mh.integral((mh.Rankine(), mh.Rankine()), e1, e2)
    assert abs(two - 2 * one) <= 1e-14 * abs(one)
    # the pullback is linear in the tuple
    rng = np.random.default_rng(2)
    Sbar = rng.normal(size=(m.nfaces, m.nfaces)) + 1j * rng.normal(size=(m.nfaces, m.nfaces))
    Dbar = rng.normal(size=(m.nfaces, m.nfaces)) + 1j * rng.normal(size=(m.nfaces, m.nfaces))
    got = mh.assemble_matrices_vjp(both, m, [k], Sbar, Dbar)
    pa = mh.assemble_matrices_vjp((mh.Rankine(), mh.RankineReflected(), mh.GFWu()), m, [k], Sbar, Dbar)
    pb = mh.assemble_matrices_vjp((mh.ExactGuevelDelhommeau(),), m, [k], Sbar, Dbar)
    for key in got:
        assert relmax(got[key], pa[key] + pb[key]) < 1e-13, key
    with pytest.raises(ValueError, match="at most once"):
        mh.freq_sweep(m, [1.0], [[0, 0, 1.0]], green_functions=both)
