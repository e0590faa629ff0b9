"""Reverse mode of the assembly and the solve (mh_assemble_vjp, mh_solve_vjp)
against central finite differences: of the FP64 oracle's assembly, of numpy's
solve, and of the forward-mode sweep gradient (reverse and forward mode must
agree on the same directional derivative).

Cotangents follow Zygote's convention for complex arrays: the real input u
receives ubar = sum Re(conj(zbar) dz/du), so for any direction du,
<ubar, du> = d/dh L(u + h du) with L = sum Re(conj(zbar) z).  Tolerance:
1e-6 relative to the size of the derivative (the reference's own criterion
for its gradients is central differences at 1e-6,
test/outputs_differentiation.jl:64-67).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mesh_like(mh, m, V=None, C=None, N=None, A=None):
    return mh.Mesh(m.vertices if V is None else V, m.faces, m.centers if C is None else C,
                   m.normals if N is None else N, m.areas if A is None else A, m.radii)


def _loss(oracle, mask, mesh, k, direct, Sb, Db):
    S, DK = oracle.assemble(mask, mesh, k, direct)
    return float(np.sum((np.conj(Sb) * S).real) + np.sum((np.conj(Db) * DK).real))


def _two_bodies(mh, box=False):
    # a hemisphere and a second body close to it: near-field pairs across
    # bodies, degenerate pole panels (a repeated vertex), shared vertices
    b1 = mh.hemisphere(1.0, ntheta=8, nphi=8)
    if box:
        b2 = mh.box(1.0, 0.8, 0.6, center=(1.9, 0.3), nx=2, ny=2, nz=1)
    else:
        b2 = mh.hemisphere(0.7, center=(1.9, 0.3, 0.0), ntheta=6, nphi=8)
    return mh.combine_meshes([b1.mesh, b2.mesh])


# Coplanar pairs (a collocation point in the plane of a source panel: the
# walls and bottom of a box) sit at GZ = 0 in the near-field gradient
# (rankine.jl:83-115), which the reference writes with abs(GZ) and
# sign(GZ).  Automatic differentiation takes sign(0) = 0 there (Zygote's rule
# for abs and sign, and the dual numbers here), while a finite difference that
# moves a panel centre or normal alone (off the plane of its own vertices)
# sees the one-sided slopes.  The box is therefore checked along vertices,
# areas and k, where both agree; centres and normals on two hemispheres.
@pytest.mark.parametrize("mask,direct,box", [(1 | 2, True, False), (1 | 2, False, False), (1 | 2 | 8, True, False),
                                             (1 | 2 | 8, False, False), (1 | 4 | 8, True, False),
                                             (1 | 2 | 16, True, False), (1 | 2 | 8, True, True),
                                             (1 | 2 | 8, False, True)])
def test_assemble_vjp_vs_oracle_finite_differences(mh, oracle, gpu, mask, direct, box):
    m = _two_bodies(mh, box)
    n, nv = m.nfaces, m.nvertices
    rng = np.random.default_rng(11 + mask + 100 * direct)
    k = 0.9
    Sb = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    Db = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    gfs = _gfs(mh, mask)
    print(f"assemble_vjp mask={mask} direct={direct} box={box} n={n}", flush=True)
    bar = mh.assemble_matrices_vjp(gfs, m, [k], Sb, Db, direct=direct)
    h = 1e-6
    dirs = {
        "vertices": rng.standard_normal((nv, 3)),
        "centers": rng.standard_normal((n, 3)),
        "normals": rng.standard_normal((n, 3)),
        "areas": rng.standard_normal(n),
    }
    if box:
        del dirs["centers"], dirs["normals"]
    for name, du in dirs.items():
        def at(s):
            kw = {"V": m.vertices, "C": m.centers, "N": m.normals, "A": m.areas}
            key = {"vertices": "V", "centers": "C", "normals": "N", "areas": "A"}[name]
            kw[key] = kw[key] + s * h * du
            return _loss(oracle, mask, _mesh_like(mh, m, **kw), k, direct, Sb, Db)

        fd = (at(1) - at(-1)) / (2 * h)
        got = float(np.sum(bar[name] * du))
        assert abs(got - fd) <= 1e-6 * max(abs(fd), 1e-3 * abs(at(0))), (name, got, fd)
    fdk = (_loss(oracle, mask, m, k + h, direct, Sb, Db) - _loss(oracle, mask, m, k - h, direct, Sb, Db)) / (2 * h)
    if mask & (8 | 16):
        assert abs(bar["wavenumbers"][0] - fdk) <= 1e-6 * abs(fdk), (bar["wavenumbers"], fdk)
    else:
        assert bar["wavenumbers"][0] == 0.0 and fdk == 0.0


def _gfs(mh, mask):
    out = []
    for cls in (mh.Rankine, mh.RankineReflected, mh.RankineReflectedNegative, mh.GFWu, mh.ExactGuevelDelhommeau):
        if mask & cls.mask:
            out.append(cls())
    return tuple(out)


def test_assemble_vjp_several_wavenumbers(mh, gpu):
    # batched pullback = sum of single-wavenumber pullbacks; dk per item
    m = mh.hemisphere(1.0, ntheta=10, nphi=8).mesh
    n = m.nfaces
    rng = np.random.default_rng(5)
    ks = [0.3, 1.1, 2.4]
    Sb = rng.standard_normal((3, n, n)) + 1j * rng.standard_normal((3, n, n))
    Db = rng.standard_normal((3, n, n)) + 1j * rng.standard_normal((3, n, n))
    gfs = (mh.Rankine(), mh.RankineReflected(), mh.GFWu())
    tot = mh.assemble_matrices_vjp(gfs, m, ks, Sb, Db)
    parts = [mh.assemble_matrices_vjp(gfs, m, [k], Sb[w], Db[w]) for w, k in enumerate(ks)]
    for key in ("vertices", "centers", "normals", "areas"):
        s = sum(p[key] for p in parts)
        assert np.max(np.abs(tot[key] - s)) <= 1e-12 * np.max(np.abs(s)), key
    assert np.allclose(tot["wavenumbers"], [p["wavenumbers"][0] for p in parts], rtol=1e-12, atol=0)


def test_vertex_jacobian_like_reference(mh, oracle, gpu):
    # test/matrix_assembly_differentiation.jl:34-64: Jacobian of vec(real(S))
    # and vec(real(D)) w.r.t. mesh.vertices with centres, normals, areas held
    m = mh.hemisphere(1.0, ntheta=4, nphi=6).mesh
    n, nv = m.nfaces, m.nvertices
    gfs = (mh.Rankine(), mh.RankineReflected(), mh.GFWu())
    k = 1.0
    h = 1e-6
    fd = []
    for q in range(3 * nv):  # column-major vec of the nv x 3 vertex matrix
        dV = np.zeros(3 * nv)
        dV[q] = h
        dV = dV.reshape(3, nv).T
        Sp, Dp = oracle.assemble(1 | 2 | 8, _mesh_like(mh, m, V=m.vertices + dV), k, True)
        Sm, Dm = oracle.assemble(1 | 2 | 8, _mesh_like(mh, m, V=m.vertices - dV), k, True)
        fd.append(((Sp - Sm).real.reshape(-1, order="F") / (2 * h), (Dp - Dm).real.reshape(-1, order="F") / (2 * h)))
    JS_fd = np.stack([f[0] for f in fd], axis=1)
    JD_fd = np.stack([f[1] for f in fd], axis=1)
    JS = np.zeros((n * n, 3 * nv))
    JD = np.zeros((n * n, 3 * nv))
    for row in range(n * n):
        i, j = row % n, row // n
        E = np.zeros((n, n))
        E[i, j] = 1.0
        JS[row] = mh.assemble_matrices_vjp(gfs, m, [k], E, None)["vertices"].reshape(-1, order="F")
        JD[row] = mh.assemble_matrices_vjp(gfs, m, [k], None, E)["vertices"].reshape(-1, order="F")
    assert not np.any(np.isnan(JS)) and not np.any(np.isnan(JD))
    assert np.max(np.abs(JS - JS_fd)) <= 1e-6 * np.max(np.abs(JS_fd))
    assert np.max(np.abs(JD - JD_fd)) <= 1e-6 * np.max(np.abs(JD_fd))


@pytest.mark.parametrize("direct", [True, False])
@pytest.mark.parametrize("n,nrhs", [(3, 1), (37, 2), (300, 3)])
def test_solve_vjp_vs_finite_differences(mh, gpu, direct, n, nrhs):
    rng = np.random.default_rng(n + nrhs + direct)
    D = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    D += np.diag(np.sum(np.abs(D), axis=1))  # diagonally dominant, as the reference test
    S = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    b = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    pbar = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))

    def phi(D_, S_, b_):
        return np.linalg.solve(D_, S_ @ b_) if direct else S_ @ np.linalg.solve(D_, b_)

    def L(D_, S_, b_):
        return float(np.sum((np.conj(pbar) * phi(D_, S_, b_)).real))

    Dbar, Sbar, bbar = mh.solve_vjp(D, S, b, pbar, direct=direct)
    Dbar, Sbar, bbar = (x.cpu().numpy() for x in (Dbar, Sbar, bbar))
    # closed forms of the pullback (ImplicitAD's implicit_linear rrule)
    if direct:
        y = np.linalg.solve(D.conj().T, pbar)
        ref = (-y @ phi(D, S, b).conj().T, y @ b.conj().T, S.conj().T @ y)
    else:
        sig = np.linalg.solve(D, b)
        bb = np.linalg.solve(D.conj().T, S.conj().T @ pbar)
        ref = (-bb @ sig.conj().T, pbar @ sig.conj().T, bb)
    for got, r in zip((Dbar, Sbar, bbar), ref):
        assert np.max(np.abs(got - r)) <= 1e-10 * np.max(np.abs(r))
    # and central differences along random complex directions
    h = 1e-6
    for which, X in ((0, D), (1, S), (2, b)):
        dX = rng.standard_normal(X.shape) + 1j * rng.standard_normal(X.shape)
        args_p = [D, S, b]
        args_m = [D, S, b]
        args_p[which] = X + h * dX
        args_m[which] = X - h * dX
        fd = (L(*args_p) - L(*args_m)) / (2 * h)
        got = float(np.sum((np.conj((Dbar, Sbar, bbar)[which]) * dX).real))
        assert abs(got - fd) <= 1e-6 * max(abs(fd), 1.0), (which, got, fd)


@pytest.mark.parametrize("direct", [True, False])
def test_reverse_chain_matches_forward_sweep_gradient(mh, gpu, direct):
    # F = c . phi, phi = solve(assemble(mesh, k), bc): the reverse chain
    # (solve_vjp, then assemble_vjp, plus the O(N) partials of bc and c)
    # gives the same derivative along the radius as mh_freq_sweep_grad's
    # forward tangents.
    rho, g, om = 1025.0, 9.81, 1.3
    body = mh.hemisphere(1.1, ntheta=12, nphi=12)
    m = body.mesh
    k = om * om / g
    dof = np.tile([0.0, 0.0, 1.0], (m.nfaces, 1))
    T = mh.shape_tangents(m, [body.dvertices["radius"]])
    A, B, _, dA, dB, _ = mh.freq_sweep_grad(m, [om], [dof], tangents=T, direct=direct, rho=rho, g=g)
    dF_fwd = om * om * dA[0, 0, 0, 0] + 1j * om * dB[0, 0, 0, 0]
    gfs = (mh.Rankine(), mh.RankineReflected(), mh.GFWu())
    S, DK = mh.assemble_matrices(gfs, m, k, direct)
    S, DK = S.cpu().numpy(), DK.cpu().numpy()
    nd = np.sum(m.normals * dof, axis=1)
    bc = -1j * om * nd
    c = 1j * rho * om * (-nd) * m.areas
    phi = mh.solve(DK, S, bc, direct).cpu().numpy()
    for Fbar in (1.0, 1j):
        phibar = Fbar * np.conj(c)
        Dbar, Sbar, bbar = (x.cpu().numpy() for x in mh.solve_vjp(DK, S, bc, phibar, direct=direct))
        bar = mh.assemble_matrices_vjp(gfs, m, [k], Sbar, Dbar, direct=direct)
        # O(N) partials along the tangent: c and bc through normals and areas
        dN, dAr = T.dnormals[0], T.dareas[0]
        dnd = np.sum(dN * dof, axis=1)
        dc = 1j * rho * om * (-dnd) * m.areas + 1j * rho * om * (-nd) * dAr
        dbc = -1j * om * dnd
        rev = (np.sum(bar["vertices"] * T.dvertices[0]) + np.sum(bar["centers"] * T.dcenters[0]) +
               np.sum(bar["normals"] * dN) + np.sum(bar["areas"] * dAr) +
               np.sum((np.conj(bbar) * dbc).real) + (np.conj(Fbar) * np.sum(dc * phi)).real)
        fwd = (np.conj(Fbar) * dF_fwd).real
        assert abs(rev - fwd) <= 1e-8 * abs(dF_fwd), (Fbar, rev, fwd)
