"""GPU gradients of A, B, F (adjoint solve + geometry tangents) against
central finite differences of the FP64 CPU oracle.

The reference's own criterion for its gradients is agreement with central
finite differences at atol = rtol = 1e-6 (test/outputs_differentiation.jl:
64-67, 139-150).  Here the step is h = 1e-5 (relative for lengths), the
finite differences come from the oracle on regenerated meshes, and the
tolerance is 1e-6 relative to the largest derivative of each quantity.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GF_WU = 1 | 2 | 8


def pair_body(mh, r1, r2, d, res):
    """Two floating hemispheres, centres (0,0,0) and (d,0,0)
    (the adjoint-gradient configuration); d vertices / d (r1, r2, d)."""
    b1 = mh.hemisphere(r1, center=(0.0, 0.0, 0.0), ntheta=res, nphi=res)
    b2 = mh.hemisphere(r2, center=(d, 0.0, 0.0), ntheta=res, nphi=res)
    mesh = mh.combine_meshes([b1.mesh, b2.mesh])
    nv1, nv2 = b1.mesh.nvertices, b2.mesh.nvertices
    z1, z2 = np.zeros((nv1, 3)), np.zeros((nv2, 3))
    dV = [np.vstack([b1.dvertices["radius"], z2]), np.vstack([z1, b2.dvertices["radius"]]),
          np.vstack([z1, b2.dvertices["x"]])]
    return mesh, dV


def fd(oracle, make, theta, m, omega_of, dofs_of, betas, direct, h, rho=1025.0):
    """central difference of the oracle's (A, B, F) along parameter m"""
    out = []
    for s in (+1, -1):
        t = list(theta)
        t[m] += s * h
        mesh = make(t)
        A, B, F = oracle.freq_solve(GF_WU, mesh, omega_of(t), dofs_of(mesh), betas, direct=direct,
                                    rho=rho)
        out.append((A, B, F))
    return [(out[0][q] - out[1][q]) / (2 * h) for q in range(3)]


def check(got, ref, tol, what, value=None):
    """|got - ref| <= tol * max|ref|, or, for derivatives that vanish (a
    translation of a single body leaves A and B unchanged), below the
    finite-difference noise floor 1e-9 * max|value| per unit step."""
    err = np.max(np.abs(got - ref))
    scale = np.max(np.abs(ref))
    floor = 1e-9 * np.max(np.abs(value)) if value is not None else 0.0
    assert err <= tol * scale or err <= floor, (what, err, scale, floor, got, ref)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("direct", [True, False])
def test_single_hemisphere_radius_position_omega(mh, oracle, gpu, direct, precision):
    res, omega = 10, 1.03
    theta0 = [1.0, 0.0, omega]  # radius, x, omega

    def make(t):
        return mh.hemisphere(t[0], center=(t[1], 0.0, 0.0), ntheta=res, nphi=res).mesh

    body = mh.hemisphere(1.0, ntheta=res, nphi=res)
    mesh = body.mesh
    dofs_of = lambda m: mh.rigid_dofs(m, ["Heave", "Surge"])
    betas = [0.0]
    T = mh.shape_tangents(mesh, [body.dvertices["radius"], body.dvertices["x"], None], nomega=1,
                          omega_param=2)
    A, B, F, dA, dB, dF = mh.freq_sweep_grad(mesh, [omega], dofs_of(mesh), betas, T, direct=direct,
                                             rho=1025.0, precision=precision)
    Ao, Bo, Fo = oracle.freq_solve(GF_WU, mesh, omega, dofs_of(mesh), betas, direct=direct, rho=1025.0)
    tol = 1e-10 if precision == "fp64" else 1e-8  # mixed: refinement to a 1e-10 residual
    check(A[0], Ao, tol, "A")
    check(F[0].T, Fo, tol, "F")
    for m in range(3):
        gA, gB, gF = fd(oracle, make, theta0, m, lambda t: t[2], dofs_of, betas, direct, 1e-5)
        check(dA[0, m], gA, 1e-6, ("dA", m), A)
        check(dB[0, m], gB, 1e-6, ("dB", m), B)
        check(dF[0, m].T, gF, 1e-6, ("dF", m), F)


def test_hemisphere_pair_seeded(mh, oracle, gpu):
    """pair of heave hemispheres, theta = (r1, r2, d, omega) drawn from a seed
    (r ~ U[0.5, 2], d ~ U[3, 10] redrawn until the bodies are 0.5 m apart,
    omega ~ U[0.5, 2])"""
    rng = np.random.default_rng(2024)
    while True:
        r1, r2 = rng.uniform(0.5, 2.0, 2)
        d = rng.uniform(3.0, 10.0)
        if d > r1 + r2 + 0.5:
            break
    omega = rng.uniform(0.5, 2.0)
    res = 8
    theta0 = [r1, r2, d, omega]
    make = lambda t: pair_body(mh, t[0], t[1], t[2], res)[0]
    mesh, dV = pair_body(mh, r1, r2, d, res)
    heave = np.zeros((mesh.nfaces, 3))
    heave[:, 2] = 1.0
    dofs_of = lambda m: [np.tile([0, 0, 1.0], (m.nfaces, 1))]
    T = mh.shape_tangents(mesh, dV + [None], nomega=1, omega_param=3)
    A, B, F, dA, dB, dF = mh.freq_sweep_grad(mesh, [omega], [heave], [0.0], T, rho=1025.0)
    for m in range(4):
        gA, gB, gF = fd(oracle, make, theta0, m, lambda t: t[3], dofs_of, [0.0], True,
                        1e-5 * max(1.0, abs(theta0[m])))
        check(dA[0, m], gA, 1e-6, ("dA", m), A)
        check(dB[0, m], gB, 1e-6, ("dB", m), B)
        check(dF[0, m].T, gF, 1e-6, ("dF", m), F)


def test_rotational_dof_and_several_frequencies(mh, oracle, gpu):
    # pitch about an offset centre: the dof field moves with the centres
    res = 8
    rc = np.array([0.3, 0.2, 0.0])
    body = mh.hemisphere(1.2, ntheta=res, nphi=res)
    mesh = body.mesh
    omegas = [0.6, 1.4]
    dofs_of = lambda m: mh.rigid_dofs(m, ["Heave", "Pitch"], rotation_center=rc)
    T = mh.shape_tangents(mesh, [body.dvertices["radius"]], dof_axes=[None, [0, 1.0, 0]])
    A, B, F, dA, dB, dF = mh.freq_sweep_grad(mesh, omegas, dofs_of(mesh), [0.4], T, rho=1025.0)
    make = lambda t: mh.hemisphere(t[0], ntheta=res, nphi=res).mesh
    for w, om in enumerate(omegas):
        gA, gB, gF = fd(oracle, make, [1.2], 0, lambda t: om, dofs_of, [0.4], True, 1e-5)
        check(dA[w, 0], gA, 1e-6, ("dA", w), A)
        check(dB[w, 0], gB, 1e-6, ("dB", w), B)
        check(dF[w, 0].T, gF, 1e-6, ("dF", w), F)


def test_wavebot_shape_gradients(mh, oracle, gpu):
    # shape gradients of the paper's WEC hull (src/meshes.jl:458-479) w.r.t.
    # its four design variables, heave and surge, against central
    # differences of the oracle on regenerated meshes (h = 1e-6)
    p0 = {"r1": 0.88, "r2": 0.35, "d1": 0.16, "d2": 0.37}
    names = list(p0)
    kw = {"n_points": (3, 6, 3), "n_theta": 24}
    b = mh.wavebot_mesh(**p0, **kw)
    dofs = [[0.0, 0.0, 1.0], [1.0, 0.0, 0.0]]
    om = 2.5
    T = mh.shape_tangents(b.mesh, [b.dvertices[k] for k in names])
    A, B, F, dA, dB, dF = mh.freq_sweep_grad(b.mesh, [om], dofs, betas=[0.0], tangents=T, rho=1025.0)
    h = 1e-6
    for m, k in enumerate(names):
        outs = []
        for s in (+1, -1):
            p = dict(p0)
            p[k] += s * h
            outs.append(oracle.freq_solve(GF_WU, mh.wavebot_mesh(**p, **kw).mesh, om, dofs, [0.0], rho=1025.0))
        fA, fB, fF = [(outs[0][q] - outs[1][q]) / (2 * h) for q in range(3)]
        check(dA[0, m], fA, 1e-6, ("dA", k), A[0])
        check(dB[0, m], fB, 1e-6, ("dB", k), B[0])
        check(dF[0, m, 0], fF[:, 0], 1e-6, ("dF", k), F[0, 0])


@pytest.mark.parametrize("direct", [True, False])
def test_six_rigid_dofs_and_two_headings(mh, oracle, gpu, direct):
    """all six rigid-body dofs about an offset centre and two headings: eight
    right-hand sides, which the pair kernels take in column blocks of at most
    four (dofs 0-2 / 3-5, right-hand sides 0-3 / 4-7).  Radius, x position
    and omega against central differences of the oracle; the heave-pitch
    block against a narrower sweep of those two dofs and one heading."""
    res, omega = 8, 1.1
    rc = [0.1, -0.2, -0.3]
    names = ["Surge", "Sway", "Heave", "Roll", "Pitch", "Yaw"]
    axes = [None, None, None, [1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0]]
    theta0 = [1.0, 0.0, omega]
    make = lambda t: mh.hemisphere(t[0], center=(t[1], 0.0, 0.0), ntheta=res, nphi=res).mesh
    body = mh.hemisphere(1.0, ntheta=res, nphi=res)
    mesh = body.mesh
    dofs_of = lambda m: mh.rigid_dofs(m, names, rotation_center=rc)
    betas = [0.0, 0.9]
    dV = [body.dvertices["radius"], body.dvertices["x"], None]
    T = mh.shape_tangents(mesh, dV, nomega=1, omega_param=2, dof_axes=axes)
    A, B, F, dA, dB, dF = mh.freq_sweep_grad(mesh, [omega], dofs_of(mesh), betas, T, direct=direct, rho=1025.0)
    assert dA.shape == (1, 3, 6, 6) and dF.shape == (1, 3, 2, 6)
    Ao, Bo, Fo = oracle.freq_solve(GF_WU, mesh, omega, dofs_of(mesh), betas, direct=direct, rho=1025.0)
    check(A[0], Ao, 1e-10, "A")
    check(F[0].T, Fo, 1e-10, "F")
    for m in range(3):
        gA, gB, gF = fd(oracle, make, theta0, m, lambda t: t[2], dofs_of, betas, direct, 1e-5)
        check(dA[0, m], gA, 1e-6, ("dA", m), A)
        check(dB[0, m], gB, 1e-6, ("dB", m), B)
        check(dF[0, m].T, gF, 1e-6, ("dF", m), F)
    sub = [2, 4]
    T2 = mh.shape_tangents(mesh, dV, nomega=1, omega_param=2, dof_axes=[None, axes[4]])
    d6 = dofs_of(mesh)
    _, _, _, dA2, dB2, dF2 = mh.freq_sweep_grad(mesh, [omega], [d6[2], d6[4]], [0.9], T2, direct=direct,
                                                rho=1025.0)
    check(dA[0][:, sub][:, :, sub], dA2[0], 1e-12, "dA block")
    check(dB[0][:, sub][:, :, sub], dB2[0], 1e-12, "dB block")
    check(dF[0][:, 1:2][:, :, sub], dF2[0], 1e-12, "dF block")


def test_device_resident_inputs_give_the_same_sweep(mh, gpu):
    """Dof fields and tangents uploaded once (device_dof_fields,
    device_tangents) and passed to the sweeps give the results of the host
    inputs, bit for bit in the forward quantities."""
    body = mh.hemisphere(1.0, ntheta=12, nphi=12)
    m = body.mesh
    om = [0.5, 1.2, 2.4]
    dofs = [[0, 0, 1.0], [1.0, 0, 0]]
    T = mh.shape_tangents(m, [body.dvertices["radius"], None], nomega=len(om), omega_param=1)
    ref = mh.freq_sweep_grad(m, om, dofs, [0.0, 0.5], T)
    got = mh.freq_sweep_grad(m, om, mh.device_dof_fields(m, dofs), [0.0, 0.5], mh.device_tangents(T, m, 2))
    for g_, r_ in zip(got, ref):
        assert np.array_equal(g_, r_) or float(np.max(np.abs(g_ - r_))) <= 1e-13 * float(np.max(np.abs(r_)))
    A, B, F = mh.freq_sweep(m, om, mh.device_dof_fields(m, dofs), [0.0, 0.5])
    A0, B0, F0 = mh.freq_sweep(m, om, dofs, [0.0, 0.5])
    assert np.array_equal(A, A0) and np.array_equal(B, B0) and np.array_equal(F, F0)
