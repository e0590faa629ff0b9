"""GPU Froude-Krylov and excitation forces, the coefficient API, and the
pins of the diffraction force that do not rest on the oracle alone
(Haskind's relation and Hulme's surge table), through libmhip."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GF_WU = 1 | 2 | 8


def relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_froude_krylov_matches_oracle(mh, oracle, gpu):
    body = mh.hemisphere(1.3, center=(0.4, -0.2, 0.0), ntheta=16, nphi=14)
    m = body.mesh
    dofs = mh.rigid_dofs(m, ["Surge", "Sway", "Heave", "Roll", "Pitch", "Yaw"], (0.1, 0.0, -0.3))
    omegas = [0.4, 1.1, 2.7]
    betas = [0.0, 0.7, math.pi]
    FK = mh.froude_krylov(m, omegas, dofs, betas=betas, rho=1025.0)
    for w, om in enumerate(omegas):
        ref = oracle.froude_krylov(m, om, dofs, betas=betas, rho=1025.0)  # [i, b]
        assert relmax(FK[w].T, ref) < 1e-12, (om, relmax(FK[w].T, ref))
    # the reference's single-heading helper
    f = mh.FroudeKrylovForce(m, 1.1, [0, 0, 1.0], 0.7)
    assert abs(f - FK[1, 1, 2] * 1000.0 / 1025.0) < 1e-10 * abs(f)


def test_froude_krylov_gradient_vs_finite_difference(mh, oracle, gpu):
    # theta = (radius, x, omega), central differences of the oracle, h = 1e-6
    res = 12
    theta0 = [1.2, 0.3, 1.4]
    body = mh.hemisphere(theta0[0], center=(theta0[1], 0.0, 0.0), ntheta=res, nphi=res)
    m = body.mesh
    T = mh.shape_tangents(m, [body.dvertices["radius"], body.dvertices["x"], None], nomega=1,
                          omega_param=2, dof_axes=[None, None, (0.0, 1.0, 0.0)])
    dofs = lambda mm: mh.rigid_dofs(mm, ["Surge", "Heave", "Pitch"])
    betas = [0.0, 1.0]
    FK, dFK = mh.froude_krylov_grad(m, [theta0[2]], dofs(m), betas=betas, tangents=T, rho=1025.0)
    h = 1e-6
    for p in range(3):
        outs = []
        for s in (+1, -1):
            t = list(theta0)
            t[p] += s * h
            b = mh.hemisphere(t[0], center=(t[1], 0.0, 0.0), ntheta=res, nphi=res)
            outs.append(oracle.froude_krylov(b.mesh, t[2], dofs(b.mesh), betas=betas, rho=1025.0))
        fd = (outs[0] - outs[1]) / (2 * h)
        assert relmax(dFK[0, p].T, fd) < 1e-6, (p, relmax(dFK[0, p].T, fd))


@pytest.mark.parametrize("res", [10, 20])
def test_haskind_relation_through_the_sweep(mh, gpu, res):
    # B_jj = k omega / (4 pi rho g^2) int |X_j|^2 d beta with X = F_FK + F_D:
    # heave B33 = k w |X3|^2 / (2 rho g^2), surge B11 = k w |X1(0)|^2 / (4 rho g^2)
    rho, g = 1025.0, 9.81
    m = mh.hemisphere(1.0, ntheta=res, nphi=res).mesh
    ks = np.array([0.3, 0.8, 1.5])
    oms = np.sqrt(ks * g)
    dofs = [[1.0, 0, 0], [0, 0, 1.0]]
    A, B, F = mh.freq_sweep(m, oms, dofs, betas=[0.0], rho=rho, g=g)
    FK = mh.froude_krylov(m, oms, dofs, betas=[0.0], rho=rho, g=g)
    X = F + FK  # [w, b, i]
    b11 = ks * oms * np.abs(X[:, 0, 0]) ** 2 / (4 * rho * g * g)
    b33 = ks * oms * np.abs(X[:, 0, 1]) ** 2 / (2 * rho * g * g)
    gap = max(np.max(np.abs(b11 / B[:, 0, 0] - 1)), np.max(np.abs(b33 / B[:, 1, 1] - 1)))
    print(f"haskind res={res} gap={gap:.4f}", flush=True)
    assert gap < (0.06 if res == 10 else 0.03), gap


def test_hulme_hemisphere_surge_gpu(mh, gpu):
    from test_oracle import HULME_A_SURGE, HULME_B_SURGE, HULME_K_SURGE, _isapprox

    rho = 1025.0
    m = mh.hemisphere(1.0, ntheta=10, nphi=10).mesh
    K = np.array(HULME_K_SURGE)
    om = np.sqrt(K * 9.8)
    A, B, _ = mh.freq_sweep(m, om, [[1.0, 0, 0]], rho=rho)
    A_b = A[:, 0, 0] / ((2 / 3) * math.pi * rho)
    B_b = B[:, 0, 0] / ((2 / 3) * math.pi * rho) / om
    assert _isapprox(A_b, HULME_A_SURGE, 0.1)
    assert _isapprox(B_b, HULME_B_SURGE, 0.1)


def test_compute_hydrodynamic_coefficients_vs_oracle(mh, oracle, gpu):
    # src/problems_and_results.jl:105-243 through the batched sweep, against
    # the oracle problem by problem
    m = mh.hemisphere(1.0, ntheta=12, nphi=12).mesh
    body = mh.FloatingBody(m, ["Surge", "Heave", "Pitch"], "buoy", rotation_center=(0.0, 0.0, -0.2))
    params = {"wave_frequencies": [0.6, 1.2, 2.0], "wave_directions": [0.0, 0.5],
              "radiating_dofs": ["Heave", "Pitch"], "influenced_dofs": ["Surge", "Heave", "Pitch"]}
    data = mh.compute_hydrodynamic_coefficients(params, body)
    fields = [body.dofs[n] for n in params["influenced_dofs"]]
    rad = [params["influenced_dofs"].index(n) for n in params["radiating_dofs"]]
    for w, om in enumerate(params["wave_frequencies"]):
        A, B, Fd = oracle.freq_solve(GF_WU, m, om, fields, betas=params["wave_directions"])
        FK = oracle.froude_krylov(m, om, fields, betas=params["wave_directions"])
        assert relmax(data["added_mass"][:, :, w], A[:, rad]) < 1e-10
        assert relmax(data["radiation_damping"][:, :, w], B[:, rad]) < 1e-10
        assert relmax(data["diffraction_force"][:, w, :], Fd) < 1e-10
        assert relmax(data["Froude_Krylov_force"][:, w, :], FK) < 1e-12
        assert relmax(data["excitation_force"][:, w, :], Fd + FK) < 1e-10
    stack = mh.create_DimStack(data, params, body)
    assert stack["added_mass"].dims == ("influenced_dofs", "radiating_dofs", "wave_frequencies")
    a = stack["added_mass"].sel(influenced_dofs="Heave", radiating_dofs="Heave", wave_frequencies=1.2)
    assert a == data["added_mass"][1, 0, 1]
    # one problem at a time gives the same forces (src/solve.jl:4-14)
    p = mh.RadiationProblem(body, 1.2, "Heave", ["Heave"])
    r = mh.solve_problem(p)
    assert abs(r.forces["Heave"].real / 1.2 ** 2 - data["added_mass"][1, 0, 1]) < 1e-9 * abs(
        data["added_mass"][1, 0, 1])
