"""The tabulated Delhommeau term (MH_GF_DELHOMMEAU_TABLE) against the exact
term it replaces (the device quadrature, itself pinned to the oracle and to
mpmath by test_gpu_delhommeau.py and test_oracle.py): pointwise over the
(r, z) domain, including the free surface, the origin and beyond the table,
and in the assembled matrices."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _wave(mh, torch, term, R, Z):
    L = mh._lib
    dr = torch.from_numpy(np.ascontiguousarray(R)).cuda()
    dz = torch.from_numpy(np.ascontiguousarray(Z)).cuda()
    out = torch.empty(3 * len(R), dtype=torch.complex128, device="cuda")
    L.check(L.lib().mh_wave_term_eval(L.context(0), term, len(R), dr.data_ptr(), dz.data_ptr(), out.data_ptr()))
    return out.cpu().numpy().reshape(3, -1)


def test_table_pointwise_against_exact_term(mh, gpu):
    r = np.concatenate([[0.0, 1e-9, 1e-6, 1e-4, 3e-3], np.linspace(0.01, 2.0, 57), np.linspace(2.01, 31.9, 211)])
    z = -np.concatenate([[1e-8, 1e-6, 1e-4, 3e-3], np.linspace(0.01, 2.0, 53), np.linspace(2.02, 31.9, 201)])
    R, Z = np.meshgrid(r, z, indexing="ij")
    R, Z = R.ravel(), Z.ravel()
    ex = _wave(mh, gpu, 16, R, Z)
    tb = _wave(mh, gpu, 32, R, Z)
    for q, name in enumerate(("G", "dG/dr", "dG/dz")):
        rel = np.abs(tb[q] - ex[q]) / np.maximum(np.abs(ex[q]), 1.0)
        print(f"table vs exact {name}: max {rel.max():.2e}, median {np.median(rel):.1e}", flush=True)
        assert rel.max() < 1e-8, name
    # imaginary parts are the closed forms 2 pi e^z J0(r), -2 pi e^z J1(r)
    from scipy.special import j0, j1

    assert np.max(np.abs(tb[0].imag - 2 * np.pi * np.exp(Z) * j0(R))) < 1e-12
    assert np.max(np.abs(tb[1].imag + 2 * np.pi * np.exp(Z) * j1(R))) < 1e-12
    # beyond the table the exact term itself
    Rb = np.array([33.0, 40.0, 5.0, 0.5])
    Zb = np.array([-0.5, -3.0, -36.0, -50.0])
    assert np.array_equal(_wave(mh, gpu, 32, Rb, Zb), _wave(mh, gpu, 16, Rb, Zb))


@pytest.mark.parametrize("direct", [True, False])
def test_table_assembly_against_exact(mh, gpu, direct):
    gpu_ = gpu
    m = mh.hemisphere(1.0, ntheta=40, nphi=40).mesh
    ks = np.array([0.009, 0.1, 0.4, 0.92, 3.0])
    ex = mh.assemble_matrices_batched((mh.Rankine(), mh.RankineReflected(), mh.ExactGuevelDelhommeau()), m, ks,
                                      direct)
    tb = mh.assemble_matrices_batched((mh.Rankine(), mh.RankineReflected(), mh.TabulatedDelhommeau()), m, ks,
                                      direct)
    for a, b, name in zip(tb, ex, ("S", "D/K")):
        err = (a - b).abs().max().item() / b.abs().max().item()
        print(f"table assembly {name} direct={direct}: {err:.2e}", flush=True)
        assert err < 1e-10, name
    gpu_.cuda.synchronize()


def test_table_gf_eval_and_gradients_use_exact_derivatives(mh, oracle, gpu):
    # per-pair evaluation through the table, and the gradient sweep (which
    # differentiates the exact term) against the oracle's central differences
    body = mh.hemisphere(1.0, ntheta=8, nphi=8)
    m = body.mesh
    e1 = {"center": m.centers[3]}
    for k in (0.3, 2.0):
        for j in range(0, m.nfaces, 7):
            e2 = mh.element(m, j)
            a = mh.integral(mh.TabulatedDelhommeau(), e1, e2, k)
            b = mh.integral(mh.ExactGuevelDelhommeau(), e1, e2, k)
            assert abs(a - b) <= 1e-8 * max(abs(b), 1.0)
    T = mh.shape_tangents(m, [body.dvertices["radius"]])
    gfs = (mh.Rankine(), mh.RankineReflected(), mh.TabulatedDelhommeau())
    A, B, F, dA, dB, dF = mh.freq_sweep_grad(m, [1.1], [[0, 0, 1.0]], [0.0], T, green_functions=gfs)
    h = 1e-6
    outs = [oracle.freq_solve(1 | 2 | 16, mh.hemisphere(1.0 + s * h, ntheta=8, nphi=8).mesh, 1.1, [[0, 0, 1.0]],
                              [0.0]) for s in (1, -1)]
    fA = (outs[0][0] - outs[1][0]) / (2 * h)
    assert abs(dA[0, 0, 0, 0] - fA[0, 0]) <= 1e-6 * abs(fA[0, 0])
