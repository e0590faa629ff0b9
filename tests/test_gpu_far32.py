"""Far-field panel pairs in FP32 (mh_set_far_field_fp32): the Rankine and
image terms of pairs beyond 7 panel radii (src/green_functions/
rankine.jl:33-34, :74-77) evaluated in FP32 by the assembly.

Per entry: the difference to the FP64 assembly, relative to the magnitude
of the far-field terms of that pair (a_j / 4 pi (1/r + 1/r') for S and
a_j / 4 pi (1/r^2 + 1/r'^2) for D, r' the distance to the image point), on
the 2,048-panel hemisphere; entries of pairs with no far-field term are
unchanged bit for bit.  Per sweep: the mixed sweep of the FP32-far-field
matrices against the FP64 sweep, for A, B, F and the gradients."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def far_scales(mesh):
    C, A, R = mesh.centers, mesh.areas, mesh.radii
    d = C[:, None, :] - C[None, :, :]
    r = np.sqrt((d ** 2).sum(-1))
    di = d.copy()
    di[..., 2] = -C[:, None, 2] - C[None, :, 2]  # collocation point reflected across z = 0
    ri = np.sqrt((di ** 2).sum(-1))
    far = r > 7.0 * R[None, :]
    far_i = ri > 7.0 * R[None, :]
    inv = lambda x, m: np.where(m, 1.0 / np.where(m, x, 1.0), 0.0)
    sS = A[None, :] / (4 * np.pi) * (inv(r, far) + inv(ri, far_i))
    sD = A[None, :] / (4 * np.pi) * (inv(r, far) ** 2 + inv(ri, far_i) ** 2)
    return far | far_i, sS, sD


def test_far_field_entries_within_fp32_error(mh, gpu):
    body = mh.hemisphere(1.0, ntheta=64, nphi=64)
    m = body.mesh
    assert m.nfaces == 2048
    ks = np.array([0.1, 1.0, 3.0])
    gfs = (mh.Rankine(), mh.RankineReflected(), mh.GFWu())
    S64, D64 = (x.cpu().numpy() for x in mh.assemble_matrices_batched(gfs, m, ks))
    try:
        mh._lib.set_far_field_fp32(True)
        S32, D32 = (x.cpu().numpy() for x in mh.assemble_matrices_batched(gfs, m, ks))
    finally:
        mh._lib.set_far_field_fp32(False)
    anyfar, sS, sD = far_scales(m)
    # pairs with no far-field term: bit for bit
    near = ~anyfar
    assert np.array_equal(S32[:, near], S64[:, near]) and np.array_equal(D32[:, near], D64[:, near])
    eS = np.abs(S32 - S64)[:, anyfar] / sS[anyfar]
    eD = np.abs(D32 - D64)[:, anyfar] / sD[anyfar]
    print(f"\nFP32 far field, 2048 panels: {anyfar.mean():.3f} of the pairs; per-entry error relative to "
          f"the pair's far-field terms: S max {eS.max():.2e} (mean {eS.mean():.1e}), "
          f"D max {eD.max():.2e} (mean {eD.mean():.1e}); matrix-wise S {relmax(S32, S64):.1e}, "
          f"D {relmax(D32, D64):.1e}", flush=True)
    assert eS.max() < 1e-6 and eD.max() < 1e-6


@pytest.mark.parametrize("direct", [True, False])
def test_mixed_sweep_with_fp32_far_field(mh, gpu, direct):
    body = mh.hemisphere(1.0, ntheta=32, nphi=32)
    m = body.mesh
    om = np.linspace(0.3, 3.0, 8)
    dofs = [[0, 0, 1.0], [1.0, 0, 0]]
    T = mh.shape_tangents(m, [body.dvertices["radius"], None], nomega=len(om), omega_param=1)
    ref = mh.freq_sweep_grad(m, om, dofs, [0.0], T, direct=direct, precision="fp64")
    try:
        mh._lib.set_far_field_fp32(True)
        got = mh.freq_sweep_grad(m, om, dofs, [0.0], T, direct=direct, precision="mixed")
    finally:
        mh._lib.set_far_field_fp32(False)
    err = {k: relmax(g, r) for k, g, r in zip(("A", "B", "F", "dA", "dB", "dF"), got, ref)}
    print(f"\nmixed sweep with FP32 far field vs FP64 ({m.nfaces} panels, direct={direct}): " +
          str({k: f"{v:.1e}" for k, v in err.items()}), flush=True)
    assert max(err.values()) < 1e-6
    # and the switch is off again: the mixed sweep is back at the FP64 result
    back = mh.freq_sweep_grad(m, om, dofs, [0.0], T, direct=direct, precision="mixed")
    assert max(relmax(g, r) for g, r in zip(back, ref)) < 1e-10
