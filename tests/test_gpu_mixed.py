"""Mixed precision: the FP64 fallback of a chunk whose refinement does not
converge (forced from the host by allowing a single refinement sweep; every
kernel runs normally), two sweeps sufficing on ragged meshes (which pins
the assembly's FP32 copy of the matrix the FP32 LU factors), the deferred
convergence check (same results; a chunk that misses the tolerance solved
again), and non-finite input reported as an error rather than 'converged'."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def relmax(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(np.asarray(b))))


@pytest.mark.parametrize("direct", [True, False])
def test_fp64_fallback_when_refinement_cannot_converge(mh, gpu, direct):
    body = mh.hemisphere(1.0, ntheta=16, nphi=16)
    m = body.mesh
    om = [0.5, 1.2, 2.4]
    dofs = [[0, 0, 1.0], [1.0, 0, 0]]
    T = mh.shape_tangents(m, [body.dvertices["radius"], None], nomega=3, omega_param=1)
    ref = mh.freq_sweep_grad(m, om, dofs, [0.0], T, direct=direct, precision="fp64")
    try:
        mh._lib.set_refinement(max_sweeps=1, fp64_fallback=True)
        n0 = mh._lib.fallback_count()
        got = mh.freq_sweep_grad(m, om, dofs, [0.0], T, direct=direct, precision="mixed")
        assert mh._lib.fallback_count() == n0 + 1
        for g, r, name in zip(got, ref, ("A", "B", "F", "dA", "dB", "dF")):
            assert relmax(g, r) < 1e-12, name  # solved in FP64 again
        mh._lib.set_refinement(max_sweeps=1, fp64_fallback=False)
        with pytest.raises(ArithmeticError, match="did not reach"):
            mh.freq_sweep(m, om, dofs, [0.0], direct=direct, precision="mixed")
    finally:
        mh._lib.set_refinement(max_sweeps=12, fp64_fallback=True)
    # defaults restored: converges in mixed precision without a fallback
    n0 = mh._lib.fallback_count()
    got = mh.freq_sweep_grad(m, om, dofs, [0.0], T, direct=direct, precision="mixed")
    assert mh._lib.fallback_count() == n0
    for g, r in zip(got, ref):
        assert relmax(g, r) < 1e-8


@pytest.mark.parametrize("shape", [(18, 25), (30, 41)])
@pytest.mark.parametrize("direct", [True, False])
def test_two_sweeps_suffice_on_ragged_meshes(mh, gpu, shape, direct):
    # The FP32 copy of D (K) that the mixed assembly writes, both halves of
    # every 64 x 64 tile pair and the ragged last tiles and row groups
    # included (225 and 615 panels), is the matrix the FP32 LU factors; a
    # wrong entry in it would leave the refinement short of its 1e-10
    # tolerance after two sweeps, which (without the FP64 fallback) is an error.
    body = mh.hemisphere(1.0, ntheta=shape[0], nphi=shape[1])
    m = body.mesh
    om = np.linspace(0.3, 3.0, 11)
    dofs = [[0, 0, 1.0], [1.0, 0, 0]]
    T = mh.shape_tangents(m, [body.dvertices["radius"], None], nomega=len(om), omega_param=1)
    ref = mh.freq_sweep_grad(m, om, dofs, [0.0, 0.7], T, direct=direct, precision="fp64")
    try:
        mh._lib.set_refinement(max_sweeps=2, fp64_fallback=False)
        got = mh.freq_sweep_grad(m, om, dofs, [0.0, 0.7], T, direct=direct, precision="mixed")
    finally:
        mh._lib.set_refinement(max_sweeps=12, fp64_fallback=True)
    for g, r, name in zip(got, ref, ("A", "B", "F", "dA", "dB", "dF")):
        assert relmax(g, r) < 1e-9, name


@pytest.mark.parametrize("direct", [True, False])
def test_deferred_check_gives_the_checked_results(mh, gpu, direct):
    # the default deferral makes the same two sweeps as the check after every
    # sweep (which converges after the second), only its last residual pass
    # runs beside the gradient kernels: the same results
    body = mh.hemisphere(1.0, ntheta=20, nphi=22)
    m = body.mesh
    om = np.linspace(0.4, 2.8, 9)
    dofs = [[0, 0, 1.0], [1.0, 0, 0]]
    T = mh.shape_tangents(m, [body.dvertices["radius"], None], nomega=len(om), omega_param=1)
    try:
        mh._lib.set_deferred_check(0)
        ref = mh.freq_sweep_grad(m, om, dofs, [0.0, 0.5], T, direct=direct, precision="mixed")
    finally:
        mh._lib.set_deferred_check(2)
    n0 = mh._lib.deferred_rerun_count()
    got = mh.freq_sweep_grad(m, om, dofs, [0.0, 0.5], T, direct=direct, precision="mixed")
    assert mh._lib.deferred_rerun_count() == n0
    for g, r, name in zip(got, ref, ("A", "B", "F", "dA", "dB", "dF")):
        assert relmax(g, r) < 1e-14, name


@pytest.mark.parametrize("direct", [True, False])
def test_deferred_check_reruns_a_chunk_that_misses_the_tolerance(mh, gpu, direct):
    # one deferred sweep never meets 1e-10: the device verdict fails and the
    # chunk is solved again with a check after every sweep; with a single
    # sweep allowed there as well, the FP64 fallback then solves it
    body = mh.hemisphere(1.0, ntheta=16, nphi=18)
    m = body.mesh
    om = [0.5, 1.3, 2.2]
    dofs = [[0, 0, 1.0], [1.0, 0, 0]]
    T = mh.shape_tangents(m, [body.dvertices["radius"], None], nomega=len(om), omega_param=1)
    ref_mixed = mh.freq_sweep_grad(m, om, dofs, [0.0], T, direct=direct, precision="mixed")
    ref64 = mh.freq_sweep_grad(m, om, dofs, [0.0], T, direct=direct, precision="fp64")
    try:
        mh._lib.set_deferred_check(1)
        n0, f0 = mh._lib.deferred_rerun_count(), mh._lib.fallback_count()
        got = mh.freq_sweep_grad(m, om, dofs, [0.0], T, direct=direct, precision="mixed")
        assert mh._lib.deferred_rerun_count() == n0 + 1 and mh._lib.fallback_count() == f0
        for g, r, name in zip(got, ref_mixed, ("A", "B", "F", "dA", "dB", "dF")):
            assert relmax(g, r) < 1e-14, name
        mh._lib.set_refinement(max_sweeps=1, fp64_fallback=True)
        got = mh.freq_sweep_grad(m, om, dofs, [0.0], T, direct=direct, precision="mixed")
        assert mh._lib.deferred_rerun_count() == n0 + 2 and mh._lib.fallback_count() == f0 + 1
        for g, r, name in zip(got, ref64, ("A", "B", "F", "dA", "dB", "dF")):
            assert relmax(g, r) < 1e-12, name
    finally:
        mh._lib.set_deferred_check(2)
        mh._lib.set_refinement(max_sweeps=12, fp64_fallback=True)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_sweeps_in_several_frequency_chunks(mh, gpu, precision):
    # 8 frequencies in chunks of at most 3 (3 + 3 + 2): each chunk its own
    # assembly, factorisation, solves and gradients, at its offset in the
    # frequency list; the results are those of the single chunk.  Mixed:
    # with one deferred sweep every chunk's verdict fails and the chunk is
    # solved again with the check after every sweep.
    body = mh.hemisphere(1.0, ntheta=14, nphi=16)
    m = body.mesh
    om = np.linspace(0.3, 2.9, 8)
    dofs = [[0, 0, 1.0], [1.0, 0, 0]]
    T = mh.shape_tangents(m, [body.dvertices["radius"], None], nomega=len(om), omega_param=1)
    ref = mh.freq_sweep_grad(m, om, dofs, [0.0, 0.9], T, precision=precision)
    try:
        mh._lib.set_sweep_chunk(3)
        got = mh.freq_sweep_grad(m, om, dofs, [0.0, 0.9], T, precision=precision)
        for g, r, name in zip(got, ref, ("A", "B", "F", "dA", "dB", "dF")):
            assert relmax(g, r) < 1e-13, name
        if precision == "mixed":
            mh._lib.set_deferred_check(1)
            n0 = mh._lib.deferred_rerun_count()
            got = mh.freq_sweep_grad(m, om, dofs, [0.0, 0.9], T, precision=precision)
            assert mh._lib.deferred_rerun_count() == n0 + 3
            for g, r, name in zip(got, ref, ("A", "B", "F", "dA", "dB", "dF")):
                assert relmax(g, r) < 1e-13, name
    finally:
        mh._lib.set_sweep_chunk(0)
        mh._lib.set_deferred_check(2)


def test_nan_in_the_system_is_an_error_not_converged(mh, gpu):
    m = mh.hemisphere(1.0, ntheta=10, nphi=10).mesh
    dof = np.tile([0.0, 0.0, 1.0], (m.nfaces, 1))
    dof[3, 2] = np.nan  # a NaN right-hand-side entry
    n0 = mh._lib.fallback_count()
    with pytest.raises(ArithmeticError, match="non-finite"):
        mh.freq_sweep(m, [1.0], [dof], precision="mixed")
    assert mh._lib.fallback_count() == n0  # not retried in FP64
    # the context keeps working
    A, B, F = mh.freq_sweep(m, [1.0], [[0, 0, 1.0]], precision="mixed")
    assert np.all(np.isfinite(A)) and np.all(np.isfinite(B))


@pytest.mark.parametrize("direct", [True, False])
def test_fp32_factors_converge_in_two_sweeps(mh, gpu, direct):
    """The FP32 factorisation (its trailing updates on the half-precision
    matrix cores through split, per-row/column scaled operands) is accurate
    to FP32 level: one FP32 solve and one correction reach the 1e-10
    residual tolerance.  A factorisation with fp16-level errors (a dropped
    split term, a wrong scale) contracts the error by ~1e-3 per sweep and
    fails here, since the FP64 fallback is off.  1,058 panels: edge tiles
    in both directions (1058 = 8 x 128 + 34), narrow and wide updates."""
    m = mh.hemisphere(1.0, ntheta=46, nphi=46).mesh
    assert m.nfaces == 1058
    om = [0.4, 1.5, 3.0]
    dofs = [[0, 0, 1.0], [1.0, 0, 0]]
    ref = mh.freq_sweep(m, om, dofs, [0.0], direct=direct, precision="fp64")
    try:
        mh._lib.set_refinement(max_sweeps=2, fp64_fallback=False)
        got = mh.freq_sweep(m, om, dofs, [0.0], direct=direct, precision="mixed")
    finally:
        mh._lib.set_refinement(max_sweeps=12, fp64_fallback=True)
    for g, r, name in zip(got, ref, ("A", "B", "F")):
        assert relmax(g, r) < 1e-8, name



@pytest.mark.parametrize("direct", [True, False])
def test_mixed_solve_of_systems_far_outside_the_fp16_range(mh, gpu, direct):
    """mh_solve_mixed on D (K) and S scaled together by 2^40 and by 2^-40
    (phi = D^-1 S bc and S K^-1 bc are invariant): the FP32 factorisation
    scales each row of L21 and column of U12 by its own power of two before
    the fp16 split, so every entry is in range, the refinement converges
    without falling back to FP64 and the solution is the unscaled FP64 one.
    512 panels: the factorisation's narrow and wide trailing updates both
    run on the split-image kernel."""
    import torch

    m = mh.hemisphere(1.0, ntheta=32, nphi=32).mesh
    assert m.nfaces == 512
    S, DK = mh.assemble_matrices((mh.Rankine(), mh.RankineReflected(), mh.GFWu()), m, 0.4, direct=direct)
    rng = np.random.default_rng(11)
    bc = torch.from_numpy(rng.normal(size=(512, 2)) + 1j * rng.normal(size=(512, 2))).to(DK.device)
    ref = mh.solve(DK, S, bc, direct=direct).cpu().numpy()
    n0 = mh._lib.fallback_count()
    for s in (2.0 ** 40, 2.0 ** -40, 1.0):
        got = mh.solve(DK * s, S * s, bc, direct=direct, precision="mixed").cpu().numpy()
        assert relmax(got, ref) < 1e-9, s
    assert mh._lib.fallback_count() == n0


def test_solve_leaves_the_assembled_matrices_intact(mh, gpu):
    """solve() on the assembly's own output (column-major memory): D and S
    are not modified, so a second solve gives the same phi (the FP64 path
    factors a copy in place)."""
    import torch

    m = mh.hemisphere(1.0, ntheta=12, nphi=12).mesh
    S, D = mh.assemble_matrices((mh.Rankine(), mh.RankineReflected(), mh.GFWu()), m, 0.4)
    D0, S0 = D.clone(), S.clone()
    bc = torch.ones(m.nfaces, dtype=torch.complex128, device=D.device)
    p1 = mh.solve(D, S, bc).cpu().numpy()
    p2 = mh.solve(D, S, bc).cpu().numpy()
    assert torch.equal(D, D0) and torch.equal(S, S0)
    assert np.array_equal(p1, p2)
