"""Reverse mode of the wave-term pair sums (grad_rev.hip, mh_set_gradient_mode)
against the forward tangents of the same sweep: they compute the same
derivatives of A, B and F (test/outputs_differentiation.jl's quantities) in
another summation order, so they agree to rounding.  Cases: radius and
frequency on a hemisphere (direct and indirect, FP64 and mixed); 16 radial
shape modes (the case the automatic choice sends to reverse mode), also
against central differences of FP64 sweeps; six rigid dofs and two headings
(48 outputs, 24 reverse passes); the exact Delhommeau term; a box, whose
walls hold coplanar panel pairs, with its three dimensions as parameters."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def both_modes(mh, fn):
    try:
        mh._lib.set_gradient_mode("forward")
        fwd = fn()
        mh._lib.set_gradient_mode("reverse")
        rev = fn()
    finally:
        mh._lib.set_gradient_mode("auto")
    return fwd, rev


def shape_modes(U, count):
    x, y, z = U[:, 0], U[:, 1], U[:, 2]
    f = [np.ones_like(x), x, y, z, x * x, y * y, z * z, x * y, x * z, y * z, x ** 3, y ** 3, z ** 3,
         x * y * z, x * x * z, y * y * z]
    return [U * fm[:, None] for fm in f[:count]]


@pytest.mark.parametrize("direct", [True, False])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_reverse_equals_forward_radius_and_omega(mh, gpu, direct, precision):
    body = mh.hemisphere(1.0, ntheta=20, nphi=20)
    m = body.mesh
    om = [0.4, 1.1, 2.3]
    dofs = [[0, 0, 1.0], [1.0, 0, 0]]
    T = mh.shape_tangents(m, [body.dvertices["radius"], None], nomega=len(om), omega_param=1)
    fwd, rev = both_modes(mh, lambda: mh.freq_sweep_grad(m, om, dofs, [0.0, 0.9], T, direct=direct,
                                                         precision=precision))
    for name, f, r in zip(("A", "B", "F", "dA", "dB", "dF"), fwd, rev):
        assert relmax(r, f) < 1e-11, (name, relmax(r, f))


def test_reverse_sixteen_shape_modes_against_forward_and_differences(mh, gpu):
    body = mh.hemisphere(1.0, ntheta=16, nphi=16)
    m = body.mesh
    om = [0.6, 1.7]
    dofs = [[0, 0, 1.0]]
    modes = shape_modes(body.dvertices["radius"], 16)
    T = mh.shape_tangents(m, modes)
    fwd, rev = both_modes(mh, lambda: mh.freq_sweep_grad(m, om, dofs, [0.0], T))
    for name, f, r in zip(("dA", "dB", "dF"), fwd[3:], rev[3:]):
        assert relmax(r, f) < 1e-11, (name, relmax(r, f))
    auto = mh.freq_sweep_grad(m, om, dofs, [0.0], T)  # the automatic choice is reverse here
    for f, a in zip(fwd[3:], auto[3:]):
        assert relmax(a, f) < 1e-11
    # central differences of FP64 sweeps for modes whose derivative the
    # body's symmetry does not cancel (1, z, z^2, z^3), errors relative to the
    # largest derivative of each quantity
    h = 1e-6
    V0, F = m.vertices, m.faces
    qs = (0, 3, 6, 12)
    fds = {c: [] for c in range(3)}
    for q in qs:
        p = mh.freq_sweep(mh.mesh_from_vertices(V0 + h * modes[q], F), om, dofs, [0.0])
        n_ = mh.freq_sweep(mh.mesh_from_vertices(V0 - h * modes[q], F), om, dofs, [0.0])
        for c in range(3):
            fds[c].append((p[c] - n_[c]) / (2 * h))
    for c, name in enumerate(("dA", "dB", "dF")):
        fd = np.stack(fds[c], axis=1)
        got = rev[3 + c][:, list(qs)]
        assert relmax(got, fd) < 1e-6, (name, relmax(got, fd))


@pytest.mark.parametrize("direct", [True, False])
def test_reverse_many_outputs(mh, gpu, direct):
    body = mh.hemisphere(1.0, ntheta=10, nphi=12)
    m = body.mesh
    dofs = mh.rigid_dofs(m, ["Surge", "Sway", "Heave", "Roll", "Pitch", "Yaw"])
    om = [0.8, 1.9]
    T = mh.shape_tangents(m, [body.dvertices["radius"], body.dvertices["x"], None], nomega=len(om),
                          omega_param=2)
    fwd, rev = both_modes(mh, lambda: mh.freq_sweep_grad(m, om, dofs, [0.0, 1.2], T, direct=direct))
    for name, f, r in zip(("dA", "dB", "dF"), fwd[3:], rev[3:]):
        assert relmax(r, f) < 1e-11, (name, relmax(r, f))


def test_reverse_exact_delhommeau(mh, gpu):
    body = mh.hemisphere(1.0, ntheta=8, nphi=10)
    m = body.mesh
    gfs = (mh.Rankine(), mh.RankineReflected(), mh.ExactGuevelDelhommeau())
    om = [0.7, 1.5]
    T = mh.shape_tangents(m, [body.dvertices["radius"], None], nomega=len(om), omega_param=1)
    fwd, rev = both_modes(mh, lambda: mh.freq_sweep_grad(m, om, [[0, 0, 1.0]], [0.0], T, green_functions=gfs))
    for name, f, r in zip(("dA", "dB", "dF"), fwd[3:], rev[3:]):
        assert relmax(r, f) < 1e-11, (name, relmax(r, f))


@pytest.mark.parametrize("direct", [True, False])
def test_reverse_box_dimensions(mh, gpu, direct):
    body = mh.box(2.0, 1.2, 0.8, nx=6, ny=4, nz=3)
    m = body.mesh
    om = [0.5, 1.4, 2.6]
    T = mh.shape_tangents(m, [body.dvertices[k] for k in ("length", "width", "draft")])
    dofs = [[0, 0, 1.0], [1.0, 0, 0]]
    fwd, rev = both_modes(mh, lambda: mh.freq_sweep_grad(m, om, dofs, [0.0], T, direct=direct,
                                                         precision="mixed"))
    for name, f, r in zip(("dA", "dB", "dF"), fwd[3:], rev[3:]):
        assert relmax(r, f) < 1e-11, (name, relmax(r, f))
