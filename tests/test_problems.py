"""CPU checks of the problem / result bookkeeping (no GPU calls): the
layout and naming rules of src/problems_and_results.jl and src/bodies.jl."""
import numpy as np
import pytest


def _body(mh, name="b", center=(0.0, 0.0, 0.0)):
    m = mh.hemisphere(1.0, center=center, ntheta=6, nphi=6).mesh
    return mh.FloatingBody(m, ["Surge", "Heave", "Pitch"], name)


def test_problems_from_data_order_and_kinds(mh):
    body = _body(mh)
    params = {"wave_frequencies": [0.5, 1.0, 1.5], "wave_directions": [0.0, 0.3],
              "radiating_dofs": ["Heave", "Surge"]}
    probs = mh.problems_from_data(params, body)
    assert len(probs) == 3 * 2 + 3 * 2
    d = probs[:6]
    assert all(isinstance(p, mh.DiffractionProblem) for p in d)
    # headings vary fastest, then frequencies (vec of the Julia comprehension)
    assert [(p.omega, p.beta) for p in d[:3]] == [(0.5, 0.0), (0.5, 0.3), (1.0, 0.0)]
    r = probs[6:]
    assert [(p.omega, p.radiating_dof) for p in r[:3]] == [(0.5, "Heave"), (0.5, "Surge"), (1.0, "Heave")]
    assert r[0].influenced_dofs == ["Surge", "Heave", "Pitch"]  # all dofs by default
    with pytest.raises(AssertionError):
        mh.RadiationProblem(body, 1.0, "Yaw", ["Heave"])


def test_combine_floatingbodies_names_and_fields(mh):
    b1, b2 = _body(mh, "wec"), _body(mh, "wec", center=(4.0, 0.0, 0.0))
    c = b1 + b2
    assert c.body_name == "wec+wec_2"
    assert list(c.dofs)[:3] == ["wec__Surge", "wec__Heave", "wec__Pitch"]
    assert list(c.dofs)[3] == "wec_2__Surge"
    n1 = b1.mesh.nfaces
    assert np.array_equal(c.dofs["wec__Heave"][:n1], b1.dofs["Heave"])
    assert not np.any(c.dofs["wec__Heave"][n1:])
    assert np.array_equal(c.dofs["wec_2__Pitch"][n1:], b2.dofs["Pitch"])
    assert c.mesh.nfaces == 2 * n1


def test_rigid_dof_fields(mh):
    body = _body(mh)
    C = body.mesh.centers
    assert np.array_equal(body.dofs["Heave"], np.tile([0, 0, 1.0], (len(C), 1)))
    # Pitch = y-axis x (centre - rotation centre), src/bodies.jl:55-68
    assert np.allclose(body.dofs["Pitch"], np.cross([0, 1.0, 0], C))
    with pytest.raises(ValueError):
        mh.FloatingBody(body.mesh, {"bad": np.zeros((3, 3))}, "x")


def test_dimarray_sel(mh):
    a = mh.DimArray(np.arange(6.0).reshape(2, 3), ("a", "b"), {"a": ["x", "y"], "b": [1, 2, 3]})
    assert a.sel(a="y", b=2) == 4.0
    assert np.array_equal(a.sel(b=3), [2.0, 5.0])
