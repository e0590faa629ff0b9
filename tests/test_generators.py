"""CPU checks of the reference's geometry generators (src/meshes.jl:247-479):
axisymmetric_mesh, wavebot_profile, wavebot_mesh, and the face quantities of
Mesh(::SimpleMesh)."""
import numpy as np


def test_wavebot_topology_and_orientation(mh):
    b = mh.wavebot_mesh(n_points=(5, 10, 5), n_theta=72)
    m = b.mesh
    assert m.nfaces == (4 + 9 + 4) * 72 and m.nvertices == 17 * 72 + 1
    tri = m.faces[:, 2] == m.faces[:, 3]
    assert np.sum(tri) == 72  # the ring around the axis point only
    # outward normals: radially out on the walls, down on the bottom
    radial = np.sum(m.normals[:, :2] * m.centers[:, :2], axis=1)
    walls = m.normals[:, 2] > -0.99
    assert np.all(radial[walls] > 0) and np.all(m.normals[~walls, 2] < -0.99)
    # triangles: centre (a + b + c) / 3, radius = distance to the first vertex
    V = m.vertices
    f = m.faces[np.argmax(tri)]
    assert np.allclose(m.centers[np.argmax(tri)], V[f[:3]].mean(axis=0), atol=1e-15)
    assert np.isclose(m.radii[np.argmax(tri)], np.linalg.norm(V[f[0]] - m.centers[np.argmax(tri)]))


def test_wavebot_area_converges_to_surface_of_revolution(mh):
    r1, r2, d1, d2 = 0.88, 0.35, 0.16, 0.37
    exact = 2 * np.pi * r1 * d1 + np.pi * (r1 + r2) * np.hypot(r1 - r2, d2) + np.pi * r2 ** 2
    errs = [abs(mh.wavebot_mesh(r1, r2, d1, d2, n_theta=nt).mesh.areas.sum() - exact) / exact
            for nt in (18, 36, 72)]
    # inscribed polygons: the error falls as 1 / n_theta^2
    assert 3.5 < errs[0] / errs[1] < 4.5 and 3.5 < errs[1] / errs[2] < 4.5 and errs[2] < 1e-3


def test_wavebot_vertex_derivatives_match_finite_differences(mh):
    p0 = {"r1": 0.9, "r2": 0.4, "d1": 0.2, "d2": 0.35}
    b = mh.wavebot_mesh(**p0, n_points=(3, 4, 3), n_theta=12)
    h = 1e-6
    for k in p0:
        pp, pm = dict(p0), dict(p0)
        pp[k] += h
        pm[k] -= h
        fd = (mh.wavebot_mesh(**pp, n_points=(3, 4, 3), n_theta=12).mesh.vertices -
              mh.wavebot_mesh(**pm, n_points=(3, 4, 3), n_theta=12).mesh.vertices) / (2 * h)
        assert np.max(np.abs(b.dvertices[k] - fd)) < 1e-9, k


def test_reference_face_geometry_of_a_square(mh):
    V = np.array([[0, 0, 0], [0, 1.0, 0], [1.0, 1.0, 0], [1.0, 0, 0]])
    c, n, a, r = mh.reference_face_geometry(V, np.array([[0, 1, 2, 3]]))
    assert np.allclose(c, [[0.5, 0.5, 0]]) and np.allclose(a, [1.0])
    assert np.allclose(n, [[0, 0, -1.0]]) and np.isclose(r[0], np.sqrt(0.5))


def test_axisymmetric_mesh_rejects_bad_profiles(mh):
    import pytest

    with pytest.raises(ValueError):
        mh.axisymmetric_mesh([[1.0, 0.0]], 8)
    with pytest.raises(ValueError):
        mh.axisymmetric_mesh([[0.0, 0.0], [0.0, -1.0]], 8)
    with pytest.raises(ValueError):
        mh.axisymmetric_mesh([[1.0, 0.0], [0.5, -1.0]], 2)
