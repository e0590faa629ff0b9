"""CPU checks of the C-ABI boundary and the host-side logic (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    txt = open(os.path.join(ROOT, "include", "mhip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mh_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(mh):
    lib = mh._lib.lib()
    names = _declared_symbols()
    assert "mh_assemble" in names and "mh_freq_sweep" in names
    for name in names:
        assert hasattr(lib, name), f"{name} declared in include/mhip.h but not exported"
    assert lib.mh_version() == 100


def test_python_binding_covers_every_symbol(mh):
    names = set(_declared_symbols())
    assert names == set(mh._lib._SIGS), names ^ set(mh._lib._SIGS)


def test_error_reporting_without_gpu(mh):
    lib = mh._lib.lib()
    rc = lib.mh_synchronize(None)
    assert rc == 1
    assert b"null context" in lib.mh_last_error()
    with pytest.raises(ValueError):
        mh._lib.check(rc, "probe")


def test_gf_mask(mh):
    assert mh.gf_mask((mh.Rankine(), mh.RankineReflected(), mh.GFWu())) == 1 | 2 | 8
    assert mh.gf_mask(mh.ExactGuevelDelhommeau()) == 16
    with pytest.raises(ValueError):
        mh.gf_mask((mh.Rankine(), mh.Rankine()))
    with pytest.raises(ValueError):
        mh.gf_mask((mh.GFWu(), mh.ExactGuevelDelhommeau()))
    with pytest.raises(ValueError):
        mh.gf_mask(())


def test_mesh_assertions(mh):
    # src/meshes.jl:28-34
    b = mh.hemisphere(1.0, ntheta=4, nphi=4)
    m = b.mesh
    with pytest.raises(ValueError):
        mh.Mesh(m.vertices[:, :2], m.faces, m.centers, m.normals, m.areas, m.radii)
    with pytest.raises(ValueError):
        mh.Mesh(m.vertices, m.faces[:, :3], m.centers, m.normals, m.areas, m.radii)
    with pytest.raises(ValueError):
        mh.Mesh(m.vertices, m.faces, m.centers[:-1], m.normals, m.areas, m.radii)
    with pytest.raises(ValueError):
        mh.Mesh(m.vertices, m.faces + 100, m.centers, m.normals, m.areas, m.radii)


def test_synthetic_meshes(mh):
    for ntheta, nphi, nf in [(10, 10, 50), (40, 40, 800), (64, 64, 2048), (20, 20, 200)]:
        m = mh.hemisphere(1.0, ntheta=ntheta, nphi=nphi).mesh
        assert m.nfaces == nf
        assert np.all(m.centers[:, 2] < 0)
        # outward normals, unit length
        assert np.allclose(np.linalg.norm(m.normals, axis=1), 1.0)
        assert np.all(np.sum(m.normals * m.centers, axis=1) > 0)
        # area converges to the hemisphere's 2 pi R^2 from below
        assert 0.9 * 2 * np.pi < m.areas.sum() < 2 * np.pi
    c = mh.vertical_cylinder(5.0, 10.0, ntheta=80, nz=30, nr=20).mesh
    assert c.nfaces == 80 * 30 + 80 * 20
    assert abs(c.areas.sum() - (2 * np.pi * 5 * 10 + np.pi * 25)) < 0.01 * (2 * np.pi * 50 + np.pi * 25)
    bx = mh.box(2.0, 1.0, 0.5, nx=4, ny=2, nz=2).mesh
    assert abs(bx.areas.sum() - (2 * 1 + 2 * (2 * 0.5) + 2 * (1 * 0.5))) < 1e-12
    nrm_out = np.sum(bx.normals * (bx.centers - np.array([0, 0, -0.25])), axis=1)
    assert np.all(nrm_out > 0)


def test_face_geometry_jvp_matches_finite_difference(mh):
    b = mh.vertical_cylinder(1.3, 2.0, ntheta=8, nz=3, nr=2)
    m = b.mesh
    for name, dV in b.dvertices.items():
        h = 1e-6
        cp, npl, ap, _ = mh.face_geometry(m.vertices + h * dV, m.faces)
        cm, nm, am, _ = mh.face_geometry(m.vertices - h * dV, m.faces)
        dc, dn, da = mh.face_geometry_jvp(m.vertices, m.faces, dV)
        assert np.allclose(dc, (cp - cm) / (2 * h), atol=1e-7), name
        assert np.allclose(dn, (npl - nm) / (2 * h), atol=1e-7), name
        assert np.allclose(da, (ap - am) / (2 * h), atol=1e-7), name


def test_rigid_dofs(mh):
    m = mh.hemisphere(1.0, ntheta=6, nphi=6).mesh
    d = mh.rigid_dofs(m, ["Heave", "Pitch"], rotation_center=(1.0, 1.0, 0.0))
    assert np.all(d[0] == [0, 0, 1])
    pos = m.centers - np.array([1.0, 1.0, 0.0])
    assert np.allclose(d[1], np.cross([0, 1, 0], pos))


def test_green_function_tuples_split_into_library_masks(mh):
    """gf_layers: any tuple as masks that one library call evaluates (each
    term at most once, at most one wave term), summing to the tuple."""
    R, RR, W, D = mh.Rankine(), mh.RankineReflected(), mh.GFWu(), mh.ExactGuevelDelhommeau()
    assert mh.api.gf_layers(W) == [8]
    assert mh.api.gf_layers((R, RR, W)) == [1 | 2 | 8]
    assert mh.api.gf_layers((R, R)) == [1, 1]
    layers = mh.api.gf_layers((R, RR, W, D))
    assert sorted(layers) == sorted([1 | 2 | 8, 16]) or sorted(layers) == sorted([1 | 2 | 16, 8])
    assert sum(bin(x).count("1") for x in mh.api.gf_layers((R, R, RR, W, W, D))) == 6
    for x in mh.api.gf_layers((R, R, RR, W, W, D)):
        assert bin(x & (8 | 16 | 32)).count("1") <= 1
    assert mh.api.gf_mask((R, RR, W)) == 11
    with pytest.raises(ValueError, match="at most once"):
        mh.api.gf_mask((R, R))
    with pytest.raises(ValueError, match="empty"):
        mh.api.gf_layers(())
    with pytest.raises(TypeError):
        mh.api.gf_layers((R, 3))
