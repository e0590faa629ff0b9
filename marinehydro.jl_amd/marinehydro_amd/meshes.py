"""Panel meshes: the ``Mesh`` container and synthetic body generators.

``Mesh`` mirrors the reference struct (reference src/meshes.jl:12-37): the
same eight fields, the same shape checks, faces 0-based (as Capytaine writes
them and as reference src/meshes.jl:179 reads them).  Arrays are kept in
column-major (Fortran) order so they can be handed to the C-ABI without a
copy, exactly as a Julia caller would.

The generators replace the Capytaine mesh builders the reference's tests use
(reference src/meshes.jl:120-127, out of scope here because Capytaine is a
Python dependency of the reference): latitude-longitude spheres, vertical
cylinders and boxes, clipped to the immersed part z <= 0.  Face centres are
area-weighted, normals point out of the body, radii are the largest
centre-to-vertex distance (Capytaine's ``faces_radiuses``).
``face_geometry_jvp`` differentiates the face quantities w.r.t. the vertices;
body generators return d(vertices)/d(parameter) so shape gradients can be
chained without finite differences.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


class Mesh:
    """reference src/meshes.jl:12-37 (same fields and assertions)."""

    def __init__(self, vertices, faces, centers, normals, areas, radii, nvertices=None,
                 nfaces=None):
        vertices = np.asfortranarray(vertices, dtype=np.float64)
        faces = np.asfortranarray(faces, dtype=np.int64)
        centers = np.asfortranarray(centers, dtype=np.float64)
        normals = np.asfortranarray(normals, dtype=np.float64)
        areas = np.ascontiguousarray(areas, dtype=np.float64).reshape(-1)
        radii = np.ascontiguousarray(radii, dtype=np.float64).reshape(-1)
        nv = vertices.shape[0]
        nf = faces.shape[0]
        if vertices.ndim != 2 or vertices.shape != (nv, 3):
            raise ValueError("each vertex needs 3 coordinates")
        if faces.ndim != 2 or faces.shape != (nf, 4):
            raise ValueError("only quadrilateral panels are allowed")
        if centers.shape != (nf, 3):
            raise ValueError("centers must be [nfaces x 3]")
        if normals.shape != (nf, 3):
            raise ValueError("normals must be [nfaces x 3]")
        if areas.shape[0] != nf:
            raise ValueError("areas vector must have length nfaces")
        if radii.shape[0] != nf:
            raise ValueError("radii vector must have length nfaces")
        if nf and (faces.min() < 0 or faces.max() >= nv):
            raise ValueError("face vertex index out of range (faces are 0-based)")
        self.vertices, self.faces, self.centers, self.normals = vertices, faces, centers, normals
        self.areas, self.radii = areas, radii
        self.nvertices, self.nfaces = nv, nf

    def __repr__(self):
        return f"Mesh(nfaces={self.nfaces}, nvertices={self.nvertices})"


def combine_meshes(meshes):
    """reference src/meshes.jl:53-99 (face indices shifted by the vertex
    count of the meshes before)."""
    V, F, Cc, N, A, R = [], [], [], [], [], []
    nbv = 0
    for m in meshes:
        V.append(m.vertices)
        F.append(m.faces + nbv)
        Cc.append(m.centers)
        N.append(m.normals)
        A.append(m.areas)
        R.append(m.radii)
        nbv += m.nvertices
    return Mesh(np.vstack(V), np.vstack(F), np.vstack(Cc), np.vstack(N), np.concatenate(A),
                np.concatenate(R))


# --------------------------------------------------------------------------
# face geometry from vertices
# --------------------------------------------------------------------------
def _tri(a, b, c):
    cr = np.cross(b - a, c - a)
    return 0.5 * np.linalg.norm(cr, axis=1), cr


def face_geometry(vertices, faces):
    """centers, normals, areas, radii of quadrilateral faces.

    Two triangles (v0, v1, v2) and (v0, v2, v3); area-weighted centroid;
    normal = normalised cross product of the diagonals, which is the exact
    normal of a planar quad and stays defined for a degenerate one.
    """
    V = np.asarray(vertices, dtype=np.float64)
    Fc = np.asarray(faces)
    a, b, c, d = (V[Fc[:, i]] for i in range(4))
    a1, _ = _tri(a, b, c)
    a2, _ = _tri(a, c, d)
    area = a1 + a2
    c1 = (a + b + c) / 3.0
    c2 = (a + c + d) / 3.0
    center = (c1 * a1[:, None] + c2 * a2[:, None]) / area[:, None]
    nrm = np.cross(c - a, d - b)
    normal = nrm / np.linalg.norm(nrm, axis=1)[:, None]
    radius = np.max(np.stack([np.linalg.norm(v - center, axis=1) for v in (a, b, c, d)]), axis=0)
    return center, normal, area, radius


def face_geometry_jvp(vertices, faces, dvertices):
    """Forward-mode derivative of ``face_geometry`` along dvertices.

    Returns (dcenters, dnormals, dareas).  Radii only switch the near/far
    branch of the Rankine panel integral and carry no derivative.
    """
    V = np.asarray(vertices, dtype=np.float64)
    dV = np.asarray(dvertices, dtype=np.float64)
    Fc = np.asarray(faces)
    a, b, c, d = (V[Fc[:, i]] for i in range(4))
    da, db, dc, dd = (dV[Fc[:, i]] for i in range(4))

    def tri(p, q, r, dp, dq, dr):
        cr = np.cross(q - p, r - p)
        dcr = np.cross(dq - dp, r - p) + np.cross(q - p, dr - dp)
        nn = np.linalg.norm(cr, axis=1)
        safe = np.where(nn > 0, nn, 1.0)
        dn = np.where(nn > 0, np.sum(cr * dcr, axis=1) / safe, 0.0)
        return 0.5 * nn, 0.5 * dn

    a1, da1 = tri(a, b, c, da, db, dc)
    a2, da2 = tri(a, c, d, da, dc, dd)
    area = a1 + a2
    darea = da1 + da2
    c1 = (a + b + c) / 3.0
    c2 = (a + c + d) / 3.0
    dc1 = (da + db + dc) / 3.0
    dc2 = (da + dc + dd) / 3.0
    num = c1 * a1[:, None] + c2 * a2[:, None]
    dnum = dc1 * a1[:, None] + c1 * da1[:, None] + dc2 * a2[:, None] + c2 * da2[:, None]
    dcenter = dnum / area[:, None] - num * (darea / area ** 2)[:, None]
    nrm = np.cross(c - a, d - b)
    dnrm = np.cross(dc - da, d - b) + np.cross(c - a, dd - db)
    nn = np.linalg.norm(nrm, axis=1)
    dnormal = dnrm / nn[:, None] - nrm * (np.sum(nrm * dnrm, axis=1) / nn ** 3)[:, None]
    return dcenter, dnormal, darea


def mesh_from_vertices(vertices, faces):
    c, n, a, r = face_geometry(vertices, faces)
    return Mesh(vertices, faces, c, n, a, r)


# --------------------------------------------------------------------------
# body generators (immersed parts), each with d(vertices)/d(parameter)
# --------------------------------------------------------------------------
@dataclass
class Body:
    mesh: Mesh
    params: dict = field(default_factory=dict)
    # parameter name -> d(vertices)/d(parameter), shape (nv, 3)
    dvertices: dict = field(default_factory=dict)


def _orient_outward(V, F, inside_point_fn):
    """Reverse the vertex order of faces whose normal points into the body."""
    c, n, _, _ = face_geometry(V, F)
    inward = np.sum(n * (c - inside_point_fn(c)), axis=1) < 0
    F = F.copy()
    F[inward] = F[inward][:, ::-1]
    return F


def _latlong_sphere(radius, center, ntheta, nphi, a_first):
    R = float(radius)
    c0 = np.asarray(center, dtype=np.float64)
    a_idx = np.arange(a_first, ntheta + 1)
    theta = np.pi * a_idx / ntheta
    phi = 2 * np.pi * np.arange(nphi) / nphi
    T, P = np.meshgrid(theta, phi, indexing="ij")
    U = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], axis=-1)
    U[np.abs(U) < 1e-15] = 0.0
    U = U.reshape(-1, 3)
    V = c0 + R * U
    nr = len(a_idx)
    F = []
    for a in range(nr - 1):
        for b in range(nphi):
            b2 = (b + 1) % nphi
            F.append([a * nphi + b, (a + 1) * nphi + b, (a + 1) * nphi + b2, a * nphi + b2])
    F = np.array(F, dtype=np.int64)
    F = _orient_outward(V, F, lambda c: c0)
    mesh = mesh_from_vertices(V, F)
    dV = {"radius": U.copy(), "x": np.tile([1.0, 0.0, 0.0], (len(V), 1)),
          "y": np.tile([0.0, 1.0, 0.0], (len(V), 1))}
    return Body(mesh, {"radius": R, "x": c0[0], "y": c0[1]}, dV)


def hemisphere(radius=1.0, center=(0.0, 0.0, 0.0), ntheta=10, nphi=10) -> Body:
    """Immersed half of a latitude-longitude sphere.

    Vertex rings at theta_a = pi a / ntheta (a = ntheta/2 .. ntheta, equator
    to the bottom pole) and phi_b = 2 pi b / nphi; (ntheta/2) * nphi quad
    panels, the bottom ring degenerate (a repeated vertex), as for the
    immersed part of Capytaine's mesh_sphere(resolution=(ntheta, nphi)).
    """
    if ntheta % 2:
        raise ValueError("ntheta must be even so the waterline is a vertex ring")
    return _latlong_sphere(radius, center, ntheta, nphi, ntheta // 2)


def floating_sphere(radius=1.0, draft=1.0, center=(0.0, 0.0), ntheta=10, nphi=10) -> Body:
    """Immersed cap of a sphere of the given radius floating at the given
    draft (depth of its lowest point), 0 < draft < 2 radius: centre at
    z = radius - draft, vertex rings from the waterline (theta_0 with
    cos theta_0 = draft / radius - 1) to the bottom pole, theta_a = theta_0 +
    (pi - theta_0) a / ntheta, a = 0 .. ntheta, and phi_b = 2 pi b / nphi;
    ntheta * nphi quads (the bottom ring degenerate).  d vertices / d radius
    and / d draft are exact."""
    R, d = float(radius), float(draft)
    if not 0.0 < d < 2.0 * R:
        raise ValueError("draft must lie in (0, 2 radius)")
    cx, cy = float(center[0]), float(center[1])
    c0 = d / R - 1.0
    th0 = np.arccos(c0)
    s0 = np.sqrt(1.0 - c0 * c0)
    dth0_dR = -(-d / R ** 2) / s0  # d arccos(c0) = -dc0 / sqrt(1 - c0^2)
    dth0_dd = -(1.0 / R) / s0
    a = np.arange(ntheta + 1)
    w = 1.0 - a / ntheta
    th = th0 * w + np.pi * (a / ntheta)
    phi = 2 * np.pi * np.arange(nphi) / nphi
    T, P = np.meshgrid(th, phi, indexing="ij")
    W, _ = np.meshgrid(w, phi, indexing="ij")
    st, ct, cp, sp = np.sin(T), np.cos(T), np.cos(P), np.sin(P)
    zc = R - d
    V = np.stack([cx + R * st * cp, cy + R * st * sp, zc + R * ct], axis=-1)
    V[0, :, 2] = 0.0  # the waterline ring exactly at z = 0
    dT_dR, dT_dd = dth0_dR * W, dth0_dd * W
    dVR = np.stack([st * cp + R * ct * cp * dT_dR, st * sp + R * ct * sp * dT_dR,
                    1.0 + ct - R * st * dT_dR], axis=-1)
    dVd = np.stack([R * ct * cp * dT_dd, R * ct * sp * dT_dd, -1.0 - R * st * dT_dd], axis=-1)
    dVR[0, :, 2] = 0.0
    dVd[0, :, 2] = 0.0
    V, dVR, dVd = V.reshape(-1, 3), dVR.reshape(-1, 3), dVd.reshape(-1, 3)
    F = []
    for i in range(ntheta):
        for b in range(nphi):
            b2 = (b + 1) % nphi
            F.append([i * nphi + b, (i + 1) * nphi + b, (i + 1) * nphi + b2, i * nphi + b2])
    F = np.array(F, dtype=np.int64)
    ctr = np.array([cx, cy, zc])
    F = _orient_outward(V, F, lambda c: np.tile(ctr, (len(c), 1)))
    mesh = mesh_from_vertices(V, F)
    return Body(mesh, {"radius": R, "draft": d}, {"radius": dVR, "draft": dVd})


def sphere(radius=1.0, center=(0.0, 0.0, 0.0), ntheta=10, nphi=10) -> Body:
    """Full latitude-longitude sphere (both poles degenerate rings); used for
    closed-surface identities, not a floating body."""
    return _latlong_sphere(radius, center, ntheta, nphi, 0)


def vertical_cylinder(radius=1.0, draft=1.0, center=(0.0, 0.0), ntheta=16, nz=4, nr=4) -> Body:
    """Truncated vertical cylinder: side wall (ntheta x nz quads) and bottom
    disk (nr rings x ntheta, the innermost ring degenerate)."""
    R, H = float(radius), float(draft)
    cx, cy = float(center[0]), float(center[1])
    phi = 2 * np.pi * np.arange(ntheta) / ntheta
    ring = np.stack([np.cos(phi), np.sin(phi)], axis=1)
    verts, dR, dH, F = [], [], [], []
    # side wall vertices: rows z_l = -H l / nz, l = 0..nz
    for l in range(nz + 1):
        s = l / nz
        for q in range(ntheta):
            verts.append([cx + R * ring[q, 0], cy + R * ring[q, 1], -H * s])
            dR.append([ring[q, 0], ring[q, 1], 0.0])
            dH.append([0.0, 0.0, -s])
    for l in range(nz):
        for q in range(ntheta):
            q2 = (q + 1) % ntheta
            F.append([l * ntheta + q, (l + 1) * ntheta + q, (l + 1) * ntheta + q2, l * ntheta + q2])
    # bottom disk rings rho_m = R (nr - m)/nr, m = 0..nr (m = nr: centre)
    base = len(verts)
    for m in range(nr + 1):
        f = (nr - m) / nr
        for q in range(ntheta):
            verts.append([cx + R * f * ring[q, 0], cy + R * f * ring[q, 1], -H])
            dR.append([f * ring[q, 0], f * ring[q, 1], 0.0])
            dH.append([0.0, 0.0, -1.0])
    for m in range(nr):
        for q in range(ntheta):
            q2 = (q + 1) % ntheta
            F.append([base + m * ntheta + q, base + (m + 1) * ntheta + q,
                      base + (m + 1) * ntheta + q2, base + m * ntheta + q2])
    V = np.array(verts)
    F = np.array(F, dtype=np.int64)
    ctr = np.array([cx, cy, -H / 2])

    def inside(c):
        p = np.tile(ctr, (len(c), 1))
        # for side panels the axis point at the same depth; for the bottom the centre above
        side = c[:, 2] > -H + 1e-9 * max(H, 1.0)
        p[side, 2] = c[side, 2]
        p[~side, 0], p[~side, 1] = c[~side, 0], c[~side, 1]
        return p

    F = _orient_outward(V, F, inside)
    mesh = mesh_from_vertices(V, F)
    return Body(mesh, {"radius": R, "draft": H}, {"radius": np.array(dR), "draft": np.array(dH)})


def box(length=2.0, width=1.0, draft=1.0, center=(0.0, 0.0), nx=4, ny=2, nz=2) -> Body:
    """Immersed rectangular box: four walls and the bottom, z in [-draft, 0]."""
    Lx, Ly, H = float(length), float(width), float(draft)
    cx, cy = float(center[0]), float(center[1])
    verts, dL, dW, dH, F = [], [], [], [], []

    def add_grid(P0, e1, e2, n1, n2, d0, de1, de2):
        """grid P0 + s e1 + t e2; d* = tangents (per parameter) of P0, e1, e2."""
        start = len(verts)
        for i in range(n1 + 1):
            for j in range(n2 + 1):
                s, t = i / n1, j / n2
                verts.append(P0 + s * e1 + t * e2)
                for lst, (a0, a1, a2) in zip((dL, dW, dH), zip(d0, de1, de2)):
                    lst.append(a0 + s * a1 + t * a2)
        for i in range(n1):
            for j in range(n2):
                q = start + i * (n2 + 1) + j
                F.append([q, q + n2 + 1, q + n2 + 2, q + 1])

    z3 = np.zeros(3)
    ex, ey, ez = np.eye(3)
    x0, y0 = cx - Lx / 2, cy - Ly / 2
    # tangents per parameter (length, width, draft) of P0, e1, e2
    # bottom z=-H: P0=(x0,y0,-H), e1=Lx ex, e2=Ly ey
    add_grid(np.array([x0, y0, -H]), Lx * ex, Ly * ey, nx, ny,
             (-0.5 * ex, -0.5 * ey, -ez), (ex, z3, z3), (z3, ey, z3))
    # wall y=y0: P0=(x0,y0,-H), e1=Lx ex, e2=H ez
    add_grid(np.array([x0, y0, -H]), Lx * ex, H * ez, nx, nz,
             (-0.5 * ex, -0.5 * ey, -ez), (ex, z3, z3), (z3, z3, ez))
    # wall y=y0+Ly
    add_grid(np.array([x0, y0 + Ly, -H]), Lx * ex, H * ez, nx, nz,
             (-0.5 * ex, 0.5 * ey, -ez), (ex, z3, z3), (z3, z3, ez))
    # wall x=x0: P0=(x0,y0,-H), e1=Ly ey, e2=H ez
    add_grid(np.array([x0, y0, -H]), Ly * ey, H * ez, ny, nz,
             (-0.5 * ex, -0.5 * ey, -ez), (z3, ey, z3), (z3, z3, ez))
    # wall x=x0+Lx
    add_grid(np.array([x0 + Lx, y0, -H]), Ly * ey, H * ez, ny, nz,
             (0.5 * ex, -0.5 * ey, -ez), (z3, ey, z3), (z3, z3, ez))
    V = np.array(verts)
    F = np.array(F, dtype=np.int64)
    ctr = np.array([cx, cy, -H / 2])
    F = _orient_outward(V, F, lambda c: np.tile(ctr, (len(c), 1)))
    mesh = mesh_from_vertices(V, F)
    return Body(mesh, {"length": Lx, "width": Ly, "draft": H},
                {"length": np.array(dL), "width": np.array(dW), "draft": np.array(dH)})


def translated(body: Body, dx=0.0, dy=0.0) -> Body:
    m = body.mesh
    V = m.vertices + np.array([dx, dy, 0.0])
    mesh = Mesh(V, m.faces, m.centers + np.array([dx, dy, 0.0]), m.normals, m.areas, m.radii)
    return Body(mesh, dict(body.params), dict(body.dvertices))


def translational_dofs(mesh: Mesh, name: str):
    """reference src/bodies.jl:42-53: a constant field per panel."""
    d = np.zeros((mesh.nfaces, 3))
    idx = {"Surge": 0, "Sway": 1, "Heave": 2}[name]
    d[:, idx] = 1.0
    return d


def rotational_dofs(mesh: Mesh, name: str, rotation_center=(0.0, 0.0, 0.0)):
    """reference src/bodies.jl:55-68: axis x (centre - rotation centre)."""
    axis = {"Roll": [1.0, 0, 0], "Pitch": [0, 1.0, 0], "Yaw": [0, 0, 1.0]}[name]
    pos = mesh.centers - np.asarray(rotation_center, dtype=np.float64)
    return np.cross(np.asarray(axis), pos)


def rigid_dofs(mesh: Mesh, names, rotation_center=(0.0, 0.0, 0.0)):
    out = []
    for nm in names:
        if nm in ("Surge", "Sway", "Heave"):
            out.append(translational_dofs(mesh, nm))
        else:
            out.append(rotational_dofs(mesh, nm, rotation_center))
    return out


# --------------------------------------------------------------------------
# surfaces of revolution (reference src/meshes.jl:247-479)
# --------------------------------------------------------------------------
def axisymmetric_mesh(profile, n_theta=72, dprofile=None):
    """Revolve a (r, z) profile about the z axis (src/meshes.jl:247-310).

    profile: (np, 2) points from the waterline down; a point with r = 0 is a
    single vertex and the faces next to it are triangles, stored as quads
    with the last vertex repeated (src/meshes.jl:372-373).  Ring i of a
    regular point holds n_theta vertices at theta = 2 pi q / n_theta.
    dprofile: optional {name: (np, 2) d profile / d name}; the matching
    d vertices / d name are returned alongside.
    Returns (vertices (nv, 3), faces (nf, 4) 0-based, {name: dvertices}).
    """
    P = np.asarray(profile, dtype=np.float64)
    if n_theta < 3:
        raise ValueError("n_theta must be at least 3")
    if P.shape[0] < 2:
        raise ValueError("profile must contain at least two points")
    if np.any(P[:, 0] < 0):
        raise ValueError("profile radii must be nonnegative")
    th = 2 * np.pi * np.arange(n_theta) / n_theta
    ct, st = np.cos(th), np.sin(th)
    V, starts, sizes = [], [], []
    dnames = list(dprofile) if dprofile else []
    dV = {k: [] for k in dnames}
    for p, (r, z) in enumerate(P):
        starts.append(len(V))
        if r == 0.0:
            sizes.append(1)
            V.append([0.0, 0.0, z])
            for k in dnames:
                dr, dz = dprofile[k][p]
                dV[k].append([0.0, 0.0, dz])
        else:
            sizes.append(n_theta)
            for q in range(n_theta):
                V.append([r * ct[q], r * st[q], z])
                for k in dnames:
                    dr, dz = dprofile[k][p]
                    dV[k].append([dr * ct[q], dr * st[q], dz])
    F = []
    for layer in range(P.shape[0] - 1):
        s1, s2 = starts[layer], starts[layer + 1]
        n1, n2 = sizes[layer], sizes[layer + 1]
        if n1 == 1 and n2 == 1:
            raise ValueError("adjacent points cannot both have radius 0")
        for q in range(n_theta):
            q2 = (q + 1) % n_theta
            if n1 == 1:
                F.append([s1, s2 + q, s2 + q2, s2 + q2])
            elif n2 == 1:
                F.append([s1 + q, s2, s1 + q2, s1 + q2])
            else:
                F.append([s1 + q, s2 + q, s2 + q2, s1 + q2])
    return (np.array(V), np.array(F, dtype=np.int64),
            {k: np.array(v) for k, v in dV.items()})


def wavebot_profile(cylinder_height=0.16, frustum_height=0.37, top_radius=0.88, bottom_radius=0.35,
                    n_points=(5, 10, 5)):
    """WaveBot hull profile (src/meshes.jl:312-349): a cylinder below the
    waterline, a frustum, and the bottom disk to the axis.  Returns the (r, z)
    points and their derivatives w.r.t. (cylinder_height, frustum_height,
    top_radius, bottom_radius), keyed "d1", "d2", "r1", "r2"."""
    n_bottom, n_frustum, n_cylinder = n_points
    h1, h2, r1, r2 = map(float, (cylinder_height, frustum_height, top_radius, bottom_radius))
    pts, d = [], {"d1": [], "d2": [], "r1": [], "r2": []}

    def add(r, z, dr, dz):
        pts.append([r, z])
        for k in d:
            d[k].append([dr.get(k, 0.0), dz.get(k, 0.0)])

    for j in range(n_cylinder):
        t = j / (n_cylinder - 1)
        add(r1, -h1 * t, {"r1": 1.0}, {"d1": -t})
    for j in range(1, n_frustum):
        t = j / (n_frustum - 1)
        add(r1 + (r2 - r1) * t, -h1 - h2 * t, {"r1": 1.0 - t, "r2": t}, {"d1": -1.0, "d2": -t})
    zb = -h1 - h2
    for j in range(1, n_bottom):
        t = j / (n_bottom - 1)
        add(r2 * (1 - t), zb, {"r2": 1.0 - t}, {"d1": -1.0, "d2": -1.0})
    return np.array(pts), {k: np.array(v) for k, v in d.items()}


def reference_face_geometry(vertices, faces):
    """Face quantities the way the reference builds a Mesh from a mesh of
    quadrangles and triangles (src/meshes.jl:352-430): area-weighted centre
    of the triangles (a, b, c) and (a, c, d), area = their sum, normal =
    normalised cross(b - a, c - a), radius = |a - centre| (first vertex)."""
    V = np.asarray(vertices, dtype=np.float64)
    Fc = np.asarray(faces)
    a, b, c, d = (V[Fc[:, i]] for i in range(4))
    a1, _ = _tri(a, b, c)
    a2, _ = _tri(a, c, d)
    area = a1 + a2
    center = ((a + b + c) / 3.0 * a1[:, None] + (a + c + d) / 3.0 * a2[:, None]) / area[:, None]
    nrm = np.cross(b - a, c - a)
    normal = nrm / np.linalg.norm(nrm, axis=1)[:, None]
    radius = np.linalg.norm(a - center, axis=1)
    return center, normal, area, radius


def wavebot_mesh(r1=0.88, r2=0.35, d1=0.16, d2=0.37, n_points=(5, 10, 5), n_theta=72) -> Body:
    """src/meshes.jl:458-479: the WaveBot hull of top radius r1, bottom
    radius r2, cylinder height d1 and frustum height d2, as a Mesh with the
    reference's face conventions, plus d vertices / d (r1, r2, d1, d2).  Its
    quads are planar trapezoids and stay planar along these parameters, so
    face_geometry_jvp's normal (of the diagonals) has the same derivative as
    the reference's cross(b - a, c - a)."""
    prof, dprof = wavebot_profile(d1, d2, r1, r2, n_points)
    V, F, dV = axisymmetric_mesh(prof, n_theta, dprof)
    c, n, a, r = reference_face_geometry(V, F)
    return Body(Mesh(V, F, c, n, a, r), {"r1": r1, "r2": r2, "d1": d1, "d2": d2}, dV)
