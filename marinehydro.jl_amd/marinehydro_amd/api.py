"""Host-side mirror of MarineHydro.jl's hot-path API over libmhip.so.

Same names and argument meaning as the reference's exported functions
(reference src/MarineHydro.jl:13-50):

  SETTINGS, set_g, set_rho                 src/constants.jl:2-12
  Rankine, RankineReflected, RankineReflectedNegative, GFWu,
  ExactGuevelDelhommeau                    src/green_functions/*.jl
  greens, gradient_greens, integral, integral_gradient
                                           src/green_functions/abstract_greens_function.jl:16-83
  element                                  src/meshes.jl:171-179 (0-based here)
  assemble_matrices, assemble_matrix_wu    src/matrix_assembly.jl:4-42, :97-102
  solve                                    src/matrix_assembly.jl:105-133
  radiation_bc, AiryBC, airy_waves_velocity, airy_waves_potential,
  airy_waves_pressure, integrate_pressure  src/waves.jl:4-25, :112-154
  calculate_radiation_forces, DiffractionForce, diffraction_force
                                           src/waves.jl:29-59

Every numerical function on the N x N path runs in the HIP library; the O(N)
helpers (boundary conditions, pressure integration) are numpy restatements
offered for API completeness, and the batched sweep ``freq_sweep`` runs them
on the GPU.  Matrices come back as torch complex128 CUDA tensors whose memory
is column-major (the Julia layout).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .meshes import Mesh

# ---------------------------------------------------------------------------
# settings (src/constants.jl:2-12)
# ---------------------------------------------------------------------------


@dataclass
class _Constants:
    rho: float = 1000.0  # fluid density [kg/m^3]
    g: float = 9.81      # gravity [m/s^2]


SETTINGS = _Constants()


def set_g(val):
    SETTINGS.g = float(val)


def set_rho(val):
    SETTINGS.rho = float(val)


# ---------------------------------------------------------------------------
# Green-function terms
# ---------------------------------------------------------------------------
class GreensFunction:
    mask = 0

    def __repr__(self):
        return f"{type(self).__name__}()"

    def __eq__(self, other):
        return type(self) is type(other)

    def __hash__(self):
        return hash(type(self))


class Rankine(GreensFunction):
    """src/green_functions/rankine.jl"""
    mask = L.GF_RANKINE


class RankineReflected(GreensFunction):
    """src/green_functions/rankine_reflected.jl"""
    mask = L.GF_RANKINE_REFLECTED


class RankineReflectedNegative(GreensFunction):
    """src/green_functions/rankine_reflected_negative.jl"""
    mask = L.GF_RANKINE_REFLECTED_NEG


class GFWu(GreensFunction):
    """src/green_functions/wu.jl"""
    mask = L.GF_WU


class ExactGuevelDelhommeau(GreensFunction):
    """src/green_functions/exact_Guevel_Delhommeau.jl"""
    mask = L.GF_DELHOMMEAU


class TabulatedDelhommeau(GreensFunction):
    """The Guevel-Delhommeau term from a table of the exact one (no
    reference counterpart): bicubic Hermite interpolation of its real part
    with the log singularity subtracted, closed-form imaginary part; within
    about 1e-9 of ExactGuevelDelhommeau.  Derivatives use the exact term."""
    mask = L.GF_DELHOMMEAU_TABLE


_WAVE_BITS = L.GF_WU | L.GF_DELHOMMEAU | L.GF_DELHOMMEAU_TABLE


def gf_layers(gfs) -> list:
    """A GreensFunction or any tuple/vector of them (the reference sums over
    the elements, abstract_greens_function.jl:18-83, repeats and several wave
    terms included) as a list of masks whose sum is the tuple: each mask
    holds every term at most once and at most one wave term, which one
    library call evaluates.  Most tuples are one mask."""
    if isinstance(gfs, GreensFunction):
        return [gfs.mask]
    counts = {}
    for g in gfs:
        if not isinstance(g, GreensFunction):
            raise TypeError(f"not a GreensFunction: {g!r}")
        counts[g.mask] = counts.get(g.mask, 0) + 1
    if not counts:
        raise ValueError("empty Green-function tuple")
    layers = []
    while any(counts.values()):
        m = 0
        for bit in sorted(counts):
            if counts[bit] and not (bit & _WAVE_BITS and m & _WAVE_BITS):
                m |= bit
                counts[bit] -= 1
        layers.append(m)
    return layers


def gf_mask(gfs) -> int:
    """Bit mask of a GreensFunction or a tuple of them that one library call
    evaluates: every term at most once and at most one wave term.  Other
    tuples are summed term group by term group where the result is linear
    in the Green function (gf_eval and the functions on it,
    assemble_matrices, assemble_matrices_vjp); the sweeps need a single
    mask."""
    layers = gf_layers(gfs)
    if len(layers) > 1:
        raise ValueError("this path takes each Green-function term at most once and at most one wave term "
                         "(GFWu, ExactGuevelDelhommeau or TabulatedDelhommeau); assemble_matrices and "
                         "solve accept any tuple")
    return layers[0]


# ---------------------------------------------------------------------------
# elements (src/meshes.jl:157-205)
# ---------------------------------------------------------------------------
@dataclass
class Element:
    center: np.ndarray
    vertices: np.ndarray = None  # 4 x 3
    normal: np.ndarray = None
    area: float = 1.0
    radius: float = 0.0


def element(mesh: Mesh, j: int) -> Element:
    """Panel j of mesh (0-based)."""
    return Element(center=np.array(mesh.centers[j]), vertices=np.array(mesh.vertices[mesh.faces[j]]),
                   normal=np.array(mesh.normals[j]), area=float(mesh.areas[j]),
                   radius=float(mesh.radii[j]))


def _field(e, name, default=None):
    if isinstance(e, dict):
        return e.get(name, default)
    return getattr(e, name, default)


def _panel_record(e) -> np.ndarray:
    """mh_panel_t: center[3], vertices[4][3], normal[3], area, radius."""
    rec = np.zeros(20)
    rec[0:3] = np.asarray(_field(e, "center"), dtype=np.float64).reshape(3)
    v = _field(e, "vertices")
    if v is not None:
        rec[3:15] = np.asarray(v, dtype=np.float64).reshape(4, 3).reshape(-1)
    n = _field(e, "normal")
    if n is not None:
        rec[15:18] = np.asarray(n, dtype=np.float64).reshape(3)
    a = _field(e, "area")
    rec[18] = 1.0 if a is None else float(a)
    r = _field(e, "radius")
    rec[19] = 0.0 if r is None else float(r)
    return rec


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("marinehydro_amd needs a ROCm GPU (gfx950); none is visible")
    return torch


def gf_eval(gfs, points, panels, wavenumber=1.0, op=L.OP_INTEGRAL, wrt_first=False, device=0):
    """Batched per-pair evaluation on the GPU.

    points: (npairs, 3); panels: (npairs, 20) mh_panel_t records.
    Returns complex (npairs,) for value ops, (npairs, 3) for gradient ops.
    """
    torch = _torch()
    dev = f"cuda:{device}"
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    pan = np.ascontiguousarray(np.asarray(panels, dtype=np.float64).reshape(-1, 20))
    npairs = pts.shape[0]
    if pan.shape[0] != npairs:
        raise ValueError("points and panels must have the same length")
    d_pts = torch.from_numpy(np.ascontiguousarray(pts.T)).to(dev)
    d_pan = torch.from_numpy(pan).to(dev)
    grad = op in (L.OP_GRADIENT_GREENS, L.OP_INTEGRAL_GRADIENT)
    out = torch.empty((3 if grad else 1, npairs), dtype=torch.complex128, device=dev)
    ctx = L.context(device)
    res = None
    for m in gf_layers(gfs):  # the sum over the tuple, term group by term group
        L.check(L.lib().mh_gf_eval(ctx, m, op, int(bool(wrt_first)), npairs,
                                   d_pts.data_ptr(), d_pan.data_ptr(), float(wavenumber),
                                   out.data_ptr()), "gf_eval")
        part = out.cpu().numpy()
        res = part if res is None else res + part
    return res.T if grad else res[0]


def _k(w):
    return 1.0 if w is None else float(w)


def greens(gf, element_1, element_2, wavenumber=None):
    """abstract_greens_function.jl:5-16; point value between two centres."""
    rec = _panel_record(element_2)
    return gf_eval(gf, [_field(element_1, "center")], [rec], _k(wavenumber), L.OP_GREENS)[0]


def gradient_greens(gf, element_1, element_2, wavenumber=None, with_respect_to_first_variable=False):
    """abstract_greens_function.jl:19-32."""
    rec = _panel_record(element_2)
    g = gf_eval(gf, [_field(element_1, "center")], [rec], _k(wavenumber), L.OP_GRADIENT_GREENS,
                with_respect_to_first_variable)[0]
    return g


def integral(gf, element_1, element_2, wavenumber=None):
    """abstract_greens_function.jl:35-48; integral over panel element_2."""
    rec = _panel_record(element_2)
    return gf_eval(gf, [_field(element_1, "center")], [rec], _k(wavenumber), L.OP_INTEGRAL)[0]


def integral_gradient(gf, element_1, element_2, wavenumber=None, with_respect_to_first_variable=False):
    """abstract_greens_function.jl:51-64."""
    rec = _panel_record(element_2)
    return gf_eval(gf, [_field(element_1, "center")], [rec], _k(wavenumber), L.OP_INTEGRAL_GRADIENT,
                   with_respect_to_first_variable)[0]


def both_integral_and_integral_gradient(gf, element_1, element_2, wavenumber=None,
                                        with_respect_to_first_variable=False):
    """abstract_greens_function.jl:67-80."""
    return (integral(gf, element_1, element_2, wavenumber),
            integral_gradient(gf, element_1, element_2, wavenumber, with_respect_to_first_variable))


# ---------------------------------------------------------------------------
# device mesh
# ---------------------------------------------------------------------------
class DeviceMesh:
    """A Mesh copied to device memory in the reference's column-major layout."""

    def __init__(self, mesh: Mesh, device=0):
        torch = _torch()
        self.mesh = mesh
        self.device = device
        dev = f"cuda:{device}"

        def cm(a, dt):  # column-major (n, c) == row-major (c, n)
            return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=dt).T)).to(dev)

        self.vertices = cm(mesh.vertices, np.float64)
        self.faces = cm(mesh.faces, np.int64)
        self.centers = cm(mesh.centers, np.float64)
        self.normals = cm(mesh.normals, np.float64)
        self.areas = torch.from_numpy(np.ascontiguousarray(mesh.areas)).to(dev)
        self.radii = torch.from_numpy(np.ascontiguousarray(mesh.radii)).to(dev)
        self.struct = L.MeshT(mesh.nfaces, mesh.nvertices, self.vertices.data_ptr(),
                              self.faces.data_ptr(), self.centers.data_ptr(),
                              self.normals.data_ptr(), self.areas.data_ptr(),
                              self.radii.data_ptr())

    @property
    def nfaces(self):
        return self.mesh.nfaces


def _dmesh(mesh, device=0) -> DeviceMesh:
    return mesh if isinstance(mesh, DeviceMesh) else DeviceMesh(mesh, device)


def _colmajor_empty(torch, n, batch, device):
    """(batch, n, n) complex128 whose [b].T views are column-major matrices."""
    return torch.empty((batch, n, n), dtype=torch.complex128, device=f"cuda:{device}")


# ---------------------------------------------------------------------------
# assembly (src/matrix_assembly.jl)
# ---------------------------------------------------------------------------
def assemble_matrices_batched(green_functions, mesh, wavenumbers, direct=True, device=0,
                              formulation=None):
    """Assemble for many wavenumbers in one launch.

    Returns (S, DK) as (nk, n, n) tensors in logical [i, j] indexing (memory
    column-major); formulation "both" returns (S, D, K).
    """
    torch = _torch()
    dm = _dmesh(mesh, device)
    n = dm.nfaces
    ks = np.ascontiguousarray(np.asarray(wavenumbers, dtype=np.float64).reshape(-1))
    nk = ks.shape[0]
    form = formulation or (L.DIRECT if direct else L.INDIRECT)
    S = _colmajor_empty(torch, n, nk, dm.device)
    D = _colmajor_empty(torch, n, nk, dm.device) if form & L.DIRECT else None
    K = _colmajor_empty(torch, n, nk, dm.device) if form & L.INDIRECT else None
    ctx = L.context(dm.device)
    layers = gf_layers(green_functions)
    acc = None
    for q, m in enumerate(layers):  # the sum over the tuple, term group by term group
        L.check(L.lib().mh_assemble(ctx, C.byref(dm.struct), m, form, nk,
                                    ks.ctypes.data, S.data_ptr(), D.data_ptr() if D is not None else None,
                                    K.data_ptr() if K is not None else None, n, n * n),
                "assemble_matrices")
        parts = [X for X in (S, D, K) if X is not None]
        if len(layers) > 1:
            acc = [X.clone() for X in parts] if acc is None else [a + X for a, X in zip(acc, parts)]
    if acc is not None:
        # D and K carry the free term 1/2 on the diagonal once, not per group
        # (src/matrix_assembly.jl:18, :23)
        S, rest = acc[0], acc[1:]
        eye = torch.eye(n, dtype=torch.complex128, device=S.device) * (0.5 * (len(layers) - 1))
        rest = [X - eye for X in rest]
        D = rest.pop(0) if D is not None else None
        K = rest.pop(0) if K is not None else None
    tr = lambda X: X.transpose(1, 2)
    if form == L.BOTH:
        return tr(S), tr(D), tr(K)
    return tr(S), tr(D if D is not None else K)


def assemble_matrices(green_functions, mesh, wavenumber, direct=True, device=0):
    """src/matrix_assembly.jl:4-42 (the default `assemble_matrices`, :97).

    Returns (S, D) when direct, else (S, K); n x n complex128 CUDA tensors.
    """
    S, DK = assemble_matrices_batched(green_functions, mesh, [wavenumber], direct, device)
    return S[0], DK[0]


def assemble_matrix_wu(mesh, wavenumber, direct=True, device=0):
    """src/matrix_assembly.jl:100-102."""
    return assemble_matrices((Rankine(), RankineReflected(), GFWu()), mesh, wavenumber, direct,
                             device)


def _as_colmajor_batch(torch, X, device):
    """Copy an (n, n) or (b, n, n) matrix (torch or numpy, logical [i, j])
    into a fresh (b, n, n) buffer whose memory is column-major per item."""
    if not torch.is_tensor(X):
        X = torch.from_numpy(np.asarray(X, dtype=np.complex128))
    X = X.to(device=f"cuda:{device}", dtype=torch.complex128)
    if X.dim() == 2:
        X = X.unsqueeze(0)
    # memory [b][j][i] = X[b, i, j], always a fresh buffer: .contiguous()
    # alone would return X's own storage when it is already column-major
    # (the assembly's output), and the FP64 solve factors its buffer in place
    out = torch.empty((X.shape[0], X.shape[2], X.shape[1]), dtype=torch.complex128, device=X.device)
    out.copy_(X.transpose(1, 2))
    return out


def solve(D, S, bc, direct=True, device=0, precision="fp64"):
    """src/matrix_assembly.jl:105-133.  D (or K) is not modified.
    precision="mixed": FP32 factorisation of D (K) refined in FP64 to a
    1e-10 relative residual (mh_solve_mixed)."""
    torch = _torch()
    Dm = _as_colmajor_batch(torch, D, device)
    Sm = _as_colmajor_batch(torch, S, device)
    b = bc if torch.is_tensor(bc) else torch.from_numpy(np.asarray(bc, dtype=np.complex128))
    b = b.to(device=f"cuda:{device}", dtype=torch.complex128)
    vec = b.dim() == 1
    if vec:
        b = b.reshape(-1, 1)
    n = Dm.shape[1]
    if Dm.shape[0] != 1 or Sm.shape != Dm.shape or b.shape[0] != n:
        raise ValueError("solve: D, S must be n x n and bc of length n")
    nrhs = b.shape[1]
    bm = b.transpose(0, 1).contiguous()  # memory column-major n x nrhs
    phi = torch.empty_like(bm)
    ctx = L.context(device)
    fn = {"fp64": L.lib().mh_solve, "mixed": L.lib().mh_solve_mixed}[precision]
    L.check(fn(ctx, n, 1, int(bool(direct)), Dm.data_ptr(), Sm.data_ptr(), n, n * n,
               bm.data_ptr(), nrhs, phi.data_ptr(), n, n * nrhs), "solve")
    out = phi.transpose(0, 1)
    return out[:, 0] if vec else out


# ---------------------------------------------------------------------------
# waves, boundary conditions, forces (src/waves.jl) - O(N) host helpers
# ---------------------------------------------------------------------------
def radiation_bc(mesh, dof, omega):
    """src/waves.jl:4-19 and :89-102 (dof: 3-vector or nfaces x 3 field)."""
    d = np.broadcast_to(np.asarray(dof, dtype=np.float64), mesh.normals.shape)
    return -1j * omega * np.sum(mesh.normals * d, axis=1)


def integrate_pressure(mesh, pressure, dof):
    """src/waves.jl:21-25."""
    d = np.broadcast_to(np.asarray(dof, dtype=np.float64), mesh.normals.shape)
    amp = -np.sum(d * mesh.normals, axis=1)
    return np.sum(np.asarray(pressure).reshape(-1) * amp * mesh.areas)


def airy_waves_potential(points, omega, beta=0.0):
    """src/waves.jl:112-119."""
    P = np.asarray(points, dtype=np.float64)
    k = omega ** 2 / SETTINGS.g
    wbar = P[:, 0] * np.cos(beta) + P[:, 1] * np.sin(beta)
    return -1j * SETTINGS.g / omega * np.exp(k * P[:, 2]) * np.exp(1j * k * wbar)


def airy_waves_velocity(points, omega, beta=0.0):
    """src/waves.jl:121-134 (infinite depth)."""
    P = np.asarray(points, dtype=np.float64)
    k = omega ** 2 / SETTINGS.g
    wbar = P[:, 0] * np.cos(beta) + P[:, 1] * np.sin(beta)
    cih = np.exp(k * P[:, 2])
    pre = SETTINGS.g * k / omega * np.exp(1j * k * wbar)
    return pre[:, None] * np.stack([np.cos(beta) * cih, np.sin(beta) * cih, -1j * cih], axis=1)


def airy_waves_pressure(points, omega, beta=0.0):
    """src/waves.jl:150-154."""
    return 1j * omega * SETTINGS.rho * airy_waves_potential(points, omega, beta)


def AiryBC(mesh, omega, beta=0.0):
    """src/waves.jl:138-142."""
    return -np.sum(airy_waves_velocity(mesh.centers, omega, beta) * mesh.normals, axis=1)


def diffraction_force(potential, mesh, omega, dof):
    """src/waves.jl:55-59."""
    pressure = 1j * SETTINGS.rho * np.asarray(potential) * omega
    return integrate_pressure(mesh, pressure, dof)


def FroudeKrylovForce(mesh, omega, dof, beta=0.0):
    """src/waves.jl:62-66 (and :160-164 with a heading): the incident-wave
    pressure integrated over the hull, computed by libmhip."""
    FK = froude_krylov(mesh, [omega], [dof], betas=[beta])
    return complex(FK[0, 0, 0])


# ---------------------------------------------------------------------------
# batched end-to-end sweep and the reference's two physics entry points
# ---------------------------------------------------------------------------
class _Outputs:
    """Device outputs of a sweep in one zero-filled float64 buffer: one fill
    and one read-back instead of one per array.  Every segment starts at an
    even offset, so complex segments are 16-byte aligned."""

    def __init__(self, torch, dev, parts):
        self.offs, tot = {}, 0
        for name, shape, cplx in parts:
            cnt = int(np.prod(shape)) * (2 if cplx else 1)
            self.offs[name] = (tot, shape, cplx)
            tot += cnt + (cnt & 1)
        self.buf = torch.zeros(max(tot, 2), dtype=torch.float64, device=dev)
        self.host = None

    def ptr(self, name):
        return self.buf.data_ptr() + 8 * self.offs[name][0]

    def get(self, name):
        if self.host is None:
            self.host = self.buf.cpu().numpy()
        off, shape, cplx = self.offs[name]
        cnt = int(np.prod(shape)) * (2 if cplx else 1)
        a = self.host[off:off + cnt]
        return (a.view(np.complex128) if cplx else a).reshape(shape)


def device_dof_fields(mesh, dof_fields, device=0):
    """The dof fields (3-vectors or (nfaces, 3) fields) as one device tensor
    (ndof, 3, nfaces), the layout the sweeps read: pass it as dof_fields to
    keep them resident across calls."""
    torch = _torch()
    n = mesh.nfaces
    fields = [np.broadcast_to(np.asarray(f, dtype=np.float64), (n, 3)) for f in dof_fields]
    return torch.from_numpy(np.ascontiguousarray(np.stack([f.T for f in fields]))).to(f"cuda:{device}")


def _dof_input(torch, dof_fields, n, dev):
    """(device tensor (ndof, 3, n), ndof) from host fields or a tensor of
    device_dof_fields."""
    if torch.is_tensor(dof_fields):
        if dof_fields.dim() != 3 or dof_fields.shape[1:] != (3, n) or dof_fields.dtype != torch.float64:
            raise ValueError(f"device dof fields must be float64 (ndof, 3, {n})")
        return dof_fields.to(dev).contiguous(), dof_fields.shape[0]
    fields = [np.broadcast_to(np.asarray(f, dtype=np.float64), (n, 3)) for f in dof_fields]
    ndof = len(fields)
    dofs = torch.from_numpy(np.ascontiguousarray(
        np.stack([f.T for f in fields]) if ndof else np.zeros((1, 3, n)))).to(dev)
    return dofs, ndof


def freq_sweep(mesh, omegas, dof_fields, betas=(), green_functions=None, direct=True,
               precision="fp64", device=0, g=None, rho=None):
    """All frequencies in one batched pass (one LU per frequency, shared by
    every radiating dof and heading).

    dof_fields: list of 3-vectors or (nfaces, 3) fields.
    Returns numpy (A, B, F): A[w, i, j], B[w, i, j] = added mass / damping on
    influenced dof i due to radiating dof j; F[w, b, i] = diffraction force
    on dof i for heading b.
    """
    torch = _torch()
    dm = _dmesh(mesh, device)
    n = dm.nfaces
    gfs = green_functions if green_functions is not None else (Rankine(), RankineReflected(), GFWu())
    om = np.ascontiguousarray(np.asarray(omegas, dtype=np.float64).reshape(-1))
    be = np.ascontiguousarray(np.asarray(betas, dtype=np.float64).reshape(-1))
    dev = f"cuda:{dm.device}"
    dofs, ndof = _dof_input(torch, dof_fields, n, dev)
    nbeta, nw = be.shape[0], om.shape[0]
    nd = max(ndof, 1)
    out = _Outputs(torch, dev, [("A", (nw, nd, nd), False), ("B", (nw, nd, nd), False),
                                ("F", (nw, max(nbeta, 1), nd), True)])
    prec = {"fp64": L.FP64, "mixed": L.MIXED}[precision]
    ctx = L.context(dm.device)
    L.check(L.lib().mh_freq_sweep(ctx, C.byref(dm.struct), gf_mask(gfs), int(bool(direct)), prec,
                                  nw, om.ctypes.data, float(g or SETTINGS.g),
                                  float(rho or SETTINGS.rho), ndof, dofs.data_ptr(), nbeta,
                                  be.ctypes.data if nbeta else None, out.ptr("A"), out.ptr("B"),
                                  out.ptr("F")), "freq_sweep")
    # library layout A[w][j][i] -> A[w, i, j]
    A_ = np.ascontiguousarray(out.get("A").transpose(0, 2, 1)[:, :ndof, :ndof])
    B_ = np.ascontiguousarray(out.get("B").transpose(0, 2, 1)[:, :ndof, :ndof])
    F_ = np.ascontiguousarray(out.get("F")[:, :nbeta, :ndof])
    return A_, B_, F_


class SweepGraph:
    """freq_sweep(precision="fp64") of one body recorded once as a HIP graph
    and replayed by launch(omegas, betas) (mh_sweep_graph_*).  The mesh must
    be a DeviceMesh: its arrays may be updated in place between launches
    (same sizes), and the graph reads them at every launch.  The number of
    frequencies and headings is fixed at creation."""

    def __init__(self, mesh, nomega, dof_fields, nbeta=0, green_functions=None, direct=True, device=0,
                 g=None, rho=None):
        torch = _torch()
        if not isinstance(mesh, DeviceMesh):
            raise TypeError("SweepGraph: the mesh must be a DeviceMesh (its arrays are recorded)")
        self.dm, self.nomega, self.nbeta = mesh, int(nomega), int(nbeta)
        n = mesh.nfaces
        gfs = green_functions if green_functions is not None else (Rankine(), RankineReflected(), GFWu())
        fields = [np.broadcast_to(np.asarray(f, dtype=np.float64), (n, 3)) for f in dof_fields]
        self.ndof = len(fields)
        dev = f"cuda:{mesh.device}"
        self.dofs = torch.from_numpy(np.ascontiguousarray(
            np.stack([f.T for f in fields]) if self.ndof else np.zeros((1, 3, n)))).to(dev)
        nd = max(self.ndof, 1)
        self.out = _Outputs(torch, dev, [("A", (self.nomega, nd, nd), False), ("B", (self.nomega, nd, nd), False),
                                         ("F", (self.nomega, max(self.nbeta, 1), nd), True)])
        self.handle = C.c_void_p()
        L.check(L.lib().mh_sweep_graph_create(
            L.context(mesh.device), C.byref(mesh.struct), gf_mask(gfs), int(bool(direct)), self.nomega, self.ndof,
            self.dofs.data_ptr(), self.nbeta, float(g or SETTINGS.g), float(rho or SETTINGS.rho),
            self.out.ptr("A"), self.out.ptr("B"), self.out.ptr("F"), C.byref(self.handle)), "sweep_graph_create")

    def launch(self, omegas, betas=()):
        """Returns numpy (A, B, F) laid out as freq_sweep's."""
        om = np.ascontiguousarray(np.asarray(omegas, dtype=np.float64).reshape(-1))
        be = np.ascontiguousarray(np.asarray(betas, dtype=np.float64).reshape(-1))
        if om.shape[0] != self.nomega or be.shape[0] != self.nbeta:
            raise ValueError(f"SweepGraph.launch: {self.nomega} frequencies and {self.nbeta} headings expected")
        L.check(L.lib().mh_sweep_graph_launch(L.context(self.dm.device), self.handle, om.ctypes.data,
                                              be.ctypes.data if self.nbeta else None), "sweep_graph_launch")
        self.out.host = None  # read this launch's outputs
        nd = self.ndof
        A_ = np.ascontiguousarray(self.out.get("A").transpose(0, 2, 1)[:, :nd, :nd])
        B_ = np.ascontiguousarray(self.out.get("B").transpose(0, 2, 1)[:, :nd, :nd])
        F_ = np.ascontiguousarray(self.out.get("F")[:, :self.nbeta, :nd])
        return A_, B_, F_

    def close(self):
        if self.handle:
            L.lib().mh_sweep_graph_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def calculate_radiation_forces(mesh, dof, omega):
    """src/waves.jl:29-37: [A, B] for one dof 3-vector (Wu, direct)."""
    A, B, _ = freq_sweep(mesh, [omega], [dof])
    return [float(A[0, 0, 0]), float(B[0, 0, 0])]


def DiffractionForce(mesh, omega, dof):
    """src/waves.jl:40-52: complex diffraction force, heading 0 (Wu, direct)."""
    _, _, F = freq_sweep(mesh, [omega], [dof], betas=[0.0])
    return complex(F[0, 0, 0])


# ---------------------------------------------------------------------------
# gradients (adjoint of the solve, tangents of the geometry)
# ---------------------------------------------------------------------------
@dataclass
class Tangents:
    """Derivatives of the mesh fields (and frequencies) along nparam
    parameter directions.  Arrays are (nparam, n, 3) / (nparam, n); any may
    be None (zero).  domegas: (nparam, nomega) or None."""
    nparam: int
    dvertices: np.ndarray = None
    dcenters: np.ndarray = None
    dnormals: np.ndarray = None
    dareas: np.ndarray = None
    ddof_fields: np.ndarray = None  # (nparam, ndof, nfaces, 3)
    domegas: np.ndarray = None


def shape_tangents(mesh, dvertices_list, nomega=0, omega_param=None, dof_axes=None):
    """Tangents of a mesh whose face quantities are functions of its vertices
    (the synthetic bodies): dvertices_list[m] = d vertices / d theta_m (None
    for a parameter that is not geometric).  omega_param: index m of the
    frequency parameter, if any.  dof_axes: per dof, None for a translation or
    the rotation axis of a rotational dof (its field depends on the centres).
    """
    from .meshes import face_geometry_jvp

    npar = len(dvertices_list)
    nv, nf = mesh.nvertices, mesh.nfaces
    dV = np.zeros((npar, nv, 3))
    dC = np.zeros((npar, nf, 3))
    dN = np.zeros((npar, nf, 3))
    dA = np.zeros((npar, nf))
    for m, dv in enumerate(dvertices_list):
        if dv is None:
            continue
        dV[m] = dv
        dC[m], dN[m], dA[m] = face_geometry_jvp(mesh.vertices, mesh.faces, dv)
    ddof = None
    if dof_axes is not None and any(a is not None for a in dof_axes):
        ddof = np.zeros((npar, len(dof_axes), nf, 3))
        for d, ax in enumerate(dof_axes):
            if ax is not None:
                for m in range(npar):
                    ddof[m, d] = np.cross(np.asarray(ax, dtype=np.float64), dC[m])
    dom = None
    if omega_param is not None:
        dom = np.zeros((npar, nomega))
        dom[omega_param, :] = 1.0
    return Tangents(npar, dV, dC, dN, dA, ddof, dom)


def freq_sweep_grad(mesh, omegas, dof_fields, betas=(), tangents: Tangents = None,
                    green_functions=None, direct=True, precision="fp64", device=0, g=None, rho=None):
    """freq_sweep plus derivatives along the tangent directions.

    Returns (A, B, F, dA, dB, dF): A, B, F as freq_sweep; dA[w, m, i, j],
    dB[w, m, i, j], dF[w, m, b, i] = derivatives along parameter m.
    """
    torch = _torch()
    dm = _dmesh(mesh, device)
    n, nv = dm.nfaces, dm.mesh.nvertices
    gfs = green_functions if green_functions is not None else (Rankine(), RankineReflected(), GFWu())
    om = np.ascontiguousarray(np.asarray(omegas, dtype=np.float64).reshape(-1))
    be = np.ascontiguousarray(np.asarray(betas, dtype=np.float64).reshape(-1))
    dev = f"cuda:{dm.device}"
    dofs, ndof = _dof_input(torch, dof_fields, n, dev)
    nbeta, nw = be.shape[0], om.shape[0]
    T = tangents
    npar = T.nparam

    def up(a, shape):  # (nparam, n, c) -> column-major per parameter
        if a is None:
            return None
        if torch.is_tensor(a):  # already device-resident, column-major per parameter
            return a
        a = np.asarray(a, dtype=np.float64).reshape(shape)
        if a.ndim == 3:
            a = np.transpose(a, (0, 2, 1))
        elif a.ndim == 4:
            a = np.transpose(a, (0, 1, 3, 2))
        return torch.from_numpy(np.ascontiguousarray(a)).to(dev)

    dV = up(T.dvertices, (npar, nv, 3))
    dC = up(T.dcenters, (npar, n, 3))
    dN = up(T.dnormals, (npar, n, 3))
    dA_ = up(T.dareas, (npar, n))
    ddof = up(T.ddof_fields, (npar, ndof, n, 3))
    dom = None if T.domegas is None else np.ascontiguousarray(np.asarray(T.domegas, dtype=np.float64).reshape(npar, nw))
    mp, nb = max(npar, 1), max(nbeta, 1)
    out = _Outputs(torch, dev, [("A", (nw, ndof, ndof), False), ("B", (nw, ndof, ndof), False),
                                ("F", (nw, nb, ndof), True), ("gA", (nw, mp, ndof, ndof), False),
                                ("gB", (nw, mp, ndof, ndof), False), ("gF", (nw, mp, nb, ndof), True)])
    ptr = lambda t: t.data_ptr() if t is not None else None
    prec = {"fp64": L.FP64, "mixed": L.MIXED}[precision]
    ctx = L.context(dm.device)
    L.check(L.lib().mh_freq_sweep_grad(
        ctx, C.byref(dm.struct), gf_mask(gfs), int(bool(direct)), prec, nw, om.ctypes.data,
        float(g or SETTINGS.g), float(rho or SETTINGS.rho), ndof, dofs.data_ptr(), nbeta,
        be.ctypes.data if nbeta else None, npar, ptr(dV), ptr(dC), ptr(dN), ptr(dA_), ptr(ddof),
        dom.ctypes.data if dom is not None else None, out.ptr("A"), out.ptr("B"), out.ptr("F"),
        out.ptr("gA"), out.ptr("gB"), out.ptr("gF")), "freq_sweep_grad")
    A_ = np.ascontiguousarray(out.get("A").transpose(0, 2, 1))
    B_ = np.ascontiguousarray(out.get("B").transpose(0, 2, 1))
    F_ = np.ascontiguousarray(out.get("F")[:, :nbeta, :])
    gA_ = np.ascontiguousarray(out.get("gA").transpose(0, 1, 3, 2)[:, :npar])
    gB_ = np.ascontiguousarray(out.get("gB").transpose(0, 1, 3, 2)[:, :npar])
    gF_ = np.ascontiguousarray(out.get("gF")[:, :npar, :nbeta, :])
    return A_, B_, F_, gA_, gB_, gF_


# ---------------------------------------------------------------------------
# Froude-Krylov and excitation forces (src/waves.jl:62-66, :150-164;
# src/problems_and_results.jl:134-136)
# ---------------------------------------------------------------------------
def _fk_inputs(torch, dm, omegas, dof_fields, betas):
    n = dm.nfaces
    om = np.ascontiguousarray(np.asarray(omegas, dtype=np.float64).reshape(-1))
    be = np.ascontiguousarray(np.asarray(betas, dtype=np.float64).reshape(-1))
    fields = [np.broadcast_to(np.asarray(f, dtype=np.float64), (n, 3)) for f in dof_fields]
    dofs = torch.from_numpy(np.ascontiguousarray(np.stack([f.T for f in fields]))).to(
        f"cuda:{dm.device}")
    return om, be, fields, dofs


def froude_krylov(mesh, omegas, dof_fields, betas=(0.0,), device=0, g=None, rho=None):
    """Froude-Krylov forces for every frequency, heading and influenced dof
    on the GPU: FK[w, b, i] (numpy complex128)."""
    torch = _torch()
    dm = _dmesh(mesh, device)
    om, be, fields, dofs = _fk_inputs(torch, dm, omegas, dof_fields, betas)
    nw, nb, nd = om.shape[0], be.shape[0], len(fields)
    out = _Outputs(torch, f"cuda:{dm.device}", [("FK", (nw, max(nb, 1), max(nd, 1)), True)])
    L.check(L.lib().mh_froude_krylov(L.context(dm.device), C.byref(dm.struct), nw, om.ctypes.data,
                                     float(g or SETTINGS.g), float(rho or SETTINGS.rho), nd,
                                     dofs.data_ptr(), nb, be.ctypes.data, out.ptr("FK")),
            "froude_krylov")
    return np.ascontiguousarray(out.get("FK")[:, :nb, :nd])


def froude_krylov_grad(mesh, omegas, dof_fields, betas=(0.0,), tangents: "Tangents" = None,
                       device=0, g=None, rho=None):
    """froude_krylov plus derivatives along the tangent directions:
    returns (FK[w, b, i], dFK[w, m, b, i])."""
    torch = _torch()
    dm = _dmesh(mesh, device)
    n = dm.nfaces
    om, be, fields, dofs = _fk_inputs(torch, dm, omegas, dof_fields, betas)
    nw, nb, nd = om.shape[0], be.shape[0], len(fields)
    T = tangents
    npar = T.nparam
    dev = f"cuda:{dm.device}"
    dC = _dev_tangent(torch, T.dcenters, (npar, n, 3), dev)
    dN = _dev_tangent(torch, T.dnormals, (npar, n, 3), dev)
    dA = _dev_tangent(torch, T.dareas, (npar, n), dev)
    ddof = _dev_tangent(torch, T.ddof_fields, (npar, nd, n, 3), dev)
    dom = None if T.domegas is None else np.ascontiguousarray(
        np.asarray(T.domegas, dtype=np.float64).reshape(npar, nw))
    out = _Outputs(torch, dev, [("FK", (nw, max(nb, 1), max(nd, 1)), True),
                                ("dFK", (nw, max(npar, 1), max(nb, 1), max(nd, 1)), True)])
    ptr = lambda t: t.data_ptr() if t is not None else None
    L.check(L.lib().mh_froude_krylov_grad(
        L.context(dm.device), C.byref(dm.struct), nw, om.ctypes.data, float(g or SETTINGS.g),
        float(rho or SETTINGS.rho), nd, dofs.data_ptr(), nb, be.ctypes.data, npar, ptr(dC), ptr(dN),
        ptr(dA), ptr(ddof), dom.ctypes.data if dom is not None else None, out.ptr("FK"),
        out.ptr("dFK")), "froude_krylov_grad")
    return (np.ascontiguousarray(out.get("FK")[:, :nb, :nd]),
            np.ascontiguousarray(out.get("dFK")[:, :npar, :nb, :nd]))


def device_tangents(tangents, mesh, ndof, device=0):
    """The tangents with their mesh and dof-field arrays on the device in the
    layout the sweeps read (domegas stays a host array), so that repeated
    sweeps find them resident."""
    torch = _torch()
    T = tangents
    n, nv, npar = mesh.nfaces, mesh.nvertices, T.nparam
    dev = f"cuda:{device}"
    return Tangents(npar, _dev_tangent(torch, T.dvertices, (npar, nv, 3), dev),
                    _dev_tangent(torch, T.dcenters, (npar, n, 3), dev),
                    _dev_tangent(torch, T.dnormals, (npar, n, 3), dev),
                    _dev_tangent(torch, T.dareas, (npar, n), dev),
                    _dev_tangent(torch, T.ddof_fields, (npar, ndof, n, 3), dev), T.domegas)


def _dev_tangent(torch, a, shape, dev):
    """(nparam, n, c) host tangents -> device, column-major per parameter."""
    if a is None:
        return None
    if torch.is_tensor(a):
        return a
    a = np.asarray(a, dtype=np.float64).reshape(shape)
    if a.ndim == 3:
        a = np.transpose(a, (0, 2, 1))
    elif a.ndim == 4:
        a = np.transpose(a, (0, 1, 3, 2))
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


# ---------------------------------------------------------------------------
# reverse mode (the rrules Zygote uses on the reference's assembly and solve)
# ---------------------------------------------------------------------------
def _cotangent_batch(torch, X, n, nk, device):
    """None or an (n, n) / (nk, n, n) complex cotangent (logical [i, j]) ->
    column-major device batch."""
    if X is None:
        return None
    Xm = _as_colmajor_batch(torch, X, device)
    if Xm.shape != (nk, n, n):
        raise ValueError(f"cotangent of shape {tuple(Xm.shape)}, expected ({nk}, {n}, {n})")
    return Xm


def assemble_matrices_vjp(green_functions, mesh, wavenumbers, Sbar, DKbar=None, direct=True, device=0):
    """Pullback of assemble_matrices (src/matrix_assembly.jl:4-42) in Zygote's
    convention for complex arrays: returns a dict of real cotangents
    {"vertices": (nv, 3), "centers": (nf, 3), "normals": (nf, 3),
    "areas": (nf,), "wavenumbers": (nk,)} for the cotangents Sbar of S and
    DKbar of D (direct) or K (indirect), each (n, n) or (nk, n, n)."""
    torch = _torch()
    dm = _dmesh(mesh, device)
    n, nv = dm.nfaces, dm.mesh.nvertices
    ks = np.ascontiguousarray(np.asarray(wavenumbers, dtype=np.float64).reshape(-1))
    nk = ks.shape[0]
    Sb = _cotangent_batch(torch, Sbar, n, nk, dm.device)
    Db = _cotangent_batch(torch, DKbar, n, nk, dm.device)
    dev = f"cuda:{dm.device}"
    form = L.DIRECT if direct else L.INDIRECT
    total = None
    for m in gf_layers(green_functions):  # the pullbacks of the term groups add up
        dV = torch.zeros((3, nv), dtype=torch.float64, device=dev)
        dC = torch.zeros((3, n), dtype=torch.float64, device=dev)
        dN = torch.zeros((3, n), dtype=torch.float64, device=dev)
        dA = torch.zeros(n, dtype=torch.float64, device=dev)
        dk = torch.zeros(nk, dtype=torch.float64, device=dev)
        L.check(L.lib().mh_assemble_vjp(L.context(dm.device), C.byref(dm.struct), m, form,
                                        nk, ks.ctypes.data, Sb.data_ptr() if Sb is not None else None,
                                        Db.data_ptr() if Db is not None else None, n, n * n, dV.data_ptr(),
                                        dC.data_ptr(), dN.data_ptr(), dA.data_ptr(), dk.data_ptr()),
                "assemble_matrices_vjp")
        part = (dV, dC, dN, dA, dk)
        total = part if total is None else tuple(a + b for a, b in zip(total, part))
    dV, dC, dN, dA, dk = total
    return {"vertices": dV.cpu().numpy().T.copy(), "centers": dC.cpu().numpy().T.copy(),
            "normals": dN.cpu().numpy().T.copy(), "areas": dA.cpu().numpy(),
            "wavenumbers": dk.cpu().numpy()}


def rrule_assemble_matrices(green_functions, mesh, wavenumber, direct=True, device=0):
    """(S, DK), pullback: the ChainRulesCore.rrule a Julia binding attaches
    to assemble_matrices.  pullback(Sbar, DKbar) -> dict of cotangents."""
    S, DK = assemble_matrices(green_functions, mesh, wavenumber, direct, device)

    def pullback(Sbar, DKbar=None):
        out = assemble_matrices_vjp(green_functions, mesh, [wavenumber], Sbar, DKbar, direct, device)
        out["wavenumber"] = float(out.pop("wavenumbers")[0])
        return out

    return (S, DK), pullback


def solve_vjp(D, S, bc, phibar, direct=True, device=0):
    """Pullback of solve(D, S, bc; direct) (src/matrix_assembly.jl:105-133,
    the implicit_linear rrule): returns (Dbar, Sbar, bcbar) for the
    cotangent phibar of phi, in Zygote's convention for complex arrays."""
    torch = _torch()
    Dm = _as_colmajor_batch(torch, D, device)
    Sm = _as_colmajor_batch(torch, S, device)
    n = Dm.shape[1]
    b = bc if torch.is_tensor(bc) else torch.from_numpy(np.asarray(bc, dtype=np.complex128))
    pb = phibar if torch.is_tensor(phibar) else torch.from_numpy(np.asarray(phibar, dtype=np.complex128))
    b = b.to(device=f"cuda:{device}", dtype=torch.complex128)
    pb = pb.to(device=f"cuda:{device}", dtype=torch.complex128)
    vec = b.dim() == 1
    if vec:
        b, pb = b.reshape(-1, 1), pb.reshape(-1, 1)
    if Dm.shape[0] != 1 or Sm.shape != Dm.shape or b.shape[0] != n or pb.shape != b.shape:
        raise ValueError("solve_vjp: D, S must be n x n, bc and phibar of the same shape")
    nrhs = b.shape[1]
    bm = b.transpose(0, 1).contiguous()
    pm = pb.transpose(0, 1).contiguous()
    ipiv = torch.empty(n, dtype=torch.int64, device=f"cuda:{device}")
    ctx = L.context(device)
    L.check(L.lib().mh_lu_factor(ctx, n, 1, Dm.data_ptr(), n, n * n, ipiv.data_ptr()), "solve_vjp: lu")
    Db = torch.empty_like(Dm)
    Sb = torch.empty_like(Sm)
    bb = torch.empty_like(bm)
    L.check(L.lib().mh_solve_vjp(ctx, n, 1, int(bool(direct)), Dm.data_ptr(), ipiv.data_ptr(), Sm.data_ptr(), n,
                                 n * n, bm.data_ptr(), nrhs, pm.data_ptr(), n, n * nrhs, Db.data_ptr(),
                                 Sb.data_ptr(), bb.data_ptr()), "solve_vjp")
    Dbar = Db[0].transpose(0, 1)
    Sbar = Sb[0].transpose(0, 1)
    bcbar = bb.transpose(0, 1)
    return Dbar, Sbar, (bcbar[:, 0] if vec else bcbar)


def rrule_solve(D, S, bc, direct=True, device=0):
    """phi, pullback: the rrule of solve (ImplicitAD.implicit_linear's
    adjoint).  pullback(phibar) -> (Dbar, Sbar, bcbar)."""
    phi = solve(D, S, bc, direct, device)

    def pullback(phibar):
        return solve_vjp(D, S, bc, phibar, direct, device)

    return phi, pullback
