"""ctypes front end of the FP64 CPU oracle (oracle/mh_oracle.c).

TEST INFRASTRUCTURE ONLY.  The oracle is a plain-C restatement of the
reference algorithm (MarineHydro.jl); the HIP library is tested against it.
Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module; the product package never does.

Array conventions follow the reference's Julia arrays: mesh fields are
(n, 3) arrays stored column-major, complex matrices are column-major
complex128.  Any object with the attributes ``vertices, faces, centers,
normals, areas, radii`` (faces 0-based, as in reference src/meshes.jl:179)
is accepted as a mesh.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# MH_ORACLE_LIB selects another build of the same sources (the sanitizer
# build of tests/test_sanitize.py)
LIB_PATH = os.environ.get("MH_ORACLE_LIB") or os.path.join(HERE, "build", "liborc.so")

RANKINE = 1
RANKINE_REFLECTED = 2
RANKINE_REFLECTED_NEG = 4
WU = 8
DELHOMMEAU = 16

_lib = None
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int64)


class _Panel(C.Structure):
    _fields_ = [
        ("center", C.c_double * 3),
        ("vertices", (C.c_double * 3) * 4),
        ("normal", C.c_double * 3),
        ("area", C.c_double),
        ("radius", C.c_double),
    ]


def build() -> None:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(_Panel)
        sig = {
            "orc_greens": (None, [C.c_int, _dp, _dp, C.c_double, _dp]),
            "orc_gradient_greens": (None, [C.c_int, _dp, _dp, C.c_double, C.c_int, _dp]),
            "orc_integral": (None, [C.c_int, _dp, P, C.c_double, _dp]),
            "orc_integral_gradient": (None, [C.c_int, _dp, P, C.c_double, C.c_int, _dp]),
            "orc_rankine_integral": (C.c_double, [_dp, P]),
            "orc_rankine_integral_gradient": (None, [_dp, P, C.c_int, _dp]),
            "orc_wu_dimless": (None, [C.c_double, C.c_double, _dp, _dp, _dp]),
            "orc_delhommeau_dimless": (None, [C.c_double, C.c_double, _dp, _dp, _dp]),
            "orc_exp_e1": (None, [C.c_double, C.c_double, _dp]),
            "orc_struve_h0": (C.c_double, [C.c_double]),
            "orc_struve_h1": (C.c_double, [C.c_double]),
            "orc_assemble": (None, [C.c_int, C.c_int64, _dp, C.c_int64, _ip, _dp, _dp, _dp, _dp,
                                    C.c_double, C.c_int, _dp, _dp, C.c_int]),
            "orc_zgetrf": (C.c_int, [C.c_int64, _dp, _ip]),
            "orc_zgetrs": (None, [C.c_int64, _dp, _ip, C.c_int, C.c_int64, _dp]),
            "orc_solve": (C.c_int, [C.c_int64, _dp, _dp, _dp, C.c_int64, C.c_int, _dp]),
            "orc_freq_solve": (C.c_int, [C.c_int, C.c_int, C.c_int64, _dp, C.c_int64, _ip, _dp, _dp,
                                         _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_int64,
                                         _dp, C.c_int64, _dp, _dp, _dp, C.c_int]),
            "orc_froude_krylov": (None, [C.c_int64, _dp, _dp, _dp, C.c_double, C.c_double,
                                         C.c_double, C.c_int64, _dp, C.c_int64, _dp, _dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _v3(x):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(3))
    return a, a.ctypes.data_as(_dp)


def make_panel(center, vertices, normal, area, radius) -> _Panel:
    p = _Panel()
    for c in range(3):
        p.center[c] = float(center[c])
        p.normal[c] = float(normal[c])
    V = np.asarray(vertices, dtype=np.float64).reshape(4, 3)
    for v in range(4):
        for c in range(3):
            p.vertices[v][c] = V[v, c]
    p.area = float(area)
    p.radius = float(radius)
    return p


def _cplx(buf):
    return complex(buf[0], buf[1])


def greens(mask, x, xi, k=1.0):
    o = (C.c_double * 2)()
    xa, xp = _v3(x)
    ya, yp = _v3(xi)
    lib().orc_greens(mask, xp, yp, k, o)
    return _cplx(o)


def gradient_greens(mask, x, xi, k=1.0, wrt_first=False):
    o = (C.c_double * 6)()
    xa, xp = _v3(x)
    ya, yp = _v3(xi)
    lib().orc_gradient_greens(mask, xp, yp, k, int(wrt_first), o)
    return np.array([complex(o[0], o[1]), complex(o[2], o[3]), complex(o[4], o[5])])


def integral(mask, x, panel: _Panel, k=1.0):
    o = (C.c_double * 2)()
    xa, xp = _v3(x)
    lib().orc_integral(mask, xp, C.byref(panel), k, o)
    return _cplx(o)


def integral_gradient(mask, x, panel: _Panel, k=1.0, wrt_first=False):
    o = (C.c_double * 6)()
    xa, xp = _v3(x)
    lib().orc_integral_gradient(mask, xp, C.byref(panel), k, int(wrt_first), o)
    return np.array([complex(o[0], o[1]), complex(o[2], o[3]), complex(o[4], o[5])])


def rankine_integral(x, panel: _Panel):
    xa, xp = _v3(x)
    return lib().orc_rankine_integral(xp, C.byref(panel))


def rankine_integral_gradient(x, panel: _Panel, wrt_first=False):
    o = (C.c_double * 3)()
    xa, xp = _v3(x)
    lib().orc_rankine_integral_gradient(xp, C.byref(panel), int(wrt_first), o)
    return np.array(o[:])


def wu_dimless(h, v):
    """(G, dG/dh, dG/dv) of the Wu wave term in reduced coordinates."""
    a, b, c = (C.c_double * 2)(), (C.c_double * 2)(), (C.c_double * 2)()
    lib().orc_wu_dimless(h, v, a, b, c)
    return _cplx(a), _cplx(b), _cplx(c)


def delhommeau_dimless(r, z):
    """(G, dG/dr, dG/dz) of the exact Guevel-Delhommeau wave term."""
    a, b, c = (C.c_double * 2)(), (C.c_double * 2)(), (C.c_double * 2)()
    lib().orc_delhommeau_dimless(r, z, a, b, c)
    return _cplx(a), _cplx(b), _cplx(c)


def exp_e1(z: complex) -> complex:
    o = (C.c_double * 2)()
    lib().orc_exp_e1(z.real, z.imag, o)
    return _cplx(o)


def struve_h0(x):
    return lib().orc_struve_h0(float(x))


def struve_h1(x):
    return lib().orc_struve_h1(float(x))


def _mesh_args(mesh):
    V = np.asfortranarray(mesh.vertices, dtype=np.float64)
    F = np.asfortranarray(mesh.faces, dtype=np.int64)
    Cc = np.asfortranarray(mesh.centers, dtype=np.float64)
    N = np.asfortranarray(mesh.normals, dtype=np.float64)
    A = np.ascontiguousarray(mesh.areas, dtype=np.float64)
    R = np.ascontiguousarray(mesh.radii, dtype=np.float64)
    keep = (V, F, Cc, N, A, R)
    nf, nv = F.shape[0], V.shape[0]
    args = (nf, V.ctypes.data_as(_dp), nv, F.ctypes.data_as(_ip), Cc.ctypes.data_as(_dp),
            N.ctypes.data_as(_dp), A.ctypes.data_as(_dp), R.ctypes.data_as(_dp))
    return keep, args


def assemble(mask, mesh, k, direct=True, nthreads=0):
    """(S, D) for direct=True, (S, K) otherwise; column-major complex128."""
    keep, args = _mesh_args(mesh)
    nf = args[0]
    S = np.zeros((nf, nf), dtype=np.complex128, order="F")
    DK = np.zeros((nf, nf), dtype=np.complex128, order="F")
    lib().orc_assemble(mask, *args, float(k), int(bool(direct)), S.ctypes.data_as(_dp),
                       DK.ctypes.data_as(_dp), int(nthreads))
    return S, DK


def zgetrf(A):
    LU = np.array(A, dtype=np.complex128, order="F", copy=True)
    n = LU.shape[0]
    ipiv = np.zeros(n, dtype=np.int64)
    info = lib().orc_zgetrf(n, LU.ctypes.data_as(_dp), ipiv.ctypes.data_as(_ip))
    return LU, ipiv, info


def zgetrs(LU, ipiv, B, trans=False):
    X = np.array(B, dtype=np.complex128, order="F", copy=True)
    if X.ndim == 1:
        X = X.reshape(-1, 1, order="F")
    n, nrhs = X.shape
    lib().orc_zgetrs(n, np.asfortranarray(LU).ctypes.data_as(_dp), ipiv.ctypes.data_as(_ip),
                     int(trans), nrhs, X.ctypes.data_as(_dp))
    return X


def solve(DK, S, bc, direct=True):
    """reference src/matrix_assembly.jl:105-133."""
    A = np.array(DK, dtype=np.complex128, order="F", copy=True)
    Sf = np.asfortranarray(S, dtype=np.complex128)
    b = np.array(bc, dtype=np.complex128, order="F", copy=True)
    vec = b.ndim == 1
    if vec:
        b = b.reshape(-1, 1, order="F")
    n, nrhs = b.shape
    phi = np.zeros((n, nrhs), dtype=np.complex128, order="F")
    lib().orc_solve(n, A.ctypes.data_as(_dp), Sf.ctypes.data_as(_dp), b.ctypes.data_as(_dp),
                    nrhs, int(bool(direct)), phi.ctypes.data_as(_dp))
    return phi[:, 0] if vec else phi


def freq_solve(mask, mesh, omega, dof_fields, betas=(), direct=True, g=9.81, rho=1000.0,
               nthreads=0):
    """One frequency end to end.  dof_fields: list of (nf, 3) arrays.

    Returns (A[ndof, ndof], B[ndof, ndof], F[ndof, nbeta]) with A[i, j] the
    force on influenced dof i from radiating dof j.
    """
    keep, args = _mesh_args(mesh)
    nf = args[0]
    ndof = len(dof_fields)
    D = np.zeros((ndof, nf, 3), dtype=np.float64)
    for d, f in enumerate(dof_fields):
        D[d] = np.broadcast_to(np.asarray(f, dtype=np.float64), (nf, 3))
    # each dof field column-major nf x 3
    Dcm = np.ascontiguousarray(np.transpose(D, (0, 2, 1)))
    bet = np.ascontiguousarray(np.asarray(betas, dtype=np.float64).reshape(-1))
    AB = np.zeros(2 * ndof * ndof)
    F = np.zeros(2 * ndof * max(len(bet), 1))
    info = lib().orc_freq_solve(mask, int(bool(direct)), *args, float(omega), float(g), float(rho),
                                ndof, Dcm.ctypes.data_as(_dp), len(bet), bet.ctypes.data_as(_dp),
                                AB.ctypes.data_as(_dp), F.ctypes.data_as(_dp), int(nthreads))
    if info:
        raise RuntimeError(f"oracle LU: zero pivot at {info}")
    A = AB[: ndof * ndof].reshape(ndof, ndof, order="F")
    B = AB[ndof * ndof:].reshape(ndof, ndof, order="F")
    Fc = (F[0::2] + 1j * F[1::2])[: ndof * len(bet)].reshape(ndof, len(bet), order="F")
    return A, B, Fc


def froude_krylov(mesh, omega, dof_fields, betas=(0.0,), g=9.81, rho=1000.0):
    """Froude-Krylov force (reference src/waves.jl:62-66, :150-164).

    Returns FK[ndof, nbeta] (complex), influenced dof by heading."""
    nf = mesh.nfaces
    Cc = np.asfortranarray(mesh.centers, dtype=np.float64)
    N = np.asfortranarray(mesh.normals, dtype=np.float64)
    A = np.ascontiguousarray(mesh.areas, dtype=np.float64)
    ndof = len(dof_fields)
    D = np.zeros((ndof, nf, 3), dtype=np.float64)
    for d, f in enumerate(dof_fields):
        D[d] = np.broadcast_to(np.asarray(f, dtype=np.float64), (nf, 3))
    Dcm = np.ascontiguousarray(np.transpose(D, (0, 2, 1)))
    bet = np.ascontiguousarray(np.asarray(betas, dtype=np.float64).reshape(-1))
    out = np.zeros(2 * ndof * max(len(bet), 1))
    lib().orc_froude_krylov(nf, Cc.ctypes.data_as(_dp), N.ctypes.data_as(_dp), A.ctypes.data_as(_dp),
                            float(omega), float(g), float(rho), ndof, Dcm.ctypes.data_as(_dp),
                            len(bet), bet.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    return (out[0::2] + 1j * out[1::2])[: ndof * len(bet)].reshape(ndof, len(bet), order="F")
