"""ctypes binding of libmhip.so (C ABI declared in include/mhip.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C marinehydro.jl_amd/csrc``).  There is no fallback: if the library
is missing or no gfx950 device is present every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmhip.so")

MH_OK = 0
_ERRORS = {
    1: ValueError,          # MH_ERR_ARGUMENT
    2: RuntimeError,        # MH_ERR_HIP
    3: ArithmeticError,     # MH_ERR_SINGULAR
    4: NotImplementedError, # MH_ERR_NOT_SUPPORTED
    5: RuntimeError,        # MH_ERR_NO_DEVICE
    6: ArithmeticError,     # MH_ERR_NOT_CONVERGED
}

# Green-function term bits, include/mhip.h
GF_RANKINE = 1
GF_RANKINE_REFLECTED = 2
GF_RANKINE_REFLECTED_NEG = 4
GF_WU = 8
GF_DELHOMMEAU = 16
GF_DELHOMMEAU_TABLE = 32

DIRECT, INDIRECT, BOTH = 1, 2, 3
OP_GREENS, OP_GRADIENT_GREENS, OP_INTEGRAL, OP_INTEGRAL_GRADIENT = 0, 1, 2, 3
FP64, MIXED = 0, 1

_vp = C.c_void_p
_i64 = C.c_int64
_d = C.c_double
_int = C.c_int

_SIGS = {
    "mh_version": (_int, []),
    "mh_last_error": (C.c_char_p, []),
    "mh_context_create": (_int, [_int, _vp, C.POINTER(_vp)]),
    "mh_context_destroy": (_int, [_vp]),
    "mh_synchronize": (_int, [_vp]),
    "mh_profile_enable": (_int, [_vp, _int]),
    "mh_profile_report": (_int, [_vp, C.c_char_p, _i64]),
    "mh_set_refinement": (_int, [_vp, _int, _int]),
    "mh_fallback_count": (_int, [_vp, C.POINTER(_i64)]),
    "mh_set_deferred_check": (_int, [_vp, _int]),
    "mh_deferred_rerun_count": (_int, [_vp, C.POINTER(_i64)]),
    "mh_set_sweep_chunk": (_int, [_vp, _i64]),
    "mh_gf_eval": (_int, [_vp, _int, _int, _int, _i64, _vp, _vp, _d, _vp]),
    "mh_wave_term_eval": (_int, [_vp, _int, _i64, _vp, _vp, _vp]),
    "mh_assemble": (_int, [_vp, _vp, _int, _int, _i64, _vp, _vp, _vp, _vp, _i64, _i64]),
    "mh_assemble_noalloc": (_int, [_vp, _vp, _int, _int, _i64, _vp, _vp, _vp, _vp, _i64, _i64]),
    "mh_lu_factor": (_int, [_vp, _i64, _i64, _vp, _i64, _i64, _vp]),
    "mh_lu_solve": (_int, [_vp, _i64, _i64, _vp, _i64, _i64, _vp, _int, _i64, _vp, _i64, _i64]),
    "mh_solve": (_int, [_vp, _i64, _i64, _int, _vp, _vp, _i64, _i64, _vp, _i64, _vp, _i64, _i64]),
    "mh_freq_sweep": (_int, [_vp, _vp, _int, _int, _int, _i64, _vp, _d, _d, _i64, _vp, _i64, _vp,
                             _vp, _vp, _vp]),
    "mh_freq_sweep_grad": (_int, [_vp, _vp, _int, _int, _int, _i64, _vp, _d, _d, _i64, _vp, _i64,
                                  _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                  _vp, _vp]),
    "mh_assemble_vjp": (_int, [_vp, _vp, _int, _int, _i64, _vp, _vp, _vp, _i64, _i64, _vp, _vp, _vp,
                               _vp, _vp]),
    "mh_solve_vjp": (_int, [_vp, _i64, _i64, _int, _vp, _vp, _vp, _i64, _i64, _vp, _i64, _vp, _i64,
                            _i64, _vp, _vp, _vp]),
    "mh_froude_krylov": (_int, [_vp, _vp, _i64, _vp, _d, _d, _i64, _vp, _i64, _vp, _vp]),
    "mh_froude_krylov_grad": (_int, [_vp, _vp, _i64, _vp, _d, _d, _i64, _vp, _i64, _vp, _i64, _vp,
                                     _vp, _vp, _vp, _vp, _vp, _vp]),
    "mh_multi_sweep": (_int, [_vp, _i64, _vp, _int, _d, _d, _int]),
    "mh_solve_mixed": (_int, [_vp, _i64, _i64, _int, _vp, _vp, _i64, _i64, _vp, _i64, _vp, _i64, _i64]),
    "mh_sweep_graph_create": (_int, [_vp, _vp, _int, _int, _i64, _i64, _vp, _i64, _d, _d, _vp, _vp, _vp, _vp]),
    "mh_sweep_graph_launch": (_int, [_vp, _vp, _vp, _vp]),
    "mh_sweep_graph_destroy": (_int, [_vp]),
    "mh_batch_graph_create": (_int, [_vp, _i64, _vp, _int, _d, _d, _int, _vp]),
    "mh_batch_graph_launch": (_int, [_vp, _vp, _i64, _vp]),
    "mh_batch_graph_info": (_int, [_vp, _vp, _vp, _vp]),
    "mh_batch_graph_dot": (_int, [_vp, C.c_char_p]),
    "mh_batch_graph_destroy": (_int, [_vp]),
    "mh_set_far_field_fp32": (_int, [_vp, _int]),
    "mh_set_gradient_mode": (_int, [_vp, _int]),
    "mh_multi_autotune": (_int, [_vp, _i64, _vp, _int, _d, _d, C.POINTER(_int), C.POINTER(_int),
                                 C.POINTER(_d)]),
    "mh_calculate_radiation_forces_host": (_int, [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                                  _d, _d, _d, _vp]),
    "mh_diffraction_force_host": (_int, [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _d,
                                         _d, _vp]),
}


class MeshT(C.Structure):
    """mh_mesh_t"""
    _fields_ = [("nfaces", _i64), ("nvertices", _i64), ("vertices", _vp), ("faces", _vp),
                ("centers", _vp), ("normals", _vp), ("areas", _vp), ("radii", _vp)]


class SweepItemT(C.Structure):
    """mh_sweep_item_t"""
    _fields_ = [("mesh", _vp), ("gf_mask", _int), ("direct", _int), ("nomega", _i64), ("omegas", _vp),
                ("ndof", _i64), ("dof_fields", _vp), ("nbeta", _i64), ("betas", _vp), ("nparam", _i64),
                ("dvertices", _vp), ("dcenters", _vp), ("dnormals", _vp), ("dareas", _vp),
                ("ddof_fields", _vp), ("domegas", _vp), ("A", _vp), ("B", _vp), ("F", _vp), ("dA", _vp),
                ("dB", _vp), ("dF", _vp)]


_lib = None
_lock = threading.Lock()


def lib():
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise RuntimeError(
                        f"libmhip.so not found at {LIB_PATH}; build it with "
                        "`python -c 'import __graft_entry__ as g; g.build()'` or "
                        "`make -C marinehydro.jl_amd/csrc`")
                L = C.CDLL(LIB_PATH)
                for name, (res, args) in _SIGS.items():
                    f = getattr(L, name)
                    f.restype = res
                    f.argtypes = args
                _lib = L
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != MH_OK:
        msg = lib().mh_last_error().decode(errors="replace")
        raise _ERRORS.get(rc, RuntimeError)(f"{what}: {msg}" if what else msg)


_contexts: dict = {}


def destroy_contexts() -> None:
    """Releases every library context (streams, events, device workspace,
    pinned host memory) while the HIP runtime is still alive.  Registered
    with atexit on first use, so the handles never outlive the runtime or a
    profiler attached to it; safe to call more than once."""
    if _lib is None:
        return
    while _contexts:
        _, h = _contexts.popitem()
        _lib.mh_context_destroy(h)


_atexit_registered = False


def context(device: int = 0):
    """Library context bound to `device` and torch's current stream there."""
    import torch

    stream = torch.cuda.current_stream(device).cuda_stream
    key = (device, stream)
    ctx = _contexts.get(key)
    if ctx is None:
        global _atexit_registered
        if not _atexit_registered:
            import atexit

            # registered after torch is imported, so it runs before torch's
            # own exit handlers (atexit runs last-in, first-out)
            atexit.register(destroy_contexts)
            _atexit_registered = True
        h = _vp()
        check(lib().mh_context_create(device, _vp(stream) if stream else None, C.byref(h)),
              "mh_context_create")
        _contexts[key] = h
        ctx = h
    return ctx


def profile_enable(on=True, device=0):
    check(lib().mh_profile_enable(context(device), int(bool(on))), "profile_enable")


def profile_report(device=0) -> dict:
    """Per-kernel launches, milliseconds, algorithmic flops and bytes since
    the last report (kernel timer on the context's stream)."""
    import json

    buf = C.create_string_buffer(1 << 16)
    check(lib().mh_profile_report(context(device), buf, len(buf)), "profile_report")
    return json.loads(buf.value.decode())


def set_refinement(max_sweeps=12, fp64_fallback=True, device=0):
    """Mixed precision: refinement sweeps allowed and whether a chunk that
    does not converge is solved again in FP64 (mh_set_refinement)."""
    check(lib().mh_set_refinement(context(device), int(max_sweeps), int(bool(fp64_fallback))),
          "set_refinement")


def set_far_field_fp32(on=True, device=0):
    """Rankine and image far-field pairs (beyond 7 panel radii) of every
    assembly on the device's context in FP32 (mh_set_far_field_fp32)."""
    check(lib().mh_set_far_field_fp32(context(device), int(bool(on))), "set_far_field_fp32")


def set_gradient_mode(mode="auto", device=0):
    """How the sweep gradients differentiate the wave-term pair sums
    (mh_set_gradient_mode): "auto", "forward" or "reverse"."""
    m = {"auto": 0, "forward": 1, "reverse": 2}[mode]
    check(lib().mh_set_gradient_mode(context(device), m), "set_gradient_mode")


def set_deferred_check(sweeps=2, device=0):
    """Mixed sweeps with gradients: refinement sweeps made before the
    convergence check, which then runs beside the gradient kernels (0: a
    check after every sweep) (mh_set_deferred_check)."""
    check(lib().mh_set_deferred_check(context(device), int(sweeps)), "set_deferred_check")


def set_sweep_chunk(max_frequencies=0, device=0):
    """Frequencies per chunk of the sweeps at most (0: as many as fit 32 GiB)
    (mh_set_sweep_chunk)."""
    check(lib().mh_set_sweep_chunk(context(device), int(max_frequencies)), "set_sweep_chunk")


def deferred_rerun_count(device=0) -> int:
    """Frequency chunks solved again after a failed deferred check."""
    out = _i64(0)
    check(lib().mh_deferred_rerun_count(context(device), C.byref(out)), "deferred_rerun_count")
    return int(out.value)


def fallback_count(device=0) -> int:
    """Frequency chunks solved again in FP64 since the context was created."""
    out = _i64(0)
    check(lib().mh_fallback_count(context(device), C.byref(out)), "fallback_count")
    return int(out.value)
