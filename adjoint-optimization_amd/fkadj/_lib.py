"""ctypes binding of libfkadj.so (the C ABI in include/fkadj.h).

The library is the product; this module only marshals numpy arrays.  Loading fails loudly when
the shared library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libfkadj.so")

OK = 0
ERR_ARG, ERR_GEOMETRY, ERR_NOMEM, ERR_DEVICE, ERR_NONFINITE, ERR_STATE, ERR_CANCELLED, ERR_BUSY, ERR_NOTFOUND = \
    -1, -2, -3, -4, -5, -6, -7, -8, -9
ERR_SCRATCH = -10
NOALLOC_SCRATCH_BYTES = 256 << 20   # FKADJ_NOALLOC_SCRATCH_BYTES
F64, F32 = 0, 1
PULSE_KNOTS, PULSE_PIECEWISE = 0, 1

MODEL_FIELDS = ("u_c", "u_v", "tau_d", "tau_v1_minus", "tau_v2_minus", "tau_v_plus", "tau_0", "tau_r",
                "tau_si", "u_csi", "k_fk", "tau_w_minus", "tau_w_plus", "Du", "Dv", "Dw", "ksm", "h")


class Model(C.Structure):
    _fields_ = [(f, C.c_double) for f in MODEL_FIELDS]


class Problem(C.Structure):
    _fields_ = [
        ("nrows", C.c_int), ("ncols", C.c_int),
        ("active_index", C.POINTER(C.c_int32)),
        ("gh_num", C.POINTER(C.c_double)),
        ("model", Model),
        ("dt", C.c_double), ("nt", C.c_int),
        ("pulse_mode", C.c_int), ("n_amp", C.c_int),
        ("alpha", C.c_double), ("gamma", C.c_double * 3),
        ("S_colptr", C.POINTER(C.c_int64)), ("S_rowind", C.POINTER(C.c_int64)), ("S_val", C.POINTER(C.c_double)),
        ("s_diag", C.c_double),
        ("precision", C.c_int), ("ckpt", C.c_int), ("fwd_tile", C.c_int), ("adj_tile", C.c_int),
        ("lazy_workspace", C.c_int),
        ("keep_bytes", C.c_int64),
        ("fwd_depth", C.c_int), ("adj_depth", C.c_int), ("roll_rows", C.c_int),
    ]


class Timing(C.Structure):
    _fields_ = [("eval_ms", C.c_double), ("fwd_kernel_ms", C.c_double), ("adj_kernel_ms", C.c_double),
                ("fwd_launches", C.c_int64), ("adj_launches", C.c_int64), ("cell_steps", C.c_int64)]


class FkadjError(RuntimeError):
    def __init__(self, code: int, where: str):
        self.code = code
        super().__init__(f"{where}: {strerror(code)} ({code})")


_lib = None


def lib():
    """Load libfkadj.so (built by ``__graft_entry__.build()``); raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() (hipcc, gfx950)")
    L = C.CDLL(LIB_PATH)
    dp, ip32, ip64 = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int64)
    vp = C.c_void_p
    L.fkadj_strerror.argtypes = [C.c_int]
    L.fkadj_strerror.restype = C.c_char_p
    L.fkadj_model_reference.argtypes = [C.POINTER(Model)]
    L.fkadj_model_reference.restype = None
    L.fkadj_param_string.argtypes = [C.POINTER(Model), C.c_char_p, C.c_size_t]
    L.fkadj_grad_ref.argtypes = [C.c_double, dp, C.c_int, dp, ip32, C.c_int, C.c_int, C.c_int, dp, dp,
                                 ip64, ip64, dp, C.c_double, dp, dp, dp, dp, C.c_char_p, C.c_size_t]
    L.fkadj_create.argtypes = [C.POINTER(Problem), C.POINTER(vp)]
    L.fkadj_destroy.argtypes = [vp]
    L.fkadj_footprint_bytes.argtypes = [C.POINTER(Problem)]
    L.fkadj_footprint_bytes.restype = C.c_int64
    L.fkadj_request_bytes.argtypes = [vp]
    L.fkadj_request_bytes.restype = C.c_int64
    L.fkadj_mem_info.argtypes = [C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.fkadj_set_state.argtypes = [vp, dp]
    L.fkadj_get_state.argtypes = [vp, dp]
    L.fkadj_n_active.argtypes = [vp]
    L.fkadj_n_controls.argtypes = [vp]
    L.fkadj_grad.argtypes = [vp, dp, dp, dp, dp, dp]
    L.fkadj_grad_noalloc.argtypes = [vp, dp, dp, dp, dp, dp]
    L.fkadj_forward.argtypes = [vp, dp, C.c_int, C.c_int, dp, dp, C.c_int]
    L.fkadj_snapshot_count.argtypes = [C.c_int, C.c_int]
    L.fkadj_forward_ref.argtypes = [C.c_double, dp, C.c_int, dp, ip32, C.c_int, C.c_int, C.c_int, dp, C.c_int,
                                    C.c_int, dp, dp, C.c_char_p, C.c_size_t]
    L.fkadj_stimulate.argtypes = [vp, C.POINTER(C.c_uint8), C.c_double]
    L.fkadj_batch_grad.argtypes = [C.c_int, C.POINTER(vp), C.POINTER(dp), C.POINTER(dp), dp, dp, dp]
    L.fkadj_batch_forward.argtypes = [C.c_int, C.POINTER(vp), C.POINTER(dp), C.c_int]
    L.fkadj_plan_create.argtypes = [C.c_int, C.POINTER(vp), C.POINTER(vp)]
    L.fkadj_plan_grad.argtypes = [vp, C.POINTER(dp), C.POINTER(dp), dp, dp, dp]
    L.fkadj_plan_destroy.argtypes = [vp]
    L.fkadj_server_create.argtypes = [C.c_int64, C.c_int, C.POINTER(vp)]
    L.fkadj_server_destroy.argtypes = [vp]
    L.fkadj_server_submit.argtypes = [vp, vp, dp, C.POINTER(C.c_int64)]
    L.fkadj_server_cancel.argtypes = [vp, C.c_int64]
    L.fkadj_server_run.argtypes = [vp, C.c_int]
    L.fkadj_server_poll.argtypes = [vp, C.c_int64, dp, dp, dp, dp]
    L.fkadj_server_pending.argtypes = [vp]
    L.fkadj_set_profiling.argtypes = [C.c_int]
    L.fkadj_set_batch_groups.argtypes = [C.c_int]
    L.fkadj_set_rolling.argtypes = [C.c_int, C.c_int]
    L.fkadj_last_timing.argtypes = [C.POINTER(Timing)]
    _lib = L
    return L


def strerror(code: int) -> str:
    try:
        return lib().fkadj_strerror(int(code)).decode()
    except Exception:
        return "error"


def check(rc: int, where: str):
    if rc != OK:
        raise FkadjError(rc, where)


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def reference_model(h: float | None = None) -> Model:
    m = Model()
    lib().fkadj_model_reference(C.byref(m))
    if h is not None:
        m.h = float(h)
    return m


def param_string(model: Model | None = None) -> str:
    m = model if model is not None else reference_model()
    buf = C.create_string_buffer(2048)
    lib().fkadj_param_string(C.byref(m), buf, len(buf))
    return buf.value.decode()


def set_profiling(on: bool):
    lib().fkadj_set_profiling(1 if on else 0)


def set_batch_groups(n: int) -> int:
    """Groups (host threads and streams) a batch call of 2M cells or more runs as; returns the
    previous setting."""
    rc = lib().fkadj_set_batch_groups(int(n))
    check(min(rc, 0), "fkadj_set_batch_groups")
    return rc


def set_rolling(mode: int, strip_rows: int = 0) -> int:
    """Row-rolling kernels for large one-pitch FP32 groups: 0 never, 1 from 2M cells (default), 2
    whenever possible; strip_rows a multiple of 8, or 0 (default) for strip rows chosen per launch.
    Returns the previous mode."""
    rc = lib().fkadj_set_rolling(int(mode), int(strip_rows))
    check(min(rc, 0), "fkadj_set_rolling")
    return rc


def last_timing() -> dict:
    t = Timing()
    check(lib().fkadj_last_timing(C.byref(t)), "fkadj_last_timing")
    return {f: getattr(t, f) for f, _ in Timing._fields_}
