"""ctypes wrapper of the FP64 oracle (``fk_oracle.c``).

Test infrastructure only: imported by ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` -- never by the library.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
SRC = os.path.join(HERE, "fk_oracle.c")

MODEL_FIELDS = ("u_c", "u_v", "tau_d", "tau_v1_minus", "tau_v2_minus", "tau_v_plus", "tau_0", "tau_r",
                "tau_si", "u_csi", "k_fk", "tau_w_minus", "tau_w_plus", "Du", "Dv", "Dw", "ksm", "h")


class Model(C.Structure):
    _fields_ = [(f, C.c_double) for f in MODEL_FIELDS]


def build(force: bool = False) -> str:
    """Compile liboracle.so with gcc (no fast-math: the oracle relies on NaN semantics)."""
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(SRC):
        # x86-64-v3 (AVX2) runs on every host these builds meet; no FMA contraction and no reassociation,
        # so vector code gives the scalar results
        cmd = ["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared", "-o", LIB_PATH,
               SRC, "-lm"]
        subprocess.check_call(cmd)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        lp = C.POINTER(C.c_int64)
        mp = C.POINTER(Model)
        L.orc_model_reference.argtypes = [mp]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_max_threads.restype = C.c_int
        L.orc_fion.argtypes = [mp, C.c_double, C.c_double, C.c_double, dp]
        L.orc_jacT.argtypes = [mp, dp, dp]
        for f in (L.orc_lap, L.orc_lapT):
            f.argtypes = [mp, C.c_int, C.c_int, dp, dp]
        L.orc_rhs.argtypes = [mp, C.c_int, C.c_int, dp, dp, C.c_double, dp]
        L.orc_spline_set.argtypes = [C.c_int, dp, dp, dp]
        L.orc_spline_val.argtypes = [C.c_int, dp, dp, dp, C.c_double]
        L.orc_spline_val.restype = C.c_double
        L.orc_knot_tables.argtypes = [C.c_double, C.c_int, dp, dp, dp]
        L.orc_forward.argtypes = [mp, C.c_int, C.c_int, C.c_double, C.c_int, dp, dp, dp]
        L.orc_grad.argtypes = [mp, C.c_int, C.c_int, C.c_double, C.c_int, dp, dp, dp, dp, ip, C.c_int,
                               lp, lp, dp, dp, C.c_int, dp, dp, dp, C.c_int]
        L.orc_num2nan.argtypes = [dp, ip, C.c_int, C.c_int, dp]
        L.orc_grad_ref.argtypes = [C.c_double, dp, C.c_int, dp, ip, C.c_int, C.c_int, C.c_int, dp, dp,
                                   lp, lp, dp, C.c_double, mp, C.c_int, dp, dp, dp, dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def reference_model(h: float | None = None) -> Model:
    m = Model()
    lib().orc_model_reference(C.byref(m))
    if h is not None:
        m.h = h
    return m


def model_from_dict(d: dict) -> Model:
    m = reference_model()
    for k, v in d.items():
        setattr(m, k, float(v))
    return m


def fion(m: Model, u, v, w):
    out = np.zeros(3)
    lib().orc_fion(C.byref(m), u, v, w, _d(out))
    return out


def jacT(m: Model, y):
    y = _f64(y)
    out = np.zeros(9)
    lib().orc_jacT(C.byref(m), _d(y), _d(out))
    return out


def lap(m: Model, y: np.ndarray):
    """y: (3, nx, ny) NaN-padded planes."""
    y = _f64(y)
    out = np.empty_like(y)
    lib().orc_lap(C.byref(m), y.shape[1], y.shape[2], _d(y), _d(out))
    return out


def lapT(m: Model, z: np.ndarray):
    z = _f64(z)
    out = np.empty_like(z)
    lib().orc_lapT(C.byref(m), z.shape[1], z.shape[2], _d(z), _d(out))
    return out


def rhs(m: Model, y: np.ndarray, gh: np.ndarray, E: float):
    y = _f64(y); gh = _f64(gh)
    out = np.empty_like(y)
    lib().orc_rhs(C.byref(m), y.shape[1], y.shape[2], _d(y), _d(gh), E, _d(out))
    return out


def spline_set(t, y):
    t = _f64(t); y = _f64(y)
    ypp = np.zeros_like(y)
    rc = lib().orc_spline_set(len(t), _d(t), _d(y), _d(ypp))
    if rc:
        raise ValueError("spline_set failed")
    return ypp


def spline_val(t, y, ypp, tv):
    t = _f64(t); y = _f64(y); ypp = _f64(ypp)
    return lib().orc_spline_val(len(t), _d(t), _d(y), _d(ypp), float(tv))


def knot_tables(dt: float, Et):
    Et = _f64(Et)
    nt = len(Et) - 1
    Ef = np.zeros(3 * nt); Ea = np.zeros(3 * nt)
    rc = lib().orc_knot_tables(dt, nt, _d(Et), _d(Ef), _d(Ea))
    if rc:
        raise ValueError("knot tables failed")
    return Ef, Ea


def forward(m: Model, y_planes: np.ndarray, gh: np.ndarray, dt: float, Ef: np.ndarray):
    y = _f64(y_planes).copy()
    gh = _f64(gh); Ef = _f64(Ef)
    nt = len(Ef) // 3
    rc = lib().orc_forward(C.byref(m), y.shape[1], y.shape[2], dt, nt, _d(Ef), _d(gh), _d(y))
    if rc:
        raise RuntimeError("orc_forward failed")
    return y


def grad(m: Model, y_planes, gh, dt, Ef, Ea, active_index, S_csc, gamma, ckpt=50, round_state=False):
    """Generic gradient evaluation; returns (zgh[nt+1], L0, yT planes).

    round_state=True is a precision probe, not the reference's algorithm: the trajectory is rounded to
    FP32 after every forward and recompute step (FP64 arithmetic otherwise), which measures how much of
    an FP32 gradient's error the storage of the state alone accounts for."""
    y = _f64(y_planes); gh = _f64(gh); Ef = _f64(Ef); Ea = _f64(Ea)
    ai = np.ascontiguousarray(active_index, dtype=np.int32)
    colptr, rowind, val = (np.ascontiguousarray(S_csc[0], dtype=np.int64),
                           np.ascontiguousarray(S_csc[1], dtype=np.int64), _f64(S_csc[2]))
    gamma = _f64(gamma)
    nt = len(Ef) // 3
    zgh = np.zeros(nt + 1)
    L0 = np.zeros(1)
    yT = np.empty_like(y)
    n_num = int((ai > 0).sum())
    rc = lib().orc_grad(C.byref(m), y.shape[1], y.shape[2], dt, nt, _d(Ef), _d(Ea), _d(gh), _d(y),
                        ai.ctypes.data_as(C.POINTER(C.c_int32)), n_num,
                        colptr.ctypes.data_as(C.POINTER(C.c_int64)), rowind.ctypes.data_as(C.POINTER(C.c_int64)),
                        _d(val), _d(gamma), int(ckpt), _d(zgh), _d(L0), _d(yT), int(bool(round_state)))
    if rc:
        raise RuntimeError(f"orc_grad failed ({rc})")
    return zgh, float(L0[0]), yT


def grad_ref(dt, Et, nt, y_ini, active_index, nrows, ncols, gh_num, gamma, S_csc, alpha, model=None, ckpt=50):
    """Mirror of fk_grad_Et_comp: returns (gradE, Lcost, L0, LE)."""
    Et = _f64(Et); y_ini = _f64(y_ini); gh_num = _f64(gh_num); gamma = _f64(gamma)
    ai = np.ascontiguousarray(active_index, dtype=np.int32)
    colptr, rowind, val = (np.ascontiguousarray(S_csc[0], dtype=np.int64),
                           np.ascontiguousarray(S_csc[1], dtype=np.int64), _f64(S_csc[2]))
    n_num = len(y_ini) // 3
    gradE = np.zeros(nt + 1)
    L0 = np.zeros(1); LE = np.zeros(1); Lc = np.zeros(1)
    mp = C.byref(model) if model is not None else None
    rc = lib().orc_grad_ref(dt, _d(Et), nt, _d(y_ini), ai.ctypes.data_as(C.POINTER(C.c_int32)), nrows, ncols,
                            n_num, _d(gh_num), _d(gamma), colptr.ctypes.data_as(C.POINTER(C.c_int64)),
                            rowind.ctypes.data_as(C.POINTER(C.c_int64)), _d(val), alpha, mp, int(ckpt),
                            _d(L0), _d(LE), _d(Lc), _d(gradE))
    if rc:
        raise RuntimeError(f"orc_grad_ref failed ({rc})")
    return gradE, float(Lc[0]), float(L0[0]), float(LE[0])


def num2nan(packed, active_index, nrows, ncols):
    packed = _f64(packed)
    ai = np.ascontiguousarray(active_index, dtype=np.int32)
    out = np.empty(3 * nrows * ncols)
    lib().orc_num2nan(_d(packed), ai.ctypes.data_as(C.POINTER(C.c_int32)), nrows * ncols, len(packed) // 3,
                      _d(out))
    return out.reshape(3, nrows, ncols)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def max_threads() -> int:
    return int(lib().orc_max_threads())
