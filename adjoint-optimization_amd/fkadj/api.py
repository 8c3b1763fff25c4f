"""Host-side mirror of the reference's operator interface over libfkadj.

* ``fk_grad_Et_mex`` -- the MATLAB gateway ``[gradE,Lcost,L0,LE,param_str] =
  fk_grad_Et_mex(Et,dt,nt,y0,active_index,nrows,ncols,ghost_lr,gamma,S,alpha)``
  (src/cuda/fk_grad_Et_param_mex.cpp:11-92, called at src/cuda/simple_grad_desc.m:82-83),
  same argument order and meaning, errors raised as exceptions instead of aborting.
* ``Instance`` -- device-resident problem (geometry, objective, state) with gradient and forward
  calls; ``batch_grad`` evaluates many FP32 instances in shared launches.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import FkadjError, check  # noqa: F401


def _csc(S):
    """(colptr int64, rowind int64, val float64) from a scipy sparse matrix or a 3-tuple."""
    if isinstance(S, tuple):
        colptr, rowind, val = S
    else:
        S = S.tocsc()
        S.sort_indices()
        colptr, rowind, val = S.indptr, S.indices, S.data
    return (np.ascontiguousarray(colptr, dtype=np.int64), np.ascontiguousarray(rowind, dtype=np.int64),
            np.ascontiguousarray(val, dtype=np.float64))


def fk_grad_Et_mex(Et, dt, nt, y0, active_index, nrows, ncols, ghost_lr, gamma, S, alpha):
    """Reference-shaped gradient call (FP64 GPU path).  Returns (gradE, Lcost, L0, LE, param_str)."""
    Et = L._f64(np.ravel(Et))
    y0 = L._f64(np.ravel(y0))
    ai = np.ascontiguousarray(np.ravel(active_index), dtype=np.int32)
    gh = L._f64(np.ravel(ghost_lr))
    gamma = L._f64(np.ravel(gamma))
    colptr, rowind, val = _csc(S)
    nt = int(nt)
    if Et.size != nt + 1:
        raise ValueError("Et must have nt+1 samples")
    n_num = y0.size // 3
    gradE = np.zeros(nt + 1)
    L0 = C.c_double(); LE = C.c_double(); Lc = C.c_double()
    buf = C.create_string_buffer(2048)
    rc = L.lib().fkadj_grad_ref(float(dt), L._d(Et), nt, L._d(y0), ai.ctypes.data_as(C.POINTER(C.c_int32)),
                                int(nrows), int(ncols), n_num, L._d(gh), L._d(gamma),
                                colptr.ctypes.data_as(C.POINTER(C.c_int64)),
                                rowind.ctypes.data_as(C.POINTER(C.c_int64)), L._d(val), float(alpha),
                                C.byref(L0), C.byref(LE), C.byref(Lc), L._d(gradE), buf, len(buf))
    check(rc, "fkadj_grad_ref")
    return gradE, Lc.value, L0.value, LE.value, buf.value.decode()


def snapshot_count(nt: int, nstep: int) -> int:
    """Snapshots a forward run of nt steps keeps every nstep steps (the reference's compute_ns,
    src/cuda/fk_forward_Et_param_mex.cpp:7-25): nt/nstep - 1 when nstep divides nt, else floor."""
    n = L.lib().fkadj_snapshot_count(int(nt), int(nstep))
    if n < 0:
        check(n, "fkadj_snapshot_count")
    return n


def fk_forward_Et_mex(Et, dt, nt, y0, active_index, nrows, ncols, ghost_lr, nstep):
    """Reference-shaped forward call (FP64 GPU path), src/cuda/fk_forward_Et_param_mex.cpp:40-121.
    Returns (ys, yf, param_str): ys is (3*nrows*ncols, ns) with snapshot ks in column ks, yf is
    (3*nrows*ncols,); each state is the u, v and w planes (cell k = i*ncols + j), NaN at holes."""
    Et = L._f64(np.ravel(Et))
    y0 = L._f64(np.ravel(y0))
    ai = np.ascontiguousarray(np.ravel(active_index), dtype=np.int32)
    gh = L._f64(np.ravel(ghost_lr))
    nt, nstep = int(nt), int(nstep)
    if Et.size != nt + 1:
        raise ValueError("Et must have nt+1 samples")
    ns = snapshot_count(nt, nstep)
    n = int(nrows) * int(ncols)
    ys = np.zeros((max(ns, 1), 3 * n))
    yf = np.zeros(3 * n)
    buf = C.create_string_buffer(2048)
    rc = L.lib().fkadj_forward_ref(float(dt), L._d(Et), nt, L._d(y0), ai.ctypes.data_as(C.POINTER(C.c_int32)),
                                   int(nrows), int(ncols), y0.size // 3, L._d(gh), nstep, ns, L._d(ys), L._d(yf),
                                   buf, len(buf))
    check(rc, "fkadj_forward_ref")
    return ys[:ns].T.copy(), yf, buf.value.decode()


class Instance:
    """A device-resident gradient problem (one grid, one objective, one resident state).

    keep_bytes (FP32): device bytes for keeping the last checkpoint segments of the forward
    trajectory, which the backward pass then does not recompute (same results, fewer launches).
    fwd_depth, adj_depth (FP32): time steps per launch of the tiled kernels (1 or 2; 0 = the
    library's rule); roll_rows (FP32): strip rows of the row-rolling kernels (a multiple of 8; 0 =
    the library setting).  None of them changes a result."""

    def __init__(self, active_index, nrows, ncols, *, dt, nt, model=None, ghost_lr=None, pulse_mode=L.PULSE_KNOTS,
                 n_amp=0, alpha=0.5, gamma=(1.0, 0.2, 0.2), S=None, s_diag=0.0, precision=L.F32, ckpt=50,
                 fwd_tile=0, adj_tile=0, lazy_workspace=False, keep_bytes=0, fwd_depth=0, adj_depth=0,
                 roll_rows=0):
        self._keep = []
        ai = np.ascontiguousarray(np.ravel(active_index), dtype=np.int32)
        p = L.Problem()
        p.nrows, p.ncols = int(nrows), int(ncols)
        p.active_index = ai.ctypes.data_as(C.POINTER(C.c_int32))
        self._keep.append(ai)
        if ghost_lr is not None:
            gh = L._f64(np.ravel(ghost_lr))
            p.gh_num = L._d(gh)
            self._keep.append(gh)
        p.model = model if model is not None else L.reference_model()
        p.dt, p.nt = float(dt), int(nt)
        p.pulse_mode, p.n_amp = int(pulse_mode), int(n_amp)
        p.alpha = float(alpha)
        for v in range(3):
            p.gamma[v] = float(gamma[v])
        if S is not None:
            colptr, rowind, val = _csc(S)
            self._keep += [colptr, rowind, val]
            p.S_colptr = colptr.ctypes.data_as(C.POINTER(C.c_int64))
            p.S_rowind = rowind.ctypes.data_as(C.POINTER(C.c_int64))
            p.S_val = L._d(val)
        p.s_diag = float(s_diag)
        p.precision, p.ckpt, p.fwd_tile, p.adj_tile = int(precision), int(ckpt), int(fwd_tile), int(adj_tile)
        p.lazy_workspace = 1 if lazy_workspace else 0
        p.keep_bytes = int(keep_bytes)
        p.fwd_depth, p.adj_depth, p.roll_rows = int(fwd_depth), int(adj_depth), int(roll_rows)
        self.problem = p
        self.nrows, self.ncols, self.nt, self.dt = p.nrows, p.ncols, p.nt, p.dt
        self.precision = p.precision
        h = C.c_void_p()
        check(L.lib().fkadj_create(C.byref(p), C.byref(h)), "fkadj_create")
        self._keep = []
        self.h = h
        self.n_active = L.lib().fkadj_n_active(h)
        self.n_controls = L.lib().fkadj_n_controls(h)
        # device bytes one server request on this instance holds while admitted
        self.request_bytes = int(L.lib().fkadj_request_bytes(h))

    def close(self):
        """Frees the instance; raises FkadjError (ERR_BUSY) while a server request or a batch plan
        still uses it (the instance then stays valid)."""
        if getattr(self, "h", None):
            check(L.lib().fkadj_destroy(self.h), "fkadj_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_state(self, y_packed):
        y = L._f64(np.ravel(y_packed))
        if y.size != 3 * self.n_active:
            raise ValueError("state must have 3*n_active values")
        check(L.lib().fkadj_set_state(self.h, L._d(y)), "fkadj_set_state")

    def get_state(self):
        y = np.zeros(3 * self.n_active)
        check(L.lib().fkadj_get_state(self.h, L._d(y)), "fkadj_get_state")
        return y

    def grad(self, E):
        """Returns (grad, Lcost, L0, LE)."""
        E = L._f64(np.ravel(E))
        if E.size != self.n_controls:
            raise ValueError(f"E must have {self.n_controls} values")
        g = np.zeros(self.n_controls)
        Lc, L0, LE = C.c_double(), C.c_double(), C.c_double()
        check(L.lib().fkadj_grad(self.h, L._d(E), L._d(g), C.byref(Lc), C.byref(L0), C.byref(LE)), "fkadj_grad")
        return g, Lc.value, L0.value, LE.value

    def grad_noalloc(self, E):
        """``grad`` through fkadj_grad_noalloc: the evaluation buffers live in the library's static
        device array (no device allocation); same results.  Returns (grad, Lcost, L0, LE)."""
        E = L._f64(np.ravel(E))
        if E.size != self.n_controls:
            raise ValueError(f"E must have {self.n_controls} values")
        g = np.zeros(self.n_controls)
        Lc, L0, LE = C.c_double(), C.c_double(), C.c_double()
        check(L.lib().fkadj_grad_noalloc(self.h, L._d(E), L._d(g), C.byref(Lc), C.byref(L0), C.byref(LE)),
              "fkadj_grad_noalloc")
        return g, Lc.value, L0.value, LE.value

    def forward(self, E, nsteps=None, snap_every=0, advance=False):
        """Forward solve; returns (snapshots (m, 3*n_active), final packed state)."""
        E = L._f64(np.ravel(E))
        nsteps = self.nt if nsteps is None else int(nsteps)
        nsnap = snapshot_count(nsteps, snap_every) if snap_every > 0 else 0
        snaps = np.zeros((max(nsnap, 1), 3 * self.n_active))
        yf = np.zeros(3 * self.n_active)
        check(L.lib().fkadj_forward(self.h, L._d(E), nsteps, int(snap_every), L._d(snaps), L._d(yf),
                                    1 if advance else 0), "fkadj_forward")
        return snaps[:nsnap], yf

    def stimulate(self, cell_mask, value=1.0):
        m = np.ascontiguousarray(np.ravel(cell_mask), dtype=np.uint8)
        check(L.lib().fkadj_stimulate(self.h, m.ctypes.data_as(C.POINTER(C.c_uint8)), float(value)),
              "fkadj_stimulate")


def _fields(instances, Es):
    """One contiguous float64 field per instance, each of the instance's n_controls values (the C side
    reads exactly that many from every pointer)."""
    if len(Es) != len(instances):
        raise ValueError(f"need {len(instances)} fields, got {len(Es)}")
    out = [L._f64(np.ravel(E)) for E in Es]
    for k, (inst, E) in enumerate(zip(instances, out)):
        if E.size != inst.n_controls:
            raise ValueError(f"field {k} has {E.size} values, its instance needs {inst.n_controls}")
    return out


def batch_grad(instances, Es):
    """Gradients of several FP32 instances in shared launches; returns a list of (grad, Lcost, L0, LE)."""
    n = len(instances)
    Es = _fields(instances, Es)
    grads = [np.zeros(inst.n_controls) for inst in instances]
    hs = (C.c_void_p * n)(*[inst.h for inst in instances])
    ep = (C.POINTER(C.c_double) * n)(*[L._d(E) for E in Es])
    gp = (C.POINTER(C.c_double) * n)(*[L._d(g) for g in grads])
    Lc, L0, LE = np.zeros(n), np.zeros(n), np.zeros(n)
    check(L.lib().fkadj_batch_grad(n, hs, ep, gp, L._d(Lc), L._d(L0), L._d(LE)), "fkadj_batch_grad")
    return [(grads[k], Lc[k], L0[k], LE[k]) for k in range(n)]


class Plan:
    """A fixed batch of FP32 instances with one captured HIP graph of a whole evaluation
    (fkadj_plan_*); ``grad(Es)`` replays it for new fields and returns what ``batch_grad`` returns."""

    def __init__(self, instances):
        self.instances = list(instances)
        n = len(self.instances)
        hs = (C.c_void_p * n)(*[inst.h for inst in self.instances])
        h = C.c_void_p()
        check(L.lib().fkadj_plan_create(n, hs, C.byref(h)), "fkadj_plan_create")
        self.h = h.value

    def grad(self, Es):
        n = len(self.instances)
        Es = _fields(self.instances, Es)
        grads = [np.zeros(inst.n_controls) for inst in self.instances]
        ep = (C.POINTER(C.c_double) * n)(*[L._d(E) for E in Es])
        gp = (C.POINTER(C.c_double) * n)(*[L._d(g) for g in grads])
        Lc, L0, LE = np.zeros(n), np.zeros(n), np.zeros(n)
        check(L.lib().fkadj_plan_grad(self.h, ep, gp, L._d(Lc), L._d(L0), L._d(LE)), "fkadj_plan_grad")
        return [(grads[k], Lc[k], L0[k], LE[k]) for k in range(n)]

    def close(self):
        if getattr(self, "h", None):
            L.lib().fkadj_plan_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()


def batch_forward(instances, Es, nsteps):
    """Advance the resident states of several FP32 instances by nsteps in shared launches."""
    n = len(instances)
    Es = _fields(instances, Es)
    hs = (C.c_void_p * n)(*[inst.h for inst in instances])
    ep = (C.POINTER(C.c_double) * n)(*[L._d(E) for E in Es])
    check(L.lib().fkadj_batch_forward(n, hs, ep, int(nsteps)), "fkadj_batch_forward")


class Server:
    """Continuous-batching gradient server (C ABI ``fkadj_server_*``)."""

    def __init__(self, budget_bytes: int, max_slots: int = 64):
        h = C.c_void_p()
        check(L.lib().fkadj_server_create(int(budget_bytes), int(max_slots), C.byref(h)), "fkadj_server_create")
        self.h = h
        self._n = {}

    def submit(self, inst: Instance, E) -> int:
        E = L._f64(np.ravel(E))
        if E.size != inst.n_controls:
            raise ValueError(f"E must have {inst.n_controls} values")
        rid = C.c_int64()
        check(L.lib().fkadj_server_submit(self.h, inst.h, L._d(E), C.byref(rid)), "fkadj_server_submit")
        self._n[rid.value] = inst.n_controls
        return rid.value

    def cancel(self, rid: int):
        check(L.lib().fkadj_server_cancel(self.h, int(rid)), "fkadj_server_cancel")

    def run(self, max_rounds: int = 1) -> int:
        n = L.lib().fkadj_server_run(self.h, int(max_rounds))
        if n < 0:
            check(n, "fkadj_server_run")
        return n

    def poll(self, rid: int):
        """None while running; (grad, Lcost, L0, LE) when done; raises FkadjError on failure."""
        g = np.zeros(self._n.get(rid, 1))
        Lc, L0, LE = C.c_double(), C.c_double(), C.c_double()
        rc = L.lib().fkadj_server_poll(self.h, int(rid), L._d(g), C.byref(Lc), C.byref(L0), C.byref(LE))
        if rc == L.ERR_BUSY:
            return None
        self._n.pop(rid, None)
        check(rc, "fkadj_server_poll")
        return g, Lc.value, L0.value, LE.value

    def pending(self) -> int:
        return int(L.lib().fkadj_server_pending(self.h))

    def close(self):
        if getattr(self, "h", None):
            L.lib().fkadj_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def footprint_bytes(problem: L.Problem) -> int:
    return int(L.lib().fkadj_footprint_bytes(C.byref(problem)))


def mem_info() -> tuple[int, int]:
    """(free, total) device bytes of the current GPU."""
    fr, tot = C.c_int64(), C.c_int64()
    check(L.lib().fkadj_mem_info(C.byref(fr), C.byref(tot)), "fkadj_mem_info")
    return int(fr.value), int(tot.value)
