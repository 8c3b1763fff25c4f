"""C-ABI checks that need no GPU: the library loads, exports every declared entry point, and
argument / geometry validation returns error codes before any device work."""
import json
import os
import re

import numpy as np
import pytest

import fkadj
from fkadj import _lib as L
from fkadj import geometry as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(fkadj_\w+)\s*\(", src))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = fkadj.lib()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_reference_model_and_param_string():
    import oracle as O
    m = fkadj.reference_model()
    om = O.reference_model()
    for f in L.MODEL_FIELDS:
        assert getattr(m, f) == getattr(om, f), f
    s = fkadj.param_string(m)
    for f in L.MODEL_FIELDS:
        assert f"\n{f} " in s
    assert s.endswith("\n") and "\0" not in s


def test_strerror_codes():
    for c in range(0, -10, -1):
        assert fkadj.strerror(c)
    assert "geometry" in fkadj.strerror(L.ERR_GEOMETRY)


def _expect(code, **kw):
    base = dict(dt=0.1, nt=10)
    base.update(kw)
    ai = base.pop("ai")
    shape = base.pop("shape")
    with pytest.raises(fkadj.FkadjError) as e:
        fkadj.Instance(ai, *shape, **base)
    assert e.value.code == code, (e.value.code, code)


def test_validation_errors_before_device_work():
    mask = np.ones((16, 16), bool)
    ai = G.active_index_from_mask(mask)
    # a cell with both axial neighbours missing (the reference's mirror rule gives NaN there)
    bad = mask.copy()
    bad[5, 4] = bad[5, 6] = False
    _expect(L.ERR_GEOMETRY, ai=G.active_index_from_mask(bad), shape=(16, 16))
    dup = ai.copy(); dup[3] = dup[4]
    _expect(L.ERR_GEOMETRY, ai=dup, shape=(16, 16))
    _expect(L.ERR_ARG, ai=ai, shape=(16, 16), pulse_mode=L.PULSE_PIECEWISE, n_amp=3)      # 10 % 3 != 0
    _expect(L.ERR_ARG, ai=ai, shape=(16, 16), precision=7)
    _expect(L.ERR_ARG, ai=ai, shape=(16, 16), nt=2)                                       # spline needs 4 knots
    _expect(L.ERR_ARG, ai=ai, shape=(16, 16), ghost_lr=np.full(256, 0.5))
    _expect(L.ERR_ARG, ai=ai, shape=(3, 16), dt=0.1)
    _expect(L.ERR_ARG, ai=ai, shape=(16, 16), dt=-1.0)
    _expect(L.ERR_ARG, ai=ai, shape=(16, 16), keep_bytes=-1)


def test_footprint_grows_with_size():
    p = L.Problem()
    p.nrows = p.ncols = 256
    p.nt, p.ckpt, p.precision = 1000, 50, L.F32
    a = fkadj.footprint_bytes(p)
    p.nrows = p.ncols = 512
    b = fkadj.footprint_bytes(p)
    p.precision = L.F64
    c = fkadj.footprint_bytes(p)
    assert 0 < a < b and c > 0
    # 512^2 FP32, 10^4 steps, K=50: 201 checkpoints + 51 segment states + 2x50 stored slopes of 3 MiB
    p.precision, p.nt = L.F32, 10000
    assert 0.9e9 < fkadj.footprint_bytes(p) < 1.4e9


def test_footprint_one_segment_buffer_from_2M_cells():
    """FP32 instances below 2M cells keep two segment buffers (their backward pass may overlap a
    recompute with an adjoint), larger ones one: one row fewer than 2M cells costs K+1 states and K
    slopes more (101 FP32 states of 1023 x 2048 cells = 2.54 GB at K = 50)."""
    p = L.Problem()
    p.nt, p.ckpt, p.precision = 1000, 50, L.F32
    p.nrows, p.ncols = 1024, 2048
    one = fkadj.footprint_bytes(p)
    p.nrows = 1023
    two = fkadj.footprint_bytes(p)
    state = 3 * 4 * 1023 * 2048
    assert 100 * state < two - one < 102 * state


def test_footprint_counts_kept_segments():
    """Each kept segment holds K+1 states and K slopes of its own; a budget buys whole segments, at
    most every segment of the run (FP64 keeps none)."""
    p = L.Problem()
    p.nrows, p.ncols = 256, 200            # pitch 256
    p.nt, p.ckpt, p.precision = 1030, 50, L.F32   # 21 segments, the last of 30 steps
    base = fkadj.footprint_bytes(p)
    seg = (2 * 50 + 1) * 3 * 4 * 256 * 256
    for budget, kept in ((seg - 1, 0), (seg, 1), (5 * seg + seg // 2, 5), (21 * seg, 21), (1 << 50, 21)):
        p.keep_bytes = budget
        assert fkadj.footprint_bytes(p) - base == kept * seg, (budget, kept)
    p.precision = L.F64
    with_budget = fkadj.footprint_bytes(p)
    p.keep_bytes = 0
    assert fkadj.footprint_bytes(p) == with_budget


def test_batch_groups_setting():
    """The split-batch setting takes 1..8 and returns the previous value (no device work)."""
    prev = fkadj.set_batch_groups(3)
    try:
        assert 1 <= prev <= 8
        assert fkadj.set_batch_groups(1) == 3
        for bad in (0, 9, -1):
            with pytest.raises(fkadj.FkadjError) as e:
                fkadj.set_batch_groups(bad)
            assert e.value.code == L.ERR_ARG
        assert fkadj.set_batch_groups(prev) == 1
    finally:
        fkadj.set_batch_groups(prev)


def test_disc_holes_have_no_isthmus_and_closed_form_ghost():
    m = G.disc_holes(96, 128, 25, 2, 6, seed=7)
    assert not G.isthmus_cells(m).any()
    gh = G.ghost_lr_from_mask(m)
    assert set(np.unique(gh)) <= {-1.0, 0.0, 1.0}
    L_ = G.lap_matrix(m)
    assert np.allclose(L_.sum(axis=1), 0.0)   # mirror Laplacian annihilates constants


def test_plan_and_server_argument_errors_before_device_work():
    lib = fkadj.lib()
    import ctypes as C
    h = C.c_void_p()
    assert lib.fkadj_plan_create(0, None, C.byref(h)) == L.ERR_ARG
    assert lib.fkadj_plan_create(65, None, C.byref(h)) == L.ERR_ARG
    assert lib.fkadj_plan_grad(None, None, None, None, None, None) == L.ERR_ARG
    assert lib.fkadj_plan_destroy(None) == L.ERR_ARG
    assert lib.fkadj_server_create(0, 4, C.byref(h)) == L.ERR_ARG
    assert lib.fkadj_server_create(1 << 30, 65, C.byref(h)) == L.ERR_ARG
    assert lib.fkadj_server_create(1 << 30, 0, C.byref(h)) == L.ERR_ARG
    rid = C.c_int64()
    assert lib.fkadj_server_submit(None, None, None, C.byref(rid)) == L.ERR_ARG
    assert lib.fkadj_server_run(None, 0) == L.ERR_ARG
    assert lib.fkadj_server_pending(None) == L.ERR_ARG


def test_launch_shape_fields_checked_before_device_work():
    """fwd_depth and adj_depth outside 0..2 and roll_rows that is not a multiple of 8 in 0..256 are
    refused with FKADJ_ERR_ARG before any device work (and before the geometry is read)."""
    import ctypes as C
    ai = np.ones(4, np.int32)   # never read: the argument checks come first
    for field, bad in (("fwd_depth", 3), ("fwd_depth", -1), ("adj_depth", 3), ("adj_depth", -2),
                       ("roll_rows", 12), ("roll_rows", 264), ("roll_rows", -8)):
        p = L.Problem()
        p.nrows, p.ncols, p.nt, p.dt, p.precision = 64, 64, 10, 0.1, L.F32
        p.pulse_mode = L.PULSE_KNOTS
        p.active_index = ai.ctypes.data_as(C.POINTER(C.c_int32))
        setattr(p, field, bad)
        h = C.c_void_p()
        assert fkadj.lib().fkadj_create(C.byref(p), C.byref(h)) == L.ERR_ARG, (field, bad)


def test_grid_too_large_for_32_bit_offsets():
    p = L.Problem()
    p.nrows, p.ncols, p.nt, p.dt, p.precision = 15000, 15000, 10, 0.1, L.F32
    import ctypes as C
    ai = np.ones(4, np.int32)   # never read: the size check comes first
    p.active_index = ai.ctypes.data_as(C.POINTER(C.c_int32))
    h = C.c_void_p()
    assert fkadj.lib().fkadj_create(C.byref(p), C.byref(h)) == L.ERR_ARG


def test_tuned_table_lookup(tmp_path, monkeypatch):
    """The autotuner's table gives Instance keywords by the nearest tuned grid size and launch scale;
    the library's rules (no keywords) where nothing is tuned for that scale."""
    from fkadj import autotune
    t = tmp_path / "t.json"
    t.write_text(json.dumps({
        "256": {"single": {"fwd_tile": 6, "fwd_depth": 2, "adj_tile": 6, "adj_depth": 1, "ms": 1.0},
                "batch": {"roll_rows": 16, "fwd_tile": 4, "adj_tile": 4, "batch": 64, "ms": 2.0}},
        "1024": {"single": {"fwd_tile": 4, "fwd_depth": 1, "adj_tile": 4, "adj_depth": 1, "ms": 3.0},
                 "batch": {"roll_rows": 0, "fwd_tile": 4, "adj_tile": 4, "batch": 4, "ms": 4.0}},
        "2048": {"single": {"fwd_tile": 4, "fwd_depth": 1, "adj_tile": 6, "adj_depth": 2, "ms": 5.0}},
        "512": {"batch": {"choice": "roll", "roll_rows": 0, "fwd_tile": 4, "adj_tile": 4, "batch": 16, "ms": 6.0}},
        "768": {"batch": {"choice": "tiled", "roll_rows": 0, "fwd_tile": 4, "adj_tile": 4, "batch": 7, "ms": 7.0}}}))
    assert autotune.instance_kwargs(300, False, str(t)) == {"fwd_tile": 6, "fwd_depth": 2, "adj_tile": 6,
                                                            "adj_depth": 1}
    assert autotune.instance_kwargs(256, True, str(t)) == {"roll_rows": 16}
    assert autotune.instance_kwargs(900, True, str(t)) == {"fwd_tile": 4, "adj_tile": 4}
    assert autotune.instance_kwargs(4096, True, str(t)) == {"fwd_tile": 4, "adj_tile": 4}   # nearest batch entry
    assert autotune.instance_kwargs(2048, False, str(t))["adj_depth"] == 2
    assert autotune.instance_kwargs(512, True, str(t)) == {}   # rolling, strip rows chosen per launch
    assert autotune.instance_kwargs(760, True, str(t)) == {"fwd_tile": 4, "adj_tile": 4}
    monkeypatch.setattr(autotune, "load_table", lambda path=None: {})
    assert autotune.instance_kwargs(512, True) == {} and autotune.instance_kwargs(512, False) == {}


def test_forward_ref_argument_errors_before_device_work():
    """fkadj_forward_ref rejects a snapshot interval below 1, more snapshots than nt/nstep and
    missing arrays with FKADJ_ERR_ARG before it touches the device."""
    import ctypes as C
    lib = fkadj.lib()
    nr = nc = 8
    ai = np.arange(1, nr * nc + 1, dtype=np.int32)
    y0 = np.zeros(3 * nr * nc)
    gh = np.zeros(nr * nc)
    Et = np.zeros(11)
    ys = np.zeros(3 * nr * nc * 4)
    yf = np.zeros(3 * nr * nc)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    aip = ai.ctypes.data_as(C.POINTER(C.c_int32))

    def call(nstep, ns, ys_ptr=dp(ys), yf_ptr=dp(yf)):
        return lib.fkadj_forward_ref(0.1, dp(Et), 10, dp(y0), aip, nr, nc, nr * nc, dp(gh), nstep, ns, ys_ptr,
                                     yf_ptr, None, 0)

    assert call(0, 0) == L.ERR_ARG
    assert call(3, 4) == L.ERR_ARG            # floor(10/3) = 3 snapshots at most
    assert call(3, 1, ys_ptr=None) == L.ERR_ARG
    assert call(3, 1, yf_ptr=None) == L.ERR_ARG


def test_drop_in_entries_reject_n_num_mismatch_before_reading_arrays():
    """fkadj_grad_ref and fkadj_forward_ref size gh_num and S by the caller's n_num.  An n_num off by
    one from the grid's active cells returns FKADJ_ERR_GEOMETRY before either array is read (the
    reference checks S's column count against n_num before using it, G:1759-1762): the arrays here
    hold exactly the caller's n_num values, so a read past them would be a host overrun."""
    import ctypes as C
    lib = fkadj.lib()
    nr = nc = 8
    mask = np.ones((nr, nc), bool)
    ai = G.active_index_from_mask(mask).astype(np.int32)
    n_act = int(mask.sum())
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int64))
    aip = ai.ctypes.data_as(C.POINTER(C.c_int32))
    Et = np.zeros(11)
    gamma = np.array([1.0, 0.0, 0.0])
    L0, LE, Lc = C.c_double(), C.c_double(), C.c_double()
    for n_num in (n_act - 1, n_act + 1):
        y0 = np.zeros(3 * n_num)
        gh = np.zeros(n_num)
        colptr = np.arange(n_num + 1, dtype=np.int64)
        rowind = np.arange(n_num, dtype=np.int64)
        val = np.ones(n_num)
        grad = np.zeros(11)
        rc = lib.fkadj_grad_ref(0.1, dp(Et), 10, dp(y0), aip, nr, nc, n_num, dp(gh), dp(gamma), ip(colptr),
                                ip(rowind), dp(val), 0.5, C.byref(L0), C.byref(LE), C.byref(Lc), dp(grad), None, 0)
        assert rc == L.ERR_GEOMETRY, (n_num, rc)
        ys = np.zeros(3 * nr * nc)
        yf = np.zeros(3 * nr * nc)
        rc = lib.fkadj_forward_ref(0.1, dp(Et), 10, dp(y0), aip, nr, nc, n_num, dp(gh), 5, 1, dp(ys), dp(yf),
                                   None, 0)
        assert rc == L.ERR_GEOMETRY, (n_num, rc)


def test_batch_calls_check_field_lengths():
    """batch_grad, Plan.grad and batch_forward reject a field whose length differs from its instance's
    n_controls (the C side reads that many values from each field) before calling the library."""
    from fkadj import api

    class Stub:
        def __init__(self, n):
            self.n_controls = n
            self.h = None

    insts = [Stub(4), Stub(6)]
    with pytest.raises(ValueError):
        api.batch_grad(insts, [np.zeros(4), np.zeros(5)])
    with pytest.raises(ValueError):
        api.batch_forward(insts, [np.zeros(3), np.zeros(6)], 1)
    with pytest.raises(ValueError):
        api.batch_grad(insts, [np.zeros(4)])
    plan = api.Plan.__new__(api.Plan)
    plan.instances = insts
    plan.h = None
    with pytest.raises(ValueError):
        plan.grad([np.zeros(4), np.zeros(7)])


def test_rolling_setting():
    """The row-rolling setting takes modes 0..2 and strip heights that are multiples of 8, 0 meaning
    chosen per launch (no device work), and returns the previous mode."""
    prev = fkadj.set_rolling(2, 32)
    try:
        assert prev in (0, 1, 2)
        assert fkadj.set_rolling(0, 8) == 2
        assert fkadj.set_rolling(1, 0) == 0
        for mode, hs in ((3, 16), (-1, 16), (1, 12), (1, -8), (1, 264)):
            with pytest.raises(fkadj.FkadjError) as e:
                fkadj.set_rolling(mode, hs)
            assert e.value.code == L.ERR_ARG
    finally:
        fkadj.set_rolling(prev, 0)
