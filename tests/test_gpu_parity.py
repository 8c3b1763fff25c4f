"""GPU parity of libfkadj against the FP64 oracle and the reference's golden data (through the C ABI).

Tolerances:
  * FP64 GPU path vs oracle: relative 1e-9 (different tanh/FMA rounding, amplified by the dynamics);
    vs the reference's golden gradE at 256^2 / 3000 steps: relative 1e-9.
  * FP32 path vs FP64: relative 1e-4 on J and on the gradient on the short parity cases and on a
    C2-type workload (the target of the design); on the reference's own 3000-step golden case the
    gradient error is bounded by 8 times the error of FP32 state storage alone
    (test_golden_256_fp32_error_vs_fp64).
  * batched vs alone, checkpoint interval, tile height: bit-identical (same kernels, fixed order).
"""
import numpy as np
import pytest

import fkadj
from fkadj import _lib as L
from fkadj import geometry as G

from cases import (crop_case, disc_case, golden, oracle_grad_knots, oracle_grad_knots_state_rounded,
                   oracle_grad_piecewise, relerr)

pytestmark = pytest.mark.gpu


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


@pytest.fixture(scope="module")
def crop():
    return crop_case(64, 96, 32, 64)


def _Et(nt, seed=0):
    rng = np.random.default_rng(seed)
    t = np.arange(nt + 1)
    return 1.5 + np.sin(t * 0.15 + rng.uniform(0, 3)) + 0.3 * rng.standard_normal(nt + 1) * 0.1


def test_fp64_ref_entry_matches_oracle(crop):
    _print_case("fp64_ref_crop", shape=crop.mask.shape, nt=60)
    d, meta = golden()
    nt, dt = 60, 0.1
    Et = _Et(nt)
    S = G.S_matrix(crop.mask)
    g_ref, Lc_ref, L0_ref, LE_ref = oracle_grad_knots(crop, dt, Et, meta["gamma"], meta["alpha"], S)
    g, Lc, L0, LE, pstr = fkadj.fk_grad_Et_mex(Et, dt, nt, crop.y0, crop.ai, crop.nrows, crop.ncols, crop.gh,
                                               meta["gamma"], S, meta["alpha"])
    assert relerr(g, g_ref) < 1e-9
    assert abs(Lc - Lc_ref) < 1e-9 * abs(Lc_ref)
    assert abs(L0 - L0_ref) < 1e-9 * abs(L0_ref)
    assert LE == pytest.approx(LE_ref, rel=1e-14)
    assert "ksm" in pstr


def test_fp32_matches_oracle_knots(crop):
    _print_case("fp32_knots_crop", shape=crop.mask.shape, nt=60)
    _, meta = golden()
    nt, dt = 60, 0.1
    Et = _Et(nt, 1)
    S = G.S_matrix(crop.mask)
    g_ref, Lc_ref, _, _ = oracle_grad_knots(crop, dt, Et, meta["gamma"], meta["alpha"], S)
    with fkadj.Instance(crop.ai, crop.nrows, crop.ncols, dt=dt, nt=nt, ghost_lr=crop.gh, alpha=meta["alpha"],
                        gamma=meta["gamma"], S=S, precision=fkadj.F32, ckpt=16) as inst:
        inst.set_state(crop.y0)
        g, Lc, _, _ = inst.grad(Et)
    assert relerr(g, g_ref) < 1e-4
    assert abs(Lc - Lc_ref) < 1e-4 * abs(Lc_ref)


@pytest.mark.parametrize("shape,holes,seed", [((48, 128), 10, 3), ((70, 100), 14, 5)])
def test_fp32_and_fp64_piecewise_disc_holes(shape, holes, seed):
    """Config-style problem: h=0.025, dt=0.02, mean(u^2) + lambda*energy, piecewise amplitudes,
    ragged grid (ncols not a multiple of 64, nrows not a multiple of the tile height)."""
    _print_case("piecewise_disc", shape=shape, holes=holes, seed=seed)
    case = disc_case(*shape, holes, seed)
    m = fkadj.reference_model(h=0.025)
    import oracle as O
    om = O.reference_model(h=0.025)
    nt, dt, K = 120, 0.02, 6
    A = 10.0 * np.random.default_rng(seed).uniform(0, 1, K)
    lam = 1e-3
    S = G.mean_u2_S(case.n_act)
    g_ref, Lc_ref, L0_ref, _ = oracle_grad_piecewise(case, dt, nt, A, (1, 0, 0), 2 * lam, S, om)
    for prec, tol in ((fkadj.F64, 1e-9), (fkadj.F32, 1e-4)):
        with fkadj.Instance(case.ai, case.nrows, case.ncols, dt=dt, nt=nt, model=m, pulse_mode=fkadj.PULSE_PIECEWISE,
                            n_amp=K, alpha=2 * lam, gamma=(1, 0, 0), s_diag=2.0 / case.n_act, precision=prec,
                            ckpt=25) as inst:
            inst.set_state(case.y0)
            g, Lc, L0, _ = inst.grad(A)
        assert relerr(g, g_ref) < tol, (prec, relerr(g, g_ref))
        assert abs(Lc - Lc_ref) < tol * abs(Lc_ref), (prec, Lc, Lc_ref)


def test_fp32_checkpoint_and_tile_invariance(crop):
    """The recomputation replays the forward kernels bit for bit, so K does not change the result;
    the tile configuration changes only the halo bookkeeping, not the per-cell arithmetic, and every
    adjoint tile writes the same z.gh partials (one per 64 x 8 block): every configuration gives the
    same gradient bit for bit."""
    _print_case("fp32_invariance", shape=crop.mask.shape)
    _, meta = golden()
    nt, dt = 40, 0.1
    Et = _Et(nt, 2)
    S = G.S_matrix(crop.mask)
    out = []
    cfgs = ((1, 1, 1), (7, 1, 1), (40, 1, 1), (13, 2, 1), (9, 3, 3), (40, 3, 1), (11, 4, 4), (40, 5, 5), (6, 4, 2),
            (8, 6, 6), (10, 7, 7), (40, 7, 1))
    for ck, ft, at in cfgs:
        with fkadj.Instance(crop.ai, crop.nrows, crop.ncols, dt=dt, nt=nt, ghost_lr=crop.gh, alpha=meta["alpha"],
                            gamma=meta["gamma"], S=S, precision=fkadj.F32, ckpt=ck, fwd_tile=ft,
                            adj_tile=at) as inst:
            inst.set_state(crop.y0)
            out.append(inst.grad(Et))
    for o, (ck, ft, at) in zip(out[1:], cfgs[1:]):
        assert o[1] == out[0][1]                                  # J: same per-cell arithmetic, same reduction
        assert np.array_equal(o[0], out[0][0]), (ck, ft, at)


def test_batch_equals_alone():
    """Uneven grids evaluated in shared launches give exactly the results of separate calls."""
    _print_case("batch_vs_alone", sizes="(64,64),(96,128),(40,192)")
    insts, Es, alone = [], [], []
    m = fkadj.reference_model(h=0.025)
    nt, dt, K = 60, 0.02, 6
    for k, (H, W) in enumerate([(64, 64), (96, 128), (40, 192)]):
        case = disc_case(H, W, 6 + 2 * k, 10 + k)
        inst = fkadj.Instance(case.ai, H, W, dt=dt, nt=nt, model=m, pulse_mode=fkadj.PULSE_PIECEWISE, n_amp=K,
                              alpha=2e-3, gamma=(1, 0, 0), s_diag=2.0 / case.n_act, precision=fkadj.F32, ckpt=20)
        inst.set_state(case.y0)
        A = np.random.default_rng(k).uniform(0, 10, K)
        insts.append(inst)
        Es.append(A)
        alone.append(inst.grad(A))
    mixed = fkadj.batch_grad(insts, Es)
    for a, b in zip(alone, mixed):
        assert np.array_equal(a[0], b[0]) and a[1] == b[1] and a[2] == b[2]
    for inst in insts:
        inst.close()


def test_large_instance_one_segment_buffer():
    """An instance of 2M cells or more keeps one segment state / slope buffer (its backward pass never
    overlaps a recompute with an adjoint): several segments through that one buffer give the
    single-segment result bit for bit, and so do two of five segments kept (one stream, one step per
    launch); FP32 stays within 1e-4 of FP64, and a served request holds less than the instance's
    footprint."""
    H, W = 1024, 2048
    _print_case("large_one_buffer", shape=(H, W))
    case = disc_case(H, W, 40, 21)
    m = fkadj.reference_model(h=0.025)
    nt, dt, K = 24, 0.02, 4
    A = np.random.default_rng(21).uniform(0, 10, K)
    out = {}
    seg5 = (2 * 5 + 1) * 3 * 4 * H * W
    for prec, ck, keep in ((fkadj.F32, 5, 0), (fkadj.F32, 24, 0), (fkadj.F32, 5, 2 * seg5), (fkadj.F64, 6, 0)):
        with fkadj.Instance(case.ai, H, W, dt=dt, nt=nt, model=m, pulse_mode=fkadj.PULSE_PIECEWISE, n_amp=K,
                            alpha=2e-3, gamma=(1, 0, 0), s_diag=2.0 / case.n_act, precision=prec, ckpt=ck,
                            keep_bytes=keep) as inst:
            inst.set_state(case.y0)
            out[(prec, ck, keep)] = inst.grad(A)
            if prec == fkadj.F32:
                assert inst.request_bytes < fkadj.footprint_bytes(inst.problem)
    a, b, r = out[(fkadj.F32, 5, 0)], out[(fkadj.F32, 24, 0)], out[(fkadj.F64, 6, 0)]
    kept = out[(fkadj.F32, 5, 2 * seg5)]
    assert np.array_equal(a[0], b[0]) and a[1] == b[1]
    assert np.array_equal(a[0], kept[0]) and a[1] == kept[1]
    assert relerr(a[0], r[0]) < 1e-4 and abs(a[1] - r[1]) < 1e-5 * abs(r[1])


def test_graph_plan_replays_equal_eager():
    """A captured batch plan replayed several times (new fields, a new resident state, eager calls
    and synchronisations in between) gives the eager batched results bit for bit every time."""
    _print_case("graph_plan", sizes="(64,64),(96,128),(40,192)")
    insts, E1, E2 = [], [], []
    m = fkadj.reference_model(h=0.025)
    nt, dt, K = 60, 0.02, 6
    cases = []
    for k, (H, W) in enumerate([(64, 64), (96, 128), (40, 192)]):
        case = disc_case(H, W, 6 + 2 * k, 10 + k)
        inst = fkadj.Instance(case.ai, H, W, dt=dt, nt=nt, model=m, pulse_mode=fkadj.PULSE_PIECEWISE, n_amp=K,
                              alpha=2e-3, gamma=(1, 0, 0), s_diag=2.0 / case.n_act, precision=fkadj.F32, ckpt=20)
        inst.set_state(case.y0)
        insts.append(inst)
        cases.append(case)
        E1.append(np.random.default_rng(k).uniform(0, 10, K))
        E2.append(np.random.default_rng(50 + k).uniform(0, 10, K))

    def same(a, b):
        return all(np.array_equal(x[0], y[0]) and x[1] == y[1] and x[2] == y[2] for x, y in zip(a, b))

    eager1, eager2 = fkadj.batch_grad(insts, E1), fkadj.batch_grad(insts, E2)
    with fkadj.Plan(insts) as plan:
        assert same(plan.grad(E1), eager1)
        insts[1].grad(E2[1])                                   # eager work between replays
        assert same(plan.grad(E2), eager2)
        assert same(plan.grad(E1), eager1)
        with pytest.raises(fkadj.FkadjError) as e:             # held by the plan
            insts[0].close()
        assert e.value.code == L.ERR_BUSY
        y_new = cases[0].y0.copy()
        y_new[: cases[0].n_act] *= 0.5                        # a new resident state is read at replay
        insts[0].set_state(y_new)
        eager3 = fkadj.batch_grad(insts, E1)
        assert same(plan.grad(E1), eager3)
        assert not same(eager3, eager1)
    for inst in insts:
        inst.close()


def test_split_batches_equal_alone():
    """A batch call of 2M cells or more runs as up to fkadj_set_batch_groups groups, each from its own
    host thread on its own stream: 1, 2 and 3 groups give each instance's own result bit for bit (one
    member keeps all of its segments, one some); a plan made meanwhile (three groups, one captured
    graph each, replayed side by side) agrees."""
    _print_case("split_batch", sizes="(1024,768),(512,1024),(1024,1024)")
    m = fkadj.reference_model(h=0.025)
    nt, dt, K, ck = 20, 0.02, 4, 10
    insts, Es = [], []
    for k, ((H, W), keep) in enumerate(zip([(1024, 768), (512, 1024), (1024, 1024)], (1 << 40, 0, 1))):
        case = disc_case(H, W, 20 + 4 * k, 80 + k)
        seg = (2 * ck + 1) * 3 * 4 * H * ((W + 63) // 64 * 64)
        inst = fkadj.Instance(case.ai, H, W, dt=dt, nt=nt, model=m, pulse_mode=fkadj.PULSE_PIECEWISE, n_amp=K,
                              alpha=2e-3, gamma=(1, 0, 0), s_diag=2.0 / case.n_act, precision=fkadj.F32, ckpt=ck,
                              keep_bytes=keep if keep != 1 else seg)
        inst.set_state(case.y0)
        insts.append(inst)
        Es.append(np.random.default_rng(k).uniform(0, 10, K))
    alone = [inst.grad(E) for inst, E in zip(insts, Es)]
    prev = fkadj.set_batch_groups(1)
    try:
        for n in (1, 2, 3):
            fkadj.set_batch_groups(n)
            for rep in range(2):
                got = fkadj.batch_grad(insts, Es)
                for a, g in zip(alone, got):
                    assert np.array_equal(a[0], g[0]) and a[1] == g[1] and a[2] == g[2], (n, rep)
        with fkadj.Plan(insts) as plan:
            for a, g in zip(alone, plan.grad(Es)):
                assert np.array_equal(a[0], g[0]) and a[1] == g[1]
    finally:
        fkadj.set_batch_groups(prev)
    for inst in insts:
        inst.close()


def test_split_plan_at_batch_scale():
    """Eight 512^2 instances (2.1M cells, so batch calls and plans run two groups): the plan's two
    captured graphs, replayed side by side three times with eager work between, the split batch call
    and each instance alone agree bit for bit."""
    _print_case("split_plan_512x8", nt=200, K=50)
    from fkadj import workloads as W
    insts, Es = [], []
    for k in range(8):
        mask = W.holes_mask(512, 80, 200 + k)
        inst = W.make_instance(mask, 4.0, 20, precision=fkadj.F32, ckpt=50)
        inst.set_state(W.rest_state(int(mask.sum())))
        s1 = np.zeros(mask.shape, bool)
        s1[:, :200 + 8 * k] = True
        inst.stimulate(s1, 1.0)
        insts.append(inst)
        Es.append(np.random.default_rng(k).uniform(0, 10, 20))
    alone = [inst.grad(E) for inst, E in zip(insts, Es)]
    prev = fkadj.set_batch_groups(2)
    try:
        batched = fkadj.batch_grad(insts, Es)
        with fkadj.Plan(insts) as plan:
            for rep in range(3):
                got = plan.grad(Es)
                for a, b, g in zip(alone, batched, got):
                    assert np.array_equal(a[0], g[0]) and a[1] == g[1], rep
                    assert np.array_equal(a[0], b[0]) and a[1] == b[1], rep
                insts[3].grad(Es[3])                           # eager work between replays
    finally:
        fkadj.set_batch_groups(prev)
    for inst in insts:
        inst.close()


def test_graph_plan_with_kept_segments():
    """A plan over instances keeping all, some and none of their segments replays the eager batched
    results, which equal each instance alone without kept segments, bit for bit."""
    _print_case("graph_plan_kept", sizes="(64,64),(96,128),(40,192)")
    m = fkadj.reference_model(h=0.025)
    nt, dt, K, ck = 60, 0.02, 6, 20
    insts, plain, Es = [], [], []
    for k, ((H, W), keep) in enumerate(zip([(64, 64), (96, 128), (40, 192)], (1 << 40, 1, 0))):
        case = disc_case(H, W, 6 + 2 * k, 60 + k)
        seg = (2 * ck + 1) * 3 * 4 * H * ((W + 63) // 64 * 64)
        for kb, out in ((keep * seg if keep < 1 << 40 else keep, insts), (0, plain)):
            inst = fkadj.Instance(case.ai, H, W, dt=dt, nt=nt, model=m, pulse_mode=fkadj.PULSE_PIECEWISE, n_amp=K,
                                  alpha=2e-3, gamma=(1, 0, 0), s_diag=2.0 / case.n_act, precision=fkadj.F32,
                                  ckpt=ck, keep_bytes=kb)
            inst.set_state(case.y0)
            out.append(inst)
        Es.append(np.random.default_rng(k).uniform(0, 10, K))
    alone = [inst.grad(E) for inst, E in zip(plain, Es)]
    eager = fkadj.batch_grad(insts, Es)
    with fkadj.Plan(insts) as plan:
        for rep in range(3):
            got = plan.grad(Es)
            for a, g, e in zip(alone, got, eager):
                assert np.array_equal(a[0], g[0]) and a[1] == g[1] and np.array_equal(a[0], e[0]), rep
            insts[0].grad(Es[0])                                # eager work between replays
    for inst in insts + plain:
        inst.close()


def test_forward_snapshots_match_oracle(crop):
    _print_case("forward_snapshots", shape=crop.mask.shape)
    import oracle as O
    nt, dt = 30, 0.1
    Et = _Et(nt, 3)
    Ef, _ = O.knot_tables(dt, Et)
    om = O.reference_model()
    with fkadj.Instance(crop.ai, crop.nrows, crop.ncols, dt=dt, nt=nt, ghost_lr=crop.gh, precision=fkadj.F64,
                        S=G.S_matrix(crop.mask)) as inst:
        inst.set_state(crop.y0)
        snaps, yf = inst.forward(Et, snap_every=10)
    assert snaps.shape[0] == 2   # steps 10 and 20; the final state is returned separately
    y = crop.planes()
    gh = crop.gh_plane(om)
    for k, s in enumerate((10, 20, 30)):
        yk = O.forward(om, y, gh, dt, Ef[:3 * s])
        packed = np.concatenate([yk[v][crop.mask] for v in range(3)])
        got = snaps[k] if k < 2 else yf
        assert relerr(got, packed) < 1e-12


def test_errors_fail_only_the_call(crop):
    """Bad inputs come back as error codes; the instance stays usable."""
    _print_case("errors")
    nt, dt = 20, 0.1
    S = G.S_matrix(crop.mask)
    with fkadj.Instance(crop.ai, crop.nrows, crop.ncols, dt=dt, nt=nt, ghost_lr=crop.gh, S=S,
                        precision=fkadj.F32) as inst:
        with pytest.raises(fkadj.FkadjError) as e:
            inst.grad(np.ones(nt + 1))
        assert e.value.code == -6                    # no state yet
        bad = crop.y0.copy(); bad[3] = np.nan
        with pytest.raises(fkadj.FkadjError) as e:
            inst.set_state(bad)
        assert e.value.code == -5
        inst.set_state(crop.y0)
        E = np.ones(nt + 1); E[4] = np.inf
        with pytest.raises(fkadj.FkadjError) as e:
            inst.grad(E)
        assert e.value.code == -5
        g1 = inst.grad(np.ones(nt + 1))[0]
        g2 = inst.grad(np.ones(nt + 1))[0]
        assert np.array_equal(g1, g2)


@pytest.mark.slow
def test_golden_256_iteration50_fp64():
    """The reference's own case: 256^2, 3000 steps, spiral state, iterate 50 -> stored gradE/Lcost."""
    _print_case("golden256_it50_fp64")
    d, meta = golden()
    mask = G.mask_from_active_index(d["active_index"], 256, 256)
    S = G.S_matrix(mask)
    g, Lc, L0, LE, _ = fkadj.fk_grad_Et_mex(d["Et_50"], meta["dt"], 3000, d["y0"], d["active_index"], 256, 256,
                                            d["ghost_lr"].astype(float), meta["gamma"], S, meta["alpha"])
    assert relerr(g, d["gradE_50"]) < 1e-9
    assert abs(Lc - meta["iterations"]["50"]["Lcost"]) < 1e-10 * meta["iterations"]["50"]["Lcost"]


@pytest.mark.slow
def test_golden_256_gradient_descent_iterations_1_to_3():
    """Three steps of the reference's driver (E <- E - ds*gradE, src/cuda/simple_grad_desc.m:103)
    reproduce the stored Lcost/L0/LE histories of iterations 1-3."""
    _print_case("golden256_gd3_fp64")
    d, meta = golden()
    mask = G.mask_from_active_index(d["active_index"], 256, 256)
    S = G.S_matrix(mask)
    with fkadj.Instance(d["active_index"], 256, 256, dt=meta["dt"], nt=3000, ghost_lr=d["ghost_lr"].astype(float),
                        alpha=meta["alpha"], gamma=meta["gamma"], S=S, precision=fkadj.F64) as inst:
        inst.set_state(d["y0"])
        E = d["Et_seed"].copy()
        for it in range(3):
            g, Lc, L0, LE = inst.grad(E)
            assert abs(Lc - d["Lcost_vec_50"][it]) < 1e-10 * d["Lcost_vec_50"][it]
            assert abs(L0 - d["L0_vec_50"][it]) < 1e-10 * d["L0_vec_50"][it]
            assert abs(LE - d["LE_vec_50"][it]) < 1e-12 * d["LE_vec_50"][it]
            E = E - meta["ds"] * g


@pytest.mark.slow
@pytest.mark.timeout(300)
def test_golden_256_fp32_error_vs_fp64():
    """FP32 fused path on the reference's own case (256^2, dt = 0.1 ms, 3000 steps, mostly one-cell
    holes).  J stays within 1e-5 of the FP64 golden value.  The gradient cannot meet 1e-4 here with any
    FP32 state: the oracle with FP64 arithmetic throughout, rounding only the stored state to FP32
    after each step, is already 1.5e-3 from the golden gradE (perturbing the initial state alone by
    FP32 rounding moves it by under 2e-4, so it is the rounding injected at every step at the steep
    fronts, with z.gh cancelling across one-cell holes, that the case amplifies).  The FP32 path's
    own error (7.8e-3) is held to at most 8 times that storage floor; the 1e-4 gradient bound is held
    on the C2-type workloads."""
    _print_case("golden256_it50_fp32")
    d, meta = golden()
    mask = G.mask_from_active_index(d["active_index"], 256, 256)
    S = G.S_matrix(mask)
    with fkadj.Instance(d["active_index"], 256, 256, dt=meta["dt"], nt=3000, ghost_lr=d["ghost_lr"].astype(float),
                        alpha=meta["alpha"], gamma=meta["gamma"], S=S, precision=fkadj.F32) as inst:
        inst.set_state(d["y0"])
        g, Lc, _, _ = inst.grad(d["Et_50"])
    e = relerr(g, d["gradE_50"])
    print(f"fp32 golden gradE rel err {e:.3e}, Lcost rel err {abs(Lc / meta['iterations']['50']['Lcost'] - 1):.3e}",
          flush=True)
    g_floor, _ = oracle_grad_knots_state_rounded(meta["dt"], d["Et_50"], d["y0"], d["active_index"], 256, 256,
                                                 d["ghost_lr"].astype(float), meta["gamma"], S, meta["alpha"])
    floor = relerr(g_floor, d["gradE_50"])
    print(f"FP32-state floor {floor:.3e}; FP32 path at {e / floor:.1f} x floor", flush=True)
    assert 1e-3 < floor < 3e-3
    assert e < 8 * floor
    assert abs(Lc - meta["iterations"]["50"]["Lcost"]) < 1e-5 * meta["iterations"]["50"]["Lcost"]


def test_golden_256_initial_state_conditioning():
    """The golden case's sensitivity to its initial state alone: the FP64 path from the initial state
    perturbed by FP32-sized relative noise (6e-8, two seeds) moves gradE by about 2e-4, far less than the
    FP32 path's 7.8e-3, so the FP32 error is not the initial-value conditioning of the problem but the
    rounding injected at every step (the FP32-state floor of test_golden_256_fp32_error_vs_fp64)."""
    _print_case("golden256_it50_conditioning")
    d, meta = golden()
    mask = G.mask_from_active_index(d["active_index"], 256, 256)
    S = G.S_matrix(mask)
    out = {}
    for prec in (fkadj.F64, fkadj.F32):
        with fkadj.Instance(d["active_index"], 256, 256, dt=meta["dt"], nt=3000, ghost_lr=d["ghost_lr"].astype(float),
                            alpha=meta["alpha"], gamma=meta["gamma"], S=S, precision=prec) as inst:
            inst.set_state(d["y0"])
            out[prec] = inst.grad(d["Et_50"])[0]
            if prec == fkadj.F64:
                spread = []
                for seed in (0, 1):
                    noise = np.random.default_rng(seed).standard_normal(d["y0"].size)
                    inst.set_state(d["y0"] * (1.0 + 6e-8 * noise))
                    spread.append(relerr(inst.grad(d["Et_50"])[0], out[prec]))
    e32 = relerr(out[fkadj.F32], out[fkadj.F64])
    print(f"initial-state spread {max(spread):.2e} vs FP32 error {e32:.2e}", flush=True)
    assert 0 < max(spread) < 0.1 * e32


@pytest.mark.slow
def test_fp32_c2_type_error_bound():
    """C2-type workload (h = 0.025 cm, dt = 0.02 ms, disc holes, S1-S2 spiral, piecewise pulse):
    the FP32 gradient and J within 1e-4 of the FP64 path (itself pinned to the oracle above),
    here at 256^2 and T = 100 ms (5000 steps) to keep the test short."""
    _print_case("c2_type_fp32_vs_fp64", grid=256, T=100)
    from fkadj import workloads as W
    m = W.holes_mask(256, 20, 7)
    y0 = W.spiral_states([m], [7])[0]
    A = W.amplitudes(100, 7)
    out = {}
    for prec in (fkadj.F64, fkadj.F32):
        inst = W.make_instance(m, 100.0, 100, precision=prec, ckpt=50)
        inst.set_state(y0)
        out[prec] = inst.grad(A)
        inst.close()
    e = relerr(out[fkadj.F32][0], out[fkadj.F64][0])
    eJ = abs(out[fkadj.F32][1] / out[fkadj.F64][1] - 1)
    print(f"c2-type fp32 vs fp64: grad {e:.3e}, J {eJ:.3e}")
    assert e < 1e-4 and eJ < 1e-4


@pytest.mark.slow
def test_fp32_c2_full_size_error_bound():
    """The C2 configuration at its full size (512^2, 80 disc holes, S1-S2 spiral, T = 200 ms =
    10 000 steps, 200 amplitudes, checkpoints every 50 steps): FP32 gradient and J within 1e-4 of
    the FP64 path, and the instance evaluated inside a batch gives its own result bit for bit."""
    _print_case("c2_full_fp32_vs_fp64", grid=512, T=200)
    from fkadj import workloads as W
    cfg = W.CONFIGS["C2"]
    masks = [W.holes_mask(512, cfg["holes"], s) for s in (0, 1)]
    states = W.spiral_states(masks, [0, 1])
    As = [W.amplitudes(cfg["K"], s) for s in (0, 1)]
    out = {}
    for prec in (fkadj.F64, fkadj.F32):
        inst = W.make_instance(masks[0], cfg["T"], cfg["K"], precision=prec, ckpt=cfg["ckpt"])
        inst.set_state(states[0])
        out[prec] = inst.grad(As[0])
        inst.close()
    e = relerr(out[fkadj.F32][0], out[fkadj.F64][0])
    eJ = abs(out[fkadj.F32][1] / out[fkadj.F64][1] - 1)
    print(f"c2 full size fp32 vs fp64: grad {e:.3e}, J {eJ:.3e}")
    assert e < 1e-4 and eJ < 1e-4
    insts = []
    for mk, y in zip(masks, states):
        inst = W.make_instance(mk, cfg["T"], cfg["K"], precision=fkadj.F32, ckpt=cfg["ckpt"])
        inst.set_state(y)
        insts.append(inst)
    batched = fkadj.batch_grad(insts, As)
    assert np.array_equal(batched[0][0], out[fkadj.F32][0]) and batched[0][1] == out[fkadj.F32][1]
    for inst in insts:
        inst.close()


def _sweep_case(seed):
    """Seeded small problem with everything drawn at random: ragged shape (down to 4 x 4), hole
    count, steps, pulse pieces, checkpoint interval and both tile configurations."""
    rng = np.random.default_rng(1000 + seed)
    H, W = int(rng.integers(4, 141)), int(rng.integers(4, 261))
    holes = int(rng.integers(0, 1 + H * W // 500))
    nt = int(rng.integers(6, 61))
    divisors = [k for k in range(1, nt + 1) if nt % k == 0]
    K = int(rng.choice(divisors))
    ckpt = int(rng.integers(1, nt + 1))
    ft = int(rng.choice([0, 1, 2, 3, 4, 5, 6, 7]))
    at = int(rng.choice([0, 1, 3, 4, 5, 6, 7]))
    return H, W, holes, nt, K, ckpt, ft, at


@pytest.mark.parametrize("seed", range(16))
def test_random_small_problems_match_oracle(seed):
    """Random shapes, holes, horizons, pulse pieces, checkpoint intervals and tiles: FP64 within
    1e-9 and FP32 within 1e-4 of the oracle (gradient and J).  The energy weight is 0, so the whole
    gradient comes from the adjoint solve."""
    H, W, holes, nt, K, ckpt, ft, at = _sweep_case(seed)
    _print_case("random_small", seed=seed, shape=(H, W), holes=holes, nt=nt, n_amp=K, ckpt=ckpt, fwd_tile=ft,
                adj_tile=at)
    case = disc_case(H, W, holes, seed)
    if case.n_act == 0:
        pytest.skip("every cell is a hole")
    import oracle as O
    m = fkadj.reference_model(h=0.025)
    om = O.reference_model(h=0.025)
    dt, lam = 0.02, 0.0
    A = 10.0 * np.random.default_rng(seed).uniform(0, 1, K)
    S = G.mean_u2_S(case.n_act)
    g_ref, Lc_ref, _, _ = oracle_grad_piecewise(case, dt, nt, A, (1, 0, 0), 2 * lam, S, om)
    for prec, tol in ((fkadj.F64, 1e-9), (fkadj.F32, 1e-4)):
        with fkadj.Instance(case.ai, H, W, dt=dt, nt=nt, model=m, pulse_mode=fkadj.PULSE_PIECEWISE, n_amp=K,
                            alpha=2 * lam, gamma=(1, 0, 0), s_diag=2.0 / case.n_act, precision=prec, ckpt=ckpt,
                            fwd_tile=ft, adj_tile=at) as inst:
            inst.set_state(case.y0)
            g, Lc, _, _ = inst.grad(A)
        assert relerr(g, g_ref) < tol, (prec, relerr(g, g_ref))
        assert abs(Lc - Lc_ref) < tol * abs(Lc_ref), (prec, Lc, Lc_ref)


def test_two_step_forward_launches_change_nothing():
    """Small groups on the built-in forward tile sweep forward two steps per launch (64x8 tiles, an
    8-cell halo); an explicit 64x8/512 forward tile keeps one step per launch.  Both give the same
    gradient and J bit for bit, alone and batched."""
    _print_case("forward_depth2", sizes="(64,64),(96,128)")
    m = fkadj.reference_model(h=0.025)
    nt, dt, K = 60, 0.02, 6
    res = {}
    for ft in (0, 6):
        insts, Es = [], []
        for k, (H, W) in enumerate([(64, 64), (96, 128)]):
            case = disc_case(H, W, 6 + 2 * k, 30 + k)
            inst = fkadj.Instance(case.ai, H, W, dt=dt, nt=nt, model=m, pulse_mode=fkadj.PULSE_PIECEWISE, n_amp=K,
                                  alpha=2e-3, gamma=(1, 0, 0), s_diag=2.0 / case.n_act, precision=fkadj.F32,
                                  ckpt=20, fwd_tile=ft)
            inst.set_state(case.y0)
            insts.append(inst)
            Es.append(np.random.default_rng(k).uniform(0, 10, K))
        res[ft] = [insts[0].grad(Es[0])] + fkadj.batch_grad(insts, Es)
        for inst in insts:
            inst.close()
    for a, b in zip(res[0], res[6]):
        assert np.array_equal(a[0], b[0]) and a[1] == b[1]


def test_two_step_adjoint_launches_change_nothing():
    """adj_depth = 2 takes two adjoint steps per launch (64x8 tiles, an 8-cell halo, J(y_{i-1}) of the
    first step kept for the second); adj_depth = 1 keeps one step per launch.  Both give the same
    gradient and J bit for bit, alone and batched, for segments of even and odd length (nt = 63 with
    K = 10 ends with a 3-step segment, K = 7 makes every segment odd), piecewise and knot-sampled
    fields, and kept segments; adj_depth out of {0, 1, 2} is refused."""
    m = fkadj.reference_model(h=0.025)
    dt = 0.02
    shapes = [(64, 64), (96, 128)]
    for nt, K, mode in ((60, 10, fkadj.PULSE_PIECEWISE), (63, 10, fkadj.PULSE_KNOTS), (63, 7, fkadj.PULSE_KNOTS)):
        _print_case("adjoint_depth2", sizes="(64,64),(96,128)", nt=nt, K=K, mode=mode)
        res = {}
        for ad in (1, 2):
            insts, Es = [], []
            for k, (H, W) in enumerate(shapes):
                case = disc_case(H, W, 6 + 2 * k, 50 + k)
                kw = dict(pulse_mode=mode, n_amp=3) if mode == fkadj.PULSE_PIECEWISE else {}
                seg = (2 * K + 1) * 3 * 4 * H * ((W + 63) // 64 * 64)
                inst = fkadj.Instance(case.ai, H, W, dt=dt, nt=nt, model=m, alpha=2e-3, gamma=(1, 0, 0),
                                      s_diag=2.0 / case.n_act, precision=fkadj.F32, ckpt=K, adj_depth=ad,
                                      keep_bytes=2 * seg if k == 1 else 0, **kw)
                inst.set_state(case.y0)
                insts.append(inst)
                Es.append(np.random.default_rng(k).uniform(0, 10, 3) if mode == fkadj.PULSE_PIECEWISE
                          else _Et(nt, k))
            res[ad] = [insts[0].grad(Es[0]), insts[1].grad(Es[1])] + fkadj.batch_grad(insts, Es)
            for inst in insts:
                inst.close()
        for a, b in zip(res[1], res[2]):
            assert np.array_equal(a[0], b[0]) and a[1] == b[1], (nt, K, mode)
    case = disc_case(64, 64, 6, 50)
    for bad in (-1, 3):
        with pytest.raises(fkadj.FkadjError):
            fkadj.Instance(case.ai, 64, 64, dt=dt, nt=20, model=m, precision=fkadj.F32, adj_depth=bad)


def test_kept_segments_change_nothing():
    """Kept segments (keep_bytes) are filled by the forward sweep and not recomputed: every budget,
    from none to all segments, gives the same gradient and J bit for bit, alone and batched with
    members keeping different numbers of segments; the evaluation launches nt forward steps plus the
    recompute of the segments not kept.  Knot-sampled field; nt = 62 with K = 20 (a short last
    segment, two forward-sweep steps per launch, kept ones included), nt = 61 (one step per launch)
    and nt = 30 with K = 40 (one segment, shorter than K)."""
    m = fkadj.reference_model(h=0.025)
    dt = 0.02
    shapes = [(64, 64), (96, 128)]
    for nt, K in ((62, 20), (61, 20), (30, 40)):
        _print_case("kept_segments", sizes="(64,64),(96,128)", nt=nt, K=K)
        nseg = (nt + K - 1) // K
        base, insts = {}, {}
        for k, (H, W) in enumerate(shapes):
            case = disc_case(H, W, 6 + 2 * k, 40 + k)
            seg = (2 * K + 1) * 3 * 4 * H * ((W + 63) // 64 * 64)
            for kept in sorted({0, 1, 2, nseg}):
                inst = fkadj.Instance(case.ai, H, W, dt=dt, nt=nt, model=m, alpha=2e-3, gamma=(1, 0, 0),
                                      s_diag=2.0 / case.n_act, precision=fkadj.F32, ckpt=K,
                                      keep_bytes=kept * seg + (seg // 2 if kept < nseg else 1 << 40))
                inst.set_state(case.y0)
                insts[k, kept] = inst
            E = _Et(nt, k)
            base[k] = insts[k, 0].grad(E)
            for kept in sorted({1, 2, nseg}):
                fkadj.set_profiling(True)
                try:
                    g = insts[k, kept].grad(E)
                    t = fkadj.last_timing()
                finally:
                    fkadj.set_profiling(False)
                assert np.array_equal(g[0], base[k][0]) and g[1] == base[k][1], (nt, k, kept)
                recomputed = max(0, min((nseg - kept) * K, nt))
                sweep = nt // 2 if nt % 2 == 0 else nt
                assert t["fwd_launches"] == sweep + recomputed and t["adj_launches"] == nt, (nt, k, kept, t)
        # batched: members keeping 1 and all segments, and none and 2
        for pair in (((0, 1), (1, nseg)), ((0, 0), (1, 2))):
            res = fkadj.batch_grad([insts[p] for p in pair], [_Et(nt, p[0]) for p in pair])
            for (k, _), r in zip(pair, res):
                assert np.array_equal(r[0], base[k][0]) and r[1] == base[k][1], (nt, pair)
        for inst in insts.values():
            inst.close()


@pytest.mark.slow
def test_fp32_c5_grid_short_horizon_error_bound():
    """The C5 grid at its full size (2048^2, 400 disc holes, four interacting spirals) over the first
    100 ms (5000 steps, 100 amplitudes): FP32 within 1e-4 of FP64.  (The full 500 ms horizon, where the
    four-spiral state amplifies per-step rounding to 1.5e-3, is asserted in test_gpu_configs.py.)"""
    _print_case("c5_grid_fp32_vs_fp64", grid=2048, T=100)
    from fkadj import workloads as W
    cfg = W.CONFIGS["C5"]
    m = W.holes_mask(cfg["N"], cfg["holes"], 5)
    y0 = W.spiral_states([m], [5], bands=2)[0]
    A = W.amplitudes(100, 5)
    out = {}
    for prec in (fkadj.F32, fkadj.F64):
        inst = W.make_instance(m, 100.0, 100, precision=prec, ckpt=cfg["ckpt"])
        inst.set_state(y0)
        out[prec] = inst.grad(A)
        inst.close()
    e = relerr(out[fkadj.F32][0], out[fkadj.F64][0])
    eJ = abs(out[fkadj.F32][1] / out[fkadj.F64][1] - 1)
    print(f"c5 grid T=100 fp32 vs fp64: grad {e:.3e}, J {eJ:.3e}")
    assert e < 1e-4 and eJ < 1e-4


def test_batch_forward_equals_forward_alone():
    """fkadj_batch_forward advances each resident state exactly as the instance's own forward run
    with advance (shared launches; the forward tile follows the launch size and changes nothing)."""
    _print_case("batch_forward", sizes="(64,64),(96,128),(40,192)")
    m = fkadj.reference_model(h=0.025)
    nt, dt, K, nsteps = 60, 0.02, 6, 45
    a, b, Es = [], [], []
    for k, (H, W) in enumerate([(64, 64), (96, 128), (40, 192)]):
        case = disc_case(H, W, 6 + 2 * k, 40 + k)
        for lst in (a, b):
            inst = fkadj.Instance(case.ai, H, W, dt=dt, nt=nt, model=m, pulse_mode=fkadj.PULSE_PIECEWISE, n_amp=K,
                                  alpha=2e-3, gamma=(1, 0, 0), s_diag=2.0 / case.n_act, precision=fkadj.F32, ckpt=20)
            inst.set_state(case.y0)
            lst.append(inst)
        Es.append(np.random.default_rng(k).uniform(0, 10, K))
    fkadj.batch_forward(a, Es, nsteps)
    for inst, E in zip(b, Es):
        inst.forward(E, nsteps=nsteps, advance=True)
    for x, y in zip(a, b):
        assert np.array_equal(x.get_state(), y.get_state())
    for inst in a + b:
        inst.close()


def test_stimulus_and_launch_accounting():
    """fkadj_stimulate sets u on the masked active cells only; with profiling on, the last timing
    reports the launches of the evaluation: nt forward-sweep steps (two per launch for small
    groups on the built-in forward tile, one per launch with an explicit tile), nt recompute
    launches and nt adjoint steps (one per launch, two with adj_depth = 2).  The
    segments are long enough (K = 20) that each kind's launch timing gets a full sampling window."""
    _print_case("stimulus_timing", shape=(48, 80))
    H, W, nt, K = 48, 80, 40, 4
    case = disc_case(H, W, 4, 61)
    m = fkadj.reference_model(h=0.025)
    counts = {}
    for ft, ad in ((0, 0), (6, 0), (0, 2)):
        with fkadj.Instance(case.ai, H, W, dt=0.02, nt=nt, model=m, pulse_mode=fkadj.PULSE_PIECEWISE, n_amp=K,
                            alpha=2e-3, gamma=(1, 0, 0), s_diag=2.0 / case.n_act, precision=fkadj.F32, ckpt=20,
                            fwd_tile=ft, adj_depth=ad) as inst:
            # values exact in FP32, so the resident state reads back unchanged
            y = np.concatenate([np.zeros(case.n_act), np.ones(case.n_act), np.full(case.n_act, 0.75)])
            inst.set_state(y)
            mask = np.zeros((H, W), bool)
            mask[:, : W // 3] = True
            inst.stimulate(mask, 1.0)
            got = inst.get_state()
            stim = mask.ravel()[case.ai > 0]          # packed order: active cells in increasing k
            assert np.array_equal(got[: case.n_act], np.where(stim, 1.0, 0.0))
            assert np.array_equal(got[case.n_act:], y[case.n_act:])
            fkadj.set_profiling(True)
            try:
                inst.grad(np.full(K, 2.0))
                t = fkadj.last_timing()
            finally:
                fkadj.set_profiling(False)
            counts[ft, ad] = (t["fwd_launches"], t["adj_launches"], t["cell_steps"])
            assert t["eval_ms"] > 0 and t["fwd_kernel_ms"] > 0 and t["adj_kernel_ms"] > 0
    pitch = (W + 63) // 64 * 64
    assert counts[0, 0] == (nt // 2 + nt, nt, H * pitch * nt)
    assert counts[6, 0] == (2 * nt, nt, H * pitch * nt)
    assert counts[0, 2] == (nt // 2 + nt, nt // 2, H * pitch * nt)
