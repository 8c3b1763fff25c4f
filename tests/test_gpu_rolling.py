"""GPU tests of the row-rolling kernels (fk_kernels_roll.hip): for every supported pitch, strip height,
adjoint tile, pulse mode, checkpoint schedule and kept-segment budget, a gradient computed with the
rolling kernels equals the tiled kernels' bit for bit (same per-cell arithmetic, same z.gh partial
order), batched and alone."""
import numpy as np
import pytest

import fkadj
from fkadj import geometry as G
from cases import disc_case

pytestmark = pytest.mark.gpu


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


@pytest.fixture(autouse=True)
def _restore_rolling():
    prev = fkadj.set_rolling(1, 16)
    yield
    fkadj.set_rolling(prev, 0)


def _front_state(case):
    return case.y0


def _instance(H, W, nt, n_amp, ckpt, seed, adj_tile=0, keep_bytes=0, knots=False):
    m = fkadj.reference_model(h=0.025)
    case = disc_case(H, W, 4 + seed % 5, seed)
    kw = dict(pulse_mode=fkadj.PULSE_KNOTS) if knots else dict(pulse_mode=fkadj.PULSE_PIECEWISE, n_amp=n_amp)
    inst = fkadj.Instance(case.ai, H, W, dt=0.02, nt=nt, model=m, alpha=2e-3, gamma=(1, 0.2, 0.2),
                          S=G.S_matrix(case.mask), precision=fkadj.F32, ckpt=ckpt, adj_tile=adj_tile,
                          keep_bytes=keep_bytes, **kw)
    inst.set_state(case.y0)
    rng = np.random.default_rng(seed)
    E = rng.uniform(0, 10, nt + 1 if knots else n_amp)
    return inst, E


def _grad(mode, hs, insts, Es):
    fkadj.set_rolling(mode, hs)
    if len(insts) == 1:
        return [insts[0].grad(Es[0])]
    return fkadj.batch_grad(insts, Es)


def _same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(x[0], y[0]), float(np.max(np.abs(x[0] - y[0])))
        assert x[1:] == y[1:], (x[1:], y[1:])


@pytest.mark.parametrize("W,H,hs,adj_tile", [(512, 70, 16, 0), (256, 64, 8, 0), (384, 45, 32, 4),
                                             (256, 33, 16, 6), (512, 40, 8, 6), (1024, 130, 64, 0)])
def test_rolling_equals_tiled_alone(W, H, hs, adj_tile):
    """Widths 256..1024 (ragged column count inside the pitch too), rows not a multiple of the strip,
    strips of 8..64 rows, both adjoint tiles; recompute segments of 20 steps over 60 steps."""
    _print_case("roll_alone", W=W, H=H, hs=hs, adj_tile=adj_tile)
    inst, E = _instance(H, W - 7, 60, 6, 20, 3 + H, adj_tile=adj_tile)
    tiled = _grad(0, hs, [inst], [E])
    rolled = _grad(2, hs, [inst], [E])
    _same(tiled, rolled)
    inst.close()


def test_rolling_equals_tiled_knot_mode_and_kept_segments():
    """Knot-sampled field (spline stage values, slope correction at segment ends) and kept segments
    (forward steps that store states and slopes, MODE_KEEP) on a batch of three instances."""
    _print_case("roll_knots_keep", n=3)
    pairs = [_instance(48, 250, 40, 0, 10, 7 + k, keep_bytes=(k * 3) << 20, knots=True) for k in range(3)]
    insts = [p[0] for p in pairs]
    Es = [p[1] for p in pairs]
    tiled = _grad(0, 16, insts, Es)
    rolled = _grad(2, 16, insts, Es)
    _same(tiled, rolled)
    for i in insts:
        i.close()


def test_rolling_batch_of_512_instances_equals_alone():
    """Four 512-wide instances batched with the rolling kernels equal each instance alone on the tiled
    kernels (the batched C2 case: the batch takes rolling kernels, a lone instance tiled ones)."""
    _print_case("roll_batch_512", n=4)
    pairs = [_instance(96 + 16 * k, 512, 50, 5, 25, 20 + k) for k in range(4)]
    insts = [p[0] for p in pairs]
    Es = [p[1] for p in pairs]
    alone = [_grad(0, 16, [i], [e])[0] for i, e in zip(insts, Es)]
    batched = _grad(2, 16, insts, Es)
    _same(alone, batched)
    for i in insts:
        i.close()


def test_served_requests_on_rolling_rounds_equal_alone():
    """Server rounds with the rolling kernels forced: requests of pitch 256, 384 and 512 at different
    steps and phases share each rolling launch (one per pitch), join mid-flight, and every result
    equals its instance alone on the tiled kernels."""
    _print_case("roll_served", n=4)
    specs = [(40, 250, 40, 4, 10, 61), (56, 500, 60, 6, 20, 62), (33, 380, 40, 4, 10, 63), (48, 250, 60, 6, 15, 64)]
    pairs = [_instance(H, W, nt, K, ck, seed) for (H, W, nt, K, ck, seed) in specs]
    fkadj.set_rolling(0, 16)
    alone = [inst.grad(E) for inst, E in pairs]
    fkadj.set_rolling(2, 16)
    with fkadj.Server(budget_bytes=1 << 34, max_slots=8) as srv:
        ids = [srv.submit(*pairs[0]), srv.submit(*pairs[1])]
        srv.run(23)
        ids += [srv.submit(*pairs[2]), srv.submit(*pairs[3])]
        while srv.pending():
            srv.run(0)
        res = [srv.poll(i) for i in ids]
    _same(alone, res)
    for inst, _ in pairs:
        inst.close()


def test_mixed_pitch_batch_rolls_per_pitch():
    """A batch of pitches 256, 384 (both rolled, one launch each) and 320 (not supported by the rolling
    kernels: tiled) in one group, members given out of pitch order: every result equals the instance
    alone on the tiled kernels."""
    _print_case("roll_mixed_pitch", n=5)
    specs = [(40, 300, 40, 4, 10, 81), (44, 250, 40, 4, 10, 82), (36, 380, 40, 4, 10, 83), (52, 256, 40, 4, 10, 84),
             (30, 330, 40, 4, 10, 85)]
    pairs = [_instance(H, W, nt, K, ck, seed) for (H, W, nt, K, ck, seed) in specs]
    insts = [p[0] for p in pairs]
    Es = [p[1] for p in pairs]
    alone = [_grad(0, 16, [i], [e])[0] for i, e in zip(insts, Es)]
    mixed = _grad(2, 16, insts, Es)
    _same(alone, mixed)
    for i in insts:
        i.close()


@pytest.mark.parametrize("W,H,hs", [(256, 4, 8), (384, 5, 16), (512, 9, 8), (256, 17, 16), (384, 7, 8)])
def test_rolling_equals_tiled_short_grids(W, H, hs):
    """Grids of 4 to 17 rows: one strip shorter than its strip height and than the four skewed stages,
    and a last strip of one row that rolls upwards (odd strip index); single-cell holes next to the
    strip boundaries.  Rolling equals tiled bit for bit."""
    _print_case("roll_short", W=W, H=H, hs=hs)
    mask = np.ones((H, W - 3), bool)
    for r, c in ((H // 2, 40), (min(hs, H - 2), 90), (1, W // 2), (H - 2, W - 20)):
        mask[r, c] = False
    mask = G.fill_isthmuses(mask)   # no active cell without a neighbour on an axis
    ai = G.active_index_from_mask(mask)
    n = int(mask.sum())
    rng = np.random.default_rng(H * 7 + W)
    y = np.concatenate([rng.uniform(0, 1, n), rng.uniform(0.5, 1, n), rng.uniform(0.5, 1, n)])
    inst = fkadj.Instance(ai, H, W - 3, dt=0.02, nt=30, model=fkadj.reference_model(h=0.025),
                          pulse_mode=fkadj.PULSE_PIECEWISE, n_amp=3, alpha=2e-3, gamma=(1, 0.2, 0.2),
                          S=G.S_matrix(mask), precision=fkadj.F32, ckpt=10)
    inst.set_state(y)
    E = rng.uniform(0, 10, 3)
    tiled = _grad(0, hs, [inst], [E])
    rolled = _grad(2, hs, [inst], [E])
    assert np.all(np.isfinite(tiled[0][0]))
    _same(tiled, rolled)
    inst.close()
