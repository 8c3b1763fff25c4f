"""GPU test of fkadj_grad_noalloc, the entry point that allocates no device memory: through the C ABI
its gradient equals fkadj_grad's bit for bit on two calls in a row (the second on a smaller problem,
so it runs on what the first left in the library's static array), for FP32 and FP64 instances, and
an evaluation larger than the array fails with FKADJ_ERR_SCRATCH before any device work."""
import numpy as np
import pytest

import fkadj
from fkadj import _lib as L
from cases import disc_case

pytestmark = pytest.mark.gpu


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _instance(H, W, nt, n_amp, ckpt, seed, precision, lazy=True):
    m = fkadj.reference_model(h=0.025)
    case = disc_case(H, W, 6, seed)
    inst = fkadj.Instance(case.ai, H, W, dt=0.02, nt=nt, model=m, pulse_mode=fkadj.PULSE_PIECEWISE, n_amp=n_amp,
                          alpha=2e-3, gamma=(1, 0, 0), s_diag=2.0 / case.n_act, precision=precision, ckpt=ckpt,
                          lazy_workspace=lazy)
    inst.set_state(case.y0)
    A = np.random.default_rng(seed).uniform(0, 10, n_amp)
    return inst, A


@pytest.mark.parametrize("precision", [fkadj.F32, fkadj.F64])
def test_noalloc_equals_grad_on_two_calls(precision):
    # first call: 256 x 192, 400 steps; second: a smaller grid and fewer steps
    specs = [(256, 192, 400, 8, 50, 11), (96, 128, 120, 6, 20, 12)]
    for H, W, nt, K, ck, seed in specs:
        _print_case("noalloc_vs_grad", precision=precision, H=H, W=W, nt=nt, ckpt=ck)
        inst, A = _instance(H, W, nt, K, ck, seed, precision)
        want = inst.grad(A)
        got = inst.grad_noalloc(A)
        assert np.array_equal(want[0], got[0]), (precision, H, W)
        assert want[1:] == got[1:], (precision, H, W, want[1:], got[1:])
        inst.close()


def test_noalloc_repeat_and_new_field_on_one_instance():
    _print_case("noalloc_repeat", H=128, W=128, nt=200)
    inst, A = _instance(128, 128, 200, 10, 25, 21, fkadj.F32, lazy=False)
    first = inst.grad_noalloc(A)
    B = A[::-1].copy()
    want_b = inst.grad(B)
    got_b = inst.grad_noalloc(B)
    again = inst.grad_noalloc(A)
    assert np.array_equal(want_b[0], got_b[0]) and want_b[1] == got_b[1]
    assert np.array_equal(first[0], again[0]) and first[1] == again[1]
    inst.close()


def test_noalloc_too_large_fails_before_device_work():
    # 2048 x 2048 FP32 state: 48 MiB; 3 checkpoints + 51 segment states exceed the 256 MiB array
    _print_case("noalloc_too_large", H=2048, W=2048, nt=100)
    m = fkadj.reference_model(h=0.025)
    mask = np.ones((2048, 2048), bool)
    from fkadj import geometry as G
    ai = G.active_index_from_mask(mask)
    inst = fkadj.Instance(ai, 2048, 2048, dt=0.02, nt=100, model=m, pulse_mode=fkadj.PULSE_PIECEWISE, n_amp=4,
                          alpha=2e-3, gamma=(1, 0, 0), s_diag=1.0 / mask.size, precision=fkadj.F32, ckpt=50,
                          lazy_workspace=True)
    inst.set_state(np.concatenate([np.zeros(mask.size), np.ones(mask.size), np.ones(mask.size)]))
    with pytest.raises(fkadj.FkadjError) as e:
        inst.grad_noalloc(np.ones(4))
    assert e.value.code == L.ERR_SCRATCH
    assert "no-allocation" in fkadj.strerror(L.ERR_SCRATCH)
    inst.close()
