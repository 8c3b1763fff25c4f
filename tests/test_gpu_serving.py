"""GPU tests of the continuous-batching server: every result equals the request evaluated alone,
admission respects the memory budget, cancellation and oversize requests fail only themselves."""
import numpy as np
import pytest

import fkadj
from cases import disc_case

pytestmark = pytest.mark.gpu


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _instances(specs, lazy=True, keep_bytes=0):
    """specs: (H, W, nt, n_amp, ckpt, seed) -> FP32 piecewise instances with disc holes."""
    m = fkadj.reference_model(h=0.025)
    out = []
    for (H, W, nt, K, ck, seed) in specs:
        case = disc_case(H, W, 5 + seed % 4, seed)
        inst = fkadj.Instance(case.ai, H, W, dt=0.02, nt=nt, model=m, pulse_mode=fkadj.PULSE_PIECEWISE, n_amp=K,
                              alpha=2e-3, gamma=(1, 0, 0), s_diag=2.0 / case.n_act, precision=fkadj.F32, ckpt=ck,
                              lazy_workspace=lazy, keep_bytes=keep_bytes)
        inst.set_state(case.y0)
        A = np.random.default_rng(seed).uniform(0, 10, K)
        out.append((inst, A))
    return out


SPECS = [(64, 64, 60, 6, 20, 1), (96, 128, 90, 9, 25, 2), (40, 192, 40, 4, 10, 3), (80, 96, 120, 12, 50, 4),
         (128, 64, 75, 5, 15, 5)]


def test_mixed_requests_equal_alone():
    _print_case("server_mix", n=len(SPECS))
    pairs = _instances(SPECS)
    alone = [inst.grad(A) for inst, A in pairs]
    with fkadj.Server(budget_bytes=1 << 34, max_slots=8) as srv:
        ids = [srv.submit(*pairs[0]), srv.submit(*pairs[1])]
        srv.run(37)                                  # later requests join mid-flight
        ids += [srv.submit(*pairs[2]), srv.submit(*pairs[3])]
        srv.run(11)
        ids.append(srv.submit(*pairs[4]))
        while srv.pending():
            srv.run(0)
        res = [srv.poll(i) for i in ids]
    for a, r in zip(alone, res):
        assert np.array_equal(a[0], r[0]) and a[1] == r[1] and a[2] == r[2]
    for inst, _ in pairs:
        inst.close()


def test_requests_on_one_instance_run_side_by_side():
    _print_case("server_same_instance", n=3)
    (inst, A), = _instances(SPECS[1:2])
    As = [A, A[::-1].copy(), 0.5 * A]
    alone = [inst.grad(a) for a in As]
    with fkadj.Server(budget_bytes=1 << 34, max_slots=4) as srv:
        ids = [srv.submit(inst, a) for a in As]
        assert srv.run(0) == 3                       # admitted together, finished in the same round
        res = [srv.poll(i) for i in ids]
    for a, r in zip(alone, res):
        assert np.array_equal(a[0], r[0]) and a[1] == r[1] and a[2] == r[2]
    inst.close()


def test_budget_admits_one_at_a_time_and_repeats_on_one_instance():
    _print_case("server_budget")
    pairs = _instances(SPECS[:3])
    alone = [inst.grad(A) for inst, A in pairs]
    biggest = max(inst.request_bytes for inst, _ in pairs)
    with fkadj.Server(budget_bytes=biggest, max_slots=4) as srv:
        ids = [srv.submit(inst, A) for inst, A in pairs]
        ids.append(srv.submit(*pairs[0]))            # the same instance again: runs after the first
        while srv.pending():
            srv.run(0)
        res = [srv.poll(i) for i in ids]
    for a, r in zip(alone + [alone[0]], res):
        assert np.array_equal(a[0], r[0]) and a[1] == r[1]
    for inst, _ in pairs:
        inst.close()


def test_oversize_and_cancel_fail_only_themselves():
    _print_case("server_cancel")
    pairs = _instances(SPECS[:3])
    alone = [inst.grad(A) for inst, A in pairs]
    small = min(inst.request_bytes for inst, _ in pairs)
    big = max(range(3), key=lambda k: pairs[k][0].request_bytes)
    with fkadj.Server(budget_bytes=pairs[big][0].request_bytes, max_slots=4) as srv:
        ids = [srv.submit(inst, A) for inst, A in pairs]
        srv.run(5)
        victim = [k for k in range(3) if k != big][0]
        srv.cancel(ids[victim])
        while srv.pending():
            srv.run(0)
        for k in range(3):
            if k == victim:
                with pytest.raises(fkadj.FkadjError) as e:
                    srv.poll(ids[k])
                assert e.value.code == -7
            else:
                r = srv.poll(ids[k])
                assert np.array_equal(alone[k][0], r[0])
    with fkadj.Server(budget_bytes=small // 2, max_slots=2) as srv:
        rid = srv.submit(*pairs[0])
        while srv.pending():
            srv.run(0)
        with pytest.raises(fkadj.FkadjError) as e:
            srv.poll(rid)
        assert e.value.code == -3
    for inst, _ in pairs:
        inst.close()


def test_queued_request_holds_its_instance():
    """From submission on (queued, not yet admitted) the instance is busy: its state cannot be
    replaced and it cannot be destroyed until the request is cancelled or finished."""
    _print_case("server_queued_busy")
    (inst, A), = _instances(SPECS[:1])
    y0 = inst.get_state()
    with fkadj.Server(budget_bytes=1 << 34, max_slots=2) as srv:
        rid = srv.submit(inst, A)
        for call in (lambda: inst.set_state(y0), inst.close):
            with pytest.raises(fkadj.FkadjError) as e:
                call()
            assert e.value.code == -8
        srv.cancel(rid)
        inst.set_state(y0)                          # free again once the queued request is cancelled
        with pytest.raises(fkadj.FkadjError) as e:
            srv.poll(rid)
        assert e.value.code == -7
        rid = srv.submit(inst, A)
        while srv.pending():
            srv.run(0)
        srv.poll(rid)
    inst.close()


def test_knot_mode_requests_equal_alone():
    """The reference's own field form (nt+1 knot samples, not-a-knot spline): served and batched
    requests give the instance's own result bit for bit."""
    _print_case("server_knots")
    m = fkadj.reference_model(h=0.025)
    pairs = []
    for k, (H, W, nt) in enumerate([(48, 64, 40), (64, 96, 40)]):
        case = disc_case(H, W, 4 + k, 50 + k)
        inst = fkadj.Instance(case.ai, H, W, dt=0.02, nt=nt, model=m, pulse_mode=fkadj.PULSE_KNOTS, alpha=2e-3,
                              gamma=(1, 0.2, 0.2), s_diag=2.0 / case.n_act, precision=fkadj.F32, ckpt=8)
        inst.set_state(case.y0)
        Et = 2.0 + np.sin(np.arange(nt + 1) * (0.1 + 0.05 * k))
        pairs.append((inst, Et))
    alone = [inst.grad(E) for inst, E in pairs]
    batched = fkadj.batch_grad([p[0] for p in pairs], [p[1] for p in pairs])
    with fkadj.Server(budget_bytes=1 << 34, max_slots=4) as srv:
        ids = [srv.submit(inst, E) for inst, E in pairs]
        while srv.pending():
            srv.run(0)
        served = [srv.poll(i) for i in ids]
    for a, b, c in zip(alone, batched, served):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[0], c[0]) and a[1] == b[1] == c[1]
    for inst, _ in pairs:
        inst.close()


def test_requests_on_instances_with_kept_segments_equal_alone():
    """Instances keeping trajectory segments for their own evaluations: a served request (whose
    buffers keep none) gives what fkadj_grad on the instance gives."""
    _print_case("server_kept", n=3)
    pairs = _instances(SPECS[:3], lazy=False, keep_bytes=1 << 40)
    alone = [inst.grad(A) for inst, A in pairs]
    with fkadj.Server(budget_bytes=1 << 34, max_slots=8) as srv:
        ids = [srv.submit(inst, A) for inst, A in pairs]
        while srv.pending():
            srv.run(0)
        res = [srv.poll(i) for i in ids]
    for a, r in zip(alone, res):
        assert np.array_equal(a[0], r[0]) and a[1] == r[1] and a[2] == r[2]
    for inst, _ in pairs:
        inst.close()


def test_pooled_context_reused_by_another_instance_of_the_same_shape():
    """A finished request's buffers serve the next request of the same shape, on another instance:
    a mirrored hole layout (same active cells, holes elsewhere) must not see the first instance's
    values at its hole cells.  Each result equals the instance alone."""
    _print_case("server_pool_reuse")
    m = fkadj.reference_model(h=0.025)
    case = disc_case(96, 128, 7, 31)
    masks = [case.mask, case.mask[:, ::-1].copy()]
    pairs = []
    for k, mk in enumerate(masks):
        ai = fkadj.geometry.active_index_from_mask(mk)
        y = np.concatenate([case.y0[v * case.n_act:(v + 1) * case.n_act] for v in range(3)])
        inst = fkadj.Instance(ai, 96, 128, dt=0.02, nt=80, model=m, pulse_mode=fkadj.PULSE_PIECEWISE, n_amp=8,
                              alpha=2e-3, gamma=(1, 0, 0), s_diag=2.0 / case.n_act, precision=fkadj.F32, ckpt=20,
                              lazy_workspace=True)
        inst.set_state(y)
        pairs.append((inst, np.random.default_rng(40 + k).uniform(0, 10, 8)))
    alone = [inst.grad(A) for inst, A in pairs]
    with fkadj.Server(budget_bytes=pairs[0][0].request_bytes, max_slots=2) as srv:
        out = []
        for inst, A in pairs + pairs[:1]:             # one at a time: each takes the pooled buffers
            rid = srv.submit(inst, A)
            while srv.pending():
                srv.run(0)
            out.append(srv.poll(rid))
    for a, r in zip(alone + alone[:1], out):
        assert np.array_equal(a[0], r[0]) and a[1] == r[1]
    for inst, _ in pairs:
        inst.close()


def test_injected_device_error_fails_only_its_request(monkeypatch):
    """A completion failure forced from the host on one request of a running mix (no device fault is
    involved): that request reports FKADJ_ERR_DEVICE, every other request equals its instance alone."""
    _print_case("server_fault_injection")
    pairs = _instances(SPECS)
    alone = [inst.grad(A) for inst, A in pairs]
    monkeypatch.setenv("FKADJ_TEST_FAIL_REQUEST", "3")   # ids count from 1 per server
    with fkadj.Server(budget_bytes=1 << 34, max_slots=8) as srv:
        ids = [srv.submit(inst, A) for inst, A in pairs]
        assert ids[2] == 3
        while srv.pending():
            srv.run(0)
        for k, rid in enumerate(ids):
            if k == 2:
                with pytest.raises(fkadj.FkadjError) as e:
                    srv.poll(rid)
                assert e.value.code == -4 and "injected" in str(e.value)
            else:
                r = srv.poll(rid)
                assert np.array_equal(alone[k][0], r[0]) and alone[k][1] == r[1]
    for inst, _ in pairs:
        inst.close()


def test_nonfinite_request_fails_only_itself():
    """A field large enough to drive its request's FP32 state to Inf/NaN: that request reports
    FKADJ_ERR_NONFINITE, the requests sharing its launches equal their instances alone, and the same
    instance serves a normal field afterwards."""
    _print_case("server_nonfinite")
    pairs = _instances(SPECS[:3])
    alone = [inst.grad(A) for inst, A in pairs]
    huge = np.full_like(pairs[1][1], 1e30)
    with fkadj.Server(budget_bytes=1 << 34, max_slots=8) as srv:
        ids = [srv.submit(inst, A) for inst, A in pairs]
        bad = srv.submit(pairs[1][0], huge)
        while srv.pending():
            srv.run(0)
        with pytest.raises(fkadj.FkadjError) as e:
            srv.poll(bad)
        assert e.value.code == -5
        for k, rid in enumerate(ids):
            r = srv.poll(rid)
            assert np.array_equal(alone[k][0], r[0]) and alone[k][1] == r[1]
        rid = srv.submit(*pairs[1])                  # the pooled buffers of the failed request again
        while srv.pending():
            srv.run(0)
        r = srv.poll(rid)
        assert np.array_equal(alone[1][0], r[0])
    for inst, _ in pairs:
        inst.close()


@pytest.mark.parametrize("roll_mode", [0, 2])
def test_slot_reuse_and_rolling_rounds_equal_alone(roll_mode):
    """More requests than slots over several pitches, so that freed slots pass to later requests, and
    (mode 2) the pitches the row-rolling kernels take run on them in the server's rounds.  Every result
    equals its instance alone (evaluated on the tiled kernels)."""
    _print_case("server_slot_reuse", roll_mode=roll_mode)
    specs = [(48, 256, 60, 6, 20, 11), (40, 384, 50, 5, 25, 12), (32, 200, 40, 4, 10, 13), (56, 256, 45, 5, 15, 14),
             (24, 512, 30, 3, 10, 15), (40, 320, 36, 4, 12, 16), (36, 384, 44, 4, 11, 17)]
    pairs = _instances(specs)
    prev = fkadj.set_rolling(0, 16)
    try:
        alone = [inst.grad(A) for inst, A in pairs]
        fkadj.set_rolling(roll_mode, 16)
        with fkadj.Server(budget_bytes=1 << 34, max_slots=3) as srv:
            ids = [srv.submit(inst, A) for inst, A in pairs[:4]]
            srv.run(7)
            ids += [srv.submit(inst, A) for inst, A in pairs[4:]]
            while srv.pending():
                srv.run(0)
            res = [srv.poll(i) for i in ids]
    finally:
        fkadj.set_rolling(prev, 0)
    for a, r in zip(alone, res):
        assert np.array_equal(a[0], r[0]) and a[1] == r[1] and a[2] == r[2]
    for inst, _ in pairs:
        inst.close()
