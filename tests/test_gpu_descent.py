"""GPU tests of the gradient-descent driver (src/cuda/simple_grad_desc.m restated in
fkadj.descent): the reference's own 200-step run is reproduced checkpoint by checkpoint, and a run
resumed from its last checkpoint ends exactly where an uninterrupted run ends."""
import numpy as np
import pytest
import scipy.io as sio

import fkadj
from fkadj import config as CF
from fkadj import descent as D
from fkadj import geometry as G

from cases import crop_case, golden, relerr

pytestmark = pytest.mark.gpu


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _cfg(out, **kw):
    base = dict(alpha=0.5, ds=1e-3, ns=200, final_time=300.0, gamma=[1.0, 0.2, 0.2], nmod=50,
                Et_seed_file="seed.mat", icondfile="spirals.mat", grid_file="grid.mat", S_file="S.mat",
                mex_function="fk_grad_Et_mex", in_path="", out_path=str(out), jobid="000001")
    base.update(kw)
    return CF.GradDescentConfig(**base)


@pytest.mark.slow
def test_descent_reproduces_the_reference_run(tmp_path):
    """200 steps of E <- E - ds*gradE on the 256^2 golden case (FP64): the checkpoints at steps 50,
    100, 150, 200 hold the reference's E (before the update), gradE and L/M/N histories."""
    _print_case("descent_golden256", ns=200, nmod=50)
    d, meta = golden()
    mask = G.mask_from_active_index(d["active_index"], 256, 256)
    inputs = dict(Et=d["Et_seed"].copy(), y0=d["y0"], active_index=d["active_index"], nrows=256, ncols=256,
                  ghost_lr=d["ghost_lr"].astype(float), S=G.S_matrix(mask))
    cfg = _cfg(tmp_path, alpha=meta["alpha"], ds=meta["ds"], final_time=meta["final_time"], gamma=meta["gamma"],
               nmod=int(meta["nmod"]))
    run = D.Descent(cfg, inputs)
    try:
        _, files = run.run(time_str="test_000001", resume=False, log=lambda *a: None)
    finally:
        run.close()
    assert [f.rsplit("_", 1)[1] for f in files] == ["50.mat", "100.mat", "150.mat", "200.mat"]
    for f, it in zip(files, (50, 100, 150, 200)):
        c = sio.loadmat(f)
        assert relerr(c["Et"].ravel(), d[f"Et_{it}"]) < 1e-12, it
        assert relerr(c["gradE"].ravel(), d[f"gradE_{it}"]) < 1e-9, it
        for k in ("Lcost_vec", "L0_vec", "LE_vec"):
            assert np.max(np.abs(c[k].ravel() / d[f"{k}_{it}"] - 1)) < 1e-10, (it, k)
        assert "tau_w_plus" in str(c["param_str"][0])


def test_resumed_run_equals_uninterrupted(tmp_path):
    _print_case("descent_resume", shape=(32, 64), ns=6, nmod=2)
    crop = crop_case(64, 96, 32, 64)
    nt = 40
    Et = 2.0 + np.sin(np.arange(nt + 1) * 0.2)
    inputs = dict(Et=Et, y0=crop.y0, active_index=crop.ai, nrows=crop.nrows, ncols=crop.ncols, ghost_lr=crop.gh,
                  S=G.S_matrix(crop.mask))
    runs = {}
    for name, stop in (("whole", None), ("split", 3)):
        out = tmp_path / name
        cfg = _cfg(out, ns=6, nmod=2, final_time=nt * 0.1, ds=0.05)
        r = D.Descent(cfg, inputs)
        try:
            r.run(time_str="t_000001", resume=False, log=lambda *a: None, stop_after=stop)
        finally:
            r.close()
        if stop:
            r = D.Descent(cfg, inputs)
            try:
                E_end, files = r.run(resume=True, log=lambda *a: None)
            finally:
                r.close()
            assert [f.rsplit("_", 1)[1] for f in files] == ["4.mat", "6.mat"]
        runs[name] = sio.loadmat(str(out / D.checkpoint_name("t_000001", 6)))
    a, b = runs["whole"], runs["split"]
    for k in ("Et", "gradE", "Lcost_vec", "L0_vec", "LE_vec"):
        assert np.array_equal(a[k], b[k]), k
