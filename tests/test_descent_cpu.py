"""Gradient-descent driver pieces that need no GPU: the input-file parser (literals only, the
reference's read_check_2 rules), checkpoint discovery and the MAT-file input conventions."""
import os

import numpy as np
import pytest
import scipy.io as sio
import scipy.sparse as sp

from fkadj import config as CF
from fkadj import descent as D
from fkadj import geometry as G

INFILE = """% -*-octave-*-
alpha=0.25; % weight of the field term
ds=2e-3; % descent step
ns=6;
final_time=30;
gamma = [1, 0.5, 0.25]; % per-variable weights
nmod = 2;
Et_seed_file = 'seed.mat';
icondfile = 'state.mat';
grid_file = 'grid.mat';
S_file = 'S.mat';
mex_function = @fk_grad_Et_mex;
in_path = './in/';
out_path = './out/';
jobid = '000007';
unused = 3;
"""


def test_infile_parses_literals_and_reports_the_rest():
    notes = []
    v = CF.read_infile(INFILE, text=True, log=notes.append)
    assert v["alpha"] == 0.25 and v["ds"] == 2e-3 and v["ns"] == 6.0
    assert v["gamma"] == [1.0, 0.5, 0.25]
    assert v["Et_seed_file"] == "seed.mat" and v["mex_function"] == "fk_grad_Et_mex"
    assert v["jobid"] == "000007"
    assert any("unused" in n for n in notes) and any("octave" not in n for n in notes)
    cfg = CF.GradDescentConfig.from_values(v)
    assert cfg.nt == 300 and cfg.nmod == 2 and cfg.ns == 6


def test_infile_missing_names_and_no_evaluation():
    with pytest.raises(KeyError) as e:
        CF.read_infile("alpha=1;\nds=2;", text=True)
    assert "ns" in str(e.value) and "jobid" in str(e.value)
    with pytest.raises(ValueError):
        CF.parse_value("system('rm -rf /')")
    assert CF.parse_value("'it''s'") == "it's"
    assert CF.parse_value("[1 2;3]") == [1.0, 2.0, 3.0]
    with pytest.raises(ValueError):
        CF.GradDescentConfig.from_values(dict(CF.read_infile(INFILE, text=True), final_time=30.05)).nt


def test_latest_checkpoint(tmp_path):
    for i in (2, 4, 10):
        (tmp_path / D.checkpoint_name("2026-dec-03-10-00-00_000007", i)).write_bytes(b"")
    (tmp_path / D.checkpoint_name("2026-dec-03-10-00-00_000008", 12)).write_bytes(b"")
    path, step = D.latest_checkpoint(str(tmp_path), "000007")
    assert step == 10 and path.endswith("_10.mat")
    assert D.latest_checkpoint(str(tmp_path), "000009") == (None, 0)


def test_load_inputs_follows_the_reference_conventions(tmp_path):
    mask = G.disc_holes(16, 16, 3, 1, 2, seed=1)
    ai = G.active_index_from_mask(mask).reshape(16, 16)
    n = int(mask.sum())
    yf = np.full(3 * 256, np.nan)
    rng = np.random.default_rng(0)
    y0 = rng.uniform(size=3 * n)
    bx3 = np.concatenate([ai.ravel() > 0] * 3)
    yf[bx3] = y0
    S = sp.identity(n, format="csc")
    sio.savemat(tmp_path / "grid.mat", {"active_index": ai.astype(float), "ghost_lr": G.ghost_lr_from_mask(mask).reshape(-1, 1)})
    sio.savemat(tmp_path / "state.mat", {"yf": yf.reshape(-1, 1)})
    sio.savemat(tmp_path / "seed.mat", {"Et": np.arange(301.0).reshape(-1, 1)})
    sio.savemat(tmp_path / "S.mat", {"S": S})
    cfg = CF.GradDescentConfig.from_values(dict(CF.read_infile(INFILE, text=True), in_path=str(tmp_path)))
    d = D.load_inputs(cfg)
    assert (d["nrows"], d["ncols"]) == (16, 16)
    assert np.array_equal(d["active_index"], ai.ravel()) and d["active_index"].dtype == np.int32
    assert np.array_equal(d["y0"], y0) and len(d["Et"]) == cfg.nt + 1
    assert d["S"].shape == (n, n)
