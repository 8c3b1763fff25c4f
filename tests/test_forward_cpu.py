"""Forward snapshot driver pieces that need no GPU: the snapshot count of the reference's
compute_ns, float_integer, the parameter file and the output-name rules."""
import numpy as np
import pytest
import scipy.io as sio

import fkadj
from fkadj import config as CF
from fkadj import forward_run as FR

INFILE = """% -*-octave-*-
Et_file = 'pulse.mat';
final_time = 40; % ignored unless Et_file is empty
nstep = 7;
icondfile = 'state.mat';
outfile = ''; % empty: outfile_prefix + Et_file
outfile_prefix = 'snaps_';
mex_function = @fk_forward_Et_mex;
grid_file = 'grid.mat';
in_path = './in/';
out_path = './out/';
"""


def test_snapshot_count_follows_compute_ns():
    # nt/nstep - 1 when nstep divides nt (final state returned separately), floor otherwise
    cases = {(3000, 100): 29, (25, 10): 2, (30, 10): 2, (5, 10): 0, (10, 10): 0, (0, 5): 0, (1, 1): 0, (7, 2): 3}
    for (nt, nstep), ns in cases.items():
        assert fkadj.snapshot_count(nt, nstep) == ns, (nt, nstep)
    with pytest.raises(fkadj.FkadjError):
        fkadj.snapshot_count(10, 0)


def test_float_integer_is_the_references_tight_test():
    assert CF.float_integer(300 / 0.1) == 3000          # 2999.9999999999995
    assert CF.float_integer(200 / 0.02) == 10000
    assert CF.float_integer(0.0) == 0
    for bad in (30.05 / 0.1, 2.5, 1e6 + 1e-6):
        with pytest.raises(ValueError):
            CF.float_integer(bad)


def test_forward_infile_and_output_name(tmp_path):
    v = CF.read_infile(INFILE, names=FR.FORWARD_VARS, text=True)
    assert v["nstep"] == 7.0 and v["outfile"] == "" and v["mex_function"] == "fk_forward_Et_mex"
    assert FR.output_name(v) == "snaps_pulse.mat"
    assert FR.output_name(dict(v, outfile="given.mat")) == "given.mat"
    with pytest.raises(ValueError):
        FR.output_name(dict(v, outfile_prefix=""))
    with pytest.raises(ValueError):
        FR.output_name(dict(v, Et_file=""))


def test_field_from_file_or_zero(tmp_path):
    sio.savemat(tmp_path / "pulse.mat", {"Et": np.arange(401.0).reshape(1, -1), "dt": 0.1, "final_time": 40.0,
                                         "alpha": 0.5})
    v = dict(Et_file="pulse.mat", in_path=str(tmp_path), final_time=12.0)
    Et, dt, nt, T, alpha = FR.load_field(v)
    assert nt == 400 and dt == 0.1 and T == 40.0 and alpha == 0.5 and Et.shape == (401,)
    Et, dt, nt, T, alpha = FR.load_field(dict(v, Et_file=""))
    assert nt == 120 and dt == 0.1 and alpha is None and not Et.any()


def test_grid_must_be_int32_and_state_from_yf(tmp_path):
    ai = np.array([[1, 2, 0], [3, 4, 5]], np.int32)
    sio.savemat(tmp_path / "grid.mat", {"active_index": ai, "ghost_lr": np.zeros((2, 3), np.int16)})
    yf = np.arange(18.0)
    yf[[2, 8, 14]] = np.nan                              # the hole in each plane
    sio.savemat(tmp_path / "state.mat", {"yf": yf.reshape(-1, 1)})
    v = dict(grid_file="grid.mat", icondfile="state.mat", in_path=str(tmp_path))
    got_ai, nrows, ncols, gh, y0 = FR.load_grid_and_state(v)
    assert (nrows, ncols) == (2, 3) and np.array_equal(got_ai, ai.ravel())
    assert np.array_equal(y0, np.delete(yf, [2, 8, 14]))
    sio.savemat(tmp_path / "grid.mat", {"active_index": ai.astype(float), "ghost_lr": np.zeros((2, 3))})
    with pytest.raises(TypeError):
        FR.load_grid_and_state(v)
