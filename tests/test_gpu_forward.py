"""GPU tests of the forward snapshot path: the drop-in for fk_forward_Et_comp (reference layout,
compute_ns snapshot count) against the FP64 oracle, and the snapshot driver end to end."""
import numpy as np
import pytest
import scipy.io as sio

import fkadj
from fkadj import forward_run as FR
from cases import crop_case, relerr

pytestmark = pytest.mark.gpu


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _Et(nt, seed=0):
    rng = np.random.default_rng(seed)
    return 1.5 + np.sin(np.arange(nt + 1) * 0.15 + rng.uniform(0, 3))


@pytest.mark.parametrize("nt,nstep", [(25, 10), (30, 10)])
def test_forward_ref_matches_oracle_in_reference_layout(nt, nstep):
    """ys holds compute_ns(nt, nstep) full-grid states (u, v, w planes, NaN at holes) after steps
    nstep, 2*nstep, ...; yf the final state; FP64 within 1e-12 of the oracle."""
    import oracle as O
    case = crop_case(64, 96, 32, 64)
    _print_case("forward_ref", shape=case.mask.shape, nt=nt, nstep=nstep)
    dt = 0.1
    Et = _Et(nt, nstep)
    ys, yf, pstr = fkadj.fk_forward_Et_mex(Et, dt, nt, case.y0, case.ai, case.nrows, case.ncols, case.gh, nstep)
    ns = {(25, 10): 2, (30, 10): 2}[(nt, nstep)]
    n = case.nrows * case.ncols
    assert ys.shape == (3 * n, ns) and yf.shape == (3 * n,) and "tau_w_plus" in pstr
    om = O.reference_model()
    Ef, _ = O.knot_tables(dt, Et)
    holes = np.tile(case.ai == 0, 3)
    for ks, step in enumerate([nstep * (k + 1) for k in range(ns)] + [nt]):
        want = O.forward(om, case.planes(), case.gh_plane(om), dt, Ef[:3 * step]).ravel()
        got = ys[:, ks] if ks < ns else yf
        assert np.array_equal(np.isnan(got), holes)
        assert relerr(got[~holes], want[~holes]) < 1e-12, step


def test_forward_driver_writes_the_references_variables(tmp_path):
    """The snapshot driver reads its parameter file and MAT inputs, runs the forward solve and writes
    ys, yf and the run's metadata; ys and yf equal the direct call."""
    case = crop_case(64, 96, 32, 64)
    _print_case("forward_driver", shape=case.mask.shape)
    dt, T, nstep = 0.1, 3.0, 7
    nt = 30
    Et = _Et(nt, 4)
    ai = case.ai.reshape(case.nrows, case.ncols).astype(np.int32)
    sio.savemat(tmp_path / "pulse.mat", {"Et": Et.reshape(-1, 1), "dt": dt, "final_time": T})
    # ghost_lr is stored per active cell as int16, as in the reference's grid files
    sio.savemat(tmp_path / "grid.mat", {"active_index": ai, "ghost_lr": case.gh.astype(np.int16).reshape(-1, 1)})
    gh = case.gh.astype(float)
    yf_full = np.full(3 * ai.size, np.nan)
    yf_full[np.tile(ai.ravel() > 0, 3)] = case.y0
    sio.savemat(tmp_path / "state.mat", {"yf": yf_full.reshape(-1, 1)})
    (tmp_path / "infile_forward").write_text(
        f"Et_file = 'pulse.mat';\nfinal_time = 99;\nnstep = {nstep};\nicondfile = 'state.mat';\noutfile = '';\n"
        f"outfile_prefix = 'ys_';\nmex_function = @fk_forward_Et_mex;\ngrid_file = 'grid.mat';\n"
        f"in_path = '{tmp_path}/';\nout_path = '{tmp_path}/';\n")
    path = FR.run(str(tmp_path / "infile_forward"), log=lambda *a: None)
    assert path.endswith("ys_pulse.mat")
    out = sio.loadmat(path)
    ys, yf, _ = fkadj.fk_forward_Et_mex(Et, dt, nt, case.y0, ai.ravel(), case.nrows, case.ncols, gh, nstep)
    assert out["ys"].shape == (3 * ai.size, 4)               # floor(30/7) = 4 snapshots
    assert np.array_equal(out["ys"], ys, equal_nan=True) and np.array_equal(out["yf"].ravel(), yf, equal_nan=True)
    for k in ("nrows", "ncols", "dt", "nstep", "final_time"):
        assert k in out
    assert str(out["program"][0]) == "save_fk_forward_Et" and "tau_w_plus" in str(out["param_str"][0])
