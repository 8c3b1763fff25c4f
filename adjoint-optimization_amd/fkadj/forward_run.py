"""Forward snapshot run: the reference's driver ``src/cuda/save_fk_forward_Et.m:22-109``.

Reads the run parameters (``infile_forward`` names, literal values only), loads the field E(t) (or
uses a zero field), the grid and the initial state, runs the FP64 forward solve on the GPU with a
snapshot every ``nstep`` steps (``fkadj_forward_ref``, the drop-in for ``fk_forward_Et_comp``) and
writes a MAT v5 file with the reference's variables: ``ys`` (3*nrows*ncols x ns, snapshot ks in
column ks, NaN at holes), ``yf`` (the final state), the grid size, the inputs' names, ``dt``,
``nstep``, ``final_time``, the model's provenance string, the program and the host name.

Differences by design: the grid size comes from the grid file's ``active_index`` (the reference
hard-codes 256 x 256, save_fk_forward_Et.m:87-88), and the parameter file is parsed, not evaluated.

    python -m fkadj.forward_run infile_forward
"""
from __future__ import annotations

import argparse
import os
import socket

import numpy as np

from .api import fk_forward_Et_mex
from .config import float_integer, read_infile

PROGRAM = "save_fk_forward_Et"

# variables of src/cuda/save_fk_forward_Et.m:22-31, in the reference's order
FORWARD_VARS = ("Et_file", "final_time", "nstep", "icondfile", "outfile", "outfile_prefix", "mex_function",
                "grid_file", "in_path", "out_path")


def _loadmat(path, names=None):
    import scipy.io as sio

    return sio.loadmat(path, mat_dtype=False, variable_names=list(names) if names else None)


def load_field(v: dict):
    """(Et, dt, nt, alpha or None) from the Et file, or a zero field with dt = 0.1 when Et_file is
    empty (save_fk_forward_Et.m:44-63)."""
    if v["Et_file"]:
        d = _loadmat(os.path.join(v["in_path"], v["Et_file"]))
        alpha = float(np.ravel(d["alpha"])[0]) if "alpha" in d else None
        dt = float(np.ravel(d["dt"])[0])
        final_time = float(np.ravel(d["final_time"])[0])
        Et = np.ravel(d["Et"]).astype(float)
        nt = float_integer(final_time / dt)
        return Et, dt, nt, final_time, alpha
    dt = 0.1
    final_time = float(v["final_time"])
    nt = float_integer(final_time / dt)
    return np.zeros(nt + 1), dt, nt, final_time, None


def load_grid_and_state(v: dict):
    """(active_index, nrows, ncols, ghost_lr, y0): the packed initial state is yf at the active cells
    of the grid or the file's own y0 (save_fk_forward_Et.m:67-85)."""
    grid = _loadmat(os.path.join(v["in_path"], v["grid_file"]), ["active_index", "ghost_lr"])
    ai_m = grid["active_index"]
    if ai_m.dtype != np.int32:
        raise TypeError('Variable "active_index" must be of type int32')
    nrows, ncols = ai_m.shape
    active_index = ai_m.ravel()
    ghost_lr = grid["ghost_lr"].ravel().astype(float)
    st = _loadmat(os.path.join(v["in_path"], v["icondfile"]))
    if "yf" in st:
        y0 = st["yf"].ravel()[np.concatenate([active_index > 0] * 3)].astype(float)
    else:
        y0 = st["y0"].ravel().astype(float)
    return active_index, int(nrows), int(ncols), ghost_lr, y0


def output_name(v: dict) -> str:
    """outfile, or outfile_prefix + Et_file when outfile is empty (save_fk_forward_Et.m:90-101)."""
    if v["outfile"]:
        return v["outfile"]
    if not v["Et_file"]:
        raise ValueError("Et_file is empty. Cannot generate outfile")
    if not v["outfile_prefix"]:
        raise ValueError("outfile_prefix is empty. Cannot generate outfile")
    return v["outfile_prefix"] + v["Et_file"]


def run(infile: str, log=print) -> str:
    """Runs the forward snapshot job of one parameter file; returns the path of the written file."""
    import scipy.io as sio

    v = read_infile(infile, names=FORWARD_VARS, log=log)
    nstep = int(v["nstep"])
    Et, dt, nt, final_time, alpha = load_field(v)
    if alpha is not None:
        log(f"alpha={alpha}")
    active_index, nrows, ncols, ghost_lr, y0 = load_grid_and_state(v)
    outfile = output_name(v)
    ys, yf, param_str = fk_forward_Et_mex(Et, dt, nt, y0, active_index, nrows, ncols, ghost_lr, nstep)
    path = os.path.join(v["out_path"], outfile)
    sio.savemat(path, {
        "ys": ys, "yf": yf.reshape(-1, 1), "nrows": float(nrows), "ncols": float(ncols),
        "icondfile": v["icondfile"], "dt": dt, "nstep": float(nstep), "final_time": final_time,
        "grid_file": v["grid_file"], "param_str": param_str, "Et_file": v["Et_file"], "program": PROGRAM,
        "mex_function": "libfkadj fkadj_forward_ref (FP64)", "hostname": socket.gethostname(),
    }, do_compression=True)
    log(f'Saved file "{path}"')
    return path


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("infile", nargs="?", default="infile_forward")
    run(ap.parse_args().infile)


if __name__ == "__main__":
    main()
