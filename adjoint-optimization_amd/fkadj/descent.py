"""Gradient-descent driver over a resident instance, with checkpoint files and resume.

Restates the reference's driver ``src/cuda/simple_grad_desc.m:16-107``: read the run parameters,
load the seed pulse, the initial state, the grid and S, then for i = 1..ns evaluate the gradient,
record Lcost / L0 / LE in vectors of length nmod (slot (i-1) mod nmod), write a checkpoint file every
nmod steps (with E as it was *before* this step's update, as the reference stores it) and update
E <- E - ds * gradE.

Differences by design: the instance (geometry, state, S) is created once and stays on the device;
each step moves only E and the gradient.  Checkpoint files are MAT v5 files with the reference's
variable names (scipy.io.savemat), plus the model's provenance string and the step number, so a
run can resume from its last checkpoint: the resumed run continues with E - ds * gradE of that file
and its later checkpoints equal those of an uninterrupted run bit for bit.

    python -m fkadj.descent infile_grad [--precision f64|f32] [--no-resume]
"""
from __future__ import annotations

import argparse
import glob
import os
import re
import socket
import time

import numpy as np

from . import _lib as L
from .api import Instance
from .config import GradDescentConfig

PROGRAM = "simple_grad_desc"


def _loadmat(path, names):
    import scipy.io as sio

    d = sio.loadmat(path, mat_dtype=True, variable_names=list(names))
    missing = [n for n in names if n not in d]
    if missing:
        raise KeyError(f"{path}: missing {missing}")
    return d


def load_inputs(cfg: GradDescentConfig) -> dict:
    """Seed pulse, packed initial state, grid and S from the run's MAT files (data only), with the
    reference driver's conventions (simple_grad_desc.m:55-72)."""
    p = cfg.in_path
    Et = _loadmat(os.path.join(p, cfg.Et_seed_file), ["Et"])["Et"].ravel().astype(float)
    grid = _loadmat(os.path.join(p, cfg.grid_file), ["active_index", "ghost_lr"])
    ai_m = grid["active_index"]
    nrows, ncols = ai_m.shape
    active_index = ai_m.ravel().astype(np.int32)
    ghost_lr = grid["ghost_lr"].ravel().astype(float)
    yf = _loadmat(os.path.join(p, cfg.icondfile), ["yf"])["yf"].ravel()
    y0 = yf[np.concatenate([active_index > 0] * 3)].astype(float)
    S = _loadmat(os.path.join(p, cfg.S_file), ["S"])["S"]
    return dict(Et=Et, y0=y0, active_index=active_index, nrows=int(nrows), ncols=int(ncols), ghost_lr=ghost_lr, S=S)


def checkpoint_name(time_str: str, i: int) -> str:
    return f"{PROGRAM}_step_{time_str}_{i}.mat"


def latest_checkpoint(out_path: str, jobid: str):
    """(path, step) of this job's checkpoint with the highest step, or (None, 0)."""
    best, best_i = None, 0
    for f in glob.glob(os.path.join(out_path, f"{PROGRAM}_step_*_{jobid}_*.mat")):
        m = re.search(r"_(\d+)\.mat$", f)
        if m and int(m.group(1)) > best_i:
            best, best_i = f, int(m.group(1))
    return best, best_i


class Descent:
    """The descent loop on one resident instance (FP64 reference path by default)."""

    def __init__(self, cfg: GradDescentConfig, inputs: dict, precision=L.F64, ckpt: int = 50):
        self.cfg = cfg
        self.inputs = inputs
        if len(inputs["Et"]) != cfg.nt + 1:
            raise ValueError(f"seed pulse has {len(inputs['Et'])} samples, nt+1 = {cfg.nt + 1}")
        self.inst = Instance(inputs["active_index"], inputs["nrows"], inputs["ncols"], dt=cfg.dt, nt=cfg.nt,
                             ghost_lr=inputs["ghost_lr"], alpha=cfg.alpha, gamma=cfg.gamma, S=inputs["S"],
                             precision=precision, ckpt=ckpt)
        self.inst.set_state(inputs["y0"])
        self.provenance = L.param_string(self.inst.problem.model)
        self.precision = precision

    def close(self):
        self.inst.close()

    def run(self, time_str: str | None = None, resume: bool = True, log=print, stop_after: int | None = None):
        """Runs steps up to cfg.ns (or stop_after), writing checkpoints to cfg.out_path.  Returns
        (E after the last update, list of written checkpoint paths)."""
        cfg = self.cfg
        os.makedirs(cfg.out_path, exist_ok=True)
        E = self.inputs["Et"].copy()
        start = 1
        if resume:
            path, step = latest_checkpoint(cfg.out_path, cfg.jobid)
            if path:
                d = _loadmat(path, ["Et", "gradE", "ds", "time_str"])
                E = d["Et"].ravel() - float(d["ds"].ravel()[0]) * d["gradE"].ravel()
                start = step + 1
                time_str = str(d["time_str"][0])
                log(f"resuming after step {step} from {path}")
        if time_str is None:
            time_str = time.strftime("%Y-%b-%d-%H-%M-%S").lower() + "_" + cfg.jobid
        Lc_vec, L0_vec, LE_vec = np.zeros(cfg.nmod), np.zeros(cfg.nmod), np.zeros(cfg.nmod)
        written = []
        last = cfg.ns if stop_after is None else min(cfg.ns, stop_after)
        for i in range(start, last + 1):
            t0 = time.perf_counter()
            grad, Lc, L0, LE = self.inst.grad(E)
            ip = (i - 1) % cfg.nmod
            Lc_vec[ip], L0_vec[ip], LE_vec[ip] = Lc, L0, LE
            log(f"step {i} of {cfg.ns}: Lcost = {Lc:.15g} ({time.perf_counter() - t0:.3f} s)")
            if i % cfg.nmod == 0:
                path = os.path.join(cfg.out_path, checkpoint_name(time_str, i))
                self._save(path, E, grad, Lc, Lc_vec, L0_vec, LE_vec, i, time_str)
                written.append(path)
                log(f"file \"{path}\" written")
            E = E - cfg.ds * grad
        return E, written

    def _save(self, path, E, grad, Lc, Lc_vec, L0_vec, LE_vec, i, time_str):
        import scipy.io as sio

        cfg = self.cfg
        # the reference's variables (simple_grad_desc.m:93-95) plus provenance and resume data
        sio.savemat(path, {
            "alpha": cfg.alpha, "dt": cfg.dt, "final_time": cfg.final_time, "Et_seed_file": cfg.Et_seed_file,
            "Et": E.reshape(-1, 1), "gradE": grad.reshape(-1, 1), "Lcost": Lc, "ds": cfg.ds,
            "Lcost_vec": Lc_vec.reshape(-1, 1).copy(), "L0_vec": L0_vec.reshape(-1, 1).copy(),
            "LE_vec": LE_vec.reshape(-1, 1).copy(), "gamma": np.asarray(cfg.gamma, float).reshape(1, -1),
            "program": PROGRAM, "hostname": socket.gethostname(), "nmod": float(cfg.nmod),
            "icondfile": cfg.icondfile, "grid_file": cfg.grid_file, "S_file": cfg.S_file,
            "mex_function": "libfkadj " + ("FP64" if self.precision == L.F64 else "FP32"),
            "param_str": self.provenance, "step": float(i), "time_str": time_str,
        }, do_compression=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("infile")
    ap.add_argument("--precision", choices=("f64", "f32"), default="f64")
    ap.add_argument("--no-resume", action="store_true")
    args = ap.parse_args()
    cfg = GradDescentConfig.from_infile(args.infile, log=print)
    print(f"alpha={cfg.alpha:f} ds={cfg.ds:f} ns={cfg.ns} final_time={cfg.final_time:f} gamma={cfg.gamma}")
    d = Descent(cfg, load_inputs(cfg), precision=L.F64 if args.precision == "f64" else L.F32)
    try:
        d.run(resume=not args.no_resume)
    finally:
        d.close()


if __name__ == "__main__":
    main()
