"""Tile autotuner and profiling driver for the fused FP32 kernels.

The fused kernels come in tile configurations (64 x TY cells per workgroup of NT threads):
1 = 64x16/256, 2 = 64x32/256, 3 = 64x32/512, 4 = 64x16/512, 5 = 64x32/1024, 6 = 64x8/512,
7 = 64x8/256.  The forward tile never
changes a result; the adjoint tile regroups the per-tile z.gh partial sums (rounding-level changes),
so it is fixed per grid size.  Time blocking: at these launch sizes one step per launch (the kernels
are bound by vector-ALU issue, not by HBM, so a deeper halo only adds recomputation); the library
takes two forward-sweep steps per launch on its own for launches under 1M cells on the built-in
forward tile, which a tuned forward tile turns off.

    python -m fkadj.autotune --grid 512 --batch 16            # sweep one size, print a table
    python -m fkadj.autotune --sizes 256 384 512 768 1024 --write
                                    # sweep each grid size at about 4M cells per launch and store
                                    # the best pair per size in tuned_tiles.json
    python -m fkadj.autotune --grid 512 --batch 16 --once 3 4 # one evaluation (profiler target)

The synthetic batch uses the benchmark's geometry (80 disc holes per 512^2, h = 0.025 cm,
dt = 0.02 ms, 200 amplitudes) with a cheap initial state (half-plane excitation) and T = 20 ms.
Instances pick their tiles by grid size with lookup(N); the forward tile changes no result and
the adjoint tile only the grouping of the FP64 partial sums.
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

from . import _lib as L
from . import workloads as W
from .api import batch_grad

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_tiles.json")
FWD_CFGS = (1, 2, 3, 4, 5, 6, 7)
ADJ_CFGS = (1, 3, 4, 5, 6, 7)


def synthetic_batch(N: int, B: int, T: float, fwd_tile: int, adj_tile: int, holes: int | None = None):
    insts = []
    holes = holes if holes is not None else max(1, round(80 * N * N / 512 ** 2))
    for s in range(B):
        m = W.holes_mask(N, holes, s)
        inst = W.make_instance(m, T, 200, precision=L.F32, ckpt=50, fwd_tile=fwd_tile, adj_tile=adj_tile)
        inst.set_state(W.rest_state(int(m.sum())))
        s1 = np.zeros(m.shape, bool)
        s1[:, : N // 2] = True
        inst.stimulate(s1, 1.0)
        insts.append(inst)
    return insts, [W.amplitudes(200, s) for s in range(B)]


def time_pair(N: int, B: int, T: float, ft: int, at: int, reps: int = 1) -> dict:
    insts, A = synthetic_batch(N, B, T, ft, at)
    try:
        batch_grad(insts, A)
        L.set_profiling(True)
        best = None
        for _ in range(reps):
            t0 = time.perf_counter()
            batch_grad(insts, A)
            el = time.perf_counter() - t0
            tm = L.last_timing()
            if best is None or el < best["eval_s"]:
                best = {"eval_s": el, "fwd_us": 1e3 * tm["fwd_kernel_ms"], "adj_us": 1e3 * tm["adj_kernel_ms"]}
        L.set_profiling(False)
        return best
    finally:
        for i in insts:
            i.close()


def load_table(path: str = TABLE) -> dict:
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def lookup(N: int) -> tuple[int, int]:
    """Tuned (fwd_tile, adj_tile) for an N^2 grid: the entry of the nearest tuned size, or (0, 0)
    (the library's built-in choice) when nothing is tuned."""
    t = load_table()
    if not t:
        return 0, 0
    key = min(t, key=lambda k: abs(int(k) - N))
    return int(t[key]["fwd_tile"]), int(t[key]["adj_tile"])


def sweep(N: int, B: int, T: float) -> dict:
    """Forward tile first (adjoint at the built-in choice), then the adjoint tile."""
    res_f = {ft: time_pair(N, B, T, ft, 0) for ft in FWD_CFGS}
    for ft, r in res_f.items():
        print(f"{N}^2 x{B} fwd_tile={ft}: forward {r['fwd_us']:.1f} us/launch", flush=True)
    best_f = min(res_f, key=lambda c: res_f[c]["fwd_us"])
    res_a = {at: time_pair(N, B, T, best_f, at) for at in ADJ_CFGS}
    for at, r in res_a.items():
        print(f"{N}^2 x{B} adj_tile={at}: adjoint {r['adj_us']:.1f} us/launch, eval {1e3 * r['eval_s']:.1f} ms",
              flush=True)
    best_a = min(res_a, key=lambda c: res_a[c]["adj_us"])
    print(f"{N}^2 x{B} best: fwd_tile={best_f} adj_tile={best_a}", flush=True)
    return {"fwd_tile": best_f, "adj_tile": best_a, "batch": B, "fwd_us": res_f[best_f]["fwd_us"],
            "adj_us": res_a[best_a]["adj_us"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", type=int, default=512)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--sizes", type=int, nargs="+", help="sweep these grid sizes at --cells cells per launch")
    ap.add_argument("--cells", type=int, default=1 << 22)
    ap.add_argument("--T", type=float, default=20.0)
    ap.add_argument("--once", type=int, nargs=2, metavar=("FWD", "ADJ"))
    ap.add_argument("--write", action="store_true")
    ap.add_argument("--table", default=TABLE, help="table file read and written by --write")
    args = ap.parse_args()
    if args.once:
        insts, A = synthetic_batch(args.grid, args.batch, args.T, *args.once)
        t0 = time.perf_counter()
        batch_grad(insts, A)
        print(f"one evaluation: {time.perf_counter() - t0:.3f} s", flush=True)
        return
    sizes = args.sizes or [args.grid]
    t = load_table(args.table)
    for N in sizes:
        B = args.batch if not args.sizes else min(64, max(1, round(args.cells / (N * N))))   # <= 64 per launch
        t[str(N)] = sweep(N, B, args.T)
    if args.write:
        with open(args.table, "w") as f:
            json.dump(t, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
