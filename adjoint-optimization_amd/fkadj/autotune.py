"""Launch-shape autotuner for the FP32 path: tile shape, time-blocking depth and strip height.

None of these choices changes a result (the tests check that bit for bit); they only change how
the work is cut into launches.  Per grid size N the tuner times whole gradient evaluations of a
synthetic instance (the benchmark's geometry: 80 disc holes per 512^2, h = 0.025 cm, dt = 0.02 ms,
200 amplitudes, half-plane excitation, T = 20 ms by default) under each candidate and stores the
fastest in tuned_tiles.json, at two launch scales:

* "single": one instance per launch, on the tiled kernels.  Candidates: forward tile and depth
  (64x8/512 cells/threads with one or two steps per launch, 64x16/512 with one), then adjoint tile
  and depth (64x8/512 with one or two steps per launch, 64x16/512 with one), the adjoint swept with
  the best forward choice.
* "batch": instances batched to about 4M cells per launch.  Candidates: the row-rolling kernels
  (pitches 256, 384, 512, 768, 1024) with strip rows chosen per launch by the library or fixed at 8,
  16 or 32, and the tiled kernels (64x16/512) for every pitch.  A fixed strip height is stored only
  where it beats the library's per-launch choice by more than 5% (the per-launch choice also follows
  the batch's size, which a fixed height measured on one batch does not).

Tile ids: 1 = 64x16/256, 2 = 64x32/256, 3 = 64x32/512, 4 = 64x16/512, 5 = 64x32/1024,
6 = 64x8/512, 7 = 64x8/256.

    python -m fkadj.autotune --sizes 128 256 512 --write      # sweep, store in tuned_tiles.json
    python -m fkadj.autotune --grid 512 --scale single         # sweep one size, print only
    python -m fkadj.autotune --grid 512 --batch 16 --once 0 4  # one evaluation (profiler target)

instance_kwargs(N, batched) returns the stored choice as Instance / make_instance keyword
arguments (the library's built-in rules where nothing is stored).
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
ROLL_PITCHES = (256, 384, 512, 768, 1024)

# single-instance candidates: (fwd_tile, fwd_depth) and (adj_tile, adj_depth)
FWD_SINGLE = ((6, 2), (6, 1), (4, 1))
ADJ_SINGLE = ((6, 1), (4, 1), (6, 2))
# batched candidates: fixed strip rows of the rolling kernels (besides the library's per-launch choice)
ROLL_ROWS = (8, 16, 32)


def synthetic_batch(N: int, B: int, T: float, holes: int | None = None, **kw):
    """B instances of an N^2 grid (seeded disc holes, half-plane excitation), Instance keywords kw."""
    insts = []
    holes = holes if holes is not None else max(1, round(80 * N * N / 512 ** 2))
    for s in range(B):
        m = W.holes_mask(N, holes, s)
        inst = W.make_instance(m, T, 200, precision=L.F32, ckpt=50, **kw)
        inst.set_state(W.rest_state(int(m.sum())))
        s1 = np.zeros(m.shape, bool)
        s1[:, : N // 2] = True
        inst.stimulate(s1, 1.0)
        insts.append(inst)
    return insts, [W.amplitudes(200, s) for s in range(B)]


def time_config(N: int, B: int, T: float, reps: int = 2, roll_mode: int = 1, **kw) -> dict:
    """Best of `reps` wall-clock batched evaluations (after one warm-up) under the given Instance
    keywords; the rolling kernels are allowed (roll_mode 1, the library default) or not (0)."""
    prev = L.set_rolling(roll_mode, 0)
    insts, A = synthetic_batch(N, B, T, **kw)
    try:
        batch_grad(insts, A)
        best = None
        for _ in range(reps):
            t0 = time.perf_counter()
            res = batch_grad(insts, A)
            el = time.perf_counter() - t0
            best = el if best is None else min(best, el)
        return {"ms": 1e3 * best, "grad0": res[0][0]}
    finally:
        for i in insts:
            i.close()
        L.set_rolling(prev, 0)


def sweep_single(N: int, T: float) -> dict:
    """Forward (tile, depth) first with the built-in adjoint, then adjoint (tile, depth)."""
    fw = {}
    for ft, fd in FWD_SINGLE:
        fw[ft, fd] = time_config(N, 1, T, fwd_tile=ft, fwd_depth=fd)
        print(f"{N}^2 single fwd_tile={ft} fwd_depth={fd}: {fw[ft, fd]['ms']:.1f} ms/eval", flush=True)
    bf = min(fw, key=lambda c: fw[c]["ms"])
    ad = {}
    for at, adp in ADJ_SINGLE:
        ad[at, adp] = time_config(N, 1, T, fwd_tile=bf[0], fwd_depth=bf[1], adj_tile=at, adj_depth=adp)
        print(f"{N}^2 single adj_tile={at} adj_depth={adp}: {ad[at, adp]['ms']:.1f} ms/eval", flush=True)
    ba = min(ad, key=lambda c: ad[c]["ms"])
    ref = fw[FWD_SINGLE[0]]["grad0"]
    same = all(np.array_equal(r["grad0"], ref) for r in list(fw.values()) + list(ad.values()))
    print(f"{N}^2 single best: fwd {bf} adj {ba}; all candidates give the same gradient: {same}", flush=True)
    return {"fwd_tile": bf[0], "fwd_depth": bf[1], "adj_tile": ba[0], "adj_depth": ba[1], "ms": ad[ba]["ms"],
            "candidates_ms": {**{f"fwd {k}": v["ms"] for k, v in fw.items()},
                              **{f"adj {k}": v["ms"] for k, v in ad.items()}},
            "same_gradient": same}


def sweep_batch(N: int, T: float, cells: int = 1 << 22) -> dict:
    """The rolling kernels (strip rows chosen per launch, or fixed) against the tiled kernels."""
    B = min(64, max(1, round(cells / (N * N))))
    pitch = (N + 63) // 64 * 64
    res = {"tiled": time_config(N, B, T, roll_mode=0, fwd_tile=4, adj_tile=4)}
    print(f"{N}^2 x{B} tiled 64x16/512: {res['tiled']['ms']:.1f} ms/eval", flush=True)
    if pitch in ROLL_PITCHES:
        res["auto"] = time_config(N, B, T, roll_mode=2, fwd_tile=4, adj_tile=4)
        print(f"{N}^2 x{B} rolling, strip rows chosen per launch: {res['auto']['ms']:.1f} ms/eval", flush=True)
        for hs in ROLL_ROWS:
            res[hs] = time_config(N, B, T, roll_mode=2, fwd_tile=4, adj_tile=4, roll_rows=hs)
            print(f"{N}^2 x{B} rolling strips of {hs} rows: {res[hs]['ms']:.1f} ms/eval", flush=True)
    best = min(res, key=lambda c: res[c]["ms"])
    if isinstance(best, int) and res[best]["ms"] > 0.95 * res["auto"]["ms"]:
        best = "auto"   # a fixed height only where it clearly beats the per-launch choice
    same = all(np.array_equal(r["grad0"], res["tiled"]["grad0"]) for r in res.values())
    desc = {"tiled": "tiled", "auto": "rolling, strip rows chosen per launch"}.get(best, f"rolling, {best}-row strips")
    print(f"{N}^2 x{B} best: {desc}; same gradient: {same}", flush=True)
    return {"batch": B, "choice": "tiled" if best == "tiled" else "roll",
            "roll_rows": best if isinstance(best, int) else 0, "fwd_tile": 4, "adj_tile": 4, "ms": res[best]["ms"],
            "candidates_ms": {(k if isinstance(k, str) else f"roll {k}"): v["ms"] for k, v in res.items()},
            "same_gradient": same}


def load_table(path: str = TABLE) -> dict:
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def instance_kwargs(N: int, batched: bool, path: str = TABLE) -> dict:
    """Instance keywords of the tuned choice for an N^2 grid at the given launch scale (the entry of
    the nearest tuned size), or {} (the library's rules) when nothing is tuned for that scale."""
    t = {k: v for k, v in load_table(path).items() if ("batch" if batched else "single") in v}
    if not t:
        return {}
    e = t[min(t, key=lambda k: abs(int(k) - N))]["batch" if batched else "single"]
    if batched:
        if e["roll_rows"] > 0:
            return {"roll_rows": int(e["roll_rows"])}
        if e.get("choice") == "roll":
            return {}   # rolling kernels with strip rows chosen per launch
        return {"fwd_tile": int(e["fwd_tile"]), "adj_tile": int(e["adj_tile"])}
    return {k: int(e[k]) for k in ("fwd_tile", "fwd_depth", "adj_tile", "adj_depth")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", type=int, default=512)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--sizes", type=int, nargs="+", help="sweep these grid sizes")
    ap.add_argument("--scale", choices=("single", "batch", "both"), default="both")
    ap.add_argument("--T", type=float, default=20.0)
    ap.add_argument("--once", type=int, nargs=2, metavar=("FWD", "ADJ"),
                    help="one batched evaluation with these forward and adjoint tiles")
    ap.add_argument("--write", action="store_true")
    ap.add_argument("--table", default=TABLE, help="table file read and written by --write")
    args = ap.parse_args()
    if args.once:
        insts, A = synthetic_batch(args.grid, args.batch, args.T, fwd_tile=args.once[0], adj_tile=args.once[1])
        t0 = time.perf_counter()
        batch_grad(insts, A)
        print(f"one evaluation: {time.perf_counter() - t0:.3f} s", flush=True)
        return
    t = load_table(args.table)
    for N in args.sizes or [args.grid]:
        e = t.setdefault(str(N), {})
        if args.scale in ("single", "both"):
            e["single"] = sweep_single(N, args.T)
        if args.scale in ("batch", "both"):
            e["batch"] = sweep_batch(N, args.T)
    if args.write:
        with open(args.table, "w") as f:
            json.dump(t, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
