"""Seeded synthetic workloads for the benchmark configurations.

Shared physics: the reference's FK constants (src/cuda/fk_parameters_00_05_alpha_0p8_dx_0p035.h:2-21)
with h = 0.025 cm, dt = 0.02 ms, Du = 1e-3 cm^2/ms (Dv = Dw = 1e-5), no-flux walls and disc holes
of radius 2-6 cells.  Pulse: K piecewise-constant amplitudes A_k ~ U[0, 1] V/cm, passed to the
model in dV/cm (x10, the reference's unit, src/scripting/python/smooth_biphasic_pulse.py:18).
Objective: mean over active cells of u(T)^2 + lambda * integral E^2 dt with lambda = 1e-3, i.e.
S = (2/n_active) I, gamma = (1, 0, 0), alpha = 2 lambda in the reference's L0/2 + alpha LE/2.

Initial states: an S1-S2 cross-field protocol.  S1 excites a strip at the left wall; the planar
wave (about 1.2 cells/ms, action potential about 180 ms at these constants) crosses the tissue.
When its back has passed mid-width, S2 excites the lower half behind the back; the S2 wave's
free end at the half-height line curls into a spiral.  The S2 time is jittered from the seed.
With two S2 bands (rows N/8..3N/8 and 5N/8..7N/8) each band has two free ends, giving four
interacting spirals (the large-grid configuration).
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from . import geometry as G
from .api import Instance, batch_forward

H_CM = 0.025
DT_MS = 0.02
LAMBDA = 1e-3
DV_PER_V = 10.0          # model field unit dV/cm per V/cm
CV_CELLS_PER_MS = 1.2    # planar conduction velocity at h = 0.025 (measured with the oracle)
APD_MS = 180.0           # action potential duration (u > 0.5)

CONFIGS = {
    "C1": dict(N=128, holes=20, T=100.0, K=100, precision=L.F64),
    "C2": dict(N=512, holes=80, T=200.0, K=200, precision=L.F32, ckpt=50),
    "C3": dict(sizes=(256, 384, 512, 768, 1024), count=64, T=150.0, K=150, precision=L.F32, ckpt=50),
    "C5": dict(N=2048, holes=400, T=500.0, K=500, precision=L.F32, ckpt=50),
}


def model():
    return L.reference_model(h=H_CM)


def holes_mask(N: int, n_holes: int, seed: int, ncols: int | None = None) -> np.ndarray:
    return G.disc_holes(N, ncols or N, n_holes, 2, 6, seed)


def amplitudes(K: int, seed: int) -> np.ndarray:
    """K amplitudes in the model unit (dV/cm)."""
    return DV_PER_V * np.random.default_rng(1000 + seed).uniform(0.0, 1.0, K)


def rest_state(n_act: int) -> np.ndarray:
    return np.concatenate([np.zeros(n_act), np.ones(n_act), np.ones(n_act)])


def make_instance(mask: np.ndarray, T: float, K: int, precision=L.F32, ckpt=50, fwd_tile=0, adj_tile=0,
                  keep_bytes=0, lazy_workspace=False, fwd_depth=0, adj_depth=0, roll_rows=0) -> Instance:
    nt = int(round(T / DT_MS))
    n_act = int(mask.sum())
    return Instance(G.active_index_from_mask(mask), mask.shape[0], mask.shape[1], dt=DT_MS, nt=nt, model=model(),
                    pulse_mode=L.PULSE_PIECEWISE, n_amp=K, alpha=2 * LAMBDA, gamma=(1.0, 0.0, 0.0),
                    s_diag=2.0 / n_act, precision=precision, ckpt=ckpt, fwd_tile=fwd_tile,
                    adj_tile=adj_tile, keep_bytes=keep_bytes, lazy_workspace=lazy_workspace, fwd_depth=fwd_depth,
                    adj_depth=adj_depth, roll_rows=roll_rows)


def s2_region(nrows: int, ncols: int, back_col: int, bands: int) -> np.ndarray:
    """S2 stimulus: the lower half (one free end) or two horizontal bands (four free ends)."""
    s2 = np.zeros((nrows, ncols), bool)
    if bands == 1:
        s2[nrows // 2:, :back_col] = True
    else:
        for lo, hi in ((nrows // 8, 3 * nrows // 8), (5 * nrows // 8, 7 * nrows // 8)):
            s2[lo:hi, :back_col] = True
    return s2


def spiral_states(masks, seeds, chunk_ms: float = 5.0, settle_ms: float = 150.0, bands: int = 1):
    """S1-S2 spiral initial states (packed) for grids of one shape, prepared in shared launches."""
    nrows, ncols = masks[0].shape
    chunk = int(round(chunk_ms / DT_MS))
    prep = []
    for m in masks:
        inst = Instance(G.active_index_from_mask(m), nrows, ncols, dt=DT_MS, nt=chunk, model=model(),
                        pulse_mode=L.PULSE_PIECEWISE, n_amp=1, alpha=0.0, gamma=(1.0, 0.0, 0.0),
                        s_diag=1.0, precision=L.F32, ckpt=chunk)
        inst.set_state(rest_state(int(m.sum())))
        s1 = np.zeros((nrows, ncols), bool)
        s1[:, :6] = True
        inst.stimulate(s1, 1.0)
        prep.append(inst)
    t_back = APD_MS + 0.5 * ncols / CV_CELLS_PER_MS
    t_s2 = [t_back + np.random.default_rng(seed).uniform(-10.0, 10.0) for seed in seeds]
    back_col = [int(0.5 * ncols - 12) for _ in seeds]
    done = [False] * len(masks)
    t = 0.0
    t_end = max(t_s2) + settle_ms
    zero = [np.zeros(1)] * len(masks)
    while t < t_end - 1e-9:
        batch_forward(prep, zero, chunk)
        t += chunk_ms
        for k, inst in enumerate(prep):
            if not done[k] and t >= t_s2[k]:
                inst.stimulate(s2_region(nrows, ncols, back_col[k], bands), 1.0)
                done[k] = True
    states = [inst.get_state() for inst in prep]
    for inst in prep:
        inst.close()
    return states


def holes_for(N: int) -> int:
    """Disc-hole count at the C2 density (80 holes on 512^2)."""
    return max(1, int(round(80 * N * N / 512 ** 2)))


def prepare(sizes, seeds):
    """(masks, states): disc holes at the C2 density and S1-S2 spiral states for instances of the
    given square sizes (states prepared in shared launches per size)."""
    masks = [holes_mask(N, holes_for(N), s) for N, s in zip(sizes, seeds)]
    states = [None] * len(sizes)
    for N in sorted(set(sizes)):
        idx = [k for k, n in enumerate(sizes) if n == N]
        st = spiral_states([masks[k] for k in idx], [seeds[k] for k in idx])
        for k, y in zip(idx, st):
            states[k] = y
    return masks, states


def phase_singularities(y: np.ndarray, mask: np.ndarray) -> int:
    """Spiral cores of a packed state: clusters of 2x2 plaquettes of active cells around which the
    phase atan2(v - 1/2, u - 1/2) winds by a full turn."""
    from scipy import ndimage

    n = int(mask.sum())
    u = np.zeros(mask.shape)
    v = np.zeros(mask.shape)
    u[mask] = y[:n]
    v[mask] = y[n:2 * n]
    th = np.arctan2(v - 0.5, u - 0.5)

    def d(a, b):
        return np.angle(np.exp(1j * (b - a)))

    a, b, c, e = th[:-1, :-1], th[:-1, 1:], th[1:, 1:], th[1:, :-1]
    wind = np.rint((d(a, b) + d(b, c) + d(c, e) + d(e, a)) / (2 * np.pi)).astype(int)
    ok = mask[:-1, :-1] & mask[:-1, 1:] & mask[1:, 1:] & mask[1:, :-1]
    _, count = ndimage.label((wind != 0) & ok, structure=np.ones((3, 3)))
    return int(count)
