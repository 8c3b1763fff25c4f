"""Seeded problem builders shared by the tests (small sizes the oracle finishes in seconds)."""
from __future__ import annotations

import os

import numpy as np

import oracle as O
from fkadj import geometry as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def golden():
    import json

    d = dict(np.load(os.path.join(ROOT, "tests", "golden", "ref256.npz")))
    with open(os.path.join(ROOT, "tests", "golden", "ref256.json")) as f:
        meta = json.load(f)
    return d, meta


class Case:
    """A grid, a packed initial state and the objective pieces."""

    def __init__(self, mask, y_planes, h=0.035):
        self.mask = mask
        self.nrows, self.ncols = mask.shape
        self.ai = G.active_index_from_mask(mask)
        self.gh = G.ghost_lr_from_mask(mask)
        self.y0 = np.concatenate([y_planes[v][mask] for v in range(3)])
        self.n_act = int(mask.sum())
        self.h = h

    def planes(self):
        return O.num2nan(self.y0, self.ai, self.nrows, self.ncols)

    def gh_plane(self, model):
        g = np.full(self.nrows * self.ncols, np.nan)
        g[self.ai > 0] = model.Du * 2 / model.h * self.gh
        return g.reshape(self.nrows, self.ncols)


def crop_case(r0, c0, H, W):
    """Window of the reference's 256x256 spiral state (realistic fronts and holes)."""
    d, _ = golden()
    planes = O.num2nan(d["y0"], d["active_index"], 256, 256)[:, r0:r0 + H, c0:c0 + W].copy()
    mask = G.fill_isthmuses(~np.isnan(planes[0]))
    return Case(mask, np.nan_to_num(planes))


def disc_case(H, W, n_holes, seed, rmin=1, rmax=3):
    """Seeded disc holes and a curved excitation front (u), recovery gates (v, w)."""
    mask = G.disc_holes(H, W, n_holes, rmin, rmax, seed)
    rng = np.random.default_rng(seed + 1)
    I, J = np.mgrid[0:H, 0:W]
    front = 0.35 * W + 0.15 * W * np.sin(2 * np.pi * I / H + rng.uniform(0, 6))
    with np.errstate(over="ignore"):   # far behind the front exp overflows to inf and u is 0
        u = 1.0 / (1.0 + np.exp((J - front) / 1.5))
    v = 1.0 - 0.9 * u
    w = 0.95 - 0.3 * u
    return Case(mask, np.stack([u, v, w]))


def oracle_grad_knots(case, dt, Et, gamma, alpha, S, model=None, ckpt=50):
    """Oracle gradient in knot mode: (gradE, Lcost, L0, LE)."""
    nt = len(Et) - 1
    return O.grad_ref(dt, Et, nt, case.y0, case.ai, case.nrows, case.ncols, case.gh, gamma, G.csc_arrays(S), alpha,
                      model=model, ckpt=ckpt)


def oracle_grad_knots_state_rounded(dt, Et, y_ini, active_index, nrows, ncols, gh_num, gamma, S, alpha, model=None,
                                    ckpt=50, round_state=True):
    """The knot-mode oracle gradient (gradE, Lcost) assembled from orc_grad; with round_state the
    trajectory is rounded to FP32 after every step and FP64 arithmetic is kept otherwise, which gives
    the error that FP32 storage of the state alone puts into gradE."""
    model = model or O.reference_model()
    nt = len(Et) - 1
    planes = O.num2nan(y_ini, active_index, nrows, ncols)
    ai = np.asarray(active_index)
    gh = np.full(nrows * ncols, np.nan)
    gh[ai > 0] = model.Du * 2 / model.h * np.asarray(gh_num, float)[ai[ai > 0] - 1]
    Ef, Ea = O.knot_tables(dt, Et)
    zgh, L0, _ = O.grad(model, planes, gh.reshape(nrows, ncols), dt, Ef, Ea, active_index, G.csc_arrays(S), gamma,
                        ckpt=ckpt, round_state=round_state)
    Et = np.asarray(Et, float)
    LE = dt * float(np.sum(Et * Et))
    return alpha * Et - zgh, 0.5 * L0 + 0.5 * alpha * LE


def oracle_grad_piecewise(case, dt, nt, A, gamma, alpha, S, model):
    """Oracle gradient in piecewise mode: E constant over each step; dJ/dA_k by trapezoid of
    alpha*E - z.gh over the steps of piece k.  Returns (gradA, Lcost, L0, LE)."""
    K = len(A)
    P = nt // K
    Ef = np.repeat(np.asarray(A, float), 3 * P)
    Ea = Ef.copy()
    zgh, L0, _ = O.grad(model, case.planes(), case.gh_plane(model), dt, Ef, Ea, case.ai, G.csc_arrays(S), gamma, ckpt=50)
    Tp = P * dt
    g = np.array([alpha * A[k] * Tp - dt * np.sum(0.5 * (zgh[k * P:(k + 1) * P] + zgh[k * P + 1:(k + 1) * P + 1]))
                  for k in range(K)])
    LE = Tp * float(np.sum(np.square(A)))
    return g, 0.5 * L0 + 0.5 * alpha * LE, L0, LE


def relerr(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
