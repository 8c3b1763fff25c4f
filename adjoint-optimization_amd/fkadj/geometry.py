"""Tissue geometry for the Fenton-Karma grid: hole masks, ``active_index``, ``ghost_lr``,
the mirror Laplacian, the gradient-penalty matrix ``S = D^T D`` and seeded disc-hole layouts.

Grid convention (the reference's C layout, ``src/cuda/fk_grad_Et_mono_PARAMETERS.cu:25-28``):
cell ``k = i*ncols + j``; ``j`` is the contiguous axis.  ``active_index[k]`` is the 1-based
packed index of cell ``k`` or 0 for a hole (``src/cuda/fk_grad_Et_mono_PARAMETERS.cu:1533-1551``);
packed state vectors list active cells in increasing ``k``.

Everything here is host-side set-up, run once per problem.
"""
from __future__ import annotations

import numpy as np

try:  # scipy is used for the sparse operators only
    import scipy.sparse as sp
except Exception:  # pragma: no cover
    sp = None


def active_index_from_mask(mask: np.ndarray) -> np.ndarray:
    """1-based packed index per cell, 0 at holes (row-major ``k = i*ncols + j``)."""
    m = np.asarray(mask, dtype=bool).ravel()
    ai = np.zeros(m.size, dtype=np.int32)
    ai[m] = np.arange(1, int(m.sum()) + 1, dtype=np.int32)
    return ai


def mask_from_active_index(active_index: np.ndarray, nrows: int, ncols: int) -> np.ndarray:
    return (np.asarray(active_index).reshape(nrows, ncols) > 0)


def _shift(mask: np.ndarray, di: int, dj: int) -> np.ndarray:
    """present[i, j] = mask[i+di, j+dj], False outside the grid (walls count as missing)."""
    n0, n1 = mask.shape
    out = np.zeros_like(mask)
    src_i = slice(max(di, 0), n0 + min(di, 0))
    dst_i = slice(max(-di, 0), n0 + min(-di, 0))
    src_j = slice(max(dj, 0), n1 + min(dj, 0))
    dst_j = slice(max(-dj, 0), n1 + min(-dj, 0))
    out[dst_i, dst_j] = mask[src_i, src_j]
    return out


def ghost_lr_from_mask(mask: np.ndarray) -> np.ndarray:
    """Packed ``ghost_lr`` = [cell i+1 missing] - [cell i-1 missing] along the slow (C ``i``) axis.

    ``gh = 2*Du/h*ghost_lr`` is the virtual-electrode coupling of a uniform field
    (``src/cuda/fk_grad_Et_mono_PARAMETERS.cu:1588-1595``).  Walls count as missing.
    """
    m = np.asarray(mask, dtype=bool)
    plus_missing = ~_shift(m, 1, 0)
    minus_missing = ~_shift(m, -1, 0)
    g = plus_missing.astype(np.int8) - minus_missing.astype(np.int8)
    return g[m].astype(np.float64)


def isthmus_cells(mask: np.ndarray) -> np.ndarray:
    """Active cells whose two neighbours on one axis are both missing (the mirror rule is undefined)."""
    m = np.asarray(mask, dtype=bool)
    bad_i = ~_shift(m, 1, 0) & ~_shift(m, -1, 0)
    bad_j = ~_shift(m, 0, 1) & ~_shift(m, 0, -1)
    return m & (bad_i | bad_j)


def fill_isthmuses(mask: np.ndarray) -> np.ndarray:
    """Turn active cells with both axial neighbours missing into holes until none are left."""
    m = np.asarray(mask, dtype=bool).copy()
    while True:
        bad = isthmus_cells(m)
        if not bad.any():
            return m
        m &= ~bad


def lap_matrix(mask: np.ndarray):
    """Packed mirror Laplacian (grid units, diagonal -4): the operator of
    ``lap_func`` (``src/cuda/fk_grad_Et_mono_PARAMETERS.cu:244-440``).

    Along each axis a missing neighbour (hole or wall) is replaced by the opposite one.
    """
    m = np.asarray(mask, dtype=bool)
    if isthmus_cells(m).any():
        raise ValueError("mask has active cells with both axial neighbours missing")
    n0, n1 = m.shape
    ai = active_index_from_mask(m).reshape(n0, n1) - 1
    rows, cols, vals = [], [], []
    ii, jj = np.nonzero(m)
    me = ai[ii, jj]
    rows.append(me); cols.append(me); vals.append(np.full(me.size, -4.0))
    for (di, dj) in ((1, 0), (0, 1)):
        has_p = _shift(m, di, dj)[ii, jj]
        has_m = _shift(m, -di, -dj)[ii, jj]
        # + neighbour: weight 1, plus 1 more if the - neighbour is missing
        sel = has_p
        rows.append(me[sel]); cols.append(ai[ii[sel] + di, jj[sel] + dj])
        vals.append(np.where(has_m[sel], 1.0, 2.0))
        sel = has_m
        rows.append(me[sel]); cols.append(ai[ii[sel] - di, jj[sel] - dj])
        vals.append(np.where(has_p[sel], 1.0, 2.0))
    n = int(m.sum())
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))


def grad_matrices(mask: np.ndarray):
    """Per-axis second-order finite differences D_i, D_j over active cells (grid units).

    Central ``(y+ - y-)/2`` when both neighbours exist, otherwise the one-sided
    ``(-3 y0 + 4 y1 - y2)/2`` pointing into the tissue (sign mirrored for the - side), or
    ``y1 - y0`` when the open side has a single cell.
    """
    m = np.asarray(mask, dtype=bool)
    n0, n1 = m.shape
    ai = active_index_from_mask(m).reshape(n0, n1) - 1
    ii, jj = np.nonzero(m)
    me = ai[ii, jj]
    n = int(m.sum())
    mats = []
    for (di, dj) in ((1, 0), (0, 1)):
        has_p = _shift(m, di, dj)[ii, jj]
        has_m = _shift(m, -di, -dj)[ii, jj]
        has_pp = _shift(m, 2 * di, 2 * dj)[ii, jj]
        has_mm = _shift(m, -2 * di, -2 * dj)[ii, jj]
        r, c, v = [], [], []
        both = has_p & has_m
        r += [me[both], me[both]]
        c += [ai[ii[both] + di, jj[both] + dj], ai[ii[both] - di, jj[both] - dj]]
        v += [np.full(both.sum(), 0.5), np.full(both.sum(), -0.5)]
        # one-sided second order where two cells exist on the open side; first order
        # (y1 - y0) where only one does (never the case on the reference's 256^2 grid)
        fwd = has_p & ~has_m & has_pp
        r += [me[fwd]] * 3
        c += [me[fwd], ai[ii[fwd] + di, jj[fwd] + dj], ai[ii[fwd] + 2 * di, jj[fwd] + 2 * dj]]
        v += [np.full(fwd.sum(), -1.5), np.full(fwd.sum(), 2.0), np.full(fwd.sum(), -0.5)]
        fwd1 = has_p & ~has_m & ~has_pp
        r += [me[fwd1]] * 2
        c += [me[fwd1], ai[ii[fwd1] + di, jj[fwd1] + dj]]
        v += [np.full(fwd1.sum(), -1.0), np.full(fwd1.sum(), 1.0)]
        bwd = has_m & ~has_p & has_mm
        r += [me[bwd]] * 3
        c += [me[bwd], ai[ii[bwd] - di, jj[bwd] - dj], ai[ii[bwd] - 2 * di, jj[bwd] - 2 * dj]]
        v += [np.full(bwd.sum(), 1.5), np.full(bwd.sum(), -2.0), np.full(bwd.sum(), 0.5)]
        bwd1 = has_m & ~has_p & ~has_mm
        r += [me[bwd1]] * 2
        c += [me[bwd1], ai[ii[bwd1] - di, jj[bwd1] - dj]]
        v += [np.full(bwd1.sum(), 1.0), np.full(bwd1.sum(), -1.0)]
        mats.append(sp.csr_matrix((np.concatenate(v), (np.concatenate(r), np.concatenate(c))), shape=(n, n)))
    return mats


def S_matrix(mask: np.ndarray):
    """Gradient-penalty matrix ``S = D_i^T D_i + D_j^T D_j`` (so ``y^T S y`` is the discrete
    ``sum |grad y|^2`` in grid units), as CSC."""
    Di, Dj = grad_matrices(mask)
    return (Di.T @ Di + Dj.T @ Dj).tocsc()


def mean_u2_S(n_active: int):
    """``S = (2/n_active) I``: with gamma=(1,0,0) the reference objective ``1/2 y^T S y``
    becomes ``mean(u^2)`` over active cells."""
    return (sp.identity(n_active, format="csc") * (2.0 / n_active)).tocsc()


def disc_holes(nrows: int, ncols: int, n_holes: int, rmin: int, rmax: int, seed: int) -> np.ndarray:
    """Seeded mask with ``n_holes`` non-conducting discs of radius U{rmin..rmax} cells, then
    isthmus-filled so that every active cell has at least one neighbour per axis."""
    rng = np.random.default_rng(seed)
    m = np.ones((nrows, ncols), dtype=bool)
    I, J = np.mgrid[0:nrows, 0:ncols]
    for _ in range(n_holes):
        r = int(rng.integers(rmin, rmax + 1))
        ci = rng.uniform(0, nrows)
        cj = rng.uniform(0, ncols)
        m &= (I - ci) ** 2 + (J - cj) ** 2 > r * r
    return fill_isthmuses(m)


def csc_arrays(S):
    """(colptr int64, rowind int64, values float64) of a scipy sparse matrix in CSC form."""
    S = S.tocsc()
    S.sort_indices()
    return (S.indptr.astype(np.int64), S.indices.astype(np.int64), S.data.astype(np.float64))


def sparse_fingerprint(A) -> dict:
    """Order-independent summary of a sparse matrix used to compare it with a stored copy."""
    A = A.tocoo()
    r = A.row.astype(np.float64)
    c = A.col.astype(np.float64)
    v = A.data.astype(np.float64)
    keep = v != 0
    r, c, v = r[keep], c[keep], v[keep]
    return {
        "shape": [int(A.shape[0]), int(A.shape[1])],
        "nnz": int(keep.sum()),
        "sum": float(v.sum()),
        "sum_abs": float(np.abs(v).sum()),
        "sum_rcv": float(((r + 1.0) * v).sum() + ((c + 1.0) * 3.0 * v).sum()),
        "sum_v2r": float((v * v * (r + 1.0) * (c + 1.0) % 1e9).sum()),
    }
