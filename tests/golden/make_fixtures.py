"""Extract the reference's golden data into small fixtures under tests/golden/.

Run once, in a container where the reference checkout is mounted at /root/reference:

    python tests/golden/make_fixtures.py

Only MAT-file *data* is read (scipy.io.loadmat; nothing is executed).  Inputs are
``src/cuda/data/*.mat`` of the reference: the 256x256 grid (``active_index``, ``ghost_lr``,
``lap``), the gradient-penalty matrix ``S``, the spiral initial state ``yf``, the seed pulse
``Et`` and the four golden gradient-descent snapshots written by
``src/cuda/simple_grad_desc.m:93-101`` with ``src/cuda/infile_grad`` settings.

The large sparse matrices are not stored; instead this script checks that the package's own
constructions (mirror Laplacian and ``S = D^T D`` from the hole mask) reproduce them exactly and
stores fingerprints, so the tests can re-check the constructions without the reference.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import scipy.io as sio

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "adjoint-optimization_amd"))

from fkadj import geometry  # noqa: E402

DATA = "/root/reference/src/cuda/data"
RUN = "simple_grad_desc_step_2024-nov-02-17-32-50_000001_{}.mat"


def load(name):
    return sio.loadmat(os.path.join(DATA, name), mat_dtype=True)


def main():
    grid = load("lap_Cfile_2nd_order_grad_dx_0p035_N_256_hole_dens_16.mat")
    active_index = grid["active_index"].ravel().astype(np.int32)
    ghost_lr = grid["ghost_lr"].ravel()
    nrows = ncols = 256
    mask = geometry.mask_from_active_index(active_index, nrows, ncols)
    n_act = int(mask.sum())

    # our constructions vs the stored operators
    assert np.array_equal(geometry.active_index_from_mask(mask), active_index)
    assert np.array_equal(geometry.ghost_lr_from_mask(mask), ghost_lr), "ghost_lr closed form"
    lap_ref = grid["lap"].tocsr()
    lap_ours = geometry.lap_matrix(mask)
    assert abs(lap_ref - lap_ours).max() == 0.0, "mirror Laplacian"
    S_ref = load("S_1o4_N_256_hole_dens_16.mat")["S"].tocsc()
    S_ours = geometry.S_matrix(mask)
    assert abs(S_ref - S_ours).max() == 0.0, "S = D^T D"

    yf = load("spirals.mat")["yf"].ravel()
    bx3 = np.concatenate([active_index > 0] * 3)
    y0 = yf[bx3]                        # src/cuda/simple_grad_desc.m:62-64
    assert np.array_equal(np.isnan(yf[: nrows * ncols]), active_index == 0)
    assert not np.isnan(y0).any()

    seed = load("Et_pulse_N_1_E0_5.00_t0_20.0_300ms.mat")
    Et_seed = seed["Et"].ravel()

    out = dict(
        active_index=active_index,
        ghost_lr=ghost_lr.astype(np.int8),
        y0=y0,
        Et_seed=Et_seed,
    )
    meta = dict(
        nrows=nrows, ncols=ncols, n_active=n_act,
        lap=geometry.sparse_fingerprint(lap_ref),
        S=geometry.sparse_fingerprint(S_ref),
        iterations={},
    )
    for it in (50, 100, 150, 200):
        d = load(RUN.format(it))
        if it == 50:
            for k in ("alpha", "ds", "dt", "final_time", "nmod"):
                meta[k] = float(d[k].ravel()[0])
            meta["gamma"] = [float(x) for x in d["gamma"].ravel()]
        out[f"Et_{it}"] = d["Et"].ravel()
        out[f"gradE_{it}"] = d["gradE"].ravel()
        out[f"Lcost_vec_{it}"] = d["Lcost_vec"].ravel()
        out[f"L0_vec_{it}"] = d["L0_vec"].ravel()
        out[f"LE_vec_{it}"] = d["LE_vec"].ravel()
        meta["iterations"][str(it)] = dict(Lcost=float(d["Lcost"].ravel()[0]),
                                           gradE_norm=float(np.linalg.norm(d["gradE"])))

    np.savez_compressed(os.path.join(HERE, "ref256.npz"), **out)
    with open(os.path.join(HERE, "ref256.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print("wrote ref256.npz / ref256.json; n_active =", n_act)


if __name__ == "__main__":
    main()
