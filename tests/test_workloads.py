"""Workload generators (CPU): S2 stimulus regions and the spiral-core counter used to describe the
benchmark states."""
import numpy as np

from fkadj import workloads as W


def _packed(u, v, mask):
    return np.concatenate([u[mask], v[mask], np.ones(int(mask.sum()))])


def test_s2_regions():
    one = W.s2_region(64, 64, 20, 1)
    assert one[32:, :20].all() and not one[:32].any() and not one[:, 20:].any()
    two = W.s2_region(64, 64, 20, 2)
    assert two[8:24, :20].all() and two[40:56, :20].all()
    assert two.sum() == 2 * 16 * 20
    # each band edge inside the grid gives the S2 wave a free end: 1 for the half plane, 4 for two bands
    def inner_edges(r):
        col = r[:, 0].astype(int)
        return int(np.abs(np.diff(col)).sum())
    assert inner_edges(one) == 1 and inner_edges(two) == 4


def test_phase_singularities_count_cores():
    N = 64
    yy, xx = np.mgrid[0:N, 0:N]
    mask = np.ones((N, N), bool)
    cores = [(20.3, 30.7), (45.2, 10.1), (50.5, 50.5)]
    for k in range(len(cores) + 1):
        th = sum(np.arctan2(yy - cy, xx - cx) for cy, cx in cores[:k]) + 0.0 * yy
        y = _packed(0.5 + 0.4 * np.cos(th), 0.5 + 0.4 * np.sin(th), mask)
        assert W.phase_singularities(y, mask) == k
    # a hole over a core removes it (plaquettes need four active cells)
    m2 = mask.copy()
    m2[18:23, 28:34] = False
    th = np.arctan2(yy - 20.3, xx - 30.7)
    assert W.phase_singularities(_packed(0.5 + 0.4 * np.cos(th), 0.5 + 0.4 * np.sin(th), m2), m2) == 0
    rest = W.rest_state(int(mask.sum()))
    assert W.phase_singularities(rest, mask) == 0
