"""Pins of the FP64 oracle (oracle/fk_oracle.c) against the reference's golden data and against
independent mathematics.  CPU only.

What pins what:
  * golden end-to-end (iteration 1 of src/cuda/simple_grad_desc.m with src/cuda/infile_grad):
    forward RK4 + F_ion + Laplacian + spline + objective  -> Lcost, M (L0), N (LE)
  * golden iteration 50: the whole adjoint (Hermite midpoint, L^T, J^T, adjoint RK4, z.gh) -> gradE
  * lap / lapT against the sparse mirror Laplacian, which equals the reference's stored `lap`
  * jacT against central finite differences of F_ion
  * spline against scipy's independent not-a-knot CubicSpline and exactness on cubics
  * RK4 against its 4th-order convergence on a small grid
  * the adjoint gradient against central differences of J in the pulse amplitudes
"""
import numpy as np
import pytest

import oracle as O
from fkadj import geometry as G


@pytest.fixture(scope="module")
def m_ref():
    O.build()
    return O.reference_model()


def test_geometry_matches_reference_operators(golden):
    d, meta = golden
    mask = G.mask_from_active_index(d["active_index"], 256, 256)
    assert int(mask.sum()) == meta["n_active"] == len(d["y0"]) // 3
    assert np.array_equal(G.ghost_lr_from_mask(mask), d["ghost_lr"].astype(float))
    assert G.sparse_fingerprint(G.lap_matrix(mask)) == meta["lap"]
    assert G.sparse_fingerprint(G.S_matrix(mask)) == meta["S"]
    assert not G.isthmus_cells(mask).any()


def _random_planes(rng, mask):
    y = rng.standard_normal((3,) + mask.shape)
    y[:, ~mask] = np.nan
    return y


def test_lap_equals_sparse_mirror_laplacian(m_ref):
    rng = np.random.default_rng(1)
    mask = G.disc_holes(48, 64, 12, 1, 3, seed=3)
    L = G.lap_matrix(mask)
    y = _random_planes(rng, mask)
    out = O.lap(m_ref, y)
    D = [m_ref.Du, m_ref.Dv, m_ref.Dw]
    for s in range(3):
        exp = D[s] / m_ref.h ** 2 * (L @ y[s][mask])
        np.testing.assert_allclose(out[s][mask], exp, rtol=1e-13, atol=1e-13)
        assert np.isnan(out[s][~mask]).all()


def test_lapT_equals_transposed_laplacian(m_ref):
    rng = np.random.default_rng(2)
    mask = G.disc_holes(40, 56, 10, 1, 3, seed=5)
    L = G.lap_matrix(mask)
    z = _random_planes(rng, mask)
    out = O.lapT(m_ref, z)
    D = [m_ref.Du, m_ref.Dv, m_ref.Dw]
    for s in range(3):
        exp = D[s] / m_ref.h ** 2 * (L.T @ z[s][mask])
        np.testing.assert_allclose(out[s][mask], exp, rtol=1e-13, atol=1e-13)


def test_fion_rest_state(m_ref):
    # resting tissue u=0, v=w=1 is a fixed point up to the smoothed-step tails
    f = O.fion(m_ref, 0.0, 1.0, 1.0)
    assert np.all(np.abs(f) < 1e-3)


@pytest.mark.parametrize("y", [(0.02, 0.9, 0.95), (0.13, 0.5, 0.7), (0.6, 0.1, 0.4), (0.9, 0.0, 0.2),
                               (0.041, 0.8, 0.8), (0.55, 0.3, 0.6)])
def test_jacT_matches_finite_differences(m_ref, y):
    y = np.array(y)
    JT = O.jacT(m_ref, y).reshape(3, 3)   # JT[a, b] = dF_b/dy_a
    eps = 1e-7
    for a in range(3):
        e = np.zeros(3); e[a] = eps
        fd = (O.fion(m_ref, *(y + e)) - O.fion(m_ref, *(y - e))) / (2 * eps)
        np.testing.assert_allclose(JT[a], fd, rtol=2e-5, atol=1e-6)


def test_spline_not_a_knot_matches_scipy():
    from scipy.interpolate import CubicSpline
    rng = np.random.default_rng(0)
    t = np.arange(40) * 0.1
    y = rng.standard_normal(40)
    ypp = O.spline_set(t, y)
    cs = CubicSpline(t, y, bc_type="not-a-knot")
    np.testing.assert_allclose(ypp, cs(t, 2), rtol=1e-10, atol=1e-10)
    for tv in (0.05, 0.333, 1.25, 3.86):
        assert abs(O.spline_val(t, y, ypp, tv) - cs(tv)) < 1e-12


def test_spline_exact_on_cubic():
    t = np.arange(12) * 0.25
    p = lambda x: 0.3 - 1.2 * x + 0.7 * x ** 2 - 0.11 * x ** 3
    ypp = O.spline_set(t, p(t))
    for tv in (0.1, 1.37, 2.6):
        assert abs(O.spline_val(t, p(t), ypp, tv) - p(tv)) < 1e-12


def test_knot_tables_indexing():
    dt = 0.1
    Et = np.sin(np.arange(21) * 0.3)
    Ef, Ea = O.knot_tables(dt, Et)
    Ef = Ef.reshape(-1, 3); Ea = Ea.reshape(-1, 3)
    assert np.array_equal(Ef[:, 0], Et[:-1]) and np.array_equal(Ef[:, 2], Et[1:])
    assert np.array_equal(Ea[:, 0], Et[:-1]) and np.array_equal(Ea[:, 1], Et[1:])
    from scipy.interpolate import CubicSpline
    cs = CubicSpline(np.arange(21) * dt, Et, bc_type="not-a-knot")
    np.testing.assert_allclose(Ef[:, 1], cs((np.arange(20) + 0.5) * dt), atol=1e-13)
    np.testing.assert_allclose(Ea[:, 2], cs((np.arange(20) + 1 / 3) * dt), atol=1e-13)


def test_rk4_fourth_order_convergence(golden):
    d, _ = golden
    m = O.reference_model()
    # a 32x32 window of the reference spiral state: realistic fronts, real holes
    # (the crop's edges are walls, so isthmus cells are filled first: the mirror rule is NaN there)
    y = O.num2nan(d["y0"], d["active_index"], 256, 256)[:, 96:128, 96:128].copy()
    mask = G.fill_isthmuses(~np.isnan(y[0]))
    y[:, ~mask] = np.nan
    gh = np.where(mask, 2 * m.Du / m.h * 0.3, np.nan)
    # the ksm=100 smoothed steps need dt << 0.01 ms for the asymptotic regime
    T = 0.02

    def run(nt):
        Ef = np.full(3 * nt, 1.5)
        return O.forward(m, y, gh, T / nt, Ef)

    ref = run(512)
    e1 = np.nanmax(np.abs(run(4) - ref))
    e2 = np.nanmax(np.abs(run(8) - ref))
    assert 14.0 < e1 / e2 < 18.0, (e1, e2)


@pytest.mark.slow
def test_golden_iteration1_objective(golden):
    """Seed pulse: Lcost, M and N of iteration 1 (…_50.mat Lcost_vec/L0_vec/LE_vec[0])."""
    d, meta = golden
    ai = d["active_index"]
    mask = G.mask_from_active_index(ai, 256, 256)
    csc = G.csc_arrays(G.S_matrix(mask))
    g, Lc, L0, LE = O.grad_ref(meta["dt"], d["Et_seed"], 3000, d["y0"], ai, 256, 256,
                               d["ghost_lr"].astype(float), meta["gamma"], csc, meta["alpha"])
    assert abs(Lc - d["Lcost_vec_50"][0]) / d["Lcost_vec_50"][0] < 1e-12
    assert abs(L0 - d["L0_vec_50"][0]) / d["L0_vec_50"][0] < 1e-12
    assert abs(LE - d["LE_vec_50"][0]) / d["LE_vec_50"][0] < 1e-14


@pytest.mark.slow
def test_golden_iteration50_gradient(golden):
    """Iterate 50: full gradient vs the stored gradE (3001 values) and Lcost."""
    d, meta = golden
    ai = d["active_index"]
    mask = G.mask_from_active_index(ai, 256, 256)
    csc = G.csc_arrays(G.S_matrix(mask))
    g, Lc, L0, LE = O.grad_ref(meta["dt"], d["Et_50"], 3000, d["y0"], ai, 256, 256,
                               d["ghost_lr"].astype(float), meta["gamma"], csc, meta["alpha"])
    G50 = d["gradE_50"]
    assert np.linalg.norm(g - G50) / np.linalg.norm(G50) < 1e-10
    assert abs(Lc - meta["iterations"]["50"]["Lcost"]) / meta["iterations"]["50"]["Lcost"] < 1e-12


def test_gradient_matches_finite_differences():
    """Independent of the golden data: dJ/dA_k from the oracle's adjoint against central differences
    of J in the amplitudes (piecewise field, energy weight 0, so the whole gradient comes from the
    adjoint solve).  The mismatch is the method's own (the functional gradient at the knots,
    integrated over each piece by the trapezoid rule): second order in dt, so halving dt divides it
    by four; a wrong sign, factor or time index would not converge at all."""
    from cases import disc_case, oracle_grad_piecewise
    case = disc_case(24, 40, 2, 5)
    om = O.reference_model(h=0.025)
    S = G.mean_u2_S(case.n_act)
    A = np.array([3.0, 7.5, 1.0, 5.0])

    def worst_mismatch(dt, nt):
        g = oracle_grad_piecewise(case, dt, nt, A, (1, 0, 0), 0.0, S, om)[0]
        worst = 0.0
        for k in range(len(A)):
            h = 1e-3 * (1 + abs(A[k]))
            Ap, Am = A.copy(), A.copy()
            Ap[k] += h
            Am[k] -= h
            Jp = oracle_grad_piecewise(case, dt, nt, Ap, (1, 0, 0), 0.0, S, om)[1]
            Jm = oracle_grad_piecewise(case, dt, nt, Am, (1, 0, 0), 0.0, S, om)[1]
            worst = max(worst, abs((Jp - Jm) / (2 * h) - g[k]) / abs(g[k]))
        return worst

    e1, e2 = worst_mismatch(0.02, 40), worst_mismatch(0.01, 80)
    print(f"oracle adjoint vs central differences: max rel {e1:.2e} (dt 0.02), {e2:.2e} (dt 0.01)")
    assert e1 < 1e-3 and 3.5 < e1 / e2 < 4.5, (e1, e2)


def test_state_rounding_probe():
    """The precision probe of orc_grad (round_state): the final state it returns is exactly FP32
    representable, the gradient moves by a rounding-sized amount (not zero, not order one) and the
    knot-mode helper built on it agrees with grad_ref when the probe is off."""
    from cases import disc_case, oracle_grad_knots, oracle_grad_knots_state_rounded
    case = disc_case(24, 40, 2, 5)
    om = O.reference_model(h=0.025)
    S = G.S_matrix(case.mask)
    nt, dt = 200, 0.02
    Ef = np.repeat(np.linspace(2.0, 6.0, nt), 3)
    gh = case.gh_plane(om)
    z0, L0, yT = O.grad(om, case.planes(), gh, dt, Ef, Ef, case.ai, G.csc_arrays(S), (1, 0.2, 0.2), ckpt=20)
    z1, L1, yR = O.grad(om, case.planes(), gh, dt, Ef, Ef, case.ai, G.csc_arrays(S), (1, 0.2, 0.2), ckpt=20,
                        round_state=True)
    act = case.mask
    assert np.array_equal(yR[:, act], yR[:, act].astype(np.float32).astype(np.float64))
    assert not np.array_equal(yT[:, act], yR[:, act])
    e = np.linalg.norm(z1 - z0) / np.linalg.norm(z0)
    assert 1e-9 < e < 1e-3, e
    Et = 3.0 + np.sin(np.arange(nt + 1) * 0.05)
    args = (dt, Et, case.y0, case.ai, case.nrows, case.ncols, case.gh, (1, 0.2, 0.2), S, 2e-3)
    g0, Lc0, _, _ = oracle_grad_knots(case, dt, Et, (1, 0.2, 0.2), 2e-3, S, model=om)
    gf, Lcf = oracle_grad_knots_state_rounded(*args, model=om, round_state=False)
    assert np.linalg.norm(gf - g0) / np.linalg.norm(g0) < 1e-13 and abs(Lcf / Lc0 - 1) < 1e-13
    g, Lc = oracle_grad_knots_state_rounded(*args, model=om)
    assert 0 < np.linalg.norm(g - g0) / np.linalg.norm(g0) < 1e-3
    assert abs(Lc / Lc0 - 1) < 1e-5
