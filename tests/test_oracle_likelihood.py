"""Pins of the oracle's likelihood, prediction and gradient (CPU only)."""
import os

import numpy as np
import pytest
import scipy.linalg as sla
from scipy.stats import multivariate_normal

from oracle import datagen, likelihood, matern

GOLD = os.path.join(os.path.dirname(__file__), "golden")
K = "univariate_matern_stationary"


@pytest.fixture(scope="module")
def c1():
    return dict(np.load(os.path.join(GOLD, "c1_seed0.npz")))


def test_dmle_matches_mvn_logpdf(c1):
    for th in c1["thetas"][:3]:
        S = likelihood.sigma(K, c1["x"], c1["y"], th)
        ref = multivariate_normal(mean=np.zeros(len(c1["z"])), cov=S).logpdf(c1["z"])
        val, _, _ = likelihood.dmle(K, c1["x"], c1["y"], c1["z"], th)
        assert abs(val - ref) <= 1e-10 * abs(ref)


def test_dmle_against_high_precision_cholesky():
    """n = 40: the whole path in 40-digit mpmath."""
    import mpmath as mp
    mp.mp.dps = 40
    x, y, _ = datagen.generate_xy_loc(40, 3)
    z = np.sin(np.arange(40) * 1.3)
    th = (1.2, 0.2, 1.37)
    val, ld, dp = likelihood.dmle(K, x, y, z, th)
    n = 40
    S = mp.matrix(n, n)
    con = 1 / (mp.power(2, mp.mpf(th[2]) - 1) * mp.gamma(th[2]))
    for i in range(n):
        for j in range(n):
            d = mp.sqrt((mp.mpf(x[i]) - x[j]) ** 2 + (mp.mpf(y[i]) - y[j]) ** 2) / th[1]
            S[i, j] = th[0] if d == 0 else th[0] * con * d ** th[2] * mp.besselk(th[2], d)
    L = mp.cholesky(S)
    ld_mp = 2 * mp.fsum(mp.log(L[i, i]) for i in range(n))
    v = mp.lu_solve(L, mp.matrix(z.tolist()))
    dp_mp = mp.fsum(v[i] ** 2 for i in range(n))
    ref = -dp_mp / 2 - ld_mp / 2 - mp.mpf(n) / 2 * mp.log(2 * mp.pi)
    assert abs(val - float(ref)) <= 1e-11 * abs(float(ref))
    assert abs(ld - float(ld_mp)) <= 1e-11 * max(1.0, abs(float(ld_mp)))


def test_dmle_permutation_invariance(c1):
    p = np.random.default_rng(3).permutation(1600)
    a, _, _ = likelihood.dmle(K, c1["x"], c1["y"], c1["z"], c1["theta_true"])
    b, _, _ = likelihood.dmle(K, c1["x"][p], c1["y"][p], c1["z"][p], c1["theta_true"])
    assert abs(a - b) <= 1e-11 * abs(a)


def test_golden_c1_regenerates(c1):
    x, y, rng = datagen.generate_xy_loc(1600, 0)
    e, _ = datagen.dlarnv(3, [0, 0, 0, 1], 1600)
    assert np.array_equal(x, c1["x"]) and np.array_equal(e, c1["e"])
    z = likelihood.dzvg(K, x, y, c1["theta_true"], e)
    assert np.allclose(z, c1["z"], rtol=0, atol=1e-13)
    val, _, _ = likelihood.dmle(K, x, y, z, c1["theta_true"])
    assert abs(val - c1["loglik"][0]) <= 1e-12 * abs(val)
    # z = L e at theta_true: L^-1 z = e, so dotp = |e|^2
    assert abs(c1["dotp"][0] - e @ e) <= 1e-10 * (e @ e)


def test_predict_matches_explicit_formula(c1):
    th = c1["theta_true"]
    C22 = matern.core_dcmg(c1["xobs"], c1["yobs"], c1["xobs"], c1["yobs"], th)
    C12 = matern.core_dcmg(c1["xmiss"], c1["ymiss"], c1["xobs"], c1["yobs"], th)
    ref = C12 @ np.linalg.solve(C22, c1["zobs"])
    assert np.linalg.norm(ref - c1["zmiss"]) <= 1e-10 * np.linalg.norm(ref)
    assert c1["mse"][0] == pytest.approx(np.mean((c1["zactual"] - ref) ** 2), rel=1e-10)


def test_predict_interpolates_observed(c1):
    idx = [0, 77, 1599]
    zm, _, _, _ = likelihood.predict(K, c1["x"], c1["y"], c1["z"], c1["x"][idx], c1["y"][idx],
                                     c1["theta_true"])
    assert np.max(np.abs(zm - c1["z"][idx])) < 1e-9


def test_pick_random_points_split(c1):
    """Held-out and observed sets partition the data; both are Morton-sorted."""
    assert len(c1["zobs"]) == 1500 and len(c1["zactual"]) == 100
    allz = np.sort(np.concatenate([c1["zobs"], c1["zactual"]]))
    assert np.array_equal(allz, np.sort(c1["z"]))
    o = datagen.obs_zsort_order(c1["xobs"], c1["yobs"])
    assert np.array_equal(o, np.arange(1500))


@pytest.mark.parametrize("kernel,theta", [
    ("univariate_matern_stationary", (1.0, 0.1, 0.5)),
    ("univariate_matern_stationary", (1.3, 0.07, 1.37)),
    ("univariate_matern_nuggets_stationary", (1.0, 0.1, 0.5, 0.01)),
    ("bivariate_matern_parsimonious", (1.0, 1.3, 0.1, 0.5, 1.0, 0.5)),
])
def test_gradient_matches_central_differences(kernel, theta):
    """The dense adjoint formula vs central differences of dmle, h = 1e-6
    relative (the bar the GPU adjoint is held to as well)."""
    nloc = 200 if kernel.startswith("bivariate") else 400
    x, y, _ = datagen.generate_xy_loc(nloc, 1)
    e, _ = datagen.dlarnv(3, [1, 1, 1, 1], 400)
    z = likelihood.dzvg(kernel, x, y, theta, e)
    _, g = likelihood.gradient(kernel, x, y, z, theta)
    th = np.array(theta)
    for p in range(len(th)):
        h = 1e-6 * abs(th[p])
        tp, tm = th.copy(), th.copy()
        tp[p] += h
        tm[p] -= h
        fd = (likelihood.dmle(kernel, x, y, z, tp)[0] - likelihood.dmle(kernel, x, y, z, tm)[0]) / (2 * h)
        assert abs(g[p] - fd) <= 1e-5 * max(1.0, abs(fd)), (p, g[p], fd)


def test_gradient_golden_nugget():
    g = np.load(os.path.join(GOLD, "nugget_900.npz"))
    l, gr = likelihood.gradient("univariate_matern_nuggets_stationary", g["x"], g["y"], g["z"], g["theta"])
    assert abs(l - g["loglik"][0]) <= 1e-12 * abs(l)
    assert np.allclose(gr, g["grad"], rtol=1e-9, atol=1e-9)


def test_bivariate_golden_and_spd():
    g = np.load(os.path.join(GOLD, "bivariate_400.npz"))
    l, ld, _ = likelihood.dmle("bivariate_matern_parsimonious", g["x"], g["y"], g["z"], g["theta"])
    assert abs(l - g["loglik"][0]) <= 1e-11 * abs(l)


def test_flop_formulas():
    assert likelihood.flops_dpotrf(90000) == pytest.approx(2.430e14, rel=1e-3)
    assert likelihood.flops_dtrsm_left(90000, 1) == 90000 * 90001


def test_fisher_closed_forms():
    """Fisher information pinned without its own formula: with dS/dsigma^2 =
    S/sigma^2 (no nugget), F_00 = n / (2 sigma^4) and F_0k = -g_k(z = 0) /
    sigma^2, where g is the oracle gradient (pinned to finite differences) at
    z = 0, whose value is -1/2 tr(S^-1 dS_k); with the nugget kernel
    (dS/dtau^2 = I), F_33 = 1/2 sum 1/lambda_i^2 over the eigenvalues of S."""
    x, y, _ = datagen.generate_xy_loc(300, 3)
    theta = (1.3, 0.08, 0.9)
    F, Finv = likelihood.fisher("univariate_matern_stationary", x, y, theta)
    n = len(x)
    assert abs(F[0, 0] - n / (2 * theta[0] ** 2)) <= 1e-10 * F[0, 0]
    _, g0 = likelihood.gradient("univariate_matern_stationary", x, y, np.zeros(n), theta)
    np.testing.assert_allclose(F[0], -g0 / theta[0], rtol=1e-9)
    np.testing.assert_allclose(F @ Finv, np.eye(3), atol=1e-9)
    thn = (1.0, 0.1, 0.5, 0.05)
    Fn, _ = likelihood.fisher("univariate_matern_nuggets_stationary", x, y, thn)
    lam = np.linalg.eigvalsh(likelihood.sigma("univariate_matern_nuggets_stationary", x, y, thn))
    assert abs(Fn[3, 3] - 0.5 * np.sum(lam ** -2.0)) <= 1e-9 * Fn[3, 3]


def test_fisher_is_covariance_of_the_score():
    """F = E[s s^T] for the score s = d l / d theta at z ~ N(0, S): a Monte
    Carlo estimate from the oracle gradient (2,000 draws, n = 36) agrees
    with every entry within 4.5 standard errors."""
    x, y, _ = datagen.generate_xy_loc(36, 7)
    theta = (1.0, 0.2, 0.8)
    K = "univariate_matern_stationary"
    F, _ = likelihood.fisher(K, x, y, theta)
    L = np.linalg.cholesky(likelihood.sigma(K, x, y, theta))
    rng = np.random.default_rng(5)
    S = np.array([likelihood.gradient(K, x, y, L @ rng.standard_normal(len(x)), theta)[1] for _ in range(2000)])
    est = S.T @ S / len(S)
    prod = np.einsum("si,sj->sij", S, S)
    se = prod.std(axis=0) / np.sqrt(len(S))
    assert np.all(np.abs(est - F) <= 4.5 * se), (est, F, se)
