"""Pins of the oracle's MLOE / MMOM (CPU only).

The oracle follows the reference's per-location sequence of solves and
products.  These checks use routes independent of that sequence: the mean
square error of a linear predictor as a quadratic form of the joint
(observed + held-out) covariance, the Schur-complement identity for the
optimal predictor, closed forms at the true theta and for a rescaled
variance, MLOE >= 0, and a Monte Carlo estimate of the true MSE.
"""
import numpy as np
import pytest

from oracle import datagen, likelihood, matern

K = "univariate_matern_stationary"
KN = "univariate_matern_nuggets_stationary"


def _split(n, nmiss, seed):
    x, y, _ = datagen.generate_xy_loc(n, seed)
    rng = np.random.default_rng(seed)
    miss = np.sort(rng.choice(n, nmiss, replace=False))
    obs = np.setdiff1d(np.arange(n), miss)
    return x[obs], y[obs], x[miss], y[miss]


@pytest.fixture(scope="module")
def locs():
    return _split(260, 12, 3)


def _joint(kernel, xo, yo, x0, y0, theta, var0):
    """Covariance of (Z(s0), Z(obs)) with Var Z(s0) set to var0."""
    xs = np.concatenate([[x0], xo])
    ys = np.concatenate([[y0], yo])
    J = matern.cov_matrix(kernel, xs, ys, xs, ys, theta)
    J[0, 0] = var0
    return J


@pytest.mark.parametrize("kernel,tt", [(K, (1.0, 0.1, 0.5)), (KN, (1.0, 0.1, 0.5, 0.05))])
def test_zero_at_true_theta(locs, kernel, tt):
    mloe, mmom, loe, mom, _ = likelihood.mloe_mmom(kernel, *locs, tt, tt)
    assert abs(mloe) <= 1e-12 and abs(mmom) <= 1e-12
    assert np.max(np.abs(loe)) <= 1e-11 and np.max(np.abs(mom)) <= 1e-11


def test_rescaled_variance_closed_form(locs):
    """theta_est = (c sigma^2, beta, nu): the kriging weights are unchanged,
    so MLOE = 0 and the presumed MSE is c times the true one: MMOM = c - 1."""
    tt = (1.3, 0.08, 0.9)
    for c in (0.5, 2.5):
        te = (c * tt[0], tt[1], tt[2])
        mloe, mmom, _, _, _ = likelihood.mloe_mmom(K, *locs, tt, te)
        assert abs(mloe) <= 1e-10, mloe
        assert abs(mmom - (c - 1.0)) <= 1e-10, (mmom, c)


@pytest.mark.parametrize("kernel,tt,te", [
    (K, (1.0, 0.1, 0.5), (0.8, 0.15, 0.7)),
    (K, (1.2, 0.05, 1.2), (1.0, 0.09, 0.6)),
    (KN, (1.0, 0.1, 0.5, 0.05), (0.9, 0.07, 0.8, 0.02)),
])
def test_parts_match_joint_quadratic_forms(locs, kernel, tt, te):
    """expr2 = [-w_a; 1]^T J_t [-w_a; 1] with w_a from a dense solve under
    theta_est; expr3 = 1 / (J_t^-1)_00 and expr4 = 1 / (J_a^-1)_00 (the
    conditional variance as a Schur complement), with Var Z(s0) = theta[0]
    as the reference takes it."""
    xo, yo, xm, ym = locs
    _, _, loe, mom, parts = likelihood.mloe_mmom(kernel, xo, yo, xm, ym, tt, te)
    for p in range(len(xm)):
        Jt = _joint(kernel, xo, yo, xm[p], ym[p], tt, tt[0])
        Ja = _joint(kernel, xo, yo, xm[p], ym[p], te, te[0])
        w = np.linalg.solve(Ja[1:, 1:], Ja[1:, 0])
        v = np.concatenate([[1.0], -w])
        e2 = v @ Jt @ v
        e3 = 1.0 / np.linalg.inv(Jt)[0, 0]
        e4 = 1.0 / np.linalg.inv(Ja)[0, 0]
        assert abs(parts[0, p] - e2) <= 1e-9 * abs(e2), (p, parts[0, p], e2)
        assert abs(parts[1, p] - e3) <= 1e-9 * abs(e3), (p, parts[1, p], e3)
        assert abs(parts[2, p] - e4) <= 1e-9 * abs(e4), (p, parts[2, p], e4)
        # the optimal predictor has the smallest true MSE
        assert loe[p] >= -1e-12
    assert np.all(mom > -1.0)


def test_true_mse_monte_carlo(locs):
    """expr2 is the mean squared error of the theta_est predictor when the
    field follows theta_true: 40,000 joint draws, within 4 standard errors."""
    xo, yo, xm, ym = locs
    tt, te = (1.0, 0.1, 0.5), (0.7, 0.2, 1.1)
    _, _, _, _, parts = likelihood.mloe_mmom(K, xo, yo, xm[:3], ym[:3], tt, te)
    rng = np.random.default_rng(11)
    for p in range(3):
        Jt = _joint(K, xo, yo, xm[p], ym[p], tt, tt[0])
        Ka = matern.cov_matrix(K, xo, yo, xo, yo, te)
        ka = matern.cov_matrix(K, xo, yo, xm[p:p + 1], ym[p:p + 1], te)[:, 0]
        w = np.linalg.solve(Ka, ka)
        L = np.linalg.cholesky(Jt)
        Z = L @ rng.standard_normal((Jt.shape[0], 40000))
        err = (Z[0] - w @ Z[1:]) ** 2
        se = err.std() / np.sqrt(err.size)
        assert abs(err.mean() - parts[0, p]) <= 4 * se, (p, err.mean(), parts[0, p], se)
