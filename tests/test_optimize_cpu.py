"""The MLE drivers (L-BFGS-B on the adjoint gradient, derivative-free
COBYQA) on a small problem, with the oracle standing in for the library's
MLE_alg so the driver logic is tested without a GPU."""
import numpy as np

from exageostat_amd import optimize
from oracle import datagen, likelihood

K = "univariate_matern_stationary"


class OracleMLE:
    """Duck-type of ExactMLE.loglik backed by the dense oracle."""

    def __init__(self, x, y, z):
        self.x, self.y, self.z = x, y, z
        self.N = len(z)

    def loglik(self, theta, grad=False):
        theta = np.asarray(theta, dtype=float)
        if grad:
            try:
                return likelihood.gradient(K, self.x, self.y, self.z, theta)
            except np.linalg.LinAlgError:
                return -np.inf, np.full(len(theta), np.nan)
        try:
            return likelihood.dmle(K, self.x, self.y, self.z, theta)[0]
        except np.linalg.LinAlgError:
            return -np.inf


def _problem(n=256, theta=(1.0, 0.1, 0.5)):
    x, y, _ = datagen.generate_xy_loc(n, 4)
    e, _ = datagen.dlarnv(3, [0, 0, 0, 1], n)
    z = likelihood.dzvg(K, x, y, theta, e)
    return OracleMLE(x, y, z)


def test_lbfgsb_and_dfo_agree():
    m = _problem()
    lb, ub = [0.01, 0.01, 0.01], [5.0, 5.0, 5.0]
    start = [0.5, 0.05, 1.0]
    r_g = optimize.fit(m, lb, ub, start=start, method="lbfgsb", opt_tol=7)
    r_d = optimize.fit(m, lb, ub, start=start, method="dfo", opt_tol=7, max_evals=400)
    assert np.all(np.isfinite(r_g.theta)) and np.all(np.isfinite(r_d.theta))
    # both reach the same maximum; the gradient method with fewer evaluations
    assert abs(r_g.loglik - r_d.loglik) <= 1e-5 * abs(r_d.loglik)
    assert np.max(np.abs(r_g.theta - r_d.theta) / np.abs(r_d.theta)) <= 1e-3, (r_g.theta, r_d.theta)
    assert r_g.evals < r_d.evals
    # at an interior maximum the gradient vanishes
    _, g = m.loglik(r_g.theta, grad=True)
    assert np.max(np.abs(g)) <= 1e-2


def test_fit_survives_non_spd_region():
    """A start where Sigma is not positive definite gets the penalty and the
    driver still returns the best finite evaluation."""
    m = _problem(n=128)
    r = optimize.fit(m, [0.01, 0.01, 0.01], [5.0, 5.0, 5.0], start=[0.5, 0.05, 1.0], method="lbfgsb",
                     opt_tol=5, max_evals=60)
    assert np.isfinite(r.loglik)
