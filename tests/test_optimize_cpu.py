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


# ------------------------------------------------------------- BOBYQA ---
def _bobyqa(f, lb, ub, x0, ftol=1e-10, maxeval=2000):
    import ctypes

    from exageostat_amd import _lib
    lib = _lib.load()
    calls = []

    def cb(n, xp, gp, d):
        x = np.array([xp[i] for i in range(n)])
        calls.append(x)
        return float(f(x))
    fn = _lib.OBJECTIVE(cb)
    x = np.array(x0, dtype=float)
    lb = np.array(lb, dtype=float)
    ub = np.array(ub, dtype=float)
    fmax = ctypes.c_double(0)
    nev = ctypes.c_int(0)
    rc = lib.exageostat_bobyqa(len(x), fn, None, lb.ctypes.data_as(_lib.c_double_p),
                               ub.ctypes.data_as(_lib.c_double_p), x.ctypes.data_as(_lib.c_double_p), ftol,
                               maxeval, ctypes.byref(fmax), ctypes.byref(nev))
    return rc, x, fmax.value, nev.value, calls


def test_bobyqa_interior_maximum_of_a_concave_quadratic():
    A = np.array([[3.0, 0.5, 0.2], [0.5, 2.0, -0.3], [0.2, -0.3, 1.0]])
    c = np.array([1.2, 0.4, 2.1])
    rc, x, fmax, nev, calls = _bobyqa(lambda v: -0.5 * (v - c) @ A @ (v - c) + 7.0, [0.01] * 3, [5.0] * 3,
                                      [0.01] * 3)
    assert rc == 0
    assert np.max(np.abs(x - c)) < 1e-4 and abs(fmax - 7.0) < 1e-8
    # every evaluation stayed inside the box; the first is the start
    assert all(np.all(p >= 0.01) and np.all(p <= 5.0) for p in calls)
    assert np.array_equal(calls[0], [0.01] * 3)
    # 2n+1 initial points: start, +rhobeg and +2 rhobeg per variable
    # (rhobeg = (5 - 0.01)/4, NLOPT's default step)
    rb = (5.0 - 0.01) / 4
    assert np.allclose(calls[1], [0.01 + rb, 0.01, 0.01]) and np.allclose(calls[4], [0.01 + 2 * rb, 0.01, 0.01])
    assert nev == len(calls)


def test_bobyqa_maximum_on_the_bounds_and_fixed_variables():
    # unconstrained maximum at (-1, 6, 2): the box makes it (0.01, 5, 2)
    c = np.array([-1.0, 6.0, 2.0])
    rc, x, _, _, calls = _bobyqa(lambda v: -np.sum((v - c) ** 2), [0.01] * 3, [5.0] * 3, [0.01] * 3)
    assert np.allclose(x, [0.01, 5.0, 2.0], atol=1e-4)
    # lb == ub holds a variable fixed (NLOPT eliminates it)
    rc, x, _, _, calls = _bobyqa(lambda v: -np.sum((v - c) ** 2), [0.01, 0.3, 0.01], [5.0, 0.3, 5.0],
                                 [1.0, 0.3, 1.0])
    assert all(p[1] == 0.3 for p in calls) and np.allclose(x[[0, 2]], [0.01, 2.0], atol=1e-4)


def test_bobyqa_rosenbrock_and_maxeval():
    def f(v):
        return -(100.0 * (v[1] - v[0] ** 2) ** 2 + (1.0 - v[0]) ** 2)
    rc, x, fmax, nev, _ = _bobyqa(f, [-2.0, -2.0], [2.0, 2.0], [-1.2, 1.0], ftol=1e-14, maxeval=3000)
    assert np.max(np.abs(x - 1.0)) < 1e-3, x
    rc, x, _, nev, _ = _bobyqa(f, [-2.0, -2.0], [2.0, 2.0], [-1.2, 1.0], ftol=1e-14, maxeval=12)
    assert rc == 1 and nev == 12


def test_bobyqa_reaches_the_gradient_method_theta():
    """The reference's MLE run (BOBYQA from the lower bounds, ftol_abs =
    10^-opt_tol) on the oracle likelihood reaches the L-BFGS-B theta hat
    within 1e-3, with more evaluations than the gradient method."""
    m = _problem()
    lb, ub = [0.01, 0.01, 0.01], [5.0, 5.0, 5.0]
    r_b = optimize.fit(m, lb, ub, method="bobyqa", opt_tol=7, max_evals=2000)
    r_g = optimize.fit(m, lb, ub, method="lbfgsb", opt_tol=9)
    assert np.max(np.abs(r_b.theta - r_g.theta) / np.abs(r_g.theta)) <= 1e-3, (r_b.theta, r_g.theta)
    assert r_b.loglik >= r_g.loglik - 1e-6 * abs(r_g.loglik)
    assert r_g.evals < r_b.evals


class NoisyPrecisionMLE(OracleMLE):
    """OracleMLE with a precision switch: at precision 2 l carries a small
    deterministic perturbation (standing in for the mixed path's rounding),
    and the precision used by each evaluation is recorded."""

    def __init__(self, x, y, z):
        super().__init__(x, y, z)
        self.precision = 0
        self.calls = []

    def loglik(self, theta, grad=False):
        self.calls.append(self.precision)
        out = super().loglik(theta, grad)
        if self.precision != 2:
            return out
        bump = 1e-7 * np.sin(1e3 * float(np.sum(theta)))
        return (out[0] + bump, out[1]) if grad else out + bump


def test_lbfgsb_mixed_then_fp64():
    """method "lbfgsb_mixed": a precision-2 stage, then an FP64 stage from
    its point; theta hat is the FP64 maximum (that of plain "lbfgsb" within
    1e-4), the model's precision is restored, and both stages ran."""
    base = _problem()
    m = NoisyPrecisionMLE(base.x, base.y, base.z)
    lb, ub = [0.01, 0.01, 0.01], [5.0, 5.0, 5.0]
    start = [0.5, 0.05, 1.0]
    ref = optimize.fit(m, lb, ub, start=start, method="lbfgsb", opt_tol=7)
    m.calls.clear()
    r = optimize.fit(m, lb, ub, start=start, method="lbfgsb_mixed", opt_tol=7)
    assert m.precision == 0
    assert 2 in m.calls and m.calls[-1] == 0
    first_fp64 = m.calls.index(0)
    assert all(p == 2 for p in m.calls[:first_fp64]) and all(p == 0 for p in m.calls[first_fp64:])
    assert r.evals == len(m.calls)
    assert np.max(np.abs(r.theta - ref.theta) / np.abs(ref.theta)) <= 1e-4, (r.theta, ref.theta)
    assert abs(r.loglik - ref.loglik) <= 1e-9 * abs(ref.loglik)
