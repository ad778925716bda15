"""GPU adjoint gradient: Sigma^-1 from the tile factor (potri) and
d l / d theta through MLE_alg's grad argument.

Checks: potri against numpy's inverse; the gradient against the CPU oracle's
dense gradient (itself pinned to finite differences on the CPU); against
central finite differences of the GPU FP64 l (h = 1e-6 relative) at a size
the oracle does not reach.
"""
import os

import numpy as np
import pytest

from oracle import datagen, likelihood, matern

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
K = "univariate_matern_stationary"
KN = "univariate_matern_nuggets_stationary"


@pytest.fixture(scope="module")
def lib():
    import exageostat_amd
    exageostat_amd.load()
    assert exageostat_amd.device_count() >= 1, "no HIP device visible"
    return exageostat_amd


@pytest.mark.parametrize("n,ts", [(700, 64), (1600, 320), (2500, 960), (1000, 192), (1100, 128)])
def test_potri_matches_inverse(lib, n, ts):
    """Lower tiles (and full diagonal tiles) of Sigma^-1 within 1e-12 of
    numpy's inverse, relative to max |Sigma^-1|; padded last tile."""
    x, y, _ = datagen.generate_xy_loc(n, 5)
    theta = (1.0, 0.1, 0.5)
    S = matern.core_dcmg(x, y, x, y, theta) + 0.05 * np.eye(n)
    M = lib.DeviceMatrix(n, n, ts, lower=True)
    M.from_numpy(S)
    assert M.potrf() == 0
    M.potri()
    Z = M.to_numpy()
    R = np.linalg.inv(S)
    mt = (n + ts - 1) // ts
    mask = np.zeros((n, n), dtype=bool)
    for tn in range(mt):
        mask[tn * ts:, tn * ts:(tn + 1) * ts] = True
    err = np.max(np.abs(Z[mask] - R[mask])) / np.max(np.abs(R))
    assert err <= 1e-12, err


@pytest.fixture(scope="module")
def c1():
    return dict(np.load(os.path.join(GOLD, "c1_seed0.npz")))


@pytest.mark.parametrize("theta", [(1.0, 0.1, 0.5), (1.3, 0.07, 0.8), (0.7, 0.15, 1.37), (1.0, 0.1, 1.5)])
def test_gradient_matches_oracle(lib, c1, theta):
    """n = 1600 (ts = 320): every component within 1e-8 of the oracle's
    dense gradient, relative to the gradient's largest component."""
    l_ref, g_ref = likelihood.gradient(K, c1["x"], c1["y"], c1["z"], theta)
    m = lib.ExactMLE(c1["x"], c1["y"], 320)
    m.set_z(c1["z"])
    val, g = m.loglik(theta, grad=True)
    assert m.status == 0
    assert abs(val - l_ref) <= 1e-9 * abs(l_ref)
    scale = np.max(np.abs(g_ref))
    assert np.max(np.abs(g - g_ref)) <= 1e-8 * scale, (g, g_ref)
    # the context keeps working after the in-place inverse
    val2 = m.loglik(theta)
    assert abs(val2 - l_ref) <= 1e-9 * abs(l_ref)
    m.close()


def test_gradient_nugget_matches_oracle(lib):
    g = np.load(os.path.join(GOLD, "nugget_900.npz"))
    theta = np.array(g["theta"], dtype=float)
    l_ref, g_ref = likelihood.gradient(KN, g["x"], g["y"], g["z"], theta)
    m = lib.ExactMLE(g["x"], g["y"], 64, kernel=KN)
    m.set_z(g["z"])
    val, gr = m.loglik(theta, grad=True)
    assert len(gr) == 4
    assert abs(val - l_ref) <= 1e-9 * abs(l_ref)
    assert np.max(np.abs(gr - g_ref)) <= 1e-8 * np.max(np.abs(g_ref)), (gr, g_ref)
    m.close()


def test_gradient_bivariate_matches_oracle(lib):
    """Parsimonious bivariate kernel, all six parameters (n = 800 rows,
    ts = 64), against the oracle's dense gradient (pinned to finite
    differences of l on the CPU)."""
    g = np.load(os.path.join(GOLD, "bivariate_400.npz"))
    KB = "bivariate_matern_parsimonious"
    for theta in (np.array(g["theta"], dtype=float), np.array([0.8, 1.2, 0.12, 0.7, 1.3, -0.4])):
        l_ref, g_ref = likelihood.gradient(KB, g["x"], g["y"], g["z"], theta)
        m = lib.ExactMLE(g["x"], g["y"], 64, kernel=KB)
        m.set_z(g["z"])
        val, gr = m.loglik(theta, grad=True)
        m.close()
        assert len(gr) == 6
        assert abs(val - l_ref) <= 1e-9 * abs(l_ref)
        assert np.max(np.abs(gr - g_ref)) <= 1e-8 * np.max(np.abs(g_ref)), (theta, gr, g_ref)


def test_gradient_central_differences_large(lib):
    """n = 12000 (ts = 960, partial last tile): the adjoint gradient against
    central differences of the GPU FP64 l with h = 1e-6 * theta_p (the FD
    truncation and rounding error is ~1e-6 of the gradient's scale here)."""
    n, ts = 12000, 960
    theta = np.array([1.0, 0.1, 0.5])
    x, y, _ = datagen.generate_xy_loc(n, 2)
    e, _ = datagen.dlarnv(3, [0, 0, 0, 1], n)
    m = lib.ExactMLE(x, y, ts)
    m.zvg(theta, e)
    th = np.array([0.9, 0.12, 0.6])
    _, g = m.loglik(th, grad=True)
    fd = np.zeros(3)
    for p in range(3):
        h = 1e-6 * th[p]
        tp, tm = th.copy(), th.copy()
        tp[p] += h
        tm[p] -= h
        fd[p] = (m.loglik(tp) - m.loglik(tm)) / (2 * h)
    m.close()
    assert np.max(np.abs(g - fd)) <= 1e-5 * np.max(np.abs(fd)), (g, fd)


def test_gradient_non_spd(lib, c1):
    m = lib.ExactMLE(c1["x"], c1["y"], 320)
    m.set_z(c1["z"])
    val, g = m.loglik([-1.0, 0.1, 0.5], grad=True)
    assert val == -np.inf
    assert m.status > 0
    assert np.all(np.isnan(g))
    m.close()


def test_lbfgsb_reaches_derivative_free_theta(lib, c1):
    """C1 (n = 1,600, ts = 320, bounds [0.01, 5]^3 as in the reference's
    example): L-BFGS-B on the adjoint gradient and the derivative-free
    comparator reach the same theta-hat within 1e-3 (relative, per
    component), the gradient method with far fewer evaluations."""
    from exageostat_amd import optimize
    m = lib.ExactMLE(c1["x"], c1["y"], 320)
    m.set_z(c1["z"])
    lb, ub = [0.01] * 3, [5.0] * 3
    start = [0.5, 0.05, 1.0]
    r_g = optimize.fit(m, lb, ub, start=start, method="lbfgsb", opt_tol=7)
    r_d = optimize.fit(m, lb, ub, start=start, method="dfo", opt_tol=7, max_evals=3000)
    m.close()
    assert abs(r_g.loglik - r_d.loglik) <= 1e-7 * abs(r_d.loglik), (r_g.loglik, r_d.loglik)
    assert np.max(np.abs(r_g.theta - r_d.theta) / np.abs(r_d.theta)) <= 1e-3, (r_g.theta, r_d.theta)
    assert r_g.evals * 3 < r_d.evals, (r_g.evals, r_d.evals)


def test_nugget_lbfgsb_against_bobyqa_at_1600(lib, c1):
    """The n = 40,000 optimiser comparison (nugget kernel, true theta = (1,
    0.1, 0.5, 0.01), L-BFGS-B on the adjoint gradient from (0.5, 0.05, 1.0,
    0.05) against BOBYQA over MLE_alg from the lower bounds, bounds
    [0.01, 5]) on the C1 locations: with a tight stopping rule (ftol_abs
    1e-9) both reach the same maximum, to 1e-8 relative in l and 1e-3 per
    component in theta, the gradient method with at least three times fewer
    evaluations."""
    from exageostat_amd import optimize
    kn = "univariate_matern_nuggets_stationary"
    e, _ = datagen.dlarnv(3, [0, 0, 0, 1], len(c1["x"]))
    z = likelihood.dzvg(kn, c1["x"], c1["y"], (1.0, 0.1, 0.5, 0.01), e)
    m = lib.ExactMLE(c1["x"], c1["y"], 320, kernel=kn)
    m.set_z(z)
    lb, ub = [0.01] * 4, [5.0] * 4
    rg = optimize.fit(m, lb, ub, start=[0.5, 0.05, 1.0, 0.05], method="lbfgsb", opt_tol=9)
    rb = optimize.fit(m, lb, ub, method="bobyqa", opt_tol=9, max_evals=5000)
    m.close()
    print(f"nugget n=1600: lbfgsb {rg.theta} l={rg.loglik:.10f} {rg.evals} evals; bobyqa {rb.theta} "
          f"l={rb.loglik:.10f} {rb.evals} evals", flush=True)
    assert abs(rg.loglik - rb.loglik) <= 1e-8 * abs(rb.loglik), (rg.loglik, rb.loglik)
    assert np.max(np.abs(rg.theta - rb.theta) / np.abs(rb.theta)) <= 1e-3, (rg.theta, rb.theta)
    assert 3 * rg.evals <= rb.evals, (rg.evals, rb.evals)


@pytest.mark.parametrize("theta", [(1.0, 0.1, 0.5), (1.3, 0.07, 0.8)])
def test_fisher_matches_oracle(lib, c1, theta):
    """C1 locations (n = 1,600, ts = 320): F and F^-1 against the oracle
    (itself pinned by closed forms and the covariance of the score)."""
    F_ref, Finv_ref = likelihood.fisher(K, c1["x"], c1["y"], theta)
    m = lib.ExactMLE(c1["x"], c1["y"], 320)
    m.set_z(c1["z"])
    F, Finv = m.fisher(theta)
    m.close()
    assert np.max(np.abs(F - F_ref)) <= 1e-9 * np.max(np.abs(F_ref)), (F, F_ref)
    assert np.max(np.abs(Finv - Finv_ref)) <= 1e-7 * np.max(np.abs(Finv_ref)), (Finv, Finv_ref)


def test_fisher_nugget_ragged(lib):
    """Nugget kernel (4 x 4), n = 1,100 with ts = 128 (ragged last tile)."""
    x, y, _ = datagen.generate_xy_loc(1100, 4)
    theta = (1.0, 0.1, 0.5, 0.05)
    KN_ = "univariate_matern_nuggets_stationary"
    F_ref, _ = likelihood.fisher(KN_, x, y, theta)
    m = lib.ExactMLE(x, y, 128, kernel=KN_)
    m.set_z(np.zeros(len(x)))
    F, Finv = m.fisher(theta)
    m.close()
    assert F.shape == (4, 4)
    assert np.max(np.abs(F - F_ref)) <= 1e-9 * np.max(np.abs(F_ref)), (F, F_ref)
    np.testing.assert_allclose(F @ Finv, np.eye(4), atol=1e-8)


def test_fisher_tile_entry_point(lib, c1):
    """EXAGEOSTAT_Fisher_Tile (the reference's signature) returns F^-1 as a
    malloc'd array of 3 * 3 + 1 doubles."""
    import ctypes
    theta = np.array([1.0, 0.1, 0.5])
    _, Finv_ref = likelihood.fisher(K, c1["x"], c1["y"], theta)
    m = lib.ExactMLE(c1["x"], c1["y"], 320)
    m.set_z(c1["z"])
    r = m.lib.EXAGEOSTAT_Fisher_Tile(ctypes.byref(m.data), 1600, theta.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     320, 1, 1)
    assert r
    got = np.array([r[i] for i in range(10)])
    ctypes.CDLL(None).free(ctypes.cast(r, ctypes.c_void_p))
    m.close()
    assert got[9] == 0.0
    Finv = got[:9].reshape(3, 3, order="F")
    assert np.max(np.abs(Finv - Finv_ref)) <= 1e-7 * np.max(np.abs(Finv_ref))


def test_fisher_large_closed_forms(lib):
    """n = 12,000 (ts = 960): F_00 = n / (2 sigma^4) exactly, and F_0k equals
    -g_k / sigma^2 for the adjoint gradient g at z = 0 (two GPU paths)."""
    n, ts = 12000, 960
    theta = np.array([1.2, 0.1, 0.7])
    x, y, _ = datagen.generate_xy_loc(n, 2)
    m = lib.ExactMLE(x, y, ts)
    m.set_z(np.zeros(n))
    F, Finv = m.fisher(theta)
    _, g = m.loglik(theta, grad=True)
    m.close()
    assert abs(F[0, 0] - n / (2 * theta[0] ** 2)) <= 1e-9 * F[0, 0], F[0, 0]
    np.testing.assert_allclose(F[0], -g / theta[0], rtol=1e-8)
    np.testing.assert_allclose(F @ Finv, np.eye(3), atol=1e-7)
