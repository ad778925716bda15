"""Matern covariance kernels of the reference, restated with numpy/scipy.

Test infrastructure only (see ``oracle/__init__.py``).

K_nu is ``scipy.special.kv`` (AMOS), pinned in the tests against mpmath
``besselk`` at 30 digits and against the half-integer closed forms; it
stands in for GSL 2.6 ``gsl_sf_bessel_Knu``, which the reference calls
(core_zcmg.c:987) and which is not available here.

Pinning of the covariance formulas (tests/test_oracle_matern.py):
* nu = 1/2 reduces to sigma^2 exp(-d/beta); nu = 3/2 to
  sigma^2 (1 + d/beta) exp(-d/beta) (closed forms);
* C(0) = sigma^2 (+ tau^2 with the nugget), symmetry, and positive
  definiteness of the generated matrices;
* the bivariate parsimonious 2x2 blocks reduce to the univariate kernel
  when nu1 = nu2, sigma1 = sigma2 and rho = 1 (then all four entries agree);
* the derivative kernels match central differences of ``cov_matrix``.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.special as sp

EARTH_RADIUS_KM = 6371.0  # misc/include/MLE_misc.h:68


def distance_matrix(x1, y1, x2, y2, metric: int = 0) -> np.ndarray:
    """calculateDistance (core_zcmg.c:71-92) between every pair.

    metric 0: Euclidean sqrt((x2-x1)^2 + (y2-y1)^2).
    metric 1: haversine great-circle distance in km (distanceEarth,
    core_zcmg.c:51-60), coordinates in degrees (x = longitude-like first
    argument, as the reference passes (x1, y1, x2, y2)).
    """
    x1 = np.asarray(x1, dtype=np.float64)[:, None]
    y1 = np.asarray(y1, dtype=np.float64)[:, None]
    x2 = np.asarray(x2, dtype=np.float64)[None, :]
    y2 = np.asarray(y2, dtype=np.float64)[None, :]
    if metric == 1:
        lat1r = np.deg2rad(x1)
        lon1r = np.deg2rad(y1)
        lat2r = np.deg2rad(x2)
        lon2r = np.deg2rad(y2)
        u = np.sin((lat2r - lat1r) / 2)
        v = np.sin((lon2r - lon1r) / 2)
        return 2.0 * EARTH_RADIUS_KM * np.arcsin(np.sqrt(u * u + np.cos(lat1r) * np.cos(lat2r) * v * v))
    dx = x2 - x1
    dy = y2 - y1
    return np.sqrt(dx * dx + dy * dy)


def matern_correlation(expr: np.ndarray, nu: float) -> np.ndarray:
    """(1/(2^(nu-1) Gamma(nu))) * expr^nu * K_nu(expr); 1 where expr == 0."""
    expr = np.asarray(expr, dtype=np.float64)
    con = 1.0 / (math.pow(2.0, nu - 1.0) * math.gamma(nu))
    out = np.ones_like(expr)
    nz = expr != 0
    e = expr[nz]
    out[nz] = con * np.power(e, nu) * sp.kv(nu, e)
    return out


def core_dcmg(x1, y1, x2, y2, theta, metric: int = 0) -> np.ndarray:
    """core_dcmg (core_zcmg.c:963-994), all m x n entries at once.

    theta = (sigma^2, beta, nu): A = sigma^2 where d/beta == 0, else
    sigma^2/(2^(nu-1) Gamma(nu)) (d/beta)^nu K_nu(d/beta).
    """
    s2, beta, nu = float(theta[0]), float(theta[1]), float(theta[2])
    expr = distance_matrix(x1, y1, x2, y2, metric) / beta
    return s2 * matern_correlation(expr, nu)


def core_dcmg_nuggets(x1, y1, x2, y2, theta, metric: int = 0) -> np.ndarray:
    """core_dcmg_nuggets (core_zcmg_nuggets.c:107-156, 2-D branch):
    as core_dcmg plus theta[3] where the distance is exactly zero."""
    s2, beta, nu, tau2 = (float(t) for t in theta[:4])
    expr = distance_matrix(x1, y1, x2, y2, metric) / beta
    out = s2 * matern_correlation(expr, nu)
    out[expr == 0] = s2 + tau2
    return out


def parsimonious_constants(theta):
    """Constants of core_dcmg_bivariate_parsimonious (core_zcmg.c:1628-1655).

    theta = (sigma1^2, sigma2^2, beta, nu1, nu2, rho).  Returns
    (nu12, rho_tilde, cross sigma = rho_tilde*sqrt(sigma1^2 sigma2^2)).
    """
    s1, s2, beta, nu1, nu2, rho = (float(t) for t in theta[:6])
    nu12 = 0.5 * (nu1 + nu2)
    rho_t = rho * math.sqrt((math.gamma(nu1 + 1) * math.gamma(nu2 + 1)) /
                            (math.gamma(nu1) * math.gamma(nu2))) * math.gamma(nu12) / math.gamma(nu12 + 1)
    return nu12, rho_t, rho_t * math.sqrt(s1 * s2)


def core_dcmg_bivariate_parsimonious(x1, y1, x2, y2, theta, metric: int = 0) -> np.ndarray:
    """core_dcmg_bivariate_parsimonious (core_zcmg.c:1628-1682) for whole
    location sets: returns the (2*m1) x (2*m2) matrix whose row 2p+a and
    column 2q+b hold the covariance of variable a at location p with
    variable b at location q (variables interleaved, a, b in {0, 1})."""
    s1, s2, beta, nu1, nu2, _ = (float(t) for t in theta[:6])
    nu12, _, s12 = parsimonious_constants(theta)
    expr = distance_matrix(x1, y1, x2, y2, metric) / beta
    m1, m2 = expr.shape
    out = np.empty((2 * m1, 2 * m2))
    out[0::2, 0::2] = s1 * matern_correlation(expr, nu1)
    c12 = s12 * matern_correlation(expr, nu12)
    out[1::2, 0::2] = c12
    out[0::2, 1::2] = c12
    out[1::2, 1::2] = s2 * matern_correlation(expr, nu2)
    return out


def _matern_parts(expr: np.ndarray, nu: float, beta: float):
    """For expr = d/beta: the correlation M (1 at d = 0) and its derivatives
    dM/dbeta and dM/dnu (both 0 at d = 0), as in matern_derivatives."""
    con = 1.0 / (math.pow(2.0, nu - 1.0) * math.gamma(nu))
    nz = expr != 0
    e = expr[nz]
    knu = sp.kv(nu, e)
    xnu = np.power(e, nu)
    M = np.ones_like(expr)
    M[nz] = con * xnu * knu
    dbeta = np.zeros_like(expr)
    dbeta[nz] = con * (e * xnu * sp.kv(nu + 1, e) - 2.0 * nu * xnu * knu) / beta
    dnu = np.zeros_like(expr)
    dnu[nz] = con * xnu * ((-math.log(2.0) - sp.digamma(nu) + np.log(e)) * knu + dkv_dnu(nu, e))
    return M, dbeta, dnu


def bivariate_derivatives(x1, y1, x2, y2, theta, metric: int = 0):
    """Partial derivatives of core_dcmg_bivariate_parsimonious with respect to
    theta = (sigma1^2, sigma2^2, beta, nu1, nu2, rho), as six interleaved
    (2*m1) x (2*m2) matrices.  The reference has no derivative of this
    kernel; with rho~ = rho*sqrt(nu1 nu2)/nu12 (the gamma ratios of
    parsimonious_constants reduce to this) the cross block is
    C12 = rho~ sqrt(s1 s2) M(d; nu12), so
      dC12/ds_a  = C12 / (2 s_a),
      dC12/dnu_a = C12 (1/(2 nu_a) - 1/(2 nu12)) + rho~ sqrt(s1 s2) dM/dnu / 2,
      dC12/drho  = C12 / rho (written as sqrt(nu1 nu2)/nu12 sqrt(s1 s2) M)."""
    s1, s2, beta, nu1, nu2, rho = (float(t) for t in theta[:6])
    nu12, rho_t, s12 = parsimonious_constants(theta)
    expr = distance_matrix(x1, y1, x2, y2, metric) / beta
    m1, m2 = expr.shape
    M1, b1, n1 = _matern_parts(expr, nu1, beta)
    M2, b2, n2 = _matern_parts(expr, nu2, beta)
    M12, b12, n12 = _matern_parts(expr, nu12, beta)
    out = [np.zeros((2 * m1, 2 * m2)) for _ in range(6)]

    def put(mat, a11, a12, a22):
        mat[0::2, 0::2] = a11
        mat[1::2, 0::2] = a12
        mat[0::2, 1::2] = a12
        mat[1::2, 1::2] = a22

    z = np.zeros_like(expr)
    c12 = s12 * M12
    put(out[0], M1, c12 / (2.0 * s1), z)
    put(out[1], z, c12 / (2.0 * s2), M2)
    put(out[2], s1 * b1, s12 * b12, s2 * b2)
    put(out[3], s1 * n1, c12 * (0.5 / nu1 - 0.5 / nu12) + 0.5 * s12 * n12, z)
    put(out[4], z, c12 * (0.5 / nu2 - 0.5 / nu12) + 0.5 * s12 * n12, s2 * n2)
    put(out[5], z, math.sqrt(nu1 * nu2) / nu12 * math.sqrt(s1 * s2) * M12, z)
    return out


def core_dcmg_bivariate_parsimonious2(x1, y1, x2, y2, theta, metric: int = 0) -> np.ndarray:
    """The block layout of "bivariate_matern_parsimonious2"
    (exageostat_exact/src/compute/MLE_dexact.c:323-360): three univariate
    core_dcmg generations into the quadrants of C.  Rows 0..m1-1 are
    variable 1 at locations 0..m1-1, rows m1..2m1-1 variable 2 (columns
    likewise with m2):
      C11 = core_dcmg(sigma1^2, beta, nu1),
      C21 = C12 = core_dcmg(rho~ sqrt(sigma1^2 sigma2^2), beta, nu12),
      C22 = core_dcmg(sigma2^2, beta, nu2)."""
    s1, s2, beta, nu1, nu2, _ = (float(t) for t in theta[:6])
    nu12, _, s12 = parsimonious_constants(theta)
    m1, m2 = len(x1), len(x2)
    out = np.empty((2 * m1, 2 * m2))
    out[:m1, :m2] = core_dcmg(x1, y1, x2, y2, (s1, beta, nu1), metric)
    c12 = core_dcmg(x1, y1, x2, y2, (s12, beta, nu12), metric)
    out[m1:, :m2] = c12
    out[:m1, m2:] = c12
    out[m1:, m2:] = core_dcmg(x1, y1, x2, y2, (s2, beta, nu2), metric)
    return out


def core_dcmg_spacetime_matern(x1, y1, t1, x2, y2, t2, theta, metric: int = 0) -> np.ndarray:
    """core_dcmg_spacetime_matern (exageostat_exact/cpu_core/compute/
    core_zcmg.c:1320-1357), Gneiting (2002) eq. 16 as the reference writes
    it.  theta = (sigma^2, spatial range, nu, time range a, time smoothness
    alpha, non-separability beta, delta); t is the location's z coordinate
    (time).  With the spatial distance d (2-D: calculateDistance with
    z_flag = 1 leaves time out) and u = |t_i - t_j|:
      psi  = u^(2 alpha) / a + 1,
      x    = (d / range) / psi^(beta/2),
      A    = sigma^2 / (2^(nu-1) Gamma(nu)) x^nu K_nu(x) / psi^(beta+delta),
    and A = sigma^2 / psi^(beta+delta) where d / range == 0."""
    s2, rng, nu, a, alpha, beta, delta = (float(v) for v in theta[:7])
    ds = distance_matrix(x1, y1, x2, y2, metric) / rng
    u = np.abs(np.asarray(t1, dtype=np.float64)[:, None] - np.asarray(t2, dtype=np.float64)[None, :])
    psi = np.power(u, 2.0 * alpha) / a + 1.0
    x = ds / np.power(psi, beta / 2.0)
    corr = matern_correlation(x, nu)
    corr[ds == 0] = 1.0
    return s2 * corr / np.power(psi, beta + delta)


KERNELS = {
    "univariate_matern_stationary": (3, core_dcmg),
    "univariate_matern_nuggets_stationary": (4, core_dcmg_nuggets),
    "bivariate_matern_parsimonious": (6, core_dcmg_bivariate_parsimonious),
    "bivariate_matern_parsimonious_profile": (6, core_dcmg_bivariate_parsimonious),
    "bivariate_matern_parsimonious2": (6, core_dcmg_bivariate_parsimonious2),
    "bivariate_matern_parsimonious2_profile": (6, core_dcmg_bivariate_parsimonious2),
    "univariate_spacetime_matern_stationary": (7, None),
}


def cov_matrix(kernel: str, x1, y1, x2, y2, theta, metric: int = 0, t1=None, t2=None) -> np.ndarray:
    """Dispatch on the reference's kernel strings (codelet_zcmg.c:262-310;
    the "_profile" variants generate the same covariance).  t1/t2: the time
    coordinates of the space-time kernel."""
    try:
        _, fn = KERNELS[kernel]
    except KeyError:
        raise ValueError(f"unknown kernel {kernel!r}") from None
    if kernel == "univariate_spacetime_matern_stationary":
        if t1 is None or t2 is None:
            raise ValueError("the space-time kernel needs time coordinates")
        return core_dcmg_spacetime_matern(x1, y1, t1, x2, y2, t2, theta, metric)
    return fn(x1, y1, x2, y2, theta, metric)


def num_vars(kernel: str) -> int:
    """Covariance rows per location (2 for the bivariate kernels)."""
    return 2 if kernel.startswith("bivariate") else 1


def num_params(kernel: str) -> int:
    return KERNELS[kernel][0]


def dkv_dnu(nu: float, x: np.ndarray, h: float = 0.05) -> np.ndarray:
    """d K_nu(x) / d nu from the integral representation
    K_nu(x) = int_0^inf exp(-x cosh t) cosh(nu t) dt, differentiated under
    the integral: int_0^inf t exp(-x cosh t) sinh(nu t) dt.  The integrand is
    even and analytic, so the trapezoidal rule with step h converges
    geometrically; the range is cut where x cosh t - nu t > 800.  Agrees with
    30-digit mpmath to ~1e-15 relative on the golden table (a finite
    difference of scipy kv only reaches ~1e-7 because kv is not smooth in nu
    across its internal branches)."""
    x = np.asarray(x, dtype=np.float64)
    flat = x.ravel()
    out = np.empty_like(flat)
    if flat.size == 0:
        return out.reshape(x.shape)
    xmin = float(np.min(flat))
    T = float(np.arccosh(max(800.0 / xmin, 1.0))) + 1.0
    for _ in range(6):
        T = float(np.arccosh(max((800.0 + abs(nu) * T) / xmin, 1.0))) + 0.5
    t = np.arange(0.0, T + h, h)
    w = np.full(t.shape, h)
    w[0] = 0.5 * h
    g = w * t * np.sinh(nu * t)
    ch = np.cosh(t)
    for i in range(0, flat.size, 20000):
        out[i:i + 20000] = np.exp(-np.outer(flat[i:i + 20000], ch)) @ g
    return out.reshape(x.shape)


def matern_derivatives(x1, y1, x2, y2, theta, metric: int = 0, nugget: bool = False):
    """Partial derivatives of the (nugget) Matern covariance with respect to
    each parameter, as full m x n matrices.

    d/dsigma^2: the correlation (1 at d = 0), as core_dcmg_matern_dsigma_square
    (core_zcmg.c:315-342).
    d/dbeta: sigma^2 c [-(nu/beta) x^nu K_nu(x) - x^nu K'_nu(x) x/beta], with
    K'_nu = (nu/x) K_nu - K_{nu+1} (core_zcmg.c:189-191, 453-488); 0 at d = 0.
    d/dnu: derivative of sigma^2 x^nu K_nu(x) / (2^(nu-1) Gamma(nu)):
    sigma^2 c x^nu [ (-ln 2 - psi(nu) + ln x) K_nu + dK_nu/dnu ]
    (the quantity core_dcmg_matern_dnu, core_zcmg.c:380-416, approximates);
    0 at d = 0.
    d/dtau^2 (nugget only): 1 where d == 0.
    """
    s2, beta, nu = float(theta[0]), float(theta[1]), float(theta[2])
    d = distance_matrix(x1, y1, x2, y2, metric)
    expr = d / beta
    con = 1.0 / (math.pow(2.0, nu - 1.0) * math.gamma(nu))
    nz = expr != 0
    e = expr[nz]
    knu = sp.kv(nu, e)
    xnu = np.power(e, nu)
    dsig = np.ones_like(expr)
    dsig[nz] = con * xnu * knu
    dk_dx = (nu / e) * knu - sp.kv(nu + 1, e)
    dbeta = np.zeros_like(expr)
    dbeta[nz] = s2 * con * (-(nu / beta) * xnu * knu - xnu * dk_dx * e / beta)
    dnu = np.zeros_like(expr)
    dnu[nz] = s2 * con * xnu * ((-math.log(2.0) - sp.digamma(nu) + np.log(e)) * knu + dkv_dnu(nu, e))
    out = [dsig, dbeta, dnu]
    if nugget:
        out.append((expr == 0).astype(np.float64))
    return out
