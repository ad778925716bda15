"""Dense FP64 restatement of the likelihood, prediction and gradient path.

Test infrastructure only (see ``oracle/__init__.py``).

Pinning (tests/test_oracle_likelihood.py):
* ``dmle`` against an independent multivariate-normal log-density
  (``scipy.stats.multivariate_normal.logpdf``) and, for n <= 48, against
  a 40-digit mpmath Cholesky; invariance under a joint permutation of the
  locations and z.
* ``predict`` reproduces the observed value at an observed location
  (noise-free kriging interpolates) and matches an explicit
  C12 @ inv(C22) @ z.
* ``gradient`` against central finite differences of ``dmle``.
* ``mloe_mmom``: every per-location MSE against the quadratic form of the
  joint (observed + held-out) covariance, the optimal MSE against the
  Schur-complement identity 1 / (Sigma_joint^-1)_00, zero MLOE and MMOM at
  the true theta, MLOE >= 0, and a Monte Carlo estimate of the true MSE.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.linalg as sla

from . import matern

LOG_2PI = math.log(2.0 * math.pi)


def sigma(kernel: str, x, y, theta, metric: int = 0, t=None) -> np.ndarray:
    """Full symmetric covariance of one location set (t: time coordinates
    of the space-time kernel)."""
    return matern.cov_matrix(kernel, x, y, x, y, theta, metric, t1=t, t2=t)


def dzvg(kernel: str, x, y, theta, e, metric: int = 0, t=None) -> np.ndarray:
    """EXAGEOSTAT_MLE_dzvg_Tile (MLE_dexact.c:25-148): z = L(theta) e with
    L the lower Cholesky factor of Sigma(theta) (dcmg -> dpotrf -> dtrmm)."""
    L = np.linalg.cholesky(sigma(kernel, x, y, theta, metric, t))
    return L @ np.asarray(e, dtype=np.float64)


def loglik_from_sigma(S: np.ndarray, z: np.ndarray):
    """The steps of EXAGEOSTAT_dmle_Tile after generation
    (MLE_dexact.c:372-426): dpotrf, logdet = 2 sum log L_ii over positive
    pivots (core_zmdet.c:50-59), v = L^-1 z (dtrsm), dotp = v.v (dgemm),
    l = -dotp/2 - logdet/2 - (N/2) log 2 pi.  Returns (l, logdet, dotp)."""
    L = np.linalg.cholesky(S)
    d = np.diag(L)
    logdet = 2.0 * float(np.sum(np.log(d[d > 0])))
    v = sla.solve_triangular(L, np.asarray(z, dtype=np.float64), lower=True)
    dotp = float(v @ v)
    n = S.shape[0]
    return -0.5 * dotp - 0.5 * logdet - (n / 2.0) * LOG_2PI, logdet, dotp


def dmle(kernel: str, x, y, z, theta, metric: int = 0, t=None):
    """EXAGEOSTAT_dmle_Tile (MLE_dexact.c:231-489), dense.
    Returns (l, logdet, dotp)."""
    return loglik_from_sigma(sigma(kernel, x, y, theta, metric, t), z)


def profile_from_solve(kernel: str, v: np.ndarray, logdet: float, approx: bool = False):
    """The profile likelihood of the "_profile" bivariate kernels from
    v = L^-1 z and log|Sigma| (exageostat_exact/src/compute/
    MLE_dexact.c:403-421): with N/2 in integer arithmetic as there,
      l = -(N/2) + (N/2) log N - (N/2) log(v.v) - logdet/2 - (N/2.0) log 2 pi,
    and variance_a = |v_a|^2 / (N/2), v_a the rows of variable a (the first
    and second half of v in the block layout of
    "bivariate_matern_parsimonious2_profile", even and odd rows in the
    interleaved layout).  approx: the band-limited path's formula for the
    interleaved kernel (exageostat_approx/src/compute/MLE_approx.c:386-396),
    which has (N/2.0) log(N/2.0) in place of (N/2) log N.
    Returns (l, variance1, variance2)."""
    N = len(v)
    dotp = float(v @ v)
    half = N // 2
    if kernel.startswith("bivariate_matern_parsimonious2"):
        v1, v2 = v[:half], v[half:]
    else:
        v1, v2 = v[0::2], v[1::2]
    if approx and not kernel.startswith("bivariate_matern_parsimonious2"):
        l = -(N / 2.0) + (N / 2.0) * math.log(N / 2.0) - (N / 2.0) * math.log(dotp) - 0.5 * logdet \
            - (N / 2.0) * LOG_2PI
        return l, float(v1 @ v1) / (N / 2.0), float(v2 @ v2) / (N / 2.0)
    l = -half + half * math.log(N) - half * math.log(dotp) - 0.5 * logdet - (N / 2.0) * LOG_2PI
    return l, float(v1 @ v1) / half, float(v2 @ v2) / half


def dmle_profile(kernel: str, x, y, z, theta, metric: int = 0):
    """EXAGEOSTAT_dmle_Tile with a "_profile" kernel (MLE_dexact.c:372-421),
    dense.  Returns (l, variance1, variance2, logdet, dotp)."""
    S = sigma(kernel, x, y, theta, metric)
    L = np.linalg.cholesky(S)
    d = np.diag(L)
    logdet = 2.0 * float(np.sum(np.log(d[d > 0])))
    v = sla.solve_triangular(L, np.asarray(z, dtype=np.float64), lower=True)
    l, v1, v2 = profile_from_solve(kernel, v, logdet)
    return l, v1, v2, logdet, float(v @ v)


def band_cholesky_tiles(S: np.ndarray, ts: int, band: int) -> np.ndarray:
    """CHAM_pdpotrf_diag (exageostat_approx/src/compute/pdpotrf_diag.c:
    64-110) on the lower tiles of S (ts x ts tiles, ragged last tile):
    for k: potrf(A_kk); A_mk <- A_mk A_kk^-T for k < m < min(mt, k + band);
    for k < n < min(mt, k + band): A_nn -= A_nk A_nk^T and
    A_mn -= A_mk A_nk^T for n < m < min(mt, n + band).  Tiles outside
    |m - n| < band are never written and keep the covariance; A_mk with
    m >= k + band enters the updates unsolved, as there.  Returns the
    dense lower-triangular result (upper triangle zero)."""
    n = S.shape[0]
    mt = (n + ts - 1) // ts
    A = np.tril(np.array(S, dtype=np.float64))
    sl = [slice(i * ts, min(n, (i + 1) * ts)) for i in range(mt)]
    for k in range(mt):
        A[sl[k], sl[k]] = np.linalg.cholesky(A[sl[k], sl[k]])
        Lkk = A[sl[k], sl[k]]
        for m in range(k + 1, min(mt, k + band)):
            A[sl[m], sl[k]] = sla.solve_triangular(Lkk, A[sl[m], sl[k]].T, lower=True).T
        for nn in range(k + 1, min(mt, k + band)):
            A[sl[nn], sl[nn]] -= np.tril(A[sl[nn], sl[k]] @ A[sl[nn], sl[k]].T)
            for m in range(nn + 1, min(mt, nn + band)):
                A[sl[m], sl[nn]] -= A[sl[m], sl[k]] @ A[sl[nn], sl[k]].T
    return A


def dmle_band(kernel: str, x, y, z, theta, ts: int, band: int, metric: int = 0):
    """EXAGEOSTAT_dmle_diag_Tile (exageostat_approx/src/compute/
    MLE_approx.c:201-420): band_cholesky_tiles, logdet = 2 sum log L_ii
    over positive pivots (core_zmdet.c:50-59), v = L^-1 z with the whole
    lower triangle of the result (CHAMELEON_dtrsm_Tile, MLE_approx.c:365),
    l = -v.v/2 - logdet/2 - (N/2) log 2 pi (the profile kernels as in
    profile_from_solve with approx=True).  Returns (l, logdet, dotp) or,
    for the profile kernels, (l, variance1, variance2, logdet, dotp)."""
    S = sigma(kernel, x, y, theta, metric)
    L = band_cholesky_tiles(S, ts, band)
    d = np.diag(L)
    logdet = 2.0 * float(np.sum(np.log(d[d > 0])))
    v = sla.solve_triangular(L, np.asarray(z, dtype=np.float64), lower=True)
    dotp = float(v @ v)
    if kernel.endswith("_profile"):
        l, v1, v2 = profile_from_solve(kernel, v, logdet, approx=True)
        return l, v1, v2, logdet, dotp
    n = S.shape[0]
    return -0.5 * dotp - 0.5 * logdet - (n / 2.0) * LOG_2PI, logdet, dotp


def predict(kernel: str, xobs, yobs, zobs, xmiss, ymiss, theta, zactual=None, metric: int = 0,
            tobs=None, tmiss=None):
    """EXAGEOSTAT_dmle_Predict_Tile (MLE_dexact.c:688-874):
    C22 = Sigma(lobs, lobs), C12 = Sigma(lmiss, lobs), zmiss = C12 C22^-1 zobs
    (dposv + dgemm); mse = sum (zactual - zmiss)^2 / nZmiss
    (codelet_zmse.c:22-38, MLE_dexact.c:841-847).  For the bivariate kernel
    nZmiss counts rows (2 per location) and the per-variable errors are
    divided by nZmiss/2 (codelet_zmse_bivariate.c:22-51).
    Returns (zmiss, mse, mse1, mse2)."""
    C22 = matern.cov_matrix(kernel, xobs, yobs, xobs, yobs, theta, metric, t1=tobs, t2=tobs)
    C12 = matern.cov_matrix(kernel, xmiss, ymiss, xobs, yobs, theta, metric, t1=tmiss, t2=tobs)
    cf = sla.cho_factor(C22, lower=True)
    alpha = sla.cho_solve(cf, np.asarray(zobs, dtype=np.float64))
    zmiss = C12 @ alpha
    if zactual is None:
        return zmiss, -1.0, None, None
    err = (np.asarray(zactual, dtype=np.float64) - zmiss) ** 2
    nm = len(zmiss)
    mse = float(np.sum(err)) / nm
    if kernel.startswith("bivariate_matern_parsimonious2"):
        h = nm // 2
        return zmiss, mse, float(np.sum(err[:h])) / h, float(np.sum(err[h:])) / h
    if kernel.startswith("bivariate"):
        return zmiss, mse, float(np.sum(err[0::2])) / (nm // 2), float(np.sum(err[1::2])) / (nm // 2)
    return zmiss, mse, None, None


def gradient(kernel: str, x, y, z, theta, metric: int = 0):
    """d l / d theta = 1/2 alpha^T dS alpha - 1/2 tr(S^-1 dS), alpha = S^-1 z,
    for the univariate Matern kernels (3 parameters, or 4 with the nugget)
    and the parsimonious bivariate kernel (6 parameters).  Returns (l, grad)."""
    nugget = kernel == "univariate_matern_nuggets_stationary"
    S = sigma(kernel, x, y, theta, metric)
    cf = sla.cho_factor(S, lower=True)
    alpha = sla.cho_solve(cf, np.asarray(z, dtype=np.float64))
    Sinv = sla.cho_solve(cf, np.eye(S.shape[0]))
    if kernel == "bivariate_matern_parsimonious":
        dS = matern.bivariate_derivatives(x, y, x, y, theta, metric)
    else:
        dS = matern.matern_derivatives(x, y, x, y, theta, metric, nugget=nugget)
    g = np.array([0.5 * alpha @ D @ alpha - 0.5 * float(np.sum(Sinv * D)) for D in dS])
    l, _, _ = loglik_from_sigma(S, z)
    return l, g


def fisher(kernel: str, x, y, theta, metric: int = 0):
    """Fisher information F_jk = 1/2 tr(S^-1 dS_j S^-1 dS_k) of the
    univariate Matern kernels, as EXAGEOSTAT_Fisher_Tile forms it
    (exageostat_exact/src/compute/MLE_dexact.c:1777-1934: two triangular
    solves of each derivative matrix with the Cholesky factor, then the
    trace of each product, A[k + 3 j] = 0.5 t for k >= j).  Returns (F, F^-1)
    with F^-1 through the Cholesky factor of F (:1886-1921)."""
    nugget = kernel == "univariate_matern_nuggets_stationary"
    S = sigma(kernel, x, y, theta, metric)
    cf = sla.cho_factor(S, lower=True)
    dS = matern.matern_derivatives(x, y, x, y, theta, metric, nugget=nugget)
    B = [sla.cho_solve(cf, D) for D in dS]           # S^-1 dS_j
    npar = len(dS)
    F = np.zeros((npar, npar))
    for j in range(npar):
        for k in range(j, npar):
            F[j, k] = F[k, j] = 0.5 * float(np.sum(B[j] * B[k].T))   # tr(B_j B_k)
    Lf = np.linalg.cholesky(F)
    Finv = sla.cho_solve((Lf, True), np.eye(npar))
    return F, Finv


def mloe_mmom(kernel: str, xobs, yobs, xmiss, ymiss, theta_true, theta_est, metric: int = 0):
    """EXAGEOSTAT_dmle_mloe_mmom_Tile (MLE_dexact.c:1112-1435) for the
    univariate kernels, location by location as the reference loops:
    K_a, K_t = Sigma(lobs, lobs) under theta_est / theta_true, factored
    (:1233-1249); per held-out location k_t, k_a = Sigma(lobs, s0)
    (:1267-1270), k_a <- L_a^-T L_a^-1 k_a, k_t <- L_t^-T L_t^-1 k_t
    (:1283-1311), expr1 = k_t0^T k_a, expr4 = k_a0^T k_a, expr3 = k_t0^T k_t
    (:1313-1333), expr2 = |L_t^T k_a|^2 (:1336-1343), then
    expr1 = s_t - 2 expr1, expr2 += expr1, expr3 = s_t - expr3,
    expr4 = s_a - expr4 with s = theta[0] (:1349-1353, :1200-1203), and
    loe = expr2 / expr3 - 1, mom = expr4 / expr2 - 1, each -1 when its
    denominator is zero (codelet_zmloe_mmom.c:47-59).
    Returns (mloe, mmom, loe, mom, parts) with parts = (expr2, expr3, expr4)
    per location."""
    Ka = matern.cov_matrix(kernel, xobs, yobs, xobs, yobs, theta_est, metric)
    Kt = matern.cov_matrix(kernel, xobs, yobs, xobs, yobs, theta_true, metric)
    La = np.linalg.cholesky(Ka)
    Lt = np.linalg.cholesky(Kt)
    st, sa = float(theta_true[0]), float(theta_est[0])
    nm = len(xmiss)
    loe, mom = np.zeros(nm), np.zeros(nm)
    parts = np.zeros((3, nm))
    for p in range(nm):
        kt0 = matern.cov_matrix(kernel, xobs, yobs, xmiss[p:p + 1], ymiss[p:p + 1], theta_true, metric)[:, 0]
        ka0 = matern.cov_matrix(kernel, xobs, yobs, xmiss[p:p + 1], ymiss[p:p + 1], theta_est, metric)[:, 0]
        ka = sla.solve_triangular(La, ka0, lower=True)
        kt = sla.solve_triangular(Lt, kt0, lower=True)
        ka = sla.solve_triangular(La, ka, lower=True, trans="T")
        kt = sla.solve_triangular(Lt, kt, lower=True, trans="T")
        e1 = float(kt0 @ ka)
        e4 = float(ka0 @ ka)
        e3 = float(kt0 @ kt)
        v = Lt.T @ ka
        e2 = float(v @ v)
        e1 = st - 2.0 * e1
        e2 = e1 + e2
        e3 = st - e3
        e4 = sa - e4
        loe[p] = -1.0 if e3 == 0.0 else e2 / e3 - 1.0
        mom[p] = -1.0 if e2 == 0.0 else e4 / e2 - 1.0
        parts[:, p] = (e2, e3, e4)
    return float(np.mean(loe)), float(np.mean(mom)), loe, mom, parts


def flops_dpotrf(n: float) -> float:
    """FLOPS_DPOTRF (include/flops.h:97-98,254)."""
    n = float(n)
    fmuls = n * (((1.0 / 6.0) * n + 0.5) * n + (1.0 / 3.0))
    fadds = n * (((1.0 / 6.0) * n) * n - (1.0 / 6.0))
    return fmuls + fadds


def flops_dtrsm_left(m: float, nrhs: float) -> float:
    """FLOPS_DTRSM(ChamLeft, m, nrhs) (include/flops.h:73-81,231).  The
    header defines FADDS_TRSM as FMULS_TRMM, so both halves count
    nrhs*m(m+1)/2 and the total is nrhs*m*(m+1)."""
    m = float(m)
    nrhs = float(nrhs)
    return nrhs * m * (m + 1.0)
