"""Pins of the oracle's kernel and likelihood variants: the space-time
Matern kernel, the block layout of the parsimonious bivariate kernel, the
profile likelihood and the band-limited (diag_approx) factorisation.

Each check takes a route that does not restate the function it pins:
closed forms and reductions to the plain Matern kernel, a permutation of
the interleaved bivariate matrix, the maximum over a variance scale of the
plain likelihood, marginal quadratic forms, and exact Cholesky factors of
block-banded matrices (no fill-in outside the band)."""
import math

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import datagen, likelihood, matern

ST = "univariate_spacetime_matern_stationary"
BIV = "bivariate_matern_parsimonious"
BIV2 = "bivariate_matern_parsimonious2"


def _st_points(n, seed):
    x, y, _ = datagen.generate_xy_loc(n, seed)
    t = np.random.default_rng(seed).integers(0, 5, n).astype(np.float64)
    return x, y, t


def test_spacetime_reduces_to_matern_at_equal_times():
    """u = 0: psi = 1, so the kernel is core_dcmg(sigma^2, range, nu)."""
    x, y, _ = datagen.generate_xy_loc(150, 4)
    t = np.full(150, 2.0)
    th = (1.4, 0.12, 0.8, 0.7, 0.6, 0.4, 0.3)
    A = matern.core_dcmg_spacetime_matern(x, y, t, x, y, t, th)
    B = matern.core_dcmg(x, y, x, y, th[:3])
    assert np.max(np.abs(A - B)) <= 1e-15 * th[0]


def test_spacetime_zero_distance_and_exponential_closed_form():
    """d = 0: sigma^2 / psi^(beta + delta); nu = 1/2: sigma^2 exp(-x) /
    psi^(beta + delta) with x = d / (range psi^(beta/2))."""
    x, y, t = _st_points(120, 5)
    s2, rng, a, alpha, beta, delta = 1.3, 0.2, 0.5, 0.8, 0.7, 0.25
    th = (s2, rng, 0.5, a, alpha, beta, delta)
    A = matern.core_dcmg_spacetime_matern(x, y, t, x, y, t, th)
    u = np.abs(t[:, None] - t[None, :])
    psi = u ** (2 * alpha) / a + 1.0
    d = np.hypot(x[:, None] - x[None, :], y[:, None] - y[None, :])
    expected = s2 * np.exp(-(d / rng) / psi ** (beta / 2)) / psi ** (beta + delta)
    assert np.max(np.abs(A - expected)) <= 2e-15 * s2
    # same place, different times
    xs = np.zeros(6)
    ts = np.arange(6.0)
    A0 = matern.core_dcmg_spacetime_matern(xs, xs, ts, xs, xs, ts, (s2, rng, 1.3, a, alpha, beta, delta))
    psi0 = np.abs(ts[:, None] - ts[None, :]) ** (2 * alpha) / a + 1.0
    assert np.allclose(A0, s2 / psi0 ** (beta + delta), rtol=1e-15, atol=0)


def test_spacetime_matrix_is_symmetric_positive_definite():
    """Gneiting's class is a valid covariance for alpha in (0, 1],
    beta in [0, 1] and a non-negative remaining exponent."""
    x, y, t = _st_points(300, 6)
    th = (1.0, 0.1, 0.9, 1.5, 0.9, 0.8, 0.1)
    S = likelihood.sigma(ST, x, y, th, t=t)
    assert np.array_equal(S, S.T)
    assert np.min(np.linalg.eigvalsh(S)) > 0


def test_block_bivariate_is_permuted_interleaved():
    """Rows/columns of the interleaved matrix taken as (variable 1 rows,
    variable 2 rows) give the block layout exactly."""
    x, y, _ = datagen.generate_xy_loc(90, 7)
    th = (1.2, 0.8, 0.09, 0.6, 1.4, 0.45)
    Ai = matern.core_dcmg_bivariate_parsimonious(x, y, x, y, th)
    Ab = matern.core_dcmg_bivariate_parsimonious2(x, y, x, y, th)
    perm = np.concatenate([np.arange(0, 180, 2), np.arange(1, 180, 2)])
    assert np.max(np.abs(Ai[np.ix_(perm, perm)] - Ab)) <= 1e-15 * max(th[0], th[1])
    # rectangular (prediction) shape
    x2, y2, _ = datagen.generate_xy_loc(30, 8)
    Ai = matern.core_dcmg_bivariate_parsimonious(x2, y2, x, y, th)
    Ab = matern.core_dcmg_bivariate_parsimonious2(x2, y2, x, y, th)
    pr = np.concatenate([np.arange(0, 60, 2), np.arange(1, 60, 2)])
    assert np.max(np.abs(Ai[np.ix_(pr, perm)] - Ab)) <= 1e-15 * max(th[0], th[1])


@pytest.mark.parametrize("kernel", [BIV + "_profile", BIV2 + "_profile"])
def test_profile_likelihood_is_the_maximum_over_a_variance_scale(kernel):
    """For Sigma(s) = s Sigma(theta) (both variances scaled by s, the cross
    term follows), max_s l(s) is reached at s* = z^T Sigma^-1 z / N and
    equals the profile formula (N even)."""
    x, y, _ = datagen.generate_xy_loc(80, 9)
    th = np.array([1.1, 0.9, 0.12, 0.5, 1.0, 0.4])
    e, _ = datagen.dlarnv(3, [9, 9, 9, 1], 160)
    z = likelihood.dzvg(BIV2 if "2" in kernel else BIV, x, y, th * np.array([1.7, 1.7, 1, 1, 1, 1]), e)
    l, _, _, logdet, dotp = likelihood.dmle_profile(kernel, x, y, z, th)
    plain = BIV2 if "2" in kernel else BIV
    s = dotp / 160
    lmax, _, _ = likelihood.dmle(plain, x, y, z, th * np.array([s, s, 1, 1, 1, 1]))
    assert l == pytest.approx(lmax, rel=1e-12, abs=1e-9)
    for f in (0.97, 1.03):
        lo, _, _ = likelihood.dmle(plain, x, y, z, th * np.array([s * f, s * f, 1, 1, 1, 1]))
        assert lo < lmax


def test_profile_variances_are_marginal_quadratic_forms():
    """Block layout: the first half of L^-1 z is L11^-1 z1, so variance1 =
    z1^T Sigma11^-1 z1 / (N/2) for any rho.  Interleaved layout with rho = 0:
    Sigma decouples by row parity, so variance_a = z_a^T Sigma_a^-1 z_a / (N/2)."""
    x, y, _ = datagen.generate_xy_loc(70, 10)
    e, _ = datagen.dlarnv(3, [10, 10, 10, 1], 140)
    th = np.array([1.3, 0.7, 0.1, 0.8, 1.2, 0.6])
    z = likelihood.dzvg(BIV2, x, y, th, e)
    _, v1, v2, _, dotp = likelihood.dmle_profile(BIV2 + "_profile", x, y, z, th)
    S11 = matern.core_dcmg(x, y, x, y, (th[0], th[2], th[3]))
    assert v1 == pytest.approx(z[:70] @ np.linalg.solve(S11, z[:70]) / 70, rel=1e-11)
    assert v1 + v2 == pytest.approx(dotp / 70, rel=1e-13)
    th0 = th.copy()
    th0[5] = 0.0
    z = likelihood.dzvg(BIV, x, y, th0, e)
    _, v1, v2, _, _ = likelihood.dmle_profile(BIV + "_profile", x, y, z, th0)
    S22 = matern.core_dcmg(x, y, x, y, (th[1], th[2], th[4]))
    assert v1 == pytest.approx(z[0::2] @ np.linalg.solve(S11, z[0::2]) / 70, rel=1e-11)
    assert v2 == pytest.approx(z[1::2] @ np.linalg.solve(S22, z[1::2]) / 70, rel=1e-11)


def _block_banded_factor(n, ts, width, seed):
    """Lower factor B with nonzero tiles only for 0 <= m - n < width
    (positive diagonal): S = B B^T has Cholesky factor B."""
    rng = np.random.default_rng(seed)
    mt = (n + ts - 1) // ts
    B = np.zeros((n, n))
    for k in range(mt):
        r = slice(k * ts, min(n, (k + 1) * ts))
        for m in range(k, min(mt, k + width)):
            rr = slice(m * ts, min(n, (m + 1) * ts))
            blk = rng.standard_normal((rr.stop - rr.start, r.stop - r.start)) * 0.3
            if m == k:
                blk = np.tril(blk) + np.diag(1.0 + rng.random(r.stop - r.start))
            B[rr, r] = blk
    return B


@pytest.mark.parametrize("n,ts,width,band", [(70, 16, 2, 2), (70, 16, 2, 3), (64, 8, 3, 3), (50, 7, 1, 1)])
def test_band_cholesky_is_exact_on_block_banded_matrices(n, ts, width, band):
    """No fill-in outside the band: when S's factor has tiles only within
    |m - n| < width <= band, the band-limited factorisation returns it
    exactly (up to rounding), and the untouched outer tiles are zero."""
    B = _block_banded_factor(n, ts, width, n + ts + band)
    L = likelihood.band_cholesky_tiles(B @ B.T, ts, band)
    assert np.max(np.abs(L - B)) <= 1e-12 * np.max(np.abs(B))


def test_band_likelihood_full_band_is_exact_and_band_one_keeps_covariance():
    x, y, _ = datagen.generate_xy_loc(200, 11)
    th = (1.0, 0.1, 0.5)
    e, _ = datagen.dlarnv(3, [11, 11, 11, 1], 200)
    z = likelihood.dzvg("univariate_matern_stationary", x, y, th, e)
    ref = likelihood.dmle("univariate_matern_stationary", x, y, z, th)
    full = likelihood.dmle_band("univariate_matern_stationary", x, y, z, th, 48, 5)
    assert np.allclose(full, ref, rtol=1e-12)
    # band 1: diagonal tiles factored, every off-diagonal tile the covariance
    S = likelihood.sigma("univariate_matern_stationary", x, y, th)
    l1, logdet1, dotp1 = likelihood.dmle_band("univariate_matern_stationary", x, y, z, th, 48, 1)
    blocks = [slice(i, min(200, i + 48)) for i in range(0, 200, 48)]
    assert logdet1 == pytest.approx(sum(np.linalg.slogdet(S[b, b])[1] for b in blocks), rel=1e-12)
    H = np.tril(S)
    for b in blocks:
        H[b, b] = np.linalg.cholesky(S[b, b])
    v = sla.solve_triangular(H, z, lower=True)
    assert dotp1 == pytest.approx(v @ v, rel=1e-12)
    assert l1 == pytest.approx(-0.5 * dotp1 - 0.5 * logdet1 - 100 * math.log(2 * math.pi), rel=1e-14)
