"""GPU parity of the likelihood variants reachable through MLE_alg against
the oracle: the block layout of the parsimonious bivariate kernel
("bivariate_matern_parsimonious2"), the profile likelihood of the
"_profile" kernels, the space-time Matern kernel, the band-limited
factorisation of computation "diag_approx" and the single-precision
path (precision 1).

Tolerances: FP64 paths 1e-9 relative on l (and on the profile variances);
generation 1e-13 * sigma^2; the single-precision path 1e-5 relative on l
(FP32 off-diagonal tiles, no refinement); the mixed path 1e-6."""
import numpy as np
import pytest

from oracle import datagen, likelihood, matern

pytestmark = pytest.mark.gpu

UNI = "univariate_matern_stationary"
BIV = "bivariate_matern_parsimonious"
BIV2 = "bivariate_matern_parsimonious2"
ST = "univariate_spacetime_matern_stationary"
TH_BIV = np.array([1.0, 1.0, 0.1, 0.5, 1.0, 0.5])
TH_ST = np.array([1.0, 0.1, 0.5, 1.0, 0.7, 0.5, 0.2])


@pytest.fixture(scope="module")
def lib():
    import exageostat_amd
    exageostat_amd.load()
    assert exageostat_amd.device_count() >= 1, "no HIP device visible"
    return exageostat_amd


def _biv_field(nloc, seed, kernel):
    x, y, _ = datagen.generate_xy_loc(nloc, seed)
    e, _ = datagen.dlarnv(3, [seed, seed, seed, 1], 2 * nloc)
    plain = BIV2 if kernel.startswith(BIV2) else BIV
    return x, y, likelihood.dzvg(plain, x, y, TH_BIV, e)


def _st_field(n, seed, slots=4):
    """n locations from GenerateXYLoc_ST: n / slots spatial points observed
    at `slots` times (t = 1..slots)."""
    x, y, t, _ = datagen.generate_xy_loc_st(n // slots, slots, seed)
    e, _ = datagen.dlarnv(3, [seed, seed, seed, 1], n)
    return x, y, t, likelihood.dzvg(ST, x, y, TH_ST, e, t=t)


# ------------------------------------------------------ block bivariate ---
@pytest.mark.parametrize("nloc,ts", [(800, 320), (700, 192)])
def test_block_bivariate_loglik_matches_oracle(lib, nloc, ts):
    """Block layout (variable 1 rows, then variable 2 rows); 700 locations
    leave a ragged last tile and a variable boundary inside a tile."""
    x, y, z = _biv_field(nloc, 1, BIV2)
    ref, ld, dp = likelihood.dmle(BIV2, x, y, z, TH_BIV)
    m = lib.ExactMLE(x, y, ts, kernel=BIV2)
    m.set_z(z)
    val = m.loglik(TH_BIV)
    assert m.status == 0
    assert abs(val - ref) <= 1e-9 * abs(ref), (val, ref)
    assert abs(m.logdet - ld) <= 1e-9 * abs(ref)
    m.close()


def test_block_bivariate_generation_matches_oracle(lib):
    x, y, _ = datagen.generate_xy_loc(150, 2)
    x2, y2, _ = datagen.generate_xy_loc(70, 3)
    M = lib.DeviceMatrix(140, 300, 64, lower=False)
    M.dcmg(TH_BIV, x2, y2, x, y, kernel=BIV2)
    S = matern.core_dcmg_bivariate_parsimonious2(x2, y2, x, y, TH_BIV)
    assert np.max(np.abs(M.to_numpy() - S)) <= 1e-13
    M.close()


def test_block_bivariate_kriging_matches_oracle(lib):
    """Kriging in the block layout: per-row predictions and per-variable MSE
    (the first and second halves of the held-out rows)."""
    nloc = 600
    x, y, z = _biv_field(nloc, 4, BIV2)
    perm = np.random.default_rng(4).permutation(nloc)
    miss, obs = np.sort(perm[:80]), np.sort(perm[80:])
    zo = np.concatenate([z[:nloc][obs], z[nloc:][obs]])
    za = np.concatenate([z[:nloc][miss], z[nloc:][miss]])
    ref, mse, mse1, mse2 = likelihood.predict(BIV2, x[obs], y[obs], zo, x[miss], y[miss], TH_BIV, za)
    m = lib.ExactMLE(x[obs], y[obs], 128, kernel=BIV2)
    zm, mse_gpu = m.predict(TH_BIV, x[obs], y[obs], zo, x[miss], y[miss], za)
    d1, d2 = m.data.mserror1, m.data.mserror2
    m.close()
    assert np.max(np.abs(zm - ref)) <= 1e-9 * np.max(np.abs(ref))
    assert abs(mse_gpu - mse) <= 1e-9 * mse
    assert abs(d1 - mse1) <= 1e-9 * mse1 and abs(d2 - mse2) <= 1e-9 * mse2


# -------------------------------------------------------------- profile ---
@pytest.mark.parametrize("kernel", [BIV + "_profile", BIV2 + "_profile"])
def test_profile_likelihood_matches_oracle(lib, kernel):
    x, y, z = _biv_field(640, 5, kernel)
    ref, v1, v2, ld, dp = likelihood.dmle_profile(kernel, x, y, z, TH_BIV)
    m = lib.ExactMLE(x, y, 256, kernel=kernel)
    m.set_z(z)
    val = m.loglik(TH_BIV)
    g1, g2 = m.variances
    m.close()
    assert abs(val - ref) <= 1e-9 * abs(ref), (val, ref)
    assert abs(g1 - v1) <= 1e-9 * v1 and abs(g2 - v2) <= 1e-9 * v2, (g1, v1, g2, v2)


# ----------------------------------------------------------- space-time ---
def test_spacetime_generation_matches_oracle(lib):
    x, y, t, _ = _st_field(400, 6)
    th = np.array([1.3, 0.15, 1.2, 2.0, 0.6, 0.8, 0.3])
    M = lib.DeviceMatrix(400, 400, 64, lower=False)
    M.dcmg(th, x, y, kernel=ST, t1=t)
    S = matern.core_dcmg_spacetime_matern(x, y, t, x, y, t, th)
    assert np.max(np.abs(M.to_numpy() - S)) <= 1e-13 * th[0]
    M.close()


@pytest.mark.parametrize("n,ts", [(1600, 320), (1200, 192)])
def test_spacetime_loglik_matches_oracle(lib, n, ts):
    x, y, t, z = _st_field(n, 7)
    ref, ld, dp = likelihood.dmle(ST, x, y, z, TH_ST, t=t)
    m = lib.ExactMLE(x, y, ts, kernel=ST, t=t)
    m.set_z(z)
    val = m.loglik(TH_ST)
    assert m.status == 0
    assert abs(val - ref) <= 1e-9 * abs(ref), (val, ref)
    m.close()


def test_spacetime_kriging_matches_oracle(lib):
    x, y, t, z = _st_field(1200, 8)
    perm = np.random.default_rng(8).permutation(1200)
    miss, obs = np.sort(perm[:100]), np.sort(perm[100:])
    ref, mse, _, _ = likelihood.predict(ST, x[obs], y[obs], z[obs], x[miss], y[miss], TH_ST, z[miss],
                                        tobs=t[obs], tmiss=t[miss])
    m = lib.ExactMLE(x[obs], y[obs], 192, kernel=ST, t=t[obs])
    zm, mse_gpu = m.predict(TH_ST, x[obs], y[obs], z[obs], x[miss], y[miss], z[miss], tobs=t[obs], tmiss=t[miss])
    m.close()
    assert np.max(np.abs(zm - ref)) <= 1e-9 * np.max(np.abs(ref))
    assert abs(mse_gpu - mse) <= 1e-9 * mse


# -------------------------------------------------- band-limited (DST) ---
@pytest.mark.parametrize("n,ts,band", [(1600, 128, 1), (1600, 128, 3), (2000, 192, 4), (1600, 320, 2),
                                       (1600, 320, 9)])
def test_band_loglik_matches_oracle(lib, n, ts, band):
    """computation "diag_approx": only tiles with |m - n| < band are
    factored, the solve reads the whole lower triangle (the outer tiles
    still hold the covariance).  band >= mt is the exact factorisation."""
    th = (1.0, 0.03, 0.5)
    x, y, _ = datagen.generate_xy_loc(n, 9)
    e, _ = datagen.dlarnv(3, [9, 9, 9, 1], n)
    z = likelihood.dzvg(UNI, x, y, th, e)
    ref, ld, dp = likelihood.dmle_band(UNI, x, y, z, th, ts, band)
    m = lib.ExactMLE(x, y, ts, computation="diag_approx", diag_thick=band)
    m.set_z(z)
    val = m.loglik(th)
    assert m.status == 0
    assert abs(val - ref) <= 1e-9 * abs(ref), (val, ref)
    assert abs(m.logdet - ld) <= 1e-9 * abs(ref) and abs(m.dotp - dp) <= 1e-9 * dp
    m.close()


def test_band_profile_uses_the_band_formula(lib):
    """The band path's profile likelihood of the interleaved kernel has
    (N/2.0) log(N/2.0) where the exact path has (N/2) log N."""
    kern = BIV + "_profile"
    x, y, z = _biv_field(640, 10, kern)
    # a short range keeps the band-limited factorisation positive definite
    th = np.array([1.0, 1.0, 0.03, 0.5, 1.0, 0.5])
    ref, v1, v2, _, _ = likelihood.dmle_band(kern, x, y, z, th, 128, 2)
    m = lib.ExactMLE(x, y, 128, kernel=kern, computation="diag_approx", diag_thick=2)
    m.set_z(z)
    val = m.loglik(th)
    g1, g2 = m.variances
    m.close()
    assert abs(val - ref) <= 1e-9 * abs(ref), (val, ref)
    assert abs(g1 - v1) <= 1e-9 * v1 and abs(g2 - v2) <= 1e-9 * v2


# ----------------------------------------------------- single precision ---
@pytest.mark.parametrize("n,ts,theta", [(1600, 320, (1.0, 0.1, 0.5)), (5760, 192, (1.0, 0.1, 0.5)),
                                        (4096, 256, (1.0, 0.03, 1.0))])
def test_single_precision_loglik_within_stated_bound(lib, n, ts, theta):
    x, y, _ = datagen.generate_xy_loc(n, 0)
    e, _ = datagen.dlarnv(3, [0, 0, 0, 1], n)
    z = likelihood.dzvg(UNI, x, y, theta, e)
    ref, _, _ = likelihood.dmle(UNI, x, y, z, theta)
    m = lib.ExactMLE(x, y, ts, precision=1)
    m.set_z(z)
    val = m.loglik(theta)
    pm = m.prec_map()
    m.close()
    assert abs(val - ref) <= 1e-5 * abs(ref), (val, ref, abs(val - ref) / abs(ref))
    off = pm[np.tril_indices(pm.shape[0], -1)]
    assert np.all(off == 1) and np.all(np.diag(pm) == 0)


# ------------------------------------------------ mixed on the variants ---
def test_mixed_block_bivariate_and_spacetime(lib):
    """The mixed path with the block bivariate layout (bounding boxes per
    row -> location map of that layout) and the space-time kernel (no box
    bound: every off-diagonal tile FP32): l within 1e-6 of the oracle."""
    x, y, z = _biv_field(1920, 11, BIV2)
    ref, _, _ = likelihood.dmle(BIV2, x, y, z, TH_BIV)
    m = lib.ExactMLE(x, y, 192, kernel=BIV2, precision=2)
    m.set_mixed(bf16=True, fp32_corr=0.01, ir_steps=3)
    m.set_z(z)
    val = m.loglik(TH_BIV)
    m.close()
    assert abs(val - ref) <= 1e-6 * abs(ref), (val, ref)
    x, y, t, z = _st_field(1920, 12)
    ref, _, _ = likelihood.dmle(ST, x, y, z, TH_ST, t=t)
    m = lib.ExactMLE(x, y, 192, kernel=ST, t=t, precision=2)
    m.set_mixed(bf16=True, fp32_corr=0.01, ir_steps=3)
    m.set_z(z)
    val = m.loglik(TH_ST)
    pm = m.prec_map()
    m.close()
    assert abs(val - ref) <= 1e-6 * abs(ref), (val, ref)
    assert np.all(pm[np.tril_indices(pm.shape[0], -1)] == 1)
