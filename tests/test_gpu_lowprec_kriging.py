"""GPU parity of the single- and mixed-precision kriging paths
(EXAGEOSTAT_smle_Predict_Tile, EXAGEOSTAT_sdmle_Predict_Tile) against the
oracle's FP64 predictor.

Both factor C22 with the mixed-precision engine (per-theta tile classes;
the fused three-class update when the tile size is a multiple of 192, one
launch per class otherwise) and keep C12 and the product in FP64.  The
mixed path refines the solve C22 alpha = z_obs in FP64, so its predictions
are held to 1e-10 of the FP64 predictor; the single-precision path solves
with its FP32-tile factor alone, as the reference's FP32 prediction does,
and is held to the accuracy of that solve.
"""
import numpy as np
import pytest

from oracle import datagen, likelihood

pytestmark = pytest.mark.gpu

K = "univariate_matern_stationary"
MIXED = dict(fp32_corr=0.01, bf16=True, ir_steps=8, ir_tol=1e-12)


@pytest.fixture(scope="module")
def lib():
    import exageostat_amd
    exageostat_amd.load()
    assert exageostat_amd.device_count() >= 1, "no HIP device visible"
    return exageostat_amd


def _split(kern, n, nmiss, theta, seed, rows=1):
    """Jittered grid (Morton order) with z drawn from Sigma(theta), nmiss
    held-out locations drawn with a seeded permutation; both sets keep the
    Morton order.  Returns the sets and the oracle's predictions and MSE."""
    x, y, _ = datagen.generate_xy_loc(n, seed)
    e, _ = datagen.dlarnv(3, [seed, seed, seed, 1], rows * n)
    z = likelihood.dzvg(kern, x, y, theta, e).reshape(n, rows)
    perm = np.random.default_rng(seed + 11).permutation(n)
    miss, obs = np.sort(perm[:nmiss]), np.sort(perm[nmiss:])
    zo, za = z[obs].ravel(), z[miss].ravel()
    out = likelihood.predict(kern, x[obs], y[obs], zo, x[miss], y[miss], theta, za)
    return x[obs], y[obs], zo, x[miss], y[miss], za, out[0], out[1]


@pytest.mark.parametrize("n,ts", [(1600, 320), (3840, 192)], ids=["per_class", "fused"])
@pytest.mark.parametrize("theta", [(1.0, 0.1, 0.5), (1.0, 0.03, 1.0)], ids=["c2theta", "c3smooth"])
def test_mixed_kriging_matches_oracle(lib, n, ts, theta):
    """Mixed-precision kriging (FP32 and BF16 tiles in C22's factor, the
    solve refined in FP64): predictions within 1e-10 of the oracle (relative
    to the largest; 1e-14 to 3e-13 measured), the MSE within 1e-10
    relative, status 0."""
    xo, yo, zo, xm, ym, za, ref, mse_ref = _split(K, n, 200, theta, 1)
    m = lib.ExactMLE(xo, yo, ts, precision=2)
    m.set_mixed(**MIXED)
    zm, mse = m.predict(theta, xo, yo, zo, xm, ym, za)
    status, steps = m.status, m.data.mp_ir_done
    zm64, _ = m.predict(theta, xo, yo, zo, xm, ym, za, precision=0)
    m.close()
    scale = np.max(np.abs(ref))
    err = np.max(np.abs(zm - ref)) / scale
    print(f"mixed kriging n={n} ts={ts} theta={theta}: rel err {err:.2e} (FP64 GPU "
          f"{np.max(np.abs(zm64 - ref)) / scale:.2e}), {steps} refinement steps, MSE {mse:.6e} vs {mse_ref:.6e}",
          flush=True)
    assert status == 0
    assert err <= 1e-10, err
    assert abs(mse - mse_ref) <= 1e-10 * mse_ref


def test_single_precision_kriging_is_its_own_solve(lib):
    """Single-precision kriging (FP32 off-diagonal tiles, no refinement):
    within the accuracy of an FP32-factor solve of the oracle's predictions
    (1e-4 of the largest; 2.3e-6 measured), and not the FP64 predictor in
    disguise (it differs from it by more than FP64 rounding)."""
    theta = (1.0, 0.1, 0.5)
    xo, yo, zo, xm, ym, za, ref, mse_ref = _split(K, 1600, 200, theta, 2)
    m = lib.ExactMLE(xo, yo, 320, precision=1)
    zm, mse = m.predict(theta, xo, yo, zo, xm, ym, za)
    m.close()
    err = np.max(np.abs(zm - ref)) / np.max(np.abs(ref))
    print(f"single-precision kriging: rel err {err:.2e}, MSE {mse:.6e} vs {mse_ref:.6e}", flush=True)
    assert 1e-12 < err <= 1e-4, err
    assert abs(mse - mse_ref) <= 1e-4 * mse_ref


def test_mixed_bivariate_kriging_matches_oracle(lib):
    """Parsimonious bivariate kriging at the theta of the 50,000-location
    workload through the mixed path: 600 observed and 100 held-out
    locations, per-row predictions and the per-variable MSE within 1e-8 of
    the oracle."""
    kern = "bivariate_matern_parsimonious"
    theta = np.array([1.0, 1.0, 0.1, 0.5, 1.0, 0.5])
    x, y, _ = datagen.generate_xy_loc(700, 3)
    e, _ = datagen.dlarnv(3, [3, 3, 3, 1], 1400)
    z = likelihood.dzvg(kern, x, y, theta, e).reshape(700, 2)
    perm = np.random.default_rng(5).permutation(700)
    miss, obs = np.sort(perm[:100]), np.sort(perm[100:])
    zo, za = z[obs].ravel(), z[miss].ravel()
    ref, mse, mse1, mse2 = likelihood.predict(kern, x[obs], y[obs], zo, x[miss], y[miss], theta, za)
    m = lib.ExactMLE(x[obs], y[obs], 128, kernel=kern, precision=2)
    m.set_mixed(**MIXED)
    zm, mse_gpu = m.predict(theta, x[obs], y[obs], zo, x[miss], y[miss], za)
    d1, d2, status = m.data.mserror1, m.data.mserror2, m.status
    m.close()
    scale = np.max(np.abs(ref))
    print(f"mixed bivariate kriging: rel err {np.max(np.abs(zm - ref)) / scale:.2e}", flush=True)
    assert status == 0
    assert np.max(np.abs(zm - ref)) <= 1e-8 * scale
    assert abs(mse_gpu - mse) <= 1e-8 * mse
    assert abs(d1 - mse1) <= 1e-8 * mse1 and abs(d2 - mse2) <= 1e-8 * mse2
