"""GPU MLOE / MMOM (EXAGEOSTAT_dmle_mloe_mmom_Tile) against the oracle.

Small cases compare every per-location value with the oracle's
location-by-location restatement; ragged last tiles on both the observed
and the held-out side; the three GEMM tile shapes (ts = 320 -> 64 x 64
blocks, 192 -> 192 x 192, 128 -> 128 x 128).  At n = 12,000 the closed
forms hold: zero at the true theta, and MLOE = 0, MMOM = c - 1 when only
the variance is rescaled by c.
"""
import ctypes

import numpy as np
import pytest

from oracle import datagen, likelihood

pytestmark = pytest.mark.gpu

K = "univariate_matern_stationary"
KN = "univariate_matern_nuggets_stationary"


@pytest.fixture(scope="module")
def lib():
    import exageostat_amd
    exageostat_amd.load()
    assert exageostat_amd.device_count() >= 1, "no HIP device visible"
    return exageostat_amd


def _split(n, nmiss, seed):
    x, y, _ = datagen.generate_xy_loc(n, seed)
    rng = np.random.default_rng(seed)
    miss = np.sort(rng.choice(n, nmiss, replace=False))
    obs = np.setdiff1d(np.arange(n), miss)
    return x[obs], y[obs], x[miss], y[miss]


@pytest.mark.parametrize("n,nmiss,ts,kernel,tt,te", [
    (1600, 100, 320, K, (1.0, 0.1, 0.5), (0.8, 0.15, 0.7)),
    (1600, 400, 320, K, (1.2, 0.05, 1.2), (1.0, 0.09, 0.6)),
    (1100, 150, 192, K, (1.0, 0.1, 0.5), (1.4, 0.06, 0.45)),
    (1000, 130, 128, KN, (1.0, 0.1, 0.5, 0.05), (0.9, 0.07, 0.8, 0.02)),
])
def test_mloe_mmom_matches_oracle(lib, n, nmiss, ts, kernel, tt, te):
    xo, yo, xm, ym = _split(n, nmiss, 7)
    ref_mloe, ref_mmom, ref_loe, ref_mom, _ = likelihood.mloe_mmom(kernel, xo, yo, xm, ym, tt, te)
    m = lib.ExactMLE(xo, yo, ts, kernel=kernel)
    mloe, mmom, loe, mom = m.mloe_mmom(tt, te, xo, yo, xm, ym)
    status = m.data.status
    m.close()
    assert status == 0
    assert np.max(np.abs(loe - ref_loe)) <= 1e-8 * max(1.0, np.max(np.abs(ref_loe))), (loe[:4], ref_loe[:4])
    assert np.max(np.abs(mom - ref_mom)) <= 1e-8 * max(1.0, np.max(np.abs(ref_mom))), (mom[:4], ref_mom[:4])
    assert abs(mloe - ref_mloe) <= 1e-9 * max(1.0, abs(ref_mloe))
    assert abs(mmom - ref_mmom) <= 1e-9 * max(1.0, abs(ref_mmom))


def test_reference_entry_point(lib):
    """mloe_mmom_init + EXAGEOSTAT_dmle_mloe_mmom_Tile fill data.mloe and
    data.mmom (the reference's call sequence)."""
    xo, yo, xm, ym = _split(900, 60, 2)
    tt, te = (1.0, 0.1, 0.5), (0.7, 0.12, 0.6)
    ref_mloe, ref_mmom, _, _, _ = likelihood.mloe_mmom(K, xo, yo, xm, ym, tt, te)
    m = lib.ExactMLE(xo, yo, 64, kernel=K)
    refs = [np.ascontiguousarray(a, dtype=np.float64) for a in (xo, yo, xm, ym)]
    P = ctypes.POINTER(ctypes.c_double)
    d = m.data
    d.lobs.x, d.lobs.y = refs[0].ctypes.data_as(P), refs[1].ctypes.data_as(P)
    d.lmiss.x, d.lmiss.y = refs[2].ctypes.data_as(P), refs[3].ctypes.data_as(P)
    tta = np.array(tt)
    tea = np.array(te)
    m.lib.mloe_mmom_init(ctypes.byref(d), 60, len(xo), 64, 1, 1)
    m.lib.EXAGEOSTAT_dmle_mloe_mmom_Tile(ctypes.byref(d), tta.ctypes.data_as(P), tea.ctypes.data_as(P), 60,
                                         len(xo), len(xo) + 60)
    got = (d.mloe, d.mmom, d.status)
    m.lib.MLOE_MMOM_Finalize(ctypes.byref(d))
    m.close()
    assert got[2] == 0
    assert abs(got[0] - ref_mloe) <= 1e-9 * max(1.0, abs(ref_mloe))
    assert abs(got[1] - ref_mmom) <= 1e-9 * max(1.0, abs(ref_mmom))


def test_bivariate_is_rejected(lib):
    xo, yo, xm, ym = _split(400, 20, 1)
    m = lib.ExactMLE(xo, yo, 64, kernel="bivariate_matern_parsimonious")
    th = (1.0, 1.0, 0.1, 0.5, 1.0, 0.5)
    with pytest.raises(RuntimeError):
        m.mloe_mmom(th, th, xo, yo, xm, ym)
    m.close()


def test_closed_forms_large(lib):
    """n = 12,000 (ts = 960, 1,000 held out: two row tiles, the last one
    ragged): zero at the true theta; a rescaled variance leaves the
    predictor unchanged (MLOE = 0) and scales the presumed MSE (MMOM = c - 1)."""
    xo, yo, xm, ym = _split(12000, 1000, 5)
    tt = (1.0, 0.1, 0.5)
    m = lib.ExactMLE(xo, yo, 960, kernel=K)
    mloe0, mmom0, loe0, _ = m.mloe_mmom(tt, tt, xo, yo, xm, ym)
    c = 1.7
    mloe1, mmom1, loe1, mom1 = m.mloe_mmom(tt, (c * tt[0], tt[1], tt[2]), xo, yo, xm, ym)
    mloe2, mmom2, loe2, _ = m.mloe_mmom(tt, (0.8, 0.15, 0.7), xo, yo, xm, ym)
    m.close()
    assert abs(mloe0) <= 1e-9 and abs(mmom0) <= 1e-9, (mloe0, mmom0)
    assert np.max(np.abs(loe0)) <= 1e-8
    assert abs(mloe1) <= 1e-9, mloe1
    assert abs(mmom1 - (c - 1.0)) <= 1e-9, mmom1
    assert np.max(np.abs(mom1 - (c - 1.0))) <= 1e-8
    # the optimal predictor has the smallest true MSE
    assert np.min(loe2) >= -1e-9 and mloe2 > 0.0
