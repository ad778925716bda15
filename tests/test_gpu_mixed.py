"""GPU parity of the mixed-precision likelihood (MLE_alg with precision = 2).

The mixed factor keeps FP64 diagonal tiles, stores the other tiles in FP32
(optionally updated with BF16 MFMA operands) and refines the solve in FP64.
The bar is |dl|/|l| <= 1e-6 against the FP64 oracle; the refined quadratic
form z^T Sigma^-1 z is held to a much tighter bound, and with every tile
forced to FP64 the mixed machinery must reproduce the FP64 path.
"""
import os

import numpy as np
import pytest

from oracle import datagen, likelihood, matern

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
K = "univariate_matern_stationary"
MIXED_TOL = 1e-6


@pytest.fixture(scope="module")
def lib():
    import exageostat_amd
    exageostat_amd.load()
    assert exageostat_amd.device_count() >= 1, "no HIP device visible"
    return exageostat_amd


@pytest.fixture(scope="module")
def field():
    """n = 4096 Morton-ordered locations and z drawn from Sigma(1, 0.1, 0.5)."""
    n = 4096
    x, y, _ = datagen.generate_xy_loc(n, 11)
    e, _ = datagen.dlarnv(3, [0, 0, 0, 1], n)
    z = likelihood.dzvg(K, x, y, (1.0, 0.1, 0.5), e)
    return x, y, z


_ORACLE = {}


def _oracle(field, theta):
    """Oracle (l, logdet, dotp) on the shared field, cached per theta."""
    key = tuple(float(t) for t in theta)
    if key not in _ORACLE:
        x, y, z = field
        _ORACLE[key] = likelihood.dmle(K, x, y, z, key)
    return _ORACLE[key]


def _mixed(lib, x, y, z, ts, kernel=K, **opts):
    m = lib.ExactMLE(x, y, ts, kernel=kernel, precision=2)
    m.set_mixed(**opts)
    m.set_z(z)
    return m


OPTS = [
    dict(),                                            # every off-diagonal tile FP32
    dict(bf16=True, fp32_corr=0.01),                   # weak tiles with BF16 updates
    dict(bf16=True, fp32_corr=0.05, diag_thick=2),     # wider FP64 band
    dict(fp64_corr=0.3, bf16=True, fp32_corr=0.01),    # strongly correlated tiles FP64
    dict(bf16=True, fp32_corr=0.01, lolo=True),        # four split-BF16 products
]


@pytest.mark.parametrize("opts", OPTS, ids=["f32", "bf16", "band2", "f64corr", "lolo"])
@pytest.mark.parametrize("theta", [(1.0, 0.1, 0.5), (1.0, 0.03, 0.5), (1.5, 0.05, 1.0), (0.8, 0.07, 1.37)])
def test_mixed_loglik_matches_oracle(lib, field, theta, opts):
    x, y, z = field
    ll, ld, dp = _oracle(field, theta)
    m = _mixed(lib, x, y, z, 256, **opts)
    val = m.loglik(theta)
    assert m.status == 0
    assert abs(val - ll) <= MIXED_TOL * abs(ll), (theta, opts, val, ll)
    # Three FP64 refinement steps bring z^T Sigma^-1 z to 1e-13..1e-10; the
    # slowest case is the smooth theta (nu = 1.37, range 0.07) with the
    # BF16-heavy options (fp32_corr = 0.05), which reaches ~4e-10.
    assert abs(m.dotp - dp) <= 1e-9 * abs(dp), (m.dotp, dp)
    m.close()


def test_mixed_all_fp64_equals_fp64_path(lib, field):
    """fp64_corr below every tile's correlation bound: every tile FP64, so
    the class-dispatched factorisation must agree with the oracle to the
    FP64 tolerance (1e-9)."""
    x, y, z = field
    theta = (1.0, 0.1, 0.5)
    ll, ld, dp = _oracle(field, theta)
    m = _mixed(lib, x, y, z, 256, fp64_corr=1e-300)
    val = m.loglik(theta)
    pm = m.prec_map()
    assert np.all(pm[np.tril_indices(pm.shape[0])] == 0)
    assert abs(val - ll) <= 1e-9 * abs(ll)
    assert abs(m.logdet - ld) <= 1e-9 * abs(ld)
    m.close()


def test_precision_map_follows_correlation(lib, field):
    """Tile classes against an independent bound: the largest correlation of
    a tile pair is the Matern correlation at the smallest distance between
    the two tiles' bounding boxes."""
    x, y, z = field
    ts, theta, thr = 256, (1.0, 0.1, 0.5), 0.01
    m = _mixed(lib, x, y, z, ts, bf16=True, fp32_corr=thr, fp64_corr=0.5)
    m.loglik(theta)
    pm = m.prec_map()
    mt = pm.shape[0]
    boxes = [(x[t * ts:(t + 1) * ts].min(), x[t * ts:(t + 1) * ts].max(), y[t * ts:(t + 1) * ts].min(),
              y[t * ts:(t + 1) * ts].max()) for t in range(mt)]
    counts = {0: 0, 1: 0, 3: 0}
    for n in range(mt):
        for mm in range(n, mt):
            a, b = boxes[mm], boxes[n]
            dx = max(0.0, a[0] - b[1], b[0] - a[1])
            dy = max(0.0, a[2] - b[3], b[2] - a[3])
            c = float(matern.matern_correlation(np.array([np.hypot(dx, dy) / theta[1]]), theta[2])[0])
            want = 0 if (mm == n or c >= 0.5) else (3 if c < thr else 1)
            assert pm[mm, n] == want, (mm, n, c)
            counts[want] += 1
    # range 0.1 on 256-point tiles: all three classes occur
    assert counts[3] > 0 and counts[1] > 0
    m.close()


def test_mixed_refinement_converges(lib, field):
    """Without refinement z^T Sigma~^-1 z carries the FP32 error; each FP64
    step shrinks it, and three steps reach 1e-10."""
    x, y, z = field
    theta = (1.0, 0.1, 0.5)
    _, _, dp = _oracle(field, theta)
    errs = []
    for steps in (0, 1, 3):
        m = _mixed(lib, x, y, z, 256, bf16=True, fp32_corr=0.01, ir_steps=steps)
        m.loglik(theta)
        errs.append(abs(m.dotp - dp) / abs(dp))
        m.close()
    assert errs[0] > errs[1] > 0 or errs[1] <= 1e-12
    assert errs[2] <= 1e-10, errs


def test_mixed_non_spd_status(lib, field):
    x, y, z = field
    m = _mixed(lib, x, y, z, 256)
    val = m.loglik([-1.0, 0.1, 0.5])
    assert val == -np.inf
    assert m.status > 0
    ll, _, _ = _oracle(field, (1.0, 0.1, 0.5))
    val = m.loglik((1.0, 0.1, 0.5))
    assert abs(val - ll) <= MIXED_TOL * abs(ll)
    m.close()


def test_mixed_golden_c1(lib):
    """C1 golden thetas (n=1600, ts=320) within 1e-6; the cond-2.2e10 theta
    is excluded (FP32 storage of Sigma alone moves l there by ~1e-3)."""
    c1 = np.load(os.path.join(GOLD, "c1_seed0.npz"))
    m = _mixed(lib, c1["x"], c1["y"], c1["z"], 320, bf16=True, fp32_corr=0.01)
    for th, ll in zip(c1["thetas"][:4], c1["loglik"][:4]):
        val = m.loglik(th)
        assert abs(val - ll) <= MIXED_TOL * abs(ll), (th, val, ll)
    m.close()


def test_mixed_nugget(lib):
    g = np.load(os.path.join(GOLD, "nugget_900.npz"))
    m = _mixed(lib, g["x"], g["y"], g["z"], 64, kernel="univariate_matern_nuggets_stationary", bf16=True,
               fp32_corr=0.01)
    val = m.loglik(g["theta"])
    assert abs(val - g["loglik"][0]) <= MIXED_TOL * abs(g["loglik"][0])
    m.close()


def test_mixed_bivariate(lib):
    g = np.load(os.path.join(GOLD, "bivariate_400.npz"))
    m = _mixed(lib, g["x"], g["y"], g["z"], 64, kernel="bivariate_matern_parsimonious", bf16=True, fp32_corr=0.01)
    val = m.loglik(g["theta"])
    assert abs(val - g["loglik"][0]) <= MIXED_TOL * abs(g["loglik"][0])
    m.close()


def test_mixed_matches_fp64_gpu_large(lib):
    """n = 30720 (ts = 960, 32 tile rows): the mixed path against the FP64
    GPU path (itself pinned to the oracle at 1e-9) on the benchmark's theta."""
    n, ts, theta = 30720, 960, (1.0, 0.1, 0.5)
    x, y, _ = datagen.generate_xy_loc(n, 0)
    e, _ = datagen.dlarnv(3, [0, 0, 0, 1], n)
    ref = lib.ExactMLE(x, y, ts)
    ref.zvg(theta, e)
    z = ref.get_z()
    l64 = ref.loglik(theta)
    ref.close()
    m = _mixed(lib, x, y, z, ts, bf16=True, fp32_corr=0.01)
    val = m.loglik(theta)
    pm = m.prec_map()
    assert np.any(pm == 3)
    assert abs(val - l64) <= MIXED_TOL * abs(l64), (val, l64)
    m.close()


def test_mixed_lolo_reduces_logdet_bias(lib):
    """n = 30720 (ts = 960), exponential kernel with range 0.3 (every
    off-diagonal tile FP32): dropping the lo*lo term of the split-BF16
    products biases log|Sigma| (neighbouring rows have correlated BF16
    remainders); keeping it must shrink the log-det error against the FP64
    GPU path, and both stay within the mixed tolerance."""
    n, ts, theta = 30720, 960, (1.0, 0.3, 0.5)
    x, y, _ = datagen.generate_xy_loc(n, 0)
    e, _ = datagen.dlarnv(3, [0, 0, 0, 1], n)
    ref = lib.ExactMLE(x, y, ts)
    ref.zvg(theta, e)
    z = ref.get_z()
    l64 = ref.loglik(theta)
    ld64 = ref.logdet
    ref.close()
    errs = {}
    for lolo in (False, True):
        m = _mixed(lib, x, y, z, ts, bf16=True, fp32_corr=0.01, lolo=lolo)
        val = m.loglik(theta)
        assert abs(val - l64) <= MIXED_TOL * abs(l64), (lolo, val, l64)
        errs[lolo] = abs(m.logdet - ld64)
        m.close()
    assert errs[True] < errs[False], errs
