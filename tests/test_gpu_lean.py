"""Lean contexts (MLE_data.mp_lean = 1): no FP64 covariance on the device.

The mixed-precision and single-precision likelihoods then generate Sigma
straight into the mixed factor and regenerate its FP64 tiles, a chunk at a
time, for the refinement's products.  The generated values, the per-tile
partial products and their summation order are the stored path's, so a
lean context must give the same log-likelihood, log-determinant and
quadratic form as a full one, to the last bits, on every kernel family,
and the likelihood must still match the oracle.  The FP64 paths (FP64
likelihood, MLE_zvg, gradient, descC) must fail with status -1 and an error
string, never touch an unallocated matrix.
"""
import ctypes

import numpy as np
import pytest

from oracle import datagen, likelihood

pytestmark = pytest.mark.gpu

OPTS = dict(bf16=True, fp32_corr=0.01, ir_steps=3)


@pytest.fixture(scope="module")
def lib():
    import exageostat_amd
    exageostat_amd.load()
    assert exageostat_amd.device_count() >= 1, "no HIP device visible"
    return exageostat_amd


def _pair(lib, x, y, z, ts, kernel, theta, t=None, precision=2, **opts):
    out = []
    for lean in (False, True):
        m = lib.ExactMLE(x, y, ts, kernel=kernel, precision=precision, t=t, lean=lean)
        m.set_mixed(**dict(OPTS, **opts))
        m.set_z(z)
        val = m.loglik(theta)
        out.append((val, m.logdet, m.dotp, m.status, m.data.mp_ir_done))
        m.close()
    print(f"lean pair {kernel} n={len(z)} ts={ts} precision={precision}: full {out[0]} lean {out[1]}", flush=True)
    return out


@pytest.mark.parametrize("n,ts", [(5760, 192), (5000, 128), (2600, 320)])
def test_lean_univariate_equals_full_and_oracle(lib, n, ts):
    """Ragged last tiles included (5,000 = 39.06 x 128, 2,600 = 8.1 x 320).
    At 5,000 / 128 the 820 lower tiles take two chunks of the 768-slot
    pool; the other sizes take one."""
    theta = (1.0, 0.1, 0.5)
    x, y, _ = datagen.generate_xy_loc(n, 0)
    e, _ = datagen.dlarnv(3, [0, 0, 0, 1], n)
    z = likelihood.dzvg("univariate_matern_stationary", x, y, theta, e)
    (lf, ldf, dpf, sf, irf), (ll, ldl, dpl, sl, irl) = _pair(lib, x, y, z, ts, "univariate_matern_stationary",
                                                              theta)
    assert sf == 0 and sl == 0
    assert irl == irf
    assert abs(ll - lf) <= 1e-14 * abs(lf) and abs(ldl - ldf) <= 1e-14 * abs(ldf), (lf, ll, ldf, ldl)
    assert abs(dpl - dpf) <= 1e-13 * abs(dpf), (dpf, dpl)
    ref, _, _ = likelihood.dmle("univariate_matern_stationary", x, y, z, theta)
    assert abs(ll - ref) <= 1e-6 * abs(ref), (ll, ref)


def test_lean_nugget_single_precision(lib):
    """Single precision (precision 1, no refinement) on a lean context: the
    same value as on a full one."""
    n, ts = 3000, 256
    theta = (1.0, 0.1, 0.5, 0.05)
    x, y, _ = datagen.generate_xy_loc(n, 0)
    e, _ = datagen.dlarnv(3, [0, 0, 0, 1], n)
    z = likelihood.dzvg("univariate_matern_nuggets_stationary", x, y, theta, e)
    (lf, _, _, sf, _), (ll, _, _, sl, _) = _pair(lib, x, y, z, ts, "univariate_matern_nuggets_stationary", theta,
                                                 precision=1)
    assert sf == 0 and sl == 0
    assert abs(ll - lf) <= 1e-14 * abs(lf), (lf, ll)


@pytest.mark.parametrize("kernel", ["bivariate_matern_parsimonious", "bivariate_matern_parsimonious2"])
def test_lean_bivariate_equals_full(lib, kernel):
    """Both bivariate row layouts (interleaved, block)."""
    nloc, ts = 1500, 192
    theta = (1.0, 1.0, 0.1, 0.5, 1.0, 0.3)
    x, y, _ = datagen.generate_xy_loc(nloc, 0)
    e, _ = datagen.dlarnv(3, [0, 0, 0, 1], 2 * nloc)
    z = likelihood.dzvg(kernel, x, y, theta, e)
    (lf, ldf, dpf, sf, _), (ll, ldl, dpl, sl, _) = _pair(lib, x, y, z, ts, kernel, theta)
    assert sf == 0 and sl == 0
    assert abs(ll - lf) <= 1e-14 * abs(lf) and abs(ldl - ldf) <= 1e-14 * abs(ldf), (lf, ll)
    assert abs(dpl - dpf) <= 1e-13 * abs(dpf), (dpf, dpl)


def test_lean_spacetime_equals_full(lib):
    """The space-time kernel (third coordinate: time): 400 points at 4 times
    from GenerateXYLoc_ST."""
    st = "univariate_spacetime_matern_stationary"
    theta = np.array([1.0, 0.1, 0.5, 1.0, 0.7, 0.5, 0.2])
    x, y, t, _ = datagen.generate_xy_loc_st(400, 4, 7)
    e, _ = datagen.dlarnv(3, [7, 7, 7, 1], 1600)
    z = likelihood.dzvg(st, x, y, theta, e, t=t)
    (lf, ldf, dpf, sf, _), (ll, ldl, dpl, sl, _) = _pair(lib, x, y, z, 192, st, theta, t=t)
    assert sf == 0 and sl == 0
    assert abs(ll - lf) <= 1e-14 * abs(lf), (lf, ll)
    assert abs(dpl - dpf) <= 1e-13 * abs(dpf), (dpf, dpl)


def test_lean_refuses_fp64_paths(lib):
    """FP64 likelihood, gradient, MLE_zvg and descC on a lean context: status
    -1 with an error string; the context stays usable for the mixed path."""
    from exageostat_amd import _lib
    n, ts = 1600, 320
    theta = (1.0, 0.1, 0.5)
    x, y, _ = datagen.generate_xy_loc(n, 0)
    e, _ = datagen.dlarnv(3, [0, 0, 0, 1], n)
    z = likelihood.dzvg("univariate_matern_stationary", x, y, theta, e)
    m = lib.ExactMLE(x, y, ts, precision=0, lean=True)
    m.set_z(z)
    c = _lib.load()
    th = np.array(theta)
    g = np.zeros(3)
    dptr = ctypes.cast(ctypes.pointer(m.data), ctypes.c_void_p)
    assert c.MLE_alg(3, th.ctypes.data_as(_lib.c_double_p), None, dptr) == -np.inf
    assert b"lean" in c.exageostat_gpu_last_error()
    assert c.MLE_alg(3, th.ctypes.data_as(_lib.c_double_p), g.ctypes.data_as(_lib.c_double_p), dptr) == -np.inf
    assert b"lean" in c.exageostat_gpu_last_error()
    with pytest.raises(RuntimeError, match="lean"):
        m.zvg(theta, e)
    # descC: a dcmg into it fails instead of writing through null tiles
    desc = ctypes.cast(ctypes.c_void_p(m.data.descC), ctypes.POINTER(_lib.Desc))
    st = c.EXAGEOSTAT_MLE_dcmg_Tile_Async(_lib.ChamLower, desc, ctypes.byref(m.data.l1), ctypes.byref(m.data.l1),
                                          ctypes.byref(m.data.lm), th.ctypes.data_as(_lib.c_double_p), m.data.dm,
                                          m.data.kernel_fun, None, None)
    assert st == -1
    assert b"not allocated" in c.exageostat_gpu_last_error()
    m.precision = 2
    m.set_mixed(**OPTS)
    val = m.loglik(theta)
    ref, _, _ = likelihood.dmle("univariate_matern_stationary", x, y, z, theta)
    assert m.status == 0 and abs(val - ref) <= 1e-6 * abs(ref)
    m.close()
