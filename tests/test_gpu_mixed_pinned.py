"""GPU parity of the mixed-precision tile classes that the n = 90,000 and
n = 160,000 workloads run, against the FP64 oracle.

The fused three-class update (one launch per step) is taken when the tile
size is a multiple of 192; other tile sizes take one launch per class with
128 x 128 or 64 x 64 blocks.  These tests pin both against the oracle's
dense FP64 likelihood: ts = 192 and 384 (fused launch), ts = 1,024 (the
tile size of the n = 160,000 workload, per-class launches), at the
benchmark's theta and at the two n = 160,000 thetas, with every tile class
present.  They also pin kriging with the parsimonious bivariate kernel
against the oracle's predictor, the mixed path's gradient against the
oracle's, and the single-precision path's refusal to return one.
"""
import numpy as np
import pytest

from oracle import datagen, likelihood

pytestmark = pytest.mark.gpu

K = "univariate_matern_stationary"
MIXED_TOL = 1e-6
GRAD_TOL = 1e-4
THETAS = [(1.0, 0.1, 0.5), (1.0, 0.03, 1.0), (1.0, 0.3, 0.5)]
# the benchmark's option set: FP32 tiles, BF16 updates below correlation 0.01
OPTS = dict(bf16=True, fp32_corr=0.01, ir_steps=3)


@pytest.fixture(scope="module")
def lib():
    import exageostat_amd
    exageostat_amd.load()
    assert exageostat_amd.device_count() >= 1, "no HIP device visible"
    return exageostat_amd


_FIELDS = {}


def _field(n, theta):
    """Morton-ordered jittered grid (seed 0) with z drawn from Sigma(theta);
    the oracle's (l, logdet, dotp) at theta."""
    key = (n, tuple(theta))
    if key not in _FIELDS:
        x, y, _ = datagen.generate_xy_loc(n, 0)
        e, _ = datagen.dlarnv(3, [0, 0, 0, 1], n)
        z = likelihood.dzvg(K, x, y, theta, e)
        _FIELDS[key] = (x, y, z, likelihood.dmle(K, x, y, z, theta))
    return _FIELDS[key]


def _check(lib, n, ts, theta, **extra):
    x, y, z, (ll, ld, dp) = _field(n, theta)
    m = lib.ExactMLE(x, y, ts, precision=2)
    m.set_mixed(**dict(OPTS, **extra))
    m.set_z(z)
    val = m.loglik(theta)
    pm = m.prec_map()
    status, logdet, dotp = m.status, m.logdet, m.dotp
    m.close()
    print(f"mixed n={n} ts={ts} theta={theta} {extra}: rel dl {abs(val - ll) / abs(ll):.2e} "
          f"dlogdet {logdet - ld:+.2e} rel ddotp {abs(dotp - dp) / abs(dp):.2e}", flush=True)
    assert status == 0
    assert abs(val - ll) <= MIXED_TOL * abs(ll), (n, ts, theta, val, ll)
    # three FP64 refinement steps: the quadratic form to FP64-level accuracy
    assert abs(dotp - dp) <= 1e-8 * abs(dp), (dotp, dp)
    assert abs(logdet - ld) <= 2e-6 * abs(ll), (logdet, ld)
    return pm


@pytest.mark.parametrize("ts", [192, 384])
@pytest.mark.parametrize("theta", THETAS, ids=["c2theta", "c3smooth", "c3rough"])
def test_fused_update_matches_oracle(lib, ts, theta):
    """n = 5,760 (30 or 15 tile rows): the fused launch of the benchmark,
    with the FP64, FP32 and BF16 classes all present at the benchmark's
    theta."""
    pm = _check(lib, 5760, ts, theta)
    low = pm[np.tril_indices(pm.shape[0])]
    assert np.all(np.diag(pm) == 0)
    if theta == THETAS[0]:
        assert np.any(low == 1) and np.any(low == 3), np.unique(low, return_counts=True)


@pytest.mark.parametrize("theta", THETAS[1:], ids=["c3smooth", "c3rough"])
def test_ts1024_matches_oracle(lib, theta):
    """n = 6,144, ts = 1,024 (the tile size of the n = 160,000 workload)."""
    _check(lib, 6144, 1024, theta)


@pytest.mark.parametrize("ts", [192, 1024])
def test_fused_update_lolo_matches_oracle(lib, ts):
    """The four-product FP32 class (lo*lo kept) on both launch shapes."""
    _check(lib, 6144 if ts == 1024 else 5760, ts, THETAS[2], lolo=True)


def test_fused_update_fp64_band_matches_oracle(lib):
    """Strongly correlated tiles kept in FP64 (fp64_corr) in the fused
    launch: the FP64 class then also updates off-diagonal tiles from FP32
    panel tiles."""
    pm = _check(lib, 5760, 192, THETAS[2], fp64_corr=0.9)
    low = pm[np.tril_indices(pm.shape[0])]
    off = pm[np.tril_indices(pm.shape[0], -1)]
    assert np.any(off == 0) and np.any(low == 1)


def test_refinement_stopping_rule(lib):
    """The benchmark's refinement (at most 3 steps, mp_ir_tol = 1e-10) at
    n = 5,760, ts = 192: it stops as soon as the change of z^T Sigma^-1 z,
    or the changes extrapolated from the last two steps, are below the
    tolerance, with the quadratic form still at 1e-9 of the oracle's and
    status 0; with a tolerance no step can reach, all steps run and the
    status is -3, the likelihood still returned."""
    x, y, z, (ll, ld, dp) = _field(5760, THETAS[0])
    m = lib.ExactMLE(x, y, 192, precision=2)
    m.set_mixed(**dict(OPTS, ir_tol=1e-10))
    m.set_z(z)
    val = m.loglik(THETAS[0])
    done, change = m.data.mp_ir_done, m.data.mp_ir_change
    print(f"ir_tol 1e-10: {done} steps, last change {change:.2e}, rel ddotp {abs(m.dotp - dp) / dp:.2e}", flush=True)
    assert m.status == 0 and 1 <= done <= 3
    assert abs(m.dotp - dp) <= 1e-9 * abs(dp) and abs(val - ll) <= MIXED_TOL * abs(ll)
    m.set_mixed(**dict(OPTS, ir_steps=2, ir_tol=1e-300))
    val = m.loglik(THETAS[0])
    assert m.status == -3 and m.data.mp_ir_done == 2
    assert abs(val - ll) <= MIXED_TOL * abs(ll)
    m.close()


@pytest.mark.parametrize("grad_f32", [False, True], ids=["fp64_sweep", "fp32_sweep"])
@pytest.mark.parametrize("n,ts", [(1600, 320), (5760, 192)], ids=["per_class", "fused"])
def test_mixed_gradient_matches_oracle(lib, n, ts, grad_f32):
    """precision 2 with a gradient request: alpha from the refined solve,
    the trace term from the Takahashi sweep of the mixed factor (its column
    stage in FP64, or on the FP32 matrix cores with grad_f32: 64 x 64 blocks
    at ts = 320, 192 x 192 at ts = 192); every component against the
    oracle's dense FP64 gradient, relative to the gradient's largest
    component."""
    x, y, z, (ll, _, _) = _field(5760, THETAS[0])
    x, y, z = x[:n], y[:n], z[:n]
    l_ref, g_ref = likelihood.gradient(K, x, y, z, np.array(THETAS[0]))
    m = lib.ExactMLE(x, y, ts, precision=2)
    m.set_mixed(**dict(OPTS, ir_steps=8, ir_tol=1e-12, grad_f32=grad_f32))
    m.set_z(z)
    val, g = m.loglik(np.array(THETAS[0]), grad=True)
    status, check = m.status, m.data.mp_grad_check
    # the context keeps working after the in-place inverse
    val2 = m.loglik(np.array(THETAS[0]))
    m.close()
    err = np.max(np.abs(g - g_ref)) / np.max(np.abs(g_ref))
    print(f"mixed gradient n={n} ts={ts} grad_f32={grad_f32}: rel err {err:.2e}, g={g}, oracle={g_ref}",
          flush=True)
    assert status == 0
    assert abs(val - l_ref) <= MIXED_TOL * abs(l_ref) and abs(val2 - l_ref) <= MIXED_TOL * abs(l_ref)
    assert err <= GRAD_TOL, err
    # the swept inverse's residual is reported for the FP32 column stage only
    assert (0.0 < check < 1e-4) if grad_f32 else check == 0.0, check


@pytest.mark.parametrize("theta,falls_back", [((1.0, 0.1, 0.5), False), ((1.0, 0.1, 1.5), True)],
                         ids=["keeps_fp32", "falls_back"])
def test_mixed_gradient_auto_sweep(lib, theta, falls_back):
    """grad_f32 = 2: the FP32 column stage's residual along z (Z z against
    the solve with the mixed factor) decides.  At nu = 0.5 it is 2e-7, below
    the 1e-6 bound, and the gradient is the FP32 sweep's, bit for bit; at
    nu = 1.5 (a worse-conditioned Sigma) it is 1e-5, and the gradient is
    recomputed with the FP64 sweep, bit for bit the grad_f32 = 0 result.
    Either way the reported residual is the FP32 sweep's."""
    x, y, _ = datagen.generate_xy_loc(1600, 0)
    e, _ = datagen.dlarnv(3, [0, 0, 0, 1], 1600)
    th = np.array(theta)
    z = likelihood.dzvg(K, x, y, th, e)
    l_ref, g_ref = likelihood.gradient(K, x, y, z, th)
    m = lib.ExactMLE(x, y, 320, precision=2)
    m.set_z(z)
    out, checks = {}, {}
    for mode in (0, 1, 2):
        m.set_mixed(**dict(OPTS, ir_steps=8, ir_tol=1e-12, grad_f32=mode))
        _, g = m.loglik(th, grad=True)
        assert m.status == 0
        out[mode], checks[mode] = np.array(g), m.data.mp_grad_check
    m.close()
    errs = {k: np.max(np.abs(v - g_ref)) / np.max(np.abs(g_ref)) for k, v in out.items()}
    print(f"auto sweep theta={theta}: residual {checks[1]:.2e}, gradient errors {errs}", flush=True)
    assert checks[0] == 0.0 and checks[2] == checks[1] > 0.0
    assert (checks[1] > 1e-6) == falls_back
    assert np.array_equal(out[2], out[0] if falls_back else out[1])


@pytest.mark.parametrize("grad_f32", [False, True], ids=["fp64_sweep", "fp32_sweep"])
def test_mixed_gradient_bivariate_matches_oracle(lib, grad_f32):
    """precision 2 with a gradient request for the parsimonious bivariate
    kernel, all six parameters (n = 800 rows, ts = 128: per-class updates
    with 128 x 128 blocks, a partial last tile, and with grad_f32 the
    sweep's FP32 column stage on the same blocks), against the oracle's
    dense FP64 gradient, at the fixture's theta and at one with a negative
    rho."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "bivariate_400.npz"))
    kb = "bivariate_matern_parsimonious"
    for theta in (np.array(g["theta"], dtype=float), np.array([0.8, 1.2, 0.12, 0.7, 1.3, -0.4])):
        l_ref, g_ref = likelihood.gradient(kb, g["x"], g["y"], g["z"], theta)
        m = lib.ExactMLE(g["x"], g["y"], 128, kernel=kb, precision=2)
        m.set_mixed(**dict(OPTS, ir_steps=8, ir_tol=1e-12, grad_f32=grad_f32))
        m.set_z(g["z"])
        val, gr = m.loglik(theta, grad=True)
        status = m.status
        m.close()
        err = np.max(np.abs(gr - g_ref)) / np.max(np.abs(g_ref))
        print(f"mixed bivariate gradient theta={theta} grad_f32={grad_f32}: rel err {err:.2e}", flush=True)
        assert status == 0 and len(gr) == 6
        assert abs(val - l_ref) <= MIXED_TOL * abs(l_ref)
        assert err <= GRAD_TOL, (theta, gr, g_ref)


def test_single_precision_gradient_is_refused(lib):
    """precision 1 with a gradient request: status -1 and an error string,
    never a NaN gradient with status 0."""
    import ctypes

    from exageostat_amd import _lib
    x, y, z, _ = _field(5760, THETAS[0])
    m = lib.ExactMLE(x[:1600], y[:1600], 320, precision=1)
    m.set_z(z[:1600])
    th = np.array(THETAS[0])
    g = np.zeros(3)
    c = _lib.load()
    val = c.MLE_alg(3, th.ctypes.data_as(_lib.c_double_p), g.ctypes.data_as(_lib.c_double_p),
                    ctypes.cast(ctypes.pointer(m.data), ctypes.c_void_p))
    assert val == -np.inf
    assert m.data.status == -1
    assert "gradient" in c.exageostat_gpu_last_error().decode()
    # the same object still evaluates l without a gradient
    ll = m.loglik(th)
    assert np.isfinite(ll) and m.status == 0
    m.close()


def test_bivariate_kriging_matches_oracle(lib):
    """Parsimonious bivariate kriging at the theta of the 50,000-location
    workload: 600 observed and 100 held-out locations (200 held-out rows),
    per-row predictions and the per-variable MSE against the oracle."""
    kern = "bivariate_matern_parsimonious"
    theta = np.array([1.0, 1.0, 0.1, 0.5, 1.0, 0.5])
    nloc = 700
    x, y, _ = datagen.generate_xy_loc(nloc, 3)
    e, _ = datagen.dlarnv(3, [3, 3, 3, 1], 2 * nloc)
    z = likelihood.dzvg(kern, x, y, theta, e).reshape(nloc, 2)
    perm = np.random.default_rng(5).permutation(nloc)
    miss, obs = np.sort(perm[:100]), np.sort(perm[100:])
    zo, za = z[obs].ravel(), z[miss].ravel()
    ref, mse, mse1, mse2 = likelihood.predict(kern, x[obs], y[obs], zo, x[miss], y[miss], theta, za)
    m = lib.ExactMLE(x[obs], y[obs], 128, kernel=kern)
    zm, mse_gpu = m.predict(theta, x[obs], y[obs], zo, x[miss], y[miss], za)
    d1, d2 = m.data.mserror1, m.data.mserror2
    m.close()
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(zm - ref)) <= 1e-9 * scale
    assert abs(mse_gpu - mse) <= 1e-9 * mse
    assert abs(d1 - mse1) <= 1e-9 * mse1 and abs(d2 - mse2) <= 1e-9 * mse2
