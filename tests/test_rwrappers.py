"""The reference's second caller of MLE_alg: the R `.C` entry points
(r-wrappers/compute/rwrappers.c) and the flat-file readers of real data
(misc/compute/flat_file.c), called through the C ABI.

CPU: the readers parse the golden CSV fixtures (written with 17
significant digits) back to the same doubles as the fixture's arrays, bit
for bit; a missing file returns NULL.  GPU: gen_z_exact and
gen_z_givenlocs_exact reproduce the golden locations and z, exact_predict
the golden kriging predictions (1e-9), exact_mloe_mmom the oracle's MLOE
and MMOM (1e-8), and mle_exact / mle_dst / mle_mp run the reference's
BOBYQA loop over MLE_alg: the FP64 theta hat within 1e-3 of the L-BFGS-B
theta hat, the mixed-precision theta hat within 1e-3 of the FP64 one."""
import ctypes
import os

import numpy as np
import pytest

from oracle import datagen, likelihood

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
K = "univariate_matern_stationary"


@pytest.fixture(scope="module")
def clib():
    from exageostat_amd import _lib
    return _lib.load()


def _ip(v):
    return ctypes.byref(ctypes.c_int(int(v)))


def _dp(a):
    from exageostat_amd import _lib
    return a.ctypes.data_as(_lib.c_double_p)


def _free(p):
    libc = ctypes.CDLL("libc.so.6")
    libc.free.argtypes = [ctypes.c_void_p]
    libc.free(ctypes.cast(p, ctypes.c_void_p))


def test_flat_file_readers_parse_the_fixtures_exactly(clib):
    g = np.load(os.path.join(GOLD, "c1_seed0.npz"))
    loc = clib.readLocsFile(os.path.join(GOLD, "c1_locs.csv").encode(), 1600)
    x = np.ctypeslib.as_array(loc.contents.x, shape=(1600,)).copy()
    y = np.ctypeslib.as_array(loc.contents.y, shape=(1600,)).copy()
    assert not loc.contents.z
    assert np.array_equal(x, g["x"]) and np.array_equal(y, g["y"])
    _free(loc.contents.x)
    _free(loc.contents.y)
    _free(loc)
    zp = clib.readObsFile(os.path.join(GOLD, "c1_obs.csv").encode(), 1600)
    z = np.ctypeslib.as_array(zp, shape=(1600,)).copy()
    assert np.array_equal(z, g["z"])
    _free(zp)
    tp = clib.readTimeFile(os.path.join(GOLD, "c1_obs.csv").encode(), 10)
    assert np.array_equal(np.ctypeslib.as_array(tp, shape=(10,)), g["z"][:10])
    _free(tp)
    assert not clib.readLocsFile(b"/nonexistent/locs.csv", 10)
    assert not clib.readObsFile(b"/nonexistent/obs.csv", 10)


def test_flat_file_3d_reader(clib, tmp_path):
    p = tmp_path / "l3.csv"
    rows = np.random.default_rng(0).random((7, 3))
    p.write_text("".join(f"{a:.17g},{b:.17g},{c:.17g}\n" for a, b, c in rows))
    loc = clib.readLocsFile3d(str(p).encode(), 7)
    got = np.stack([np.ctypeslib.as_array(getattr(loc.contents, f), shape=(7,)) for f in "xyz"], axis=1)
    assert np.array_equal(got, rows)
    for f in "xyz":
        _free(getattr(loc.contents, f))
    _free(loc)


def test_init_data_values_defaults(clib):
    from exageostat_amd._lib import MLEData
    d = MLEData()
    d.precision = 7
    clib.init_data_values(ctypes.byref(d))
    assert d.computation == b"exact" and d.dm == b"ed" and d.precision == 0 and d.variance == 1.0
    assert d.recovery_file == b"" and not d.gpu


# ------------------------------------------------------------------ GPU ---
def _init(clib, ts):
    clib.rexageostat_init(_ip(1), _ip(1), _ip(ts), _ip(0))


@pytest.mark.gpu
def test_gen_z_exact_and_givenlocs_match_golden(clib):
    g = np.load(os.path.join(GOLD, "c1_seed0.npz"))
    _init(clib, 320)
    th = np.array([1.0, 0.1, 0.5])
    out = np.zeros(3 * 1600)
    clib.gen_z_exact(_ip(0), _dp(th), _ip(3), _ip(0), _ip(1600), _ip(0), _ip(1), _ip(1), _ip(320), _ip(1), _ip(1),
                     _ip(len(out)), _dp(out))
    assert np.array_equal(out[:1600], g["x"]) and np.array_equal(out[1600:3200], g["y"])
    assert np.max(np.abs(out[3200:] - g["z"])) <= 1e-12 * np.max(np.abs(g["z"]))
    x, y = g["x"].copy(), g["y"].copy()
    z = np.zeros(1600)
    clib.gen_z_givenlocs_exact(_dp(x), _ip(1600), _dp(y), _ip(1600), _ip(0), _dp(th), _ip(3), _ip(0), _ip(1600),
                               _ip(1), _ip(1), _ip(320), _ip(1), _ip(1), _ip(1600), _dp(z))
    # e = dlarnv(3, {0, 0, 0, 1}), the same draws as the fixture's seed 0
    assert np.max(np.abs(z - g["z"])) <= 1e-12 * np.max(np.abs(g["z"]))


@pytest.mark.gpu
def test_exact_predict_matches_golden(clib):
    g = np.load(os.path.join(GOLD, "c1_seed0.npz"))
    _init(clib, 320)
    xo, yo, zo = g["xobs"].copy(), g["yobs"].copy(), g["zobs"].copy()
    xm, ym = g["xmiss"].copy(), g["ymiss"].copy()
    th = g["theta_true"].copy()
    out = np.zeros(100)
    clib.exact_predict(_dp(xo), _ip(1500), _dp(yo), _ip(1500), _dp(zo), _ip(1500), _dp(xm), _ip(100), _dp(ym),
                       _ip(100), _ip(1500), _ip(100), _ip(0), _dp(th), _ip(3), _ip(0), _ip(0), _ip(1), _ip(1),
                       _ip(320), _ip(1), _ip(1), _dp(out))
    assert np.max(np.abs(out - g["zmiss"])) <= 1e-9 * np.max(np.abs(g["zmiss"]))


@pytest.mark.gpu
def test_exact_mloe_mmom_matches_oracle(clib):
    x, y, _ = datagen.generate_xy_loc(900, 2)
    rng = np.random.default_rng(2)
    miss = np.sort(rng.permutation(900)[:60])
    obs = np.setdiff1d(np.arange(900), miss)
    tt, te = np.array([1.0, 0.1, 0.5]), np.array([0.9, 0.12, 0.55])
    ref = likelihood.mloe_mmom(K, x[obs], y[obs], x[miss], y[miss], tt, te)
    _init(clib, 192)
    xo, yo, xm, ym = x[obs].copy(), y[obs].copy(), x[miss].copy(), y[miss].copy()
    zo = np.zeros(len(obs))
    out = np.zeros(2)
    clib.exact_mloe_mmom(_dp(xo), _ip(len(obs)), _dp(yo), _ip(len(obs)), _dp(zo), _ip(len(obs)), _dp(xm),
                         _ip(60), _dp(ym), _ip(60), _ip(len(obs)), _ip(60), _ip(0), _dp(te), _dp(tt), _ip(3), _ip(0),
                         _ip(1), _ip(1), _ip(192), _ip(1), _ip(1), _dp(out))
    assert abs(out[0] - ref[0]) <= 1e-8 * max(1.0, abs(ref[0])) and abs(out[1] - ref[1]) <= 1e-8 * max(1.0, abs(ref[1]))


def _mle(clib, fn, x, y, z, extra_int=None, tol=1e-9, max_iters=0, ts=320):
    n = len(z)
    clb, cub = np.full(3, 0.01), np.full(3, 5.0)
    out = np.zeros(6)
    xs, ys, zs = x.copy(), y.copy(), z.copy()
    tolp = ctypes.byref(ctypes.c_double(tol))
    if fn == "mle_exact":
        clib.mle_exact(_dp(xs), _ip(n), _dp(ys), _ip(n), _dp(zs), _ip(n), _dp(clb), _ip(3), _dp(cub), _ip(3), _ip(0),
                       _ip(0), _ip(n), tolp, _ip(max_iters), _ip(1), _ip(1), _ip(ts), _ip(1), _ip(1), _dp(out))
    else:
        getattr(clib, fn)(_dp(xs), _ip(n), _dp(ys), _ip(n), _dp(zs), _ip(n), _ip(0), _dp(clb), _ip(3), _dp(cub),
                          _ip(3), _ip(extra_int), _ip(0), _ip(n), tolp, _ip(max_iters), _ip(1), _ip(1), _ip(ts),
                          _ip(1), _ip(1), _dp(out))
    return out


@pytest.mark.gpu
def test_mle_exact_dst_and_mp_reach_theta_hat(clib):
    """The reference's MLE run through the R entry points at n = 1,600
    (BOBYQA from the lower bounds, ftol_abs 1e-9: the likelihood is flat
    along sigma^2 / beta, where a step that changes l by less than 1e-7 can
    still move sigma^2 by 1e-3): mle_exact reaches the
    L-BFGS-B theta hat within 1e-3; mle_mp (mixed precision, FP64 band 1)
    stops on that ridge where its own rounding (|dl|/|l| about 1e-7) leads
    BOBYQA, so its theta hat is checked by the exact likelihood: l at the
    mixed theta hat is within 1e-6 |l| (the mixed tier's accuracy bar) of l
    at the FP64 theta hat, and every component within 1e-2; mle_dst with a
    band covering every tile is the exact likelihood and gives the same
    theta hat."""
    import exageostat_amd
    from exageostat_amd import optimize
    g = np.load(os.path.join(GOLD, "c1_seed0.npz"))
    _init(clib, 320)
    print("mle_exact n=1600", flush=True)
    ex = _mle(clib, "mle_exact", g["x"], g["y"], g["z"])
    m = exageostat_amd.ExactMLE(g["x"], g["y"], 320)
    m.set_z(g["z"])
    ref = optimize.fit(m, [0.01] * 3, [5.0] * 3, method="lbfgsb", opt_tol=9)
    m.close()
    th_ex = ex[:3]
    print(f"mle_exact theta {th_ex} evals {ex[5]:.0f}; lbfgsb {ref.theta} evals {ref.evals}", flush=True)
    assert np.max(np.abs(th_ex - ref.theta) / ref.theta) <= 1e-3
    assert ex[5] >= 7 and ex[4] > 0 and abs(ex[3] - ex[4] / ex[5]) <= 1e-12 * ex[4]
    print("mle_mp n=1600", flush=True)
    mp = _mle(clib, "mle_mp", g["x"], g["y"], g["z"], extra_int=1)
    m = exageostat_amd.ExactMLE(g["x"], g["y"], 320)
    m.set_z(g["z"])
    l_ex, l_mp = m.loglik(th_ex), m.loglik(mp[:3])
    m.close()
    rel = np.max(np.abs(mp[:3] - th_ex) / th_ex)
    print(f"mle_mp theta {mp[:3]} evals {mp[5]:.0f}: max rel diff {rel:.2e}, exact l {l_mp:.10f} "
          f"against {l_ex:.10f} at the FP64 theta hat", flush=True)
    assert l_ex - l_mp <= 1e-6 * abs(l_ex)
    assert rel <= 1e-2
    print("mle_dst n=1600 band 5", flush=True)
    dst = _mle(clib, "mle_dst", g["x"], g["y"], g["z"], extra_int=5, max_iters=400)
    assert np.max(np.abs(dst[:3] - th_ex) / th_ex) <= 1e-6


@pytest.mark.gpu
def test_bobyqa_against_gradient_method_at_1600(clib):
    """The reference's optimiser (BOBYQA over MLE_alg from the lower
    bounds, ftol_abs 1e-9, tight enough for theta along the flat
    sigma^2 / beta direction) and L-BFGS-B over the adjoint gradient reach
    the same theta hat within 1e-3 at n = 1,600; the gradient method needs
    at least three times fewer likelihood evaluations."""
    import exageostat_amd
    from exageostat_amd import optimize
    g = np.load(os.path.join(GOLD, "c1_seed0.npz"))
    m = exageostat_amd.ExactMLE(g["x"], g["y"], 320)
    m.set_z(g["z"])
    rb = optimize.fit(m, [0.01] * 3, [5.0] * 3, method="bobyqa", opt_tol=9, max_evals=3000)
    rg = optimize.fit(m, [0.01] * 3, [5.0] * 3, method="lbfgsb", opt_tol=9)
    m.close()
    print(f"bobyqa {rb.theta} {rb.evals} evals {rb.wall_s:.2f}s; lbfgsb {rg.theta} {rg.evals} evals "
          f"{rg.wall_s:.2f}s", flush=True)
    assert np.max(np.abs(rb.theta - rg.theta) / rg.theta) <= 1e-3
    assert 3 * rg.evals <= rb.evals
