"""Benchmark: exact-dense Matern log-likelihood evaluations per second on one
MI355X.

A step is one evaluation l(theta) = generate Sigma(theta) + tile Cholesky +
log-determinant + forward solve + dot product (the phase set the reference
times as "Total Time", exageostat_exact/src/compute/MLE_dexact.c:474), at
n = 90,000 locations (300 x 300 jittered grid, seed 0), tile size 960,
theta = (1, 0.1, 0.5), with locations and z = L(theta) e already resident in
HBM.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6  # MI355X dense FP64 matrix peak
PEAK_HBM_GBS = 8000.0


def flops_dpotrf(n: float) -> float:
    """FLOPS_DPOTRF (include/flops.h:97-98,254 of the reference)."""
    return n * (((1.0 / 6.0) * n + 0.5) * n + 1.0 / 3.0) + n * (((1.0 / 6.0) * n) * n - 1.0 / 6.0)


def cpu_baseline(n_sample: int, n_target: int, theta, threads: int, reps: int = 3) -> dict:
    """The CPU oracle (numpy/scipy restatement of the reference path) timed on
    this host: generation once, then `reps` timed Cholesky + log-determinant
    + solve passes at n_sample (median), extrapolated to n_target by n^2
    (generation, solve) and n^3 (factorisation)."""
    from concurrent.futures import ThreadPoolExecutor

    import scipy.linalg as sla
    from threadpoolctl import threadpool_limits

    from oracle import datagen, likelihood, matern

    x, y, _ = datagen.generate_xy_loc(n_sample, 0)
    e, _ = datagen.dlarnv(3, [0, 0, 0, 1], n_sample)
    S = np.zeros((n_sample, n_sample), order="F")
    blocks = [(i, min(i + 1000, n_sample)) for i in range(0, n_sample, 1000)]

    def gen_block(b):
        i0, i1 = b
        S[i0:i1, :i1] = matern.core_dcmg(x[i0:i1], y[i0:i1], x[:i1], y[:i1], theta)

    with threadpool_limits(limits=threads):
        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=threads) as ex:
            list(ex.map(gen_block, blocks))
        t_gen = time.perf_counter() - t0
        t_fac, t_sol = [], []
        for _ in range(reps):
            W = np.array(S, order="F", copy=True)
            t1 = time.perf_counter()
            L = sla.cholesky(W, lower=True, overwrite_a=True, check_finite=False)
            t2 = time.perf_counter()
            logdet = 2.0 * float(np.sum(np.log(np.diag(L))))
            v = sla.solve_triangular(L, e, lower=True, check_finite=False)
            dotp = float(v @ v)
            t3 = time.perf_counter()
            t_fac.append(t2 - t1)
            t_sol.append(t3 - t2)
            del W, L
    ll = -0.5 * dotp - 0.5 * logdet - (n_sample / 2.0) * math.log(2 * math.pi)
    tf, tsv = float(np.median(t_fac)), float(np.median(t_sol))
    r = n_target / n_sample
    t_target = t_gen * r ** 2 + tf * r ** 3 + tsv * r ** 2
    return {
        "value": 1.0 / t_target,
        "unit": "evals/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle (numpy/scipy OpenBLAS, thread pool) l(theta) at n={n_sample}: gen {t_gen:.2f}s once, "
                   f"dpotrf {tf:.2f}s and logdet+solve {tsv:.3f}s (median of {reps}), l={ll:.6f}; "
                   f"extrapolated to n={n_target} by n^2 (gen, solve) and n^3 (dpotrf)"),
    }


PEAK_BF16_TFLOPS = 2500.0  # MI355X dense BF16 matrix peak
# FP32 tiles are updated with split BF16 operands (hi*hi + hi*lo + lo*hi):
# three BF16 matrix products per useful product, so the useful-flop peak of
# that class is a third of the BF16 peak.
PEAK_SPLIT_BF16_TFLOPS = PEAK_BF16_TFLOPS / 3.0

# Mixed-precision options of the benchmark: tiles whose largest correlation
# is below fp32_corr get BF16 MFMA updates, the other off-diagonal tiles are
# FP32, diagonal tiles FP64; two FP64 refinement steps on the solve.  At this
# workload a third step changes z^T Sigma^-1 z by 2e-11 (relative): the
# log-determinant's error (2.4e-8 relative) is what sets |dl|, and the bench
# line reports both.
MIXED = dict(fp32_corr=1e-2, bf16=True, fp64_corr=0.0, diag_thick=1, ir_steps=2, ir_tol=0.0)


def _timed(m, theta, steps: int, warmup: int, names):
    from exageostat_amd import mle
    for _ in range(warmup):
        m.loglik(theta)
    mle.synchronize()
    mle.profile_enable(True)
    phases = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ll = m.loglik(theta)
        phases.append(m.phase_times())
    mle.synchronize()
    elapsed = time.perf_counter() - t0
    prof = {k: mle.profile_get(k) for k in names}
    mle.profile_enable(False)
    return {
        "elapsed": elapsed, "ll": ll, "status": m.status, "logdet": m.logdet, "dotp": m.dotp,
        "phases": {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}, "prof": prof,
    }


def _measured_traffic(kernel: str, n: int, ts: int):
    """HBM bytes per launch of a kernel at this workload, as measured with
    rocprofv3 FETCH_SIZE/WRITE_SIZE (scaled by a plain copy) and recorded in
    profiles/hbm_traffic.json; None when no measurement exists."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(f"{kernel}|n={n}|ts={ts}")
    except (OSError, ValueError):
        return None
    return None if rec is None else rec["bytes_per_launch"] / 1e9


def _rate(p, key):
    """(achieved units/s, per-launch ms) of one profiled kernel group."""
    if p["ms"] <= 0:
        return None, None
    return p[key] / (p["ms"] * 1e-3), p["ms"] / max(p["launches"], 1)


def _fp64_vs_mixed(m, theta, steps: int, variants: dict) -> dict:
    """Times FP64 and mixed evaluations of one ExactMLE (FP64 first, then
    each mixed option set of `variants` on the same data) and compares l and
    log det of each mixed variant with the FP64 result."""
    from exageostat_amd import mle
    out = {}
    for path in ["fp64"] + list(variants):
        if path == "fp64":
            m.precision = 0
        else:
            m.precision = 2
            m.set_mixed(**variants[path])
        m.loglik(theta)  # warm-up
        mle.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ll = m.loglik(theta)
        dt = (time.perf_counter() - t0) / steps
        rec = {"s_per_eval": dt, "evals_per_s": 1.0 / dt, "loglik": ll, "logdet": m.logdet, "dotp": m.dotp,
               "status": m.status}
        if path != "fp64":
            a = out["fp64"]
            rec["options"] = variants[path]
            rec["rel_dl"] = abs(ll - a["loglik"]) / abs(a["loglik"])
            rec["dlogdet"] = m.logdet - a["logdet"]
            rec["rel_dlogdet"] = abs(m.logdet - a["logdet"]) / abs(a["logdet"])
            rec["rel_ddotp"] = abs(m.dotp - a["dotp"]) / abs(a["dotp"])
            rec["speedup"] = a["s_per_eval"] / dt
        out[path] = rec
        print(f"bench: {path} {dt:.3f}s/eval l={ll:.10f}", file=sys.stderr, flush=True)
    m.precision = 0
    return out


# Mixed option sets measured at the larger configurations: the benchmark's
# set, the same with the lo*lo product kept in the FP32 tiles, and per
# configuration the set that keeps |dl|/|l| <= 1e-6: the log-det error left
# with lo*lo is the FP32 storage of the tiles (a three-term split of the
# products, 24 bits, measured the same error), so the remaining lever is
# FP64 storage for the most correlated tiles.  Up to 8 refinement steps (the
# refinement stops once a step changes z^T Sigma^-1 z by <= 1e-10).
_M8 = dict(MIXED, ir_steps=8, ir_tol=1e-10)
VARIANTS = {"mixed": _M8, "mixed_lolo": dict(_M8, lolo=True)}
VARIANTS_C3_ROUGH = dict(VARIANTS, mixed_lolo_fp64corr_0_7=dict(_M8, lolo=True, fp64_corr=0.7))
VARIANTS_C5 = dict(VARIANTS, mixed_lolo_nobf16_fp64corr_0_3=dict(_M8, lolo=True, bf16=False, fp64_corr=0.3))


def run_c3(steps: int) -> None:
    """Mixed-precision accuracy and speed at n = 160,000 (400 x 400 jittered
    grid), ts = 1,024, for theta = (1, 0.03, 1.0) and (1, 0.3, 0.5); z drawn
    at each theta; FP64 diagonal tiles, FP32 tiles updated with split BF16,
    BF16 far tiles (correlation < 0.01), up to 8 FP64 refinement steps; at
    theta = (1, 0.3, 0.5) also the tiles of correlation >= 0.7 in FP64."""
    import exageostat_amd
    from exageostat_amd import mle
    n, ts = 160000, 1024
    x, y = exageostat_amd.generate_xy_loc(n, 0)
    e, _ = mle.dlarnv(3, [0, 0, 0, 1], n)
    m = exageostat_amd.ExactMLE(x, y, ts)
    res = []
    for theta in ([1.0, 0.03, 1.0], [1.0, 0.3, 0.5]):
        m.zvg(theta, e)
        r = _fp64_vs_mixed(m, theta, steps, VARIANTS if theta[1] < 0.1 else VARIANTS_C3_ROUGH)
        m.precision = 2
        m.set_mixed(**MIXED)
        m.loglik(theta)
        pm = m.prec_map()
        low = pm[np.tril_indices(pm.shape[0])]
        r["tiles"] = {"f64": int(np.sum(low == 0)), "f32": int(np.sum(low == 1)), "bf16": int(np.sum(low == 3))}
        m.precision = 0
        r["theta"] = theta
        res.append(r)
    m.close()
    print(json.dumps({"config": "C3: n=160000, ts=1024, mixed vs FP64", "results": res}), flush=True)


def run_c5(steps: int) -> None:
    """Parsimonious bivariate Matern, 50,000 locations x 2 variables
    (n = 100,000 rows; 224 x 224 jittered grid truncated), theta = (1, 1, 0.1,
    0.5, 1.0, 0.5): FP64 vs mixed l, the six-parameter adjoint gradient with
    central differences (h = 1e-6 relative) for two of its components, the
    same gradient on the mixed path against the FP64 one, and kriging of 2,000 held-out locations per variable (seeded random split of
    the locations; both sets keep the Morton order)."""
    import exageostat_amd
    from exageostat_amd import mle
    nloc, ts = 50000, 960
    kern = "bivariate_matern_parsimonious"
    theta = np.array([1.0, 1.0, 0.1, 0.5, 1.0, 0.5])
    x, y = exageostat_amd.generate_xy_loc(nloc, 0)
    e, _ = mle.dlarnv(3, [0, 0, 0, 1], 2 * nloc)
    m = exageostat_amd.ExactMLE(x, y, ts, kernel=kern)
    m.zvg(theta, e)
    z = m.get_z()
    out = {"config": "C5: bivariate parsimonious, 50000 locations x 2", "theta": theta.tolist()}
    out["fp64_vs_mixed"] = _fp64_vs_mixed(m, theta, steps, VARIANTS_C5)
    t0 = time.perf_counter()
    _, g = m.loglik(theta, grad=True)
    out["grad_s"] = time.perf_counter() - t0
    out["grad"] = list(g)
    fd = {}
    for p in (4, 5):
        h = 1e-6 * abs(theta[p])
        tp, tm = theta.copy(), theta.copy()
        tp[p] += h
        tm[p] -= h
        fd[p] = (m.loglik(tp) - m.loglik(tm)) / (2 * h)
    out["grad_fd_check"] = {str(p): {"adjoint": g[p], "central_fd": v, "rel": abs(g[p] - v) / abs(v)}
                            for p, v in fd.items()}
    print(f"bench: gradient {out['grad_s']:.1f}s", file=sys.stderr, flush=True)
    # the six-parameter gradient on the mixed path (alpha from the refined
    # solve, the Takahashi sweep of the mixed factor), with the option set
    # that holds l to 1e-6 here: the sweep in FP64, then its column stage on
    # the FP32 matrix cores
    l64 = out["fp64_vs_mixed"]["fp64"]["loglik"]
    m.precision = 2
    for name, f32 in (("mixed_grad", False), ("mixed_grad_f32_sweep", True)):
        opts = dict(VARIANTS_C5["mixed_lolo_nobf16_fp64corr_0_3"], grad_f32=f32)
        m.set_mixed(**opts)
        t0 = time.perf_counter()
        lm, gm = m.loglik(theta, grad=True)
        out[name] = {"s": time.perf_counter() - t0, "options": opts, "status": m.status, "grad": list(gm),
                     "sweep_residual": m.data.mp_grad_check,
                     "max_rel_diff_vs_fp64": float(np.max(np.abs(np.asarray(gm) - np.asarray(g)))
                                                   / np.max(np.abs(g))),
                     "rel_dl_vs_fp64": abs(lm - l64) / abs(l64)}
        print(f"bench: {name} {out[name]['s']:.1f}s", file=sys.stderr, flush=True)
    m.precision = 0
    m.close()
    # kriging: 2,000 held-out locations (both variables)
    zz = z.reshape(nloc, 2)
    perm = np.random.default_rng(0).permutation(nloc)
    miss, obs = np.sort(perm[:2000]), np.sort(perm[2000:])
    xo, yo, xm, ym = x[obs], y[obs], x[miss], y[miss]
    zo = zz[obs].ravel()
    za = zz[miss].ravel()
    m = exageostat_amd.ExactMLE(xo, yo, ts, kernel=kern)
    t0 = time.perf_counter()
    zm, mse = m.predict(theta, xo, yo, zo, xm, ym, za)
    out["kriging"] = {"s": time.perf_counter() - t0, "n_obs_locations": len(xo), "n_miss_locations": len(xm),
                      "mse_var1": m.data.mserror1, "mse_var2": m.data.mserror2,
                      "mse_check_var1": float(np.mean((za[0::2] - zm[0::2]) ** 2)),
                      "mse_check_var2": float(np.mean((za[1::2] - zm[1::2]) ** 2))}
    # the same kriging through the mixed path, with the option set that
    # holds the likelihood to 1e-6 here
    opts = VARIANTS_C5["mixed_lolo_nobf16_fp64corr_0_3"]
    m.set_mixed(**opts)
    t0 = time.perf_counter()
    zm2, _ = m.predict(theta, xo, yo, zo, xm, ym, za, precision=2)
    out["kriging"]["mixed"] = {"s": time.perf_counter() - t0, "options": opts, "status": m.status,
                               "refinement_steps": m.data.mp_ir_done,
                               "max_rel_diff_vs_fp64": float(np.max(np.abs(zm2 - zm)) / np.max(np.abs(zm))),
                               "mse_var1": m.data.mserror1, "mse_var2": m.data.mserror2}
    m.close()
    print(json.dumps(out), flush=True)


def run_kriging(steps: int) -> None:
    """Kriging at n = 90,000: 1,000 held-out locations drawn at random
    (seeded) from GenerateXYLoc(90000, 0), 89,000 observed (both sets keep
    the Morton order), ts = 960, theta = (1, 0.1, 0.5), z = L(theta) e: the
    FP64 predictor (EXAGEOSTAT_dmle_Predict_Tile) against the mixed-precision
    one (EXAGEOSTAT_sdmle_Predict_Tile: C22 factored with the benchmark's
    tile classes, the solve refined in FP64 to 1e-12) and the single-
    precision one (EXAGEOSTAT_smle_Predict_Tile); one warm-up and `steps`
    timed calls each, the largest prediction difference against FP64 and the
    MSEs."""
    import exageostat_amd
    from exageostat_amd import mle
    n, nmiss, ts = 90000, 1000, 960
    theta = np.array([1.0, 0.1, 0.5])
    x, y = exageostat_amd.generate_xy_loc(n, 0)
    e, _ = mle.dlarnv(3, [0, 0, 0, 1], n)
    m = exageostat_amd.ExactMLE(x, y, ts)
    m.zvg(theta, e)
    z = m.get_z()
    m.close()
    rng = np.random.default_rng(0)
    miss = np.sort(rng.choice(n, nmiss, replace=False))
    obs = np.setdiff1d(np.arange(n), miss)
    xo, yo, zo, xm, ym, za = x[obs], y[obs], z[obs], x[miss], y[miss], z[miss]
    m = exageostat_amd.ExactMLE(xo, yo, ts)
    m.set_mixed(**dict(MIXED, ir_steps=8, ir_tol=1e-12))
    out = {"config": "kriging: n=90000, 89,000 observed, 1,000 held out, ts=960", "theta": theta.tolist()}
    ref = None
    for name, prec in (("fp64", 0), ("mixed", 2), ("single", 1)):
        m.predict(theta, xo, yo, zo, xm, ym, za, precision=prec)
        mle.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            zm, mse = m.predict(theta, xo, yo, zo, xm, ym, za, precision=prec)
        dt = (time.perf_counter() - t0) / steps
        row = {"s_per_call": dt, "mse": mse, "status": m.status}
        if ref is None:
            ref = zm
        else:
            row["max_rel_diff_vs_fp64"] = float(np.max(np.abs(zm - ref)) / np.max(np.abs(ref)))
            row["speedup_vs_fp64"] = out["fp64"]["s_per_call"] / dt
        if prec == 2:
            row["refinement_steps"] = m.data.mp_ir_done
        out[name] = row
        print(f"bench: kriging {name} {dt:.3f}s mse={mse:.8f}", file=sys.stderr, flush=True)
    m.close()
    print(json.dumps(out), flush=True)


C4_FITS = ("lbfgsb", "lbfgsb_mixed", "lbfgsb_mixed_then_fp64", "bobyqa_ftol1e-5", "bobyqa_ftol1e-7")
# the mixed path's options for the gradient-based fit: the benchmark's tile
# classes, the solve refined to 1e-12, the Takahashi sweep's column stage on
# the FP32 matrix cores unless its residual along z exceeds 1e-6 (then the
# gradient is recomputed with the FP64 sweep; with the nugget, Sigma is well
# enough conditioned for the FP32 sweep to pass: residual 4e-7 at theta0)
C4_MIXED = dict(MIXED, ir_steps=8, ir_tol=1e-12, grad_f32=2)


def run_c4(steps: int, max_evals: int = 3000, which=C4_FITS) -> None:
    """Adjoint gradient and gradient-based MLE at n = 40,000 (200 x 200
    jittered grid), nugget kernel, theta = (1, 0.1, 0.5, 0.01), ts = 960:
    one l evaluation and one l + gradient, the gradient against central
    differences (h = 1e-6 relative), then L-BFGS-B over the adjoint gradient
    from theta0 = (0.5, 0.05, 1.0, 0.05) against the reference's optimiser
    (BOBYQA over MLE_alg from the lower bounds, ftol_abs 1e-5 and, for a
    converged comparison, 1e-7), bounds [0.01, 5].  `which` selects the
    fits (the three take tens of minutes together)."""
    import exageostat_amd
    from exageostat_amd import mle, optimize
    n, ts = 40000, 960
    kern = "univariate_matern_nuggets_stationary"
    theta = np.array([1.0, 0.1, 0.5, 0.01])
    start = np.array([0.5, 0.05, 1.0, 0.05])
    x, y = exageostat_amd.generate_xy_loc(n, 0)
    e, _ = mle.dlarnv(3, [0, 0, 0, 1], n)
    # not silent: the library prints the reference's line per evaluation (on
    # stdout, before the JSON line), so a long derivative-free run shows
    # progress
    m = exageostat_amd.ExactMLE(x, y, ts, kernel=kern, silent=False)
    m.zvg(theta, e)
    out = {"config": "C4: n=40000 nugget kernel, L-BFGS-B (adjoint) vs BOBYQA", "theta_true": theta.tolist()}
    m.loglik(start)
    t0 = time.perf_counter()
    for _ in range(steps):
        m.loglik(start)
    out["loglik_s"] = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        l0, g0 = m.loglik(start, grad=True)
    out["loglik_grad_s"] = (time.perf_counter() - t0) / steps
    fd = np.zeros(4)
    for p in range(4):
        h = 1e-6 * abs(start[p])
        tp, tm = start.copy(), start.copy()
        tp[p] += h
        tm[p] -= h
        fd[p] = (m.loglik(tp) - m.loglik(tm)) / (2 * h)
    out["grad_check_max_rel"] = float(np.max(np.abs(g0 - fd)) / np.max(np.abs(fd)))
    print(f"bench: c4 l {out['loglik_s']:.3f}s, l+grad {out['loglik_grad_s']:.3f}s", file=sys.stderr, flush=True)
    # the same at precision 2: l from the mixed path, the gradient from its
    # refined solve and the Takahashi sweep of its factor
    m.precision = 2
    m.set_mixed(**C4_MIXED)
    m.loglik(start, grad=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        lm, gm = m.loglik(start, grad=True)
    mixed = {"loglik_grad_s": (time.perf_counter() - t0) / steps, "options": C4_MIXED, "status": m.status,
             "sweep_residual": m.data.mp_grad_check, "rel_dl_vs_fp64": abs(lm - l0) / abs(l0),
             "grad_max_rel_diff_vs_fp64": float(np.max(np.abs(gm - g0)) / np.max(np.abs(g0)))}
    t0 = time.perf_counter()
    for _ in range(steps):
        m.loglik(start)
    mixed["loglik_s"] = (time.perf_counter() - t0) / steps
    out["mixed"] = mixed
    m.precision = 0
    print(f"bench: c4 mixed l {mixed['loglik_s']:.3f}s, l+grad {mixed['loglik_grad_s']:.3f}s", file=sys.stderr,
          flush=True)
    lb, ub = [0.01] * 4, [5.0] * 4
    fits = {}
    for name, meth, st, tol in [("lbfgsb", "lbfgsb", start, 5), ("lbfgsb_mixed", "lbfgsb", start, 5),
                                ("lbfgsb_mixed_then_fp64", "lbfgsb_mixed", start, 5),
                                ("bobyqa_ftol1e-5", "bobyqa", None, 5), ("bobyqa_ftol1e-7", "bobyqa", None, 7),
                                ("bobyqa_ftol1e-9", "bobyqa", None, 9)]:
        if name not in which:
            continue
        m.precision = 2 if name.endswith("_mixed") else 0
        r = optimize.fit(m, lb, ub, start=st, method=meth, opt_tol=tol, max_evals=max_evals)
        m.precision = 0
        fits[name] = {"theta": r.theta.tolist(), "loglik": r.loglik, "evals": r.evals, "wall_s": r.wall_s,
                      "message": r.message}
        print(f"bench: c4 {name}: {r.wall_s:.1f}s {r.evals} evals theta={np.round(r.theta, 6)} l={r.loglik:.6f}",
              file=sys.stderr, flush=True)
    m.close()
    for k in ("lbfgsb_mixed", "lbfgsb_mixed_then_fp64", "bobyqa_ftol1e-5", "bobyqa_ftol1e-7", "bobyqa_ftol1e-9"):
        if k not in fits or "lbfgsb" not in fits:
            continue
        a = np.array(fits["lbfgsb"]["theta"])
        b = np.array(fits[k]["theta"])
        fits[k]["theta_max_rel_diff_vs_lbfgsb"] = float(np.max(np.abs(a - b) / np.abs(b)))
        fits[k]["wall_ratio_vs_lbfgsb"] = fits[k]["wall_s"] / fits["lbfgsb"]["wall_s"]
    out["fits"] = fits
    print(json.dumps(out), flush=True)


def run_mloe(steps: int) -> None:
    """MLOE / MMOM at n = 90,000: 1,000 held-out locations drawn at random
    (seeded) from GenerateXYLoc(90000, 0), 89,000 observed, ts = 960,
    theta_true = (1, 0.1, 0.5), theta_est = (0.9, 0.12, 0.55); one warm-up,
    `steps` timed calls, then one call with the per-group kernel timing."""
    import exageostat_amd
    from exageostat_amd import mle
    n, nmiss, ts = 90000, 1000, 960
    x, y = exageostat_amd.generate_xy_loc(n, 0)
    rng = np.random.default_rng(0)
    miss = np.sort(rng.choice(n, nmiss, replace=False))
    obs = np.setdiff1d(np.arange(n), miss)
    xo, yo, xm, ym = x[obs], y[obs], x[miss], y[miss]
    tt, te = (1.0, 0.1, 0.5), (0.9, 0.12, 0.55)
    m = exageostat_amd.ExactMLE(xo, yo, ts)
    m.mloe_mmom(tt, te, xo, yo, xm, ym)
    mle.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        mloe, mmom, loe, _ = m.mloe_mmom(tt, te, xo, yo, xm, ym)
    dt = (time.perf_counter() - t0) / steps
    mle.profile_enable(True)
    m.mloe_mmom(tt, te, xo, yo, xm, ym)
    groups = {}
    for g in ("tile_update", "mloe_solve", "mloe_symm"):
        p = mle.profile_get(g)
        groups[g] = {"launches": p["launches"], "ms": p["ms"],
                     "tflops": p["flops"] / (p["ms"] * 1e-3) / 1e12 if p["ms"] > 0 else None}
    mle.profile_enable(False)
    m.close()
    print(json.dumps({"config": "MLOE/MMOM: 89,000 observed, 1,000 held out, ts=960", "s_per_call": dt,
                      "mloe": mloe, "mmom": mmom, "min_loe": float(np.min(loe)), "theta_true": tt,
                      "theta_est": te, "groups": groups}), flush=True)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=90000)
    ap.add_argument("--ts", type=int, default=960)
    ap.add_argument("--cpu-sample", type=int, default=16000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp64-only", action="store_true", help="time the FP64 path only")
    ap.add_argument("--fits", default=",".join(C4_FITS), help="--config c4: comma-separated fits to run")
    ap.add_argument("--max-evals", type=int, default=3000, help="--config c4: evaluation cap of each fit")
    ap.add_argument("--config", default="c2", choices=["c2", "c3", "c4", "c5", "mloe", "kriging"],
                    help="c2 (default): the benchmark line; c3, c4, c5: the other configurations' measurements; "
                         "mloe: MLOE/MMOM at n = 90,000; kriging: FP64, mixed and single-precision kriging at "
                         "n = 90,000")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy kernel of BYTES (scales the HBM counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        nel = args.copy_calibration // 8
        a = torch.ones(nel, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": nel * 8, "bytes_written": nel * 8}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build targets one GPU; multi-GPU runs are not part of it "
              "(the tile layout shards block-cyclically over a GPU grid; one GPU is measured here)", file=sys.stderr)
        sys.exit(2)

    if args.config == "c3":
        run_c3(args.steps)
        return
    if args.config == "c4":
        run_c4(args.steps, max_evals=args.max_evals, which=args.fits.split(","))
        return
    if args.config == "c5":
        run_c5(args.steps)
        return
    if args.config == "mloe":
        run_mloe(args.steps)
        return
    if args.config == "kriging":
        run_kriging(args.steps)
        return

    import exageostat_amd
    from exageostat_amd import mle

    n, ts = args.n, args.ts
    theta = np.array([1.0, 0.1, 0.5])
    x, y = exageostat_amd.generate_xy_loc(n, 0)
    e, _ = mle.dlarnv(3, [0, 0, 0, 1], n)
    m = exageostat_amd.ExactMLE(x, y, ts)
    t0 = time.perf_counter()
    m.zvg(theta, e)
    t_zvg = time.perf_counter() - t0
    print(f"bench: n={n} ts={ts} zvg {t_zvg:.2f}s", file=sys.stderr, flush=True)

    f64 = _timed(m, theta, args.steps, args.warmup, ["tile_update", "gen", "trsv", "potrf_diag", "trsm_panel"])
    print(f"bench: fp64 {f64['elapsed'] / args.steps:.3f}s/eval l={f64['ll']:.10f}", file=sys.stderr, flush=True)
    mix = None
    if not args.fp64_only:
        m.precision = 2
        m.set_mixed(**MIXED)
        mix = _timed(m, theta, args.steps, args.warmup,
                     ["tile_update_mixed", "mixed_flops_f64", "mixed_flops_f32", "mixed_flops_bf16", "gen", "convert",
                      "trsv", "symv", "potrf_diag", "trsm_panel"])
        pm = m.prec_map()
        low = pm[np.tril_indices(pm.shape[0])]
        mix["tiles"] = {"f64": int(np.sum(low == 0)), "f32": int(np.sum(low == 1)), "bf16": int(np.sum(low == 3))}
        mix["ir"] = {"steps_done": int(m.data.mp_ir_done), "last_change": float(m.data.mp_ir_change)}
        print(f"bench: mixed {mix['elapsed'] / args.steps:.3f}s/eval l={mix['ll']:.10f}", file=sys.stderr,
              flush=True)
    m.close()

    eval_flops = flops_dpotrf(n) + n * (n + 1.0)
    f64_value = args.steps / f64["elapsed"]
    # z = L(theta) e, so at the same theta L^-1 z = e exactly in exact
    # arithmetic: a size-independent check of factor + solve.
    ee = float(e @ e)
    upd = f64["prof"]["tile_update"]
    a64, l64 = _rate(upd, "flops")
    roof64 = {
        "kernel": "k_tile_update (FP64 MFMA trailing update, 192x192 blocks)",
        "bound": "mfma", "achieved": a64 / 1e12, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
        "frac": a64 / 1e12 / PEAK_FP64_TFLOPS, "traffic": _measured_traffic("k_tile_update", n, ts),
        "traffic_unit": "GB/launch", "launches": upd["launches"], "avg_launch_ms": l64,
    }
    fp64_rec = {
        "value": f64_value, "ms_per_step": 1000.0 * f64["elapsed"] / args.steps,
        "ms_per_step_excl_gen": 1000.0 * (f64["elapsed"] / args.steps - f64["phases"]["gen"]), "loglik": f64["ll"],
        "status": f64["status"], "phases_s": f64["phases"],
        "tflops_effective": eval_flops * f64_value / 1e12,
        "solve_check_rel": abs(f64["dotp"] - ee) / ee,
        "roofline": roof64,
        "gen_GBps": (_rate(f64["prof"]["gen"], "bytes")[0] or 0) / 1e9,
        # lower-triangle entries generated per second (8 bytes each)
        "gen_Gentries_per_s": (_rate(f64["prof"]["gen"], "bytes")[0] or 0) / 8e9,
        "trsv_GBps": (_rate(f64["prof"]["trsv"], "bytes")[0] or 0) / 1e9,
    }
    if mix is None:
        value, ms_per_step, path, dtype = f64_value, fp64_rec["ms_per_step"], "fp64", "f64"
        roofline = roof64
        extra = {}
    else:
        value = args.steps / mix["elapsed"]
        ms_per_step = 1000.0 * mix["elapsed"] / args.steps
        path, dtype = "mixed", "f64/f32/bf16"
        # One fused launch per update covers the three tile classes; its peak
        # is the class-weighted MFMA peak: total flops / sum(class flops /
        # class peak), i.e. frac = (time at each class's peak) / time.
        classes = {}
        t_peak = 0.0
        for cname, peak, label in [("f64", PEAK_FP64_TFLOPS, "FP64 tiles, FP64 MFMA"),
                                   ("f32", PEAK_SPLIT_BF16_TFLOPS,
                                    "FP32 tiles, split-BF16 MFMA (3 BF16 products), FP32 accumulation"),
                                   ("bf16", PEAK_BF16_TFLOPS, "far tiles, BF16 MFMA, FP32 accumulation")]:
            fl = mix["prof"][f"mixed_flops_{cname}"]["flops"] / args.steps
            classes[cname] = {"tiles": label, "flops_per_eval": fl, "peak_tflops": peak}
            t_peak += fl / (peak * 1e12)
        p = mix["prof"]["tile_update_mixed"]
        a, lm = _rate(p, "flops")
        total_fl = sum(c["flops_per_eval"] for c in classes.values())
        peak_mix = total_fl / t_peak / 1e12 if t_peak > 0 else None
        roofline = {"kernel": "k_update_fused (FP64, split-BF16 and BF16 tile classes in one launch)",
                    "bound": "mfma", "achieved": a / 1e12, "peak": peak_mix, "unit": "TFLOP/s",
                    "frac": a / 1e12 / peak_mix, "peak_rule": "class-weighted: total flops / sum(class flops / class "
                    "peak); FP64 78.6, split BF16 833 (BF16 / 3), BF16 2500 TFLOP/s",
                    "traffic": _measured_traffic("tile_update_mixed", n, ts), "traffic_unit": "GB/launch",
                    "launches": p["launches"], "avg_launch_ms": lm, "total_ms": p["ms"] / args.steps}
        extra = {
            "rel_dl_vs_fp64": abs(mix["ll"] - f64["ll"]) / abs(f64["ll"]),
            "rel_dlogdet_vs_fp64": abs(mix["logdet"] - f64["logdet"]) / abs(f64["logdet"]),
            "rel_ddotp_vs_fp64": abs(mix["dotp"] - f64["dotp"]) / abs(f64["dotp"]),
            "speedup_vs_fp64": value / f64_value,
            # the evaluation without Sigma's generation (and its conversion
            # into the mixed factor)
            "ms_per_step_excl_gen": 1000.0 * (mix["elapsed"] / args.steps - mix["phases"]["gen"]),
            "loglik": mix["ll"], "status": mix["status"], "phases_s": mix["phases"], "tiles": mix["tiles"],
            "refinement": mix["ir"],
            "mixed_options": MIXED, "update_classes": classes,
            "symv_GBps": (_rate(mix["prof"]["symv"], "bytes")[0] or 0) / 1e9,
            "trsv_ms_total": mix["prof"]["trsv"]["ms"] / args.steps,
            "symv_ms_total": mix["prof"]["symv"]["ms"] / args.steps,
            "convert_ms": mix["prof"]["convert"]["ms"] / args.steps,
        }
    rec = {
        "metric": "log-lik evals/s at n=90k, FP64 vs mixed-prec+adjoint on 1 MI355X at stated |Δℓ|",
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (GenerateXYLoc seed 0, z = L(theta) e with e = dlarnv(3,{0,0,0,1}))",
        "config": {
            "workload": f"{path} exact Matern log-likelihood, n={n}, ts={ts}, theta=(1,0.1,0.5), "
                        "one evaluation = gen + dpotrf + logdet + solve + dot",
            "n": n, "ts": ts, "theta": theta.tolist(), "path": path,
        },
        "roofline": roofline,
        "fp64": fp64_rec,
        "zvg_s": t_zvg,
        "cpu_baseline": None,
    }
    rec.update(extra)
    if not args.no_cpu:
        # the host's CPU share for one GPU (16 cores on the MI355X boxes)
        threads = min(16, os.cpu_count() or 1)
        rec["cpu_baseline"] = cpu_baseline(args.cpu_sample, n, theta, threads)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
