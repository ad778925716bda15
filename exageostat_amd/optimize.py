"""Maximum-likelihood fitting over MLE_alg.

The reference maximises l(theta) with NLOPT's derivative-free BOBYQA
(examples/synthetic_dmle_test.c:140-142,250,283-284): bounds [lb, ub], start
at lb, stopping on an absolute change of l below 10^-opt_tol
(misc/compute/MLE_misc.c:1352-1366) or after opt_max_iters evaluations.

NLOPT is not available to this package, so three drivers are provided:

* ``fit(..., method="bobyqa")``: the reference's run with the library's own
  BOBYQA (``exageostat_bobyqa``, Powell 2009, with NLOPT's conventions:
  2n+1 points, default initial step, start at lb, stop when an evaluation
  is within 10^-opt_tol of the best or after max_evals).  For an ExactMLE
  the objective handed to it is the C function MLE_alg itself with the
  MLE_data pointer, as the reference hands MLE_alg to NLOPT
  (synthetic_dmle_test.c:283), so no Python runs inside the loop;
* ``fit(..., method="dfo")``: scipy's COBYQA (Ragonneau & Zhang, a
  trust-region method on quadratic interpolation models) with the
  reference's bounds, start and evaluation budget;
* ``fit(..., method="lbfgsb")``: L-BFGS-B driven by the adjoint gradient that
  MLE_alg returns in its grad argument.

All evaluate l through the same C entry point as the reference's optimiser.
"""
from __future__ import annotations

import ctypes
import dataclasses
import time

import numpy as np
from scipy import optimize

from . import _lib

# l is -inf (and the gradient NaN) where Sigma(theta) is not positive
# definite; the minimisers get this finite penalty instead.
_PENALTY = 1e30


@dataclasses.dataclass
class FitResult:
    theta: np.ndarray
    loglik: float
    evals: int          # likelihood evaluations (with or without gradient)
    grad_evals: int     # of which with gradient
    wall_s: float
    method: str
    message: str
    history: list       # (theta, loglik) per evaluation


def fit(m, lb, ub, start=None, method: str = "lbfgsb", opt_tol: int = 5, max_evals: int = 500,
        gtol: float = 1e-6, log_params: bool | None = None) -> FitResult:
    """Maximise l(theta) over the box [lb, ub] for the data held by the
    ExactMLE ``m`` (precision 0 or 2).

    method "bobyqa": the library's BOBYQA, derivative-free, stops when an
    evaluation lies within 10^-opt_tol of the best value (NLOPT's ftol_abs,
    as the reference sets it) or after max_evals evaluations.
    method "dfo": COBYQA, derivative-free, stops when the trust region falls
    below 10^-(opt_tol+1) of the bound box or after max_evals evaluations.
    method "lbfgsb": L-BFGS-B with the adjoint gradient, stops when the
    relative decrease of -l falls below 10^-opt_tol / |l| (the reference's
    absolute ftol scaled by l) or the projected gradient below gtol * n.
    method "lbfgsb_mixed": L-BFGS-B over the mixed-precision likelihood and
    gradient (precision 2 under the model's set_mixed options) until the
    relative decrease of -l falls below 10^-min(opt_tol, 3) / |l| (above
    the mixed likelihood's own rounding), then "lbfgsb" over the FP64 ones
    from that point; theta hat and l come from the FP64 stage.
    start defaults to lb, as in the reference.
    log_params: optimise over log(theta) (all bounds must be positive); the
    covariance parameters span orders of magnitude (variance ~1, range
    ~0.1, nugget ~0.01), which the quasi-Newton model handles far better on
    a log scale.  Default: on for "lbfgsb", off for "dfo" (the reference's
    BOBYQA works on theta itself).
    """
    if method == "lbfgsb_mixed":
        prec = m.precision
        try:
            m.precision = 2
            r1 = fit(m, lb, ub, start, "lbfgsb", opt_tol=min(opt_tol, 3), max_evals=max_evals, gtol=gtol,
                     log_params=log_params)
            m.precision = 0
            r2 = fit(m, lb, ub, r1.theta, "lbfgsb", opt_tol=opt_tol, max_evals=max(1, max_evals - r1.evals),
                     gtol=gtol, log_params=log_params)
        finally:
            m.precision = prec
        return FitResult(theta=r2.theta, loglik=r2.loglik, evals=r1.evals + r2.evals,
                         grad_evals=r1.grad_evals + r2.grad_evals, wall_s=r1.wall_s + r2.wall_s, method=method,
                         message=f"mixed stage ({r1.evals} evaluations): {r1.message}; FP64 stage "
                                 f"({r2.evals}): {r2.message}", history=r1.history + r2.history)
    if log_params is None:
        log_params = method == "lbfgsb"
    lb = np.asarray(lb, dtype=float)
    ub = np.asarray(ub, dtype=float)
    x0 = lb.copy() if start is None else np.asarray(start, dtype=float)
    n = float(m.N)
    history = []
    counts = {"evals": 0, "grads": 0}

    if log_params and (np.any(lb <= 0) or np.any(x0 <= 0)):
        raise ValueError("log_params needs positive bounds and start")
    to_theta = np.exp if log_params else (lambda u: u)
    # chain rule for d/du with theta = exp(u): dl/du = theta * dl/dtheta
    jac_scale = (lambda th: th) if log_params else (lambda th: np.ones_like(th))
    u0 = np.log(x0) if log_params else x0
    ulb = np.log(lb) if log_params else lb
    uub = np.log(ub) if log_params else ub

    def f_only(u):
        th = np.clip(to_theta(np.asarray(u)), lb, ub)
        val = m.loglik(th)
        counts["evals"] += 1
        history.append((np.array(th), float(val)))
        return _PENALTY if not np.isfinite(val) else -val / n

    good = {"u": None, "f": None, "penalized": False}

    def f_and_g(u):
        u = np.asarray(u, dtype=float)
        th = np.clip(to_theta(u), lb, ub)
        val, g = m.loglik(th, grad=True)
        counts["evals"] += 1
        counts["grads"] += 1
        history.append((np.array(th), float(val)))
        if not np.isfinite(val) or not np.all(np.isfinite(g)):
            # Sigma(theta) not positive definite: a smooth bowl around the
            # last feasible point, above every value seen there, so that the
            # line search backtracks towards it instead of stalling.
            good["penalized"] = True
            if good["u"] is None:
                return _PENALTY, np.zeros_like(u)
            scale = 1.0 + abs(good["f"])
            d = u - good["u"]
            return good["f"] + scale * (1.0 + 1e3 * float(d @ d)), 2e3 * scale * d
        f = -val / n
        good["u"], good["f"] = u.copy(), f
        return f, -np.asarray(g) * jac_scale(th) / n

    t0 = time.perf_counter()
    if method == "bobyqa":
        lib = _lib.load()
        k = len(lb)
        x = np.ascontiguousarray(x0, dtype=np.float64).copy()
        lbc = np.ascontiguousarray(lb, dtype=np.float64)
        ubc = np.ascontiguousarray(ub, dtype=np.float64)
        fmax = ctypes.c_double(0.0)
        nev = ctypes.c_int(0)
        if getattr(m, "data", None) is not None and hasattr(m, "lib"):
            # the C objective: MLE_alg(n, theta, grad, &data)
            fn = _lib.OBJECTIVE(ctypes.cast(lib.MLE_alg, ctypes.c_void_p).value)
            fdata = ctypes.cast(ctypes.pointer(m.data), ctypes.c_void_p)
        else:
            def pyobj(nn, xp, gp, _d):
                th = np.array([xp[i] for i in range(nn)])
                val = float(m.loglik(th))
                history.append((th, val))
                return val if np.isfinite(val) else -np.inf
            fn = _lib.OBJECTIVE(pyobj)
            fdata = None
        rc = lib.exageostat_bobyqa(k, fn, fdata, lbc.ctypes.data_as(_lib.c_double_p),
                                   ubc.ctypes.data_as(_lib.c_double_p), x.ctypes.data_as(_lib.c_double_p),
                                   10.0 ** (-opt_tol), int(max_evals), ctypes.byref(fmax), ctypes.byref(nev))
        if rc < 0:
            raise RuntimeError(f"exageostat_bobyqa failed ({rc})")
        counts["evals"] = nev.value
        if not history:
            history.append((x.copy(), fmax.value))

        class _Res:
            message = {0: "ftol_abs reached", 1: "max_evals reached", 2: "trust region at its floor"}[rc]
        res = _Res()
    elif method == "dfo":
        # Variables scaled to the bound box (BOBYQA's steps are relative to
        # the box as well); stop when the trust region is below
        # 10^-(opt_tol+1) of the box.
        res = optimize.minimize(f_only, u0, method="COBYQA", bounds=optimize.Bounds(ulb, uub),
                                options={"maxfev": max_evals, "scale": True,
                                         "final_tr_radius": 10.0 ** (-opt_tol - 1)})
    elif method == "lbfgsb":
        ll_scale = 10.0 ** (-opt_tol) / n
        ustart = u0
        for _ in range(4):
            # A run that met an infeasible point may stop early: restart from
            # the best point with a fresh quasi-Newton memory.
            good["penalized"] = False
            res = optimize.minimize(f_and_g, ustart, jac=True, method="L-BFGS-B", bounds=list(zip(ulb, uub)),
                                    options={"maxfun": max(1, max_evals - counts["evals"]), "ftol": ll_scale,
                                             "gtol": gtol})
            if not good["penalized"] or counts["evals"] >= max_evals:
                break
            best_th = max(history, key=lambda h: h[1] if np.isfinite(h[1]) else -np.inf)[0]
            ustart = np.log(best_th) if log_params else best_th
    else:
        raise ValueError(f"unknown method {method!r}")
    wall = time.perf_counter() - t0
    best = max(history, key=lambda h: h[1] if np.isfinite(h[1]) else -np.inf)
    return FitResult(theta=best[0], loglik=best[1], evals=counts["evals"], grad_evals=counts["grads"],
                     wall_s=wall, method=method, message=str(res.message), history=history)


def _floats(s: str):
    return [float(v) for v in s.split(",")]


def main(argv=None) -> None:
    """Synthetic-data MLE in the manner of the reference's example driver
    (examples/synthetic_dmle_test.c): GenerateXYLoc + dlarnv + MLE_zvg at
    theta, then both drivers from the same start, one JSON line with theta
    hat, evaluations and wall time of each, and a central-difference check
    of the adjoint gradient at the start point."""
    import argparse
    import json
    import sys

    from . import mle

    ap = argparse.ArgumentParser(prog="python -m exageostat_amd.optimize")
    ap.add_argument("--n", type=int, default=40000)
    ap.add_argument("--ts", type=int, default=960)
    ap.add_argument("--kernel", default="univariate_matern_nuggets_stationary")
    ap.add_argument("--theta", type=_floats, default=[1.0, 0.1, 0.5, 0.01])
    ap.add_argument("--start", type=_floats, default=[0.5, 0.05, 1.0, 0.05])
    ap.add_argument("--lb", type=float, default=0.01)
    ap.add_argument("--ub", type=float, default=5.0)
    ap.add_argument("--opt-tol", type=int, default=5)
    ap.add_argument("--max-evals", type=int, default=500)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--methods", default="lbfgsb,bobyqa")
    ap.add_argument("--bobyqa-start", default="lb", choices=["lb", "start"],
                    help="BOBYQA's start: the lower bounds (the reference's, examples/examples.c:263-264) or --start")
    args = ap.parse_args(argv)

    x, y = mle.generate_xy_loc(args.n, args.seed)
    e, _ = mle.dlarnv(3, [args.seed] * 3 + [1], args.n)
    # not silent: the library prints the reference's per-evaluation line
    m = mle.ExactMLE(x, y, args.ts, kernel=args.kernel, silent=False)
    m.zvg(args.theta, e)
    k = len(args.theta)
    lb, ub = [args.lb] * k, [args.ub] * k
    start = np.asarray(args.start, dtype=float)
    out = {"n": args.n, "ts": args.ts, "kernel": args.kernel, "theta_true": args.theta, "start": list(start),
           "bounds": [args.lb, args.ub], "opt_tol": args.opt_tol}
    # adjoint gradient vs central differences (h = 1e-6 relative) at the start
    t0 = time.perf_counter()
    l0, g0 = m.loglik(start, grad=True)
    out["grad_eval_s"] = time.perf_counter() - t0
    fd = np.zeros(k)
    t0 = time.perf_counter()
    for p in range(k):
        h = 1e-6 * abs(start[p])
        tp, tm = start.copy(), start.copy()
        tp[p] += h
        tm[p] -= h
        fd[p] = (m.loglik(tp) - m.loglik(tm)) / (2.0 * h)
    out["loglik_eval_s"] = (time.perf_counter() - t0) / (2 * k)
    out["grad_check"] = {"adjoint": list(g0), "central_fd": list(fd),
                         "max_rel": float(np.max(np.abs(g0 - fd)) / np.max(np.abs(fd)))}
    print(f"gradient check {out['grad_check']['max_rel']:.2e}", file=sys.stderr, flush=True)
    res = {}
    for meth in args.methods.split(","):
        st = None if (meth == "bobyqa" and args.bobyqa_start == "lb") else start
        r = fit(m, lb, ub, start=st, method=meth, opt_tol=args.opt_tol, max_evals=args.max_evals)
        res[meth] = {"theta": list(r.theta), "loglik": r.loglik, "evals": r.evals, "grad_evals": r.grad_evals,
                     "wall_s": r.wall_s, "message": r.message}
        print(f"{meth}: {r.wall_s:.1f}s {r.evals} evals theta={np.round(r.theta, 6)} l={r.loglik:.6f}",
              file=sys.stderr, flush=True)
    out["fits"] = res
    for other in ("bobyqa", "dfo"):
        if "lbfgsb" in res and other in res:
            a, b = np.array(res["lbfgsb"]["theta"]), np.array(res[other]["theta"])
            out[f"theta_max_rel_diff_{other}"] = float(np.max(np.abs(a - b) / np.abs(b)))
            out[f"wall_ratio_{other}_over_lbfgsb"] = res[other]["wall_s"] / res["lbfgsb"]["wall_s"]
            out[f"loglik_diff_{other}"] = res["lbfgsb"]["loglik"] - res[other]["loglik"]
    m.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
