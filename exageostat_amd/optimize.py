"""Maximum-likelihood fitting over MLE_alg.

The reference maximises l(theta) with NLOPT's derivative-free BOBYQA
(examples/synthetic_dmle_test.c:140-142,250,283-284): bounds [lb, ub], start
at lb, stopping on an absolute change of l below 10^-opt_tol
(misc/compute/MLE_misc.c:1352-1366) or after opt_max_iters evaluations.

NLOPT is not available to this package, so two drivers are provided:

* ``fit(..., method="dfo")``: the derivative-free comparator, scipy's COBYQA
  (Ragonneau & Zhang: a trust-region method on quadratic interpolation
  models, the bound-constrained successor of Powell's BOBYQA) with the
  reference's bounds, start and evaluation budget;
* ``fit(..., method="lbfgsb")``: L-BFGS-B driven by the adjoint gradient that
  MLE_alg returns in its grad argument.

Both evaluate l through the same C entry point as the reference's optimiser.
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np
from scipy import optimize

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
        gtol: float = 1e-6) -> FitResult:
    """Maximise l(theta) over the box [lb, ub] for the data held by the
    ExactMLE ``m`` (precision 0 or 2).

    method "dfo": COBYQA, derivative-free, stops when the trust region falls
    below 10^-opt_tol (relative to the box) or after max_evals evaluations.
    method "lbfgsb": L-BFGS-B with the adjoint gradient, stops when the
    relative decrease of -l falls below 10^-opt_tol / |l| (the reference's
    absolute ftol scaled by l) or the projected gradient below gtol * n.
    start defaults to lb, as in the reference.
    """
    lb = np.asarray(lb, dtype=float)
    ub = np.asarray(ub, dtype=float)
    x0 = lb.copy() if start is None else np.asarray(start, dtype=float)
    n = float(m.N)
    history = []
    counts = {"evals": 0, "grads": 0}

    def f_only(th):
        val = m.loglik(th)
        counts["evals"] += 1
        history.append((np.array(th), float(val)))
        return _PENALTY if not np.isfinite(val) else -val / n

    def f_and_g(th):
        val, g = m.loglik(th, grad=True)
        counts["evals"] += 1
        counts["grads"] += 1
        history.append((np.array(th), float(val)))
        if not np.isfinite(val) or not np.all(np.isfinite(g)):
            return _PENALTY, np.zeros_like(th)
        return -val / n, -np.asarray(g) / n

    t0 = time.perf_counter()
    if method == "dfo":
        res = optimize.minimize(f_only, x0, method="COBYQA", bounds=optimize.Bounds(lb, ub),
                                options={"maxfev": max_evals, "final_tr_radius": 10.0 ** (-opt_tol),
                                         "initial_tr_radius": 0.1 * float(np.min(ub - lb))})
    elif method == "lbfgsb":
        ll_scale = 10.0 ** (-opt_tol) / n
        res = optimize.minimize(f_and_g, x0, jac=True, method="L-BFGS-B", bounds=list(zip(lb, ub)),
                                options={"maxfun": max_evals, "ftol": ll_scale, "gtol": gtol})
    else:
        raise ValueError(f"unknown method {method!r}")
    wall = time.perf_counter() - t0
    best = max(history, key=lambda h: h[1] if np.isfinite(h[1]) else -np.inf)
    return FitResult(theta=best[0], loglik=best[1], evals=counts["evals"], grad_evals=counts["grads"],
                     wall_s=wall, method=method, message=str(res.message), history=history)
