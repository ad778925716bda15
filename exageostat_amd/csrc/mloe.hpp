// Prediction-efficiency assessment: mean loss of efficiency (MLOE) and mean
// misspecification of the mean square error (MMOM) of kriging with an
// estimated theta against the true theta.
#pragma once

#include "engine.hpp"

namespace exa {

// Over the held-out locations ctx.lmiss (nmiss) given the observed
// locations ctx.lobs (nobs), univariate Matern kernels.  Per location s0,
// with k_t, k_a = Sigma(obs, s0) and K_t, K_a = Sigma(obs, obs) under the
// true and the estimated theta, w_a = K_a^-1 k_a:
//   true MSE of the estimated predictor  E2 = s_t - 2 k_t^T w_a + w_a^T K_t w_a
//   MSE of the optimal predictor         E3 = s_t - k_t^T K_t^-1 k_t
//   MSE presumed under the estimate      E4 = s_a - k_a^T w_a
// (s = theta[0]), loe = E2 / E3 - 1 (-1 when E3 == 0) and
// mom = E4 / E2 - 1 (-1 when E2 == 0); MLOE and MMOM are their means.
// loe and mom (nmiss each) may be null.  Returns 0, or the first failing
// column + 1 of K_a (then K_t) when it is not positive definite.
int mloe_mmom(Context& ctx, KernelId kid, const double* theta_true, const double* theta_est, int metric, int nobs,
              int nmiss, double* loe, double* mom, double* mloe_out, double* mmom_out);

}  // namespace exa
