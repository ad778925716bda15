// Adjoint gradient of the log-likelihood.
#pragma once

#include "engine.hpp"

namespace exa {

// l(theta) and d l / d theta = 1/2 alpha^T dS alpha - 1/2 tr(S^-1 dS).
// Sigma^-1 in place of the factor in A (lower tiles, full diagonal tiles);
// A must be the matrix ws last factored.  P and Q: mt tiles of scratch each.
void tile_potri_inplace(TileMatrix& A, const CholWorkspace& ws, double* P, double* Q, hipStream_t s);

double adjoint_loglik_grad(Context& ctx, KernelId kid, const double* theta, int metric, double* grad,
                           double* logdet_out, double* dotp_out, int* info);

}  // namespace exa
