// Adjoint gradient of the log-likelihood.
#pragma once

#include "engine.hpp"
#include "mixed.hpp"

namespace exa {

// l(theta) and d l / d theta = 1/2 alpha^T dS alpha - 1/2 tr(S^-1 dS).
// Sigma^-1 in place of the factor in A (lower tiles, full diagonal tiles);
// A must be the matrix ws last factored.  P and Q: mt tiles of scratch each.
// The factor is gone afterwards: ws no longer names A as factored, so a
// later potri or factor-dependent call on A fails instead of reading Sigma^-1
// as if it were L.
// W (may be null): slabs of max_split * (mt - 1) tiles that let
// the column stage split its sum over j (fills the GPU evenly).
// U (may be null): a lower tile matrix of A's grid that receives Z_mk^T, so
// that the sweep reads no transposed operands (the memory of a second
// lower-triangular matrix for a faster sweep).
// f32 (with U): the column stage on the FP32 matrix cores (operands
// rounded to FP32, FP32 sums; the trmm and diagonal stages stay FP64).
void tile_potri_inplace(TileMatrix& A, CholWorkspace& ws, double* P, double* Q, hipStream_t s, double* W = nullptr,
                        int max_split = 1, TileMatrix* U = nullptr, bool f32 = false);

double adjoint_loglik_grad(Context& ctx, KernelId kid, const double* theta, int metric, double* grad,
                           double* logdet_out, double* dotp_out, int* info);
// l from the mixed-precision path and its gradient with alpha = the refined
// Sigma^-1 z and the trace term from the Takahashi sweep of the mixed factor
// (converted to FP64 in place of the FP64 Sigma, which a lean context does
// not hold; the sweep's column stage on the FP32 matrix cores when
// opt.grad_f32 is set, in FP64 otherwise).
double mixed_adjoint_loglik_grad(Context& ctx, KernelId kid, const double* theta, int metric,
                                 const MixedOptions& opt, double* grad, double* logdet_out, double* dotp_out,
                                 int* info, MixedReport* rep);

// Fisher information F_jk = 1/2 tr(Sigma^-1 dSigma_j Sigma^-1 dSigma_k) of
// the univariate Matern kernels at theta (np x np, row-major, symmetric).
// Returns 0, or the first failing column + 1 when Sigma(theta) is not
// positive definite.  Leaves Sigma^-1 in place of the factor.
int fisher_information(Context& ctx, KernelId kid, const double* theta, int metric, double* F);

}  // namespace exa
