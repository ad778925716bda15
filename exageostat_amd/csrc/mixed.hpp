// Mixed-precision likelihood: FP64 diagonal tiles, FP32/BF16 off-diagonal
// tiles chosen per theta, FP64 iterative refinement of the solve.
#pragma once

#include "engine.hpp"

namespace exa {

struct MixedOptions {
  double fp64_corr = 0.0;  // off-diagonal tiles with max correlation >= this stay FP64 (0: none)
  double fp32_corr = 0.0;  // ... >= this are FP32, the rest BF16 (when bf16)
  bool bf16 = false;
  int ir_steps = 3;   // at most this many FP64 refinement steps
  // > 0: stop once a step changes z^T x by <= ir_tol (relative), or once the
  // changes still to come, extrapolated from the last two (linear
  // convergence with ratio < 1/2), sum to <= ir_tol
  double ir_tol = 0;
  int diag_thick = 1;
  bool lolo = false;  // FP32 class: hi*hi + hi*lo + lo*hi (+ lo*lo when set)
  // gradient: 0 runs the Takahashi sweep in FP64; 1 runs its column stage on
  // the FP32 matrix cores (its error grows with the condition number of
  // Sigma) and reports the sweep's residual; 2 does the same and recomputes
  // the gradient with the FP64 sweep when the residual exceeds
  // kSweepResidualMax
  int grad_f32 = 0;
};
constexpr double kSweepResidualMax = 1e-6;

// What the refinement did in the last mixed evaluation.
struct MixedReport {
  int ir_done = 0;         // refinement steps run
  double ir_change = 0.0;  // |z^T c| / |z^T x| of the last step's correction c (0: no step)
  bool ir_converged = true;  // false: ir_tol > 0 and the stopping test was never met
  // gradient with the FP32 column stage: ||Z z - x0|| / ||x0|| (Z the swept
  // inverse, x0 the solve with the factor it inverted); 0 when not run
  double sweep_residual = 0.0;
};

double mixed_loglik(Context& ctx, KernelId kid, const double* theta, int metric, const MixedOptions& opt,
                    double* logdet_out, double* dotp_out, int* info, MixedReport* rep = nullptr);
// dst (FP64, the mixed factor's tile grid) <- the mixed factor of the last
// mixed evaluation, every tile widened to FP64; ws then names dst as the
// factored matrix (its inverted diagonal tiles are the mixed factor's).
void mixed_factor_to_fp64(Context& ctx, TileMatrix& dst, hipStream_t s);
// The check of an FP32-column-stage sweep: mixed_probe_solve keeps
// x0 = F^-T F^-1 z for the widened factor F (call it before the sweep),
// mixed_probe_residual returns ||Z z - x0|| / ||x0|| for the swept Z.
void mixed_probe_solve(Context& ctx, TileMatrix& F, hipStream_t s);
double mixed_probe_residual(Context& ctx, TileMatrix& Zinv, hipStream_t s);
// Copies the tile precision map (TilePrec, lower-packed tile order) of the
// last mixed evaluation into out[0:cap); returns the tile count (0: none).
long mixed_prec_map(const Context& ctx, uint8_t* out, long cap);

}  // namespace exa
