// Adjoint gradient of the exact log-likelihood:
//   d l / d theta_p = 1/2 alpha^T dSigma_p alpha - 1/2 tr(Sigma^-1 dSigma_p)
//                   = -1/2 <Sigma^-1 - alpha alpha^T, dSigma_p>_F,
// with alpha = Sigma^-1 z from the same factor as l, Sigma^-1 from the block
// Takahashi sweep over that factor (in place), and dSigma_p regenerated
// entry by entry inside the reduction (never stored).  d/dnu uses the
// nu-derivative of K_nu carried through the Bessel evaluation as a dual
// number.
//
// The reference computes l only (EXAGEOSTAT_dmle_Tile, exageostat_exact/
// src/compute/MLE_dexact.c:231-489) and leaves MLE_alg's grad argument
// untouched; the derivative formulas of d/dsigma^2 and d/dbeta follow its
// derivative tiles (core_zcmg.c:315-342, 453-488).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "adjoint.hpp"
#include "profile.hpp"

namespace exa {

struct AdjointState {
  DeviceVector alpha;
  // Fisher information: dSigma_p, Z dSigma_p Z (lower tiles), Z dSigma_p
  // (full tile matrix) and a zero vector.
  TileMatrix D, M;
  double* Y = nullptr;
  long y_tiles = 0;
  DeviceVector zero;
  double* P = nullptr;  // mt tiles: L_jk X_kk of the current column
  double* Q = nullptr;  // mt tiles: split-k terms of the diagonal tile
  double* W = nullptr;  // kPotriSplit * mt tiles: split-j slabs of the column stage
  TileMatrix U;         // Z_mk^T of the sweep (lower tiles), read in place of transposed Z tiles
  long tiles = 0;
  double* partial = nullptr;
  long partial_blocks = 0;
  double* sums = nullptr;  // 8 * kGroups
  KnuOrder<Dual>* ko = nullptr;  // device: K_nu constants of the bivariate kernel
  ~AdjointState() {
    if (Y) (void)hipFree(Y);
    if (ko) (void)hipFree(ko);
    if (P) (void)hipFree(P);
    if (Q) (void)hipFree(Q);
    if (W) (void)hipFree(W);
    if (partial) (void)hipFree(partial);
    if (sums) (void)hipFree(sums);
  }
};

static constexpr int kGroups = 256;
static constexpr int kPotriSplit = 4;

void destroy_adjoint(AdjointState* a) { delete a; }

// Z = Sigma^-1 in place of the factor held in A (tile_potrf output, with the
// inverted diagonal tiles in ws).
void tile_potri_inplace(TileMatrix& A, CholWorkspace& ws, double* P, double* Q, hipStream_t s, double* W,
                        int max_split, TileMatrix* U, bool f32) {
  TileView uv{};
  if (U) {
    if (!U->allocated() || U->mt != A.mt || U->ts != A.ts || !U->lower)
      throw std::invalid_argument("potri: the transposed copy must have the factor's tile grid");
    uv = U->view();
  }
  const TileView* up = U ? &uv : nullptr;
  if (ws.factored != A.tile_raw(0, 0) || ws.mt_ < A.mt || ws.ts_ != A.ts)
    throw std::invalid_argument("potri: the matrix is not the last one factored");
  const TileView v = A.view();
  const int mt = A.mt;
  auto rows = [&](int t) { return (double)std::min<long>(A.ts, (long)A.m - (long)t * A.ts); };
  for (int k = mt - 1; k >= 0; --k) {
    const double below = (double)A.m - (double)(k + 1) * A.ts;
    const double rk = rows(k);
    if (k + 1 < mt) {
      profiled("potri_trmm", below * rk * rk, 0.0, s,
               [&] { launch_takahashi_stage(v, k, ws.xinv(k), P, Q, A.tile(k, k), 0, s); });
      profiled("potri_col", 2.0 * below * below * rk, 0.0, s,
               [&] { launch_takahashi_stage(v, k, ws.xinv(k), P, Q, A.tile(k, k), 1, s, W, max_split, up, f32); });
    }
    profiled("potri_diag", 2.0 * below * rk * rk + rk * rk * rk / 3.0, 0.0, s,
             [&] { launch_takahashi_stage(v, k, ws.xinv(k), P, Q, A.tile(k, k), 2, s, nullptr, 1, up); });
  }
  ws.factored = nullptr;
}

// Scratch of the Takahashi sweep and of the dSigma reductions, sized for
// ctx.C and kept across calls.
static AdjointState& adjoint_buffers(Context& ctx) {
  TileMatrix& A = ctx.fp64("the gradient and the Fisher information");
  if (!ctx.adjoint) ctx.adjoint = new AdjointState();
  AdjointState& st = *ctx.adjoint;
  const long tile_elems = (long)A.ts * A.ts;
  if (st.tiles < A.mt) {
    if (st.P) (void)hipFree(st.P);
    if (st.Q) (void)hipFree(st.Q);
    if (st.W) (void)hipFree(st.W);
    st.P = st.Q = st.W = nullptr;
    EXA_HIP(hipMalloc(&st.P, sizeof(double) * tile_elems * A.mt));
    EXA_HIP(hipMalloc(&st.Q, sizeof(double) * tile_elems * A.mt));
    EXA_HIP(hipMalloc(&st.W, sizeof(double) * tile_elems * A.mt * kPotriSplit));
    st.tiles = A.mt;
  }
  const long nb = grad_partial_blocks(A.view());
  if (st.partial_blocks < nb) {
    if (st.partial) (void)hipFree(st.partial);
    st.partial = nullptr;
    EXA_HIP(hipMalloc(&st.partial, sizeof(double) * 8 * nb));
    st.partial_blocks = nb;
  }
  if (!st.sums) EXA_HIP(hipMalloc(&st.sums, sizeof(double) * 8 * kGroups));
  if (!st.U.allocated() || st.U.m != A.m || st.U.ts != A.ts) {
    st.U.release();
    st.U.allocate(A.m, A.m, A.ts, true);
  }
  return st;
}

static MaternGradParams matern_grad_params(const Context& ctx, KernelId kid, const double* theta, int metric) {
  MaternGradParams G;
  G.x = ctx.l1.dx;
  G.y = ctx.l1.dy;
  G.n = ctx.N;
  G.metric = metric;
  G.nugget = kid == kNugget ? 1 : 0;
  G.sigma2 = theta[0];
  G.beta = theta[1];
  G.con = host_matern_con(theta[2]);
  knu_prepare(theta[2], G.ko);
  return G;
}

// Sums the per-group partials of a reduction launch: out[p], p < nsum.
static void reduce_sums(AdjointState& st, int nsum, hipStream_t s, double* out) {
  std::vector<double> h((size_t)nsum * kGroups);
  EXA_HIP(hipMemcpyAsync(h.data(), st.sums, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
  EXA_HIP(hipStreamSynchronize(s));
  for (int p = 0; p < nsum; ++p) {
    out[p] = 0.0;
    for (int g = 0; g < kGroups; ++g) out[p] += h[(size_t)p * kGroups + g];
  }
}

// The gradient from a factor of Sigma held in FP64 in A (ws.factored ==
// A): alpha = Sigma^-1 z (ctx.work holds L^-1 z, or with work_is_alpha
// Sigma^-1 z itself), Z = Sigma^-1 in place of the factor by the Takahashi
// sweep, then the fused dSigma reduction.
static void gradient_from_factor(Context& ctx, KernelId kid, const double* theta, int metric, TileMatrix& A,
                                 bool work_is_alpha, double* grad, bool f32_sweep = false) {
  const int np = kernel_num_params(kid);
  AdjointState& st = adjoint_buffers(ctx);
  const TileView v = A.view();
  hipStream_t s = ctx.s1;

  st.alpha.allocate(ctx.N, ctx.work.padded);
  EXA_HIP(hipMemcpyAsync(st.alpha.d, ctx.work.d, sizeof(double) * ctx.work.padded, hipMemcpyDeviceToDevice, s));
  if (!work_is_alpha)  // alpha = L^-T (L^-1 z)
    profiled("trsv", (double)ctx.N * (ctx.N + 1.0), 4.0 * (double)ctx.N * (ctx.N + 1.0), s,
             [&] { tile_trsv(A, ctx.ws, st.alpha.d, 1, s); });
  tile_potri_inplace(A, ctx.ws, st.P, st.Q, s, st.W, kPotriSplit, &st.U, f32_sweep);

  const int nsum = kid == kBivariate ? 6 : 4;
  if (kid == kBivariate) {
    // theta = (s1, s2, beta, nu1, nu2, rho); rho~ = rho sqrt(nu1 nu2)/nu12
    // (the gamma ratios of the parsimonious constants, core_zcmg.c:1614-1655).
    BivariateGradParams B;
    B.x = ctx.l1.dx;
    B.y = ctx.l1.dy;
    B.n = ctx.N;
    B.metric = metric;
    B.s1 = theta[0];
    B.s2 = theta[1];
    B.beta = theta[2];
    B.nu1 = theta[3];
    B.nu2 = theta[4];
    B.nu12 = 0.5 * (theta[3] + theta[4]);
    B.rho_ratio = std::sqrt(B.nu1 * B.nu2) / B.nu12 * std::sqrt(B.s1 * B.s2);
    B.s12 = theta[5] * B.rho_ratio;
    B.con1 = host_matern_con(B.nu1);
    B.con2 = host_matern_con(B.nu2);
    B.con12 = host_matern_con(B.nu12);
    KnuOrder<Dual> hko[3];
    knu_prepare(B.nu1, hko[0]);
    knu_prepare(B.nu2, hko[1]);
    knu_prepare(B.nu12, hko[2]);
    if (!st.ko) EXA_HIP(hipMalloc(&st.ko, sizeof(hko)));
    EXA_HIP(hipMemcpyAsync(st.ko, hko, sizeof(hko), hipMemcpyHostToDevice, s));
    profiled("grad_reduce", 0.0, 4.0 * (double)ctx.N * (ctx.N + 1.0), s,
             [&] { launch_grad_bivariate(v, B, st.ko, st.alpha.d, st.partial, st.sums, kGroups, s); });
  } else {
    const MaternGradParams G = matern_grad_params(ctx, kid, theta, metric);
    profiled("grad_reduce", 0.0, 4.0 * (double)ctx.N * (ctx.N + 1.0), s,
             [&] { launch_grad_matern(v, G, st.alpha.d, st.partial, st.sums, kGroups, s); });
  }
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  reduce_sums(st, nsum, s, acc);
  if (Profiler::get().enabled()) Profiler::get().collect();
  for (int p = 0; p < np; ++p) grad[p] = -0.5 * acc[p];
}

double adjoint_loglik_grad(Context& ctx, KernelId kid, const double* theta, int metric, double* grad,
                           double* logdet_out, double* dotp_out, int* info) {
  if (kid != kUnivariate && kid != kNugget && kid != kBivariate) throw std::invalid_argument("unsupported kernel");
  const int np = kernel_num_params(kid);
  const double t0 = wall_seconds();
  int inf = 0;
  double logdet = 0.0, dotp = 0.0;
  const double ll = ctx.loglik(kid, theta, metric, &logdet, &dotp, &inf);
  PhaseTimes lt = ctx.last;
  if (info) *info = inf;
  if (logdet_out) *logdet_out = logdet;
  if (dotp_out) *dotp_out = dotp;
  if (inf != 0) {
    for (int p = 0; p < np; ++p) grad[p] = std::nan("");
    return ll;
  }
  gradient_from_factor(ctx, kid, theta, metric, ctx.fp64("the gradient and the Fisher information"), false, grad);
  const double t1 = wall_seconds();
  ctx.last = lt;
  ctx.last.grad = t1 - t0 - lt.total;
  ctx.last.total = t1 - t0;
  return ll;
}

double mixed_adjoint_loglik_grad(Context& ctx, KernelId kid, const double* theta, int metric,
                                 const MixedOptions& opt, double* grad, double* logdet_out, double* dotp_out,
                                 int* info, MixedReport* rep) {
  if (kid != kUnivariate && kid != kNugget && kid != kBivariate) throw std::invalid_argument("unsupported kernel");
  const int np = kernel_num_params(kid);
  TileMatrix& A = ctx.fp64("the mixed-precision gradient");
  const double t0 = wall_seconds();
  int inf = 0;
  double logdet = 0.0, dotp = 0.0;
  const double ll = mixed_loglik(ctx, kid, theta, metric, opt, &logdet, &dotp, &inf, rep);
  PhaseTimes lt = ctx.last;
  if (info) *info = inf;
  if (logdet_out) *logdet_out = logdet;
  if (dotp_out) *dotp_out = dotp;
  if (inf != 0) {
    for (int p = 0; p < np; ++p) grad[p] = std::nan("");
    return ll;
  }
  // The FP64 Sigma has served the refinement; it now receives the mixed
  // factor, whose Takahashi sweep gives the trace term.  With opt.grad_f32
  // the sweep's column stage runs on the FP32 matrix cores: 0.74 against
  // 1.01 s for l + gradient at n = 40,000 with a 0.01 nugget, the gradient
  // as accurate as with the FP64 sweep there; on a worse-conditioned Sigma
  // (the parsimonious bivariate kernel at 50,000 locations) its rounding
  // error is 3.7e-3 of the largest component against 1.8e-4 from the mixed
  // factor alone.  The swept inverse is then checked along z against the
  // triangular solves with the factor it inverted; with opt.grad_f32 == 2 a
  // residual above kSweepResidualMax recomputes the gradient with the FP64
  // sweep.  ctx.work holds the refined Sigma^-1 z.
  mixed_factor_to_fp64(ctx, A, ctx.s1);
  if (opt.grad_f32) mixed_probe_solve(ctx, A, ctx.s1);
  gradient_from_factor(ctx, kid, theta, metric, A, true, grad, opt.grad_f32 != 0);
  if (opt.grad_f32) {
    const double resid = mixed_probe_residual(ctx, A, ctx.s1);
    if (opt.grad_f32 == 2 && !(resid <= kSweepResidualMax)) {
      MixedOptions o64 = opt;
      o64.grad_f32 = 0;
      const double ll64 = mixed_adjoint_loglik_grad(ctx, kid, theta, metric, o64, grad, logdet_out, dotp_out, info, rep);
      if (rep) rep->sweep_residual = resid;
      ctx.last.total = wall_seconds() - t0;  // both attempts
      return ll64;
    }
    if (rep) rep->sweep_residual = resid;
  }
  const double t1 = wall_seconds();
  ctx.last = lt;
  ctx.last.grad = t1 - t0 - lt.total;
  ctx.last.total = t1 - t0;
  return ll;
}

// Fisher information (reference: EXAGEOSTAT_Fisher_Tile, exageostat_exact/
// src/compute/MLE_dexact.c:1777-1934, which forms Sigma^-1 dSigma_j by two
// triangular solves per parameter on stored derivative matrices and the
// trace of each product).  Here Z = Sigma^-1 comes from the Takahashi sweep,
// M_j = Z D_j Z from two tile products per parameter, and <M_j, D_k> for
// every k from one pass of the fused derivative reduction, which
// regenerates D_k entry by entry.
int fisher_information(Context& ctx, KernelId kid, const double* theta, int metric, double* F) {
  if (kid != kUnivariate && kid != kNugget)
    throw std::invalid_argument("Fisher information: univariate Matern kernels only");
  const int np = kernel_num_params(kid);
  int inf = 0;
  double logdet = 0.0, dotp = 0.0;
  ctx.loglik(kid, theta, metric, &logdet, &dotp, &inf);
  if (inf != 0) return inf;
  AdjointState& st = adjoint_buffers(ctx);
  TileMatrix& A = ctx.fp64("the gradient and the Fisher information");
  hipStream_t s = ctx.s1;
  if (!st.D.allocated() || st.D.mt != A.mt || st.D.ts != A.ts) {
    st.D.release();
    st.M.release();
    st.D.allocate(A.m, A.n, A.ts, true);
    st.M.allocate(A.m, A.n, A.ts, true);
  }
  if (st.y_tiles < (long)A.mt * A.mt) {
    if (st.Y) (void)hipFree(st.Y);
    st.Y = nullptr;
    EXA_HIP(hipMalloc(&st.Y, sizeof(double) * (size_t)A.ts * A.ts * A.mt * A.mt));
    st.y_tiles = (long)A.mt * A.mt;
  }
  st.zero.allocate(ctx.N, ctx.work.padded);
  tile_potri_inplace(A, ctx.ws, st.P, st.Q, s, st.W, kPotriSplit, &st.U);
  const MaternGradParams G = matern_grad_params(ctx, kid, theta, metric);
  const double n = (double)ctx.N;
  for (int j = 0; j < np; ++j) {
    profiled("fisher_gen", 0.0, 4.0 * n * (n + 1.0), s, [&] { launch_gen_dmatern(st.D.view(), G, j, s); });
    profiled("fisher_zdz", 3.0 * n * n * n, 0.0, s,
             [&] { launch_fisher_zdz(A.view(), st.D.view(), st.Y, st.M.view(), s); });
    profiled("fisher_reduce", 0.0, 4.0 * n * (n + 1.0), s,
             [&] { launch_grad_matern(st.M.view(), G, st.zero.d, st.partial, st.sums, kGroups, s); });
    double acc[8];
    reduce_sums(st, 4, s, acc);
    // row j from pass j for k >= j, as the reference fills its lower triangle
    for (int k = j; k < np; ++k) F[j * np + k] = F[k * np + j] = 0.5 * acc[k];
  }
  if (Profiler::get().enabled()) Profiler::get().collect();
  return 0;
}

}  // namespace exa
