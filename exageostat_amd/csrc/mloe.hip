// MLOE / MMOM driver (exageostat_exact/src/compute/MLE_dexact.c:1112-1435,
// reduction runtime/starpu/codelets/codelet_zmloe_mmom.c:22-60).
//
// The reference loops over the held-out locations with vector solves.
// Here all of them are right-hand sides of one solve (rows of R), and the
// order of the work lets one n x n matrix serve both thetas:
//   K <- K_a, factor;  R <- C_a = Sigma_a(miss, obs)
//   U_a = rows of R mapped by L_a^-1        E4' = |u_a|^2 = k_a^T K_a^-1 k_a
//   W_a = rows of U_a mapped by L_a^-T      (w_a = K_a^-1 k_a)
//   R <- C_t                                E1' = k_t^T w_a
//   K <- K_t (unfactored); Y = W_a K_t      E2' = w_a^T K_t w_a
//   factor K_t; U_t = rows of C_t by L_t^-1 E3' = |u_t|^2 = k_t^T K_t^-1 k_t
#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>
#include <vector>

#include "kernels.hpp"
#include "mloe.hpp"
#include "profile.hpp"

namespace exa {

struct MloeState {
  TileMatrix K;       // Sigma(obs, obs), lower tiles
  TileMatrix R, U, W;  // nmiss x nobs, rows are right-hand sides
  double* part = nullptr;
  long part_len = 0;
  DeviceVector e;     // 4 x (row tiles * ts) dot products
  ~MloeState() {
    if (part) (void)hipFree(part);
  }
};

void destroy_mloe(MloeState* m) { delete m; }

static MloeState& mloe_buffers(Context& ctx, int nobs, int nmiss) {
  if (!ctx.mloe) ctx.mloe = new MloeState();
  MloeState& st = *ctx.mloe;
  const int ts = ctx.ts;
  if (!st.K.allocated() || st.K.m != nobs || st.K.ts != ts) {
    st.K.release();
    st.K.allocate(nobs, nobs, ts, true);
  }
  if (!st.R.allocated() || st.R.m != nmiss || st.R.n != nobs || st.R.ts != ts) {
    st.R.release();
    st.U.release();
    st.W.release();
    st.R.allocate(nmiss, nobs, ts, false);
    st.U.allocate(nmiss, nobs, ts, false);
    st.W.allocate(nmiss, nobs, ts, false);
  }
  const long rows = (long)st.R.mt * ts;
  const long need = (long)st.R.nt * rows;
  if (st.part_len < need) {
    if (st.part) (void)hipFree(st.part);
    st.part = nullptr;
    EXA_HIP(hipMalloc(&st.part, sizeof(double) * need));
    st.part_len = need;
  }
  st.e.allocate(4 * rows, 4 * rows);
  ctx.ws.ensure(st.K.mt, ts);
  return st;
}

int mloe_mmom(Context& ctx, KernelId kid, const double* theta_true, const double* theta_est, int metric, int nobs,
              int nmiss, double* loe, double* mom, double* mloe_out, double* mmom_out) {
  if (kid != kUnivariate && kid != kNugget)
    throw std::invalid_argument("MLOE/MMOM: univariate Matern kernels only");
  if (nobs <= 0 || nmiss <= 0) throw std::invalid_argument("MLOE/MMOM: empty location set");
  if (ctx.lobs.n != nobs || ctx.lmiss.n != nmiss)
    throw std::invalid_argument("MLOE/MMOM: location sets not uploaded");
  MloeState& st = mloe_buffers(ctx, nobs, nmiss);
  hipStream_t s = ctx.s1;
  const long rows = (long)st.R.mt * ctx.ts;
  double* e1 = st.e.d;
  double* e2 = st.e.d + rows;
  double* e3 = st.e.d + 2 * rows;
  double* e4 = st.e.d + 3 * rows;
  const double lo = (double)nobs, lm = (double)nmiss;
  const double solve_flops = lm * lo * lo;
  const double t0 = wall_seconds();

  generate_covariance(st.K, kid, theta_est, ctx.lobs, ctx.lobs, metric, s);
  int info = tile_potrf(st.K, ctx.ws, s, ctx.s2, true);
  if (info != 0) return info;
  generate_covariance(st.R, kid, theta_est, ctx.lmiss, ctx.lobs, metric, s);
  profiled("mloe_solve", solve_flops, 0.0, s,
           [&] { launch_rhs_solve(st.K.view(), ctx.ws.xinv_base(), st.R.view(), st.U.view(), true, s); });
  launch_row_dot(st.U.view(), st.U.view(), st.part, e4, s);
  profiled("mloe_solve", solve_flops, 0.0, s,
           [&] { launch_rhs_solve(st.K.view(), ctx.ws.xinv_base(), st.U.view(), st.W.view(), false, s); });
  generate_covariance(st.R, kid, theta_true, ctx.lmiss, ctx.lobs, metric, s);
  launch_row_dot(st.R.view(), st.W.view(), st.part, e1, s);
  generate_covariance(st.K, kid, theta_true, ctx.lobs, ctx.lobs, metric, s);
  profiled("mloe_symm", 2.0 * solve_flops, 0.0, s,
           [&] { launch_rect_symm(st.W.view(), st.K.view(), st.U.view(), s); });
  launch_row_dot(st.W.view(), st.U.view(), st.part, e2, s);
  info = tile_potrf(st.K, ctx.ws, s, ctx.s2, true);
  if (info != 0) return info;
  profiled("mloe_solve", solve_flops, 0.0, s,
           [&] { launch_rhs_solve(st.K.view(), ctx.ws.xinv_base(), st.R.view(), st.U.view(), true, s); });
  launch_row_dot(st.U.view(), st.U.view(), st.part, e3, s);

  std::vector<double> h(4 * rows);
  EXA_HIP(hipMemcpyAsync(h.data(), st.e.d, sizeof(double) * 4 * rows, hipMemcpyDeviceToHost, s));
  EXA_HIP(hipStreamSynchronize(s));
  const double st_t = theta_true[0], st_a = theta_est[0];
  double sum_loe = 0.0, sum_mom = 0.0;
  for (int j = 0; j < nmiss; ++j) {
    const double expr2 = st_t - 2.0 * h[j] + h[rows + j];
    const double expr3 = st_t - h[2 * rows + j];
    const double expr4 = st_a - h[3 * rows + j];
    const double l = expr3 == 0.0 ? -1.0 : expr2 / expr3 - 1.0;
    const double m = expr2 == 0.0 ? -1.0 : expr4 / expr2 - 1.0;
    if (loe) loe[j] = l;
    if (mom) mom[j] = m;
    sum_loe += l;
    sum_mom += m;
  }
  *mloe_out = sum_loe / nmiss;
  *mmom_out = sum_mom / nmiss;
  ctx.last.total = wall_seconds() - t0;
  if (Profiler::get().enabled()) Profiler::get().collect();
  return 0;
}

}  // namespace exa
