// Mixed-precision log-likelihood.
//
// The reference's mixed tier (MLE_sdmle in exageostat_exact/src/compute/
// MLE_sdexact.c:44-224, chameleon_psdpotrf in psdportf.c:44-187) factors
// Sigma with a band of FP64 tiles around the diagonal and FP32 tiles
// elsewhere, and takes the solve and the determinant straight from that
// factor.  Here:
//   * the precision of every off-diagonal tile is chosen per theta from the
//     largest correlation the tile can hold (bounding boxes of its row and
//     column location sets): strongly correlated tiles stay FP64, weakly
//     correlated ones become FP32, and with bf16 on the weakest become FP32
//     storage with BF16 MFMA updates;
//   * the solve Sigma^-1 z is refined in FP64 (x += L~^-T L~^-1 (z - Sigma
//     x)), so z^T Sigma^-1 z reaches FP64 accuracy whatever the off-diagonal
//     precision.  Sigma x reads the FP64 Sigma, or, in a lean context (no
//     FP64 Sigma), regenerates Sigma's tiles in FP64 a chunk at a time into
//     a small pool and multiplies them there: the same values and the same
//     summation order, so both give the same result;
//   * log det Sigma comes from the FP64 diagonal tiles of the mixed factor.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "mixed.hpp"
#include "profile.hpp"

namespace exa {

struct MixedState {
  TileMatrix Cm;       // mixed factor, each tile in a slot of its precision
  ClassLists lists;    // update pairs per precision class
  uint16_t* split = nullptr;  // two panels in split BF16 form
  long split_tiles = 0;
  DeviceVector x, r;   // solution and correction
  double* part = nullptr;  // partial sums of Sigma x: 2 * mt columns of mt*ts
  long part_len = 0;
  // Per-tile bounding boxes of the location set (xmin, xmax, ymin, ymax),
  // cached for one host location array and its fingerprint.
  std::vector<double> bbox;
  uint64_t bbox_fp = 0;
  const double* bbox_x = nullptr;
  const double* bbox_y = nullptr;
  long bbox_n = 0;
  int bbox_ts = 0;
  int bbox_layout = -1;
  // Lean contexts: FP64 tiles of Sigma regenerated for Sigma x, `regen_slots`
  // tiles at a time; regen_ptr maps every lower tile to its slot (tile p of
  // the row-by-row enumeration to slot p mod regen_slots).
  double* regen = nullptr;
  void** regen_ptr = nullptr;
  long regen_slots = 0;
  int regen_mt = 0, regen_ts = 0;
  ~MixedState() {
    if (part) (void)hipFree(part);
    if (split) (void)hipFree(split);
    if (regen) (void)hipFree(regen);
    if (regen_ptr) (void)hipFree(regen_ptr);
  }
};

// Tiles regenerated per chunk in a lean context: enough work-groups for the
// product (one per tile) to fill the GPU; 5.7 GB at ts = 960.
static constexpr long kRegenSlots = 768;

void destroy_mixed(MixedState* m) { delete m; }

// Row -> location layouts of the covariance rows.
enum RowLayout { kRowPlain = 0, kRowInterleaved = 1, kRowBlock = 2 };

// Location of covariance row i: i (one variable), i / 2 (two variables
// interleaved by row parity), i mod nloc (variable 1 rows, then variable 2).
static inline long row_location(long i, int layout, long nloc) {
  return layout == kRowInterleaved ? i / 2 : (layout == kRowBlock ? (i < nloc ? i : i - nloc) : i);
}

// Per-tile bounding boxes: tile t covers rows [t*ts, (t+1)*ts).
static void tile_bboxes(MixedState& st, const double* x, const double* y, uint64_t fp, long nloc, int layout,
                        long nrows, int ts, int mt) {
  if (st.bbox_x == x && st.bbox_y == y && st.bbox_fp == fp && st.bbox_n == nloc && st.bbox_ts == ts &&
      st.bbox_layout == layout && (long)st.bbox.size() == 4L * mt)
    return;
  st.bbox.assign(4L * mt, 0.0);
  for (int t = 0; t < mt; ++t) {
    double x0 = HUGE_VAL, x1 = -HUGE_VAL, y0 = HUGE_VAL, y1 = -HUGE_VAL;
    const long lo = (long)t * ts, hi = std::min<long>(nrows, lo + ts);
    for (long i = lo; i < hi; ++i) {
      const long p = row_location(i, layout, nloc);
      x0 = std::min(x0, x[p]);
      x1 = std::max(x1, x[p]);
      y0 = std::min(y0, y[p]);
      y1 = std::max(y1, y[p]);
    }
    st.bbox[4L * t + 0] = x0;
    st.bbox[4L * t + 1] = x1;
    st.bbox[4L * t + 2] = y0;
    st.bbox[4L * t + 3] = y1;
  }
  st.bbox_x = x;
  st.bbox_y = y;
  st.bbox_fp = fp;
  st.bbox_n = nloc;
  st.bbox_ts = ts;
  st.bbox_layout = layout;
}

// Largest correlation any entry of the Matern kernel with smoothness nu and
// range beta can take at distance >= d.
static double corr_bound(double d, double beta, const std::vector<KnuOrder<double>>& ko,
                         const std::vector<double>& con) {
  if (d <= 0.0) return 1.0;
  double c = 0.0;
  for (size_t i = 0; i < ko.size(); ++i) c = std::max(c, matern_corr(ko[i], con[i], d / beta));
  return std::min(c, 1.0);
}

// Precision class of every tile of the lower factor for this theta.
static std::vector<uint8_t> precision_map(MixedState& st, const Context& ctx, KernelId kid, const double* theta,
                                          int metric, const MixedOptions& opt, int mt, int ts) {
  const bool two = kid == kBivariate || kid == kBivariateBlock;
  const int layout = kid == kBivariate ? kRowInterleaved : (kid == kBivariateBlock ? kRowBlock : kRowPlain);
  std::vector<KnuOrder<double>> ko;
  std::vector<double> con;
  std::vector<double> nus;
  double beta;
  if (two) {
    beta = theta[2];
    nus = {theta[3], theta[4], 0.5 * (theta[3] + theta[4])};
  } else {
    beta = theta[1];
    nus = {theta[2]};
  }
  for (double nu : nus) {
    KnuOrder<double> o;
    knu_prepare(nu, o);
    ko.push_back(o);
    con.push_back(1.0 / (std::pow(2.0, nu - 1.0) * std::tgamma(nu)));
  }
  // Box bounds need Euclidean 2-D distances and a kernel that decreases
  // with them; the space-time kernel also depends on the time lag, and
  // great-circle or 3-D distances have no box bound here.
  const bool boxes = metric == 0 && ctx.l1.hx && ctx.l1.hy && !ctx.l1.hz && kid != kSpaceTime;
  if (boxes) tile_bboxes(st, ctx.l1.hx, ctx.l1.hy, ctx.l1.fp, ctx.l1.n, layout, ctx.N, ts, mt);
  std::vector<uint8_t> prec((size_t)mt * (mt + 1) / 2);
  const int thick = std::max(1, opt.diag_thick);
  long idx = 0;
  for (int n = 0; n < mt; ++n) {
    for (int m = n; m < mt; ++m, ++idx) {
      uint8_t c = kF32;
      if (m - n < thick) {
        c = kF64;
      } else if (boxes) {
        const double* a = &st.bbox[4L * m];
        const double* b = &st.bbox[4L * n];
        const double dx = std::max(0.0, std::max(a[0] - b[1], b[0] - a[1]));
        const double dy = std::max(0.0, std::max(a[2] - b[3], b[2] - a[3]));
        const double corr = corr_bound(std::sqrt(dx * dx + dy * dy), beta, ko, con);
        if (opt.fp64_corr > 0.0 && corr >= opt.fp64_corr) c = kF64;
        else if (opt.bf16 && corr < opt.fp32_corr) c = kF32B;
      } else if (opt.fp64_corr > 0.0) {
        // No box bound: every off-band tile is treated as strongly
        // correlated.
        c = kF64;
      }
      prec[idx] = c;
    }
  }
  return prec;
}

// y = Sigma x for the FP64 lower Sigma held in ctx.C (full diagonal tiles).
static void symv(const TileMatrix& C, const double* x, double* y, double* part, hipStream_t s) {
  launch_symv(C.view(), x, y, part, s);
}

// y = Sigma x without a stored FP64 Sigma: the lower tiles are generated in
// FP64 into the slot pool a chunk of the row-by-row enumeration at a time,
// and each chunk's partial products are taken before the next chunk
// overwrites the slots (one stream, in order).  The generated values, the
// per-tile partials and their final sum are those of symv() on the stored
// Sigma.
static void symv_regen(Context& ctx, MixedState& st, KernelId kid, const double* theta, int metric, const double* x,
                       double* y, hipStream_t s) {
  const int mt = ctx.mt, ts = ctx.ts;
  const long ntiles = (long)mt * (mt + 1) / 2;
  const long tt = (long)ts * ts;
  if (!st.regen || st.regen_mt != mt || st.regen_ts != ts) {
    if (st.regen) (void)hipFree(st.regen);
    if (st.regen_ptr) (void)hipFree(st.regen_ptr);
    st.regen = nullptr;
    st.regen_ptr = nullptr;
    st.regen_slots = std::min(ntiles, kRegenSlots);
    EXA_HIP(hipMalloc(&st.regen, sizeof(double) * tt * st.regen_slots));
    EXA_HIP(hipMalloc(&st.regen_ptr, sizeof(void*) * ntiles));
    std::vector<void*> h(ntiles);
    long p = 0;
    for (int m = 0; m < mt; ++m)
      for (int n = 0; n <= m; ++n, ++p) {
        const long t = (long)n * mt - ((long)n * (n - 1)) / 2 + (m - n);  // lower-packed tile index
        h[t] = st.regen + (p % st.regen_slots) * tt;
      }
    EXA_HIP(hipMemcpy(st.regen_ptr, h.data(), sizeof(void*) * ntiles, hipMemcpyHostToDevice));
    st.regen_mt = mt;
    st.regen_ts = ts;
  }
  TileView v{};
  v.ptr = st.regen_ptr;
  v.prec = nullptr;
  v.m = v.n = ctx.N;
  v.mt = v.nt = mt;
  v.ts = ts;
  v.lower = 1;
  for (long p0 = 0; p0 < ntiles; p0 += st.regen_slots) {
    const long cnt = std::min(st.regen_slots, ntiles - p0);
    generate_covariance_tiles(v, kid, theta, ctx.l1, ctx.l1, metric, p0, cnt, s);
    launch_symv_tiles(v, x, st.part, p0, cnt, s);
  }
  launch_symv_reduce(v, st.part, y, s);
}

double mixed_loglik(Context& ctx, KernelId kid, const double* theta, int metric, const MixedOptions& opt,
                    double* logdet_out, double* dotp_out, int* info, MixedReport* rep) {
  if (opt.ir_steps < 0) throw std::invalid_argument("ir_steps must be >= 0");
  if (!(opt.ir_tol >= 0.0)) throw std::invalid_argument("ir_tol must be >= 0");
  if (!ctx.mixed) ctx.mixed = new MixedState();
  MixedState& st = *ctx.mixed;
  const int N = ctx.N, ts = ctx.ts, mt = ctx.mt;
  const long padded = (long)mt * ts;
  hipStream_t s1 = ctx.s1;

  const double t0 = wall_seconds();
  // Each tile's slot holds exactly its precision (FP32 tiles take 4 bytes
  // per entry); a map that needs a larger slot somewhere than the current
  // allocation gives it repacks the factor's pool.
  std::vector<uint8_t> map = precision_map(st, ctx, kid, theta, metric, opt, mt, ts);
  if (!st.Cm.fits(map) || st.Cm.m != N || st.Cm.ts != ts) {
    st.Cm.release();
    st.Cm.allocate(N, N, ts, true, map, false);
  }
  st.x.allocate(N, padded);
  st.r.allocate(N, padded);
  if (st.part_len < 2L * mt * padded) {
    if (st.part) (void)hipFree(st.part);
    st.part = nullptr;
    EXA_HIP(hipMalloc(&st.part, sizeof(double) * 2L * mt * padded));
    st.part_len = 2L * mt * padded;
  }
  if (st.split_tiles < mt) {
    if (st.split) (void)hipFree(st.split);
    st.split = nullptr;
    EXA_HIP(hipMalloc(&st.split, sizeof(uint16_t) * 4L * mt * ts * ts));
    st.split_tiles = mt;
  }
  st.Cm.set_prec(map);
  st.lists.build(st.Cm);
  st.lists.split[0] = st.split;
  st.lists.split[1] = st.split + 2L * mt * ts * ts;
  st.lists.lolo = opt.lolo;
  if (ctx.lean) {
    // straight into the mixed factor's slots, each tile in its precision
    // (the values the conversion below would give)
    profiled("gen", 0.0, 4.0 * (double)N * (N + 1.0), s1,
             [&] { generate_covariance(st.Cm, kid, theta, ctx.l1, ctx.l1, metric, s1); });
  } else {
    profiled("gen", 0.0, 4.0 * (double)N * (N + 1.0), s1,
             [&] { generate_covariance(ctx.C, kid, theta, ctx.l1, ctx.l1, metric, s1); });
    profiled("convert", 0.0, 0.0, s1,
             [&] { launch_convert_mixed(ctx.C.view(), st.Cm.view(), ctx.C.num_tiles(), s1); });
  }
  EXA_HIP(hipStreamSynchronize(s1));
  const double t1 = wall_seconds();
  const int inf = tile_potrf(st.Cm, ctx.ws, s1, ctx.s2, true, &st.lists);
  const double t2 = wall_seconds();
  if (info) *info = inf;
  const double logdet = 2.0 * tile_logdet(st.Cm, ctx.d_partial, s1);
  const double t3 = wall_seconds();

  // Solve Sigma x = z with FP64 refinement.  Padding entries of z, x and r
  // stay zero (the padded rows of the factor are the identity).
  const double* z = ctx.Zcpy.d;
  double* x = st.x.d;
  double* r = st.r.d;
  auto solve = [&](double* v) {
    profiled("trsv", (double)N * (N + 1.0), 4.0 * (double)N * (N + 1.0), s1,
             [&] { tile_trsv(st.Cm, ctx.ws, v, 0, s1); });
    profiled("trsv", (double)N * (N + 1.0), 4.0 * (double)N * (N + 1.0), s1,
             [&] { tile_trsv(st.Cm, ctx.ws, v, 1, s1); });
  };
  EXA_HIP(hipMemcpyAsync(x, z, sizeof(double) * padded, hipMemcpyDeviceToDevice, s1));
  MixedReport report;
  if (inf == 0) {
    double prev_change = 0.0;
    solve(x);
    for (int it = 0; it < opt.ir_steps; ++it) {
      if (ctx.lean)
        profiled("symv_regen", 2.0 * (double)N * N, 0.0, s1,
                 [&] { symv_regen(ctx, st, kid, theta, metric, x, r, s1); });
      else
        profiled("symv", 2.0 * (double)N * N, 8.0 * (double)N * (N + 1.0) / 2.0, s1,
                 [&] { symv(ctx.C, x, r, st.part, s1); });
      launch_axpby(1.0, z, -1.0, r, r, N, s1);
      solve(r);
      launch_axpby(1.0, x, 1.0, r, x, N, s1);
      ++report.ir_done;
      if (opt.ir_tol > 0.0) {
        // The step changed z^T x by z^T c (c = r, the correction just
        // added).  Stop once that change is below the tolerance, or once
        // the change still to come is: the refinement converges linearly,
        // so with the ratio rho = change / previous change (< 1) the
        // remaining changes sum to about change * rho / (1 - rho).
        const double dq = device_dot(z, r, N, ctx.d_partial, s1);
        const double q = device_dot(z, x, N, ctx.d_partial, s1);
        const double change = q != 0.0 ? std::fabs(dq) / std::fabs(q) : std::fabs(dq);
        report.ir_change = change;
        const double rho = prev_change > 0.0 ? change / prev_change : 1.0;
        report.ir_converged = change <= opt.ir_tol || (rho < 0.5 && change * rho / (1.0 - rho) <= opt.ir_tol);
        if (report.ir_converged) break;
        prev_change = change;
      }
    }
    if (!(opt.ir_tol > 0.0 && opt.ir_steps > 0)) report.ir_converged = true;
  }
  if (rep) *rep = report;
  const double dotp = device_dot(z, x, N, ctx.d_partial, s1);
  // ctx.work receives Sigma^-1 z (the FP64 path leaves L^-1 z there).
  EXA_HIP(hipMemcpyAsync(ctx.work.d, x, sizeof(double) * padded, hipMemcpyDeviceToDevice, s1));
  EXA_HIP(hipStreamSynchronize(s1));
  const double t4 = wall_seconds();
  if (Profiler::get().enabled()) Profiler::get().collect();
  ctx.last.gen = t1 - t0;
  ctx.last.facto = t2 - t1;
  ctx.last.logdet = t3 - t2;
  ctx.last.solve = t4 - t3;
  ctx.last.total = t4 - t0;
  if (logdet_out) *logdet_out = logdet;
  if (dotp_out) *dotp_out = dotp;
  return -0.5 * dotp - 0.5 * logdet - (N / 2.0) * std::log(2.0 * M_PI);
}

void mixed_factor_to_fp64(Context& ctx, TileMatrix& dst, hipStream_t s) {
  if (!ctx.mixed || !ctx.mixed->Cm.allocated() || ctx.ws.factored != ctx.mixed->Cm.tile_raw(0, 0))
    throw std::invalid_argument("no mixed-precision factor to widen");
  const TileMatrix& src = ctx.mixed->Cm;
  if (!dst.allocated() || dst.mt != src.mt || dst.ts != src.ts || !dst.lower)
    throw std::invalid_argument("the FP64 matrix must have the mixed factor's tile grid");
  launch_mixed_to_fp64(src.view(), dst.view(), src.num_tiles(), s);
  ctx.ws.factored = dst.tile_raw(0, 0);
}

// x0 = F^-T F^-1 z with F the widened mixed factor (before its Takahashi
// sweep), kept in the refinement's correction vector.
void mixed_probe_solve(Context& ctx, TileMatrix& F, hipStream_t s) {
  MixedState& st = *ctx.mixed;
  EXA_HIP(hipMemcpyAsync(st.r.d, ctx.Zcpy.d, sizeof(double) * ctx.work.padded, hipMemcpyDeviceToDevice, s));
  tile_trsv(F, ctx.ws, st.r.d, 0, s);
  tile_trsv(F, ctx.ws, st.r.d, 1, s);
}

// ||Z z - x0|| / ||x0||, Z the swept inverse in the lower tiles of Zinv and
// x0 from mixed_probe_solve: the sweep's error along z, measured against the
// triangular solves with the factor it inverted.
double mixed_probe_residual(Context& ctx, TileMatrix& Zinv, hipStream_t s) {
  MixedState& st = *ctx.mixed;
  const long N = ctx.N;
  symv(Zinv, ctx.Zcpy.d, st.x.d, st.part, s);
  launch_axpby(1.0, st.x.d, -1.0, st.r.d, st.x.d, N, s);
  const double dd = device_dot(st.x.d, st.x.d, N, ctx.d_partial, s);
  const double xx = device_dot(st.r.d, st.r.d, N, ctx.d_partial, s);
  return xx > 0.0 ? std::sqrt(dd / xx) : 0.0;
}

// Tile precision map of the last mixed evaluation (lower-packed order);
// returns the number of tiles or 0 when no mixed factor exists.
long mixed_prec_map(const Context& ctx, uint8_t* out, long cap) {
  if (!ctx.mixed || !ctx.mixed->Cm.allocated()) return 0;
  const auto& p = ctx.mixed->Cm.prec_map();
  if (out) std::copy(p.begin(), p.begin() + std::min<long>(cap, (long)p.size()), out);
  return (long)p.size();
}

}  // namespace exa
