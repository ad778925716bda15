// Host engine: device tile matrices, covariance generation dispatch, the
// single-GPU tile Cholesky scheduler and the likelihood driver.
#include "engine.hpp"

#include <algorithm>

#include "profile.hpp"

#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

namespace exa {

double wall_seconds() {
  using clk = std::chrono::steady_clock;
  return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

KernelId kernel_from_name(const char* name) {
  if (!name) return kUnknownKernel;
  if (!strcmp(name, "univariate_matern_stationary")) return kUnivariate;
  if (!strcmp(name, "univariate_matern_nuggets_stationary")) return kNugget;
  if (!strcmp(name, "bivariate_matern_parsimonious") || !strcmp(name, "bivariate_matern_parsimonious_profile"))
    return kBivariate;
  if (!strcmp(name, "bivariate_matern_parsimonious2") || !strcmp(name, "bivariate_matern_parsimonious2_profile"))
    return kBivariateBlock;
  if (!strcmp(name, "univariate_spacetime_matern_stationary")) return kSpaceTime;
  return kUnknownKernel;
}

bool kernel_is_profile(const char* name) {
  return name && (!strcmp(name, "bivariate_matern_parsimonious_profile") ||
                  !strcmp(name, "bivariate_matern_parsimonious2_profile"));
}

int kernel_num_params(KernelId k) {
  switch (k) {
    case kUnivariate: return 3;
    case kNugget: return 4;
    case kBivariate: return 6;
    case kBivariateBlock: return 6;
    case kSpaceTime: return 7;
    default: return 0;
  }
}

static size_t prec_size(uint8_t p) { return p == kF64 ? 8 : ((p == kF32 || p == kF32B) ? 4 : 2); }

// ------------------------------------------------------------- TileMatrix --
TileMatrix::~TileMatrix() { release(); }

void TileMatrix::release() {
  if (pool_) (void)hipFree(pool_);
  if (d_ptr_) (void)hipFree(d_ptr_);
  if (d_prec_) (void)hipFree(d_prec_);
  pool_ = nullptr;
  d_ptr_ = nullptr;
  d_prec_ = nullptr;
  h_ptr_.clear();
  h_prec_.clear();
  h_slot_.clear();
  bytes_ = 0;
}

bool TileMatrix::fits(const std::vector<uint8_t>& prec) const {
  if (!pool_ || prec.size() != h_slot_.size()) return false;
  for (size_t t = 0; t < prec.size(); ++t)
    if (prec_size(prec[t]) > h_slot_[t]) return false;
  return true;
}

void TileMatrix::allocate(int m_, int n_, int ts_, bool lower_, const std::vector<uint8_t>& prec,
                          bool fp64_slots) {
  if (ts_ <= 0 || ts_ % kNB != 0 || ts_ > 1024)
    throw std::invalid_argument("tile size must be a multiple of 64 and at most 1024");
  if (lower_ && m_ != n_) throw std::invalid_argument("lower-packed matrices must be square");
  release();
  m = m_;
  n = n_;
  ts = ts_;
  lower = lower_;
  mt = (m + ts - 1) / ts;
  nt = (n + ts - 1) / ts;
  const long ntiles = lower ? (long)mt * (mt + 1) / 2 : (long)mt * nt;
  h_prec_.assign(ntiles, (uint8_t)kF64);
  if (!prec.empty()) {
    if ((long)prec.size() != ntiles) throw std::invalid_argument("precision map size mismatch");
    h_prec_ = prec;
  }
  std::vector<size_t> off(ntiles);
  h_slot_.resize(ntiles);
  size_t total = 0;
  for (long t = 0; t < ntiles; ++t) {
    off[t] = total;
    h_slot_[t] = (uint8_t)(fp64_slots ? 8 : prec_size(h_prec_[t]));
    total += ((size_t)ts * ts * h_slot_[t] + 255) & ~(size_t)255;
  }
  EXA_HIP(hipMalloc(&pool_, total));
  bytes_ = total;
  h_ptr_.resize(ntiles);
  for (long t = 0; t < ntiles; ++t) h_ptr_[t] = static_cast<char*>(pool_) + off[t];
  EXA_HIP(hipMalloc(&d_ptr_, sizeof(void*) * ntiles));
  EXA_HIP(hipMemcpy(d_ptr_, h_ptr_.data(), sizeof(void*) * ntiles, hipMemcpyHostToDevice));
  EXA_HIP(hipMalloc(&d_prec_, ntiles));
  EXA_HIP(hipMemcpy(d_prec_, h_prec_.data(), ntiles, hipMemcpyHostToDevice));
  fp64_slots_ = fp64_slots;
}

void TileMatrix::set_prec(const std::vector<uint8_t>& prec) {
  if (prec.size() != h_prec_.size()) throw std::invalid_argument("precision map size mismatch");
  if (!fits(prec)) throw std::invalid_argument("tile slot too small");
  if (prec == h_prec_) return;  // unchanged map: the device table is current
  h_prec_ = prec;
  // The table is read by kernels of earlier launches on the non-blocking
  // streams; the null-stream copy below does not wait for them.
  EXA_HIP(hipDeviceSynchronize());
  EXA_HIP(hipMemcpy(d_prec_, h_prec_.data(), h_prec_.size(), hipMemcpyHostToDevice));
}

TileView TileMatrix::view() const {
  TileView v;
  v.ptr = d_ptr_;
  v.prec = d_prec_;
  v.m = m;
  v.n = n;
  v.mt = mt;
  v.nt = nt;
  v.ts = ts;
  v.lower = lower ? 1 : 0;
  return v;
}

void* TileMatrix::tile_raw(int mm, int nn) const {
  TileView v = view();
  return h_ptr_[tile_index(v, mm, nn)];
}

double* TileMatrix::tile(int mm, int nn) const { return static_cast<double*>(tile_raw(mm, nn)); }

uint8_t TileMatrix::prec(int mm, int nn) const {
  TileView v = view();
  return h_prec_[tile_index(v, mm, nn)];
}

void TileMatrix::to_host(double* A, long ld, hipStream_t s) const {
  std::vector<double> buf((size_t)ts * ts);
  std::vector<float> fbuf((size_t)ts * ts);
  std::vector<uint16_t> hbuf((size_t)ts * ts);
  for (int tn = 0; tn < nt; ++tn) {
    for (int tm = lower ? tn : 0; tm < mt; ++tm) {
      const uint8_t p = prec(tm, tn);
      if (p == kF64) {
        EXA_HIP(hipMemcpyAsync(buf.data(), tile_raw(tm, tn), buf.size() * 8, hipMemcpyDeviceToHost, s));
      } else if (p == kF32 || p == kF32B) {
        EXA_HIP(hipMemcpyAsync(fbuf.data(), tile_raw(tm, tn), fbuf.size() * 4, hipMemcpyDeviceToHost, s));
      } else {
        EXA_HIP(hipMemcpyAsync(hbuf.data(), tile_raw(tm, tn), hbuf.size() * 2, hipMemcpyDeviceToHost, s));
      }
      EXA_HIP(hipStreamSynchronize(s));
      for (int j = 0; j < ts && (long)tn * ts + j < n; ++j)
        for (int i = 0; i < ts && (long)tm * ts + i < m; ++i) {
          double val;
          const size_t e = i + (size_t)j * ts;
          if (p == kF64) {
            val = buf[e];
          } else if (p == kF32 || p == kF32B) {
            val = fbuf[e];
          } else {
            uint32_t u = (uint32_t)hbuf[e] << 16;
            float f;
            memcpy(&f, &u, 4);
            val = f;
          }
          A[((long)tm * ts + i) + ((long)tn * ts + j) * ld] = val;
        }
    }
  }
}

void TileMatrix::from_host(const double* A, long ld, hipStream_t s) {
  std::vector<double> buf((size_t)ts * ts);
  for (int tn = 0; tn < nt; ++tn) {
    for (int tm = lower ? tn : 0; tm < mt; ++tm) {
      if (prec(tm, tn) != kF64) throw std::invalid_argument("from_host supports FP64 tiles only");
      for (int j = 0; j < ts; ++j)
        for (int i = 0; i < ts; ++i) {
          const long gi = (long)tm * ts + i, gj = (long)tn * ts + j;
          buf[i + (size_t)j * ts] = (gi < m && gj < n) ? A[gi + gj * ld] : (gi == gj ? 1.0 : 0.0);
        }
      EXA_HIP(hipMemcpyAsync(tile_raw(tm, tn), buf.data(), buf.size() * 8, hipMemcpyHostToDevice, s));
      EXA_HIP(hipStreamSynchronize(s));
    }
  }
}

// ---------------------------------------------------------- small helpers --
void DeviceVector::allocate(long n_, long padded_) {
  if (d && padded == padded_) {
    n = n_;
    return;
  }
  release();
  n = n_;
  padded = padded_;
  EXA_HIP(hipMalloc(&d, sizeof(double) * padded));
  // hipMemset runs on the null stream, which does not order with the
  // library's non-blocking streams: wait for it, or a copy queued next on
  // another stream can land before the zeros.
  EXA_HIP(hipMemset(d, 0, sizeof(double) * padded));
  EXA_HIP(hipDeviceSynchronize());
}

void DeviceVector::release() {
  if (d) (void)hipFree(d);
  d = nullptr;
  n = padded = 0;
}

// FNV-1a over the 64-bit patterns of the coordinates.
static uint64_t location_fingerprint(const double* x, const double* y, long n) {
  uint64_t h = 1469598103934665603ull;
  for (long i = 0; i < n; ++i) {
    uint64_t a, b;
    std::memcpy(&a, x + i, sizeof(a));
    std::memcpy(&b, y + i, sizeof(b));
    h = (h ^ a) * 1099511628211ull;
    h = (h ^ b) * 1099511628211ull;
  }
  return h;
}

void DeviceLocations::upload(const double* x, const double* y, long n_, hipStream_t s, const double* z) {
  uint64_t f = location_fingerprint(x, y, n_);
  if (z) f = (f ^ location_fingerprint(z, z, n_)) * 1099511628211ull;
  if (x == hx && y == hy && z == hz && n_ == n && dx && f == fp) return;
  if (n_ != n || !dx) {
    release();
    n = n_;
    EXA_HIP(hipMalloc(&dx, sizeof(double) * (n > 0 ? n : 1)));
    EXA_HIP(hipMalloc(&dy, sizeof(double) * (n > 0 ? n : 1)));
  }
  if (z && !dz) EXA_HIP(hipMalloc(&dz, sizeof(double) * (n > 0 ? n : 1)));
  if (!z && dz) {
    (void)hipFree(dz);
    dz = nullptr;
  }
  hx = x;
  hy = y;
  hz = z;
  fp = f;
  EXA_HIP(hipMemcpyAsync(dx, x, sizeof(double) * n, hipMemcpyHostToDevice, s));
  EXA_HIP(hipMemcpyAsync(dy, y, sizeof(double) * n, hipMemcpyHostToDevice, s));
  if (z) EXA_HIP(hipMemcpyAsync(dz, z, sizeof(double) * n, hipMemcpyHostToDevice, s));
  EXA_HIP(hipStreamSynchronize(s));
}

void DeviceLocations::release() {
  if (dx) (void)hipFree(dx);
  if (dy) (void)hipFree(dy);
  if (dz) (void)hipFree(dz);
  dx = dy = dz = nullptr;
  hx = hy = hz = nullptr;
  n = 0;
  fp = 0;
}

// ------------------------------------------------------------- generation --
void generate_covariance(TileMatrix& A, KernelId kid, const double* theta, const DeviceLocations& rows,
                         const DeviceLocations& cols, int metric, hipStream_t s) {
  if (!A.allocated()) throw std::invalid_argument("covariance generation into an unallocated tile matrix");
  generate_covariance_tiles(A.view(), kid, theta, rows, cols, metric, 0, A.num_tiles(), s);
}

void generate_covariance_tiles(const TileView& v, KernelId kid, const double* theta, const DeviceLocations& rows,
                               const DeviceLocations& cols, int metric, long p0, long ntiles, hipStream_t s) {
  const bool three_d = rows.dz && cols.dz && kid != kSpaceTime;
  if (three_d && metric == 1)
    throw std::invalid_argument("great-circle distances are 2-D only (locations carry a z coordinate)");
  if (kid == kUnivariate || kid == kNugget) {
    MaternParams P;
    P.x1 = rows.dx;
    P.y1 = rows.dy;
    P.x2 = cols.dx;
    P.y2 = cols.dy;
    P.z1 = three_d ? rows.dz : nullptr;
    P.z2 = three_d ? cols.dz : nullptr;
    P.metric = metric;
    P.sigma2 = theta[0];
    P.beta = theta[1];
    P.nugget = (kid == kNugget) ? theta[3] : 0.0;
    P.con = 1.0 / (std::pow(2.0, theta[2] - 1.0) * std::tgamma(theta[2]));
    knu_prepare(theta[2], P.ko);
    launch_gen_matern(v, P, p0, ntiles, s);
  } else if (kid == kBivariate || kid == kBivariateBlock) {
    if (kid == kBivariate && v.ts % 2 != 0) throw std::invalid_argument("bivariate kernel needs an even tile size");
    BivariateParams P;
    P.x1 = rows.dx;
    P.y1 = rows.dy;
    P.x2 = cols.dx;
    P.y2 = cols.dy;
    P.z1 = three_d ? rows.dz : nullptr;
    P.z2 = three_d ? cols.dz : nullptr;
    P.nloc1 = rows.n;
    P.nloc2 = cols.n;
    P.metric = metric;
    const double s1 = theta[0], s2 = theta[1], nu1 = theta[3], nu2 = theta[4], rho = theta[5];
    P.beta = theta[2];
    const double nu12 = 0.5 * (nu1 + nu2);
    const double rho_t = rho * std::sqrt((std::tgamma(nu1 + 1) * std::tgamma(nu2 + 1)) /
                                         (std::tgamma(nu1) * std::tgamma(nu2))) *
                         std::tgamma(nu12) / std::tgamma(nu12 + 1);
    P.sigma11 = s1;
    P.sigma22 = s2;
    P.sigma12 = rho_t * std::sqrt(s1 * s2);
    P.con11 = 1.0 / (std::pow(2.0, nu1 - 1.0) * std::tgamma(nu1));
    P.con22 = 1.0 / (std::pow(2.0, nu2 - 1.0) * std::tgamma(nu2));
    P.con12 = 1.0 / (std::pow(2.0, nu12 - 1.0) * std::tgamma(nu12));
    knu_prepare(nu1, P.ko11);
    knu_prepare(nu2, P.ko22);
    knu_prepare(nu12, P.ko12);
    launch_gen_bivariate(v, P, p0, ntiles, kid == kBivariateBlock, s);
  } else if (kid == kSpaceTime) {
    if (!rows.dz || !cols.dz) throw std::invalid_argument("space-time kernel: locations need z (time)");
    SpaceTimeParams P;
    P.x1 = rows.dx;
    P.y1 = rows.dy;
    P.t1 = rows.dz;
    P.x2 = cols.dx;
    P.y2 = cols.dy;
    P.t2 = cols.dz;
    P.metric = metric;
    P.sigma2 = theta[0];
    P.beta_s = theta[1];
    P.a = theta[3];
    P.alpha = theta[4];
    P.beta = theta[5];
    P.delta = theta[6];
    P.con = 1.0 / (std::pow(2.0, theta[2] - 1.0) * std::tgamma(theta[2]));
    knu_prepare(theta[2], P.ko);
    launch_gen_spacetime(v, P, p0, ntiles, s);
  } else {
    throw std::invalid_argument("unsupported kernel");
  }
}

// --------------------------------------------------------------- Cholesky --
CholWorkspace::~CholWorkspace() {
  if (dinv_) (void)hipFree(dinv_);
  if (xinv_) (void)hipFree(xinv_);
  if (yvec_) (void)hipFree(yvec_);
  if (tpart_) (void)hipFree(tpart_);
  if (ticket_) (void)hipFree(ticket_);
  if (info_) (void)hipFree(info_);
  if (h_info_) (void)hipHostFree(h_info_);
  for (auto e : ev_) (void)hipEventDestroy(e);
}

void CholWorkspace::ensure(int mt, int ts) {
  if (dinv_ && mt <= mt_ && ts == ts_) return;
  if (dinv_) (void)hipFree(dinv_);
  if (xinv_) (void)hipFree(xinv_);
  if (yvec_) (void)hipFree(yvec_);
  if (tpart_) (void)hipFree(tpart_);
  dinv_ = xinv_ = yvec_ = tpart_ = nullptr;
  mt_ = mt;
  ts_ = ts;
  EXA_HIP(hipMalloc(&dinv_, sizeof(double) * (size_t)mt * (ts / kNB) * 4096));
  EXA_HIP(hipMalloc(&xinv_, sizeof(double) * (size_t)mt * ts * ts));
  EXA_HIP(hipMalloc(&yvec_, sizeof(double) * (size_t)mt * ts));
  EXA_HIP(hipMalloc(&tpart_, sizeof(double) * (size_t)(ts / kNB) * ts));
  if (!ticket_) {
    EXA_HIP(hipMalloc(&ticket_, sizeof(unsigned)));
    // synchronous on the null stream, which does not order with the
    // non-blocking streams the kernels run on: wait for it here
    EXA_HIP(hipMemset(ticket_, 0, sizeof(unsigned)));
    EXA_HIP(hipDeviceSynchronize());
  }
  if (!info_) EXA_HIP(hipMalloc(&info_, sizeof(int)));
  if (!h_info_) EXA_HIP(hipHostMalloc(&h_info_, sizeof(int), hipHostMallocDefault));
}

hipEvent_t CholWorkspace::ev(int i) {
  while ((int)ev_.size() <= i) {
    hipEvent_t e;
    EXA_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    ev_.push_back(e);
  }
  return ev_[i];
}

// Algorithmic flops of A(m,n) -= A(m,k) A(n,k)^T over n in [lo, hi],
// m in [n, mt) on the logical (unpadded) sizes: 2 r_m r_n K per
// off-diagonal tile, r_n (r_n + 1) K per diagonal tile (FLOPS_DSYRK).
static double update_flops(const TileMatrix& A, int k, int lo, int hi) {
  auto rows = [&](int t) { return (double)std::min<long>(A.ts, (long)A.m - (long)t * A.ts); };
  const double K = rows(k);
  double f = 0.0;
  for (int n = lo; n <= hi && n < A.mt; ++n) {
    const double rn = rows(n);
    f += rn * (rn + 1.0) * K;
    double below = 0.0;
    for (int m = n + 1; m < A.mt; ++m) below += rows(m);
    f += 2.0 * below * rn * K;
  }
  return f;
}

static double potrf_flops(double n) {
  return n * (((1.0 / 6.0) * n + 0.5) * n + 1.0 / 3.0) + n * (((1.0 / 6.0) * n) * n - 1.0 / 6.0);
}

ClassLists::~ClassLists() {
  for (auto p : d_pairs)
    if (p) (void)hipFree(p);
}

void ClassLists::build(const TileMatrix& A, int band) {
  // The lists depend only on the tile grid, the precision map and the band:
  // an evaluation whose map equals the previous one reuses them.
  if (built_mt == A.mt && built_m == A.m && built_ts == A.ts && built_band == band && built_for == A.prec_map())
    return;
  const int mt = A.mt;
  auto rows = [&](int t) { return (double)std::min<long>(A.ts, (long)A.m - (long)t * A.ts); };
  std::vector<int> pairs[4];
  for (int c = 0; c < 4; ++c) {
    off[c].assign(mt + 2, 0);
    flops[c].assign(1, 0.0);
  }
  for (int n = 0; n < mt; ++n) {
    for (int c = 0; c < 4; ++c) off[c][n] = (long)pairs[c].size();
    for (int m = n; m < mt && (band <= 0 || m - n < band); ++m) {
      const int c = A.prec(m, n);
      pairs[c].push_back((m << 16) | n);
      const double f = (m == n) ? rows(n) * (rows(n) + 1.0) : 2.0 * rows(m) * rows(n);
      flops[c].push_back(flops[c].back() + f);  // times K at use
    }
  }
  for (int c = 0; c < 4; ++c) {
    off[c][mt] = off[c][mt + 1] = (long)pairs[c].size();
    if (pairs[c].empty()) continue;
    // Device lists keep their high-water size across maps.
    if ((long)pairs[c].size() > cap[c]) {
      if (d_pairs[c]) (void)hipFree(d_pairs[c]);
      d_pairs[c] = nullptr;
      cap[c] = 0;
      EXA_HIP(hipMalloc(&d_pairs[c], sizeof(int) * pairs[c].size()));
      cap[c] = (long)pairs[c].size();
    }
    EXA_HIP(hipMemcpy(d_pairs[c], pairs[c].data(), sizeof(int) * pairs[c].size(), hipMemcpyHostToDevice));
  }
  built_for = A.prec_map();
  built_mt = A.mt;
  built_m = A.m;
  built_ts = A.ts;
  built_band = band;
}

int tile_potrf(TileMatrix& A, CholWorkspace& ws, hipStream_t s1, hipStream_t s2, bool sync,
               const ClassLists* lists, int band) {
  const int mt = A.mt;
  const int ts = A.ts;
  if (band > 0 && !lists) throw std::invalid_argument("band-limited factorisation needs band lists");
  if (band >= mt) band = 0;  // the band covers the whole matrix
  // Panel rows of step k: m < row_end(k); updated columns of step k:
  // n <= col_end(k).
  auto row_end = [&](int k) { return band > 0 ? std::min(mt, k + band) : mt; };
  auto col_end = [&](int k) { return band > 0 ? std::min(mt - 1, k + band - 1) : mt - 1; };
  ws.ensure(mt, ts);
  int* d_info = ws.info();
  EXA_HIP(hipMemsetAsync(d_info, 0, sizeof(int), s1));
  TileView v = A.view();
  auto rows = [&](int t) { return (double)std::min<long>(ts, (long)A.m - (long)t * ts); };
  auto factor_panel = [&](int k) {
    const double rk = rows(k);
    profiled("potrf_diag", potrf_flops(rk), 0.0, s1,
             [&] { launch_potrf_tile(A.tile(k, k), ts, ws.dinv(k), d_info, (long)k * ts, s1); });
    const double below = std::min((double)A.m, (double)row_end(k) * ts) - (double)(k + 1) * ts;
    if (k + 1 < row_end(k))
      profiled("trsm_panel", below * rk * rk, 0.0, s1,
               [&] { launch_trsm_panel(v, k, ws.dinv(k), s1, row_end(k)); });
    // The BF16 update classes read the panel in split BF16 form.  The buffer
    // of step k is reused at step k + 2, whose panel (on s1) waits for the
    // column update of step k + 1, queued on s2 after the rest of step k.
    if (lists && lists->split[k & 1] && k + 1 < mt)
      profiled("split_panel", 0.0, 12.0 * below * A.ts, s1,
               [&] { launch_split_panel(v, k, lists->split[k & 1], s1); });
  };
  auto update = [&](int k, int lo, int hi, hipStream_t s) {
    if (!lists) {
      profiled("tile_update", update_flops(A, k, lo, hi), 0.0, s, [&] { launch_tile_update(v, k, 1, lo, hi, s); });
      return;
    }
    // One fused launch for the three classes; each class's share of the
    // algorithmic flops is counted under its own name as well.
    static const char* names[4] = {"mixed_flops_f64", "mixed_flops_f32", "", "mixed_flops_bf16"};
    FusedSpans sp;
    double fl = 0.0;
    int i = 0;
    // Each XCD walks the classes in this order: the split-BF16 pairs first,
    // then FP64, then the far tiles.  At n = 90,000 this order takes 5% less
    // update time than walking FP64 first (895 against 946 ms per
    // evaluation).
    const int order[3] = {(int)kF32, (int)kF64, (int)kF32B};
    for (int c : order) {
      sp.cls[i] = c == kF64 ? 0 : (c == kF32 ? 1 : 2);
      const long beg = lists->off[c][lo];
      const long end = lists->off[c][std::min(hi + 1, mt)];
      sp.pairs[i] = (lists->d_pairs[c] && end > beg) ? lists->d_pairs[c] + beg : nullptr;
      sp.nblocks[i] = end > beg ? end - beg : 0;
      const double f = (lists->flops[c][end] - lists->flops[c][beg]) * rows(k);
      if (f > 0) Profiler::get().count(names[c], f, 0.0);
      fl += f;
      ++i;
    }
    if (sp.nblocks[0] + sp.nblocks[1] + sp.nblocks[2] == 0) return;
    profiled("tile_update_mixed", fl, 0.0, s,
             [&] { launch_update_fused(v, k, sp, lists->split[k & 1], lists->lolo, s); });
  };
  int nev = 0;
  auto record = [&](hipStream_t st) {
    hipEvent_t e = ws.ev(nev++);
    EXA_HIP(hipEventRecord(e, st));
    return e;
  };
  // s1 carries the panel factorisations (small kernels on the CUs the
  // persistent update leaves free); s2 carries the trailing updates: first
  // the column that unblocks the next panel, then the rest, which overlaps
  // with that panel on s1.
  EXA_HIP(hipStreamWaitEvent(s2, record(s1), 0));
  factor_panel(0);
  hipEvent_t ev_panel = record(s1);
  if (lists) {
    for (int k = 0; k + 1 < mt; ++k) {
      EXA_HIP(hipStreamWaitEvent(s2, ev_panel, 0));
      if (k + 1 <= col_end(k)) update(k, k + 1, k + 1, s2);
      hipEvent_t ev_col = record(s2);
      if (k + 2 <= col_end(k)) update(k, k + 2, col_end(k), s2);
      EXA_HIP(hipStreamWaitEvent(s1, ev_col, 0));
      factor_panel(k + 1);
      ev_panel = record(s1);
    }
  } else {
    // FP64 factor: the trailing matrix is updated with two panels at once
    // (K = 2 ts per block, one read and write of C per two steps).  For the
    // pair (k, k+1):
    //   s2: column k+1 with panel k;            s1: panel k+1
    //   s2: the deferred part of the previous pair's trailing update
    //   s2: column k+2 with panels k, k+1;      s1: panel k+2
    //   s2: column k+3, then the first part of columns >= k+4 with panels
    //       k, k+1; the rest is deferred to run during panel k+3.
    // Updates of a column commute; each column receives all of its updates
    // on s2 before its panel starts on s1, and every panel overlaps update
    // work on s2.
    auto update2 = [&](int kk, int lo, int hi) {
      profiled("tile_update", update_flops(A, kk, lo, hi) + update_flops(A, kk + 1, lo, hi), 0.0, s2,
               [&] { launch_tile_update(v, kk, 2, lo, hi, s2); });
    };
    struct Deferred {
      int k = -1, lo = 0, hi = -1;
    } pend;
    auto run_pending = [&] {
      if (pend.k >= 0 && pend.lo <= pend.hi) update2(pend.k, pend.lo, pend.hi);
      pend.k = -1;
    };
    // columns [lo, hi] split into two parts of about equal update work
    // (column c of the trailing matrix has mt - c tiles)
    auto split_cols = [&](int lo, int hi) {
      double total = 0.0;
      for (int c = lo; c <= hi; ++c) total += mt - c;
      double acc = 0.0;
      int mid = lo;
      for (; mid < hi; ++mid) {
        acc += mt - mid;
        if (acc >= 0.5 * total) break;
      }
      return mid;
    };
    int k = 0;
    while (k + 1 < mt) {
      EXA_HIP(hipStreamWaitEvent(s2, ev_panel, 0));
      update(k, k + 1, k + 1, s2);
      EXA_HIP(hipStreamWaitEvent(s1, record(s2), 0));
      factor_panel(k + 1);
      ev_panel = record(s1);
      run_pending();
      if (k + 2 >= mt) break;
      EXA_HIP(hipStreamWaitEvent(s2, ev_panel, 0));
      update2(k, k + 2, k + 2);
      hipEvent_t ev_col = record(s2);
      EXA_HIP(hipStreamWaitEvent(s1, ev_col, 0));
      factor_panel(k + 2);
      ev_panel = record(s1);
      if (k + 3 < mt) update2(k, k + 3, k + 3);
      if (k + 4 < mt) {
        const int mid = split_cols(k + 4, mt - 1);
        update2(k, k + 4, mid);
        pend.k = k;
        pend.lo = mid + 1;
        pend.hi = mt - 1;
      }
      k += 2;
    }
    run_pending();
  }
  // Join: s1 waits for the last updates on s2, s2 for the last panel on s1.
  EXA_HIP(hipStreamWaitEvent(s1, record(s2), 0));
  EXA_HIP(hipStreamWaitEvent(s2, record(s1), 0));
  launch_zero_upper_diag(v, s1);
  profiled("trtri_diag", (double)mt * ts * (double)ts * ts / 3.0, 0.0, s1,
           [&] { launch_trtri_tiles(v, ws.dinv_base(), ws.xinv_base(), s1); });
  EXA_HIP(hipMemcpyAsync(ws.host_info(), d_info, sizeof(int), hipMemcpyDeviceToHost, s1));
  ws.factored = A.tile_raw(0, 0);
  if (!sync) return 0;
  EXA_HIP(hipStreamSynchronize(s1));
  return *ws.host_info();
}

void tile_trsv(const TileMatrix& L, const CholWorkspace& ws, double* z, int trans, hipStream_t s) {
  tile_trsv_scratch(L, ws, z, trans, s, ws.yvec(), ws.tpart(), ws.ticket());
}

void tile_trsv_scratch(const TileMatrix& L, const CholWorkspace& ws, double* z, int trans, hipStream_t s, double* y,
                       double* tpart, unsigned* ticket) {
  // Blocked substitution with the inverted diagonal tiles: y_k = X_k z_k
  // (resp. X_k^T z_k), then the off-diagonal tiles update the remaining
  // right-hand side; y is assembled in a scratch vector and copied back.
  TileView v = L.view();
  const int ts = L.ts;
  if (!trans) {
    for (int k = 0; k < L.mt; ++k) {
      launch_trmv_diag(ws.xinv(k), ts, z + (long)k * ts, y + (long)k * ts, 0, tpart, ticket, s);
      launch_gemv_panel(v, k, y + (long)k * ts, z, s);
    }
  } else {
    for (int k = L.mt - 1; k >= 0; --k) {
      launch_trmv_diag(ws.xinv(k), ts, z + (long)k * ts, y + (long)k * ts, 1, tpart, ticket, s);
      launch_gemv_t_row(v, k, y + (long)k * ts, z, s);
    }
  }
  EXA_HIP(hipMemcpyAsync(z, y, sizeof(double) * (size_t)L.mt * ts, hipMemcpyDeviceToDevice, s));
}

double tile_logdet(const TileMatrix& L, double* d_partial, hipStream_t s) {
  launch_logdet(L.view(), d_partial, s);
  std::vector<double> h(L.mt);
  EXA_HIP(hipMemcpyAsync(h.data(), d_partial, sizeof(double) * L.mt, hipMemcpyDeviceToHost, s));
  EXA_HIP(hipStreamSynchronize(s));
  double acc = 0.0;
  for (double x : h) acc += x;
  return acc;
}

double device_dot(const double* a, const double* b, long n, double* d_partial, hipStream_t s) {
  const long nb = (n + 4095) / 4096;
  launch_dot(a, b, n, d_partial, s);
  std::vector<double> h(nb);
  EXA_HIP(hipMemcpyAsync(h.data(), d_partial, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
  EXA_HIP(hipStreamSynchronize(s));
  double acc = 0.0;
  for (double x : h) acc += x;
  return acc;
}

// ---------------------------------------------------------------- Context --
Context::Context(int N_, int ts_, int nZobs, int nZmiss, bool lean_) : N(N_), ts(ts_), lean(lean_) {
  (void)nZobs;
  (void)nZmiss;
  if (N <= 0) throw std::invalid_argument("matrix size must be positive");
  if (ts <= 0 || ts % kNB != 0 || ts > 1024)
    throw std::invalid_argument("tile size must be a multiple of 64 and at most 1024");
  mt = (int)(((long)N + ts - 1) / ts);
  EXA_HIP(hipStreamCreateWithFlags(&s1, hipStreamNonBlocking));
  EXA_HIP(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
  if (!lean) C.allocate(N, N, ts, true);
  const long padded = (long)mt * ts;
  Z.allocate(N, padded);
  Zcpy.allocate(N, padded);
  work.allocate(N, padded);
  ws.ensure(mt, ts);
  ensure_partial(padded / 4096 + mt + 16);
}

TileMatrix& Context::fp64(const char* what) {
  if (lean)
    throw std::invalid_argument(std::string(what) +
                                " needs the FP64 covariance, which a lean context (mp_lean = 1) does not hold");
  return C;
}

Context::~Context() {
  delete pred;
  pred = nullptr;
  destroy_mixed(mixed);
  destroy_adjoint(adjoint);
  destroy_mloe(mloe);
  mloe = nullptr;
  mixed = nullptr;
  if (d_partial) (void)hipFree(d_partial);
  if (s1) (void)hipStreamDestroy(s1);
  if (s2) (void)hipStreamDestroy(s2);
}

void Context::ensure_partial(long n) {
  if (n <= partial_len) return;
  if (d_partial) (void)hipFree(d_partial);
  EXA_HIP(hipMalloc(&d_partial, sizeof(double) * n));
  partial_len = n;
}

double Context::loglik(KernelId kid, const double* theta, int metric, double* logdet_out, double* dotp_out,
                       int* info, int band) {
  if (band < 0) throw std::invalid_argument("diag_thick must be positive");
  fp64("the FP64 likelihood");
  const bool banded = band > 0 && band < C.mt;
  const double t0 = wall_seconds();
  EXA_HIP(hipMemcpyAsync(work.d, Zcpy.d, sizeof(double) * work.padded, hipMemcpyDeviceToDevice, s1));
  profiled("gen", 0.0, 4.0 * (double)N * (N + 1.0), s1,
           [&] { generate_covariance(C, kid, theta, l1, l1, metric, s1); });
  EXA_HIP(hipStreamSynchronize(s1));
  const double t1 = wall_seconds();
  if (banded) band_lists.build(C, band);
  const int inf = banded ? tile_potrf(C, ws, s1, s2, true, &band_lists, band) : tile_potrf(C, ws, s1, s2, true);
  const double t2 = wall_seconds();
  if (info) *info = inf;
  const double logdet = 2.0 * tile_logdet(C, d_partial, s1);
  const double t3 = wall_seconds();
  profiled("trsv", (double)N * (N + 1.0), 4.0 * (double)N * (N + 1.0), s1,
           [&] { tile_trsv(C, ws, work.d, 0, s1); });
  const double dotp = device_dot(work.d, work.d, N, d_partial, s1);
  const double t4 = wall_seconds();
  if (Profiler::get().enabled()) Profiler::get().collect();
  last.gen = t1 - t0;
  last.facto = t2 - t1;
  last.logdet = t3 - t2;
  last.solve = t4 - t3;
  last.total = t4 - t0;
  if (logdet_out) *logdet_out = logdet;
  if (dotp_out) *dotp_out = dotp;
  return -0.5 * dotp - 0.5 * logdet - (N / 2.0) * std::log(2.0 * M_PI);
}

}  // namespace exa
