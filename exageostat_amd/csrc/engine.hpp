// Host-side engine: device tile matrices, the single-GPU tile Cholesky
// scheduler and the likelihood / prediction drivers.  Everything here is
// C++; the C entry points of the reference live in capi.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

namespace exa {

// Kernel families selected by the reference's kernel strings
// (runtime/starpu/codelets/codelet_zcmg.c:262-310).
enum KernelId {
  kUnivariate = 0,       // "univariate_matern_stationary"
  kBivariate = 3,        // "bivariate_matern_parsimonious" (variables interleaved by row parity)
  kNugget = 4,           // "univariate_matern_nuggets_stationary"
  kBivariateBlock = 5,   // "bivariate_matern_parsimonious2" (variable 1 rows, then variable 2 rows)
  kSpaceTime = 6,        // "univariate_spacetime_matern_stationary" (location z = time)
  kUnknownKernel = -1,
};
KernelId kernel_from_name(const char* name);
int kernel_num_params(KernelId k);
// Rows of the covariance per location (2 for the bivariate kernels).
inline int kernel_vars(KernelId k) { return (k == kBivariate || k == kBivariateBlock) ? 2 : 1; }
// The "_profile" variants of the bivariate kernels: the same covariance,
// the profile likelihood (MLE_dexact.c:403-421).
bool kernel_is_profile(const char* name);

// A tiled matrix in device memory.  Tiles are ts x ts, column-major; all
// padding rows/columns are part of the tiles.
class TileMatrix {
 public:
  TileMatrix() = default;
  ~TileMatrix();
  TileMatrix(const TileMatrix&) = delete;
  TileMatrix& operator=(const TileMatrix&) = delete;

  // prec: per-tile precision map (empty: all FP64), indexed like tile_index.
  // fp64_slots: every tile gets an FP64-sized slot, so set_prec() can later
  // change any tile's precision without reallocating.
  void allocate(int m, int n, int ts, bool lower, const std::vector<uint8_t>& prec = {}, bool fp64_slots = false);
  void set_prec(const std::vector<uint8_t>& prec);
  // Every tile's slot can hold the precision `prec` gives it.
  bool fits(const std::vector<uint8_t>& prec) const;
  const std::vector<uint8_t>& prec_map() const { return h_prec_; }
  void release();
  bool allocated() const { return pool_ != nullptr; }
  TileView view() const;
  long num_tiles() const { return (long)h_ptr_.size(); }
  double* tile(int m, int n) const;  // FP64 tiles only
  void* tile_raw(int m, int n) const;
  uint8_t prec(int m, int n) const;
  size_t bytes() const { return bytes_; }
  // Host <-> device copies of the logical m x n matrix (column-major, ld).
  void to_host(double* A, long ld, hipStream_t s) const;
  void from_host(const double* A, long ld, hipStream_t s);

  int m = 0, n = 0, ts = 0, mt = 0, nt = 0;
  bool lower = false;

 private:
  void* pool_ = nullptr;
  size_t bytes_ = 0;
  bool fp64_slots_ = false;
  std::vector<void*> h_ptr_;
  std::vector<uint8_t> h_prec_;
  std::vector<uint8_t> h_slot_;  // bytes per entry of each tile's slot
  void** d_ptr_ = nullptr;
  uint8_t* d_prec_ = nullptr;
};

// A device vector padded to mt*ts entries.
struct DeviceVector {
  double* d = nullptr;
  long n = 0, padded = 0;
  void allocate(long n_, long padded_);
  void release();
  ~DeviceVector() { release(); }
};

struct PhaseTimes {
  double gen = 0, facto = 0, logdet = 0, solve = 0, grad = 0, total = 0;
};

// Device copy of a location set (x, y).  upload() copies again only when
// the host pointers, the count or a fingerprint of the host values differ
// from the last upload, so unchanged locations stay resident across calls
// and changed ones (also at the same address) are sent again.
struct DeviceLocations {
  const double* hx = nullptr;
  const double* hy = nullptr;
  const double* hz = nullptr;
  long n = 0;
  uint64_t fp = 0;
  double* dx = nullptr;
  double* dy = nullptr;
  double* dz = nullptr;  // third coordinate (3-D locations, or time), or null
  void upload(const double* x, const double* y, long n_, hipStream_t s, const double* z = nullptr);
  void release();
  ~DeviceLocations() { release(); }
};

// Covariance generation for a kernel family and theta.
void generate_covariance(TileMatrix& A, KernelId kid, const double* theta, const DeviceLocations& rows,
                         const DeviceLocations& cols, int metric, hipStream_t s);
// The tiles p0 .. p0 + ntiles - 1 of v's enumeration (lower: p = m (m + 1) / 2
// + n, row by row), each stored in its tile's precision.
void generate_covariance_tiles(const TileView& v, KernelId kid, const double* theta, const DeviceLocations& rows,
                               const DeviceLocations& cols, int metric, long p0, long ntiles, hipStream_t s);

// Workspace of the tile Cholesky (inverses of 64x64 diagonal blocks, info).
class CholWorkspace {
 public:
  ~CholWorkspace();
  void ensure(int mt, int ts);
  double* dinv(int k) const { return dinv_ + (long)k * (ts_ / kNB) * 4096; }
  double* dinv_base() const { return dinv_; }
  // Inverse of diagonal tile k (ts x ts, lower), filled by tile_potrf.
  double* xinv(int k) const { return xinv_ + (long)k * ts_ * ts_; }
  double* xinv_base() const { return xinv_; }
  // Scratch vector of mt*ts doubles for the triangular solves.
  double* yvec() const { return yvec_; }
  // Partial sums ((ts/64) * ts doubles) and the zeroed completion ticket of
  // the diagonal-tile products in the solves.
  double* tpart() const { return tpart_; }
  unsigned* ticket() const { return ticket_; }
  int* info() const { return info_; }      // device
  int* host_info() const { return h_info_; }  // pinned host mirror
  hipEvent_t ev(int i);
  int mt_ = 0, ts_ = 0;
  // First tile of the matrix whose factorisation filled dinv/xinv last.
  const void* factored = nullptr;

 private:
  double* dinv_ = nullptr;
  double* xinv_ = nullptr;
  double* yvec_ = nullptr;
  double* tpart_ = nullptr;
  unsigned* ticket_ = nullptr;
  int* info_ = nullptr;
  int* h_info_ = nullptr;
  std::vector<hipEvent_t> ev_;
};

// Tile pairs (m, n) of a mixed-precision factor grouped by the precision
// class of tile (m, n): pairs[c] (device, packed m << 16 | n) sorted by
// column then row, off[c][n] = index of the first pair of column >= n,
// flops[c] = prefix sums of the pairs' algorithmic update flops per step.
// split[2]: BF16 parts of the panels of even / odd steps (owned by the
// caller, mt tiles of 2 ts^2 BF16 values each).
struct ClassLists {
  int* d_pairs[4] = {nullptr, nullptr, nullptr, nullptr};
  std::vector<long> off[4];
  std::vector<double> flops[4];
  uint16_t* split[2] = {nullptr, nullptr};
  bool lolo = false;  // kF32 updates with four split-BF16 products
  ~ClassLists();
  // Rebuilds the lists unless A's grid, precision map and band are the ones
  // they were last built for.  band > 0 keeps only the pairs m - n < band.
  void build(const TileMatrix& A, int band = 0);

 private:
  long cap[4] = {0, 0, 0, 0};  // device list capacities (pairs)
  std::vector<uint8_t> built_for;
  int built_mt = -1, built_m = -1, built_ts = -1, built_band = -1;
};

// Right-looking tile Cholesky with one-step look-ahead on two streams.
// Returns 0 or the first failing global column + 1 (the matrix is then
// not positive definite).  Synchronises s_main at the end.
// With lists, the trailing updates run per precision class (mixed factor).
// band > 0 (with lists built for that band): the band-limited
// factorisation of CHAM_pdpotrf_diag (exageostat_approx/src/compute/
// pdpotrf_diag.c:64-110): step k solves only the panel tiles m < k + band
// and updates only the columns n < k + band, tiles m < n + band; tiles
// outside the band keep the covariance.
int tile_potrf(TileMatrix& A, CholWorkspace& ws, hipStream_t s_main, hipStream_t s_aux, bool sync = true,
               const ClassLists* lists = nullptr, int band = 0);
// z <- L^-1 z (trans=0) or L^-T z (trans=1) with the factor from tile_potrf.
void tile_trsv(const TileMatrix& L, const CholWorkspace& ws, double* z, int trans, hipStream_t s);
// The same with caller-provided scratch: y (mt*ts doubles), tpart
// ((ts/64)*ts doubles) and ticket (zero on entry; left zero).
void tile_trsv_scratch(const TileMatrix& L, const CholWorkspace& ws, double* z, int trans, hipStream_t s, double* y,
                       double* tpart, unsigned* ticket);
// sum_i log L_ii (deterministic order).
double tile_logdet(const TileMatrix& L, double* d_partial, hipStream_t s);
double device_dot(const double* a, const double* b, long n, double* d_partial, hipStream_t s);

struct MixedState;
void destroy_mixed(MixedState* m);
struct AdjointState;
void destroy_adjoint(AdjointState* a);
struct MloeState;
void destroy_mloe(MloeState* m);

// Everything one MLE_data needs on the device.
class Context {
 public:
  // lean: no FP64 Sigma is allocated; only the mixed-precision and
  // single-precision evaluations (which generate straight into their own
  // factor and regenerate Sigma's tiles for the FP64 refinement) and the
  // prediction and MLOE paths (own matrices) run on it.
  explicit Context(int N, int ts, int nZobs, int nZmiss, bool lean = false);
  ~Context();
  hipStream_t s1 = nullptr, s2 = nullptr;
  int N, ts;
  int mt;                      // tile rows of Sigma
  bool lean = false;
  TileMatrix C;                // Sigma (lower); unallocated when lean
  DeviceVector Z, Zcpy, work;  // observations, pristine copy, solve vector
  DeviceLocations l1;
  CholWorkspace ws;
  double* d_partial = nullptr;  // reduction partial sums
  long partial_len = 0;
  // prediction
  TileMatrix C22, C12;
  DeviceVector Zobs, Zmiss, Zactual, gemv_part;
  DeviceLocations lobs, lmiss;
  PhaseTimes last;
  MixedState* mixed = nullptr;  // mixed-precision factor, created on first use
  AdjointState* adjoint = nullptr;  // gradient buffers, created on first use
  MloeState* mloe = nullptr;        // MLOE / MMOM buffers, created on first use
  ClassLists band_lists;            // update pairs of the band-limited factorisation
  // The observed locations as a context of their own, for the kriging of
  // the single- and mixed-precision paths (their factor of C22 and the
  // refined C22^-1 z_obs); created on first use.
  Context* pred = nullptr;

  // band > 0: the band-limited factorisation (computation "diag_approx");
  // the log-determinant and the solve then read the whole lower triangle,
  // whose tiles outside the band still hold the covariance.
  double loglik(KernelId kid, const double* theta, int metric, double* logdet_out, double* dotp_out, int* info,
                int band = 0);
  void ensure_partial(long n);
  // C, or std::invalid_argument naming `what` when the context is lean.
  TileMatrix& fp64(const char* what);
};

double wall_seconds();

}  // namespace exa
