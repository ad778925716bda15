// Kernels of the adjoint gradient: the inverse Z = Sigma^-1 from the tile
// Cholesky factor (block Takahashi recurrence, FP64 MFMA) and the fused
// regeneration of dSigma/dtheta reduced against Z - alpha alpha^T.
//
// The reference has no gradient: its optimiser is derivative-free (NLOPT
// BOBYQA over MLE_alg, src/compute/MLE_exact.c / examples/examples.c:
// 263-264), and its derivative tiles (core_dcmg_matern_dsigma_square,
// core_dcmg_matern_dbeta, core_dcmg_matern_dnu in exageostat_exact/
// cpu_core/compute/core_zcmg.c:315-488) are not wired to a gradient.
//
// Takahashi recurrence.  With Sigma = L L^T, Z L = L^-T is upper
// triangular, which for tile column k (from the last to the first) gives
//   P_jk = L_jk X_kk                       (X_kk = L_kk^-1, j > k; kept as P_jk^T)
//   Z_mk = - sum_{j>k} Z(m,j) P_jk         (m > k; Z(m,j) = Z_jm^T for j > m)
//   Z_kk = X_kk^T X_kk - sum_{j>k} Z_jk^T P_jk.
// Z overwrites L in place (lower tiles; diagonal tiles full).
#include "besselk.hpp"
#include "common.hpp"
#include "kernels.hpp"
#include "mfma.hpp"

namespace exa {

using C192 = GemmCfg<192, 192, 2, 4, 16>;
using C128 = GemmCfg<128, 128, 2, 2, 16>;
using CT64 = GemmCfg<64, 64, 2, 2, 16>;

// P_j = L_jk X_kk for j in (k, mt), stored transposed (tile j of P holds
// P_j^T, column-major), so the two consumers read it row-major.
template <class C>
__global__ void __launch_bounds__(C::NT) k_trmm_panel(TileView v, int k, const double* X, double* P) {
  __shared__ __attribute__((aligned(16))) double lds[C::LDS_DOUBLES];
  const int qm = v.ts / C::BM, qn = v.ts / C::BN;
  const long b = blockIdx.x;
  const int j = k + 1 + (int)(b / (qm * qn));
  const int rem = (int)(b % (qm * qn));
  const int bi = rem % qm, bj = rem / qm;
  const long ts = v.ts;
  // X_kk is lower triangular: rows below bj*BN only.
  const int k0 = bj * C::BN;
  const double* A = tile_f64(v, j, k) + bi * C::BM + (long)k0 * ts;
  const double* B = X + k0 + (long)bj * C::BN * ts;  // B(c, kk) = X[k0 + kk, bj*BN + c]
  dx4 acc[C::FM][C::FN];
  zero_acc<C>(acc);
  gemm_nt_acc<C, true>(acc, A, ts, B, ts, (int)(ts - k0), lds, 1.0);
  acc_store_t<C>(acc, P + (long)j * ts * ts + bj * C::BN + (long)bi * C::BM * ts, ts);
}

// Z_mk = - sum_{j>k} Z(m,j) P_j for m in (k, mt).  With nsplit > 1 the
// sum over j is cut into nsplit contiguous ranges; work unit (s, m, block)
// writes its partial sum of range s to slab s of W (below tiles of ts x ts
// each, tile m - k - 1 at W + (s * below + m - k - 1) ts^2) and k_sum_slabs
// adds the slabs into Z_mk in a fixed order.  More, smaller work units fill
// the GPU's CUs evenly when (mt - k - 1) * (ts / BM)^2 units alone would
// leave a partial last round.
// HAS_U: tile (j, m), j > m, of U holds Z_jm^T, written by the epilogues of
// the column stages before it is read, so every operand is read in its
// plain (row-contiguous) layout and each j step is the same GEMM on
// different pointers (one instance of the GEMM body: the kernel keeps its
// accumulators in registers).  Without U, Z_jm is read transposed.
// Persistent launch, as in the Cholesky update: work-group w runs on XCD
// w % nx and the XCDs take contiguous slices of the unit list (units of one
// range s and row m share their A tiles and every P_j of the range).
template <class C, bool HAS_U, bool F32>
__device__ __forceinline__ void takahashi_unit(const TileView& v, int k, const double* P, double* W, int nsplit,
                                               const TileView& U, long u, double* lds) {
  const int qm = v.ts / C::BM, qn = v.ts / C::BN;
  const long below = v.mt - k - 1;
  const long per_split = below * qm * qn;
  const int sp = (int)(u / per_split);
  const long b = u - (long)sp * per_split;
  const int m = k + 1 + (int)(b / (qm * qn));
  const int rem = (int)(b % (qm * qn));
  const int bi = rem % qm, bj = rem / qm;
  const long ts = v.ts;
  const int chunk = (int)((below + nsplit - 1) / nsplit);
  const int j_lo = k + 1 + sp * chunk;
  const int j_hi = j_lo + chunk < v.mt ? j_lo + chunk : v.mt;
  if constexpr (F32) {
    // FP32 matrix cores: operands rounded to FP32 in LDS, FP32 sums
    f32x4 acc[C::FM][C::FN];
    zero_acc32<C>(acc);
#pragma unroll 1
    for (int j = j_lo; j < j_hi; ++j) {
      const double* B = P + (long)j * ts * ts + bj * C::BN;
      const double* A = (j <= m ? tile_f64(v, m, j) : tile_f64(U, j, m)) + bi * C::BM;
      gemm_nt_acc32<C>(acc, A, ts, B, ts, (int)ts, reinterpret_cast<float*>(lds), -1.f);
    }
    double* out = nsplit > 1 ? W + ((long)sp * below + (m - k - 1)) * ts * ts : tile_f64(v, m, k);
    acc32_store<C>(acc, out + bi * C::BM + (long)bj * C::BN * ts, ts);
    if (nsplit == 1) acc32_store_t<C>(acc, tile_f64(U, m, k) + bj * C::BN + (long)bi * C::BM * ts, ts);
    return;
  }
  dx4 acc[C::FM][C::FN];
  zero_acc<C>(acc);
#pragma unroll 1
  for (int j = j_lo; j < j_hi; ++j) {
    const double* B = P + (long)j * ts * ts + bj * C::BN;  // B(c, kk) = P_j^T[bj*BN + c, kk]
    if constexpr (HAS_U) {
      // A(r, kk) = Z(m, j)[bi*BM + r, kk]: tile (m, j) of Z for j <= m, Z_jm^T from U for j > m
      const double* A = (j <= m ? tile_f64(v, m, j) : tile_f64(U, j, m)) + bi * C::BM;
      gemm_nt_ring<C>(acc, A, ts, B, ts, (int)ts, lds, -1.0);
    } else if (j <= m) {
      gemm_nt_acc<C>(acc, tile_f64(v, m, j) + bi * C::BM, ts, B, ts, (int)ts, lds, -1.0);
    } else {
      const double* A = tile_f64(v, j, m) + (long)bi * C::BM * ts;  // A(r, kk) = Z_jm[kk, bi*BM + r]
      gemm_nt_acc<C, false, double, double, true>(acc, A, ts, B, ts, (int)ts, lds, -1.0);
    }
  }
  double* out = nsplit > 1 ? W + ((long)sp * below + (m - k - 1)) * ts * ts : tile_f64(v, m, k);
  acc_update<C>(acc, out + bi * C::BM + (long)bj * C::BN * ts, ts, 0.0);
  if constexpr (HAS_U)
    if (nsplit == 1) acc_store_t<C>(acc, tile_f64(U, m, k) + bj * C::BN + (long)bi * C::BM * ts, ts);
}

template <class C, bool HAS_U, bool F32 = false>
__global__ void __launch_bounds__(C::NT) k_takahashi_col(TileView v, int k, const double* P, double* W, int nsplit,
                                                         TileView U, long units, int nx) {
  static_assert(!F32 || HAS_U, "the FP32 column stage reads the transposed copy");
  __shared__ __attribute__((aligned(16))) double lds[F32 ? C::LDS_DOUBLES / 2
                                                         : (HAS_U ? RingLds<C>::DOUBLES : C::LDS_DOUBLES)];
  const int w = blockIdx.x;
  const int x = w % nx;
  const int gx = gridDim.x / nx;
  const long per = (units + nx - 1) / nx;
  const long start = (long)x * per;
  const long end = start + per < units ? start + per : units;
#pragma unroll 1
  for (long u = start + w / nx; u < end; u += gx) takahashi_unit<C, HAS_U, F32>(v, k, P, W, nsplit, U, u, lds);
}

// Z_mk = sum_s W[s][m - k - 1] for m in (k, mt), s ascending, and (with
// U) Z_mk^T into U's tile (m, k).  One work-group per 64 x 64 block: the sum
// is written column by column and the transpose goes through LDS so that
// both stores are contiguous.
__global__ void __launch_bounds__(256) k_sum_slabs(TileView v, int k, const double* W, int nsplit, TileView U) {
  __shared__ double blk[64][65];
  const int ts = v.ts;
  const int q = ts / 64;
  const long len = (long)ts * ts;
  const long below = v.mt - k - 1;
  const long b = blockIdx.x;
  const long t = b / (q * q);
  const int rem = (int)(b - t * q * q);
  const int r0 = (rem % q) * 64, c0 = (rem / q) * 64;
  const int lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
  double* Z = tile_f64(v, k + 1 + (int)t, k);
  for (int c = grp; c < 64; c += 4) {
    const long e = (long)(r0 + lane) + (long)(c0 + c) * ts;
    double acc = 0.0;
    for (int sp = 0; sp < nsplit; ++sp) acc += W[((long)sp * below + t) * len + e];
    Z[e] = acc;
    blk[lane][c] = acc;
  }
  if (!U.ptr) return;
  __syncthreads();
  double* Ut = tile_f64(U, k + 1 + (int)t, k);
  for (int r = grp; r < 64; r += 4) Ut[(long)(c0 + lane) + (long)(r0 + r) * ts] = blk[r][lane];
}

// Split-k terms of Z_kk: Q_0 = X_kk^T X_kk, Q_i = -Z_{k+i,k}^T P_{k+i}.
template <class C>
__global__ void __launch_bounds__(C::NT) k_takahashi_diag(TileView v, int k, const double* X, const double* P,
                                                          double* Q, TileView U) {
  __shared__ __attribute__((aligned(16))) double lds[C::LDS_DOUBLES];
  const int qm = v.ts / C::BM, qn = v.ts / C::BN;
  const long b = blockIdx.x;
  const int i = (int)(b / (qm * qn));
  const int rem = (int)(b % (qm * qn));
  const int bi = rem % qm, bj = rem / qm;
  const long ts = v.ts;
  dx4 acc[C::FM][C::FN];
  zero_acc<C>(acc);
  if (i == 0) {
    // X^T X: X(kk, r) = 0 for kk < r, so the sum starts at the larger block.
    const int k0 = (bi * C::BM > bj * C::BN ? bi * C::BM : bj * C::BN);
    gemm_nt_acc<C, true, double, double, true>(acc, X + k0 + (long)bi * C::BM * ts, ts,
                                               X + k0 + (long)bj * C::BN * ts, ts, (int)(ts - k0), lds, 1.0);
  } else if (U.ptr) {
    const int j = k + i;
    gemm_nt_acc<C>(acc, tile_f64(U, j, k) + bi * C::BM, ts, P + (long)j * ts * ts + bj * C::BN, ts, (int)ts, lds,
                   -1.0);
  } else {
    const int j = k + i;
    gemm_nt_acc<C, false, double, double, true>(acc, tile_f64(v, j, k) + (long)bi * C::BM * ts, ts,
                                                P + (long)j * ts * ts + bj * C::BN, ts, (int)ts, lds, -1.0);
  }
  acc_update<C>(acc, Q + (long)i * ts * ts + bi * C::BM + (long)bj * C::BN * ts, ts, 0.0);
}

// out[e] = sum_{i < cnt} Q[i*len + e] (fixed order), e < len.
__global__ void k_sum_tiles(const double* Q, int cnt, long len, double* out) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= len) return;
  double s = 0.0;
  for (int i = 0; i < cnt; ++i) s += Q[(long)i * len + e];
  out[e] = s;
}

// Number of j ranges of the column stage: the split (at most 4, and at most
// one tile per range, and only as many as the slab buffer holds) whose work
// units fill whole rounds of the concurrent work-groups best; each extra
// range costs a write and a read of its slab, counted as 2% per range.
template <class C, bool F32 = false>
static int takahashi_splits(const TileView& v, int k, int max_split, int* slots_out) {
  static int slots = 0;
  if (!slots) {
    int dev = 0, cus = 0, occ = 1;
    EXA_HIP(hipGetDevice(&dev));
    EXA_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    EXA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_takahashi_col<C, true, F32>, C::NT, 0));
    slots = cus * (occ > 0 ? occ : 1);
    if (slots >= 8) slots -= slots % 8;
  }
  const long per = (long)(v.ts / C::BM) * (v.ts / C::BN);
  const long below = v.mt - k - 1;
  int best = 1;
  double best_cost = 1e300;
  for (int sp = 1; sp <= 4 && sp <= below && sp <= max_split; ++sp) {
    const long units = below * per * sp;
    const double rounds = (double)((units + slots - 1) / slots);
    // time in units of one unsplit block's work
    const double cost = rounds / sp * (1.0 + 0.02 * (sp - 1));
    if (cost < best_cost - 1e-12) {
      best_cost = cost;
      best = sp;
    }
  }
  *slots_out = slots;
  return best;
}

template <class C>
static void takahashi_cfg(const TileView& v, int k, const double* X, double* P, double* Q, double* W, int max_split,
                          const TileView& U, int stage, hipStream_t s, bool f32) {
  const long per = (long)(v.ts / C::BM) * (v.ts / C::BN);
  const long below = v.mt - k - 1;
  if (stage == 0 && below > 0)
    hipLaunchKernelGGL(k_trmm_panel<C>, dim3((unsigned)(below * per)), dim3(C::NT), 0, s, v, k, X, P);
  if (stage == 1 && below > 0) {
    int slots = 0;
    const bool f32u = f32 && U.ptr;
    const int ns = f32u ? takahashi_splits<C, true>(v, k, W ? max_split : 1, &slots)
                        : takahashi_splits<C>(v, k, W ? max_split : 1, &slots);
    const long units = below * per * ns;
    // persistent grid: at most one round of resident work-groups, a
    // multiple of the 8 XCDs when it is that large
    long grid = units < slots ? units : slots;
    int nx = 8;
    if (grid >= 8) grid -= grid % 8;
    else nx = 1;
    if (f32u)
      hipLaunchKernelGGL((k_takahashi_col<C, true, true>), dim3((unsigned)grid), dim3(C::NT), 0, s, v, k,
                         (const double*)P, W, ns, U, units, nx);
    else if (U.ptr)
      hipLaunchKernelGGL((k_takahashi_col<C, true>), dim3((unsigned)grid), dim3(C::NT), 0, s, v, k,
                         (const double*)P, W, ns, U, units, nx);
    else
      hipLaunchKernelGGL((k_takahashi_col<C, false>), dim3((unsigned)grid), dim3(C::NT), 0, s, v, k,
                         (const double*)P, W, ns, U, units, nx);
    if (ns > 1) {
      const long q = v.ts / 64;
      hipLaunchKernelGGL(k_sum_slabs, dim3((unsigned)(below * q * q)), dim3(256), 0, s, v, k, (const double*)W, ns,
                         U);
    }
  }
  if (stage == 2)
    hipLaunchKernelGGL(k_takahashi_diag<C>, dim3((unsigned)((below + 1) * per)), dim3(C::NT), 0, s, v, k, X,
                       (const double*)P, Q, U);
  EXA_HIP(hipGetLastError());
}

void launch_takahashi_stage(const TileView& v, int k, const double* X, double* P, double* Q, double* Zkk,
                            int stage, hipStream_t s, double* W, int max_split, const TileView* U, bool f32) {
  TileView u{};
  if (U) u = *U;
  if (v.ts % 192 == 0) takahashi_cfg<C192>(v, k, X, P, Q, W, max_split, u, stage, s, f32);
  else if (v.ts % 128 == 0) takahashi_cfg<C128>(v, k, X, P, Q, W, max_split, u, stage, s, f32);
  else takahashi_cfg<CT64>(v, k, X, P, Q, W, max_split, u, stage, s, f32);
  if (stage == 2) {
    const long len = (long)v.ts * v.ts;
    hipLaunchKernelGGL(k_sum_tiles, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, (const double*)Q,
                       v.mt - k, len, Zkk);
    EXA_HIP(hipGetLastError());
  }
}

// ------------------------------------------------ dSigma/dtheta reduction --
__device__ inline double grad_distance(double x1, double y1, double x2, double y2, int metric) {
  if (metric == 1) {
    const double d2r = M_PI / 180.0;
    const double lat1 = x1 * d2r, lon1 = y1 * d2r, lat2 = x2 * d2r, lon2 = y2 * d2r;
    const double u = sin((lat2 - lat1) * 0.5);
    const double w = sin((lon2 - lon1) * 0.5);
    return 2.0 * 6371.0 * asin(sqrt(u * u + cos(lat1) * cos(lat2) * w * w));
  }
  const double dx = x2 - x1, dy = y2 - y1;
  return sqrt(dx * dx + dy * dy);
}

// One work-group per 64 x 64 block of a lower tile (m, n), 64 x 4 threads:
// sum over its entries (i >= j on diagonal tiles, off-diagonal entries
// counted twice) of (Z_ij - alpha_i alpha_j) * dSigma_ij/dtheta_p.
__global__ void __launch_bounds__(256) k_grad_matern(TileView v, MaternGradParams P, const double* alpha,
                                                     double* partial) {
  __shared__ double red[4][4];
  const int q = v.ts / kNB;
  const long b = blockIdx.x;
  const long t = b / (q * q);
  const int rem = (int)(b - t * (q * q));
  const int bj = rem / q, bi = rem - (rem / q) * q;
  int m, n;
  tri_decode(t, m, n);
  const int ts = v.ts;
  const double* Z = tile_f64(v, m, n);
  const int i = bi * kNB + threadIdx.x;
  const long gi = (long)m * ts + i;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (gi < P.n) {
    const double xi = P.x[gi], yi = P.y[gi], ai = alpha[gi];
    for (int jj = threadIdx.y; jj < kNB; jj += 4) {
      const int j = bj * kNB + jj;
      const long gj = (long)n * ts + j;
      if (gj >= P.n) break;
      if (m == n && i < j) continue;
      const double w = ((m == n && i == j) ? 1.0 : 2.0) * (Z[i + (long)j * ts] - ai * alpha[gj]);
      const double d = grad_distance(xi, yi, P.x[gj], P.y[gj], P.metric);
      if (d == 0.0) {
        acc[0] += w;
        if (P.nugget) acc[3] += w;
        continue;
      }
      const double xx = d / P.beta;
      Dual kn, kn1;
      knu_pair(P.ko, xx, kn, kn1);
      const Dual xnu = exp(P.ko.nu * ::log(xx));  // x^nu with d/dnu = x^nu ln x
      const Dual corr = P.con * xnu * kn;
      acc[0] += w * corr.v;
      acc[1] += w * P.sigma2 * P.con.v * (xnu.v * xx * kn1.v - 2.0 * P.ko.nu.v * xnu.v * kn.v) / P.beta;
      acc[2] += w * P.sigma2 * corr.d;
    }
  }
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    double s = acc[p];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (threadIdx.x == 0) red[threadIdx.y][p] = s;
  }
  __syncthreads();
  if (threadIdx.y == 0 && threadIdx.x < 4) {
    const int p = threadIdx.x;
    partial[b * 4 + p] = (red[0][p] + red[1][p]) + (red[2][p] + red[3][p]);
  }
}

// Correlation M = c x^nu K_nu(x) at x = d/beta with dM/dbeta and dM/dnu.
__device__ inline void matern_parts(const KnuOrder<Dual>& ko, Dual con, double x, double beta, double& M,
                                    double& dbeta, double& dnu) {
  Dual kn, kn1;
  knu_pair(ko, x, kn, kn1);
  const Dual xnu = exp(ko.nu * ::log(x));
  const Dual corr = con * xnu * kn;
  M = corr.v;
  dnu = corr.d;
  dbeta = con.v * xnu.v * (x * kn1.v - 2.0 * ko.nu.v * kn.v) / beta;
}

// Parsimonious bivariate kernel, rows/columns interleaved by variable
// (row 2p+a is variable a at location p).  Each thread takes 2 x 2 blocks
// (one location pair); sums over the lower entries as in k_grad_matern, for
// theta = (s1, s2, beta, nu1, nu2, rho); partial holds 8 per block.
__global__ void __launch_bounds__(256) k_grad_bivariate(TileView v, BivariateGradParams P,
                                                        const KnuOrder<Dual>* ko, const double* alpha,
                                                        double* partial) {
  __shared__ double red[4][6];
  const int q = v.ts / kNB;
  const long b = blockIdx.x;
  const long t = b / (q * q);
  const int rem = (int)(b - t * (q * q));
  const int bj = rem / q, bi = rem - (rem / q) * q;
  int m, n;
  tri_decode(t, m, n);
  const int ts = v.ts;
  const double* Z = tile_f64(v, m, n);
  const int rp = threadIdx.x & 31;
  const int i = bi * kNB + 2 * rp;
  const long gi = (long)m * ts + i;
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (gi < P.n) {
    const double xi = P.x[gi >> 1], yi = P.y[gi >> 1];
    const double a0 = alpha[gi], a1 = alpha[gi + 1];
    for (int cp = (threadIdx.x >> 5) + 2 * threadIdx.y; cp < kNB / 2; cp += 8) {
      const int j = bj * kNB + 2 * cp;
      const long gj = (long)n * ts + j;
      if (gj >= P.n) break;
      if (m == n && i < j) continue;
      const double b0 = alpha[gj], b1 = alpha[gj + 1];
      // weights of entries (0,0), (1,0), (0,1), (1,1) of the 2 x 2 block
      const bool diag = (m == n && i == j);
      const double w00 = (diag ? 1.0 : 2.0) * (Z[i + (long)j * ts] - a0 * b0);
      const double w10 = 2.0 * (Z[(i + 1) + (long)j * ts] - a1 * b0);
      const double w01 = diag ? 0.0 : 2.0 * (Z[i + (long)(j + 1) * ts] - a0 * b1);
      const double w11 = (diag ? 1.0 : 2.0) * (Z[(i + 1) + (long)(j + 1) * ts] - a1 * b1);
      const double wc = w10 + w01;  // both cross entries carry C12
      const double d = grad_distance(xi, yi, P.x[gj >> 1], P.y[gj >> 1], P.metric);
      double M1 = 1.0, M2 = 1.0, M12 = 1.0, db1 = 0.0, db2 = 0.0, db12 = 0.0, dn1 = 0.0, dn2 = 0.0, dn12 = 0.0;
      if (d != 0.0) {
        const double xx = d / P.beta;
        matern_parts(ko[0], P.con1, xx, P.beta, M1, db1, dn1);
        matern_parts(ko[1], P.con2, xx, P.beta, M2, db2, dn2);
        matern_parts(ko[2], P.con12, xx, P.beta, M12, db12, dn12);
      }
      const double c12 = P.s12 * M12;
      acc[0] += w00 * M1 + wc * c12 / (2.0 * P.s1);
      acc[1] += w11 * M2 + wc * c12 / (2.0 * P.s2);
      acc[2] += w00 * P.s1 * db1 + wc * P.s12 * db12 + w11 * P.s2 * db2;
      acc[3] += w00 * P.s1 * dn1 + wc * (c12 * (0.5 / P.nu1 - 0.5 / P.nu12) + 0.5 * P.s12 * dn12);
      acc[4] += w11 * P.s2 * dn2 + wc * (c12 * (0.5 / P.nu2 - 0.5 / P.nu12) + 0.5 * P.s12 * dn12);
      acc[5] += wc * P.rho_ratio * M12;
    }
  }
#pragma unroll
  for (int p = 0; p < 6; ++p) {
    double s = acc[p];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (threadIdx.x == 0) red[threadIdx.y][p] = s;
  }
  __syncthreads();
  if (threadIdx.y == 0 && threadIdx.x < 8) {
    const int p = threadIdx.x;
    partial[b * 8 + p] = p < 6 ? (red[0][p] + red[1][p]) + (red[2][p] + red[3][p]) : 0.0;
  }
}

// ------------------------------------------------------ Fisher information --
// F_jk = 1/2 tr(Z D_j Z D_k) with Z = Sigma^-1 and D_p = dSigma/dtheta_p is
// evaluated as 1/2 <M_j, D_k>_F with M_j = Z D_j Z: D_j is generated into
// lower tiles, Y = Z D_j (full tile matrix), M_j = Y Z (lower tiles), and
// the gradient reduction over W = M_j (alpha = 0) gives <M_j, D_k> for
// every k at once.

// D_p into the lower tiles of v (diagonal tiles full); zero outside the
// logical n x n matrix.  One work-group per 64 x 64 block, 64 x 4 threads.
__global__ void __launch_bounds__(256) k_gen_dmatern(TileView v, MaternGradParams P, int p) {
  const int q = v.ts / kNB;
  const long b = blockIdx.x;
  const long t = b / (q * q);
  const int rem = (int)(b - t * (q * q));
  const int bj = rem / q, bi = rem - (rem / q) * q;
  int m, n;
  tri_decode(t, m, n);
  const int ts = v.ts;
  double* D = tile_f64(v, m, n);
  const int i = bi * kNB + threadIdx.x;
  const long gi = (long)m * ts + i;
  for (int jj = threadIdx.y; jj < kNB; jj += 4) {
    const int j = bj * kNB + jj;
    const long gj = (long)n * ts + j;
    double val = 0.0;
    if (gi < P.n && gj < P.n) {
      const double d = grad_distance(P.x[gi], P.y[gi], P.x[gj], P.y[gj], P.metric);
      if (d == 0.0) {
        val = (p == 0 || (p == 3 && P.nugget)) ? 1.0 : 0.0;
      } else if (p < 3) {
        const double xx = d / P.beta;
        Dual kn, kn1;
        knu_pair(P.ko, xx, kn, kn1);
        const Dual xnu = exp(P.ko.nu * ::log(xx));
        const Dual corr = P.con * xnu * kn;
        if (p == 0) val = corr.v;
        else if (p == 1) val = P.sigma2 * P.con.v * (xnu.v * xx * kn1.v - 2.0 * P.ko.nu.v * xnu.v * kn.v) / P.beta;
        else val = P.sigma2 * corr.d;
      }
    }
    D[i + (long)j * ts] = val;
  }
}

// Y(m,n) = sum_l Zs(m,l) Ds(l,n) over the l of one operand layout, for the
// full tile matrix Y (tile (m,n) at Y + (n mt + m) ts^2); Zs and Ds are the
// symmetric matrices held in the lower tiles of Z and D.  AK: l > m (A read
// K-major from Z(l,m)); BK: l >= n (B read K-major from D(l,n)).  The launch
// with AK = BK = false covers l < min(m+1, n) and writes every block; the
// others accumulate.
template <class C, bool AK, bool BK>
__global__ void __launch_bounds__(C::NT) k_fisher_zd(TileView z, TileView d, double* Y) {
  __shared__ __attribute__((aligned(16))) double lds[C::LDS_DOUBLES];
  const int qm = z.ts / C::BM, qn = z.ts / C::BN;
  const long b = blockIdx.x;
  const long tp = b / (qm * qn);
  const int rem = (int)(b % (qm * qn));
  const int bi = rem % qm, bj = rem / qm;
  const int mt = z.mt;
  const int m = (int)(tp % mt), n = (int)(tp / mt);
  const long ts = z.ts;
  int l0, l1;  // [l0, l1)
  if (!AK && !BK) { l0 = 0; l1 = min(m + 1, n); }
  else if (!AK && BK) { l0 = n; l1 = m + 1; }
  else if (AK && BK) { l0 = max(m + 1, n); l1 = mt; }
  else { l0 = m + 1; l1 = n; }
  const bool first = !AK && !BK;
  if (!first && l0 >= l1) return;
  dx4 acc[C::FM][C::FN];
  zero_acc<C>(acc);
  for (int l = l0; l < l1; ++l) {
    const double* A = AK ? tile_f64(z, l, m) + (long)bi * C::BM * ts : tile_f64(z, m, l) + bi * C::BM;
    const double* B = BK ? tile_f64(d, l, n) + (long)bj * C::BN * ts : tile_f64(d, n, l) + bj * C::BN;
    gemm_nt_acc<C, BK, double, double, AK>(acc, A, ts, B, ts, (int)ts, lds, 1.0);
  }
  acc_update<C>(acc, Y + ((long)n * mt + m) * ts * ts + bi * C::BM + (long)bj * C::BN * ts, ts, first ? 0.0 : 1.0);
}

// M(m,n) = sum_l Y(m,l) Zs(l,n) for the lower tiles m >= n of M.  BK: l >= n
// (B read K-major from Z(l,n)); the launch with BK = false (l < n) writes
// every block.
template <class C, bool BK>
__global__ void __launch_bounds__(C::NT) k_fisher_yz(const double* Y, TileView z, TileView mo) {
  __shared__ __attribute__((aligned(16))) double lds[C::LDS_DOUBLES];
  const int qm = z.ts / C::BM, qn = z.ts / C::BN;
  const long b = blockIdx.x;
  const long tp = b / (qm * qn);
  const int rem = (int)(b % (qm * qn));
  const int bi = rem % qm, bj = rem / qm;
  int m, n;
  tri_decode(tp, m, n);
  const int mt = z.mt;
  const long ts = z.ts;
  const int l0 = BK ? n : 0, l1 = BK ? mt : n;
  if (BK && l0 >= l1) return;
  dx4 acc[C::FM][C::FN];
  zero_acc<C>(acc);
  for (int l = l0; l < l1; ++l) {
    const double* A = Y + ((long)l * mt + m) * ts * ts + bi * C::BM;
    const double* B = BK ? tile_f64(z, l, n) + (long)bj * C::BN * ts : tile_f64(z, n, l) + bj * C::BN;
    gemm_nt_acc<C, BK>(acc, A, ts, B, ts, (int)ts, lds, 1.0);
  }
  acc_update<C>(acc, tile_f64(mo, m, n) + bi * C::BM + (long)bj * C::BN * ts, ts, BK ? 1.0 : 0.0);
}

template <class C>
static void fisher_cfg(const TileView& z, const TileView& d, double* Y, const TileView& mo, hipStream_t s) {
  const long per = (long)(z.ts / C::BM) * (z.ts / C::BN);
  const long full = (long)z.mt * z.mt * per;
  const long lower = (long)z.mt * (z.mt + 1) / 2 * per;
  hipLaunchKernelGGL((k_fisher_zd<C, false, false>), dim3((unsigned)full), dim3(C::NT), 0, s, z, d, Y);
  hipLaunchKernelGGL((k_fisher_zd<C, false, true>), dim3((unsigned)full), dim3(C::NT), 0, s, z, d, Y);
  hipLaunchKernelGGL((k_fisher_zd<C, true, true>), dim3((unsigned)full), dim3(C::NT), 0, s, z, d, Y);
  hipLaunchKernelGGL((k_fisher_zd<C, true, false>), dim3((unsigned)full), dim3(C::NT), 0, s, z, d, Y);
  hipLaunchKernelGGL((k_fisher_yz<C, false>), dim3((unsigned)lower), dim3(C::NT), 0, s, (const double*)Y, z, mo);
  hipLaunchKernelGGL((k_fisher_yz<C, true>), dim3((unsigned)lower), dim3(C::NT), 0, s, (const double*)Y, z, mo);
  EXA_HIP(hipGetLastError());
}

void launch_gen_dmatern(const TileView& d, const MaternGradParams& P, int p, hipStream_t s) {
  hipLaunchKernelGGL(k_gen_dmatern, dim3((unsigned)grad_partial_blocks(d)), dim3(64, 4), 0, s, d, P, p);
  EXA_HIP(hipGetLastError());
}

void launch_fisher_zdz(const TileView& z, const TileView& d, double* Y, const TileView& mo, hipStream_t s) {
  if (z.ts % 192 == 0) fisher_cfg<C192>(z, d, Y, mo, s);
  else if (z.ts % 128 == 0) fisher_cfg<C128>(z, d, Y, mo, s);
  else fisher_cfg<CT64>(z, d, Y, mo, s);
}

// out[p * G + g] = sum over b = g, g + G, ... < nb of partial[b*stride + p].
__global__ void __launch_bounds__(256) k_reduce_partials(const double* partial, long nb, int stride, double* out) {
  __shared__ double red[4];
  const int p = blockIdx.y;
  const int G = gridDim.x;
  double s = 0.0;
  for (long b = (long)blockIdx.x + (long)threadIdx.x * G; b < nb; b += (long)G * 256) s += partial[b * stride + p];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[p * G + blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

long grad_partial_blocks(const TileView& v) {
  const long q = v.ts / kNB;
  return (long)v.mt * (v.mt + 1) / 2 * q * q;
}

void launch_grad_matern(const TileView& v, const MaternGradParams& P, const double* alpha, double* partial,
                        double* out, int groups, hipStream_t s) {
  const long nb = grad_partial_blocks(v);
  hipLaunchKernelGGL(k_grad_matern, dim3((unsigned)nb), dim3(64, 4), 0, s, v, P, alpha, partial);
  EXA_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_partials, dim3((unsigned)groups, 4), dim3(256), 0, s, (const double*)partial, nb, 4,
                     out);
  EXA_HIP(hipGetLastError());
}

void launch_grad_bivariate(const TileView& v, const BivariateGradParams& P, const KnuOrder<Dual>* ko,
                           const double* alpha, double* partial, double* out, int groups, hipStream_t s) {
  const long nb = grad_partial_blocks(v);
  hipLaunchKernelGGL(k_grad_bivariate, dim3((unsigned)nb), dim3(64, 4), 0, s, v, P, ko, alpha, partial);
  EXA_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_partials, dim3((unsigned)groups, 6), dim3(256), 0, s, (const double*)partial, nb, 8,
                     out);
  EXA_HIP(hipGetLastError());
}

}  // namespace exa
