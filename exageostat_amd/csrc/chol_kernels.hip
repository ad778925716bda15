// FP64 tile Cholesky kernels: diagonal-tile factorisation, panel triangular
// solve and trailing update, all on FP64 MFMA.
//
// Replaces Chameleon's CHAMELEON_dpotrf_Tile(ChamLower, C) as called from
// EXAGEOSTAT_dmle_Tile (exageostat_exact/src/compute/MLE_dexact.c:372);
// Chameleon is not vendored, the algorithm is the right-looking tile
// Cholesky (potrf of A(k,k), trsm of A(m,k), syrk/gemm of A(m,n), m>=n>k).
#include "common.hpp"
#include "kernels.hpp"
#include "mfma.hpp"

#include <cstdlib>

namespace exa {

using C64 = GemmCfg<64, 64, 2, 2, 16>;

// ---------------------------------------------------------------------------
// Unblocked Cholesky of the 64x64 diagonal block j of a diagonal tile, plus
// the explicit inverse of the resulting lower-triangular block, by one wave
// (no multi-wave barriers on the critical path of the panel).  Lane r owns
// row r.  Column c is finished left-looking,
//   L(r,c) = (A(r,c) - sum_{k<c} L(r,k) L(c,k)) / L(c,c),
// and lane r then forms column r of X = L^-1 by forward substitution,
//   X(i,r) = (delta_ir - sum_{k<i} L(i,k) X(k,r)) / L(i,i);
// each sum is taken as four interleaved partial sums (shorter dependent
// chains).
// Both matrices live in LDS column-major (element (i,k) at k*LD + i), so a
// lane's own entries of one column are contiguous across the wave and
// L(c,k) / L(i,k) are broadcast reads.
__global__ void __launch_bounds__(64) k_potrf_block(double* T, int ts, int j, double* dinv, int* info,
                                                    long col_base) {
  constexpr int LD = 65;
  __shared__ double Ls[64 * LD];  // L(i, k) at Ls[k * LD + i]
  __shared__ double Xs[64 * LD];  // X(i, k) at Xs[k * LD + i] (written transposed, see below)
  const int r = threadIdx.x;
  double* blk = T + (long)j * kNB * (ts + 1);
  {
    // all 64 column loads in flight before the first LDS store
    double col[64];
#pragma unroll
    for (int k = 0; k < 64; ++k) col[k] = blk[r + (long)k * ts];
#pragma unroll
    for (int k = 0; k < 64; ++k) Ls[k * LD + r] = col[k];
  }
  __syncthreads();
  for (int c = 0; c < 64; ++c) {
    // four partial sums over k mod 4 break the dependent FMA chain
    double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
    int k = 0;
#pragma unroll 4
    for (; k + 4 <= c; k += 4) {
      p0 += Ls[k * LD + r] * Ls[k * LD + c];
      p1 += Ls[(k + 1) * LD + r] * Ls[(k + 1) * LD + c];
      p2 += Ls[(k + 2) * LD + r] * Ls[(k + 2) * LD + c];
      p3 += Ls[(k + 3) * LD + r] * Ls[(k + 3) * LD + c];
    }
    for (; k < c; ++k) p0 += Ls[k * LD + r] * Ls[k * LD + c];
    const double s = Ls[c * LD + r] - ((p0 + p1) + (p2 + p3));
    if (r == c) {
      const bool bad = !(s > 0.0);
      if (bad) atomicCAS(info, 0, (int)(col_base + c + 1));
      Ls[c * LD + c] = bad ? __builtin_nan("") : sqrt(s);
    }
    __syncthreads();
    if (r > c) Ls[c * LD + r] = s / Ls[c * LD + c];
    __syncthreads();
  }
  for (int k = 0; k < 64; ++k) blk[r + (long)k * ts] = (r >= k) ? Ls[k * LD + r] : 0.0;
  // Lane r keeps X(i, r) at Xs[i * LD + r] (its own column, contiguous across
  // lanes for each i).
  for (int i = 0; i < 64; ++i) {
    double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
    int k = 0;
#pragma unroll 4
    for (; k + 4 <= i; k += 4) {
      p0 += Ls[k * LD + i] * Xs[k * LD + r];
      p1 += Ls[(k + 1) * LD + i] * Xs[(k + 1) * LD + r];
      p2 += Ls[(k + 2) * LD + i] * Xs[(k + 2) * LD + r];
      p3 += Ls[(k + 3) * LD + i] * Xs[(k + 3) * LD + r];
    }
    for (; k < i; ++k) p0 += Ls[k * LD + i] * Xs[k * LD + r];
    Xs[i * LD + r] = (((i == r) ? 1.0 : 0.0) - ((p0 + p1) + (p2 + p3))) / Ls[i * LD + i];
  }
  __syncthreads();
  // dinv column-major: dinv[i + c*64] = X(i, c) = Xs[i * LD + c].
  for (int c = 0; c < 64; ++c) dinv[r + c * 64] = Xs[r * LD + c];
}

// ---------------------------------------------------------------------------
// X L^T = B for one 64-row slab X (in place of B), L lower with nL diagonal
// 64-blocks whose inverses are in dinv.  Column sweep: for block jb,
// X_jb = (B_jb - X_<jb L_jb,<jb^T) inv(L_jb,jb)^T.
template <class TX>
__device__ void trsm_slab(TX* X, long ldx, const double* L, long ldl, int nL, const double* dinv,
                          double* lds) {
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave / C64::WGN;
  const int wn = wave - wm * C64::WGN;
  constexpr int LDT = 68;
  for (int jb = 0; jb < nL; ++jb) {
    dx4 acc[C64::FM][C64::FN];
    TX* Xj = X + (long)jb * kNB * ldx;
    acc_load<C64, TX>(acc, Xj, ldx);
    gemm_nt_acc<C64, false, TX, double>(acc, X, ldx, L + jb * kNB, ldl, jb * kNB, lds, -1.0);
    double* Ts = lds;
    double* Is = lds + 64 * LDT;
#pragma unroll
    for (int i = 0; i < C64::FM; ++i)
#pragma unroll
      for (int jj = 0; jj < C64::FN; ++jj) {
        const int c = wn * C64::WN + jj * 16 + frag_col(lane);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = wm * C64::WM + i * 16 + frag_row(lane, v);
          Ts[c * LDT + r] = acc[i][jj][v];
        }
      }
    const double* inv = dinv + (long)jb * 4096;
    for (int e = tid; e < 64 * 32; e += 256) {
      const int kk = e >> 5;
      const int c2 = e & 31;
      *reinterpret_cast<double2*>(Is + kk * LDT + 2 * c2) =
          *reinterpret_cast<const double2*>(inv + 2 * c2 + kk * 64);
    }
    __syncthreads();
    dx4 x[C64::FM][C64::FN];
    zero_acc<C64>(x);
#pragma unroll 4
    for (int kk = 0; kk < 64; kk += 4) {
      const int krow = kk + (lane >> 4);
      double a[C64::FM], b[C64::FN];
#pragma unroll
      for (int i = 0; i < C64::FM; ++i) a[i] = Ts[krow * LDT + wm * C64::WM + i * 16 + (lane & 15)];
#pragma unroll
      for (int jj = 0; jj < C64::FN; ++jj) b[jj] = Is[krow * LDT + wn * C64::WN + jj * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < C64::FM; ++i)
#pragma unroll
        for (int jj = 0; jj < C64::FN; ++jj) x[i][jj] = mfma_f64(a[i], b[jj], x[i][jj]);
    }
    acc_update<C64, TX>(x, Xj, ldx, 0.0);
    // X_jb is read back from global memory by the next column block of this
    // same work-group: the barrier's workgroup-scope release/acquire makes
    // the stores visible inside the CU (no device-scope fence, which would
    // write back the XCD's L2).
    __syncthreads();
  }
}

// Panel: slabs over tiles m in (k, mt) and their 64-row blocks.
__global__ void __launch_bounds__(256) k_trsm_panel(TileView v, int k, const double* dinv) {
  __shared__ __attribute__((aligned(16))) double lds[2 * 64 * 68];
  const int q = v.ts / kNB;
  const int s = blockIdx.x;
  const int m = k + 1 + s / q;
  const int rb = s - (s / q) * q;
  const double* L = tile_f64(v, k, k);
  const long t = tile_index(v, m, k);
  // Mixed-precision factors store off-diagonal tiles in FP32; the solve
  // itself runs in FP64 either way.
  if (v.prec && v.prec[t] != kF64) {
    float* X = static_cast<float*>(v.ptr[t]) + rb * kNB;
    trsm_slab<float>(X, v.ts, L, v.ts, q, dinv, lds);
  } else {
    double* X = static_cast<double*>(v.ptr[t]) + rb * kNB;
    trsm_slab<double>(X, v.ts, L, v.ts, q, dinv, lds);
  }
}

// Inside a diagonal tile after block j is factored: blocks (i, j), i > j.
__global__ void __launch_bounds__(256) k_trsm_intra(double* T, int ts, int j, const double* dinv_j) {
  __shared__ __attribute__((aligned(16))) double lds[2 * 64 * 68];
  const int i = j + 1 + blockIdx.x;
  double* X = T + (long)i * kNB + (long)j * kNB * ts;
  trsm_slab<double>(X, ts, nullptr, ts, 1, dinv_j, lds);
}

// Inside a diagonal tile: blocks (i, l), j < l <= i: B_il -= B_ij B_lj^T.
__global__ void __launch_bounds__(256) k_syrk_intra(double* T, int ts, int j) {
  __shared__ __attribute__((aligned(16))) double lds[C64::LDS_DOUBLES];
  int r, c;
  tri_decode(blockIdx.x, r, c);
  const int i = j + 1 + r;
  const int l = j + 1 + c;
  const double* A = T + (long)i * kNB + (long)j * kNB * ts;
  const double* B = T + (long)l * kNB + (long)j * kNB * ts;
  double* Cp = T + (long)i * kNB + (long)l * kNB * ts;
  dx4 acc[C64::FM][C64::FN];
  zero_acc<C64>(acc);
  gemm_nt_acc<C64>(acc, A, ts, B, ts, kNB, lds, -1.0);
  acc_update<C64>(acc, Cp, ts, 1.0);
}

// ---------------------------------------------------------------------------
// Explicit inverse X = L_kk^-1 of every diagonal tile (for the triangular
// solves, which then become matrix-vector products).  One work-group per
// (tile k, 64-column block j): X_jj = inv(L_jj) from the potrf, then for
// i > j: X_ij = -inv(L_ii) sum_{l=j}^{i-1} L_il X_lj.
__global__ void __launch_bounds__(256) k_trtri_tiles(TileView v, const double* dinv, double* xinv) {
  __shared__ __attribute__((aligned(16))) double lds[2 * 64 * 68];
  constexpr int LDT = 68;
  const int ts = v.ts;
  const int q = ts / kNB;
  const int k = blockIdx.x / q;
  const int j = blockIdx.x - k * q;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave / C64::WGN;
  const int wn = wave - wm * C64::WGN;
  const double* L = tile_f64(v, k, k);
  double* X = xinv + (long)k * ts * ts;
  const double* dk = dinv + (long)k * q * 4096;
  for (int e = tid; e < 64 * 64; e += 256) {
    const int r = e & 63, c = e >> 6;
    X[(j * kNB + r) + (long)(j * kNB + c) * ts] = dk[(long)j * 4096 + r + c * 64];
  }
  for (int e = tid; e < j * kNB * kNB; e += 256) {
    const int r = e % (j * kNB), c = e / (j * kNB);
    X[r + (long)(j * kNB + c) * ts] = 0.0;
  }
  __syncthreads();
  for (int i = j + 1; i < q; ++i) {
    dx4 acc[C64::FM][C64::FN];
    zero_acc<C64>(acc);
    gemm_nt_acc<C64, true>(acc, L + i * kNB + (long)j * kNB * ts, ts, X + j * kNB + (long)j * kNB * ts, ts,
                           (i - j) * kNB, lds, 1.0);
    double* As = lds;             // inv(L_ii): As[k][r] = inv(r, k)
    double* Bs = lds + 64 * LDT;  // acc:       Bs[k][c] = acc(k, c)
#pragma unroll
    for (int a = 0; a < C64::FM; ++a)
#pragma unroll
      for (int b = 0; b < C64::FN; ++b) {
        const int c = wn * C64::WN + b * 16 + frag_col(lane);
#pragma unroll
        for (int vv = 0; vv < 4; ++vv) {
          const int r = wm * C64::WM + a * 16 + frag_row(lane, vv);
          Bs[r * LDT + c] = acc[a][b][vv];
        }
      }
    const double* inv = dk + (long)i * 4096;
    for (int e = tid; e < 64 * 32; e += 256) {
      const int kk = e >> 5;
      const int r2 = e & 31;
      *reinterpret_cast<double2*>(As + kk * LDT + 2 * r2) = *reinterpret_cast<const double2*>(inv + 2 * r2 + kk * 64);
    }
    __syncthreads();
    dx4 y[C64::FM][C64::FN];
    zero_acc<C64>(y);
#pragma unroll 4
    for (int kk = 0; kk < 64; kk += 4) {
      const int krow = kk + (lane >> 4);
      double a[C64::FM], b[C64::FN];
#pragma unroll
      for (int ii = 0; ii < C64::FM; ++ii) a[ii] = -As[krow * LDT + wm * C64::WM + ii * 16 + (lane & 15)];
#pragma unroll
      for (int jj = 0; jj < C64::FN; ++jj) b[jj] = Bs[krow * LDT + wn * C64::WN + jj * 16 + (lane & 15)];
#pragma unroll
      for (int ii = 0; ii < C64::FM; ++ii)
#pragma unroll
        for (int jj = 0; jj < C64::FN; ++jj) y[ii][jj] = mfma_f64(a[ii], b[jj], y[ii][jj]);
    }
    acc_update<C64>(y, X + i * kNB + (long)j * kNB * ts, ts, 0.0);
    // X_ij is read back by the next i of this work-group (same CU).
    __syncthreads();
  }
}

void launch_trtri_tiles(const TileView& v, const double* dinv, double* xinv, hipStream_t s) {
  const int q = v.ts / kNB;
  hipLaunchKernelGGL(k_trtri_tiles, dim3((unsigned)(v.mt * q)), dim3(256), 0, s, v, dinv, xinv);
  EXA_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Trailing update A(m,n) -= sum_{p < np} A(m,k+p) A(n,k+p)^T over BM x BN
// blocks (np = 2: two factorisation steps at once, K = 2 ts per block and
// one read and write of C).
// Tile pairs (m, n) with n in [col_lo, col_hi] and m in [n, mt), column by
// column (consecutive pairs share the panel tile n).
//
// Persistent launch: gridDim.x work-groups (at most one per CU, fewer than
// the CU count so the panel kernels of the critical path always find free
// CUs) loop over the block list.  Work-groups are dispatched round-robin
// over the XCDs, so work-group w runs on XCD w % nx; each XCD gets a
// contiguous slice of the block list and its work-groups walk it together,
// which keeps the A/B row panels they share hot in that XCD's L2.  The
// operands go through the three-stage direct-to-LDS ring (gemm_nt_ring):
// 2.5% faster at n = 90,000 than register-staged double buffering, with
// the same products in the same order.
template <class C>
__device__ __forceinline__ void update_block(const TileView& v, int k, int np, int col_lo, int col_hi, long b,
                                             double* lds) {
  const int qm = v.ts / C::BM;
  const int qn = v.ts / C::BN;
  const long p = b / (qm * qn);
  const int rem = (int)(b - p * (qm * qn));
  const int bi = rem % qm;
  const int bj = rem / qm;
  // column n: the last c in [col_lo, col_hi] with pairs_before(c) <= p
  const int mt = v.mt;
  auto pairs_before = [&](int c) {
    return (long)(c - col_lo) * mt - ((long)c * (c - 1) / 2 - (long)col_lo * (col_lo - 1) / 2);
  };
  int lo = col_lo, hi = col_hi;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (pairs_before(mid) <= p) lo = mid;
    else hi = mid - 1;
  }
  const int n = lo;
  const int m = n + (int)(p - pairs_before(n));
  if (m == n && (bi + 1) * C::BM <= bj * C::BN) return;
  double* Cp = tile_f64(v, m, n) + bi * C::BM + (long)bj * C::BN * v.ts;
  dx4 acc[C::FM][C::FN];
  zero_acc<C>(acc);
#pragma unroll 1
  for (int p = 0; p < np; ++p) {
    const double* A = tile_f64(v, m, k + p) + bi * C::BM;
    const double* B = tile_f64(v, n, k + p) + bj * C::BN;
    gemm_nt_ring<C>(acc, A, v.ts, B, v.ts, v.ts, lds, -1.0);
  }
  acc_update<C>(acc, Cp, v.ts, 1.0);
}

template <class C>
__global__ void __launch_bounds__(C::NT) k_tile_update(TileView v, int k, int np, int col_lo, int col_hi,
                                                      long nblocks, int nx) {
  __shared__ __attribute__((aligned(16))) double lds[RingLds<C>::DOUBLES];
  const int w = blockIdx.x;
  const int x = w % nx;
  const int gx = gridDim.x / nx;
  const long per = (nblocks + nx - 1) / nx;
  const long start = (long)x * per;
  const long end = start + per < nblocks ? start + per : nblocks;
  for (long b = start + w / nx; b < end; b += gx) update_block<C>(v, k, np, col_lo, col_hi, b, lds);
}

static int g_update_wgs = 0;

// Work-groups of the persistent update: CUs minus a reserve for the panel
// factorisation (EXAGEOSTAT_RESERVE_CUS, default 8: at n = 90,000 the FP64
// factorisation takes 4.40 s with 4 or 8 reserved CUs, 4.52 s with 16 and
// 4.66 s with 24), times the occupancy,
// rounded down to a multiple of 8 (XCDs).
template <class C>
static int update_grid_cap() {
  static int cap = 0;
  if (cap) return cap;
  int dev = 0, cus = 0, occ = 1;
  EXA_HIP(hipGetDevice(&dev));
  EXA_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  EXA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_tile_update<C>, C::NT, 0));
  int reserve = 8;
  if (const char* e = getenv("EXAGEOSTAT_RESERVE_CUS")) reserve = atoi(e);
  if (reserve < 0 || reserve >= cus) reserve = 0;
  cap = (cus - reserve) * (occ > 0 ? occ : 1);
  if (cap >= 8) cap -= cap % 8;
  if (cap < 1) cap = 1;
  return cap;
}

template <class C>
static void launch_update_cfg(const TileView& v, int k, int np, int col_lo, int col_hi, hipStream_t s) {
  long pairs = 0;
  for (int c = col_lo; c <= col_hi && c < v.mt; ++c) pairs += v.mt - c;
  if (pairs <= 0) return;
  const long blocks = pairs * (v.ts / C::BM) * (v.ts / C::BN);
  const int cap = update_grid_cap<C>();
  int grid = blocks < cap ? (int)blocks : cap;
  int nx = 8;
  if (grid % 8 != 0) nx = 1;
  g_update_wgs = grid;
  hipLaunchKernelGGL(k_tile_update<C>, dim3((unsigned)grid), dim3(C::NT), 0, s, v, k, np, col_lo, col_hi, blocks,
                     nx);
  EXA_HIP(hipGetLastError());
}

void launch_tile_update(const TileView& v, int k, int np, int col_lo, int col_hi, hipStream_t s) {
  if (col_hi >= v.mt) col_hi = v.mt - 1;
  if (col_lo >= v.mt || col_hi < col_lo) return;
  if (np < 1 || np > 2 || col_lo < k + np) throw std::invalid_argument("tile_update: bad panel count or columns");
  if (v.ts % 192 == 0) {
    launch_update_cfg<GemmCfg<192, 192, 2, 4, 16>>(v, k, np, col_lo, col_hi, s);
  } else if (v.ts % 128 == 0) {
    launch_update_cfg<GemmCfg<128, 128, 2, 2, 16>>(v, k, np, col_lo, col_hi, s);
  } else {
    launch_update_cfg<C64>(v, k, np, col_lo, col_hi, s);
  }
}

// ---------------------------------------------------------------------------
void launch_potrf_tile(double* T, int ts, double* dinv_k, int* info, long col_base, hipStream_t s) {
  const int q = ts / kNB;
  for (int j = 0; j < q; ++j) {
    hipLaunchKernelGGL(k_potrf_block, dim3(1), dim3(64), 0, s, T, ts, j, dinv_k + (long)j * 4096, info,
                       col_base + (long)j * kNB);
    if (j + 1 < q) {
      const int nb = q - j - 1;
      hipLaunchKernelGGL(k_trsm_intra, dim3(nb), dim3(256), 0, s, T, ts, j, dinv_k + (long)j * 4096);
      hipLaunchKernelGGL(k_syrk_intra, dim3(nb * (nb + 1) / 2), dim3(256), 0, s, T, ts, j);
    }
  }
  EXA_HIP(hipGetLastError());
}

void launch_trsm_panel(const TileView& v, int k, const double* dinv_k, hipStream_t s, int m_end) {
  const int q = v.ts / kNB;
  if (m_end < 0 || m_end > v.mt) m_end = v.mt;
  const long slabs = (long)(m_end - k - 1) * q;
  if (slabs <= 0) return;
  hipLaunchKernelGGL(k_trsm_panel, dim3((unsigned)slabs), dim3(256), 0, s, v, k, dinv_k);
  EXA_HIP(hipGetLastError());
}

}  // namespace exa
