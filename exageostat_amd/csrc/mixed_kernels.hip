// Kernels of the mixed-precision tile Cholesky: FP64 -> mixed conversion of
// the generated covariance, trailing updates per precision class of the
// updated tile (FP64 MFMA; split-BF16 MFMA for FP32 tiles; single-BF16 MFMA
// for far tiles; FP32 accumulation) and
// the transposed half of the symmetric matrix-vector product used by the
// FP64 iterative refinement.
//
// The reference's mixed tier (chameleon_psdpotrf, exageostat_exact/src/
// compute/psdportf.c:44-187) keeps a band of |m-n| < diag_thick tiles in
// FP64 and the rest in FP32; here the class of every tile is chosen per theta
// from the largest correlation the tile can hold, and far tiles are updated
// with BF16 operands.
#include "common.hpp"
#include "kernels.hpp"
#include "mfma.hpp"
#include "mfma_lp.hpp"

#include <cstdlib>
#include <stdexcept>
#include <type_traits>

namespace exa {

// ------------------------------------------------------------ conversion --
// dst (mixed) <- src (FP64, same tiling); 4 work-groups per tile.
__global__ void __launch_bounds__(256) k_convert_mixed(TileView src, TileView dst) {
  const long t = blockIdx.x >> 2;
  const int part = blockIdx.x & 3;
  const long nel = (long)src.ts * src.ts;
  const long beg = part * (nel / 4), end = beg + nel / 4;
  const double* s = static_cast<const double*>(src.ptr[t]);
  const uint8_t p = dst.prec ? dst.prec[t] : (uint8_t)kF64;
  if (p == kF64) {
    double* d = static_cast<double*>(dst.ptr[t]);
    for (long i = beg + 2 * threadIdx.x; i < end; i += 512)
      *reinterpret_cast<double2*>(d + i) = *reinterpret_cast<const double2*>(s + i);
  } else {
    float* d = static_cast<float*>(dst.ptr[t]);
    for (long i = beg + 2 * threadIdx.x; i < end; i += 512) {
      const double2 v = *reinterpret_cast<const double2*>(s + i);
      *reinterpret_cast<float2*>(d + i) = float2{(float)v.x, (float)v.y};
    }
  }
}

void launch_convert_mixed(const TileView& src, const TileView& dst, long ntiles, hipStream_t s) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_convert_mixed, dim3((unsigned)(ntiles * 4)), dim3(256), 0, s, src, dst);
  EXA_HIP(hipGetLastError());
}

// dst (FP64) <- src (mixed, same tiling): the reverse of k_convert_mixed.
__global__ void __launch_bounds__(256) k_mixed_to_fp64(TileView src, TileView dst) {
  const long t = blockIdx.x >> 2;
  const int part = blockIdx.x & 3;
  const long nel = (long)src.ts * src.ts;
  const long beg = part * (nel / 4), end = beg + nel / 4;
  double* d = static_cast<double*>(dst.ptr[t]);
  const uint8_t p = src.prec ? src.prec[t] : (uint8_t)kF64;
  if (p == kF64) {
    const double* s = static_cast<const double*>(src.ptr[t]);
    for (long i = beg + 2 * threadIdx.x; i < end; i += 512)
      *reinterpret_cast<double2*>(d + i) = *reinterpret_cast<const double2*>(s + i);
  } else {
    const float* s = static_cast<const float*>(src.ptr[t]);
    for (long i = beg + 2 * threadIdx.x; i < end; i += 512) {
      const float2 v = *reinterpret_cast<const float2*>(s + i);
      *reinterpret_cast<double2*>(d + i) = double2{(double)v.x, (double)v.y};
    }
  }
}

void launch_mixed_to_fp64(const TileView& src, const TileView& dst, long ntiles, hipStream_t s) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_mixed_to_fp64, dim3((unsigned)(ntiles * 4)), dim3(256), 0, s, src, dst);
  EXA_HIP(hipGetLastError());
}

// ---------------------------------------------------------- panel split --
// Panel tiles (m, k), m in (k, mt), converted once per step into BF16 hi and
// lo parts for the BF16 classes: tile m of `split` holds hi at [0, ts^2) and
// lo at [ts^2, 2 ts^2), each laid out [k-group g][row r][8] (the layout the
// update kernels copy straight into LDS).  One thread per (row, group).
__global__ void __launch_bounds__(256) k_split_panel(TileView v, int k, uint16_t* split) {
  const int ts = v.ts;
  const long per_tile = (long)ts * (ts / 8);
  const long u = (long)blockIdx.x * 256 + threadIdx.x;
  const int m = k + 1 + (int)(u / per_tile);
  if (m >= v.mt) return;
  const long w = u - (long)(m - k - 1) * per_tile;
  const int g = (int)(w / ts);
  const int r = (int)(w - (long)g * ts);
  const long t = tile_index(v, m, k);
  float f[8];
  if (v.prec[t] == kF64) {
    const double* T = static_cast<const double*>(v.ptr[t]);
#pragma unroll
    for (int e = 0; e < 8; ++e) f[e] = (float)T[r + (long)(8 * g + e) * ts];
  } else {
    const float* T = static_cast<const float*>(v.ptr[t]);
#pragma unroll
    for (int e = 0; e < 8; ++e) f[e] = T[r + (long)(8 * g + e) * ts];
  }
  uint32_t hi[4], lo[4];
#pragma unroll
  for (int h = 0; h < 4; ++h) split_pair<3>(f[2 * h], f[2 * h + 1], hi[h], lo[h]);
  uint16_t* dst = split + (long)m * 2 * ts * ts + ((long)g * ts + r) * 8;
  *reinterpret_cast<uint4*>(dst) = uint4{hi[0], hi[1], hi[2], hi[3]};
  *reinterpret_cast<uint4*>(dst + (long)ts * ts) = uint4{lo[0], lo[1], lo[2], lo[3]};
}

void launch_split_panel(const TileView& v, int k, uint16_t* split, hipStream_t s) {
  const long units = (long)(v.mt - k - 1) * v.ts * (v.ts / 8);
  if (units <= 0) return;
  hipLaunchKernelGGL(k_split_panel, dim3((unsigned)((units + 255) / 256)), dim3(256), 0, s, v, k, split);
  EXA_HIP(hipGetLastError());
}

// --------------------------------------------------------- class updates --
// One block of C(m,n) -= A(m,k) B(n,k)^T, C of precision class CLS.
template <int CLS, int BS>
struct ClassCfg;
// FP64 class: the FP64 core.
template <int BS>
struct ClassCfg<kF64, BS> {
  using C = typename std::conditional<BS == 192, GemmCfg<192, 192, 2, 4, 16>,
            typename std::conditional<BS == 128, GemmCfg<128, 128, 2, 2, 16>, GemmCfg<64, 64, 2, 2, 16>>::type>::type;
  static constexpr int LDS_BYTES = C::LDS_DOUBLES * 8;
};
// FP32 class: FP32 storage, split-BF16 (three-product) MFMA.
template <int BS>
struct ClassCfg<kF32, BS> {
  using C = typename std::conditional<BS == 192, LpCfg<192, 192, 2, 4, 32>,
            typename std::conditional<BS == 128, LpCfg<128, 128, 2, 2, 32>, LpCfg<64, 64, 2, 2, 32>>::type>::type;
  static constexpr int NSPLIT = 3;
  static constexpr int LDS_BYTES = Bf16dLds<C, NSPLIT>::HALVES * 2;
};
// FP32 class with the lo*lo product as well (four split-BF16 products); a
// kernel variant of kF32, not a storage class.
constexpr int kF32Q = 4;
template <int BS>
struct ClassCfg<kF32Q, BS> {
  using C = typename ClassCfg<kF32, BS>::C;
  static constexpr int NSPLIT = 4;
  static constexpr int LDS_BYTES = Bf16dLds<C, NSPLIT>::HALVES * 2;
};
// Far class: FP32 storage, single BF16 MFMA operands.
template <int BS>
struct ClassCfg<kF32B, BS> {
  using C = typename std::conditional<BS == 192, LpCfg<192, 192, 2, 4, 64>,
            typename std::conditional<BS == 128, LpCfg<128, 128, 2, 2, 64>, LpCfg<64, 64, 2, 2, 64>>::type>::type;
  static constexpr int NSPLIT = 1;
  static constexpr int LDS_BYTES = Bf16dLds<C, NSPLIT>::HALVES * 2;
};

// FP64 class: one block of C(m,n) -= A(m,k) B(n,k)^T on the FP64 matrix
// cores, operands read from the tiles (FP64 or FP32 storage).
template <int BS, class TA, class TB>
__device__ __forceinline__ void f64_block(const TA* A, const TB* B, double* Cp, int ts, unsigned char* lds) {
  using C = typename ClassCfg<kF64, BS>::C;
  dx4 acc[C::FM][C::FN];
  zero_acc<C>(acc);
  gemm_nt_acc<C, false, TA, TB>(acc, A, ts, B, ts, ts, reinterpret_cast<double*>(lds), -1.0);
  acc_update<C, double>(acc, Cp, ts, 1.0);
}

// One BM x BN block (bi, bj) of C(m, n) -= A(m, k) A(n, k)^T with tile
// (m, n) of precision class CLS.  The FP64 class reads the panel tiles; the
// BF16 classes read the panel's BF16 parts from `split` (tile j at
// split + 2 j ts^2, see k_split_panel).
template <int CLS, int BS>
__device__ __forceinline__ void class_update(const TileView& v, int k, int m, int n, int bi, int bj,
                                             const uint16_t* split, unsigned char* lds) {
  using C = typename ClassCfg<CLS, BS>::C;
  const long tc = tile_index(v, m, n);
  const long co = bi * C::BM + (long)bj * C::BN * v.ts;
  if constexpr (CLS == kF64) {
    const long ta = tile_index(v, m, k), tb = tile_index(v, n, k);
    const bool fa = v.prec[ta] != kF64, fb = v.prec[tb] != kF64;
    const long ao = bi * C::BM, bo = bj * C::BN;
    double* Cp = static_cast<double*>(v.ptr[tc]) + co;
    if (fa) {
      const float* A = static_cast<const float*>(v.ptr[ta]) + ao;
      if (fb) f64_block<BS>(A, static_cast<const float*>(v.ptr[tb]) + bo, Cp, v.ts, lds);
      else f64_block<BS>(A, static_cast<const double*>(v.ptr[tb]) + bo, Cp, v.ts, lds);
    } else {
      const double* A = static_cast<const double*>(v.ptr[ta]) + ao;
      if (fb) f64_block<BS>(A, static_cast<const float*>(v.ptr[tb]) + bo, Cp, v.ts, lds);
      else f64_block<BS>(A, static_cast<const double*>(v.ptr[tb]) + bo, Cp, v.ts, lds);
    }
  } else {
    const long tt = (long)v.ts * v.ts;
    const uint16_t* Ah = split + 2 * (long)m * tt + (long)bi * C::BM * 8;
    const uint16_t* Bh = split + 2 * (long)n * tt + (long)bj * C::BN * 8;
    fx4 acc[C::FM][C::FN];
    zero_acc_lp<C>(acc);
    gemm_nt_bf16d<C, ClassCfg<CLS, BS>::NSPLIT>(acc, Ah, Bh, tt, v.ts, v.ts, reinterpret_cast<uint16_t*>(lds));
    acc_add_lp<C, float>(acc, static_cast<float*>(v.ptr[tc]) + co, v.ts, -1.f);
  }
}

// One precision class per launch (tile sizes that are not a multiple of 192:
// the 128 x 128 and 64 x 64 block shapes).  Work-group w runs on XCD w % nx
// and the XCDs take contiguous slices of the class's block list.
template <int CLS, int BS>
__global__ void __launch_bounds__((ClassCfg<CLS, BS>::C::NT))
    k_update_class(TileView v, int k, const int* pairs, long nblocks, int nx, const uint16_t* split) {
  __shared__ __attribute__((aligned(16))) unsigned char lds[ClassCfg<CLS, BS>::LDS_BYTES];
  const int q = v.ts / BS;
  const int w = blockIdx.x;
  const int x = w % nx;
  const int gx = gridDim.x / nx;
  const long per = (nblocks + nx - 1) / nx;
  const long start = (long)x * per;
  const long end = start + per < nblocks ? start + per : nblocks;
  for (long b = start + w / nx; b < end; b += gx) {
    const long p = b / (q * q);
    const int rem = (int)(b - p * (q * q));
    const int bi = rem % q;
    const int bj = rem / q;
    const int pk = pairs[p];
    const int m = pk >> 16;
    const int n = pk & 0xFFFF;
    if (m == n && bi < bj) continue;  // strictly upper block of a diagonal tile
    class_update<CLS, BS>(v, k, m, n, bi, bj, split, lds);
  }
}

template <int CLS, int BS>
static void launch_class_bs(const TileView& v, int k, const int* pairs, long npairs, const uint16_t* split,
                            hipStream_t s) {
  static int cap = 0;
  if (!cap) {
    int dev = 0, cus = 0, occ = 1;
    EXA_HIP(hipGetDevice(&dev));
    EXA_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    EXA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_update_class<CLS, BS>, ClassCfg<CLS, BS>::C::NT, 0));
    int reserve = 16;
    if (const char* e = getenv("EXAGEOSTAT_RESERVE_CUS")) reserve = atoi(e);
    if (reserve < 0 || reserve >= cus) reserve = 0;
    cap = (cus - reserve) * (occ > 0 ? occ : 1);
    if (cap >= 8) cap -= cap % 8;
    if (cap < 1) cap = 1;
  }
  const long q = v.ts / BS;
  const long blocks = npairs * q * q;
  const int grid = blocks < cap ? (int)blocks : cap;
  const int nx = (grid % 8 == 0) ? 8 : 1;
  hipLaunchKernelGGL((k_update_class<CLS, BS>), dim3((unsigned)grid), dim3(ClassCfg<CLS, BS>::C::NT), 0, s, v, k,
                     pairs, blocks, nx, split);
  EXA_HIP(hipGetLastError());
}

template <int BS, int F32CLS>
static void launch_per_class(const TileView& v, int k, const FusedSpans& sp, const uint16_t* split, hipStream_t s) {
  for (int c = 0; c < 3; ++c) {
    if (sp.nblocks[c] <= 0) continue;
    if (sp.cls[c] == 0) launch_class_bs<kF64, BS>(v, k, sp.pairs[c], sp.nblocks[c], split, s);
    else if (sp.cls[c] == 1) launch_class_bs<F32CLS, BS>(v, k, sp.pairs[c], sp.nblocks[c], split, s);
    else launch_class_bs<kF32B, BS>(v, k, sp.pairs[c], sp.nblocks[c], split, s);
  }
}

// The trailing update of one step over all three precision classes in one
// persistent launch.  XCD x takes slice x of the FP64 class's block list,
// then slice x of the FP32 list, then of the far list (the slowest blocks
// first; contiguous slices keep the A and B panels hot in the XCD's L2), and
// its work-groups stride through that sequence.  One launch instead of one
// per class removes two launch tails per update.  Blocks are q x q per tile
// pair (packed m << 16 | n); the strictly upper blocks of diagonal tiles
// are skipped.
template <int BS, int F32CLS>
__global__ void __launch_bounds__((ClassCfg<kF64, BS>::C::NT))
    k_update_fused(TileView v, int k, FusedSpans sp, int nx, const uint16_t* split) {
  static_assert(ClassCfg<kF64, BS>::C::NT == ClassCfg<F32CLS, BS>::C::NT &&
                    ClassCfg<kF64, BS>::C::NT == ClassCfg<kF32B, BS>::C::NT,
                "the classes share one work-group shape");
  constexpr int LB = ClassCfg<F32CLS, BS>::LDS_BYTES > ClassCfg<kF64, BS>::LDS_BYTES
                         ? ClassCfg<F32CLS, BS>::LDS_BYTES
                         : ClassCfg<kF64, BS>::LDS_BYTES;
  __shared__ __attribute__((aligned(16))) unsigned char lds[LB];
  const int q = v.ts / BS;
  const int w = blockIdx.x;
  const int x = w % nx;
  const int gx = gridDim.x / nx;
  long st[3], cnt[3], tot = 0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const long per = (sp.nblocks[c] + nx - 1) / nx;
    st[c] = (long)x * per;
    const long e = st[c] + per < sp.nblocks[c] ? st[c] + per : sp.nblocks[c];
    cnt[c] = e > st[c] ? e - st[c] : 0;
    tot += cnt[c];
  }
  for (long i = w / nx; i < tot; i += gx) {
    int c = 0;
    long j = i;
    if (j >= cnt[0]) {
      j -= cnt[0];
      c = 1;
      if (j >= cnt[1]) {
        j -= cnt[1];
        c = 2;
      }
    }
    const long b = st[c] + j;
    const long p = b / (q * q);
    const int rem = (int)(b - p * (q * q));
    const int bi = rem % q;
    const int bj = rem / q;
    const int pk = sp.pairs[c][p];
    const int m = pk >> 16;
    const int n = pk & 0xFFFF;
    if (m == n && bi < bj) continue;
    const int cl = sp.cls[c];
    if (cl == 0) class_update<kF64, BS>(v, k, m, n, bi, bj, split, lds);
    else if (cl == 1) class_update<F32CLS, BS>(v, k, m, n, bi, bj, split, lds);
    else class_update<kF32B, BS>(v, k, m, n, bi, bj, split, lds);
  }
}

// Work-groups of the fused update: one per CU less the CUs left to the
// panel factorisation on the other stream (32 by default, 4 per XCD; with
// fewer the diagonal factorisation starves), times the occupancy.
template <int BS, int F32CLS>
static int fused_grid_cap() {
  static int cap = 0;
  if (cap) return cap;
  int dev = 0, cus = 0, occ = 1;
  EXA_HIP(hipGetDevice(&dev));
  EXA_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  EXA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_update_fused<BS, F32CLS>,
                                                       ClassCfg<kF64, BS>::C::NT, 0));
  int reserve = 32;
  if (const char* e = getenv("EXAGEOSTAT_RESERVE_CUS")) reserve = atoi(e);
  if (reserve < 0 || reserve >= cus) reserve = 0;
  cap = (cus - reserve) * (occ > 0 ? occ : 1);
  if (cap >= 8) cap -= cap % 8;
  if (cap < 1) cap = 1;
  return cap;
}

template <int BS, int F32CLS>
static void launch_fused_bs(const TileView& v, int k, const FusedSpans& sp0, const uint16_t* split, hipStream_t s) {
  const long q = v.ts / BS;
  FusedSpans sp = sp0;
  long total = 0;
  for (int c = 0; c < 3; ++c) {
    sp.nblocks[c] = sp0.nblocks[c] * q * q;  // pairs -> blocks
    total += sp.nblocks[c];
  }
  if (total <= 0) return;
  const int cap = fused_grid_cap<BS, F32CLS>();
  const int grid = total < cap ? (int)total : cap;
  const int nx = (grid % 8 == 0) ? 8 : 1;
  hipLaunchKernelGGL((k_update_fused<BS, F32CLS>), dim3((unsigned)grid), dim3(ClassCfg<kF64, BS>::C::NT), 0, s, v,
                     k, sp, nx, split);
  EXA_HIP(hipGetLastError());
}

// 192 x 192 blocks: one fused launch.  The 128 and 64 block shapes keep one
// launch per class: fused, every class inherits the FP64 path's register
// allocation (measured 10% slower at ts = 1,024).
template <int F32CLS>
static void launch_fused(const TileView& v, int k, const FusedSpans& sp, const uint16_t* split, hipStream_t s) {
  if (v.ts % 192 == 0) launch_fused_bs<192, F32CLS>(v, k, sp, split, s);
  else if (v.ts % 128 == 0) launch_per_class<128, F32CLS>(v, k, sp, split, s);
  else launch_per_class<64, F32CLS>(v, k, sp, split, s);
}

void launch_update_fused(const TileView& v, int k, const FusedSpans& sp, const uint16_t* split, bool lolo,
                         hipStream_t s) {
  for (int c = 0; c < 3; ++c)
    if (sp.nblocks[c] > 0 && sp.cls[c] != 0 && split == nullptr)
      throw std::invalid_argument("BF16 update classes need the split panel");
  if (lolo) launch_fused<kF32Q>(v, k, sp, split, s);
  else launch_fused<kF32>(v, k, sp, split, s);
}

// ------------------------------------------------ symmetric product (IR) --
// y = Sigma x for a lower-packed FP64 matrix whose diagonal tiles are full.
// One work-group per lower tile (m, n) reads the tile once for both
//   partN[n][m*ts + i] = sum_j A(m,n)[i][j] x_n[j]   (m >= n) and
//   partT[m][n*ts + j] = sum_i A(m,n)[i][j] x_m[i]   (m > n),
// with partN = part, partT = part + mt*len, len = mt*ts.  Wave g takes a
// quarter of the columns, 16 at a time (16 loads in flight per lane); lane
// r takes rows r, r + 64, ...  Work-group b takes tile p_off + b of the
// row-by-row enumeration p = m (m + 1) / 2 + n.
__global__ void __launch_bounds__(256) k_symv_tile(TileView v, const double* x, double* part, long p_off) {
  __shared__ double xn[1024], xm[1024];
  __shared__ double red[4][1024];
  const int tid = threadIdx.x;
  const int r = tid & 63;
  const int g = tid >> 6;
  const int ts = v.ts;
  const int q = ts / kNB;
  int m, n;
  tri_decode(p_off + blockIdx.x, m, n);
  const long len = (long)v.mt * ts;
  for (int i = tid; i < ts; i += 256) {
    xn[i] = x[(long)n * ts + i];
    xm[i] = x[(long)m * ts + i];
  }
  __syncthreads();
  const double* A = tile_f64(v, m, n);
  double racc[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) racc[i] = 0.0;
  const int cw = ts / 4;
  for (int c0 = g * cw; c0 < (g + 1) * cw; c0 += 16) {
    double xc[16], cacc[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      xc[j] = xn[c0 + j];
      cacc[j] = 0.0;
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      if (i < q) {
        const int row = r + kNB * i;
        double a[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) a[j] = A[row + (long)(c0 + j) * ts];
        const double xr = xm[row];
        double rs0 = 0.0, rs1 = 0.0;
#pragma unroll
        for (int j = 0; j < 16; j += 2) {
          rs0 += a[j] * xc[j];
          rs1 += a[j + 1] * xc[j + 1];
          cacc[j] += a[j] * xr;
          cacc[j + 1] += a[j + 1] * xr;
        }
        racc[i] += rs0 + rs1;
      }
    }
    if (m > n) {
#pragma unroll
      for (int j = 0; j < 16; ++j) cacc[j] = wave_sum(cacc[j]);
      if (r == 0) {
        double* out = part + (long)(v.mt + m) * len + (long)n * ts + c0;
#pragma unroll
        for (int j = 0; j < 16; ++j) out[j] = cacc[j];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 16; ++i)
    if (i < q) red[g][r + kNB * i] = racc[i];
  __syncthreads();
  double* outn = part + (long)n * len + (long)m * ts;
  for (int i = tid; i < ts; i += 256) outn[i] = (red[0][i] + red[1][i]) + (red[2][i] + red[3][i]);
}

// y[i] (row i of tile p) = sum_{n <= p} partN[n][i] + sum_{m > p} partT[m][i].
__global__ void k_symv_reduce(const double* part, int mt, int ts, long rows, double* y) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  const long len = (long)mt * ts;
  const int p = (int)(i / ts);
  double s = 0.0;
  for (int n = 0; n <= p; ++n) s += part[(long)n * len + i];
  for (int m = p + 1; m < mt; ++m) s += part[(long)(mt + m) * len + i];
  y[i] = s;
}

void launch_symv_tiles(const TileView& v, const double* x, double* part, long p0, long ntiles, hipStream_t s) {
  if (!v.lower || v.ts > 1024 || v.ts % kNB) throw std::invalid_argument("symv: lower matrix, ts <= 1024, ts % 64 == 0");
  if (p0 < 0 || ntiles < 0 || p0 + ntiles > (long)v.mt * (v.mt + 1) / 2) throw std::invalid_argument("symv: tile range");
  if (ntiles == 0) return;
  hipLaunchKernelGGL(k_symv_tile, dim3((unsigned)ntiles), dim3(256), 0, s, v, x, part, p0);
  EXA_HIP(hipGetLastError());
}

void launch_symv_reduce(const TileView& v, const double* part, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_symv_reduce, dim3((unsigned)((v.m + 255) / 256)), dim3(256), 0, s, part, v.mt, v.ts,
                     (long)v.m, y);
  EXA_HIP(hipGetLastError());
}

void launch_symv(const TileView& v, const double* x, double* y, double* part, hipStream_t s) {
  launch_symv_tiles(v, x, part, 0, (long)v.mt * (v.mt + 1) / 2, s);
  launch_symv_reduce(v, part, y, s);
}

__global__ void k_axpby(double a, const double* x, double b, const double* y, double* out, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a * x[i] + b * y[i];
}

void launch_axpby(double a, const double* x, double b, const double* y, double* out, long n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axpby, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, x, b, y, out, n);
  EXA_HIP(hipGetLastError());
}

}  // namespace exa
