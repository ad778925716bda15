// FP64 MFMA work-group GEMM core for CDNA4 (gfx950).
//
// acc += sign * A * B^T, A: BM x K, B: BN x K, both column-major in global
// memory.  A and B are staged through LDS in BK-deep slices, either with
// register-staged double buffering (gemm_nt_acc: global loads of slice k+1
// are in flight while slice k feeds the matrix cores; FP32 and K-major
// operands) or with direct-to-LDS copies into a three-stage ring
// (gemm_nt_ring: plain FP64 operands).  Each 64-lane wave owns a WM x WN
// sub-block made of 16x16 fragments of v_mfma_f64_16x16x4f64.
#pragma once

#include "common.hpp"

#include <cstdint>

namespace exa {

typedef double dx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ dx4 mfma_f64(double a, double b, dx4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// v_mfma_f64_16x16x4f64 operand/result layouts (64 lanes):
//   A (16x4): lane l supplies A[l%16][l/16];  B (4x16): B[l/16][l%16];
//   D (16x16): lane l, element v holds D[l/16 + 4*v][l%16]
//   (measured on gfx950 with an in-bounds single-wave probe).
__device__ __forceinline__ int frag_row(int lane, int v) { return (lane >> 4) + 4 * v; }
__device__ __forceinline__ int frag_col(int lane) { return lane & 15; }

template <int BM_, int BN_, int WGM_, int WGN_, int BK_>
struct GemmCfg {
  static constexpr int BM = BM_, BN = BN_, WGM = WGM_, WGN = WGN_, BK = BK_;
  static constexpr int NW = WGM * WGN;
  static constexpr int NT = NW * 64;
  static constexpr int WM = BM / WGM, WN = BN / WGN;
  static constexpr int FM = WM / 16, FN = WN / 16;
  static constexpr int PAD = 4;
  static constexpr int LDA = BM + PAD, LDB = BN + PAD;
  static constexpr int A_STAGE = BK * LDA, B_STAGE = BK * LDB;
  static constexpr int LDS_DOUBLES = 2 * (A_STAGE + B_STAGE);
  static constexpr int A_LOADS = BM * BK / 2 / NT;
  static constexpr int B_LOADS = BN * BK / 2 / NT;
  static_assert(WM % 16 == 0 && WN % 16 == 0, "wave tile must be 16-aligned");
  static_assert(A_LOADS * NT * 2 == BM * BK, "A slice must split evenly");
  static_assert(B_LOADS * NT * 2 == BN * BK, "B slice must split evenly");
};

// Loads through an explicit global (address space 1) pointer: the tile
// pointers come from a table, so without the cast the compiler emits flat
// loads whose completion it can only wait for together with LDS traffic.
__device__ __forceinline__ double2 ld_global2(const double* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef __attribute__((address_space(1))) const double2 gdouble2;
  return *reinterpret_cast<gdouble2*>(reinterpret_cast<uintptr_t>(p));
#else
  return *reinterpret_cast<const double2*>(p);
#endif
}

__device__ __forceinline__ float2 ld_global_f2(const float* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef __attribute__((address_space(1))) const float2 gfloat2;
  return *reinterpret_cast<gfloat2*>(reinterpret_cast<uintptr_t>(p));
#else
  return *reinterpret_cast<const float2*>(p);
#endif
}

// Two consecutive elements of a tile stored as T, loaded raw (Pair<T>) and
// widened to FP64 only when stored to LDS, so the loads of the next slice
// stay in flight while the current one is multiplied.
template <class T>
struct Pair;
template <>
struct Pair<double> {
  using type = double2;
};
template <>
struct Pair<float> {
  using type = float2;
};
__device__ __forceinline__ double2 ld_raw(const double* p) { return ld_global2(p); }
__device__ __forceinline__ float2 ld_raw(const float* p) { return ld_global_f2(p); }
__device__ __forceinline__ double2 widen(double2 v) { return v; }
__device__ __forceinline__ double2 widen(float2 v) { return double2{(double)v.x, (double)v.y}; }

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) void glb_void_t;

// Waits until at most n vector-memory operations of this wave are pending
// (expcnt and lgkmcnt left at their maxima).  For the direct-to-LDS copies
// the counter drops when the data is in LDS.
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

template <class C>
__device__ __forceinline__ void zero_acc(dx4 (&acc)[C::FM][C::FN]) {
#pragma unroll
  for (int i = 0; i < C::FM; ++i)
#pragma unroll
    for (int j = 0; j < C::FN; ++j) acc[i][j] = dx4{0.0, 0.0, 0.0, 0.0};
}

// K must be a multiple of BK; lda/ldb even and A/B 16-byte aligned.
// B_KMAJOR: B is given as K x BN column-major (B(k, c) at B[k + c*ldb])
// instead of BN x K; A_KMAJOR likewise for A (A(r, k) at A[k + r*lda]).
// TA/TB: storage type of A/B (FP32 tiles of the mixed-precision factor are
// widened on load).  Ends with a barrier, so the caller may reuse the LDS.
template <class C, bool B_KMAJOR = false, class TA = double, class TB = double, bool A_KMAJOR = false>
__device__ __forceinline__ void gemm_nt_acc(dx4 (&acc)[C::FM][C::FN], const TA* __restrict__ A, long lda,
                                            const TB* __restrict__ B, long ldb, int K, double* lds,
                                            double sign) {
  if (K <= 0) return;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave / C::WGN;
  const int wn = wave - wm * C::WGN;
  double* As = lds;
  double* Bs = lds + 2 * C::A_STAGE;
  typename Pair<TA>::type ra[C::A_LOADS];
  typename Pair<TB>::type rb[C::B_LOADS];

  auto gload = [&](int k0) {
#pragma unroll
    for (int t = 0; t < C::A_LOADS; ++t) {
      const int e = tid + t * C::NT;
      if (A_KMAJOR) {
        const int r = e / (C::BK / 2);
        const int k2 = e - r * (C::BK / 2);
        ra[t] = ld_raw(A + (long)r * lda + k0 + 2 * k2);
      } else {
        const int c = e / (C::BM / 2);
        const int r2 = e - c * (C::BM / 2);
        ra[t] = ld_raw(A + (long)(k0 + c) * lda + 2 * r2);
      }
    }
#pragma unroll
    for (int t = 0; t < C::B_LOADS; ++t) {
      const int e = tid + t * C::NT;
      if (B_KMAJOR) {
        const int c = e / (C::BK / 2);
        const int k2 = e - c * (C::BK / 2);
        rb[t] = ld_raw(B + (long)c * ldb + k0 + 2 * k2);
      } else {
        const int c = e / (C::BN / 2);
        const int r2 = e - c * (C::BN / 2);
        rb[t] = ld_raw(B + (long)(k0 + c) * ldb + 2 * r2);
      }
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int t = 0; t < C::A_LOADS; ++t) {
      const int e = tid + t * C::NT;
      if (A_KMAJOR) {
        const int r = e / (C::BK / 2);
        const int k2 = e - r * (C::BK / 2);
        const double2 w = widen(ra[t]);
        As[buf * C::A_STAGE + (2 * k2) * C::LDA + r] = w.x;
        As[buf * C::A_STAGE + (2 * k2 + 1) * C::LDA + r] = w.y;
      } else {
        const int c = e / (C::BM / 2);
        const int r2 = e - c * (C::BM / 2);
        *reinterpret_cast<double2*>(As + buf * C::A_STAGE + c * C::LDA + 2 * r2) = widen(ra[t]);
      }
    }
#pragma unroll
    for (int t = 0; t < C::B_LOADS; ++t) {
      const int e = tid + t * C::NT;
      if (B_KMAJOR) {
        const int c = e / (C::BK / 2);
        const int k2 = e - c * (C::BK / 2);
        const double2 w = widen(rb[t]);
        Bs[buf * C::B_STAGE + (2 * k2) * C::LDB + c] = w.x;
        Bs[buf * C::B_STAGE + (2 * k2 + 1) * C::LDB + c] = w.y;
      } else {
        const int c = e / (C::BN / 2);
        const int r2 = e - c * (C::BN / 2);
        *reinterpret_cast<double2*>(Bs + buf * C::B_STAGE + c * C::LDB + 2 * r2) = widen(rb[t]);
      }
    }
  };

  const int nk = K / C::BK;
  gload(0);
  sstore(0);
  __syncthreads();
  for (int kc = 0; kc < nk; ++kc) {
    const int buf = kc & 1;
    if (kc + 1 < nk) gload((kc + 1) * C::BK);
    const double* Ab = As + buf * C::A_STAGE + wm * C::WM + (lane & 15);
    const double* Bb = Bs + buf * C::B_STAGE + wn * C::WN + (lane & 15);
#pragma unroll
    for (int kk = 0; kk < C::BK; kk += 4) {
      const int krow = kk + (lane >> 4);
      double a[C::FM], b[C::FN];
#pragma unroll
      for (int i = 0; i < C::FM; ++i) a[i] = sign * Ab[krow * C::LDA + i * 16];
#pragma unroll
      for (int j = 0; j < C::FN; ++j) b[j] = Bb[krow * C::LDB + j * 16];
#pragma unroll
      for (int i = 0; i < C::FM; ++i)
#pragma unroll
        for (int j = 0; j < C::FN; ++j) acc[i][j] = mfma_f64(a[i], b[j], acc[i][j]);
    }
    // Keep the LDS stores of the prefetched slice after the matrix-core work
    // (otherwise they are hoisted next to the global loads and every slice
    // waits for memory before its MFMAs).
    __builtin_amdgcn_sched_barrier(0);
    if (kc + 1 < nk) sstore(buf ^ 1);
    __syncthreads();
  }
}

// The same product with the operands copied straight into a three-stage LDS
// ring by global_load_lds_dwordx4 (no staging registers, no LDS stores from
// registers), as the BF16 classes do.  Stage layout: A as BK columns of BM
// doubles (column-major, no padding: each 16-lane quarter of a fragment
// read takes 16 consecutive doubles), then B likewise; a wave-wide copy
// fills 1,024 contiguous LDS bytes, lane l taking the two doubles
// 128 q + 2 l of its operand's stage array, gathered from their (row,
// column) in the tile.  At step kc each wave waits for its own copies of
// slice kc, a bare barrier makes every wave's copies visible and ends the
// reads of slice kc-1, whose buffer then receives slice kc+2.  Plain FP64
// operands (not K-major); K a multiple of BK; A, B 16-byte aligned, lda and
// ldb even.
template <class C>
struct RingLds {
  static constexpr int A_ST = C::BK * C::BM, B_ST = C::BK * C::BN;  // doubles per stage
  static constexpr int STAGE = A_ST + B_ST;
  static constexpr int DOUBLES = 3 * STAGE;
  static constexpr int A_COPIES = A_ST / 128, B_COPIES = B_ST / 128;
  static constexpr int PER_WAVE = (A_COPIES + B_COPIES) / C::NW;
  static_assert(A_ST % 128 == 0 && B_ST % 128 == 0, "stage arrays must be whole 1 KB copies");
  static_assert(PER_WAVE * C::NW == A_COPIES + B_COPIES, "copies must split evenly over the waves");
  static_assert(PER_WAVE < 32, "vmcnt range");
};

template <class C>
__device__ __forceinline__ void gemm_nt_ring(dx4 (&acc)[C::FM][C::FN], const double* __restrict__ A, long lda,
                                             const double* __restrict__ B, long ldb, int K, double* lds,
                                             double sign) {
  using L = RingLds<C>;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave / C::WGN;
  const int wn = wave - wm * C::WGN;
  const int wave_s = __builtin_amdgcn_readfirstlane(wave);
  auto issue = [&](int kc, int buf) {
    const int k0 = kc * C::BK;
#pragma unroll
    for (int j = 0; j < L::PER_WAVE; ++j) {
      int c = wave_s * L::PER_WAVE + j;
      const bool is_b = c >= L::A_COPIES;
      if (is_b) c -= L::A_COPIES;
      const int rows = is_b ? C::BN : C::BM;
      const int e = c * 128 + 2 * lane;  // element of the operand's stage array
      const int kk = e / rows;
      const int r = e - kk * rows;
      const double* src = is_b ? B + r + (long)(k0 + kk) * ldb : A + r + (long)(k0 + kk) * lda;
      double* dst = lds + buf * L::STAGE + (is_b ? L::A_ST : 0) + c * 128;
      __builtin_amdgcn_global_load_lds((glb_void_t*)(uintptr_t)src, (lds_void_t*)dst, 16, 0, 0);
    }
  };
  auto compute = [&](int buf) {
    const double* Ab = lds + buf * L::STAGE + wm * C::WM + (lane & 15);
    const double* Bb = lds + buf * L::STAGE + L::A_ST + wn * C::WN + (lane & 15);
#pragma unroll
    for (int kk = 0; kk < C::BK; kk += 4) {
      const int krow = kk + (lane >> 4);
      double a[C::FM], b[C::FN];
#pragma unroll
      for (int i = 0; i < C::FM; ++i) a[i] = sign * Ab[krow * C::BM + i * 16];
#pragma unroll
      for (int j = 0; j < C::FN; ++j) b[j] = Bb[krow * C::BN + j * 16];
#pragma unroll
      for (int i = 0; i < C::FM; ++i)
#pragma unroll
        for (int j = 0; j < C::FN; ++j) acc[i][j] = mfma_f64(a[i], b[j], acc[i][j]);
    }
  };
  const int nk = K / C::BK;
  if (nk <= 0) return;
  // The ring may still be read by the previous block's last steps.
  __builtin_amdgcn_s_barrier();
  issue(0, 0);
  if (nk > 1) issue(1, 1);
  int buf = 0;
  for (int kc = 0; kc < nk; ++kc) {
    if (kc + 1 < nk) wait_vmcnt<L::PER_WAVE>();
    else wait_vmcnt<0>();
    __builtin_amdgcn_s_barrier();
    if (kc + 2 < nk) issue(kc + 2, buf == 0 ? 2 : buf - 1);
    compute(buf);
    buf = buf == 2 ? 0 : buf + 1;
  }
}

// C[r + c*ldc] = beta*C + acc for the block owned by this work-group.
template <class C, class T = double>
__device__ __forceinline__ void acc_update(dx4 (&acc)[C::FM][C::FN], T* Cp, long ldc, double beta) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave / C::WGN;
  const int wn = wave - wm * C::WGN;
  // Global-address-space accesses, one fragment row at a time with all its
  // reads issued before its writes: with generic read-modify-writes the
  // compiler must assume that a store may alias the next load and
  // serialises the block's memory traffic.
  typedef __attribute__((address_space(1))) T gT;
  gT* G = (gT*)(uintptr_t)Cp;
#pragma unroll
  for (int i = 0; i < C::FM; ++i) {
    if (beta != 0.0) {
      T cv[C::FN][4];
#pragma unroll
      for (int j = 0; j < C::FN; ++j) {
        const int c = wn * C::WN + j * 16 + frag_col(lane);
#pragma unroll
        for (int v = 0; v < 4; ++v) cv[j][v] = G[wm * C::WM + i * 16 + frag_row(lane, v) + (long)c * ldc];
      }
#pragma unroll
      for (int j = 0; j < C::FN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[i][j][v] = beta * (double)cv[j][v] + acc[i][j][v];
    }
#pragma unroll
    for (int j = 0; j < C::FN; ++j) {
      const int c = wn * C::WN + j * 16 + frag_col(lane);
#pragma unroll
      for (int v = 0; v < 4; ++v) G[wm * C::WM + i * 16 + frag_row(lane, v) + (long)c * ldc] = (T)acc[i][j][v];
    }
  }
}

// Cp[c + r*ldc] = acc(r, c): the block stored transposed.  Each lane keeps
// one base address; the fragment offsets are uniform (a multiple of ldc plus
// a constant), so the stores need no per-element address registers next to
// the live accumulators.
template <class C>
__device__ __forceinline__ void acc_store_t(const dx4 (&acc)[C::FM][C::FN], double* Cp, long ldc) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave / C::WGN;
  const int wn = wave - wm * C::WGN;
  typedef __attribute__((address_space(1))) double gD;
  gD* G = (gD*)(uintptr_t)(Cp + wn * C::WN + frag_col(lane) + (long)(wm * C::WM + frag_row(lane, 0)) * ldc);
#pragma unroll
  for (int i = 0; i < C::FM; ++i)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const long ro = (long)(i * 16 + 4 * v) * ldc;
#pragma unroll
      for (int j = 0; j < C::FN; ++j) G[ro + j * 16] = acc[i][j][v];
    }
}

template <class C, class T = double>
__device__ __forceinline__ void acc_load(dx4 (&acc)[C::FM][C::FN], const T* Cp, long ldc) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave / C::WGN;
  const int wn = wave - wm * C::WGN;
#pragma unroll
  for (int i = 0; i < C::FM; ++i)
#pragma unroll
    for (int j = 0; j < C::FN; ++j) {
      const int c = wn * C::WN + j * 16 + frag_col(lane);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = wm * C::WM + i * 16 + frag_row(lane, v);
        acc[i][j][v] = (double)Cp[r + (long)c * ldc];
      }
    }
}

// ---------------------------------------------------- FP32 matrix cores --
// v_mfma_f32_16x16x4f32: operands as for the FP64 instruction (lane l
// supplies A[l%16][l/16] and B[l/16][l%16]); result: lane l, element v holds
// D[4*(l/16)+v][l%16].
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <class C>
__device__ __forceinline__ void zero_acc32(f32x4 (&acc)[C::FM][C::FN]) {
#pragma unroll
  for (int i = 0; i < C::FM; ++i)
#pragma unroll
    for (int j = 0; j < C::FN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
}

// acc += sign * A * B^T with FP64 operands rounded to FP32 when staged into
// LDS (the register-staged double buffering of gemm_nt_acc, plain operand
// layout) and FP32 matrix cores at twice the FP64 rate, FP32 accumulation.
// lds: C::LDS_DOUBLES doubles, used as twice as many floats' worth of half.
template <class C>
__device__ __forceinline__ void gemm_nt_acc32(f32x4 (&acc)[C::FM][C::FN], const double* __restrict__ A, long lda,
                                              const double* __restrict__ B, long ldb, int K, float* lds,
                                              float sign) {
  if (K <= 0) return;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave / C::WGN;
  const int wn = wave - wm * C::WGN;
  float* As = lds;
  float* Bs = lds + 2 * C::A_STAGE;
  double2 ra[C::A_LOADS];
  double2 rb[C::B_LOADS];
  auto gload = [&](int k0) {
#pragma unroll
    for (int t = 0; t < C::A_LOADS; ++t) {
      const int e = tid + t * C::NT;
      const int c = e / (C::BM / 2);
      const int r2 = e - c * (C::BM / 2);
      ra[t] = ld_global2(A + (long)(k0 + c) * lda + 2 * r2);
    }
#pragma unroll
    for (int t = 0; t < C::B_LOADS; ++t) {
      const int e = tid + t * C::NT;
      const int c = e / (C::BN / 2);
      const int r2 = e - c * (C::BN / 2);
      rb[t] = ld_global2(B + (long)(k0 + c) * ldb + 2 * r2);
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int t = 0; t < C::A_LOADS; ++t) {
      const int e = tid + t * C::NT;
      const int c = e / (C::BM / 2);
      const int r2 = e - c * (C::BM / 2);
      *reinterpret_cast<float2*>(As + buf * C::A_STAGE + c * C::LDA + 2 * r2) =
          float2{(float)ra[t].x, (float)ra[t].y};
    }
#pragma unroll
    for (int t = 0; t < C::B_LOADS; ++t) {
      const int e = tid + t * C::NT;
      const int c = e / (C::BN / 2);
      const int r2 = e - c * (C::BN / 2);
      *reinterpret_cast<float2*>(Bs + buf * C::B_STAGE + c * C::LDB + 2 * r2) =
          float2{(float)rb[t].x, (float)rb[t].y};
    }
  };
  const int nk = K / C::BK;
  gload(0);
  sstore(0);
  __syncthreads();
  for (int kc = 0; kc < nk; ++kc) {
    const int buf = kc & 1;
    if (kc + 1 < nk) gload((kc + 1) * C::BK);
    const float* Ab = As + buf * C::A_STAGE + wm * C::WM + (lane & 15);
    const float* Bb = Bs + buf * C::B_STAGE + wn * C::WN + (lane & 15);
#pragma unroll
    for (int kk = 0; kk < C::BK; kk += 4) {
      const int krow = kk + (lane >> 4);
      float a[C::FM], b[C::FN];
#pragma unroll
      for (int i = 0; i < C::FM; ++i) a[i] = sign * Ab[krow * C::LDA + i * 16];
#pragma unroll
      for (int j = 0; j < C::FN; ++j) b[j] = Bb[krow * C::LDB + j * 16];
#pragma unroll
      for (int i = 0; i < C::FM; ++i)
#pragma unroll
        for (int j = 0; j < C::FN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
    if (kc + 1 < nk) sstore(buf ^ 1);
    __syncthreads();
  }
}

// Cp[r + c*ldc] = acc(r, c) as FP64 (the FP32 result layout).
template <class C>
__device__ __forceinline__ void acc32_store(const f32x4 (&acc)[C::FM][C::FN], double* Cp, long ldc) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave / C::WGN;
  const int wn = wave - wm * C::WGN;
  typedef __attribute__((address_space(1))) double gD;
  gD* G = (gD*)(uintptr_t)(Cp + wm * C::WM + 4 * (lane >> 4) + (long)(wn * C::WN + (lane & 15)) * ldc);
#pragma unroll
  for (int i = 0; i < C::FM; ++i)
#pragma unroll
    for (int j = 0; j < C::FN; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) G[i * 16 + v + (long)j * 16 * ldc] = (double)acc[i][j][v];
}

// Cp[c + r*ldc] = acc(r, c) as FP64: the block stored transposed.
template <class C>
__device__ __forceinline__ void acc32_store_t(const f32x4 (&acc)[C::FM][C::FN], double* Cp, long ldc) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave / C::WGN;
  const int wn = wave - wm * C::WGN;
  typedef __attribute__((address_space(1))) double gD;
  gD* G = (gD*)(uintptr_t)(Cp + wn * C::WN + (lane & 15) + (long)(wm * C::WM + 4 * (lane >> 4)) * ldc);
#pragma unroll
  for (int i = 0; i < C::FM; ++i)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const long ro = (long)(i * 16 + v) * ldc;
#pragma unroll
      for (int j = 0; j < C::FN; ++j) G[ro + j * 16] = (double)acc[i][j][v];
    }
}

}  // namespace exa
