// Lower-precision MFMA work-group GEMM cores for the mixed-precision tile
// Cholesky on CDNA4 (gfx950): FP32 (v_mfma_f32_16x16x4_f32) and BF16 with
// FP32 accumulation (v_mfma_f32_16x16x32_bf16).
//
// Both compute acc -= A * B^T for a BM x BN block, A: BM x K and B: BN x K
// column-major tiles stored in FP64 or FP32 (narrowed on load), staged
// through LDS with double buffering like the FP64 core in mfma.hpp.
// Result layout of both instructions (measured on gfx950 with an in-bounds
// single-wave probe): lane l, element v holds D[4*(l/16)+v][l%16].
#pragma once

#include "common.hpp"
#include "mfma.hpp"

namespace exa {

typedef float fx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ int frag_row_lp(int lane, int v) { return 4 * (lane >> 4) + v; }

__device__ __forceinline__ float4 ld_quad(const float* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef __attribute__((address_space(1))) const float4 gfloat4;
  return *reinterpret_cast<gfloat4*>(reinterpret_cast<uintptr_t>(p));
#else
  return *reinterpret_cast<const float4*>(p);
#endif
}
__device__ __forceinline__ float4 ld_quad(const double* p) {
  const double2 a = ld_global2(p);
  const double2 b = ld_global2(p + 2);
  return float4{(float)a.x, (float)a.y, (float)b.x, (float)b.y};
}

__device__ __forceinline__ uint16_t bf16_bits_rne(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFF) ? 0x40 : 0));
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

template <int BM_, int BN_, int WGM_, int WGN_, int BK_>
struct LpCfg {
  static constexpr int BM = BM_, BN = BN_, WGM = WGM_, WGN = WGN_, BK = BK_;
  static constexpr int NW = WGM * WGN;
  static constexpr int NT = NW * 64;
  static constexpr int WM = BM / WGM, WN = BN / WGN;
  static constexpr int FM = WM / 16, FN = WN / 16;
  static_assert(WM % 16 == 0 && WN % 16 == 0, "wave tile must be 16-aligned");
};

template <class C>
__device__ __forceinline__ void zero_acc_lp(fx4 (&acc)[C::FM][C::FN]) {
#pragma unroll
  for (int i = 0; i < C::FM; ++i)
#pragma unroll
    for (int j = 0; j < C::FN; ++j) acc[i][j] = fx4{0.f, 0.f, 0.f, 0.f};
}

// C[r + c*ldc] += acc (C stored as T).
template <class C, class T>
__device__ __forceinline__ void acc_add_lp(const fx4 (&acc)[C::FM][C::FN], T* Cp, long ldc) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave / C::WGN;
  const int wn = wave - wm * C::WGN;
#pragma unroll
  for (int i = 0; i < C::FM; ++i)
#pragma unroll
    for (int j = 0; j < C::FN; ++j) {
      const int c = wn * C::WN + j * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = wm * C::WM + i * 16 + frag_row_lp(lane, v);
        T* p = Cp + r + (long)c * ldc;
        *p = (T)((float)*p + acc[i][j][v]);
      }
    }
}

// ---------------------------------------------------------------- FP32 --
// LDS holds float slices [k][m] (row stride BM + 4); one slice = BK columns.
template <class C>
struct F32Lds {
  static constexpr int LDA = C::BM + 4, LDB = C::BN + 4;
  static constexpr int A_STAGE = C::BK * LDA, B_STAGE = C::BK * LDB;
  static constexpr int FLOATS = 2 * (A_STAGE + B_STAGE);
  static constexpr int A_LOADS = C::BM * C::BK / 4 / C::NT;
  static constexpr int B_LOADS = C::BN * C::BK / 4 / C::NT;
  static_assert(A_LOADS * C::NT * 4 == C::BM * C::BK, "A slice must split evenly");
  static_assert(B_LOADS * C::NT * 4 == C::BN * C::BK, "B slice must split evenly");
};

template <class C, class TA, class TB>
__device__ __forceinline__ void gemm_nt_sub_f32(fx4 (&acc)[C::FM][C::FN], const TA* __restrict__ A, long lda,
                                                const TB* __restrict__ B, long ldb, int K, float* lds) {
  using L = F32Lds<C>;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave / C::WGN;
  const int wn = wave - wm * C::WGN;
  float* As = lds;
  float* Bs = lds + 2 * L::A_STAGE;
  float4 ra[L::A_LOADS], rb[L::B_LOADS];
  auto gload = [&](int k0) {
#pragma unroll
    for (int t = 0; t < L::A_LOADS; ++t) {
      const int e = tid + t * C::NT;
      const int c = e / (C::BM / 4), r4 = e - c * (C::BM / 4);
      ra[t] = ld_quad(A + (long)(k0 + c) * lda + 4 * r4);
    }
#pragma unroll
    for (int t = 0; t < L::B_LOADS; ++t) {
      const int e = tid + t * C::NT;
      const int c = e / (C::BN / 4), r4 = e - c * (C::BN / 4);
      rb[t] = ld_quad(B + (long)(k0 + c) * ldb + 4 * r4);
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int t = 0; t < L::A_LOADS; ++t) {
      const int e = tid + t * C::NT;
      const int c = e / (C::BM / 4), r4 = e - c * (C::BM / 4);
      *reinterpret_cast<float4*>(As + buf * L::A_STAGE + c * L::LDA + 4 * r4) = ra[t];
    }
#pragma unroll
    for (int t = 0; t < L::B_LOADS; ++t) {
      const int e = tid + t * C::NT;
      const int c = e / (C::BN / 4), r4 = e - c * (C::BN / 4);
      *reinterpret_cast<float4*>(Bs + buf * L::B_STAGE + c * L::LDB + 4 * r4) = rb[t];
    }
  };
  const int nk = K / C::BK;
  gload(0);
  sstore(0);
  __syncthreads();
  for (int kc = 0; kc < nk; ++kc) {
    const int buf = kc & 1;
    if (kc + 1 < nk) gload((kc + 1) * C::BK);
    const float* Ab = As + buf * L::A_STAGE + wm * C::WM + (lane & 15);
    const float* Bb = Bs + buf * L::B_STAGE + wn * C::WN + (lane & 15);
#pragma unroll
    for (int kk = 0; kk < C::BK; kk += 4) {
      const int krow = kk + (lane >> 4);
      float a[C::FM], b[C::FN];
#pragma unroll
      for (int i = 0; i < C::FM; ++i) a[i] = -Ab[krow * L::LDA + i * 16];
#pragma unroll
      for (int j = 0; j < C::FN; ++j) b[j] = Bb[krow * L::LDB + j * 16];
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

// ---------------------------------------------------------------- BF16 --
// LDS holds bf16 slices [m][k] (row stride BK + 8), so a lane's eight
// consecutive k values of one row are one 16-byte read.  Global loads are
// units of 4 rows x 8 k (eight 4-row column pieces), narrowed to bf16 and
// written as four 16-byte rows; A is negated on the way.
template <class C>
struct Bf16Lds {
  static constexpr int LDK = C::BK + 8;
  static constexpr int A_STAGE = C::BM * LDK, B_STAGE = C::BN * LDK;  // in bf16 elements
  static constexpr int HALVES = 2 * (A_STAGE + B_STAGE);
  static constexpr int A_UNITS = (C::BM / 4) * (C::BK / 8);
  static constexpr int B_UNITS = (C::BN / 4) * (C::BK / 8);
  static constexpr int A_SLOTS = (A_UNITS + C::NT - 1) / C::NT;
  static constexpr int B_SLOTS = (B_UNITS + C::NT - 1) / C::NT;
};

template <class T>
__device__ __forceinline__ void load_unit_bf16(const T* base, long ld, uint4 (&rows)[4], bool negate) {
  float v[4][8];
#pragma unroll
  for (int kk = 0; kk < 8; ++kk) {
    const float4 q = ld_quad(base + (long)kk * ld);
    v[0][kk] = q.x;
    v[1][kk] = q.y;
    v[2][kk] = q.z;
    v[3][kk] = q.w;
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    uint32_t w[4];
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const float lo = negate ? -v[r][2 * h] : v[r][2 * h];
      const float hi = negate ? -v[r][2 * h + 1] : v[r][2 * h + 1];
      w[h] = (uint32_t)bf16_bits_rne(lo) | ((uint32_t)bf16_bits_rne(hi) << 16);
    }
    rows[r] = uint4{w[0], w[1], w[2], w[3]};
  }
}

template <class C, class TA, class TB>
__device__ __forceinline__ void gemm_nt_sub_bf16(fx4 (&acc)[C::FM][C::FN], const TA* __restrict__ A, long lda,
                                                 const TB* __restrict__ B, long ldb, int K, uint16_t* lds) {
  using L = Bf16Lds<C>;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave / C::WGN;
  const int wn = wave - wm * C::WGN;
  uint16_t* As = lds;
  uint16_t* Bs = lds + 2 * L::A_STAGE;
  uint4 ra[L::A_SLOTS][4], rb[L::B_SLOTS][4];
  auto gload = [&](int k0) {
#pragma unroll
    for (int t = 0; t < L::A_SLOTS; ++t) {
      const int e = tid + t * C::NT;
      if (e < L::A_UNITS) {
        const int mq = e % (C::BM / 4), kg = e / (C::BM / 4);
        load_unit_bf16(A + (long)(k0 + 8 * kg) * lda + 4 * mq, lda, ra[t], true);
      }
    }
#pragma unroll
    for (int t = 0; t < L::B_SLOTS; ++t) {
      const int e = tid + t * C::NT;
      if (e < L::B_UNITS) {
        const int mq = e % (C::BN / 4), kg = e / (C::BN / 4);
        load_unit_bf16(B + (long)(k0 + 8 * kg) * ldb + 4 * mq, ldb, rb[t], false);
      }
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int t = 0; t < L::A_SLOTS; ++t) {
      const int e = tid + t * C::NT;
      if (e < L::A_UNITS) {
        const int mq = e % (C::BM / 4), kg = e / (C::BM / 4);
#pragma unroll
        for (int r = 0; r < 4; ++r)
          *reinterpret_cast<uint4*>(As + buf * L::A_STAGE + (4 * mq + r) * L::LDK + 8 * kg) = ra[t][r];
      }
    }
#pragma unroll
    for (int t = 0; t < L::B_SLOTS; ++t) {
      const int e = tid + t * C::NT;
      if (e < L::B_UNITS) {
        const int mq = e % (C::BN / 4), kg = e / (C::BN / 4);
#pragma unroll
        for (int r = 0; r < 4; ++r)
          *reinterpret_cast<uint4*>(Bs + buf * L::B_STAGE + (4 * mq + r) * L::LDK + 8 * kg) = rb[t][r];
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
    const uint16_t* Ab = As + buf * L::A_STAGE + (wm * C::WM + (lane & 15)) * L::LDK + 8 * (lane >> 4);
    const uint16_t* Bb = Bs + buf * L::B_STAGE + (wn * C::WN + (lane & 15)) * L::LDK + 8 * (lane >> 4);
#pragma unroll
    for (int kk = 0; kk < C::BK; kk += 32) {
      bf16x8 a[C::FM], b[C::FN];
#pragma unroll
      for (int i = 0; i < C::FM; ++i) a[i] = *reinterpret_cast<const bf16x8*>(Ab + i * 16 * L::LDK + kk);
#pragma unroll
      for (int j = 0; j < C::FN; ++j) b[j] = *reinterpret_cast<const bf16x8*>(Bb + j * 16 * L::LDK + kk);
#pragma unroll
      for (int i = 0; i < C::FM; ++i)
#pragma unroll
        for (int j = 0; j < C::FN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
    if (kc + 1 < nk) sstore(buf ^ 1);
    __syncthreads();
  }
}

}  // namespace exa
