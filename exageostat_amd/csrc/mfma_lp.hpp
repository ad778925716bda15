// Lower-precision MFMA work-group GEMM core for the mixed-precision tile
// Cholesky on CDNA4 (gfx950): BF16 operands (single or split in two BF16
// terms) with FP32 accumulation (v_mfma_f32_16x16x32_bf16).
//
// It computes acc -= A * B^T for a BM x BN block, A: BM x K and B: BN x K
// column-major tiles stored in FP64 or FP32 (narrowed on load), staged
// through LDS with double buffering like the FP64 core in mfma.hpp.
// Result layout (measured on gfx950 with an in-bounds single-wave probe):
// lane l, element v holds D[4*(l/16)+v][l%16]; operand layout: lane l holds
// A[l%16][8*(l/16)+e] and B^T[l%16][8*(l/16)+e], e = 0..7.
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

// ------------------------------------------------------- split BF16 --
// acc -= A * B^T on the BF16 matrix cores (v_mfma_f32_16x16x32_bf16, FP32
// accumulation).  NSPLIT = 1: operands rounded to BF16.  NSPLIT = 3: each
// operand is split as x = hi + lo (hi = bf16(x), lo = bf16(x - hi)) and the
// product is hi*hi + hi*lo + lo*hi, which keeps about 16 significant bits of
// each operand (products of BF16 values are exact in FP32) at one third of
// the BF16 rate, several times the FP32 MFMA rate.
//
// LDS holds, per stage, [A_hi | B_hi | A_lo | B_lo] slices [row][k] with
// row stride BK + 8 (a lane's eight consecutive k values of one row are one
// 16-byte read; the 16-byte pad keeps the eight rows of a read phase on
// distinct banks).  Global loads come in units of 4 rows x 8 k (eight
// 16-byte column pieces), narrowed and split in registers; A units go to
// the first threads of the work-group, B units to the next ones.
#ifndef EXA_LDS_SWIZZLE_SHIFT
#define EXA_LDS_SWIZZLE_SHIFT 3
#endif
// 16-byte k-group kg of row `row` is stored at group kg ^ ((row >> S) & 3):
// the unit stores write rows 4 apart from consecutive lanes, which without
// the swizzle fall on the same LDS banks (shift 3 measured best on gfx950);
// the fragment reads (16 consecutive
// rows, one k-group) stay conflict-free.
__device__ __forceinline__ int lds_kgroup(int row, int kg) { return kg ^ ((row >> EXA_LDS_SWIZZLE_SHIFT) & 3); }

template <class C, int NSPLIT>
struct Bf16sLds {
  static_assert(NSPLIT == 1 || NSPLIT == 3, "one or three BF16 products");
  static_assert(C::BK % 32 == 0, "BK must be a multiple of the MFMA depth");
  static constexpr int LDK = C::BK + 8;
  static constexpr int A_ARR = C::BM * LDK, B_ARR = C::BN * LDK;  // bf16 elements per array
  static constexpr int STAGE = (NSPLIT == 3 ? 2 : 1) * (A_ARR + B_ARR);
  static constexpr int HALVES = 2 * STAGE;
  static constexpr int A_UNITS = (C::BM / 4) * (C::BK / 8);
  static constexpr int B_UNITS = (C::BN / 4) * (C::BK / 8);
  static constexpr int SLOTS = (A_UNITS + B_UNITS + C::NT - 1) / C::NT;
};

typedef float f2v __attribute__((ext_vector_type(2)));
typedef __bf16 b2v __attribute__((ext_vector_type(2)));

// Two FP32 values rounded to BF16 (hardware round-to-nearest-even, one
// packed conversion) and, for NSPLIT = 3, their BF16 remainders.
template <int NSPLIT>
__device__ __forceinline__ void split_pair(float a, float b, uint32_t& hi, uint32_t& lo) {
  const f2v v = {a, b};
  const b2v h = __builtin_convertvector(v, b2v);
  hi = __builtin_bit_cast(uint32_t, h);
  if (NSPLIT == 3) {
    const f2v back = __builtin_convertvector(h, f2v);
    lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(v - back, b2v));
  }
}

// One 4-row x 8-k unit of a column-major source (leading dimension ld):
// eight 16-byte column pieces, loaded raw (FP32, or FP64 narrowed on
// load) so that the loads of the next slice stay in flight while the
// current slice is multiplied; the BF16 conversion happens at store time.
template <class T>
__device__ __forceinline__ void load_unit_raw(const T* base, long ld, float4 (&q)[8]) {
#pragma unroll
  for (int kk = 0; kk < 8; ++kk) q[kk] = ld_quad(base + (long)kk * ld);
}

// The unit as BF16 rows (hi) and, for NSPLIT = 3, the BF16 remainders
// (lo); negate flips the sign bits of both (exact).
template <int NSPLIT>
__device__ __forceinline__ void split_unit(const float4 (&q)[8], bool negate, uint4 (&hi)[4], uint4 (&lo)[4]) {
  const uint32_t sign = negate ? 0x80008000u : 0u;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    uint32_t wh[4], wl[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const float4& a = q[2 * h];
      const float4& b = q[2 * h + 1];
      const float fa = r == 0 ? a.x : (r == 1 ? a.y : (r == 2 ? a.z : a.w));
      const float fb = r == 0 ? b.x : (r == 1 ? b.y : (r == 2 ? b.z : b.w));
      split_pair<NSPLIT>(fa, fb, wh[h], wl[h]);
      wh[h] ^= sign;
      wl[h] ^= sign;
    }
    hi[r] = uint4{wh[0], wh[1], wh[2], wh[3]};
    if (NSPLIT == 3) lo[r] = uint4{wl[0], wl[1], wl[2], wl[3]};
  }
}

template <class C, int NSPLIT, class TA, class TB>
__device__ __forceinline__ void gemm_nt_sub_bf16s(fx4 (&acc)[C::FM][C::FN], const TA* __restrict__ A, long lda,
                                                  const TB* __restrict__ B, long ldb, int K, uint16_t* lds) {
  using L = Bf16sLds<C, NSPLIT>;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave / C::WGN;
  const int wn = wave - wm * C::WGN;
  float4 raw[L::SLOTS][8];
  auto gload = [&](int k0) {
#pragma unroll
    for (int t = 0; t < L::SLOTS; ++t) {
      const int u = tid + t * C::NT;
      if (u < L::A_UNITS) {
        const int mq = u % (C::BM / 4), kg = u / (C::BM / 4);
        load_unit_raw(A + (long)(k0 + 8 * kg) * lda + 4 * mq, lda, raw[t]);
      } else if (u < L::A_UNITS + L::B_UNITS) {
        const int w = u - L::A_UNITS;
        const int mq = w % (C::BN / 4), kg = w / (C::BN / 4);
        load_unit_raw(B + (long)(k0 + 8 * kg) * ldb + 4 * mq, ldb, raw[t]);
      }
    }
  };
  auto sstore = [&](int buf) {
    uint16_t* st = lds + buf * L::STAGE;
#pragma unroll
    for (int t = 0; t < L::SLOTS; ++t) {
      const int u = tid + t * C::NT;
      int row, kg;
      uint16_t* hi_arr;
      uint16_t* lo_arr;
      uint4 rh[4], rl[4];
      if (u < L::A_UNITS) {
        row = 4 * (u % (C::BM / 4));
        kg = u / (C::BM / 4);
        hi_arr = st;
        lo_arr = st + L::A_ARR + L::B_ARR;
        split_unit<NSPLIT>(raw[t], true, rh, rl);
      } else if (u < L::A_UNITS + L::B_UNITS) {
        const int w = u - L::A_UNITS;
        row = 4 * (w % (C::BN / 4));
        kg = w / (C::BN / 4);
        hi_arr = st + L::A_ARR;
        lo_arr = st + 2 * L::A_ARR + L::B_ARR;
        split_unit<NSPLIT>(raw[t], false, rh, rl);
      } else {
        continue;
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int off = (row + r) * L::LDK + 8 * lds_kgroup(row + r, kg);
        *reinterpret_cast<uint4*>(hi_arr + off) = rh[r];
        if (NSPLIT == 3) *reinterpret_cast<uint4*>(lo_arr + off) = rl[r];
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
    const uint16_t* st = lds + buf * L::STAGE;
#pragma unroll
    for (int kk = 0; kk < C::BK; kk += 32) {
      const int kg = kk / 8 + (lane >> 4);
      int ao[C::FM], bo[C::FN];
#pragma unroll
      for (int i = 0; i < C::FM; ++i) {
        const int row = wm * C::WM + i * 16 + (lane & 15);
        ao[i] = row * L::LDK + 8 * lds_kgroup(row, kg);
      }
#pragma unroll
      for (int j = 0; j < C::FN; ++j) {
        const int row = wn * C::WN + j * 16 + (lane & 15);
        bo[j] = row * L::LDK + 8 * lds_kgroup(row, kg);
      }
      bf16x8 ah[C::FM], bh[C::FN];
#pragma unroll
      for (int i = 0; i < C::FM; ++i) ah[i] = *reinterpret_cast<const bf16x8*>(st + ao[i]);
#pragma unroll
      for (int j = 0; j < C::FN; ++j) bh[j] = *reinterpret_cast<const bf16x8*>(st + L::A_ARR + bo[j]);
      if (NSPLIT == 3) {
        bf16x8 al[C::FM], bl[C::FN];
#pragma unroll
        for (int i = 0; i < C::FM; ++i)
          al[i] = *reinterpret_cast<const bf16x8*>(st + L::A_ARR + L::B_ARR + ao[i]);
#pragma unroll
        for (int j = 0; j < C::FN; ++j)
          bl[j] = *reinterpret_cast<const bf16x8*>(st + 2 * L::A_ARR + L::B_ARR + bo[j]);
        // small cross terms first, then the leading product
#pragma unroll
        for (int i = 0; i < C::FM; ++i)
#pragma unroll
          for (int j = 0; j < C::FN; ++j) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          }
      }
#pragma unroll
      for (int i = 0; i < C::FM; ++i)
#pragma unroll
        for (int j = 0; j < C::FN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
    if (kc + 1 < nk) sstore(buf ^ 1);
    __syncthreads();
  }
}

}  // namespace exa
