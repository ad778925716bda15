// Lower-precision MFMA work-group GEMM core for the mixed-precision tile
// Cholesky on CDNA4 (gfx950): BF16 operands (single or split in two BF16
// terms) with FP32 accumulation (v_mfma_f32_16x16x32_bf16).
//
// It computes acc += A * B^T for a BM x BN block from panel tiles that were
// converted once per step into BF16 parts (k_split_panel), staged through
// LDS with double buffering like the FP64 core in mfma.hpp.
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

// C[r + c*ldc] += sign * acc (C stored as T).
template <class C, class T>
__device__ __forceinline__ void acc_add_lp(const fx4 (&acc)[C::FM][C::FN], T* Cp, long ldc, float sign = 1.f) {
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
        *p = (T)((float)*p + sign * acc[i][j][v]);
      }
    }
}

// ------------------------------------------------------- split BF16 --
// The BF16 update classes multiply BF16 operands on the matrix cores
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation).  NSPLIT = 1: operands
// rounded to BF16.  NSPLIT = 3 or 4: each operand is split as x = hi + lo
// (hi = bf16(x), lo = bf16(x - hi), about 16 significant bits together) and
// the product is hi*hi + hi*lo + lo*hi (+ lo*lo for NSPLIT = 4); products of
// BF16 values are exact in FP32.  The lo*lo term is small (2^-16 relative)
// but not random: rows of neighbouring locations have strongly correlated
// remainders, so dropping it biases A*B^T, and through the Schur complements
// the log-determinant, in one direction.  NSPLIT = 4 runs at a quarter of
// the BF16 rate, several times the FP32 MFMA rate.

typedef float f2v __attribute__((ext_vector_type(2)));
typedef __bf16 b2v __attribute__((ext_vector_type(2)));

// Two FP32 values rounded to BF16 (hardware round-to-nearest-even, one
// packed conversion) and, for NSPLIT >= 3, their BF16 remainders.
template <int NSPLIT>
__device__ __forceinline__ void split_pair(float a, float b, uint32_t& hi, uint32_t& lo) {
  const f2v v = {a, b};
  const b2v h = __builtin_convertvector(v, b2v);
  hi = __builtin_bit_cast(uint32_t, h);
  if (NSPLIT >= 3) {
    const f2v back = __builtin_convertvector(h, f2v);
    lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(v - back, b2v));
  }
}

// ---------------------------------------------------- pre-split panels --
// Panel tiles converted once per step into BF16 hi (and lo) parts, laid out
// per tile as [k-group g][row r][8 k values] (g = column / 8): one 16-byte
// load is eight k values of one row, consecutive rows are consecutive in
// memory, and consecutive lanes store consecutive LDS rows (no conflicts).
// A part of tile (row block at ro) is at base + ro * 8 with group stride
// ts * 8 elements.
template <class C, int NSPLIT>
struct Bf16pLds {
  static_assert(C::BK % 32 == 0, "BK must be a multiple of the MFMA depth");
  static constexpr int LDK = C::BK + 8;
  static constexpr int A_ARR = C::BM * LDK, B_ARR = C::BN * LDK;
  static constexpr int NARR = NSPLIT >= 3 ? 2 : 1;
  static constexpr int STAGE = NARR * (A_ARR + B_ARR);
  static constexpr int HALVES = 2 * STAGE;
  static constexpr int G = C::BK / 8;                          // k-groups per slice
  static constexpr int UNITS = NARR * G * (C::BM + C::BN);     // 16-byte pieces per slice
  static constexpr int SLOTS = (UNITS + C::NT - 1) / C::NT;
};

__device__ __forceinline__ uint4 ld_global16(const uint16_t* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef __attribute__((address_space(1))) const uint4 guint4;
  return *reinterpret_cast<guint4*>(reinterpret_cast<uintptr_t>(p));
#else
  return *reinterpret_cast<const uint4*>(p);
#endif
}

// acc += A * B^T from pre-split operands: Ah/Al (rows of A, hi/lo parts;
// lo unused for NSPLIT = 1) and Bh/Bl, group stride gs rows.  A and B come
// from the same panel buffer, so the caller subtracts the result.
template <class C, int NSPLIT>
__device__ __forceinline__ void gemm_nt_bf16p(fx4 (&acc)[C::FM][C::FN], const uint16_t* __restrict__ Ah,
                                                  const uint16_t* __restrict__ Al, const uint16_t* __restrict__ Bh,
                                                  const uint16_t* __restrict__ Bl, long gs, int K, uint16_t* lds) {
  using L = Bf16pLds<C, NSPLIT>;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave / C::WGN;
  const int wn = wave - wm * C::WGN;
  uint4 reg[L::SLOTS];
  // unit u: array a (0: A hi, 1: B hi, 2: A lo, 3: B lo), group g, row r;
  // consecutive u are consecutive rows of one (array, group).
  auto locate = [&](int u, int& arr, int& g, int& r) {
    const int per_a = L::G * C::BM, per_b = L::G * C::BN;
    const int pair = u / (per_a + per_b);  // 0: hi, 1: lo
    int w = u - pair * (per_a + per_b);
    if (w < per_a) {
      arr = 2 * pair;
      g = w / C::BM;
      r = w - g * C::BM;
    } else {
      w -= per_a;
      arr = 2 * pair + 1;
      g = w / C::BN;
      r = w - g * C::BN;
    }
  };
  auto gload = [&](int k0) {
#pragma unroll
    for (int t = 0; t < L::SLOTS; ++t) {
      const int u = tid + t * C::NT;
      if (u < L::UNITS) {
        int arr, g, r;
        locate(u, arr, g, r);
        const uint16_t* base = arr == 0 ? Ah : (arr == 1 ? Bh : (arr == 2 ? Al : Bl));
        reg[t] = ld_global16(base + ((long)(k0 / 8 + g) * gs + r) * 8);
      }
    }
  };
  auto sstore = [&](int buf) {
    uint16_t* st = lds + buf * L::STAGE;
#pragma unroll
    for (int t = 0; t < L::SLOTS; ++t) {
      const int u = tid + t * C::NT;
      if (u < L::UNITS) {
        int arr, g, r;
        locate(u, arr, g, r);
        const int off = (arr & 1) ? L::A_ARR : 0;
        const int lo = (arr >= 2) ? (L::A_ARR + L::B_ARR) : 0;
        *reinterpret_cast<uint4*>(st + lo + off + r * L::LDK + 8 * g) = reg[t];
      }
    }
  };
  auto compute = [&](int buf) {
    const uint16_t* st = lds + buf * L::STAGE;
#pragma unroll
    for (int kk = 0; kk < C::BK; kk += 32) {
      const int kcol = kk + 8 * (lane >> 4);
      bf16x8 ah[C::FM], bh[C::FN];
#pragma unroll
      for (int i = 0; i < C::FM; ++i)
        ah[i] = *reinterpret_cast<const bf16x8*>(st + (wm * C::WM + i * 16 + (lane & 15)) * L::LDK + kcol);
#pragma unroll
      for (int j = 0; j < C::FN; ++j)
        bh[j] = *reinterpret_cast<const bf16x8*>(st + L::A_ARR + (wn * C::WN + j * 16 + (lane & 15)) * L::LDK +
                                                 kcol);
      if (NSPLIT >= 3) {
        bf16x8 al[C::FM], bl[C::FN];
        const uint16_t* sl = st + L::A_ARR + L::B_ARR;
#pragma unroll
        for (int i = 0; i < C::FM; ++i)
          al[i] = *reinterpret_cast<const bf16x8*>(sl + (wm * C::WM + i * 16 + (lane & 15)) * L::LDK + kcol);
#pragma unroll
        for (int j = 0; j < C::FN; ++j)
          bl[j] = *reinterpret_cast<const bf16x8*>(sl + L::A_ARR + (wn * C::WN + j * 16 + (lane & 15)) * L::LDK +
                                                   kcol);
#pragma unroll
        for (int i = 0; i < C::FM; ++i)
#pragma unroll
          for (int j = 0; j < C::FN; ++j) {
            // smallest terms first
            if (NSPLIT == 4)
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bl[j], acc[i][j], 0, 0, 0);
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
  };
  const int nk = K / C::BK;
  gload(0);
  sstore(0);
  __syncthreads();
  for (int kc = 0; kc < nk; ++kc) {
    const int buf = kc & 1;
    if (kc + 1 < nk) gload((kc + 1) * C::BK);
    compute(buf);
    __builtin_amdgcn_sched_barrier(0);
    if (kc + 1 < nk) sstore(buf ^ 1);
    __syncthreads();
  }
}

}  // namespace exa
