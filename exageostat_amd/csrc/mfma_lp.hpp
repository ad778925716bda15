// Lower-precision MFMA work-group GEMM core for the mixed-precision tile
// Cholesky on CDNA4 (gfx950): BF16 operands (single or split in two BF16
// terms) with FP32 accumulation (v_mfma_f32_16x16x32_bf16).
//
// It computes acc += A * B^T for a BM x BN block from panel tiles that were
// converted once per step into BF16 parts (k_split_panel), copied into a
// three-stage LDS ring with direct-to-LDS loads.
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

// C[r + c*ldc] += sign * acc (C stored as T); acc is overwritten.  Every
// read of C precedes the first write, through global-address-space
// pointers, so the block's reads are issued together instead of one
// read-modify-write at a time.
template <class C, class T>
__device__ __forceinline__ void acc_add_lp(fx4 (&acc)[C::FM][C::FN], T* Cp, long ldc, float sign = 1.f) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave / C::WGN;
  const int wn = wave - wm * C::WGN;
  typedef __attribute__((address_space(1))) T gT;
  gT* G = (gT*)(uintptr_t)Cp;
#pragma unroll
  for (int i = 0; i < C::FM; ++i)
#pragma unroll
    for (int j = 0; j < C::FN; ++j) {
      const int c = wn * C::WN + j * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = wm * C::WM + i * 16 + frag_row_lp(lane, v);
        acc[i][j][v] = (float)G[r + (long)c * ldc] + sign * acc[i][j][v];
      }
    }
#pragma unroll
  for (int i = 0; i < C::FM; ++i)
#pragma unroll
    for (int j = 0; j < C::FN; ++j) {
      const int c = wn * C::WN + j * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = wm * C::WM + i * 16 + frag_row_lp(lane, v);
        G[r + (long)c * ldc] = (T)acc[i][j][v];
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

// ------------------------------------------- direct-to-LDS operand loads --
// Panel tiles are converted once per step into BF16 hi (and lo) parts (see
// k_split_panel), laid out per tile as [k-group g][row r][8 k values]
// (g = column / 8): one 16-byte load is eight k values of one row, and
// consecutive rows are consecutive in memory.  The lo parts of a tile follow
// its hi parts.
//
// The product copies the operands with
// global_load_lds_dwordx4 (each wave-wide instruction writes 64 lanes x 16
// bytes contiguously into LDS, no registers involved) into a three-stage
// ring.  The LDS layout of a stage is the panel's own [k-group][row][8]:
// array a (A hi, B hi, A lo, B lo) holds G = BK/8 groups of its rows.  A
// 16-lane quarter of a fragment read covers 16 consecutive rows of one
// group, 256 contiguous bytes, so neither the copies nor the reads conflict
// on LDS banks.
template <class C, int NSPLIT>
struct Bf16dLds {
  static constexpr int G = C::BK / 8;
  static constexpr int NARR = NSPLIT >= 3 ? 2 : 1;
  static constexpr int A_ARR = G * C::BM * 8;  // halves
  static constexpr int B_ARR = G * C::BN * 8;
  static constexpr int STAGE = NARR * (A_ARR + B_ARR);
  static constexpr int NSTAGE = 3;
  static constexpr int HALVES = NSTAGE * STAGE;
  static constexpr int CHUNKS = NARR * G * (C::BM + C::BN) / 64;  // wave-wide copies per stage
  static constexpr int PER_WAVE = CHUNKS / C::NW;
  static_assert(C::BM % 64 == 0 && C::BN % 64 == 0, "rows must be whole 64-row chunks");
  static_assert(CHUNKS % C::NW == 0, "copies must split evenly over the waves");
};



// acc += A * B^T (A and B rows from the split panels, so the caller
// subtracts the result), for K a multiple of BK.  The ring holds
// slices kc, kc+1, kc+2: at step kc each wave waits for its own copies of
// slice kc, a barrier makes every wave's copies visible and ends every read
// of slice kc-1, whose buffer then receives slice kc+2.  The barrier is a
// bare s_barrier: the copies are ordered by the explicit counter waits, and
// LDS reads complete before the MFMAs that consume them.
template <class C, int NSPLIT>
__device__ __forceinline__ void gemm_nt_bf16d(fx4 (&acc)[C::FM][C::FN], const uint16_t* __restrict__ Ah,
                                                  const uint16_t* __restrict__ Bh, long lo_off, long gs, int K,
                                                  uint16_t* lds) {
  using L = Bf16dLds<C, NSPLIT>;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave / C::WGN;
  const int wn = wave - wm * C::WGN;
  // The copy decomposition is uniform per wave: keep it in scalar registers.
  const int wave_s = __builtin_amdgcn_readfirstlane(wave);
  // copy j of this wave: chunk c = wave * PER_WAVE + j of array (hi/lo,
  // A/B), k-group g, 64-row block rc; lo parts sit lo_off halves after hi.
  auto issue = [&](int kc, int buf) {
#pragma unroll
    for (int j = 0; j < L::PER_WAVE; ++j) {
      constexpr int per_a = L::G * (C::BM / 64), per_b = L::G * (C::BN / 64);
      int c = wave_s * L::PER_WAVE + j;
      const int pair = c / (per_a + per_b);  // 0: hi, 1: lo
      c -= pair * (per_a + per_b);
      const bool is_b = c >= per_a;
      if (is_b) c -= per_a;
      const int rows = is_b ? C::BN : C::BM;
      const int g = c / (rows / 64);
      const int rc = c - g * (rows / 64);
      const uint16_t* base = is_b ? Bh : Ah;
      const long goff = pair * lo_off + ((long)(kc * L::G + g) * gs + rc * 64) * 8;
      const int loff = buf * L::STAGE + pair * (L::A_ARR + L::B_ARR) + (is_b ? L::A_ARR : 0) + (g * rows + rc * 64) * 8;
      __builtin_amdgcn_global_load_lds((glb_void_t*)(uintptr_t)(base + goff + lane * 8), (lds_void_t*)(lds + loff), 16,
                                       0, 0);
    }
  };
  auto compute = [&](int buf) {
    const uint16_t* st = lds + buf * L::STAGE;
#pragma unroll
    for (int kk = 0; kk < C::BK; kk += 32) {
      const int gq = kk / 8 + (lane >> 4);
      bf16x8 ah[C::FM], bh[C::FN];
#pragma unroll
      for (int i = 0; i < C::FM; ++i)
        ah[i] = *reinterpret_cast<const bf16x8*>(st + (gq * C::BM + wm * C::WM + i * 16 + (lane & 15)) * 8);
#pragma unroll
      for (int j = 0; j < C::FN; ++j)
        bh[j] = *reinterpret_cast<const bf16x8*>(st + L::A_ARR + (gq * C::BN + wn * C::WN + j * 16 + (lane & 15)) * 8);
      if (NSPLIT >= 3) {
        bf16x8 al[C::FM], bl[C::FN];
        const uint16_t* sl = st + L::A_ARR + L::B_ARR;
#pragma unroll
        for (int i = 0; i < C::FM; ++i)
          al[i] = *reinterpret_cast<const bf16x8*>(sl + (gq * C::BM + wm * C::WM + i * 16 + (lane & 15)) * 8);
#pragma unroll
        for (int j = 0; j < C::FN; ++j)
          bl[j] = *reinterpret_cast<const bf16x8*>(sl + L::A_ARR + (gq * C::BN + wn * C::WN + j * 16 + (lane & 15)) * 8);
        // Each 32-deep chunk is summed by its own MFMA chain from zero
        // (smallest terms first) and added to the running sum with one
        // round-to-nearest FP32 add: the MFMA's internal accumulation is
        // not correctly rounded (about 1 ulp of error per instruction,
        // measured on gfx950), and chaining all 3 K / 32 instructions into
        // one accumulator lets that error build up at the running sum's
        // magnitude.
#pragma unroll
        for (int i = 0; i < C::FM; ++i)
#pragma unroll
          for (int j = 0; j < C::FN; ++j) {
            fx4 t = {0.f, 0.f, 0.f, 0.f};
            if (NSPLIT == 4) t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bl[j], t, 0, 0, 0);
            t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], t, 0, 0, 0);
            t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], t, 0, 0, 0);
            t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], t, 0, 0, 0);
            acc[i][j] += t;
          }
      } else {
#pragma unroll
        for (int i = 0; i < C::FM; ++i)
#pragma unroll
          for (int j = 0; j < C::FN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
      }
    }
  };
  const int nk = K / C::BK;
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

}  // namespace exa
