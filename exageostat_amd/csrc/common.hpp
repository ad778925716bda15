// Shared definitions for the HIP kernels and the host scheduler.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace exa {

// Throws on any HIP error; the C-ABI layer turns exceptions into status
// codes (the reference calls exit() instead, MLE_misc.h:343-348).
#define EXA_HIP(call)                                                        \
  do {                                                                       \
    hipError_t err_ = (call);                                                \
    if (err_ != hipSuccess) {                                                \
      throw std::runtime_error(std::string("HIP error ") +                  \
                               hipGetErrorString(err_) + " at " __FILE__ ":" + \
                               std::to_string(__LINE__) + ": " #call);       \
    }                                                                        \
  } while (0)

// Per-tile storage / update precision: kF64 (FP64 storage and MFMA), kF32
// (FP32 storage and MFMA), kF32B (FP32 storage, BF16 MFMA operands with FP32
// accumulation), kBF16 (BF16 storage, generation output only).
enum TilePrec : uint8_t { kF64 = 0, kF32 = 1, kBF16 = 2, kF32B = 3 };

// Sub-block size used inside a tile by the diagonal factorisation, the
// triangular solves and the generation kernel.  Tile sizes must be a
// multiple of it.
constexpr int kNB = 64;

// Device view of a tiled matrix: square ts x ts tiles, column-major inside
// a tile with leading dimension ts.  Lower-packed matrices store only the
// tiles m >= n, column after column; full matrices store mt*nt tiles
// column after column.  Every tile is a separate allocation slot addressed
// through ptr[], so tiles of one matrix may have different precisions.
struct TileView {
  void* const* ptr;        // device array of tile base pointers
  const uint8_t* prec;     // device array of TilePrec per tile (may be null)
  int m, n;                // logical size (rows/cols >= m/n are padding)
  int mt, nt;              // tile counts
  int ts;                  // tile size
  int lower;               // 1: lower-packed square matrix
};

__host__ __device__ inline long tile_index(const TileView& v, int m, int n) {
  if (v.lower) return (long)n * v.mt - ((long)n * (n - 1)) / 2 + (m - n);
  return (long)n * v.mt + m;
}

__device__ inline double* tile_f64(const TileView& v, int m, int n) {
  return static_cast<double*>(v.ptr[tile_index(v, m, n)]);
}

// Sum over the 64 lanes of a wave (every lane gets the total).
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Row index r of the lower-triangle enumeration p = r(r+1)/2 + c, 0<=c<=r.
__device__ inline void tri_decode(long p, int& r, int& c) {
  long rr = (long)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
  while (rr * (rr + 1) / 2 > p) --rr;
  while ((rr + 1) * (rr + 2) / 2 <= p) ++rr;
  r = (int)rr;
  c = (int)(p - rr * (rr + 1) / 2);
}

}  // namespace exa
