// Matern covariance tile generation on the GPU.
//
// Replaces the per-tile CPU task core_dcmg / core_dcmg_nuggets /
// core_dcmg_bivariate_parsimonious (exageostat_exact/cpu_core/compute/
// core_zcmg.c:963-994, 1628-1682; core_zcmg_nuggets.c:107-156) that the
// reference schedules once per tile from EXAGEOSTAT_MLE_dcmg_Tile_Async
// (runtime/starpu/codelets/codelet_zcmg.c:246-347).  One launch covers
// every tile of a matrix; each work-group writes a 64 x 64 block of one
// tile with 64 consecutive rows per column (coalesced 512-byte stores).
// Padding rows/columns (global index >= logical size) are written as the
// identity so that padded tiles factor to [[L, 0], [0, I]].
#include "besselk.hpp"
#include "common.hpp"
#include "kernels.hpp"

namespace exa {

__device__ inline uint16_t f32_to_bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFF) ? 0x40 : 0));
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ inline void store_prec(void* base, long idx, double v, uint8_t prec) {
  if (prec == kF64) {
    static_cast<double*>(base)[idx] = v;
  } else if (prec == kF32 || prec == kF32B) {
    static_cast<float*>(base)[idx] = (float)v;
  } else {
    static_cast<uint16_t*>(base)[idx] = f32_to_bf16_rne((float)v);
  }
}

// Euclidean distance in 3-D when both location sets carry a third
// coordinate (calculateDistance, core_zcmg.c:71-92); great-circle
// distances are 2-D only.
__device__ inline double pair_distance3(double x1, double y1, double z1, double x2, double y2, double z2) {
  const double dx = x2 - x1, dy = y2 - y1, dz = z2 - z1;
  return sqrt(dx * dx + dy * dy + dz * dz);
}

__device__ inline double pair_distance(double x1, double y1, double x2, double y2, int metric) {
  if (metric == 1) {
    // Haversine great-circle distance in km (core_zcmg.c:51-60).
    const double d2r = M_PI / 180.0;
    const double lat1 = x1 * d2r, lon1 = y1 * d2r, lat2 = x2 * d2r, lon2 = y2 * d2r;
    const double u = sin((lat2 - lat1) * 0.5);
    const double v = sin((lon2 - lon1) * 0.5);
    return 2.0 * 6371.0 * asin(sqrt(u * u + cos(lat1) * cos(lat2) * v * v));
  }
  const double dx = x2 - x1, dy = y2 - y1;
  return sqrt(dx * dx + dy * dy);
}

// Decode blockIdx.x into (tile m, tile n, sub-block row bi, sub-block col
// bj) for the tiles p_off, p_off + 1, ... of the matrix's enumeration
// (lower matrices: p = m (m + 1) / 2 + n, row by row; full matrices:
// p = n mt + m).
__device__ inline bool decode_gen_block(const TileView& v, long p_off, int& m, int& n, int& bi, int& bj) {
  const int q = v.ts / kNB;
  const long b = blockIdx.x;
  const long pb = b / (q * q);
  const int rem = (int)(b - pb * (q * q));
  const long p = p_off + pb;
  bj = rem / q;
  bi = rem - bj * q;
  if (v.lower) {
    tri_decode(p, m, n);
  } else {
    m = (int)(p % v.mt);
    n = (int)(p / v.mt);
  }
  return true;
}

__global__ void __launch_bounds__(256) k_gen_matern(TileView v, MaternParams P, long p_off) {
  int m, n, bi, bj;
  decode_gen_block(v, p_off, m, n, bi, bj);
  const long t = tile_index(v, m, n);
  void* base = v.ptr[t];
  const uint8_t prec = v.prec ? v.prec[t] : (uint8_t)kF64;
  const int ts = v.ts;
  const int i = bi * kNB + threadIdx.x;  // row inside the tile
  const long gi = (long)m * ts + i;
  double xi = 0.0, yi = 0.0, zi = 0.0;
  if (gi < v.m) {
    xi = P.x1[gi];
    yi = P.y1[gi];
    if (P.z1) zi = P.z1[gi];
  }
  for (int cc = threadIdx.y; cc < kNB; cc += 4) {
    const int j = bj * kNB + cc;
    const long gj = (long)n * ts + j;
    double val;
    if (gi >= v.m || gj >= v.n) {
      val = (gi == gj) ? 1.0 : 0.0;
    } else {
      const double d = P.z1 ? pair_distance3(xi, yi, zi, P.x2[gj], P.y2[gj], P.z2[gj])
                            : pair_distance(xi, yi, P.x2[gj], P.y2[gj], P.metric);
      const double x = d / P.beta;
      if (x == 0.0) {
        val = P.sigma2 + P.nugget;
      } else {
        val = P.sigma2 * matern_corr(P.ko, P.con, x);
      }
    }
    store_prec(base, (long)i + (long)j * ts, val, prec);
  }
}

// Bivariate parsimonious Matern, variables interleaved by row/column parity:
// row 2p+a, column 2q+b hold cov(variable a at location p, variable b at
// location q).  Each thread writes one 2x2 block, as the reference's
// double loop does (core_zcmg.c:1657-1679).
__global__ void __launch_bounds__(256) k_gen_bivariate(TileView v, BivariateParams P, long p_off) {
  int m, n, bi, bj;
  decode_gen_block(v, p_off, m, n, bi, bj);
  const long t = tile_index(v, m, n);
  void* base = v.ptr[t];
  const uint8_t prec = v.prec ? v.prec[t] : (uint8_t)kF64;
  const int ts = v.ts;
  // 64 rows per block = 32 locations; threadIdx.x in [0,64): 32 row pairs x
  // 2 column-pair lanes.
  const int rp = threadIdx.x & 31;
  const int i = bi * kNB + 2 * rp;
  const long gi = (long)m * ts + i;
  double xi = 0.0, yi = 0.0, zi = 0.0;
  if (gi < v.m) {
    xi = P.x1[gi >> 1];
    yi = P.y1[gi >> 1];
    if (P.z1) zi = P.z1[gi >> 1];
  }
  for (int cp = (threadIdx.x >> 5) + 2 * threadIdx.y; cp < kNB / 2; cp += 8) {
    const int j = bj * kNB + 2 * cp;
    const long gj = (long)n * ts + j;
    double a00, a10, a01, a11;
    if (gi >= v.m || gj >= v.n) {
      a00 = (gi == gj) ? 1.0 : 0.0;
      a11 = (gi + 1 == gj + 1) ? 1.0 : 0.0;
      a10 = (gi + 1 == gj) ? 1.0 : 0.0;
      a01 = (gi == gj + 1) ? 1.0 : 0.0;
    } else {
      const long q = gj >> 1;
      const double d = P.z1 ? pair_distance3(xi, yi, zi, P.x2[q], P.y2[q], P.z2[q])
                            : pair_distance(xi, yi, P.x2[q], P.y2[q], P.metric);
      const double x = d / P.beta;
      if (x == 0.0) {
        a00 = P.sigma11;
        a11 = P.sigma22;
        a10 = a01 = P.sigma12;
      } else {
        a00 = P.sigma11 * matern_corr(P.ko11, P.con11, x);
        a11 = P.sigma22 * matern_corr(P.ko22, P.con22, x);
        a10 = a01 = P.sigma12 * matern_corr(P.ko12, P.con12, x);
      }
    }
    store_prec(base, (long)i + (long)j * ts, a00, prec);
    store_prec(base, (long)(i + 1) + (long)j * ts, a10, prec);
    store_prec(base, (long)i + (long)(j + 1) * ts, a01, prec);
    store_prec(base, (long)(i + 1) + (long)(j + 1) * ts, a11, prec);
  }
}

// Parsimonious bivariate Matern in the block layout of the reference's
// "bivariate_matern_parsimonious2" (three univariate generations into the
// quadrants of C, MLE_dexact.c:323-360): row r < nloc1 is variable 1 at
// location r, row nloc1 + r variable 2 at location r (columns likewise with
// nloc2).  One thread per entry, 64 consecutive rows per column.
__global__ void __launch_bounds__(256) k_gen_bivariate_block(TileView v, BivariateParams P, long p_off) {
  int m, n, bi, bj;
  decode_gen_block(v, p_off, m, n, bi, bj);
  const long t = tile_index(v, m, n);
  void* base = v.ptr[t];
  const uint8_t prec = v.prec ? v.prec[t] : (uint8_t)kF64;
  const int ts = v.ts;
  const int i = bi * kNB + threadIdx.x;
  const long gi = (long)m * ts + i;
  int va = 0;
  double xi = 0.0, yi = 0.0, zi = 0.0;
  if (gi < v.m) {
    va = gi >= P.nloc1 ? 1 : 0;
    const long p = gi - va * P.nloc1;
    xi = P.x1[p];
    yi = P.y1[p];
    if (P.z1) zi = P.z1[p];
  }
  for (int cc = threadIdx.y; cc < kNB; cc += 4) {
    const int j = bj * kNB + cc;
    const long gj = (long)n * ts + j;
    double val;
    if (gi >= v.m || gj >= v.n) {
      val = (gi == gj) ? 1.0 : 0.0;
    } else {
      const int vb = gj >= P.nloc2 ? 1 : 0;
      const long q = gj - vb * P.nloc2;
      const double d = P.z1 ? pair_distance3(xi, yi, zi, P.x2[q], P.y2[q], P.z2[q])
                            : pair_distance(xi, yi, P.x2[q], P.y2[q], P.metric);
      const double x = d / P.beta;
      const int c = va + vb;  // 0: var 1 with var 1, 1: cross, 2: var 2 with var 2
      const double sg = c == 0 ? P.sigma11 : (c == 1 ? P.sigma12 : P.sigma22);
      if (x == 0.0) {
        val = sg;
      } else if (c == 0) {
        val = sg * matern_corr(P.ko11, P.con11, x);
      } else if (c == 1) {
        val = sg * matern_corr(P.ko12, P.con12, x);
      } else {
        val = sg * matern_corr(P.ko22, P.con22, x);
      }
    }
    store_prec(base, (long)i + (long)j * ts, val, prec);
  }
}

// Space-time Matern (core_zcmg.c:1320-1357): with the spatial distance d
// (2-D; the time coordinate is not part of it) and u = |t_i - t_j|,
//   psi = u^(2 alpha) / a + 1,  x = d / (beta_s psi^(beta/2)),
//   C = sigma^2 M_nu(x) / psi^(beta + delta),   C = sigma^2 / psi^(beta + delta) at d = 0.
__global__ void __launch_bounds__(256) k_gen_spacetime(TileView v, SpaceTimeParams P, long p_off) {
  int m, n, bi, bj;
  decode_gen_block(v, p_off, m, n, bi, bj);
  const long t = tile_index(v, m, n);
  void* base = v.ptr[t];
  const uint8_t prec = v.prec ? v.prec[t] : (uint8_t)kF64;
  const int ts = v.ts;
  const int i = bi * kNB + threadIdx.x;
  const long gi = (long)m * ts + i;
  double xi = 0.0, yi = 0.0, ti = 0.0;
  if (gi < v.m) {
    xi = P.x1[gi];
    yi = P.y1[gi];
    ti = P.t1[gi];
  }
  for (int cc = threadIdx.y; cc < kNB; cc += 4) {
    const int j = bj * kNB + cc;
    const long gj = (long)n * ts + j;
    double val;
    if (gi >= v.m || gj >= v.n) {
      val = (gi == gj) ? 1.0 : 0.0;
    } else {
      const double xs = pair_distance(xi, yi, P.x2[gj], P.y2[gj], P.metric) / P.beta_s;
      const double u = fabs(ti - P.t2[gj]);
      const double psi = pow(u, 2.0 * P.alpha) / P.a + 1.0;
      const double den = pow(psi, P.beta + P.delta);
      if (xs == 0.0) {
        val = P.sigma2 / den;
      } else {
        const double x = xs / pow(psi, 0.5 * P.beta);
        val = P.sigma2 * matern_corr(P.ko, P.con, x) / den;
      }
    }
    store_prec(base, (long)i + (long)j * ts, val, prec);
  }
}

void launch_gen_spacetime(const TileView& v, const SpaceTimeParams& P, long p_off, long ntiles, hipStream_t s) {
  const int q = v.ts / kNB;
  const long blocks = ntiles * q * q;
  if (blocks <= 0) return;
  hipLaunchKernelGGL(k_gen_spacetime, dim3((unsigned)blocks), dim3(64, 4), 0, s, v, P, p_off);
  EXA_HIP(hipGetLastError());
}

void launch_gen_matern(const TileView& v, const MaternParams& P, long p_off, long ntiles, hipStream_t s) {
  const int q = v.ts / kNB;
  const long blocks = ntiles * q * q;
  if (blocks <= 0) return;
  hipLaunchKernelGGL(k_gen_matern, dim3((unsigned)blocks), dim3(64, 4), 0, s, v, P, p_off);
  EXA_HIP(hipGetLastError());
}

void launch_gen_bivariate(const TileView& v, const BivariateParams& P, long p_off, long ntiles, bool block,
                          hipStream_t s) {
  const int q = v.ts / kNB;
  const long blocks = ntiles * q * q;
  if (blocks <= 0) return;
  if (block)
    hipLaunchKernelGGL(k_gen_bivariate_block, dim3((unsigned)blocks), dim3(64, 4), 0, s, v, P, p_off);
  else
    hipLaunchKernelGGL(k_gen_bivariate, dim3((unsigned)blocks), dim3(64, 4), 0, s, v, P, p_off);
  EXA_HIP(hipGetLastError());
}

}  // namespace exa
