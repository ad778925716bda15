// Kernels of the prediction-efficiency assessment (MLOE / MMOM).
//
// The reference evaluates, per held-out location s0 and with k = Sigma(obs,
// s0), two vector solves per factor, a triangular product and dot products
// (exageostat_exact/src/compute/MLE_dexact.c:1259-1385), re-reading both
// n x n factors for every location.  Here the nZmiss right-hand sides are
// the rows of a rectangular tile matrix R (nZmiss x nZobs), so every solve
// step is a tile product on the FP64 matrix cores and each factor tile is
// read once per row tile of R:
//   forward   R <- R L^-T   (each row r becomes L^-1 r)
//   backward  R <- R L^-1   (each row r becomes L^-T r)
//   symmetric Y = W S       (each row w becomes S w; S symmetric)
// followed by row-wise dot products.
#include "common.hpp"
#include "kernels.hpp"
#include "mfma.hpp"

namespace exa {

using R192 = GemmCfg<192, 192, 2, 4, 16>;
using R128 = GemmCfg<128, 128, 2, 2, 16>;
using R64 = GemmCfg<64, 64, 2, 2, 16>;

// O(i,k) = R(i,k) T for every row tile i, with T = X^T (forward) or X
// (backward), X = L_kk^-1 (lower triangular, zero above the diagonal).
template <class C, bool FWD>
__global__ void __launch_bounds__(C::NT) k_rhs_diag(TileView r, const double* X, int k, TileView o) {
  __shared__ __attribute__((aligned(16))) double lds[C::LDS_DOUBLES];
  const int qm = r.ts / C::BM, qn = r.ts / C::BN;
  const long b = blockIdx.x;
  const int i = (int)(b / (qm * qn));
  const int rem = (int)(b % (qm * qn));
  const int bi = rem % qm, bj = rem / qm;
  const long ts = r.ts;
  const double* A = tile_f64(r, i, k) + bi * C::BM;
  dx4 acc[C::FM][C::FN];
  zero_acc<C>(acc);
  if (FWD) {
    // B(c, kk) = X[bj BN + c][kk]: zero for kk > bj BN + c.
    gemm_nt_acc<C, false>(acc, A, ts, X + bj * C::BN, ts, (bj + 1) * C::BN, lds, 1.0);
  } else {
    // B(kk, c) = X[kk][bj BN + c] (K-major): zero for kk < bj BN + c.
    const long k0 = (long)bj * C::BN;
    gemm_nt_acc<C, true>(acc, A + k0 * ts, ts, X + k0 + k0 * ts, ts, (int)(ts - k0), lds, 1.0);
  }
  acc_update<C>(acc, tile_f64(o, i, k) + bi * C::BM + (long)bj * C::BN * ts, ts, 0.0);
}

// R(i,j) -= O(i,k) B_j for the tile columns j0 <= j < j0 + nj of R, every
// row tile i: B_j = L(j,k)^T (forward, j > k) or L(k,j) (backward, j < k).
template <class C, bool FWD>
__global__ void __launch_bounds__(C::NT) k_rhs_update(TileView o, TileView l, int k, int j0, TileView r) {
  __shared__ __attribute__((aligned(16))) double lds[C::LDS_DOUBLES];
  const int qm = r.ts / C::BM, qn = r.ts / C::BN;
  const long per = (long)qm * qn;
  const long b = blockIdx.x;
  const long t = b / per;
  const int rem = (int)(b - t * per);
  const int bi = rem % qm, bj = rem / qm;
  const int i = (int)(t % r.mt);
  const int j = j0 + (int)(t / r.mt);
  const long ts = r.ts;
  const double* A = tile_f64(o, i, k) + bi * C::BM;
  dx4 acc[C::FM][C::FN];
  zero_acc<C>(acc);
  if (FWD)
    gemm_nt_acc<C, false>(acc, A, ts, tile_f64(l, j, k) + bj * C::BN, ts, (int)ts, lds, -1.0);
  else
    gemm_nt_acc<C, true>(acc, A, ts, tile_f64(l, k, j) + (long)bj * C::BN * ts, ts, (int)ts, lds, -1.0);
  acc_update<C>(acc, tile_f64(r, i, j) + bi * C::BM + (long)bj * C::BN * ts, ts, 1.0);
}

// Y(i,n) = sum_l W(i,l) S(l,n) for the symmetric S held in the lower tiles
// of s (diagonal tiles full).  BK = false: l < n, S(l,n) = s(n,l)^T read
// plain; this launch writes every block.  BK = true: l >= n, s(l,n) read
// K-major, accumulated.
template <class C, bool BK>
__global__ void __launch_bounds__(C::NT) k_rect_symm(TileView w, TileView s, TileView y) {
  __shared__ __attribute__((aligned(16))) double lds[C::LDS_DOUBLES];
  const int qm = w.ts / C::BM, qn = w.ts / C::BN;
  const long per = (long)qm * qn;
  const long b = blockIdx.x;
  const long t = b / per;
  const int rem = (int)(b - t * per);
  const int bi = rem % qm, bj = rem / qm;
  const int i = (int)(t % w.mt);
  const int n = (int)(t / w.mt);
  const long ts = w.ts;
  const int l0 = BK ? n : 0, l1 = BK ? s.mt : n;
  dx4 acc[C::FM][C::FN];
  zero_acc<C>(acc);
  for (int l = l0; l < l1; ++l) {
    const double* A = tile_f64(w, i, l) + bi * C::BM;
    const double* B = BK ? tile_f64(s, l, n) + (long)bj * C::BN * ts : tile_f64(s, n, l) + bj * C::BN;
    gemm_nt_acc<C, BK>(acc, A, ts, B, ts, (int)ts, lds, 1.0);
  }
  acc_update<C>(acc, tile_f64(y, i, n) + bi * C::BM + (long)bj * C::BN * ts, ts, BK ? 1.0 : 0.0);
}

// part[c * len + row] = sum over the logical columns of tile column c of
// A(row, col) B(row, col), len = a.mt * ts.  One work-group per tile (i, c);
// thread t takes rows t, t + 256, ... (column-major tiles: a column of the
// tile is one coalesced read).
__global__ void __launch_bounds__(256) k_row_dot_part(TileView a, TileView b, double* part) {
  const int i = blockIdx.x % a.mt;
  const int c = blockIdx.x / a.mt;
  const int ts = a.ts;
  const double* A = tile_f64(a, i, c);
  const double* B = tile_f64(b, i, c);
  const int ncol = min(ts, a.n - c * ts);
  const long len = (long)a.mt * ts;
  for (int r = threadIdx.x; r < ts; r += 256) {
    double s0 = 0.0, s1 = 0.0;
    int j = 0;
    for (; j + 1 < ncol; j += 2) {
      s0 += A[r + (long)j * ts] * B[r + (long)j * ts];
      s1 += A[r + (long)(j + 1) * ts] * B[r + (long)(j + 1) * ts];
    }
    if (j < ncol) s0 += A[r + (long)j * ts] * B[r + (long)j * ts];
    part[(long)c * len + (long)i * ts + r] = s0 + s1;
  }
}

// out[row] = sum over c = 0 .. nt-1 (in this order) of part[c * len + row].
__global__ void __launch_bounds__(256) k_row_dot_reduce(const double* part, int nt, long len, long rows,
                                                        double* out) {
  const long r = (long)blockIdx.x * 256 + threadIdx.x;
  if (r >= rows) return;
  double s = 0.0;
  for (int c = 0; c < nt; ++c) s += part[(long)c * len + r];
  out[r] = s;
}

template <class C>
static void rhs_cfg(const TileView& l, const double* xinv, const TileView& r, const TileView& o, bool fwd,
                    hipStream_t s) {
  const long per = (long)(r.ts / C::BM) * (r.ts / C::BN);
  const long tts = (long)r.ts * r.ts;
  const int mt = l.mt;
  for (int step = 0; step < mt; ++step) {
    const int k = fwd ? step : mt - 1 - step;
    const double* X = xinv + (long)k * tts;
    const unsigned gd = (unsigned)(r.mt * per);
    if (fwd) hipLaunchKernelGGL((k_rhs_diag<C, true>), dim3(gd), dim3(C::NT), 0, s, r, X, k, o);
    else hipLaunchKernelGGL((k_rhs_diag<C, false>), dim3(gd), dim3(C::NT), 0, s, r, X, k, o);
    const int j0 = fwd ? k + 1 : 0;
    const int nj = fwd ? mt - k - 1 : k;
    if (nj > 0) {
      const unsigned gu = (unsigned)((long)r.mt * nj * per);
      if (fwd) hipLaunchKernelGGL((k_rhs_update<C, true>), dim3(gu), dim3(C::NT), 0, s, o, l, k, j0, r);
      else hipLaunchKernelGGL((k_rhs_update<C, false>), dim3(gu), dim3(C::NT), 0, s, o, l, k, j0, r);
    }
  }
  EXA_HIP(hipGetLastError());
}

static void check_rhs_shapes(const TileView& l, const TileView& r, const TileView& o) {
  if (!l.lower || r.lower || o.lower || r.ts != l.ts || o.ts != l.ts || r.nt != l.mt || o.nt != l.mt ||
      o.mt != r.mt || o.n != r.n || r.n != l.n)
    throw std::invalid_argument("right-hand-side solve: tile shapes do not match");
}

void launch_rhs_solve(const TileView& l, const double* xinv, const TileView& r, const TileView& o, bool fwd,
                      hipStream_t s) {
  check_rhs_shapes(l, r, o);
  if (l.ts % 192 == 0) rhs_cfg<R192>(l, xinv, r, o, fwd, s);
  else if (l.ts % 128 == 0) rhs_cfg<R128>(l, xinv, r, o, fwd, s);
  else rhs_cfg<R64>(l, xinv, r, o, fwd, s);
}

template <class C>
static void symm_cfg(const TileView& w, const TileView& sm, const TileView& y, hipStream_t s) {
  const long per = (long)(w.ts / C::BM) * (w.ts / C::BN);
  const unsigned g = (unsigned)((long)w.mt * w.nt * per);
  hipLaunchKernelGGL((k_rect_symm<C, false>), dim3(g), dim3(C::NT), 0, s, w, sm, y);
  hipLaunchKernelGGL((k_rect_symm<C, true>), dim3(g), dim3(C::NT), 0, s, w, sm, y);
  EXA_HIP(hipGetLastError());
}

void launch_rect_symm(const TileView& w, const TileView& sm, const TileView& y, hipStream_t s) {
  check_rhs_shapes(sm, w, y);
  if (w.ts % 192 == 0) symm_cfg<R192>(w, sm, y, s);
  else if (w.ts % 128 == 0) symm_cfg<R128>(w, sm, y, s);
  else symm_cfg<R64>(w, sm, y, s);
}

void launch_row_dot(const TileView& a, const TileView& b, double* part, double* out, hipStream_t s) {
  if (a.lower || b.lower || a.mt != b.mt || a.nt != b.nt || a.ts != b.ts || a.n != b.n)
    throw std::invalid_argument("row dot: tile shapes do not match");
  const long len = (long)a.mt * a.ts;
  hipLaunchKernelGGL(k_row_dot_part, dim3((unsigned)(a.mt * a.nt)), dim3(256), 0, s, a, b, part);
  hipLaunchKernelGGL(k_row_dot_reduce, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, (const double*)part,
                     a.nt, len, len, out);
  EXA_HIP(hipGetLastError());
}

}  // namespace exa
