// Kernel parameter blocks and launcher declarations.
#pragma once

#include "besselk.hpp"
#include "common.hpp"

namespace exa {

// Univariate Matern (optionally with nugget) for one (row set, column set)
// pair of location arrays; theta-derived constants are prepared on the host.
struct MaternParams {
  const double* x1;
  const double* y1;
  const double* x2;
  const double* y2;
  const double* z1;  // third coordinates (3-D Euclidean distance), or null
  const double* z2;
  int metric;     // 0 Euclidean, 1 great circle
  double beta;    // range
  double sigma2;  // variance
  double nugget;  // added where the distance is exactly 0
  double con;     // 1/(2^(nu-1) Gamma(nu))
  KnuOrder<double> ko;
};

struct BivariateParams {
  const double* x1;
  const double* y1;
  const double* x2;
  const double* y2;
  const double* z1;  // third coordinates, or null
  const double* z2;
  long nloc1, nloc2;  // locations of the row / column sets (block layout)
  int metric;
  double beta;
  double sigma11, sigma22, sigma12;  // variances and rho~ sqrt(s1 s2)
  double con11, con22, con12;
  KnuOrder<double> ko11, ko22, ko12;
};

// Space-time Matern of Gneiting (2002, eq. 16) as core_dcmg_spacetime_matern
// writes it (exageostat_exact/cpu_core/compute/core_zcmg.c:1320-1357):
// location z is time, theta = (sigma^2, spatial range, nu, time range a,
// time smoothness alpha, non-separability beta, delta).
struct SpaceTimeParams {
  const double* x1;
  const double* y1;
  const double* t1;
  const double* x2;
  const double* y2;
  const double* t2;
  int metric;
  double sigma2, beta_s, a, alpha, beta, delta;
  double con;  // 1/(2^(nu-1) Gamma(nu))
  KnuOrder<double> ko;
};

// Generation of ntiles tiles starting at tile p_off of the matrix's
// enumeration (lower matrices: p = m (m + 1) / 2 + n, row by row; full
// matrices: p = n mt + m).
void launch_gen_matern(const TileView& v, const MaternParams& P, long p_off, long ntiles, hipStream_t s);
// block = false: variables interleaved by row parity; true: the rows of
// variable 1 (P.nloc1 of them), then those of variable 2.
void launch_gen_bivariate(const TileView& v, const BivariateParams& P, long p_off, long ntiles, bool block,
                          hipStream_t s);
void launch_gen_spacetime(const TileView& v, const SpaceTimeParams& P, long p_off, long ntiles, hipStream_t s);

// ---- tile Cholesky building blocks (FP64) -------------------------------
// Factor the diagonal tile (k,k) of a lower-packed matrix in place with
// 64-wide sub-blocks; writes the inverses of the 64x64 diagonal blocks of
// L_kk to dinv_k (q blocks of 64x64, column-major) and the first failing
// global column + 1 to *info (left untouched on success).
void launch_potrf_tile(double* T, int ts, double* dinv_k, int* info, long col_base, hipStream_t s);
// A(m,k) <- A(m,k) L_kk^-T for all m in (k, m_end) (m_end < 0: mt).
void launch_trsm_panel(const TileView& v, int k, const double* dinv_k, hipStream_t s, int m_end = -1);
// A(m,n) -= sum_{p < np} A(m,k+p) A(n,k+p)^T for n in [col_lo, col_hi],
// m in [n, mt); np = 1 or 2 (col_lo >= k + np).
void launch_tile_update(const TileView& v, int k, int np, int col_lo, int col_hi, hipStream_t s);

// X_k = L_kk^-1 for every diagonal tile (xinv: mt tiles of ts x ts).
void launch_trtri_tiles(const TileView& v, const double* dinv, double* xinv, hipStream_t s);

// ---- vectors --------------------------------------------------------------
// y = X t (trans = 0) or X^T t (trans = 1) for one lower ts x ts tile X;
// part: (ts/64) * ts doubles of scratch, ticket: a zeroed device counter
// (left zeroed), both private to the stream.
void launch_trmv_diag(const double* X, int ts, const double* t, double* y, int trans, double* part,
                      unsigned* ticket, hipStream_t s);
// z_m -= L(m,k) y_k for m in (k, mt).
void launch_gemv_panel(const TileView& v, int k, const double* yk, double* z, hipStream_t s);
// z_m -= L(k,m)^T y_k for m in [0, k).
void launch_gemv_t_row(const TileView& v, int k, const double* yk, double* z, hipStream_t s);
// partial[k] = sum_i log(L_kk[i,i]) over positive pivots of each diagonal tile.
void launch_logdet(const TileView& v, double* partial, hipStream_t s);
// partial[b] = sum of z[i]*w[i] over 4096-element blocks.
void launch_dot(const double* z, const double* w, long n, double* partial, hipStream_t s);
// y[0:m) = A x for a full tile matrix, or for a lower-packed one whose
// diagonal tiles are lower triangular (zero strictly-upper part); x is
// padded to nt*ts, part holds nt*mt*ts doubles of scratch.
void launch_gemv_full(const TileView& v, const double* x, double* y, double* part, hipStream_t s);
// The first stage of launch_gemv_full alone (partials, no sum).
void launch_gemv_full_partial(const TileView& v, const double* x, double* part, hipStream_t s);
// y[i] = sum_{c < ncols} part[c*len + i], i < m.
void launch_sum_columns(const double* part, int ncols, long len, double* y, long m, hipStream_t s);
// Zero the strictly upper part of every diagonal tile of a lower matrix.
void launch_zero_upper_diag(const TileView& v, hipStream_t s);

// ---- mixed precision ---------------------------------------------------
// dst <- src tile by tile (dst tiles FP64 or FP32 by dst.prec).
void launch_convert_mixed(const TileView& src, const TileView& dst, long ntiles, hipStream_t s);
// dst (FP64) <- src (mixed precision, same tiling): every tile widened.
void launch_mixed_to_fp64(const TileView& src, const TileView& dst, long ntiles, hipStream_t s);
// BF16 hi/lo parts of the panel tiles (m, k), m > k, for the BF16 update
// classes: tile m at split + 2 m ts^2 (mt tiles of room).
void launch_split_panel(const TileView& v, int k, uint16_t* split, hipStream_t s);
// Trailing update of step k over the tile pairs (packed m << 16 | n) of the
// three precision classes of the updated tile, in one launch when ts is a
// multiple of 192 (else one launch per class): pairs[0] /
// nblocks[0] FP64 tiles (FP64 MFMA), [1] FP32 tiles (split-BF16 MFMA, with
// the lo*lo product when lolo), [2] far tiles (BF16 MFMA).  nblocks counts
// tile pairs; the BF16 classes read their operands from the split panel of
// step k.
// The three precision classes of one fused update launch, in the order each
// XCD walks them: slot c holds the pairs of class cls[c] (0: FP64 tiles,
// 1: FP32 tiles, 2: far tiles).
struct FusedSpans {
  const int* pairs[3];
  long nblocks[3];
  int cls[3];
};
void launch_update_fused(const TileView& v, int k, const FusedSpans& sp, const uint16_t* split, bool lolo,
                         hipStream_t s);
// out = a*x + b*y (n entries).
void launch_axpby(double a, const double* x, double b, const double* y, double* out, long n, hipStream_t s);
// y = Sigma x for a lower-packed FP64 matrix with full diagonal tiles, each
// tile read once; part: 2 * mt * (mt * ts) doubles of scratch.
void launch_symv(const TileView& v, const double* x, double* y, double* part, hipStream_t s);
// The two halves of launch_symv: the partial products of the tiles p0 ..
// p0 + ntiles - 1 (row-by-row enumeration p = m (m + 1) / 2 + n), then the
// sum of every tile's partials into y.
void launch_symv_tiles(const TileView& v, const double* x, double* part, long p0, long ntiles, hipStream_t s);
void launch_symv_reduce(const TileView& v, const double* part, double* y, hipStream_t s);

// ---- adjoint gradient ---------------------------------------------------
struct MaternGradParams {
  const double* x;
  const double* y;
  long n;       // logical matrix size
  int metric;   // 0 Euclidean, 1 great circle
  int nugget;   // 1: d/dtau^2 (fourth component) where the distance is 0
  double sigma2;
  double beta;
  Dual con;               // 1/(2^(nu-1) Gamma(nu)) and its nu-derivative
  KnuOrder<Dual> ko;      // K_nu constants with nu-derivatives
};
// One stage of tile column k of the Takahashi sweep Z = Sigma^-1 over the
// factor in v (0: P_j = L_jk X_kk; 1: Z_mk for m > k; 2: Z_kk into Zkk,
// with Q as split-k scratch of mt tiles).  X = L_kk^-1, P: mt tiles.
// W (may be null): slabs for splitting stage 1's sum over j into up to
// max_split ranges, max_split * (mt - 1) tiles.  U (may be null): a lower
// tile matrix of v's geometry that receives Z_mk^T in tile (m, k), so that
// the later stages read no transposed operands.  f32 (with U): stage 1 on
// the FP32 matrix cores (operands rounded to FP32, FP32 sums).
void launch_takahashi_stage(const TileView& v, int k, const double* X, double* P, double* Q, double* Zkk,
                            int stage, hipStream_t s, double* W = nullptr, int max_split = 1,
                            const TileView* U = nullptr,
                            bool f32 = false);
struct BivariateGradParams {
  const double* x;
  const double* y;
  long n;  // rows (2 x locations)
  int metric;
  double s1, s2, beta, nu1, nu2, nu12;
  double s12;        // rho~ sqrt(s1 s2), the cross variance
  double rho_ratio;  // d s12 / d rho = sqrt(nu1 nu2)/nu12 sqrt(s1 s2)
  Dual con1, con2, con12;
};
// Number of partial blocks of launch_grad_matern / launch_grad_bivariate
// (partial: 4, resp. 8, doubles per block).
long grad_partial_blocks(const TileView& v);
// dSigma/dtheta_p of the univariate Matern kernels into the lower tiles of d
// (diagonal tiles full, zero outside the logical matrix).
void launch_gen_dmatern(const TileView& d, const MaternGradParams& P, int p, hipStream_t s);
// M = Z D Z (lower tiles of mo) for symmetric Z and D held in lower tiles,
// through the full tile matrix Y = Z D (mt * mt tiles of scratch).
void launch_fisher_zdz(const TileView& z, const TileView& d, double* Y, const TileView& mo, hipStream_t s);

// MLOE / MMOM (mloe_kernels.hip).  r and o: rectangular tile matrices whose
// rows are right-hand sides (r.nt == l.mt); l: lower tile factor, xinv: its
// inverted diagonal tiles.  fwd: o = rows of r mapped by L^-1 (r <- r L^-T),
// else by L^-T; r is overwritten.
void launch_rhs_solve(const TileView& l, const double* xinv, const TileView& r, const TileView& o, bool fwd,
                      hipStream_t s);
// y = w S for the symmetric S in the lower tiles of sm (diagonal tiles full).
void launch_rect_symm(const TileView& w, const TileView& sm, const TileView& y, hipStream_t s);
// out[row] = sum over the logical columns of a(row, col) b(row, col), for
// all a.mt * ts rows; part: a.nt * a.mt * ts doubles.
void launch_row_dot(const TileView& a, const TileView& b, double* part, double* out, hipStream_t s);
// out[p*groups + g], p < 4: partial sums of sum_ij w_ij (Z_ij - a_i a_j)
// dSigma_ij/dtheta_p over the lower tiles of Z (w = 1 on the diagonal, 2
// elsewhere); p = sigma^2, beta, nu, nugget.
void launch_grad_matern(const TileView& v, const MaternGradParams& P, const double* alpha, double* partial,
                        double* out, int groups, hipStream_t s);
// Same for the parsimonious bivariate kernel: out[p*groups + g], p < 6, for
// theta = (s1, s2, beta, nu1, nu2, rho); ko: device array of the K_nu
// constants for nu1, nu2, nu12.
void launch_grad_bivariate(const TileView& v, const BivariateGradParams& P, const KnuOrder<Dual>* ko,
                           const double* alpha, double* partial, double* out, int groups, hipStream_t s);

}  // namespace exa
