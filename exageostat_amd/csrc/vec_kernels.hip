// Vector-side kernels of the likelihood: blocked triangular solves with the
// tile factor, log-determinant and dot-product partial sums, and the
// rectangular matrix-vector product of kriging.
//
// Replaces, from EXAGEOSTAT_dmle_Tile (exageostat_exact/src/compute/
// MLE_dexact.c:381-399): the dmdet reduction (codelet_zmdet.c:108-139,
// core_zmdet.c:50-59), CHAMELEON_dtrsm_Tile(Left, Lower, NoTrans, NonUnit)
// on Z and the 1x1 CHAMELEON_dgemm_Tile dot product; from
// EXAGEOSTAT_dmle_Predict_Tile (:812-820): the two triangular solves of
// dposv and the dgemm C12 * alpha.
#include "common.hpp"
#include "kernels.hpp"

namespace exa {

// sum over columns c = g, g+4, ... < ts of A[c*ts] * y[c] (A points at one
// row of a column-major tile); eight loads in flight per step.
template <class T>
__device__ inline double row_dot(const T* A, int ts, const double* y, int g) {
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  int c = g;
  for (; c + 28 < ts; c += 32) {
    T a[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] = A[(long)(c + 4 * u) * ts];
#pragma unroll
    for (int u = 0; u < 8; ++u) s[u & 3] += (double)a[u] * y[c + 4 * u];
  }
  for (; c < ts; c += 4) s[0] += (double)A[(long)c * ts] * y[c];
  return (s[0] + s[1]) + (s[2] + s[3]);
}

// Sums of 16 consecutive columns of a column-major tile against y:
// s[j] = sum over rows of A[row + j*ts] * y[row], this lane taking rows
// r, r + 64, ... < ts; the caller reduces over the wave.  16 loads in
// flight per step.
template <class T>
__device__ inline void col_dot16(const T* A, int ts, const double* y, int r, double (&s)[16]) {
#pragma unroll
  for (int j = 0; j < 16; ++j) s[j] = 0.0;
  for (int row = r; row < ts; row += 64) {
    T a[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) a[j] = A[row + (long)j * ts];
    const double yr = y[row];
#pragma unroll
    for (int j = 0; j < 16; ++j) s[j] += (double)a[j] * yr;
  }
}

// Partial products of one lower-triangular ts x ts tile X (the inverse of
// a diagonal factor tile) with t, one work-group per 64 x 64 block (rb, cb)
// of its lower block triangle (q(q+1)/2 blocks, q = ts / 64):
// trans = 0: part[cb][rb*64 + i] = sum_j X[rb*64+i][cb*64+j] t[cb*64+j];
// trans = 1: part[rb][cb*64 + j] = sum_i X[rb*64+i][cb*64+j] t[rb*64+i].
// The last work-group to finish (an atomic ticket after a device-scope
// fence) sums the partials in a fixed order into y and resets the ticket,
// so the result does not depend on the work-group schedule.
__global__ void __launch_bounds__(256) k_trmv_diag(const double* X, int ts, const double* t, double* y, int trans,
                                                   double* part, unsigned* ticket) {
  __shared__ double red[4][64];
  __shared__ int last;
  const int tid = threadIdx.x;
  const int r = tid & 63;
  const int g = tid >> 6;
  const int q = ts / kNB;
  int rb, cb;
  tri_decode(blockIdx.x, rb, cb);
  const double* B = X + (long)(cb * kNB + g * 16) * ts + rb * kNB;
  double a[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) a[j] = B[r + (long)j * ts];
  if (!trans) {
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int j = 0; j < 16; j += 2) {
      s0 += a[j] * t[cb * kNB + g * 16 + j];
      s1 += a[j + 1] * t[cb * kNB + g * 16 + j + 1];
    }
    red[g][r] = s0 + s1;
    __syncthreads();
    if (tid < 64) part[(long)cb * ts + rb * kNB + r] = (red[0][r] + red[1][r]) + (red[2][r] + red[3][r]);
  } else {
    const double tr = t[rb * kNB + r];
#pragma unroll
    for (int j = 0; j < 16; ++j) a[j] = wave_sum(a[j] * tr);
    if (r == 0) {
#pragma unroll
      for (int j = 0; j < 16; ++j) part[(long)rb * ts + cb * kNB + g * 16 + j] = a[j];
    }
  }
  __threadfence();
  __syncthreads();
  if (tid == 0) last = atomicAdd(ticket, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!last) return;
  __threadfence();
  for (int i = tid; i < ts; i += 256) {
    const int b = i / kNB;
    double s = 0.0;
    if (!trans) {
      for (int c = 0; c <= b; ++c) s += part[(long)c * ts + i];
    } else {
      for (int rr = b; rr < q; ++rr) s += part[(long)rr * ts + i];
    }
    y[i] = s;
  }
  if (tid == 0) *ticket = 0u;
}

// z_m -= L(m,k) y_k for m in (k, mt); one work-group per 64-row block.
__global__ void __launch_bounds__(256) k_gemv_panel(TileView v, int k, const double* yk, double* z) {
  __shared__ double y[1024];
  __shared__ double part[4][64];
  const int tid = threadIdx.x;
  const int r = tid & 63;
  const int g = tid >> 6;
  const int ts = v.ts;
  const int q = ts / kNB;
  const int m = k + 1 + blockIdx.x / q;
  const int rb = blockIdx.x - (blockIdx.x / q) * q;
  for (int i = tid; i < ts; i += 256) y[i] = yk[i];
  __syncthreads();
  const long ti = tile_index(v, m, k);
  // Tiles of a mixed-precision factor may be stored in FP32.
  if (v.prec && v.prec[ti] != kF64) {
    part[g][r] = row_dot(static_cast<const float*>(v.ptr[ti]) + rb * kNB + r, ts, y, g);
  } else {
    part[g][r] = row_dot(static_cast<const double*>(v.ptr[ti]) + rb * kNB + r, ts, y, g);
  }
  __syncthreads();
  if (tid < 64) {
    z[(long)m * ts + rb * kNB + r] -= (part[0][r] + part[1][r]) + (part[2][r] + part[3][r]);
  }
}

// z_m -= L(k,m)^T y_k for m in [0, k); one work-group per 64 columns of
// L(k,m), 16 columns per wave.
__global__ void __launch_bounds__(256) k_gemv_t_row(TileView v, int k, const double* yk, double* z) {
  __shared__ double y[1024];
  const int tid = threadIdx.x;
  const int r = tid & 63;
  const int g = tid >> 6;
  const int ts = v.ts;
  const int q = ts / kNB;
  const int m = blockIdx.x / q;
  const int cb = blockIdx.x - m * q;
  for (int i = tid; i < ts; i += 256) y[i] = yk[i];
  __syncthreads();
  const long ti = tile_index(v, k, m);
  const long c0 = (long)(cb * kNB + g * 16) * ts;
  double s[16];
  if (v.prec && v.prec[ti] != kF64)
    col_dot16(static_cast<const float*>(v.ptr[ti]) + c0, ts, y, r, s);
  else
    col_dot16(static_cast<const double*>(v.ptr[ti]) + c0, ts, y, r, s);
#pragma unroll
  for (int j = 0; j < 16; ++j) s[j] = wave_sum(s[j]);
  if (r == 0) {
    double* zc = z + (long)m * ts + cb * kNB + g * 16;
#pragma unroll
    for (int j = 0; j < 16; ++j) zc[j] -= s[j];
  }
}

__global__ void __launch_bounds__(256) k_logdet(TileView v, double* partial) {
  __shared__ double red[256];
  const int k = blockIdx.x;
  const double* T = tile_f64(v, k, k);
  double s = 0.0;
  for (int i = threadIdx.x; i < v.ts; i += 256) {
    const long gi = (long)k * v.ts + i;
    const double d = T[i + (long)i * v.ts];
    if (gi < v.m && d > 0.0) s += log(d);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[k] = red[0];
}

__global__ void __launch_bounds__(256) k_dot(const double* z, const double* w, long n, double* partial) {
  __shared__ double red[256];
  const long base = (long)blockIdx.x * 4096;
  double s = 0.0;
  for (int i = threadIdx.x; i < 4096; i += 256) {
    const long gi = base + i;
    if (gi < n) s += z[gi] * w[gi];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int ww = 128; ww > 0; ww >>= 1) {
    if (threadIdx.x < ww) red[threadIdx.x] += red[threadIdx.x + ww];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// y_partial[n][row] = sum over tile column n of A(m,n)[row] x_n; one
// work-group per (tile row, 64-row block, tile column).
__global__ void __launch_bounds__(256) k_gemv_full_partial(TileView v, const double* x, double* part_out) {
  __shared__ double part[4][64];
  const int tid = threadIdx.x;
  const int r = tid & 63;
  const int g = tid >> 6;
  const int ts = v.ts;
  const int q = ts / kNB;
  const int n = blockIdx.y;
  const int m = blockIdx.x / q;
  const int rb = blockIdx.x - m * q;
  double s = 0.0;
  if (!(v.lower && n > m)) {
    const double* A = tile_f64(v, m, n) + rb * kNB + r;
    const double* xn = x + (long)n * ts;
    for (int c = g; c < ts; c += 4) s += A[(long)c * ts] * xn[c];
  }
  part[g][r] = s;
  __syncthreads();
  if (tid < 64) {
    part_out[(long)n * ((long)v.mt * ts) + (long)m * ts + rb * kNB + r] =
        (part[0][r] + part[1][r]) + (part[2][r] + part[3][r]);
  }
}

__global__ void k_sum_columns(const double* part, int ncols, long len, double* y, long m) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double s = 0.0;
  for (int c = 0; c < ncols; ++c) s += part[(long)c * len + i];
  y[i] = s;
}

__global__ void __launch_bounds__(256) k_zero_upper_diag(TileView v) {
  const int k = blockIdx.y;
  double* T = tile_f64(v, k, k);
  const int c = blockIdx.x;
  for (int r = threadIdx.x; r < c; r += 256) T[r + (long)c * v.ts] = 0.0;
}

void launch_zero_upper_diag(const TileView& v, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_upper_diag, dim3(v.ts, v.mt), dim3(256), 0, s, v);
  EXA_HIP(hipGetLastError());
}

void launch_trmv_diag(const double* X, int ts, const double* t, double* y, int trans, double* part,
                      unsigned* ticket, hipStream_t s) {
  const int q = ts / kNB;
  hipLaunchKernelGGL(k_trmv_diag, dim3(q * (q + 1) / 2), dim3(256), 0, s, X, ts, t, y, trans, part, ticket);
  EXA_HIP(hipGetLastError());
}

void launch_gemv_panel(const TileView& v, int k, const double* yk, double* z, hipStream_t s) {
  const int q = v.ts / kNB;
  const long blocks = (long)(v.mt - k - 1) * q;
  if (blocks <= 0) return;
  hipLaunchKernelGGL(k_gemv_panel, dim3((unsigned)blocks), dim3(256), 0, s, v, k, yk, z);
  EXA_HIP(hipGetLastError());
}

void launch_gemv_t_row(const TileView& v, int k, const double* yk, double* z, hipStream_t s) {
  const int q = v.ts / kNB;
  const long blocks = (long)k * q;
  if (blocks <= 0) return;
  hipLaunchKernelGGL(k_gemv_t_row, dim3((unsigned)blocks), dim3(256), 0, s, v, k, yk, z);
  EXA_HIP(hipGetLastError());
}

void launch_logdet(const TileView& v, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(k_logdet, dim3(v.mt), dim3(256), 0, s, v, partial);
  EXA_HIP(hipGetLastError());
}

void launch_dot(const double* z, const double* w, long n, double* partial, hipStream_t s) {
  const long blocks = (n + 4095) / 4096;
  if (blocks <= 0) return;
  hipLaunchKernelGGL(k_dot, dim3((unsigned)blocks), dim3(256), 0, s, z, w, n, partial);
  EXA_HIP(hipGetLastError());
}

void launch_gemv_full_partial(const TileView& v, const double* x, double* part, hipStream_t s) {
  const int q = v.ts / kNB;
  hipLaunchKernelGGL(k_gemv_full_partial, dim3((unsigned)(v.mt * q), (unsigned)v.nt), dim3(256), 0, s, v, x,
                     part);
  EXA_HIP(hipGetLastError());
}

void launch_sum_columns(const double* part, int ncols, long len, double* y, long m, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_columns, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, part, ncols, len, y, m);
  EXA_HIP(hipGetLastError());
}

// part must hold nt * mt * ts doubles.
void launch_gemv_full(const TileView& v, const double* x, double* y, double* part, hipStream_t s) {
  const int q = v.ts / kNB;
  hipLaunchKernelGGL(k_gemv_full_partial, dim3((unsigned)(v.mt * q), (unsigned)v.nt), dim3(256), 0, s, v, x,
                     part);
  const long len = (long)v.mt * v.ts;
  hipLaunchKernelGGL(k_sum_columns, dim3((unsigned)((v.m + 255) / 256)), dim3(256), 0, s, part, v.nt, len, y,
                     (long)v.m);
  EXA_HIP(hipGetLastError());
}

}  // namespace exa
