// Per-point weight coefficients, reductions, the triangular-contraction operand B, and layout
// conversions of h (the N x N solve itself is in k_hsolve.hip).
// Complex matrices are row-major [n][n][2] (re, im).
#include "cy3q_internal.h"

namespace cy3q {

// ---- per-point coefficients ---------------------------------------------------------------
// c_p = w_p * (1 / max(D_p, 1e-30)) for valid weights (finite, > 0), else 0 (quintic.rs:1757,
// 1776-1781). D_p is the sum of the n_parts column-block partials, in block order.
__global__ void k_point_coeffs_f64(const double *__restrict__ w, const double *__restrict__ dpart,
                                   int n_parts, int64_t n, double *__restrict__ c) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double d = 0.0;
  for (int q = 0; q < n_parts; q++) d += dpart[(int64_t)q * n + p];
  double wp = w[p];
  bool ok = isfinite(wp) && wp > 0.0;
  c[p] = ok ? (1.0 / fmax(d, 1e-30)) * wp : 0.0;
}

void launch_point_coeffs_f64(hipStream_t st, const double *w, const double *dpart, int n_parts,
                             int64_t n, double *c) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_point_coeffs_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, w,
                     dpart, n_parts, n, c);
}

// Masked copy of the weights: c_p = w_p if valid else 0 (FS-Gram and sum-w inputs).
__global__ void k_valid_weights(const double *__restrict__ w, int64_t n, double *__restrict__ c) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double wp = w[p];
  c[p] = (isfinite(wp) && wp > 0.0) ? wp : 0.0;
}

void launch_valid_weights(hipStream_t st, const double *w, int64_t n, double *c) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_valid_weights, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, w, n,
                     c);
}

// Deterministic sum of a vector: one workgroup, fixed strided partition then a tree.
__global__ void __launch_bounds__(1024) k_sum_f64(const double *__restrict__ v, int64_t n,
                                                  double *__restrict__ out) {
  __shared__ double sh[1024];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 1024) s += v[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 512; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

void launch_sum_f64(hipStream_t st, const double *v, int64_t n, double *out) {
  hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(1024), 0, st, v, n, out);
}

// ---- B matrix for the triangular contraction ------------------------------------------------
// y_b = sum_{a<=b} V_ba s_a in the interleaved real layout:
//   B[2a][2b] = Re V_ba, B[2a+1][2b] = -Im V_ba, B[2a][2b+1] = Im V_ba, B[2a+1][2b+1] = Re V_ba.
// conj_v uses conj(V). Entries with a > b, and padding rows/cols, are zero.
__global__ void k_build_bmat(const double *__restrict__ v, int nb, int64_t m, int conj_v,
                             double *__restrict__ bm) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m * m) return;
  int64_t i = e / m, j = e % m;
  double val = 0.0;
  int64_t a = i / 2, b = j / 2;
  if (a < nb && b < nb && a <= b) {
    double vr = v[2 * (b * nb + a)], vi = v[2 * (b * nb + a) + 1];
    if (conj_v) vi = -vi;
    int ri = (int)(i & 1), rj = (int)(j & 1);
    if (ri == 0 && rj == 0) val = vr;
    else if (ri == 1 && rj == 0) val = -vi;
    else if (ri == 0 && rj == 1) val = vi;
    else val = vr;
  }
  bm[e] = val;
}

void launch_build_bmat(hipStream_t st, const double *v_cplx, int nb, int64_t m_cols, int conj_v,
                       double *bmat) {
  int64_t tot = m_cols * m_cols;
  hipLaunchKernelGGL(k_build_bmat, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, v_cplx,
                     nb, m_cols, conj_v, bmat);
}

// ---- layout conversions --------------------------------------------------------------------------
// h_block (2N x 2N real) <-> complex G: H[2a][2b] = Re, H[2a][2b+1] = -Im, H[2a+1][2b] = Im,
// H[2a+1][2b+1] = Re (quintic.rs:2148-2161). block -> complex reads the (2a+1, 2b) and (2a, 2b)
// entries, as unpack_h_block does (quintic.rs:5128-5140).
__global__ void k_cplx_to_block(const double *__restrict__ g, int n, double *__restrict__ h) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * n) return;
  int a = (int)(e / n), b = (int)(e % n);
  int64_t tn = 2 * (int64_t)n;
  double r = g[2 * e], i = g[2 * e + 1];
  h[(2 * a) * tn + 2 * b] = r;
  h[(2 * a) * tn + 2 * b + 1] = -i;
  h[(2 * a + 1) * tn + 2 * b] = i;
  h[(2 * a + 1) * tn + 2 * b + 1] = r;
}

void launch_cplx_to_block(hipStream_t st, const double *g_cplx, int nb, double *h_block) {
  int64_t tot = (int64_t)nb * nb;
  hipLaunchKernelGGL(k_cplx_to_block, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, g_cplx,
                     nb, h_block);
}

__global__ void k_block_to_cplx(const double *__restrict__ h, int n, double *__restrict__ g) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * n) return;
  int a = (int)(e / n), b = (int)(e % n);
  int64_t tn = 2 * (int64_t)n;
  g[2 * e] = h[(2 * a) * tn + 2 * b];
  g[2 * e + 1] = h[(2 * a + 1) * tn + 2 * b];
}

void launch_block_to_cplx(hipStream_t st, const double *h_block, int nb, double *g_cplx) {
  int64_t tot = (int64_t)nb * nb;
  hipLaunchKernelGGL(k_block_to_cplx, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                     h_block, nb, g_cplx);
}

__global__ void k_identity_cplx(int n, double *__restrict__ g) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * n) return;
  g[2 * e] = (e / n == e % n) ? 1.0 : 0.0;
  g[2 * e + 1] = 0.0;
}

void launch_identity_cplx(hipStream_t st, int nb, double *g_cplx) {
  int64_t tot = (int64_t)nb * nb;
  hipLaunchKernelGGL(k_identity_cplx, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, nb,
                     g_cplx);
}

// Hermitian part of a complex matrix: (A + A^H)/2 (user-supplied h_block inputs).
__global__ void k_hermitian_part(double *__restrict__ g, int n) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * n) return;
  int a = (int)(e / n), b = (int)(e % n);
  if (a > b) return;
  int64_t f = (int64_t)b * n + a;
  double r = 0.5 * (g[2 * e] + g[2 * f]), i = 0.5 * (g[2 * e + 1] - g[2 * f + 1]);
  if (a == b) i = 0.0;
  g[2 * e] = r;
  g[2 * e + 1] = i;
  g[2 * f] = r;
  g[2 * f + 1] = -i;
}

void launch_hermitian_part(hipStream_t st, double *g_cplx, int nb) {
  int64_t tot = (int64_t)nb * nb;
  hipLaunchKernelGGL(k_hermitian_part, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                     g_cplx, nb);
}

// ---- legacy un-inverted update (donaldson_iteration_quintic, quintic.rs:4140-4222) --------------
// resid = sqrt(2 sum |t - g|^2) (the 2N x 2N block difference), then g = t. One workgroup, fixed
// summation order.
__global__ void __launch_bounds__(1024) k_legacy_replace(const double *__restrict__ t, int nb,
                                                         double *__restrict__ g,
                                                         double *__restrict__ resid) {
  __shared__ double sh[1024];
  const int64_t tot = (int64_t)nb * nb;
  double acc = 0.0;
  for (int64_t e = threadIdx.x; e < tot; e += 1024) {
    const double dr = t[2 * e] - g[2 * e], di = t[2 * e + 1] - g[2 * e + 1];
    acc += 2.0 * (dr * dr + di * di);
    g[2 * e] = t[2 * e];
    g[2 * e + 1] = t[2 * e + 1];
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 512; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) *resid = sqrt(sh[0]);
}

void launch_legacy_replace(hipStream_t st, const double *t, int nb, double *g, double *resid) {
  hipLaunchKernelGGL(k_legacy_replace, dim3(1), dim3(1024), 0, st, t, nb, g, resid);
}

__global__ void k_fill_f64(double *__restrict__ x, int64_t n, double v) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) x[e] = v;
}

void launch_fill_f64(hipStream_t st, double *x, int64_t n, double v) {
  hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, n, v);
}

}  // namespace cy3q
