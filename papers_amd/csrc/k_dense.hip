// N x N FP64 work of one Donaldson step, and the per-point weight coefficients.
//
// Reference step (cy3/src/quintic.rs:1734-1894): T = (1/sum w) sum_p (w_p/K_p) conj(s) s^T,
// then T <- T * N / tr T, then G_next = T^-1 by LU of the 2N x 2N real block with a
// Hermitian projection and an identity fallback on failure, then residual = ||H_next - H||_F.
//
// Here T is Hermitian positive definite, so G_next = T^-1 = L^-H L^-1 with T = L L^H
// (complex Cholesky, FP64, on the GPU). The next step's D pass uses W = L^-1 directly,
// D_p = ||W s_p||^2, a sum of squares with no cancellation. G = W^H W is formed only for the
// residual and the API's h_block. A non-PD T takes the reference's identity fallback.
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

// ---- T from the real-ified symmetric product ------------------------------------------------
// xx = X^T diag(c) X (M x M, full). T_ab = xx[2a][2b] + xx[2a+1][2b+1]
//                                       + i (xx[2a][2b+1] - xx[2a+1][2b])   (quintic.rs:1783-1786)
// then T /= sum_w (if sum_w > 1e-12, else T = 0) and, if trace_normalise, T *= N / tr T when
// tr T > 1e-10 (quintic.rs:1790-1812). xx is first multiplied by scale (and by dscale[0]
// when given): the exact power of two of the split path's global scaling, 1 for FP64.
__global__ void __launch_bounds__(1024) k_finish_t(const double *__restrict__ xx, int64_t m,
                                                   int nb, const double *__restrict__ sum_w,
                                                   int trace_normalise, double scale,
                                                   const double *__restrict__ dscale,
                                                   double *__restrict__ t) {
  __shared__ double sh[1024];
  if (dscale) scale *= dscale[0];
  const double sw = *sum_w;
  const double inv_w = sw > 1e-12 ? 1.0 / sw : 0.0;
  double tr = 0.0;
  for (int64_t idx = threadIdx.x; idx < (int64_t)nb * nb; idx += 1024) {
    int a = (int)(idx / nb), b = (int)(idx % nb);
    double rr = xx[(2 * a) * m + 2 * b] * scale, ii = xx[(2 * a + 1) * m + 2 * b + 1] * scale;
    double ri = xx[(2 * a) * m + 2 * b + 1] * scale, ir = xx[(2 * a + 1) * m + 2 * b] * scale;
    double re = (rr + ii) * inv_w, im = (ri - ir) * inv_w;
    t[2 * idx] = re;
    t[2 * idx + 1] = im;
    if (a == b) tr += re;
  }
  sh[threadIdx.x] = tr;
  __syncthreads();
  for (int off = 512; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  const double trace = sh[0];
  if (!trace_normalise || !(trace > 1e-10)) return;
  const double s = (double)nb / trace;
  for (int64_t idx = threadIdx.x; idx < (int64_t)nb * nb; idx += 1024) {
    t[2 * idx] *= s;
    t[2 * idx + 1] *= s;
  }
}

void launch_finish_t(hipStream_t st, const double *xx, int64_t m_cols, int nb,
                     const double *sum_w, int trace_normalise, double scale,
                     const double *dscale, double *t_cplx) {
  hipLaunchKernelGGL(k_finish_t, dim3(1), dim3(1024), 0, st, xx, m_cols, nb, sum_w,
                     trace_normalise, scale, dscale, t_cplx);
}

// ---- Cholesky -------------------------------------------------------------------------------
// In-place lower Cholesky A = L L^H of a Hermitian matrix (lower triangle read), one workgroup,
// right-looking by columns (the column recurrences of cholesky_complex_hermitian,
// quintic.rs:3002-3049). The strict upper triangle is zeroed. status = 0, or ST_ERR_NOT_PD when a
// pivot is <= 0 or not finite.
__global__ void __launch_bounds__(1024) k_cholesky(double *__restrict__ a, int n,
                                                   int32_t *__restrict__ status) {
  __shared__ double ljj;
  __shared__ int bad;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  for (int j = 0; j < n; j++) {
    if (threadIdx.x == 0) {
      double d = a[2 * ((int64_t)j * n + j)];
      if (!(d > 0.0) || !isfinite(d)) {
        bad = 1;
        ljj = 1.0;
      } else {
        ljj = sqrt(d);
      }
      a[2 * ((int64_t)j * n + j)] = ljj;
      a[2 * ((int64_t)j * n + j) + 1] = 0.0;
    }
    __syncthreads();
    if (bad) break;
    const double inv = 1.0 / ljj;
    for (int i = j + 1 + threadIdx.x; i < n; i += 1024) {
      a[2 * ((int64_t)i * n + j)] *= inv;
      a[2 * ((int64_t)i * n + j) + 1] *= inv;
    }
    __syncthreads();
    // Trailing update of the lower triangle: A_ik -= L_ij conj(L_kj), j < k <= i.
    const int m = n - j - 1;
    const int64_t cnt = (int64_t)m * (m + 1) / 2;
    for (int64_t e = threadIdx.x; e < cnt; e += 1024) {
      // e -> (ii, kk) with 0 <= kk <= ii < m, row-major over the packed lower triangle.
      int ii = (int)((sqrt(8.0 * (double)e + 1.0) - 1.0) * 0.5);
      while ((int64_t)(ii + 1) * (ii + 2) / 2 <= e) ii++;
      while ((int64_t)ii * (ii + 1) / 2 > e) ii--;
      int kk = (int)(e - (int64_t)ii * (ii + 1) / 2);
      int i = j + 1 + ii, k = j + 1 + kk;
      double lir = a[2 * ((int64_t)i * n + j)], lii = a[2 * ((int64_t)i * n + j) + 1];
      double lkr = a[2 * ((int64_t)k * n + j)], lki = a[2 * ((int64_t)k * n + j) + 1];
      // L_ij * conj(L_kj)
      a[2 * ((int64_t)i * n + k)] -= lir * lkr + lii * lki;
      a[2 * ((int64_t)i * n + k) + 1] -= lii * lkr - lir * lki;
    }
    __syncthreads();
  }
  __syncthreads();
  for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += 1024) {
    int i = (int)(e / n), k = (int)(e % n);
    if (k > i) {
      a[2 * e] = 0.0;
      a[2 * e + 1] = 0.0;
    }
  }
  if (threadIdx.x == 0) *status = bad ? ST_ERR_NOT_PD : ST_OK;
}

void launch_cholesky(hipStream_t st, double *a_cplx, int nb, int32_t *status) {
  hipLaunchKernelGGL(k_cholesky, dim3(1), dim3(1024), 0, st, a_cplx, nb, status);
}

// ---- Triangular inverse ----------------------------------------------------------------------
// V = L^-1 for lower-triangular L. Columns of V are independent: wave c solves
// L v = e_c by forward substitution. Lane l holds the rows r = l (mod 64) of v in LDS-free
// registers... the column lives in LDS, one slice per wave, and each dot product is split across
// the 64 lanes and reduced with cross-lane shuffles.
static constexpr int TRI_WAVES = 4;
static constexpr int TRI_NMAX = 512;

__global__ void __launch_bounds__(256) k_tri_inverse(const double *__restrict__ l, int n,
                                                     double *__restrict__ v) {
  __shared__ double col[TRI_WAVES][TRI_NMAX][2];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int c = blockIdx.x * TRI_WAVES + wv;
  if (c >= n) return;
  for (int r = lane; r < n; r += 64) {
    col[wv][r][0] = 0.0;
    col[wv][r][1] = 0.0;
  }
  __builtin_amdgcn_wave_barrier();
  for (int i = c; i < n; i++) {
    // s = delta_ic - sum_{k=c}^{i-1} L_ik v_k
    double sr = 0.0, si = 0.0;
    for (int k = c + lane; k < i; k += 64) {
      double ar = l[2 * ((int64_t)i * n + k)], ai = l[2 * ((int64_t)i * n + k) + 1];
      double br = col[wv][k][0], bi = col[wv][k][1];
      sr += ar * br - ai * bi;
      si += ar * bi + ai * br;
    }
    for (int off = 32; off > 0; off >>= 1) {
      sr += __shfl_xor(sr, off);
      si += __shfl_xor(si, off);
    }
    if (lane == 0) {
      double lii = l[2 * ((int64_t)i * n + i)];
      double rr = (i == c ? 1.0 : 0.0) - sr, ri = -si;
      col[wv][i][0] = rr / lii;
      col[wv][i][1] = ri / lii;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  }
  __builtin_amdgcn_wave_barrier();
  for (int r = lane; r < n; r += 64) {
    v[2 * ((int64_t)r * n + c)] = col[wv][r][0];
    v[2 * ((int64_t)r * n + c) + 1] = col[wv][r][1];
  }
}

void launch_tri_inverse(hipStream_t st, const double *l_cplx, int nb, double *linv_cplx) {
  hipLaunchKernelGGL(k_tri_inverse, dim3((unsigned)((nb + TRI_WAVES - 1) / TRI_WAVES)), dim3(256),
                     0, st, l_cplx, nb, linv_cplx);
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

// ---- G = W^H W --------------------------------------------------------------------------------
// G_ab = sum_{k >= max(a,b)} conj(W_ka) W_kb for lower-triangular W. Each (a, b) and (b, a) is
// evaluated in the same order, so G is exactly Hermitian with an exactly real diagonal.
__global__ void k_gram_from_w(const double *__restrict__ w, int n, double *__restrict__ g) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * n) return;
  int a = (int)(e / n), b = (int)(e % n);
  int k0 = a > b ? a : b;
  double sr = 0.0, si = 0.0;
  for (int k = k0; k < n; k++) {
    double xr = w[2 * ((int64_t)k * n + a)], xi = w[2 * ((int64_t)k * n + a) + 1];
    double yr = w[2 * ((int64_t)k * n + b)], yi = w[2 * ((int64_t)k * n + b) + 1];
    sr += xr * yr + xi * yi;
    si += xr * yi - xi * yr;
  }
  g[2 * e] = sr;
  g[2 * e + 1] = si;
}

void launch_gram_from_w(hipStream_t st, const double *w_cplx, int nb, double *g_cplx) {
  int64_t tot = (int64_t)nb * nb;
  hipLaunchKernelGGL(k_gram_from_w, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, w_cplx,
                     nb, g_cplx);
}

// ---- residual, fallback, swap -------------------------------------------------------------------
// If the Cholesky failed, the new G and W become the identity (quintic.rs:1874-1883).
// residual = ||H_new - H_old||_F over the 2N x 2N block = sqrt(2 sum |G_new - G_old|^2)
// (each complex entry appears twice in re and twice in im). Then G_old <- G_new.
__global__ void __launch_bounds__(1024) k_residual_and_swap(double *__restrict__ g_new,
                                                            double *__restrict__ w_new,
                                                            double *__restrict__ g_old, int n,
                                                            const int32_t *__restrict__ status,
                                                            double *__restrict__ residual) {
  __shared__ double sh[1024];
  const bool fail = status && *status != ST_OK;
  double s = 0.0;
  for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += 1024) {
    int a = (int)(e / n), b = (int)(e % n);
    double nr = g_new[2 * e], ni = g_new[2 * e + 1];
    if (fail) {
      nr = (a == b) ? 1.0 : 0.0;
      ni = 0.0;
      g_new[2 * e] = nr;
      g_new[2 * e + 1] = ni;
      if (w_new) {
        w_new[2 * e] = nr;
        w_new[2 * e + 1] = 0.0;
      }
    }
    double dr = nr - g_old[2 * e], di = ni - g_old[2 * e + 1];
    s += 2.0 * (dr * dr + di * di);
    g_old[2 * e] = nr;
    g_old[2 * e + 1] = ni;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 512; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) *residual = sqrt(sh[0]);
}

void launch_residual_and_swap(hipStream_t st, double *g_new, double *w_new, double *g_old, int nb,
                              const int32_t *chol_status, double *residual_out) {
  hipLaunchKernelGGL(k_residual_and_swap, dim3(1), dim3(1024), 0, st, g_new, w_new, g_old, nb,
                     chol_status, residual_out);
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

}  // namespace cy3q
