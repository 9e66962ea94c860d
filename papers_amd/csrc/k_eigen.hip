// FP64 Hermitian eigensolver and complex GEMM for the Galerkin eigenproblems.
//
// Reference: hermitian_jacobi_n (route34/metric_laplacian.rs:616-693), the cyclic complex
// Jacobi method, one rotation at a time. The GPU form is the parallel ("tournament") Jacobi
// method. In each round, the n/2 disjoint slot pairs (2u, 2u+1) rotate at once. Every 2 x 2
// block of A' = J^H A J depends only on the 2 x 2 block of A and the two pairs' rotations, so a
// round is one pass over the matrix. The output is written with the circle-method slot
// permutation, so the next round's pairs are again adjacent slots. Over n-1 rounds (one sweep),
// every pair of indices meets exactly once. The rotation for a pair (a, d real, b = |b| e^{i phi})
// is the reference's: tan 2 theta = 2|b| / (a - d), via t = tan theta = sign(tau) /
// (|tau| + sqrt(1 + tau^2)), tau = (a - d) / (2|b|), so
//   J = [[c, -s e^{i phi}], [s e^{-i phi}, c]]   (right factor, columns p, q).
// Eigenvectors, when wanted, are V J accumulated the same way (rows = basis coordinates).
#include "cy3q_internal.h"

namespace cy3q {

static constexpr int JB = 16;  // 16 x 16 thread tiles

__device__ inline int slot_next(int x, int h) {
  if (x == 0 || h == 1) return x;
  const int k = x >> 1;
  if ((x & 1) == 0) return k < h - 1 ? x + 2 : x + 1;  // t_k -> t_k+1, t_h-1 -> b_h-1
  return k >= 1 ? x - 2 : 2;                           // b_k -> b_k-1, b_0 -> t_1
}

struct Rot {
  double c, s, pr, pi;  // cos, sin, e^{i phi}
};

__device__ inline Rot rot_of(const double2 *__restrict__ A, int np, int u) {
  const int p = 2 * u, q = p + 1;
  const double a = A[(int64_t)p * np + p].x, d = A[(int64_t)q * np + q].x;
  const double2 b = A[(int64_t)p * np + q];
  const double ab = hypot(b.x, b.y);
  Rot r{1.0, 0.0, 1.0, 0.0};
  if (!(ab > 1e-300)) return r;
  const double tau = (a - d) / (2.0 * ab);
  const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
  r.c = 1.0 / sqrt(1.0 + t * t);
  r.s = t * r.c;
  r.pr = b.x / ab;
  r.pi = b.y / ab;
  return r;
}

__device__ inline double2 cm(double2 a, double br, double bi) {
  return make_double2(a.x * br - a.y * bi, a.x * bi + a.y * br);
}
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }

// One round. blockIdx.y < n_uy: A tiles (u rows x v cols of 2 x 2 blocks); otherwise V tiles
// (rows i, v pairs). tag tracks which basis index sits in each slot.
__global__ void __launch_bounds__(JB * JB) k_jacobi_round(const double2 *__restrict__ A,
                                                          double2 *__restrict__ A2,
                                                          const double2 *__restrict__ V,
                                                          double2 *__restrict__ V2,
                                                          const int *__restrict__ tag,
                                                          int *__restrict__ tag2, int np,
                                                          int n_uy) {
  __shared__ Rot s_rv[JB], s_ru[JB];
  const int h = np / 2;
  const int tx = threadIdx.x % JB, ty = threadIdx.x / JB;
  const int v = blockIdx.x * JB + tx;
  const bool a_part = (int)blockIdx.y < n_uy;
  if (ty == 0 && blockIdx.x * JB + tx < h) s_rv[tx] = rot_of(A, np, blockIdx.x * JB + tx);
  if (a_part && ty == 1 && (int)blockIdx.y * JB + tx < h)
    s_ru[tx] = rot_of(A, np, blockIdx.y * JB + tx);
  __syncthreads();
  if (v >= h) return;
  const Rot rv = s_rv[tx];
  const int pv = 2 * v, qv = pv + 1;
  const int npv = slot_next(pv, h), nqv = slot_next(qv, h);
  if (a_part) {
    const int u = blockIdx.y * JB + ty;
    if (u >= h) return;
    if (u == 0) {
      tag2[npv] = tag[pv];
      tag2[nqv] = tag[qv];
    }
    const Rot ru = s_ru[ty];
    const int pu = 2 * u, qu = pu + 1;
    const double2 a00 = A[(int64_t)pu * np + pv], a01 = A[(int64_t)pu * np + qv];
    const double2 a10 = A[(int64_t)qu * np + pv], a11 = A[(int64_t)qu * np + qv];
    // B = A_blk J_v
    const double jr = rv.s * rv.pr, ji = rv.s * rv.pi;  // s e^{i phi}
    const double2 b00 = cadd(cm(a00, rv.c, 0.0), cm(a01, jr, -ji));
    const double2 b01 = cadd(cm(a00, -jr, -ji), cm(a01, rv.c, 0.0));
    const double2 b10 = cadd(cm(a10, rv.c, 0.0), cm(a11, jr, -ji));
    const double2 b11 = cadd(cm(a10, -jr, -ji), cm(a11, rv.c, 0.0));
    // A' = J_u^H B: row0 = c B0 + s e^{i phi} B1, row1 = -s e^{-i phi} B0 + c B1
    const double ur = ru.s * ru.pr, ui = ru.s * ru.pi;
    double2 c00 = cadd(cm(b00, ru.c, 0.0), cm(b10, ur, ui));
    double2 c01 = cadd(cm(b01, ru.c, 0.0), cm(b11, ur, ui));
    double2 c10 = cadd(cm(b00, -ur, ui), cm(b10, ru.c, 0.0));
    double2 c11 = cadd(cm(b01, -ur, ui), cm(b11, ru.c, 0.0));
    if (u == v) {  // the rotated pair: exact zero off-diagonal, real diagonal
      c00.y = 0.0;
      c11.y = 0.0;
      c01 = make_double2(0.0, 0.0);
      c10 = make_double2(0.0, 0.0);
    }
    const int npu = slot_next(pu, h), nqu = slot_next(qu, h);
    A2[(int64_t)npu * np + npv] = c00;
    A2[(int64_t)npu * np + nqv] = c01;
    A2[(int64_t)nqu * np + npv] = c10;
    A2[(int64_t)nqu * np + nqv] = c11;
  } else {
    const int i = (blockIdx.y - n_uy) * JB + ty;
    if (i >= np) return;
    const double jr = rv.s * rv.pr, ji = rv.s * rv.pi;
    const double2 x0 = V[(int64_t)i * np + pv], x1 = V[(int64_t)i * np + qv];
    V2[(int64_t)i * np + npv] = cadd(cm(x0, rv.c, 0.0), cm(x1, jr, -ji));
    V2[(int64_t)i * np + nqv] = cadd(cm(x0, -jr, -ji), cm(x1, rv.c, 0.0));
  }
}

void launch_jacobi_round(hipStream_t st, const double2 *A, double2 *A2, const double2 *V,
                         double2 *V2, const int *tag, int *tag2, int np, int want_v) {
  const int h = np / 2;
  const int nvx = (h + JB - 1) / JB, n_uy = (h + JB - 1) / JB;
  const int n_iy = want_v ? (np + JB - 1) / JB : 0;
  hipLaunchKernelGGL(k_jacobi_round, dim3((unsigned)nvx, (unsigned)(n_uy + n_iy)), dim3(JB * JB),
                     0, st, A, A2, V, V2, tag, tag2, np, n_uy);
}

// out[0] = sum_{i != j} |A_ij|^2, out[1] = sum |A_ij|^2; two passes with a fixed order.
__global__ void __launch_bounds__(256) k_offnorm_part(const double2 *__restrict__ A, int np,
                                                      double *__restrict__ part) {
  __shared__ double s0[256], s1[256];
  const int64_t tot = (int64_t)np * np;
  double off = 0.0, all = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < tot; e += (int64_t)gridDim.x * 256) {
    const double2 a = A[e];
    const double v = a.x * a.x + a.y * a.y;
    all += v;
    if (e / np != e % np) off += v;
  }
  s0[threadIdx.x] = off;
  s1[threadIdx.x] = all;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s0[threadIdx.x] += s0[threadIdx.x + o];
      s1[threadIdx.x] += s1[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s0[0];
    part[2 * blockIdx.x + 1] = s1[0];
  }
}

__global__ void k_offnorm_final(const double *__restrict__ part, int nparts,
                                double *__restrict__ out) {
  if (threadIdx.x != 0) return;
  double a = 0.0, b = 0.0;
  for (int i = 0; i < nparts; i++) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  out[0] = a;
  out[1] = b;
}

int offnorm_parts() { return 256; }

void launch_offnorm(hipStream_t st, const double2 *A, int np, double *part, double *out2) {
  hipLaunchKernelGGL(k_offnorm_part, dim3(256), dim3(256), 0, st, A, np, part);
  hipLaunchKernelGGL(k_offnorm_final, dim3(1), dim3(64), 0, st, part, 256, out2);
}

// dst (np x np) = src (n x n) zero-padded, plus shift on the real diagonal; V = I; tag = slot.
__global__ void k_jacobi_init(const double2 *__restrict__ src, int n, int np, double shift,
                              double2 *__restrict__ A, double2 *__restrict__ V,
                              int *__restrict__ tag) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)np * np) return;
  const int i = (int)(e / np), j = (int)(e % np);
  double2 v = make_double2(0.0, 0.0);
  if (i < n && j < n) v = src[(int64_t)i * n + j];
  if (i == j) {
    v.y = 0.0;
    if (i < n) v.x += shift;
  }
  A[e] = v;
  if (V) V[e] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  if (i == 0) tag[j] = j;
}

void launch_jacobi_init(hipStream_t st, const double2 *src, int n, int np, double shift,
                        double2 *A, double2 *V, int *tag) {
  const int64_t tot = (int64_t)np * np;
  hipLaunchKernelGGL(k_jacobi_init, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, src, n,
                     np, shift, A, V, tag);
}

// diag[x] = Re A[x][x]
__global__ void k_diag_re(const double2 *__restrict__ A, int np, double *__restrict__ d) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x < np) d[x] = A[(int64_t)x * np + x].x;
}

void launch_diag_re(hipStream_t st, const double2 *A, int np, double *d) {
  hipLaunchKernelGGL(k_diag_re, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, A, np, d);
}

// out[i][j] = V[i][col[j]] * scale[j], i < n (rows of the padded V), j < r.
__global__ void k_gather_cols(const double2 *__restrict__ V, int np, int n,
                              const int *__restrict__ col, const double *__restrict__ scale, int r,
                              double2 *__restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * r) return;
  const int i = (int)(e / r), j = (int)(e % r);
  const double2 v = V[(int64_t)i * np + col[j]];
  const double s = scale ? scale[j] : 1.0;
  out[e] = make_double2(v.x * s, v.y * s);
}

void launch_gather_cols(hipStream_t st, const double2 *V, int np, int n, const int *col,
                        const double *scale, int r, double2 *out) {
  const int64_t tot = (int64_t)n * r;
  hipLaunchKernelGGL(k_gather_cols, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, V, np, n,
                     col, scale, r, out);
}

// Hermitian projection in place (metric_laplacian.rs:356-368): real diagonal, A_ij = A_ji* =
// (A_ij + conj(A_ji)) / 2 for i < j.
__global__ void k_herm_project(double2 *__restrict__ A, int n) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * n) return;
  const int i = (int)(e / n), j = (int)(e % n);
  if (i > j) return;
  if (i == j) {
    A[e].y = 0.0;
    return;
  }
  const double2 a = A[(int64_t)i * n + j], b = A[(int64_t)j * n + i];
  const double2 avg = make_double2((a.x + b.x) * 0.5, (a.y - b.y) * 0.5);
  A[(int64_t)i * n + j] = avg;
  A[(int64_t)j * n + i] = make_double2(avg.x, -avg.y);
}

void launch_herm_project(hipStream_t st, double2 *A, int n) {
  const int64_t tot = (int64_t)n * n;
  hipLaunchKernelGGL(k_herm_project, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, A, n);
}

// ---------------------------------------------------------------------------------------------
// C = op(A) op(B), complex FP64. op 0: as stored (row-major, ld = columns), op 1: conjugate
// transpose. 64 x 64 output tile, 256 threads with 4 x 4 outputs each, K tile 16.
// ---------------------------------------------------------------------------------------------
static constexpr int ZG_T = 64, ZG_K = 16;

__global__ void __launch_bounds__(256) k_zgemm(const double2 *__restrict__ A,
                                               const double2 *__restrict__ B,
                                               double2 *__restrict__ C, int M, int N, int K,
                                               int lda, int ldb, int ldc, int opa, int opb) {
  __shared__ double2 sa[ZG_K][ZG_T + 1], sb[ZG_K][ZG_T + 1];
  const int tid = threadIdx.x;
  const int tr = tid / 16, tc = tid % 16;
  const int m0 = blockIdx.y * ZG_T, n0 = blockIdx.x * ZG_T;
  double2 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) acc[a][b] = make_double2(0.0, 0.0);
  for (int k0 = 0; k0 < K; k0 += ZG_K) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int e = tid + 256 * u;  // 1024 = 16 x 64
      const int kk = e / ZG_T, mm = e % ZG_T;
      const int gk = k0 + kk, gm = m0 + mm, gn = n0 + mm;
      double2 va = make_double2(0.0, 0.0), vb = make_double2(0.0, 0.0);
      if (gk < K && gm < M) {
        if (opa == 0) {
          va = A[(int64_t)gm * lda + gk];
        } else {
          va = A[(int64_t)gk * lda + gm];
          va.y = -va.y;
        }
      }
      if (gk < K && gn < N) {
        if (opb == 0) {
          vb = B[(int64_t)gk * ldb + gn];
        } else {
          vb = B[(int64_t)gn * ldb + gk];
          vb.y = -vb.y;
        }
      }
      sa[kk][mm] = va;
      sb[kk][mm] = vb;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < ZG_K; kk++) {
      double2 av[4], bv[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        av[q] = sa[kk][tr + 16 * q];
        bv[q] = sb[kk][tc + 16 * q];
      }
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
          acc[a][b].x += av[a].x * bv[b].x - av[a].y * bv[b].y;
          acc[a][b].y += av[a].x * bv[b].y + av[a].y * bv[b].x;
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const int gm = m0 + tr + 16 * a, gn = n0 + tc + 16 * b;
      if (gm < M && gn < N) C[(int64_t)gm * ldc + gn] = acc[a][b];
    }
}

void launch_zgemm(hipStream_t st, const double2 *A, const double2 *B, double2 *C, int M, int N,
                  int K, int lda, int ldb, int ldc, int opa, int opb) {
  hipLaunchKernelGGL(k_zgemm, dim3((unsigned)((N + ZG_T - 1) / ZG_T), (unsigned)((M + ZG_T - 1) / ZG_T)),
                     dim3(256), 0, st, A, B, C, M, N, K, lda, ldb, ldc, opa, opb);
}

// a[i][i] += shift (real), i < n.
__global__ void k_add_diag(double2 *__restrict__ a, int n, double shift) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[(int64_t)i * n + i].x += shift;
}

void launch_add_diag(hipStream_t st, double2 *a, int n, double shift) {
  hipLaunchKernelGGL(k_add_diag, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, n, shift);
}

}  // namespace cy3q
