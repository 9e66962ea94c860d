// The N x N FP64 part of a Donaldson step, on the GPU: T from the contraction, then the
// inverse Cholesky factor W = L^-1 of T (T = L L^H), then G = W^H W and the step residual.
//
// Reference (cy3/src/quintic.rs:1790-1894): T /= sum w; T *= N / tr T; G_next = T^-1 by LU of the
// 2N x 2N real block; Hermitian projection; identity on failure; residual ||H_next - H||_F.
// Here T is Hermitian positive definite, so T^-1 = W^H W with W = L^-1 lower triangular. W is
// what the next step's contractions consume (D_p = |W s_p|^2). G = W^H W is formed for the
// residual and the caller's h_block. It is exactly Hermitian: G_ab and G_ba are summed in the
// same order.
//
// W comes from one blocked elimination of the augmented matrix [T | I] in 32 x 32 complex tiles.
// For each block column J:
//   k_elim_panel  (1 workgroup): factors the diagonal tile and inverts it in the same 32-step
//                 sweep (one barrier per column; the tile lives in registers, one complex value
//                 per element, and only the pivot column and row go through LDS): W_JJ = L_JJ^-1.
//   k_elim_update (one workgroup per tile): L_IJ = A_IJ L_JJ^-H, then A_IK -= L_IJ L_KJ^H for
//                 J < K <= I, and R_IK -= L_IJ W_JK for K <= J, where R is the updated right
//                 half and W_JK = L_JJ^-1 R_JK (K < J) is formed in the same kernel. The tile
//                 products run on FP64 MFMA.
//   k_elim_lastrow: the last block row of W.
// The chain is latency-bound (one panel and one update per 32 columns), so every kernel fetches
// all its tiles at once and keeps LDS traffic low.
// A non-positive or non-finite pivot sets status = ST_ERR_NOT_PD. Later kernels then return
// early, and the gram kernel installs the reference's identity fallback (G = W = I).
// Complex matrices are row-major [n][n] of (re, im).
#include "cy3q_internal.h"

namespace cy3q {

static constexpr int TB = 32;  // tile edge

struct Cz {
  double r, i;
};

__device__ inline Cz cmul(Cz a, Cz b) { return Cz{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }

// Loads tile (r0, c0) of an n x n complex matrix into LDS, zero outside the matrix (or the
// identity on the diagonal when pad_identity). With 256 threads, all four global loads of a
// thread are issued before any LDS store, so the tile costs one memory latency.
__device__ inline void load_tile(Cz (*dst)[TB + 1], const double *__restrict__ src, int n, int r0,
                                 int c0, bool pad_identity, int tid, int nthr) {
  if (nthr == 256) {
    double2 v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int q = tid + 256 * u, r = q / TB, c = q % TB;
      v[u] = double2{(pad_identity && r == c) ? 1.0 : 0.0, 0.0};
      if (r0 + r < n && c0 + c < n)
        v[u] = *reinterpret_cast<const double2 *>(src + 2 * ((int64_t)(r0 + r) * n + c0 + c));
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int q = tid + 256 * u;
      dst[q / TB][q % TB] = Cz{v[u].x, v[u].y};
    }
    return;
  }
  for (int q = tid; q < TB * TB; q += nthr) {
    const int r = q / TB, c = q % TB;
    Cz z{(pad_identity && r == c) ? 1.0 : 0.0, 0.0};
    if (r0 + r < n && c0 + c < n) {
      const double2 v = *reinterpret_cast<const double2 *>(src + 2 * ((int64_t)(r0 + r) * n + c0 + c));
      z = Cz{v.x, v.y};
    }
    dst[r][c] = z;
  }
}

__device__ inline void store_tile(double *__restrict__ dst, int n, int r0, int c0,
                                  const Cz (*src)[TB + 1], int tid, int nthr) {
  for (int q = tid; q < TB * TB; q += nthr) {
    const int r = q / TB, c = q % TB;
    if (r0 + r < n && c0 + c < n) {
      double2 v;
      v.x = src[r][c].r;
      v.y = src[r][c].i;
      *reinterpret_cast<double2 *>(dst + 2 * ((int64_t)(r0 + r) * n + c0 + c)) = v;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// T from xx = X^T diag(c) X:  T_ab = s * (xx[2a][2b] + xx[2a+1][2b+1] + i (xx[2a][2b+1] -
// xx[2a+1][2b])), with s = hscale * dscale * inv_w * (N / tr) (quintic.rs:1783-1812;
// inv_w = 1/sum_w if sum_w > 1e-12 else 0; the trace factor only if trace_normalise and
// tr > 1e-10). Every workgroup of the fill computes s itself (one launch instead of two).
// ---------------------------------------------------------------------------------------------
// The scale s, computed by every workgroup of k_t_fill in the same order (so every workgroup
// gets the same bits): a strided sum over 256 threads, then a tree in LDS.
__device__ double t_scale_value(const double *__restrict__ xx, int64_t m, int nb,
                                const double *__restrict__ sum_w, double hscale,
                                const double *__restrict__ dscale, int trace_normalise,
                                double *sh) {
  double base = hscale;
  if (dscale) base *= dscale[0];
  const double sw = *sum_w;
  const double inv_w = sw > 1e-12 ? 1.0 / sw : 0.0;
  double tr = 0.0;
  for (int a = threadIdx.x; a < nb; a += 256)
    tr += (xx[(2 * a) * m + 2 * a] * base + xx[(2 * a + 1) * m + 2 * a + 1] * base) * inv_w;
  sh[threadIdx.x] = tr;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  const double trace = sh[0];
  const double s = (trace_normalise && trace > 1e-10) ? (double)nb / trace : 1.0;
  return base * inv_w * s;
}

// t (the caller-visible T), a (the elimination's left half), r (its right half, = I).
__global__ void __launch_bounds__(256) k_t_fill(const double *__restrict__ xx, int64_t m, int nb,
                                                const double *__restrict__ sum_w, double hscale,
                                                const double *__restrict__ dscale,
                                                int trace_normalise, double *__restrict__ t,
                                                double *__restrict__ a, double *__restrict__ r) {
  __shared__ double sh[256];
  const double s = t_scale_value(xx, m, nb, sum_w, hscale, dscale, trace_normalise, sh);
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nb * nb) return;
  const int ia = (int)(e / nb), ib = (int)(e % nb);
  const double rr = xx[(2 * ia) * m + 2 * ib], ii = xx[(2 * ia + 1) * m + 2 * ib + 1];
  const double ri = xx[(2 * ia) * m + 2 * ib + 1], ir = xx[(2 * ia + 1) * m + 2 * ib];
  const double re = (rr + ii) * s, im = (ri - ir) * s;
  if (t) {
    t[2 * e] = re;
    t[2 * e + 1] = im;
  }
  a[2 * e] = re;
  a[2 * e + 1] = im;
  r[2 * e] = ia == ib ? 1.0 : 0.0;
  r[2 * e + 1] = 0.0;
}

void launch_t_from_xx(hipStream_t st, const double *xx, int64_t m_cols, int nb,
                      const double *sum_w, double hscale, const double *dscale,
                      int trace_normalise, double *scale_buf, double *t, double *a, double *r) {
  (void)scale_buf;  // the scale is recomputed inside k_t_fill
  const int64_t tot = (int64_t)nb * nb;
  hipLaunchKernelGGL(k_t_fill, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, xx, m_cols,
                     nb, sum_w, hscale, dscale, trace_normalise, t, a, r);
}

// Copies a complex matrix into the elimination buffers (a = src, r = I).
__global__ void k_elim_init(const double *__restrict__ src, int nb, double *__restrict__ a,
                            double *__restrict__ r) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nb * nb) return;
  a[2 * e] = src[2 * e];
  a[2 * e + 1] = src[2 * e + 1];
  r[2 * e] = (e / nb == e % nb) ? 1.0 : 0.0;
  r[2 * e + 1] = 0.0;
}

void launch_elim_init(hipStream_t st, const double *src, int nb, double *a, double *r) {
  const int64_t tot = (int64_t)nb * nb;
  hipLaunchKernelGGL(k_elim_init, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, src, nb,
                     a, r);
}

// ---------------------------------------------------------------------------------------------
// Panel of block column J.
// ---------------------------------------------------------------------------------------------
// 1 / sqrt(x) for a positive normal x: the hardware estimate and two Newton steps
// y <- y + y (1 - x y^2) / 2 (each squares the relative error), to about 1 ulp.
__device__ inline double rsqrt_f64(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double t = fma(-x * y, y, 1.0);
  y = fma(0.5 * y, t, y);
  t = fma(-x * y, y, 1.0);
  return fma(0.5 * y, t, y);
}

// Factors the diagonal tile of block column jblk (in LDS, identity outside the matrix) and
// inverts it: W_JJ = E = L_JJ^-1 into einv and w. 256 threads; colb, rowb and e are LDS
// scratch. A bad pivot sets status = ST_ERR_NOT_PD (later kernels return early).
// Thread tid holds the elements (i, k) = (i0 + 8 u, k), i0 = tid / 32, k = tid % 32, u = 0..3,
// each as one complex value: the left element A_ik up to step k, where column k is the pivot
// column and is published, then the right element R_ik (still the identity's then: R_ik
// changes only at steps j >= k). So every step updates each element by one complex product,
// x -= d_ij b_k / d_jj, with b_k = conj(d_kj) (k > j) or r_jk (k <= j), with no selects.
// The 32 steps are unrolled (static buffer, row and register indices): about 120 instructions
// per wave and step. A step is a dependent chain (publish, barrier, read, 1/sqrt, update, next
// publish) whose LDS round trips set its length, about 1,050 cycles on MI355X; each step
// delivers the pivot column to all 1,024 elements as broadcast reads.
__device__ __attribute__((always_inline)) inline void panel_core(const Cz (*tile)[TB + 1], int n, int jblk, double *__restrict__ einv,
                           double *__restrict__ w, int32_t *__restrict__ status,
                           Cz (*colb)[TB], Cz (*rowb)[TB], Cz (*e)[TB + 1]) {
  const int tid = threadIdx.x;
  const int k = tid % TB, i0 = tid / TB;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);  // rows i0 = 2 wv, 2 wv + 1
  Cz x[4];
#pragma unroll
  for (int u = 0; u < 4; u++) x[u] = tile[i0 + 8 * u][k];
  // Pivots past the matrix (the last block column) belong to the identity padding: nothing
  // couples them to the matrix, so their rows of E are unit rows and their steps are skipped.
  const int jmax = n - jblk * TB < TB ? n - jblk * TB : TB;
  for (int q = tid; q < TB * TB; q += 256) {
    const int ii = q / TB, kk = q % TB;
    e[ii][kk] = Cz{(ii == kk && ii >= jmax) ? 1.0 : 0.0, 0.0};
  }
  // Step j: l_jj = sqrt(d_jj); row j of L^-1 is r_j / l_jj; rows i > j lose d_ij b_k / d_jj.
  // (Left elements above the diagonal change too; they are never read.) Buffer j & 1 is
  // rewritten at step j + 2, after the barrier of step j + 1, so its readers at step j are done.
  bool ok = true;
#pragma unroll
  for (int j = 0; j < TB; j++) {
    if (j < jmax) {
      const int b = j & 1, ji = j % 8, jo = j / 8;
      if (k == j) {  // publish column j of the left tile, then hold the right element
#pragma unroll
        for (int u = 0; u < 4; u++) {
          colb[b][i0 + 8 * u] = x[u];
          x[u] = Cz{i0 + 8 * u == j ? 1.0 : 0.0, 0.0};
        }
      }
      if (i0 == ji) rowb[b][k] = k <= j ? x[jo] : Cz{0.0, 0.0};  // row j of the right tile
      __syncthreads();
      // Row group u of this wave (rows 8 u + 2 w, 8 u + 2 w + 1) is final once
      // 8 u + 2 w + 1 <= j: its d_ij is neither read nor used.
      const bool lf = k > j;
      const double piv = colb[b][j].r;
      Cz bk = lf ? colb[b][k] : rowb[b][k];
      Cz dl[4];
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (8 * u + 2 * wv + 1 > j) dl[u] = colb[b][i0 + 8 * u];
      // A bad pivot (the same for every thread) marks the factorisation failed; the remaining
      // steps run on a unit pivot and their results are discarded.
      ok = ok && piv > 0.0 && isfinite(piv);
      const double inv_l = rsqrt_f64(ok ? piv : 1.0), ipiv = inv_l * inv_l;
      if (i0 == ji && !lf) e[j][k] = Cz{bk.r * inv_l, bk.i * inv_l};
      if (lf) bk.i = -bk.i;
      const Cz bs = Cz{bk.r * ipiv, bk.i * ipiv};
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (8 * u + 2 * wv + 1 > j && i0 + 8 * u > j) {
          x[u].r = fma(-dl[u].r, bs.r, fma(dl[u].i, bs.i, x[u].r));
          x[u].i = fma(-dl[u].r, bs.i, fma(-dl[u].i, bs.r, x[u].i));
        }
    }
  }
  __syncthreads();
  if (!ok) {
    if (tid == 0) *status = ST_ERR_NOT_PD;
    return;
  }
  // L_JJ^-1 for the update kernel (a 32 x 32 tile per block column), and W_JJ = L_JJ^-1.
  for (int q = tid; q < TB * TB; q += 256) {
    const int i = q / TB, kk = q % TB;
    double2 v;
    v.x = e[i][kk].r;
    v.y = e[i][kk].i;
    *reinterpret_cast<double2 *>(einv + 2 * ((int64_t)jblk * TB * TB + q)) = v;
  }
  store_tile(w, n, jblk * TB, jblk * TB, e, tid, 256);
}

// The first block column's panel (the update kernel of block column J factors panel J + 1).
__global__ void __launch_bounds__(256) k_elim_panel(const double *__restrict__ a, int n,
                                                    double *__restrict__ einv,
                                                    double *__restrict__ w,
                                                    int32_t *__restrict__ status) {
  __shared__ Cz colb[2][TB], rowb[2][TB];
  __shared__ Cz tile[TB][TB + 1], e[TB][TB + 1];
  if (threadIdx.x == 0) *status = ST_OK;
  load_tile(tile, a, n, 0, 0, true, threadIdx.x, 256);
  __syncthreads();
  panel_core(tile, n, 0, einv, w, status, colb, rowb, e);
}

// ---------------------------------------------------------------------------------------------
// 32 x 32 x 32 complex tile products on v_mfma_f64_16x16x4f64. Wave w computes the 16 x 16
// quadrant (qr, qc) = (w / 2, w % 2) of C = op(A) op(B), op 0 = as stored, 1 = conjugate
// transpose, as four real products per 4-deep k step (re += ar br - ai bi, im += ar bi + ai br).
// MFMA f64 layout: A[i][k] at lane i + 16k, B[k][j] at lane j + 16k; output element e of lane l
// is (row 4e + l/16, col l%16) of the quadrant.
// ---------------------------------------------------------------------------------------------
__device__ inline void cgemm32(const Cz (*A)[TB + 1], int opa, const Cz (*B)[TB + 1], int opb,
                               int qr, int qc, int lane, double4_t &cr, double4_t &ci) {
  const int li = lane & 15, lk = lane >> 4;
  const int ra = 16 * qr + li, cb = 16 * qc + li;
#pragma unroll
  for (int kk = 0; kk < TB / 4; kk++) {
    const int kx = 4 * kk + lk;
    Cz av = opa ? A[kx][ra] : A[ra][kx];
    if (opa) av.i = -av.i;
    Cz bv = opb ? B[cb][kx] : B[kx][cb];
    if (opb) bv.i = -bv.i;
    cr = __builtin_amdgcn_mfma_f64_16x16x4f64(av.r, bv.r, cr, 0, 0, 0);
    cr = __builtin_amdgcn_mfma_f64_16x16x4f64(-av.i, bv.i, cr, 0, 0, 0);
    ci = __builtin_amdgcn_mfma_f64_16x16x4f64(av.r, bv.i, ci, 0, 0, 0);
    ci = __builtin_amdgcn_mfma_f64_16x16x4f64(av.i, bv.r, ci, 0, 0, 0);
  }
}

// Register staging of a tile: 4 elements per thread (q = tid + 256 u), zero outside the matrix.
struct TileRegs {
  double2 v[4];
};
__device__ inline void tile_fetch(TileRegs &t, const double *__restrict__ src, int n, int r0,
                                  int c0, int tid) {
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int q = tid + 256 * u, rr = q / TB, cc = q % TB;
    t.v[u] = double2{0.0, 0.0};
    if (r0 + rr < n && c0 + cc < n)
      t.v[u] = *reinterpret_cast<const double2 *>(src + 2 * ((int64_t)(r0 + rr) * n + c0 + cc));
  }
}
__device__ inline void tile_fetch_packed(TileRegs &t, const double *__restrict__ src, int tid) {
#pragma unroll
  for (int u = 0; u < 4; u++)
    t.v[u] = *reinterpret_cast<const double2 *>(src + 2 * (tid + 256 * u));
}
__device__ inline void tile_put(Cz (*dst)[TB + 1], const TileRegs &t, int tid) {
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int q = tid + 256 * u;
    dst[q / TB][q % TB] = Cz{t.v[u].x, t.v[u].y};
  }
}
// Quadrant results back into an LDS tile (dst = acc) or into global memory (dst = base - acc).
__device__ inline void quad_to_lds(Cz (*dst)[TB + 1], const double4_t &cr, const double4_t &ci,
                                   int qr, int qc, int lane) {
#pragma unroll
  for (int e4 = 0; e4 < 4; e4++) dst[16 * qr + 4 * e4 + (lane >> 4)][16 * qc + (lane & 15)] =
      Cz{cr[e4], ci[e4]};
}
__device__ inline void quad_sub_store(double *__restrict__ dst, int n, int r0, int c0,
                                      const Cz (*base)[TB + 1], const double4_t &cr,
                                      const double4_t &ci, int qr, int qc, int lane) {
#pragma unroll
  for (int e4 = 0; e4 < 4; e4++) {
    const int rr = 16 * qr + 4 * e4 + (lane >> 4), cc = 16 * qc + (lane & 15);
    if (r0 + rr < n && c0 + cc < n) {
      double2 v;
      v.x = base[rr][cc].r - cr[e4];
      v.y = base[rr][cc].i - ci[e4];
      *reinterpret_cast<double2 *>(dst + 2 * ((int64_t)(r0 + rr) * n + c0 + cc)) = v;
    }
  }
}
__device__ inline void quad_store(double *__restrict__ dst, int n, int r0, int c0,
                                  const double4_t &cr, const double4_t &ci, int qr, int qc,
                                  int lane) {
#pragma unroll
  for (int e4 = 0; e4 < 4; e4++) {
    const int rr = 16 * qr + 4 * e4 + (lane >> 4), cc = 16 * qc + (lane & 15);
    if (r0 + rr < n && c0 + cc < n) {
      double2 v;
      v.x = cr[e4];
      v.y = ci[e4];
      *reinterpret_cast<double2 *>(dst + 2 * ((int64_t)(r0 + rr) * n + c0 + cc)) = v;
    }
  }
}

// One workgroup per tile of step J. E = L_JJ^-1 is lower triangular with zeros above the
// diagonal, so the triangular products are plain tile products.
//   type 1 (J < K <= I):  A_IK -= L_IJ L_KJ^H with L_XJ = A_XJ E^H;
//   type 2 (K <= J < I):  R_IK -= L_IJ W_JK with W_JK = E R_JK (K < J; the workgroup with
//                         I = J + 1 also stores it) or W_JJ = E.
// All tiles a workgroup needs are fetched at once, so the kernel pays one memory latency.
__global__ void __launch_bounds__(256) k_elim_update(double *__restrict__ a,
                                                     double *__restrict__ r,
                                                     double *__restrict__ einv,
                                                     double *__restrict__ w, int n, int jblk,
                                                     int32_t *__restrict__ status) {
  __shared__ Cz e[TB][TB + 1], t1[TB][TB + 1], t2[TB][TB + 1], t3[TB][TB + 1];
  __shared__ Cz colb[2][TB], rowb[2][TB];
  if (*status != ST_OK) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int qr = wave >> 1, qc = wave & 1;
  const int nblk = (n + TB - 1) / TB;
  const int mrem = nblk - jblk - 1;
  const int n_type1 = mrem * (mrem + 1) / 2;
  int b = blockIdx.x;
  int I, K;
  bool type1;
  if (b < n_type1) {  // (ii, kk), 0 <= kk <= ii < mrem
    int ii = 0;
    while (b > ii) {
      b -= ii + 1;
      ii++;
    }
    I = jblk + 1 + ii;
    K = jblk + 1 + b;
    type1 = true;
  } else {
    b -= n_type1;
    I = jblk + 1 + b / (jblk + 1);
    K = b % (jblk + 1);
    type1 = false;
  }
  const int j0 = jblk * TB;
  TileRegs te, ta, tb, tc;
  tile_fetch_packed(te, einv + 2 * (int64_t)jblk * TB * TB, tid);
  tile_fetch(ta, a, n, I * TB, j0, tid);
  const bool need_b = type1 ? (K != I) : (K < jblk);
  if (need_b) {
    if (type1)
      tile_fetch(tb, a, n, K * TB, j0, tid);
    else
      tile_fetch(tb, r, n, j0, K * TB, tid);
  }
  if (type1)
    tile_fetch(tc, a, n, I * TB, K * TB, tid);
  else
    tile_fetch(tc, r, n, I * TB, K * TB, tid);
  tile_put(e, te, tid);
  tile_put(t1, ta, tid);
  if (need_b) tile_put(t2, tb, tid);
  tile_put(t3, tc, tid);
  __syncthreads();
  double4_t cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
  double4_t dr = {0.0, 0.0, 0.0, 0.0}, di = {0.0, 0.0, 0.0, 0.0};
  // L_IJ = A_IJ E^H (and L_KJ = A_KJ E^H, or W_JK = E R_JK).
  cgemm32(t1, 0, e, 1, qr, qc, lane, cr, ci);
  if (need_b) {
    if (type1)
      cgemm32(t2, 0, e, 1, qr, qc, lane, dr, di);
    else
      cgemm32(e, 0, t2, 0, qr, qc, lane, dr, di);
  }
  __syncthreads();
  quad_to_lds(t1, cr, ci, qr, qc, lane);
  if (need_b) quad_to_lds(t2, dr, di, qr, qc, lane);
  if (!type1 && need_b && I == jblk + 1) quad_store(w, n, j0, K * TB, dr, di, qr, qc, lane);
  __syncthreads();
  double4_t xr = {0.0, 0.0, 0.0, 0.0}, xi = {0.0, 0.0, 0.0, 0.0};
  if (type1) {
    cgemm32(t1, 0, need_b ? t2 : t1, 1, qr, qc, lane, xr, xi);  // L_IJ L_KJ^H
    quad_sub_store(a, n, I * TB, K * TB, t3, xr, xi, qr, qc, lane);
  } else {
    cgemm32(t1, 0, need_b ? t2 : e, 0, qr, qc, lane, xr, xi);  // L_IJ W_JK
    quad_sub_store(r, n, I * TB, K * TB, t3, xr, xi, qr, qc, lane);
  }
  // The next diagonal tile (J + 1, J + 1) is final once this workgroup has updated it: factor
  // it here (the next block column's panel) instead of in a launch of its own. Its elements
  // outside the matrix are the identity.
  if (type1 && I == jblk + 1 && K == jblk + 1) {
    __syncthreads();
#pragma unroll
    for (int e4 = 0; e4 < 4; e4++) {
      const int rr = 16 * qr + 4 * e4 + (lane >> 4), cc = 16 * qc + (lane & 15);
      const bool in = I * TB + rr < n && K * TB + cc < n;
      t2[rr][cc] = in ? Cz{t3[rr][cc].r - xr[e4], t3[rr][cc].i - xi[e4]}
                      : Cz{rr == cc ? 1.0 : 0.0, 0.0};
    }
    __syncthreads();
    panel_core(t2, n, jblk + 1, einv, w, status, colb, rowb, t1);
  }
}

// Last block row of W: W_JK = L_JJ^-1 R_JK for K < J = nblk - 1 (no update kernel follows it).
__global__ void __launch_bounds__(256) k_elim_lastrow(const double *__restrict__ r,
                                                      const double *__restrict__ einv,
                                                      double *__restrict__ w, int n,
                                                      const int32_t *__restrict__ status) {
  __shared__ Cz e[TB][TB + 1], y[TB][TB + 1];
  if (*status != ST_OK) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int qr = wave >> 1, qc = wave & 1;
  const int jblk = (n + TB - 1) / TB - 1, j0 = jblk * TB, K = blockIdx.x;
  TileRegs te, ty;
  tile_fetch_packed(te, einv + 2 * (int64_t)jblk * TB * TB, tid);
  tile_fetch(ty, r, n, j0, K * TB, tid);
  tile_put(e, te, tid);
  tile_put(y, ty, tid);
  __syncthreads();
  double4_t cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
  cgemm32(e, 0, y, 0, qr, qc, lane, cr, ci);
  quad_store(w, n, j0, K * TB, cr, ci, qr, qc, lane);
}

// Runs the elimination on (a, r) prepared by launch_t_from_xx or launch_elim_init; W -> w.
void launch_inverse_cholesky(hipStream_t st, double *a, double *r, double *einv, double *w, int nb,
                             int32_t *status) {
  const int nblk = (nb + TB - 1) / TB;
  hipLaunchKernelGGL(k_elim_panel, dim3(1), dim3(256), 0, st, a, nb, einv, w, status);
  for (int j = 0; j + 1 < nblk; j++) {
    const int mrem = nblk - j - 1;
    const int tiles = mrem * (mrem + 1) / 2 + mrem * (j + 1);
    hipLaunchKernelGGL(k_elim_update, dim3(tiles), dim3(256), 0, st, a, r, einv, w, nb, j,
                       status);
  }
  if (nblk > 1)
    hipLaunchKernelGGL(k_elim_lastrow, dim3(nblk - 1), dim3(256), 0, st, r, einv, w, nb, status);
}

// ---------------------------------------------------------------------------------------------
// G = W^H W per 32 x 32 tile: G_IK = sum_{J >= max(I, K)} W_JI^H W_JK on f64 MFMA, with the next
// pair of W tiles fetched into registers while the current pair is multiplied. With
// residual != nullptr the tile also yields 2 sum |G_new - G_old|^2 into partial[tile] and G_new
// replaces g in place; otherwise G_new is written to g. If status shows a failed factorisation,
// G = W = I (the reference's identity fallback, quintic.rs:1874-1883). Each tile reads and
// writes only its own G entries, so the in-place update has no race. G_ab and G_ba come from
// the same products in the same order, so G is exactly Hermitian.
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_gram_tiles(double *__restrict__ w, int n,
                                                    double *__restrict__ g,
                                                    const int32_t *__restrict__ status,
                                                    double *__restrict__ partial,
                                                    unsigned *__restrict__ counter,
                                                    double *__restrict__ residual) {
  __shared__ Cz wi[2][TB][TB + 1], wk[2][TB][TB + 1];
  __shared__ double red[256];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int qr = wave >> 1, qc = wave & 1;
  const int nblk = (n + TB - 1) / TB;
  const int I = blockIdx.x / nblk, K = blockIdx.x % nblk;
  const bool fail = status && *status != ST_OK;
  double4_t cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
  if (!fail) {
    const int jb = max(I, K);
    TileRegs ti, tk;
    tile_fetch(ti, w, n, jb * TB, I * TB, tid);
    tile_fetch(tk, w, n, jb * TB, K * TB, tid);
    for (int J = jb; J < nblk; J++) {
      const int buf = (J - jb) & 1;
      tile_put(wi[buf], ti, tid);
      tile_put(wk[buf], tk, tid);
      __syncthreads();
      if (J + 1 < nblk) {
        tile_fetch(ti, w, n, (J + 1) * TB, I * TB, tid);
        tile_fetch(tk, w, n, (J + 1) * TB, K * TB, tid);
      }
      cgemm32(wi[buf], 1, wk[buf], 0, qr, qc, lane, cr, ci);  // W_JI^H W_JK
    }
  }
  double res = 0.0;
#pragma unroll
  for (int e4 = 0; e4 < 4; e4++) {
    const int gi = I * TB + 16 * qr + 4 * e4 + (lane >> 4);
    const int gk = K * TB + 16 * qc + (lane & 15);
    if (gi >= n || gk >= n) continue;
    double nr = cr[e4], ni = ci[e4];
    if (fail) {
      nr = gi == gk ? 1.0 : 0.0;
      ni = 0.0;
      w[2 * ((int64_t)gi * n + gk)] = nr;
      w[2 * ((int64_t)gi * n + gk) + 1] = 0.0;
    }
    double *pg = g + 2 * ((int64_t)gi * n + gk);
    if (partial) {
      const double dr = nr - pg[0], di = ni - pg[1];
      res += 2.0 * (dr * dr + di * di);
    }
    pg[0] = nr;
    pg[1] = ni;
  }
  if (partial) {
    red[tid] = res;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
      if (tid < off) red[tid] += red[tid + off];
      __syncthreads();
    }
    // With a counter, the last tile to finish sums the per-tile partials in tile order (the
    // counter wraps back to zero for the next launch); otherwise k_residual_final does.
    __shared__ bool last;
    if (tid == 0) {
      partial[blockIdx.x] = red[0];
      last = false;
      if (counter) {
        __threadfence();
        last = atomicInc(counter, gridDim.x - 1) == gridDim.x - 1;
      }
    }
    __syncthreads();
    if (last && tid == 0) {
      __threadfence();
      double acc = 0.0;
      for (unsigned i = 0; i < gridDim.x; i++)
        acc += __hip_atomic_load(partial + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      *residual = sqrt(acc);
    }
  }
}

__global__ void __launch_bounds__(64) k_residual_final(const double *__restrict__ partial, int cnt,
                                                       double *__restrict__ residual) {
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < cnt; i++) s += partial[i];
    *residual = sqrt(s);
  }
}

void launch_gram_residual(hipStream_t st, double *w, int nb, double *g, const int32_t *status,
                          double *partial, double *residual, unsigned *counter) {
  const int nblk = (nb + TB - 1) / TB;
  const bool fused = partial && residual && counter;
  hipLaunchKernelGGL(k_gram_tiles, dim3(nblk * nblk), dim3(256), 0, st, w, nb, g, status, partial,
                     fused ? counter : nullptr, fused ? residual : nullptr);
  if (partial && residual && !fused)
    hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(64), 0, st, partial, nblk * nblk, residual);
}

}  // namespace cy3q
