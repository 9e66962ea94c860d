// Split-precision (fp16 hi/lo) Hermitian rank-k update of the Galerkin assembly.
//
// The Laplacian's mass and stiffness matrices are sums over points of outer products of
// per-point rows (reference: build_mass_matrix / build_stiffness_matrix,
// route34/metric_laplacian.rs:424-524, on the rows of k_laplacian.hip). This is the FP32-MFMA
// form of that sum on gfx950's fp16 MFMA: every row value x, scaled by an exact power of two,
// is split as x = hi + lo with hi = fp16(x), lo = fp16(x - hi), and a product is
// hi hi' + hi lo' + lo hi' (about 22 bits), accumulated in FP32 on v_mfma_f32_32x32x16_f16.
// The FP32 partials are added into FP64 accumulators every HS_FLUSH stages, so no FP32 sum runs
// over more than 4096 points; the FP64 sums are deterministic (each workgroup owns its output
// elements, and the point range is split into fixed parts with one accumulator copy each).
//
// Operands: per buffer, fp16 planes hi and lo of layout [stage][column][16 points], so the MFMA
// fragment of a lane (one column, 8 consecutive points) is 16 contiguous bytes.
#include "cy3q_internal.h"

namespace cy3q {

static constexpr int HS_TB = 256;     // output tile edge (real columns)
static constexpr int HS_PTS = 16;     // points per stage, the MFMA k dimension
static constexpr int HS_WAVES = 16;   // 4 x 4 waves, each a 64 x 64 quadrant
static constexpr int HS_FLUSH = 256;  // stages per FP32 partial (4096 points)

// ---------------------------------------------------------------------------------------------
// Range: one power-of-two scale for all rows of a chunk, so that max |x| 2^s < 2^14. scale[0] =
// 2^s (applied before the split), scale[1] = 4^-s (restores the products). All-zero rows give
// scale 1.
// ---------------------------------------------------------------------------------------------
// Per-block maxima of |x| over two arrays (a[0..na), then b[0..nb)).
__global__ void __launch_bounds__(256) k_absmax_part(const double *__restrict__ a, int64_t na,
                                                     const double *__restrict__ b, int64_t nb,
                                                     double *__restrict__ part) {
  __shared__ double sh[256];
  double m = 0.0;
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < na + nb; i += stride)
    m = fmax(m, fabs(i < na ? a[i] : b[i - na]));
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(256) k_absmax_scale(const double *__restrict__ part, int nparts,
                                                      double *__restrict__ scale) {
  __shared__ double sh[256];
  double m = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 256) m = fmax(m, part[i]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    int e = 14;
    if (sh[0] > 0.0 && isfinite(sh[0])) frexp(sh[0], &e);  // max |x| < 2^e
    scale[0] = ldexp(1.0, 14 - e);
    scale[1] = ldexp(1.0, 2 * (e - 14));
  }
}

static constexpr int HS_PART_BLOCKS = 1024;

void launch_rows_scale_pair(hipStream_t st, const double *a, int64_t na, const double *b,
                            int64_t nb, double *part, double *scale) {
  hipLaunchKernelGGL(k_absmax_part, dim3(HS_PART_BLOCKS), dim3(256), 0, st, a, na, b, nb, part);
  hipLaunchKernelGGL(k_absmax_scale, dim3(1), dim3(256), 0, st, part, HS_PART_BLOCKS, scale);
}

int rows_scale_parts() { return HS_PART_BLOCKS; }

// The same scale from per-block maxima computed elsewhere (nparts of them).
void launch_rows_scale_from_parts(hipStream_t st, const double *part, int nparts, double *scale) {
  hipLaunchKernelGGL(k_absmax_scale, dim3(1), dim3(256), 0, st, part, nparts, scale);
}

// x [rows][mc] FP64 -> hi, lo [stages][mc][16] fp16 of fp32(x 2^s); zero past `rows`. One thread
// per (stage, column): 16 loads down the column (coalesced across threads), 32 bytes per plane.
__global__ void __launch_bounds__(256) k_split_rows(const double *__restrict__ x, int64_t rows,
                                                    int64_t stages, int64_t mc,
                                                    const double *__restrict__ scale,
                                                    half8_t *__restrict__ xh,
                                                    half8_t *__restrict__ xl) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= stages * mc) return;
  const int64_t s = q / mc, c = q % mc;
  const double sc = scale[0];
  half8_t h[2], l[2];
#pragma unroll
  for (int j = 0; j < HS_PTS; j++) {
    const int64_t p = s * HS_PTS + j;
    const float v = p < rows ? (float)(x[p * mc + c] * sc) : 0.f;
    const _Float16 hv = (_Float16)v;
    h[j >> 3][j & 7] = hv;
    l[j >> 3][j & 7] = (_Float16)(v - (float)hv);
  }
  xh[q * 2] = h[0];
  xh[q * 2 + 1] = h[1];
  xl[q * 2] = l[0];
  xl[q * 2 + 1] = l[1];
}

void launch_split_rows(hipStream_t st, const double *x, int64_t rows, int64_t stages, int64_t mc,
                       const double *scale, half8_t *xh, half8_t *xl) {
  const int64_t tot = stages * mc;
  if (tot <= 0) return;
  hipLaunchKernelGGL(k_split_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, x, rows,
                     stages, mc, scale, xh, xl);
}

// ---------------------------------------------------------------------------------------------
// C_b,k[lower 256 x 256 tiles] += sum over the stages of part k of X^T X, b = blockIdx.z (buffer),
// k = blockIdx.y (point part), tile = blockIdx.x. 16 waves; wave (qr, qc) owns the 64 x 64
// quadrant of the tile as 2 x 2 MFMA tiles of 32 x 32 (the quadrants above the diagonal of a
// diagonal tile are skipped: only the lower triangle is read later). Each stage's two 256-column
// panels (hi and lo, 32 KB) are loaded into registers one stage ahead and written to an LDS
// double buffer; a diagonal tile loads one panel.
// 32x32 MFMA output: element e of lane l is (row 8 (e / 4) + 4 (l / 32) + e % 4, column l % 32).
// ---------------------------------------------------------------------------------------------
struct HerkSplitArgs {
  const half8_t *xh[4];
  const half8_t *xl[4];
  double *c[4];
};

__global__ void __launch_bounds__(64 * HS_WAVES) k_herk_split16(HerkSplitArgs args, int64_t stages,
                                                               int64_t mc, int nparts,
                                                               const double *__restrict__ scale) {
  // [buffer][panel I/J][plane hi/lo][column * 2 + half]
  __shared__ half8_t lds[2][2][2][HS_TB * 2];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int qr = wave >> 2, qc = wave & 3;
  const int b = blockIdx.z, kp = blockIdx.y;
  const int t = blockIdx.x;
  int I = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= t) I++;
  while (I * (I + 1) / 2 > t) I--;
  const int J = t - I * (I + 1) / 2;
  const bool diag = I == J;
  const bool active = !(diag && qr < qc);
  const int64_t s0 = stages * kp / nparts, s1 = stages * (kp + 1) / nparts;
  const half8_t *__restrict__ xh = args.xh[b];
  const half8_t *__restrict__ xl = args.xl[b];

  float16_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int e = 0; e < 16; e++) acc[a][c][e] = 0.f;

  // Thread tid moves the 16-byte units q = tid and tid + 1024 of the stage: panel q >> 10,
  // plane (q >> 9) & 1, unit q & 511 (column unit >> 1, half unit & 1).
  half8_t reg[2];
  auto fetch = [&](int64_t s) {
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int q = tid + 1024 * u, panel = q >> 10, plane = (q >> 9) & 1, unit = q & 511;
      if (diag && panel == 1) continue;
      const int64_t col = (int64_t)(panel ? J : I) * HS_TB + (unit >> 1);
      reg[u] = (plane ? xl : xh)[(s * mc + col) * 2 + (unit & 1)];
    }
  };
  auto put = [&](int buf) {
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int q = tid + 1024 * u, panel = q >> 10, plane = (q >> 9) & 1, unit = q & 511;
      if (diag && panel == 1) continue;
      lds[buf][panel][plane][unit] = reg[u];
    }
  };
  const double restore = scale[1];
  auto flush = [&]() {
    if (!active) return;
    double *__restrict__ cm = args.c[b] + (int64_t)kp * mc * mc;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int c = 0; c < 2; c++)
#pragma unroll
        for (int e = 0; e < 16; e++) {
          const int64_t gi = (int64_t)I * HS_TB + qr * 64 + a * 32 + 8 * (e / 4) + 4 * (lane >> 5) +
                             (e % 4);
          const int64_t gj = (int64_t)J * HS_TB + qc * 64 + c * 32 + (lane & 31);
          cm[gi * mc + gj] += (double)acc[a][c][e] * restore;
          acc[a][c][e] = 0.f;
        }
  };

  if (s1 > s0) {
    fetch(s0);
    put(0);
  }
  __syncthreads();
  const int jp = diag ? 0 : 1;
  const int fl = lane & 31, fh = lane >> 5;
  for (int64_t s = s0; s < s1; s++) {
    const int buf = (int)((s - s0) & 1);
    if (s + 1 < s1) fetch(s + 1);
    if (active) {
      half8_t ah[2], al[2], bh[2], bl[2];
#pragma unroll
      for (int a = 0; a < 2; a++) {
        const int ca = (qr * 64 + a * 32 + fl) * 2 + fh;
        ah[a] = lds[buf][0][0][ca];
        al[a] = lds[buf][0][1][ca];
        const int cb = (qc * 64 + a * 32 + fl) * 2 + fh;
        bh[a] = lds[buf][jp][0][cb];
        bl[a] = lds[buf][jp][1][cb];
      }
#pragma unroll
      for (int pr = 0; pr < 3; pr++)
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
          for (int c = 0; c < 2; c++)
            acc[a][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(pr == 2 ? al[a] : ah[a],
                                                               pr == 1 ? bl[c] : bh[c], acc[a][c],
                                                               0, 0, 0);
    }
    if ((s - s0 + 1) % HS_FLUSH == 0 || s + 1 == s1) flush();
    if (s + 1 < s1) put(buf ^ 1);
    __syncthreads();
  }
}

void launch_herk_split16(hipStream_t st, const half8_t *const *xh, const half8_t *const *xl,
                         double *const *c, int nbuf, int64_t stages, int64_t mc, int nparts,
                         const double *scale) {
  HerkSplitArgs a{};
  for (int q = 0; q < nbuf; q++) {
    a.xh[q] = xh[q];
    a.xl[q] = xl[q];
    a.c[q] = c[q];
  }
  const int nt = (int)(mc / HS_TB);
  const int n_tiles = nt * (nt + 1) / 2;
  hipLaunchKernelGGL(k_herk_split16, dim3((unsigned)n_tiles, (unsigned)nparts, (unsigned)nbuf),
                     dim3(64 * HS_WAVES), 0, st, a, stages, mc, nparts, scale);
}

// The accumulator copies of the point parts summed into copy 0, in part order.
__global__ void k_sum_parts(double *__restrict__ c, int64_t len, int nparts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  double s = c[i];
  for (int k = 1; k < nparts; k++) s += c[(int64_t)k * len + i];
  c[i] = s;
}

void launch_sum_parts(hipStream_t st, double *c, int64_t len, int nparts) {
  if (nparts <= 1 || len <= 0) return;
  hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, c, len,
                     nparts);
}

int herk_split_tile() { return HS_TB; }

// ---------------------------------------------------------------------------------------------
// Split metric contraction of the intrinsic rows: Y = X B, X [rows][m] the FP64 section (or
// derivative) rows, B [m][m] the real-ified W^T (block upper triangular: column c needs only
// k < 2 (c / 2) + 2, so column block J stops at k = min(256 (J + 1), m)). Y_r. is then W x_r,
// whose inner products give the point's K, dK and ddK (the metric). The FP64 form is
// k_trigemm_f64. Here:
//   * row r is scaled by 2^(14 - ex[r]) (max |x| < 2^14, read from launch_row_exponent), split
//     into fp16 hi/lo as the stage is staged into LDS;
//   * B comes as fp16 hi/lo planes [m / 16][m][16] of B 2^sB (launch_split_rows with one
//     power-of-two scale, bscale[0] = 2^sB);
//   * a product is hi hi' + hi lo' + lo hi' on v_mfma_f32_32x32x16_f16 with FP32 accumulation
//     over the k range, and Y = acc 2^(ex[r] - 14) / bscale[0] is stored in FP64.
// Workgroup: a 256 x 256 output tile, 16 waves as in k_herk_split16 (wave (qr, qc) the 64 x 64
// quadrant as 2 x 2 tiles of 32 x 32); threads 0..511 stage the A panel (row tid / 2, 8 k
// values), threads 512..1023 the B panel (two 16-byte units each), one stage ahead in
// registers. Workgroups are dealt to XCDs so that the column blocks of one row tile share an
// L2 (X is read from HBM about once).
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(64 * HS_WAVES) k_rowgemm_split16(
    const double *__restrict__ x, int64_t rows, int64_t m, const int *__restrict__ ex,
    const half8_t *__restrict__ bh, const half8_t *__restrict__ bl,
    const double *__restrict__ bscale, double *__restrict__ y, int64_t ldy, int64_t n_rt,
    int n_cb) {
  __shared__ half8_t lds[2][2][2][HS_TB * 2];  // [buffer][panel A/B][plane hi/lo][unit]
  const int64_t g = blockIdx.x;
  const int64_t xcd = g & 7, qq = g >> 3;
  const int cb = (int)(qq % n_cb);
  const int64_t rt = (qq / n_cb) * 8 + xcd;
  if (rt >= n_rt) return;
  const int64_t r0 = rt * HS_TB;
  const int64_t c0 = (int64_t)cb * HS_TB;
  const int64_t kend = c0 + HS_TB < m ? c0 + HS_TB : m;
  const int64_t nst = kend / HS_PTS;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int qr = wave >> 2, qc = wave & 3;

  float16_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int e = 0; e < 16; e++) acc[a][c][e] = 0.f;

  // Every thread stages a quarter stage-row of A (row tid / 4, k values 4 (tid % 4) .. + 3,
  // loaded as FP64 and split in put, after the stage's MFMAs, so the loads' latency is hidden)
  // and one 16-byte unit of B (plane tid / 512, unit tid % 512 = column * 2 + k half).
  const int ar = tid >> 2, aq = tid & 3;
  const int64_t arow = r0 + ar;
  const bool a_in = arow < rows;
  const double a_sc = a_in ? ldexp(1.0, 14 - ex[arow]) : 0.0;
  const double *__restrict__ xrow = x + (a_in ? arow : 0) * m + 4 * aq;
  const int bpl = tid >> 9, bunit = tid & 511;
  const int64_t bcol = c0 + (bunit >> 1);
  const half8_t *__restrict__ bsrc = (bpl ? bl : bh) + (bcol < m ? bcol : 0) * 2 + (bunit & 1);
  double2 ra[2];
  half8_t rb;
  auto fetch = [&](int64_t s) {
    ra[0] = ra[1] = double2{0.0, 0.0};
    if (a_in) {
      ra[0] = *reinterpret_cast<const double2 *>(xrow + s * HS_PTS);
      ra[1] = *reinterpret_cast<const double2 *>(xrow + s * HS_PTS + 2);
    }
    rb = half8_t{};
    if (bcol < m) rb = bsrc[s * m * 2];
  };
  auto put = [&](int buf) {
    typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
    half4_t rh, rl;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const float f = (float)((j & 1 ? ra[j >> 1].y : ra[j >> 1].x) * a_sc);
      const _Float16 h = (_Float16)f;
      rh[j] = h;
      rl[j] = (_Float16)(f - (float)h);
    }
    half4_t *uh = reinterpret_cast<half4_t *>(&lds[buf][0][0][ar * 2 + (aq >> 1)]);
    half4_t *ul = reinterpret_cast<half4_t *>(&lds[buf][0][1][ar * 2 + (aq >> 1)]);
    uh[aq & 1] = rh;
    ul[aq & 1] = rl;
    lds[buf][1][bpl][bunit] = rb;
  };

  if (nst > 0) {
    fetch(0);
    put(0);
  }
  __syncthreads();
  const int fl = lane & 31, fh = lane >> 5;
  for (int64_t s = 0; s < nst; s++) {
    const int buf = (int)(s & 1);
    if (s + 1 < nst) fetch(s + 1);
    half8_t ah[2], al[2], bhv[2], blv[2];
#pragma unroll
    for (int a = 0; a < 2; a++) {
      const int ca = (qr * 64 + a * 32 + fl) * 2 + fh;
      ah[a] = lds[buf][0][0][ca];
      al[a] = lds[buf][0][1][ca];
      const int cb2 = (qc * 64 + a * 32 + fl) * 2 + fh;
      bhv[a] = lds[buf][1][0][cb2];
      blv[a] = lds[buf][1][1][cb2];
    }
#pragma unroll
    for (int pr = 0; pr < 3; pr++)
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++)
          acc[a][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(pr == 2 ? al[a] : ah[a],
                                                             pr == 1 ? blv[c] : bhv[c], acc[a][c],
                                                             0, 0, 0);
    if (s + 1 < nst) put(buf ^ 1);
    __syncthreads();
  }

  const double bre = 1.0 / bscale[0];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int e4 = 0; e4 < 4; e4++)
#pragma unroll
      for (int e1 = 0; e1 < 4; e1++) {
        const int64_t gr = r0 + qr * 64 + a * 32 + 8 * e4 + 4 * fh + e1;
        if (gr >= rows) continue;
        const double rs = ldexp(bre, ex[gr] - 14);
#pragma unroll
        for (int c = 0; c < 2; c++) {
          const int64_t gc = c0 + qc * 64 + c * 32 + fl;
          if (gc < m) y[gr * ldy + gc] = (double)acc[a][c][4 * e4 + e1] * rs;
        }
      }
}

void launch_rowgemm_split16(hipStream_t st, const double *x, int64_t rows, int64_t m,
                            const int *ex, const half8_t *bh, const half8_t *bl,
                            const double *bscale, double *y, int64_t ldy) {
  if (rows <= 0) return;
  const int64_t n_rt = (rows + HS_TB - 1) / HS_TB;
  const int64_t n_rt8 = (n_rt + 7) / 8 * 8;
  const int n_cb = (int)((m + HS_TB - 1) / HS_TB);
  hipLaunchKernelGGL(k_rowgemm_split16, dim3((unsigned)(n_rt8 * n_cb)), dim3(64 * HS_WAVES), 0,
                     st, x, rows, m, ex, bh, bl, bscale, y, ldy, n_rt, n_cb);
}

}  // namespace cy3q
