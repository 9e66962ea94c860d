// Split-precision (fp16 hi/lo) MFMA contractions of the Donaldson T-operator.
//
// The two dense contractions of one T-iteration (reference: quintic.rs:1745-1789) run on gfx950's
// fp16 MFMA with FP32 accumulation. Each FP32 operand x is split as x = hi + lo with
// hi = fp16(x) and lo = fp16(x - hi); a product is hi*hi' + hi*lo' + lo*hi', about 22 bits per
// product. Overflow and subnormal loss are avoided by exact power-of-two scalings:
//   * every section row is stored scaled by 2^-e_p (max |x| in [0.5, 1)), once per workspace;
//   * W (the D-pass operand) is scaled globally by 2^-f;
//   * z_p = sqrt(c_p) x_p (the T-pass operand) is scaled globally by 2^g.
// The host-side bookkeeping undoes each scaling exactly. D_p and the chunk sums are accumulated in
// FP64. T is summed across point chunks in FP64, in a fixed order.
//
// D pass  (k_dpass_split16): v_mfma_f32_16x16x32_f16, 128 points x all M columns per workgroup,
//                            so D_p = sum_j Y_pj^2 completes inside the workgroup.
// T pass  (k_syrk_split16):  v_mfma_f32_32x32x16_f16 on the lower 32x32 sub-tiles of Z^T Z,
//                            one workgroup per (point chunk, part).
#include "cy3q_internal.h"

namespace cy3q {

__device__ inline void split_f32(float x, _Float16 &h, _Float16 &l) {
  h = (_Float16)x;
  l = (_Float16)(x - (float)h);
}

// ---------------------------------------------------------------------------------------------
// B fragments of the D pass. B is the real-ified transpose of the lower-triangular complex W
// (see k_build_bmat): B[2a][2b] = Re W_ba, B[2a+1][2b] = -Im W_ba, B[2a][2b+1] = Im W_ba,
// B[2a+1][2b+1] = Re W_ba for a <= b. Fragment (ks, cg) holds B[32ks + 8(l/16) + j][16cg + l%16],
// j = 0..7, for lane l: the B operand layout of v_mfma_f32_16x16x32_f16.
// Layout: bfrag[((ks * n_cg + cg) * 2 + {0: hi, 1: lo}) * 64 + lane]. bscale[0] = 4^f, the factor
// that restores D from the scaled contraction.
// ---------------------------------------------------------------------------------------------
__device__ inline double bmat_entry(const double *__restrict__ w, int nb, int64_t i, int64_t j) {
  int64_t a = i / 2, b = j / 2;
  if (a >= nb || b >= nb || a > b) return 0.0;
  double vr = w[2 * (b * nb + a)], vi = w[2 * (b * nb + a) + 1];
  int ri = (int)(i & 1), rj = (int)(j & 1);
  if (ri == 0 && rj == 0) return vr;
  if (ri == 1 && rj == 0) return -vi;
  if (ri == 0 && rj == 1) return vi;
  return vr;
}

__global__ void __launch_bounds__(1024) k_build_bfrag16(const double *__restrict__ w, int nb,
                                                        int64_t m, half8_t *__restrict__ bfrag,
                                                        double *__restrict__ bscale) {
  __shared__ double sh[1024];
  double mx = 0.0;
  for (int64_t e = threadIdx.x; e < 2 * (int64_t)nb * nb; e += 1024) mx = fmax(mx, fabs(w[e]));
  sh[threadIdx.x] = mx;
  __syncthreads();
  for (int off = 512; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + off]);
    __syncthreads();
  }
  int f = 0;
  if (sh[0] > 0.0) frexp(sh[0], &f);
  const double sc = ldexp(1.0, -f);
  if (threadIdx.x == 0) bscale[0] = ldexp(1.0, 2 * f);
  const int n_ks = (int)(m / 32), n_cg = (int)(m / 16);
  const int64_t total = (int64_t)n_ks * n_cg * 64;
  for (int64_t e = threadIdx.x; e < total; e += 1024) {
    const int lane = (int)(e % 64);
    const int64_t frag = e / 64;
    const int cg = (int)(frag % n_cg), ks = (int)(frag / n_cg);
    half8_t hi, lo;
    for (int j = 0; j < 8; j++) {
      int64_t kr = 32 * ks + 8 * (lane / 16) + j;
      int64_t col = 16 * cg + (lane % 16);
      float v = (float)(bmat_entry(w, nb, kr, col) * sc);
      _Float16 h, l;
      split_f32(v, h, l);
      hi[j] = h;
      lo[j] = l;
    }
    bfrag[(frag * 2 + 0) * 64 + lane] = hi;
    bfrag[(frag * 2 + 1) * 64 + lane] = lo;
  }
}

void launch_build_bfrag16(hipStream_t st, const double *v_cplx, int nb, int64_t m_cols,
                          half8_t *bfrag, double *b_scale) {
  hipLaunchKernelGGL(k_build_bfrag16, dim3(1), dim3(1024), 0, st, v_cplx, nb, m_cols, bfrag,
                     b_scale);
}

// ---------------------------------------------------------------------------------------------
// D pass. Workgroup: 128 points, 8 waves. Wave w owns the column groups cglist[w][0..max_cg)
// (balanced on the host by their triangular k-step counts). Per k-step of 32 columns, the point
// tile is staged in LDS already split into hi/lo halves, [128][40] each (row pad keeps the
// 16-byte fragment reads bank-conflict free). Output: dsum[p] = sum_j Y_pj^2 (FP64), for the
// row-scaled x and the 2^-f scaled W.
// ---------------------------------------------------------------------------------------------
static constexpr int DP_PTS = 128;
static constexpr int DP_LD = 40;
static constexpr int DP_MAXCG = 5;

__global__ void __launch_bounds__(512) k_dpass_split16(const float *__restrict__ x32, int64_t n,
                                                       int64_t m, const half8_t *__restrict__ bfrag,
                                                       const int16_t *__restrict__ cglist,
                                                       double *__restrict__ dsum) {
  __shared__ _Float16 sh_hi[DP_PTS * DP_LD];
  __shared__ _Float16 sh_lo[DP_PTS * DP_LD];
  __shared__ double red[8][DP_PTS];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int64_t p0 = (int64_t)blockIdx.x * DP_PTS;
  const int n_cg = (int)(m / 16), n_ks = (int)(m / 32);
  int cgs[DP_MAXCG];
  int ncg = 0, ks_end = 0;
#pragma unroll
  for (int c = 0; c < DP_MAXCG; c++) {
    cgs[c] = cglist[wave * DP_MAXCG + c];
    if (cgs[c] >= 0) {
      ncg = c + 1;
      ks_end = max(ks_end, cgs[c] / 2 + 1);
    }
  }
  float4_t acc[8][DP_MAXCG];
#pragma unroll
  for (int r = 0; r < 8; r++)
#pragma unroll
    for (int c = 0; c < DP_MAXCG; c++) acc[r][c] = (float4_t){0.f, 0.f, 0.f, 0.f};

  // Every wave takes part in staging; the k-loop bound is the workgroup maximum.
  const int ks_all = n_ks;
  for (int ks = 0; ks < ks_all; ks++) {
    // Stage X[p0 .. p0+128)[32ks .. 32ks+32): 4096 floats, 8 per thread (two float4).
    for (int e = tid; e < DP_PTS * 8; e += 512) {
      int r = e >> 3, q = (e & 7) * 4;
      int64_t p = p0 + r;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (p < n) v = *reinterpret_cast<const float4 *>(x32 + p * m + 32 * ks + q);
      _Float16 h0, l0, h1, l1, h2, l2, h3, l3;
      split_f32(v.x, h0, l0);
      split_f32(v.y, h1, l1);
      split_f32(v.z, h2, l2);
      split_f32(v.w, h3, l3);
      _Float16 *dh = sh_hi + r * DP_LD + q, *dl = sh_lo + r * DP_LD + q;
      dh[0] = h0;
      dh[1] = h1;
      dh[2] = h2;
      dh[3] = h3;
      dl[0] = l0;
      dl[1] = l1;
      dl[2] = l2;
      dl[3] = l3;
    }
    __syncthreads();
    if (ks < ks_end) {
      half8_t bh[DP_MAXCG], bl[DP_MAXCG];
#pragma unroll
      for (int c = 0; c < DP_MAXCG; c++) {
        if (c < ncg && ks <= cgs[c] / 2) {
          int64_t frag = (int64_t)ks * n_cg + cgs[c];
          bh[c] = bfrag[(frag * 2 + 0) * 64 + lane];
          bl[c] = bfrag[(frag * 2 + 1) * 64 + lane];
        }
      }
#pragma unroll
      for (int r = 0; r < 8; r++) {
        const int row = r * 16 + (lane & 15);
        const int kof = 8 * (lane >> 4);
        half8_t ah = *reinterpret_cast<const half8_t *>(sh_hi + row * DP_LD + kof);
        half8_t al = *reinterpret_cast<const half8_t *>(sh_lo + row * DP_LD + kof);
#pragma unroll
        for (int c = 0; c < DP_MAXCG; c++) {
          if (c < ncg && ks <= cgs[c] / 2) {
            acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh[c], acc[r][c], 0, 0, 0);
            acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl[c], acc[r][c], 0, 0, 0);
            acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh[c], acc[r][c], 0, 0, 0);
          }
        }
      }
    }
    __syncthreads();
  }
  // Row sums of squares: element e of lane l is point 16r + 4(l/16) + e, column 16cg + l%16.
#pragma unroll
  for (int r = 0; r < 8; r++) {
#pragma unroll
    for (int e = 0; e < 4; e++) {
      double s = 0.0;
#pragma unroll
      for (int c = 0; c < DP_MAXCG; c++) {
        if (c < ncg) {
          double y = (double)acc[r][c][e];
          s += y * y;
        }
      }
      for (int off = 1; off < 16; off <<= 1) s += __shfl_xor(s, off);
      if ((lane & 15) == 0) red[wave][r * 16 + 4 * (lane >> 4) + e] = s;
    }
  }
  __syncthreads();
  if (tid < DP_PTS) {
    double s = 0.0;
    for (int w = 0; w < 8; w++) s += red[w][tid];
    int64_t p = p0 + tid;
    if (p < n) dsum[p] = s;
  }
}

void launch_dpass_split16(hipStream_t st, const float *x32, int64_t n, int64_t m_cols,
                          const half8_t *bfrag, const int16_t *cglist, double *dsum) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_dpass_split16, dim3((unsigned)((n + DP_PTS - 1) / DP_PTS)), dim3(512), 0, st,
                     x32, n, m_cols, bfrag, cglist, dsum);
}

// ---------------------------------------------------------------------------------------------
// Per-point coefficients of the split path. D_p = dsum_p * bscale / rs2_p restores the unscaled
// s_p^H G s_p. The weight on the scaled row is c'_p = w_p / (rs2_p * max(D_p, 1e-30)), for valid
// w (finite, > 0), else 0: the same T as the reference's w_p / max(K_p, 1e-30) on the unscaled
// row. u_p = sqrt(c'_p) (FP32). The largest u (as float bits, an order-independent max) sets the
// global 2^g of the T pass.
// ---------------------------------------------------------------------------------------------
__global__ void k_coeffs_split16(const double *__restrict__ w, const double *__restrict__ dsum,
                                 const float *__restrict__ rs2, const double *__restrict__ bscale,
                                 int64_t n, float *__restrict__ u,
                                 unsigned int *__restrict__ umax_bits) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float uv = 0.f;
  if (p < n) {
    double wp = w[p];
    if (isfinite(wp) && wp > 0.0) {
      double r = (double)rs2[p];
      double d = dsum[p] * bscale[0] / r;
      double c = wp / (r * fmax(d, 1e-30));
      uv = (float)sqrt(c);
    }
    u[p] = uv;
  }
  // wave max then one atomic per wave
  float mv = uv;
  for (int off = 32; off > 0; off >>= 1) mv = fmaxf(mv, __shfl_xor(mv, off));
  if ((threadIdx.x & 63) == 0 && mv > 0.f) atomicMax(umax_bits, __float_as_uint(mv));
}

// zscale[0] = 2^g (float, applied to z), tscale[0] = 4^-g (double, applied to T); resets the max.
__global__ void k_finalize_scale(unsigned int *__restrict__ umax_bits, float *__restrict__ zscale,
                                 double *__restrict__ tscale) {
  float um = __uint_as_float(*umax_bits);
  int e = 0;
  if (um > 0.f) frexpf(um, &e);  // um < 2^e
  int g = 14 - e;                // max |z| < 2^14
  zscale[0] = ldexpf(1.f, g);
  tscale[0] = ldexp(1.0, -2 * g);
  *umax_bits = 0u;
}

void launch_coeffs_split16(hipStream_t st, const double *w, const double *dsum, const float *rs2,
                           const double *bscale, int64_t n, float *u, unsigned int *umax_bits,
                           float *zscale, double *tscale) {
  if (n > 0)
    hipLaunchKernelGGL(k_coeffs_split16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, w,
                       dsum, rs2, bscale, n, u, umax_bits);
  hipLaunchKernelGGL(k_finalize_scale, dim3(1), dim3(1), 0, st, umax_bits, zscale, tscale);
}

// ---------------------------------------------------------------------------------------------
// T pass: lower 32x32 sub-tiles of Z^T Z with z_p = 2^g u_p x_p, over one point chunk per
// workgroup. A stage of 16 points is staged transposed and split in LDS: zT_hi/zT_lo [M][24]
// halves (column-major per point, padded), so an MFMA operand fragment (8 consecutive points of
// one column) is one 16-byte read. Double-buffered. v_mfma_f32_32x32x16_f16 operand layout:
// A[i][k] at lane i + 32(k/8), element k%8; B[k][j] at lane j + 32(k/8).
// Output: FP32 chunk partials [chunk][part][wave][t][1024] (MFMA output layout).
// ---------------------------------------------------------------------------------------------
static constexpr int SP_KP = 16;
static constexpr int SP_LD = 24;
static constexpr int SP_MAXT = 12;

__global__ void __launch_bounds__(512) k_syrk_split16(const float *__restrict__ x32,
                                                      const float *__restrict__ u,
                                                      const float *__restrict__ zscale, int64_t n,
                                                      int64_t m, int64_t chunk, int n_parts,
                                                      int max_tiles, int64_t n_chunks,
                                                      const int16_t *__restrict__ tiles,
                                                      float *__restrict__ partials) {
  extern __shared__ _Float16 zt[];  // [2 buffers][2 (hi, lo)][m][SP_LD]
  const int64_t plane = m * SP_LD;
  const int64_t g = blockIdx.x, xcd = g & 7, qq = g >> 3;
  const int part = (int)(qq % n_parts);
  const int64_t chunk_id = (qq / n_parts) * 8 + xcd;
  if (chunk_id >= n_chunks) return;
  const int tid = threadIdx.x, lane = tid & 63;
  // Wave-uniform index: keeps the tile list in scalar registers.
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int16_t *tl = tiles + ((int64_t)(part * 8 + wave) * max_tiles) * 2;
  int trg[SP_MAXT], tcg[SP_MAXT];
  int ntile = 0;
#pragma unroll
  for (int t = 0; t < SP_MAXT; t++) {
    trg[t] = t < max_tiles ? tl[2 * t] : -1;
    tcg[t] = t < max_tiles ? tl[2 * t + 1] : -1;
    if (trg[t] >= 0) ntile = t + 1;
  }
  float16_t acc[SP_MAXT];
#pragma unroll
  for (int t = 0; t < SP_MAXT; t++)
#pragma unroll
    for (int e = 0; e < 16; e++) acc[t][e] = 0.f;

  const float zs = zscale[0];
  const int64_t p_begin = chunk_id * chunk;
  const int64_t p_end = p_begin + chunk < n ? p_begin + chunk : n;
  const int n_stage = (int)((p_end - p_begin + SP_KP - 1) / SP_KP);

  auto stage = [&](int s, int buf) {
    const int64_t ps = p_begin + (int64_t)s * SP_KP;
    _Float16 *hi = zt + (int64_t)buf * 2 * plane, *lo = hi + plane;
    for (int64_t col = tid; col < m; col += 512) {
#pragma unroll
      for (int half = 0; half < 2; half++) {
        half8_t h0, l0;
#pragma unroll
        for (int j = 0; j < 8; j++) {
          int64_t p = ps + half * 8 + j;
          float z = 0.f;
          if (p < p_end) z = x32[p * m + col] * (u[p] * zs);
          _Float16 hh, ll;
          split_f32(z, hh, ll);
          h0[j] = hh;
          l0[j] = ll;
        }
        *reinterpret_cast<half8_t *>(hi + col * SP_LD + 8 * half) = h0;
        *reinterpret_cast<half8_t *>(lo + col * SP_LD + 8 * half) = l0;
      }
    }
  };

  if (n_stage > 0) stage(0, 0);
  __syncthreads();
  for (int s = 0; s < n_stage; s++) {
    const int buf = s & 1;
    if (s + 1 < n_stage) stage(s + 1, buf ^ 1);
    const _Float16 *hi = zt + (int64_t)buf * 2 * plane, *lo = hi + plane;
    const int kof = 8 * (lane >> 5);
    int last_rg = -1;
    half8_t ah = {}, al = {};
#pragma unroll
    for (int t = 0; t < SP_MAXT; t++) {
      if (t < ntile) {
        if (trg[t] != last_rg) {
          const int64_t row = (int64_t)trg[t] * 32 + (lane & 31);
          ah = *reinterpret_cast<const half8_t *>(hi + row * SP_LD + kof);
          al = *reinterpret_cast<const half8_t *>(lo + row * SP_LD + kof);
          last_rg = trg[t];
        }
        const int64_t colr = (int64_t)tcg[t] * 32 + (lane & 31);
        half8_t bh = *reinterpret_cast<const half8_t *>(hi + colr * SP_LD + kof);
        half8_t bl = *reinterpret_cast<const half8_t *>(lo + colr * SP_LD + kof);
        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, acc[t], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  float *out = partials + ((chunk_id * n_parts + part) * 8 + wave) * (int64_t)max_tiles * 1024;
#pragma unroll
  for (int t = 0; t < SP_MAXT; t++) {
    if (t < ntile) {
#pragma unroll
      for (int e = 0; e < 16; e++) out[t * 1024 + e * 64 + lane] = acc[t][e];
    }
  }
}

size_t syrk_split16_lds(int64_t m_cols) { return sizeof(_Float16) * 2 * 2 * m_cols * SP_LD; }

void launch_syrk_split16(hipStream_t st, const SyrkPlan &plan, const float *x32, const float *u,
                         const float *zscale, int64_t n, float *partials) {
  int64_t n_ch8 = (plan.n_chunks + 7) / 8 * 8;
  int64_t grid = n_ch8 * plan.n_parts;
  hipLaunchKernelGGL(k_syrk_split16, dim3((unsigned)grid), dim3(512), syrk_split16_lds(plan.m_cols),
                     st, x32, u, zscale, n, plan.m_cols, plan.chunk, plan.n_parts, plan.max_tiles,
                     plan.n_chunks, plan.d_tiles, partials);
}

// FP32 chunk partials -> FP64 sums in chunk order, scattered to the full M x M (lower kept on
// diagonal sub-tiles). 32x32 MFMA output layout: element e of lane l is
// (row 8(e/4) + 4(l/32) + e%4, col l%32).
__global__ void k_syrk_reduce_f32(const float *__restrict__ partials,
                                  const int16_t *__restrict__ tiles, int n_parts, int waves,
                                  int max_tiles, int64_t n_chunks, int64_t m,
                                  double *__restrict__ xx) {
  const int64_t per_chunk = (int64_t)n_parts * waves * max_tiles * 1024;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= per_chunk) return;
  int64_t slot = idx / 1024;
  int el = (int)(idx % 1024);
  int rg = tiles[2 * slot], cg = tiles[2 * slot + 1];
  if (rg < 0) return;
  double s = 0.0;
  for (int64_t c = 0; c < n_chunks; c++) s += (double)partials[c * per_chunk + idx];
  int e = el / 64, ln = el % 64;
  int r = 8 * (e / 4) + 4 * (ln / 32) + (e % 4), col = ln % 32;
  if (rg == cg && r < col) return;
  int64_t gi = (int64_t)rg * 32 + r, gj = (int64_t)cg * 32 + col;
  xx[gi * m + gj] = s;
  xx[gj * m + gi] = s;
}

void launch_syrk_reduce_f32(hipStream_t st, const SyrkPlan &plan, const float *partials,
                            double *xx_full) {
  int64_t per_chunk = (int64_t)plan.n_parts * plan.waves * plan.max_tiles * 1024;
  hipLaunchKernelGGL(k_syrk_reduce_f32, dim3((unsigned)((per_chunk + 255) / 256)), dim3(256), 0,
                     st, partials, plan.d_tiles, plan.n_parts, plan.waves, plan.max_tiles,
                     plan.n_chunks, plan.m_cols, xx_full);
}

}  // namespace cy3q
