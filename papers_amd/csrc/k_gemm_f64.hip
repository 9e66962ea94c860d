// FP64 MFMA contractions of the Donaldson T-operator (the plain FP64 GPU path), also used
// for FS-Gram orthonormalisation and the sigma pre-contraction.
//
// The reference computes K_p = s^T H s with a 2N x 2N quadratic form per point
// (cy3/src/quintic.rs:1745-1759). It then accumulates T = sum_p (w_p/K_p) conj(s_p) s_p^T
// sequentially (quintic.rs:1771-1789). Here both are dense contractions on
// v_mfma_f64_16x16x4f64:
//   * k_trigemm_f64: Y = X * B, with X the real-ified section rows and B the real-ified
//     transpose of a lower-triangular complex V (B is block upper triangular, so the K loop
//     stops at the column block). With V = L^-1 and T = L L^H, the row sum of squares of Y is
//     D_p = s_p^H T^-1 s_p = K_p. With V = conj(L_gram)^-1, Y is the FS-orthonormalised basis.
//   * k_syrk_f64: the lower 16x16 tiles of X^T diag(c) X, split over point chunks, with one
//     FP64 partial per chunk. k_syrk_reduce sums the chunks in a fixed order, so the result
//     is deterministic run to run.
#include "cy3q_internal.h"

namespace cy3q {

// ----------------------------------------------------------------------------------------
// Y = X * B. Workgroup tile: 128 rows x 32 columns, 4 waves. Wave w owns rows
// [32w, 32w+32) as a 2x2 grid of 16x16 MFMA tiles. LDS: A [128][KT+1], B [KT][32], double
// buffered; the next k-block is fetched into registers while the current one runs.
// MFMA f64 16x16x4 operand layout: A[i][k] at lane i + 16k; B[k][j] at lane j + 16k.
// D[i][j] sits at lane j + 16*(i%4), element i/4 (so element e of lane l is row 4e + l/16;
// this differs from the f32-output MFMAs and was checked with a one-hot probe).
// ----------------------------------------------------------------------------------------
static constexpr int TG_ROWS = 128;
static constexpr int TG_COLS = 32;
static constexpr int TG_KT = 16;

__global__ void __launch_bounds__(256) k_trigemm_f64(const double *__restrict__ x, int64_t rows,
                                                     int64_t m_cols, const double *__restrict__ bm,
                                                     double *__restrict__ y, int mode, int full,
                                                     int64_t n_rowtiles) {
  __shared__ double sa[2][TG_ROWS][TG_KT + 1];
  __shared__ double sb[2][TG_KT][TG_COLS];
  const int n_cb = (int)(m_cols / TG_COLS);
  // XCD-aware order: consecutive workgroups are dealt round-robin over the 8 XCDs, so the
  // column blocks of one row tile are grouped on one XCD and share its L2.
  const int64_t g = blockIdx.x;
  const int64_t xcd = g & 7, q = g >> 3;
  const int cb = (int)(q % n_cb);
  const int64_t rt = (q / n_cb) * 8 + xcd;
  if (rt >= n_rowtiles) return;
  const int64_t r0 = rt * TG_ROWS;
  const int c0 = cb * TG_COLS;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t kend = full ? m_cols : (int64_t)(c0 + TG_COLS);
  const int nk = (int)(kend / TG_KT);

  double4_t acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};

  // Register staging: A 128 x 16 (4 double2 per thread), B 16 x 32 (1 double2 per thread).
  double2 ra[4], rb;
  auto fetch = [&](int kk) {
    const int64_t k0 = (int64_t)kk * TG_KT;
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int e = tid + 256 * u;
      const int r = e >> 3, c2 = (e & 7) * 2;
      const int64_t gr = r0 + r;
      double2 v = {0.0, 0.0};
      if (gr < rows) v = *reinterpret_cast<const double2 *>(x + gr * m_cols + k0 + c2);
      ra[u] = v;
    }
    {
      const int kr = tid >> 4, c2 = (tid & 15) * 2;
      rb = *reinterpret_cast<const double2 *>(bm + (k0 + kr) * m_cols + c0 + c2);
    }
  };
  auto put = [&](int buf) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int e = tid + 256 * u;
      const int r = e >> 3, c2 = (e & 7) * 2;
      sa[buf][r][c2] = ra[u].x;
      sa[buf][r][c2 + 1] = ra[u].y;
    }
    const int kr = tid >> 4, c2 = (tid & 15) * 2;
    sb[buf][kr][c2] = rb.x;
    sb[buf][kr][c2 + 1] = rb.y;
  };
  if (nk > 0) {
    fetch(0);
    put(0);
  }
  __syncthreads();
  for (int kk = 0; kk < nk; kk++) {
    const int buf = kk & 1;
    if (kk + 1 < nk) fetch(kk + 1);
#pragma unroll
    for (int ks = 0; ks < TG_KT / 4; ks++) {
      const int k = ks * 4 + (lane >> 4);
      double a0 = sa[buf][wave * 32 + (lane & 15)][k];
      double a1 = sa[buf][wave * 32 + 16 + (lane & 15)][k];
      double b0 = sb[buf][k][lane & 15];
      double b1 = sb[buf][k][16 + (lane & 15)];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    if (kk + 1 < nk) put(buf ^ 1);
    __syncthreads();
  }

  if (mode == 0) {
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++)
        for (int e = 0; e < 4; e++) {
          int64_t gr = r0 + wave * 32 + i * 16 + 4 * e + (lane >> 4);
          int gc = c0 + j * 16 + (lane & 15);
          if (gr < rows) y[gr * m_cols + gc] = acc[i][j][e];
        }
  } else {
    // Row sum of squares over this block's 32 columns -> y[cb * rows + row].
    for (int i = 0; i < 2; i++)
      for (int e = 0; e < 4; e++) {
        double sq = acc[i][0][e] * acc[i][0][e] + acc[i][1][e] * acc[i][1][e];
        for (int off = 1; off < 16; off <<= 1) sq += __shfl_xor(sq, off);
        int64_t gr = r0 + wave * 32 + i * 16 + 4 * e + (lane >> 4);
        if ((lane & 15) == 0 && gr < rows) y[(int64_t)cb * rows + gr] = sq;
      }
  }
}

void launch_trigemm_f64(hipStream_t st, const double *x, int64_t rows, int64_t m_cols,
                        const double *bmat, double *y, int mode, int full) {
  if (rows <= 0) return;
  int64_t n_rt = (rows + TG_ROWS - 1) / TG_ROWS;
  int64_t n_rt8 = (n_rt + 7) / 8 * 8;
  int64_t n_cb = m_cols / TG_COLS;
  int64_t grid = n_rt8 * n_cb;
  hipLaunchKernelGGL(k_trigemm_f64, dim3((unsigned)grid), dim3(256), 0, st, x, rows, m_cols, bmat,
                     y, mode, full, n_rt);
}

// ----------------------------------------------------------------------------------------
// Weighted SYRK partials. Each workgroup (8 waves) handles one point chunk and one "part" (a
// set of lower 16x16 sub-tiles of the M x M result). Wave w of part q takes its own list of
// up to max_tiles sub-tiles, ordered by row group so the A fragment is reused.
// Per stage, KP points of X are staged in LDS (row-major [KP][M]) with their weights c.
// Output: partials[chunk][part][wave][t][256], one 16x16 tile per (wave, t) slot.
// ----------------------------------------------------------------------------------------
static constexpr int SY_WAVES = 8;
// Register budget: MAXT 16x16 accumulators (4 doubles each) plus STG staged double2 per thread;
// STG * 1024 >= KP * m_cols must hold: <22, 7, 16> for m_cols <= 448, <12, 8, 8> up to 1024.

template <int SY_MAXT, int SY_STG, int SY_KP>
__global__ void __launch_bounds__(512) k_syrk_f64(const double *__restrict__ x,
                                                  const double *__restrict__ cw, int64_t n,
                                                  int64_t m_cols, int64_t chunk, int n_parts,
                                                  int max_tiles, int64_t n_chunks,
                                                  const int16_t *__restrict__ tiles,
                                                  double *__restrict__ partials) {
  extern __shared__ double sx[];  // 2 buffers of [SY_KP][m_cols + 1] then c[2][SY_KP]
  const int64_t ld = m_cols + 1;
  const int64_t bufsz = SY_KP * ld;
  double *sc = sx + 2 * bufsz;
  // XCD-aware: the parts of one chunk share an XCD (and its L2).
  const int64_t g = blockIdx.x, xcd = g & 7, qq = g >> 3;
  const int part = (int)(qq % n_parts);
  const int64_t chunk_id = (qq / n_parts) * 8 + xcd;
  if (chunk_id >= n_chunks) return;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int16_t *tl = tiles + ((int64_t)(part * SY_WAVES + wave) * max_tiles) * 2;
  int trg[SY_MAXT], tcg[SY_MAXT];
  int ntile = 0;
#pragma unroll
  for (int t = 0; t < SY_MAXT; t++) {
    trg[t] = t < max_tiles ? tl[2 * t] : -1;
    tcg[t] = t < max_tiles ? tl[2 * t + 1] : -1;
    if (trg[t] >= 0) ntile = t + 1;
  }

  double4_t acc[SY_MAXT];
#pragma unroll
  for (int t = 0; t < SY_MAXT; t++) acc[t] = (double4_t){0.0, 0.0, 0.0, 0.0};

  const int64_t p_begin = chunk_id * chunk;
  const int64_t p_end = p_begin + chunk < n ? p_begin + chunk : n;
  const int n_stage = (int)((p_end - p_begin + SY_KP - 1) / SY_KP);
  const int64_t half = m_cols / 2;
  const int64_t per = SY_KP * half;  // double2 per stage
  double2 rx[SY_STG];
  double rc = 0.0;
  auto fetch = [&](int s) {
    const int64_t p0 = p_begin + (int64_t)s * SY_KP;
#pragma unroll
    for (int u = 0; u < SY_STG; u++) {
      const int64_t e = tid + 512 * u;
      double2 v = {0.0, 0.0};
      if (e < per) {
        const int r = (int)(e / half);
        const int64_t cc = (e % half) * 2;
        if (p0 + r < p_end) v = *reinterpret_cast<const double2 *>(x + (p0 + r) * m_cols + cc);
      }
      rx[u] = v;
    }
    if (tid < SY_KP) rc = (p0 + tid < p_end) ? cw[p0 + tid] : 0.0;
  };
  auto put = [&](int buf) {
    double *b = sx + buf * bufsz;
#pragma unroll
    for (int u = 0; u < SY_STG; u++) {
      const int64_t e = tid + 512 * u;
      if (e < per) {
        const int r = (int)(e / half);
        const int64_t cc = (e % half) * 2;
        b[r * ld + cc] = rx[u].x;
        b[r * ld + cc + 1] = rx[u].y;
      }
    }
    if (tid < SY_KP) sc[buf * SY_KP + tid] = rc;
  };
  if (n_stage > 0) {
    fetch(0);
    put(0);
  }
  __syncthreads();
  for (int s = 0; s < n_stage; s++) {
    const int buf = s & 1;
    if (s + 1 < n_stage) fetch(s + 1);
    const double *b = sx + buf * bufsz;
#pragma unroll
    for (int ks = 0; ks < SY_KP / 4; ks++) {
      const int kp = ks * 4 + (lane >> 4);
      const double cpt = sc[buf * SY_KP + kp];
      const double *row = b + kp * ld;
      int last_rg = -1;
      double a = 0.0;
#pragma unroll
      for (int t = 0; t < SY_MAXT; t++) {
        if (t < ntile) {
          if (trg[t] != last_rg) {
            a = row[trg[t] * 16 + (lane & 15)];
            last_rg = trg[t];
          }
          double bv = cpt * row[tcg[t] * 16 + (lane & 15)];
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc[t], 0, 0, 0);
        }
      }
    }
    if (s + 1 < n_stage) put(buf ^ 1);
    __syncthreads();
  }
  double *out = partials + ((chunk_id * n_parts + part) * SY_WAVES + wave) * (int64_t)max_tiles * 256;
#pragma unroll
  for (int t = 0; t < SY_MAXT; t++) {
    if (t < ntile) {
      for (int e = 0; e < 4; e++) out[t * 256 + e * 64 + lane] = acc[t][e];
    }
  }
}

int syrk_f64_max_tiles(int64_t m_cols) { return m_cols <= 448 ? 22 : 12; }

void launch_syrk_f64(hipStream_t st, const SyrkPlan &plan, const double *x, const double *c,
                     int64_t n, double *partials) {
  int64_t n_ch8 = (plan.n_chunks + 7) / 8 * 8;
  int64_t grid = n_ch8 * plan.n_parts;
  if (plan.m_cols <= 448) {
    size_t lds = sizeof(double) * 2 * (16 * (plan.m_cols + 1) + 16);
    hipLaunchKernelGGL((k_syrk_f64<22, 7, 16>), dim3((unsigned)grid), dim3(512), lds, st, x, c, n,
                       plan.m_cols, plan.chunk, plan.n_parts, plan.max_tiles, plan.n_chunks,
                       plan.d_tiles, partials);
  } else {  // m_cols <= 1024: 8-point stages keep the two LDS buffers under 160 KB
    size_t lds = sizeof(double) * 2 * (8 * (plan.m_cols + 1) + 8);
    hipLaunchKernelGGL((k_syrk_f64<12, 8, 8>), dim3((unsigned)grid), dim3(512), lds, st, x, c, n,
                       plan.m_cols, plan.chunk, plan.n_parts, plan.max_tiles, plan.n_chunks,
                       plan.d_tiles, partials);
  }
}

// Sum of the chunk partials, in chunk order, scattered into the full M x M matrix (lower tiles
// and their mirror). One thread per (part, wave, tile slot, element).
// Element (e, lane) of a 16x16 f64 tile is (row 4e + lane/16, col lane%16), the MFMA output layout.
__global__ void k_syrk_reduce(const double *__restrict__ partials, const int16_t *__restrict__ tiles,
                              int n_parts, int waves, int max_tiles, int64_t n_chunks, int tile,
                              int64_t m_cols, double *__restrict__ xx) {
  const int64_t per_chunk = (int64_t)n_parts * waves * max_tiles * tile * tile;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= per_chunk) return;
  int64_t slot = idx / (tile * tile);
  int el = (int)(idx % (tile * tile));
  int rg = tiles[2 * slot], cg = tiles[2 * slot + 1];
  if (rg < 0) return;
  double s = 0.0;
  for (int64_t c = 0; c < n_chunks; c++) s += partials[c * per_chunk + idx];
  int r, col;
  if (tile == 16) {
    int e = el / 64, ln = el % 64;
    r = 4 * e + ln / 16;
    col = ln % 16;
  } else {  // 32x32 MFMA output layout: element e of lane ln -> row 8*(e/4) + 4*(ln/32) + e%4
    int e = el / 64, ln = el % 64;
    r = 8 * (e / 4) + 4 * (ln / 32) + (e % 4);
    col = ln % 32;
  }
  // A diagonal sub-tile holds both (r, col) and (col, r); the lower one is kept, so every
  // element of xx has exactly one writer.
  if (rg == cg && r < col) return;
  int64_t gi = (int64_t)rg * tile + r, gj = (int64_t)cg * tile + col;
  xx[gi * m_cols + gj] = s;
  xx[gj * m_cols + gi] = s;
}

void launch_syrk_reduce(hipStream_t st, const SyrkPlan &plan, const double *partials,
                        double *xx_full) {
  int64_t per_chunk = (int64_t)plan.n_parts * plan.waves * plan.max_tiles * plan.tile * plan.tile;
  hipLaunchKernelGGL(k_syrk_reduce, dim3((unsigned)((per_chunk + 255) / 256)), dim3(256), 0, st,
                     partials, plan.d_tiles, plan.n_parts, plan.waves, plan.max_tiles,
                     plan.n_chunks, plan.tile, plan.m_cols, xx_full);
}

}  // namespace cy3q
