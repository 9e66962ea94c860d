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
static constexpr int HS_WAVES = 8;    // 2 x 4 waves, each a 128 x 64 block
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

// x [rows][ldx] FP64 -> hi, lo [stages][mc][16] fp16 of fp32(x 2^s); zero past `rows` and past
// column ldx. One thread per (stage, column): 16 loads down the column (coalesced across
// threads), 32 bytes per plane.
__global__ void __launch_bounds__(256) k_split_rows(const double *__restrict__ x, int64_t rows,
                                                    int64_t ldx, int64_t stages, int64_t mc,
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
    const float v = p < rows && c < ldx ? (float)(x[p * ldx + c] * sc) : 0.f;
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
                       const double *scale, half8_t *xh, half8_t *xl, int64_t ldx) {
  const int64_t tot = stages * mc;
  if (tot <= 0) return;
  if (ldx <= 0) ldx = mc;
  hipLaunchKernelGGL(k_split_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, x, rows,
                     ldx, stages, mc, scale, xh, xl);
}

// ---------------------------------------------------------------------------------------------
// C_b,k[lower 256 x 256 tiles] += sum over the stages of part k of X^T X, b = blockIdx.z (buffer),
// k = blockIdx.y (point part), tile = blockIdx.x. 8 waves; wave (qr, qc) owns the 128 x 64
// block (rows 128 qr.., columns 64 qc..) of the tile as 4 x 2 MFMA tiles of 32 x 32, so a stage
// reads 12 fragments from LDS for 24 MFMAs (the two blocks above the diagonal of a diagonal tile
// are skipped: only the lower triangle is read later).
// A stage is the two 256-column panels I and J of 16 points, hi and lo planes (32 KB; a
// diagonal tile copies its one panel twice). Stages stream into a ring of HS_RING LDS slots,
// HS_RING - 1 stages ahead, by 16-byte global_load_lds copies: each wave copies four 1 KB runs
// (64 consecutive units of one panel plane) per stage, so no staging registers are used and
// three stages (96 KB) are in flight per CU. The stage loop is unrolled by HS_RING and every
// slot is its own __shared__ array, so the copies into the other slots are told apart from the
// reads of the current one and the loop keeps them in flight. The barrier is a bare s_barrier
// after an explicit wait for this stage's copies (a workgroup fence would drain all copies).
// 32x32 MFMA output: element e of lane l is (row 8 (e / 4) + 4 (l / 32) + e % 4, column l % 32).
// ---------------------------------------------------------------------------------------------
struct HerkSplitArgs {
  const half8_t *xh[4];
  const half8_t *xl[4];
  double *c[4];
};

static constexpr int HS_RING = 4;
static constexpr int HS_SLOT = 4 * HS_TB * 2;  // half8 units per slot: [panel][plane][512]

// s_waitcnt with vmcnt = v (0..63) and the other counters at their maxima (gfx9 encoding:
// vmcnt[3:0] in bits 3:0, vmcnt[5:4] in bits 15:14, expcnt in 6:4, lgkmcnt in 11:8).
#define HS_WAIT_VM(v) \
  __builtin_amdgcn_s_waitcnt(((v)&15) | (((v) >> 4) << 14) | (7 << 4) | (15 << 8))

// Wait for the copies of the current stage (four per wave per stage; `ahead` later stages, at
// most HS_RING - 2, may stay in flight; copies complete in issue order), then the barrier.
__device__ __forceinline__ void hs_stage_sync(int ahead) {
  static_assert(HS_RING == 4 && HS_WAVES == 8, "wait immediates: two stages of 4 copies ahead");
  if (ahead >= 2)
    HS_WAIT_VM(8);
  else if (ahead == 1)
    HS_WAIT_VM(4);
  else
    HS_WAIT_VM(0);
  __builtin_amdgcn_s_barrier();
}

// Copy c = w, w + 8, w + 16, w + 24 (of 32) of a stage: panel c >> 4, plane (c >> 3) & 1, units
// 64 (c & 7) + lane. Panel 0 is read from planes (ih, il) at unit off_i, panel 1 from (jh, jl)
// at off_j.
__device__ __forceinline__ void hs_issue(half8_t *slot, const half8_t *__restrict__ ih,
                                         const half8_t *__restrict__ il, int64_t off_i,
                                         const half8_t *__restrict__ jh,
                                         const half8_t *__restrict__ jl, int64_t off_j, int wave,
                                         int lane) {
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int cp = wave + 8 * u;
    const int panel = cp >> 4, plane = (cp >> 3) & 1;
    const half8_t *src = (panel ? (plane ? jl : jh) : (plane ? il : ih)) +
                         (panel ? off_j : off_i) + 64 * (cp & 7) + lane;
    __builtin_amdgcn_global_load_lds((const void *)src,
                                     (__attribute__((address_space(3))) void *)(slot + cp * 64),
                                     16, 0, 0);
  }
}

__global__ void __launch_bounds__(64 * HS_WAVES) k_herk_split16(HerkSplitArgs args, int64_t stages,
                                                               int64_t mc, int nparts, int nbuf,
                                                               int n_tiles,
                                                               const double *__restrict__ scale) {
  __shared__ half8_t ring0[HS_SLOT], ring1[HS_SLOT], ring2[HS_SLOT], ring3[HS_SLOT];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int qr = wave >> 2, qc = wave & 3;
  // Workgroup L is dealt to XCD L % 8. Each XCD takes whole (buffer, part) groups, g = xcd,
  // xcd + 8, ..., one after another, and a group's tiles in order, so the workgroups that run
  // together on one XCD share the panels of their stage range in its L2 (about 8 tiles per panel
  // instead of almost none when a group's tiles are spread over the XCDs).
  const int L = blockIdx.x, xcd = L & 7, jx = L >> 3;
  const int grp = (jx / n_tiles) * 8 + xcd;
  if (grp >= nbuf * nparts) return;
  const int t = jx % n_tiles;
  const int b = grp / nparts, kp = grp % nparts;
  int I = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= t) I++;
  while (I * (I + 1) / 2 > t) I--;
  const int J = t - I * (I + 1) / 2;
  const bool diag = I == J;
  const bool active = !(diag && qr == 0 && qc >= 2);
  const int64_t s0 = stages * kp / nparts, s1 = stages * (kp + 1) / nparts;
  const int nst = (int)(s1 - s0);
  const half8_t *__restrict__ xh = args.xh[b];
  const half8_t *__restrict__ xl = args.xl[b];

  float16_t acc[4][2];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int e = 0; e < 16; e++) acc[a][c][e] = 0.f;

  const double restore = scale[1];
  const int fl = lane & 31, fh = lane >> 5;
  auto flush = [&]() {
    if (!active) return;
    // The lane's element (a, c, e) is at cm + (32 a + 8 (e / 4) + e % 4) mc + 32 c. The asm
    // makes cm opaque here, so its 64-bit addresses are formed at the flush instead of being
    // hoisted out of the stage loop (where they would not fit in registers).
    double *cm = args.c[b] + (int64_t)kp * mc * mc +
                 ((int64_t)I * HS_TB + qr * 128 + 4 * fh) * mc + (int64_t)J * HS_TB + qc * 64 + fl;
    asm volatile("" : "+v"(cm));
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int c = 0; c < 2; c++)
#pragma unroll
        for (int e = 0; e < 16; e++) {
          double *q = cm + (int64_t)(32 * a + 8 * (e / 4) + e % 4) * mc + 32 * c;
          *q += (double)acc[a][c][e] * restore;
          acc[a][c][e] = 0.f;
        }
  };
  // Unit 0 of panel I / J at stage s: ((s0 + s) * mc + column base) * 2.
  auto issue = [&](int s, half8_t *slot) {
    const int64_t st = (s0 + s) * mc;
    hs_issue(slot, xh, xl, (st + (int64_t)I * HS_TB) * 2, xh, xl, (st + (int64_t)J * HS_TB) * 2,
             wave, lane);
  };
  auto compute = [&](const half8_t *slot) {
    if (!active) return;
    half8_t ah[4], al[4], bh[2], bl[2];
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const int ca = (qr * 128 + a * 32 + fl) * 2 + fh;
      ah[a] = slot[ca];
      al[a] = slot[HS_TB * 2 + ca];
    }
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const int cb = (qc * 64 + c * 32 + fl) * 2 + fh;
      bh[c] = slot[2 * HS_TB * 2 + cb];
      bl[c] = slot[3 * HS_TB * 2 + cb];
    }
#pragma unroll
    for (int pr = 0; pr < 3; pr++)
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int c = 0; c < 2; c++)
          acc[a][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(pr == 2 ? al[a] : ah[a],
                                                             pr == 1 ? bl[c] : bh[c], acc[a][c],
                                                             0, 0, 0);
  };
  // Stage s: wait for its slot, refill the slot of stage s - 1 with stage s + 3, compute.
  auto step = [&](int s, const half8_t *cur, half8_t *refill) {
    if (s >= nst) return;
    hs_stage_sync(min(nst - 1 - s, HS_RING - 2));
    if (s + HS_RING - 1 < nst) issue(s + HS_RING - 1, refill);
    compute(cur);
  };

  static_assert(HS_FLUSH % HS_RING == 0, "flushes fall on ring boundaries");
  if (nst > 0) issue(0, ring0);
  if (nst > 1) issue(1, ring1);
  if (nst > 2) issue(2, ring2);
  for (int sb = 0; sb < nst; sb += HS_RING) {
    step(sb, ring0, ring3);
    step(sb + 1, ring1, ring0);
    step(sb + 2, ring2, ring1);
    step(sb + 3, ring3, ring2);
    if ((sb + HS_RING) % HS_FLUSH == 0 && sb + HS_RING < nst) flush();
  }
  flush();
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
  const int groups8 = (nbuf * nparts + 7) / 8;  // groups per XCD
  hipLaunchKernelGGL(k_herk_split16, dim3((unsigned)(8 * groups8 * n_tiles)), dim3(64 * HS_WAVES),
                     0, st, a, stages, mc, nparts, nbuf, n_tiles, scale);
}

// Point parts for nbuf buffers of n_tiles tiles on n_cu CUs (one workgroup per CU): the count in
// 1..8 whose workgroups fill the XCDs' rounds best (each XCD runs whole groups of n_tiles).
int herk_split_parts(int nbuf, int n_tiles, int n_cu) {
  const int per_xcd = n_cu / 8 > 0 ? n_cu / 8 : 1;
  int best = 1;
  double best_eff = -1.0;
  for (int np = 1; np <= 8; np++) {
    const int groups8 = (nbuf * np + 7) / 8;
    const int64_t wg = (int64_t)groups8 * n_tiles;  // workgroups of the busiest XCD
    const int64_t rounds = (wg + per_xcd - 1) / per_xcd;
    const double eff = (double)nbuf * np * n_tiles / (8.0 * rounds * per_xcd);
    if (eff > best_eff + 1e-9) {
      best_eff = eff;
      best = np;
    }
  }
  return best;
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
//   * k_rows_split_planes scales row r by 2^(14 - ex[r]) (its largest |x| below 2^14) and
//     writes the fp16 hi/lo planes [m / 16][rows_pad][16] (zero rows up to rows_pad, a multiple
//     of 256), and ex[r];
//   * B comes as fp16 hi/lo planes [m / 16][mcp][16] of B 2^sB (launch_split_rows with one
//     power-of-two scale, bscale[0] = 2^sB, mcp = m rounded up to 256, zero columns past m);
//   * k_rowgemm_split16 forms each 256 x 256 tile of Y as k_herk_split16 forms its tiles (the
//     same 8 waves, stage ring and MFMA order; panel I from the row planes, panel J from B's),
//     with FP32 accumulation over the tile's k range, and stores Y = acc 2^(ex[r] - 14) /
//     bscale[0] in FP64. Workgroups are dealt to XCDs so that the column blocks of one row
//     tile share an L2.
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_rows_split_planes(const double *__restrict__ x,
                                                            int64_t rows, int64_t rows_pad,
                                                            int64_t m, int *__restrict__ ex,
                                                            half8_t *__restrict__ xh,
                                                            half8_t *__restrict__ xl) {
  // One wave per row; lane l owns the 16-byte units u = l, l + 64, ... (k = 8u .. 8u + 7).
  const int lane = threadIdx.x & 63;
  const int64_t r = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 6);
  if (r >= rows_pad) return;
  const int nu = (int)(m / 8);
  constexpr int MAXU = 2;  // m <= 1024
  double v[MAXU][8];
  double mx = 0.0;
#pragma unroll
  for (int i = 0; i < MAXU; i++) {
    const int u = lane + 64 * i;
#pragma unroll
    for (int j = 0; j < 8; j += 2) {
      double2 d = double2{0.0, 0.0};
      if (r < rows && u < nu) d = *reinterpret_cast<const double2 *>(x + r * m + 8 * u + j);
      v[i][j] = d.x;
      v[i][j + 1] = d.y;
      mx = fmax(mx, fmax(fabs(d.x), fabs(d.y)));
    }
  }
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
  int e = 0;
  if (mx > 0.0) frexp(mx, &e);
  if (lane == 0 && r < rows) ex[r] = e;
  const double sc = ldexp(1.0, 14 - e);
#pragma unroll
  for (int i = 0; i < MAXU; i++) {
    const int u = lane + 64 * i;
    if (u >= nu) continue;
    half8_t h, l;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const float f = (float)(v[i][j] * sc);
      h[j] = (_Float16)f;
      l[j] = (_Float16)(f - (float)h[j]);
    }
    const int64_t q = ((int64_t)(u >> 1) * rows_pad + r) * 2 + (u & 1);
    xh[q] = h;
    xl[q] = l;
  }
}

__global__ void __launch_bounds__(64 * HS_WAVES) k_rowgemm_split16(
    const half8_t *__restrict__ ah, const half8_t *__restrict__ al, int64_t rows,
    int64_t rows_pad, int64_t m, int64_t mcp, const int *__restrict__ ex,
    const half8_t *__restrict__ bh, const half8_t *__restrict__ bl,
    const double *__restrict__ bscale, double *__restrict__ y, int64_t ldy, int64_t n_rt,
    int n_cb) {
  __shared__ half8_t ring0[HS_SLOT], ring1[HS_SLOT], ring2[HS_SLOT], ring3[HS_SLOT];
  const int64_t g = blockIdx.x;
  const int64_t xcd = g & 7, qq = g >> 3;
  const int cb = (int)(qq % n_cb);
  const int64_t rt = (qq / n_cb) * 8 + xcd;
  if (rt >= n_rt) return;
  const int64_t r0 = rt * HS_TB;
  const int64_t c0 = (int64_t)cb * HS_TB;
  const int64_t kend = c0 + HS_TB < m ? c0 + HS_TB : m;
  const int nst = (int)(kend / HS_PTS);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int qr = wave >> 2, qc = wave & 3;

  float16_t acc[4][2];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int e = 0; e < 16; e++) acc[a][c][e] = 0.f;

  auto issue = [&](int s, half8_t *slot) {
    hs_issue(slot, ah, al, ((int64_t)s * rows_pad + r0) * 2, bh, bl, ((int64_t)s * mcp + c0) * 2,
             wave, lane);
  };
  const int fl = lane & 31, fh = lane >> 5;
  auto compute = [&](const half8_t *slot) {
    half8_t xa[4], xb[4], ya[2], yb[2];
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const int ca = (qr * 128 + a * 32 + fl) * 2 + fh;
      xa[a] = slot[ca];
      xb[a] = slot[HS_TB * 2 + ca];
    }
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const int cb2 = (qc * 64 + c * 32 + fl) * 2 + fh;
      ya[c] = slot[2 * HS_TB * 2 + cb2];
      yb[c] = slot[3 * HS_TB * 2 + cb2];
    }
#pragma unroll
    for (int pr = 0; pr < 3; pr++)
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int c = 0; c < 2; c++)
          acc[a][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(pr == 2 ? xb[a] : xa[a],
                                                             pr == 1 ? yb[c] : ya[c], acc[a][c],
                                                             0, 0, 0);
  };
  auto step = [&](int s, const half8_t *cur, half8_t *refill) {
    if (s >= nst) return;
    hs_stage_sync(min(nst - 1 - s, HS_RING - 2));
    if (s + HS_RING - 1 < nst) issue(s + HS_RING - 1, refill);
    compute(cur);
  };
  if (nst > 0) issue(0, ring0);
  if (nst > 1) issue(1, ring1);
  if (nst > 2) issue(2, ring2);
  for (int sb = 0; sb < nst; sb += HS_RING) {
    step(sb, ring0, ring3);
    step(sb + 1, ring1, ring0);
    step(sb + 2, ring2, ring1);
    step(sb + 3, ring3, ring2);
  }

  const double bre = 1.0 / bscale[0];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int e4 = 0; e4 < 4; e4++)
#pragma unroll
      for (int e1 = 0; e1 < 4; e1++) {
        const int64_t gr = r0 + qr * 128 + a * 32 + 8 * e4 + 4 * fh + e1;
        if (gr >= rows) continue;
        const double rs = ldexp(bre, ex[gr] - 14);
#pragma unroll
        for (int c = 0; c < 2; c++) {
          const int64_t gc = c0 + qc * 64 + c * 32 + fl;
          if (gc < m) y[gr * ldy + gc] = (double)acc[a][c][4 * e4 + e1] * rs;
        }
      }
}

int64_t rowgemm_rows_pad(int64_t rows) { return (rows + HS_TB - 1) / HS_TB * HS_TB; }
int64_t rowgemm_cols_pad(int64_t m) { return (m + HS_TB - 1) / HS_TB * HS_TB; }

void launch_rows_split_planes(hipStream_t st, const double *x, int64_t rows, int64_t m, int *ex,
                              half8_t *xh, half8_t *xl) {
  if (rows <= 0) return;
  const int64_t rp = rowgemm_rows_pad(rows);
  hipLaunchKernelGGL(k_rows_split_planes, dim3((unsigned)(rp / 16)), dim3(1024), 0, st, x, rows,
                     rp, m, ex, xh, xl);
}

void launch_rowgemm_split16(hipStream_t st, const half8_t *xh, const half8_t *xl, int64_t rows,
                            int64_t m, const int *ex, const half8_t *bh, const half8_t *bl,
                            const double *bscale, double *y, int64_t ldy) {
  if (rows <= 0) return;
  const int64_t rp = rowgemm_rows_pad(rows), mcp = rowgemm_cols_pad(m);
  const int64_t n_rt = rp / HS_TB;
  const int64_t n_rt8 = (n_rt + 7) / 8 * 8;
  const int n_cb = (int)(mcp / HS_TB);
  hipLaunchKernelGGL(k_rowgemm_split16, dim3((unsigned)(n_rt8 * n_cb)), dim3(64 * HS_WAVES), 0,
                     st, xh, xl, rows, rp, m, mcp, ex, bh, bl, bscale, y, ldy, n_rt, n_cb);
}

}  // namespace cy3q
