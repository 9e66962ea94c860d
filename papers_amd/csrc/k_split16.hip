// Split-precision (fp16 hi/lo) MFMA contractions of the Donaldson T-operator.
//
// The two dense contractions of one T-iteration (reference: quintic.rs:1745-1789) run on gfx950's
// fp16 MFMA with FP32 accumulation. Each FP32 operand x is split as x = hi + lo with
// hi = fp16(x) and lo = fp16(x - hi); a product is hi*hi' + hi*lo' + lo*hi', about 22 bits per
// product. Overflow and subnormal loss are avoided by exact power-of-two scalings:
//   * every section row is stored scaled by 2^-e_p (max |x| in [0.5, 1)) and split into fp16
//     hi/lo planes once per workspace;
//   * W (the D-pass operand) is scaled globally by 2^-f;
//   * z_p = sqrt(c_p) x_p (the T-pass operand) is scaled globally by 2^g.
// The host-side bookkeeping undoes each scaling exactly. D_p and the chunk sums are accumulated in
// FP64. T is summed across point chunks in FP64, in a fixed order.
//
// D pass  (k_dpass_split16): v_mfma_f32_16x16x32_f16, 128 points x all M columns per workgroup,
//                            so D_p = sum_j Y_pj^2 completes inside the workgroup.
// T pass: k_tpass.hip.
#include <algorithm>
#include <utility>
#include <vector>

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
// Layout: bfrag[((ks * n_cg + cg) * 2 + {0: hi, 1: lo}) * 64 + lane], followed by one all-zero
// fragment pair (index n_ks * n_cg) for the D pass's idle slots. bscale[0] = 4^f restores D
// from the scaled contraction; bscale[1] = 2^-f is the scale applied to W.
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

// Scale: bscale[0] = 4^f and bscale[1] = 2^-f with 2^f > max |W| (one workgroup). W is lower
// triangular and only its lower triangle (row >= column) is defined; the rest is not read.
__global__ void __launch_bounds__(1024) k_bfrag_scale(const double *__restrict__ w, int nb,
                                                      double *__restrict__ bscale) {
  __shared__ double sh[1024];
  double mx = 0.0;
  // 32 rows at a time, 32 lanes along each row's lower part (no index division).
  for (int row = (int)(threadIdx.x >> 5); row < nb; row += 32) {
    const double *wr = w + 2 * (int64_t)row * nb;
    for (int col = (int)(threadIdx.x & 31); col <= row; col += 32)
      mx = fmax(mx, fmax(fabs(wr[2 * col]), fabs(wr[2 * col + 1])));
  }
  sh[threadIdx.x] = mx;
  __syncthreads();
  for (int off = 512; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    int f = 0;
    if (sh[0] > 0.0) frexp(sh[0], &f);
    bscale[0] = ldexp(1.0, 2 * f);
    bscale[1] = ldexp(1.0, -f);
  }
}

// One thread per (fragment, lane).
__global__ void k_bfrag_fill(const double *__restrict__ w, int nb, int64_t m,
                             const double *__restrict__ bscale, half8_t *__restrict__ bfrag) {
  const int n_ks = (int)(m / 32), n_cg = (int)(m / 16);
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ((int64_t)n_ks * n_cg + 1) * 64) return;
  const double sc = bscale[1];
  const int lane = (int)(e % 64);
  const int64_t frag = e / 64;
  if (frag == (int64_t)n_ks * n_cg) {  // the all-zero fragment pair
    bfrag[(frag * 2 + 0) * 64 + lane] = half8_t{};
    bfrag[(frag * 2 + 1) * 64 + lane] = half8_t{};
    return;
  }
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

void launch_build_bfrag16(hipStream_t st, const double *v_cplx, int nb, int64_t m_cols,
                          half8_t *bfrag, double *b_scale) {
  hipLaunchKernelGGL(k_bfrag_scale, dim3(1), dim3(1024), 0, st, v_cplx, nb, b_scale);
  const int64_t tot = ((m_cols / 32) * (m_cols / 16) + 1) * 64;
  hipLaunchKernelGGL(k_bfrag_fill, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, v_cplx,
                     nb, m_cols, b_scale, bfrag);
}

// ---------------------------------------------------------------------------------------------
// Static operand planes of the split path, built once per workspace from the (orthonormalised)
// FP64 sections. Row p is scaled by 2^-e_p (max |x| in [0.5, 1)); rs2[p] = 4^-e_p.
//  xd: D-pass A fragments of v_mfma_f32_16x16x32_f16. For point tile t (16 points), k-step ks
//      (32 columns), plane h (hi, lo), lane l: half8 of X[16t + l%16][32ks + 8(l/16) + j].
//      Index ((t * n_ks + ks) * 2 + h) * 64 + l.
//  xt: T-pass stage records (FP32), per column segment of sw columns (sw = m up to 416, else
//      192; the last segment holds the rest, ms columns). For stage s (16 points), quad q
//      (4 points), column c of segment a: the 4 floats X[16s + 4q + j][a sw + c], j = 0..3, at
//      n_pad a sw + ((s * 4 + q) * ms + c) * 4 + j, so that one 16-byte load per lane over
//      consecutive columns reads a contiguous run.
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_row_exponent(const double *__restrict__ sv, int64_t n,
                                                      int64_t m, int *__restrict__ ex,
                                                      float *__restrict__ rs2) {
  int lane = threadIdx.x & 63;
  int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (p >= n) return;
  double mx = 0.0;
  for (int64_t j = lane; j < m; j += 64) mx = fmax(mx, fabs(sv[p * m + j]));
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
  int e = 0;
  if (mx > 0.0) frexp(mx, &e);
  if (lane == 0) {
    ex[p] = e;
    rs2[p] = (float)ldexp(1.0, -2 * e);
  }
}

__global__ void k_split_planes(const double *__restrict__ sv, const int *__restrict__ ex,
                               int64_t n, int64_t n_pad, int64_t m, int64_t sw,
                               _Float16 *__restrict__ xd, float *__restrict__ xt) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_pad * m) return;
  const int64_t p = q / m, c = q % m;
  float x = 0.f;
  if (p < n) x = (float)(sv[q] * ldexp(1.0, -ex[p]));
  _Float16 h, l;
  split_f32(x, h, l);
  const int64_t n_ks = m / 32;
  const int64_t t = p / 16, ks = c / 32;
  const int lane = (int)(p % 16) + 16 * (int)((c % 32) / 8), j = (int)(c % 8);
  xd[((((t * n_ks + ks) * 2 + 0) * 64 + lane) * 8) + j] = h;
  xd[((((t * n_ks + ks) * 2 + 1) * 64 + lane) * 8) + j] = l;
  const int64_t a = c / sw, cl = c - a * sw, ms = m - a * sw < sw ? m - a * sw : sw;
  xt[n_pad * a * sw + (((p / 16) * 4 + (p % 16) / 4) * ms + cl) * 4 + p % 4] = x;
}

void launch_split_planes(hipStream_t st, const double *sv, int64_t n, int64_t n_pad, int64_t m,
                         int *ex, float *rs2, _Float16 *xd, float *xt) {
  hipLaunchKernelGGL(k_row_exponent, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, sv, n, m, ex,
                     rs2);
  int64_t tot = n_pad * m;
  hipLaunchKernelGGL(k_split_planes, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, sv, ex,
                     n, n_pad, m, tpass_segment_cols(m), xd, xt);
}

// ---------------------------------------------------------------------------------------------
// D pass. A block is 128 points (8 point tiles). The grid is persistent: one workgroup per CU
// (the LDS ring fills most of a CU's 160 KB), each looping over blocks blockIdx.x,
// blockIdx.x + gridDim.x, ... Waves 0..6 compute: wave w owns the column groups cglist[w][..]
// (round-robin, so the groups still active at each k-step are spread evenly) and keeps their
// accumulators for all 8 tiles of the current block. Wave 7 only moves data: it streams the A
// fragments (16 KB per k-step) straight from HBM into a DP_RING-slot LDS ring with
// global_load_lds_dwordx4, DP_RING - 1 k-steps ahead and across block boundaries, so the next
// block's first k-steps are already resident when the current block's reduction runs and no
// block waits for a cold HBM fetch. Its vector-memory counter tracks these copies alone, while
// the compute waves' counters track only their B-fragment loads (register prefetch one k-step
// ahead, also across blocks). One barrier per k-step and one per block (the row reduction).
// The three split products of one accumulator are spread out so consecutive MFMAs write
// different accumulators.
// Output: dsum[p] = sum_j Y_pj^2 (FP64), for the row-scaled x and the 2^-f scaled W.
// Requires xd padded with zeros to a multiple of 128 points.
// ---------------------------------------------------------------------------------------------
static constexpr int DP_TILES = 8;
static constexpr int DP_PTS = 16 * DP_TILES;
static constexpr int DP_WAVES = 8;
static constexpr int DP_CW = DP_WAVES - 1;  // compute waves
static constexpr int DP_MAXCG = 4;
static constexpr int DP_RING = 8;  // a power of two
static constexpr int DP_SLOT = DP_TILES * 2 * 64;  // half8 per ring slot (16 KB)
static constexpr int DP_LOADS = DP_TILES * 2;      // global_load_lds per k-step and lane
static_assert((DP_RING & (DP_RING - 1)) == 0, "ring index uses a mask");

int dpass_split16_waves() { return DP_CW; }
int dpass_split16_max_cg() { return DP_MAXCG; }

// s_waitcnt with vmcnt = v (0..63) and the other counters at their maxima (gfx9 encoding:
// vmcnt[3:0] in bits 3:0, vmcnt[5:4] in bits 15:14, expcnt in 6:4, lgkmcnt in 11:8).
#define DP_WAIT_VM(v) \
  __builtin_amdgcn_s_waitcnt(((v)&15) | (((v) >> 4) << 14) | (7 << 4) | (15 << 8))

// The data mover's wait before the barrier of k-step g: the copies of k-step g must have
// landed, while those of the `ahead` later k-steps already issued may stay in flight. Copies
// complete in issue order for the counter, so allowing ahead * DP_LOADS outstanding is exact;
// the counter field holds at most 63, and beyond that the wait is merely earlier than needed.
__device__ __forceinline__ void dp_mover_wait(int64_t ahead) {
  static_assert(DP_LOADS == 16, "wait immediates assume 16 copies per k-step");
  if (ahead >= 4) {
    DP_WAIT_VM(63);
  } else if (ahead == 3) {
    DP_WAIT_VM(48);
  } else if (ahead == 2) {
    DP_WAIT_VM(32);
  } else if (ahead == 1) {
    DP_WAIT_VM(16);
  } else {
    DP_WAIT_VM(0);
  }
}

// Largest number of active column groups (cg >= 2 ks) held by any compute wave at k-step ks,
// for the groups cg0 .. cg0 + n_cg - 1 dealt round-robin, cg -> wave (cg - cg0) % DP_CW.
__host__ __device__ constexpr int dp_slots(int cg0, int n_cg, int ks) {
  int best = 0;
  for (int w = 0; w < DP_CW; w++) {
    int cnt = 0;
    for (int cg = cg0 + w; cg < cg0 + n_cg; cg += DP_CW) cnt += cg >= 2 * ks ? 1 : 0;
    best = cnt > best ? cnt : best;
  }
  return best;
}

// Column range: the NCG column groups CG0 .. CG0 + NCG - 1 (at most DP_CW * DP_MAXCG) of an
// M = 32 NKS_ALL column contraction, which need the k-steps 0 .. KS_END - 1 (group cg is active
// in k-steps ks <= cg / 2). With one range (CG0 = 0, NCG = M / 16, KS_END = NKS_ALL) dsum is
// D_p itself; with several, each range writes its own partial sum and the caller adds them.
// The k-step count is a compile-time constant: the k-step loops unroll completely, so the
// compiler's memory-counter waits are exact (no loop-carried merging of pending loads).
template <int NKS_ALL, int KS_END, int CG0, int NCG>
__global__ void __launch_bounds__(64 * DP_WAVES) k_dpass_split16(const half8_t *__restrict__ xd,
                                                                 int64_t n, int64_t nblk,
                                                                 const half8_t *__restrict__ bfrag,
                                                                 const int16_t *__restrict__ cglist,
                                                                 double *__restrict__ dsum) {
  static_assert(NCG <= DP_CW * DP_MAXCG && CG0 + NCG <= 2 * NKS_ALL, "column range");
  static_assert(KS_END == (CG0 + NCG - 1) / 2 + 1, "k-steps of the range");
  __shared__ half8_t sa[DP_RING][DP_SLOT];  // slot: [(tile * 2 + plane) * 64 + lane]
  __shared__ double red[DP_CW][DP_PTS];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  constexpr int n_ks = KS_END, ks_all = NKS_ALL, n_cg_all = 2 * NKS_ALL;
  const int64_t blk0 = blockIdx.x, bstride = gridDim.x;
  const int64_t nmine = blk0 < nblk ? (nblk - 1 - blk0) / bstride + 1 : 0;

  if (wave == DP_CW) {
    // Data mover. Global k-step g is k-step g % n_ks of this workgroup's block g / n_ks. Slot
    // (g + DP_RING - 1) % DP_RING was last read in k-step g - 1, which every compute wave
    // finished before the barrier of k-step g.
    const int64_t total = nmine * n_ks;
    auto issue = [&](int64_t g) {
      const int64_t blk = blk0 + (g / n_ks) * bstride;
      const int ks = (int)(g % n_ks);
      half8_t *slot = sa[(int)(g & (DP_RING - 1))];
      const half8_t *src = xd + (blk * DP_TILES * ks_all * 2) * 64 + lane;
#pragma unroll
      for (int b = 0; b < DP_LOADS; b++) {
        const int rg = b >> 1, pl = b & 1;
        __builtin_amdgcn_global_load_lds(
            (const void *)(src + (((int64_t)rg * ks_all + ks) * 2 + pl) * 64),
            (__attribute__((address_space(3))) void *)(slot + b * 64), 16, 0, 0);
      }
    };
    for (int64_t g = 0; g < DP_RING - 1 && g < total; g++) issue(g);
    for (int64_t g = 0; g < total; g++) {
      dp_mover_wait(min((int64_t)(DP_RING - 2), total - 1 - g));
      // A bare barrier: a workgroup fence here would wait for every copy in flight.
      __builtin_amdgcn_s_barrier();
      if (g + DP_RING - 1 < total) issue(g + DP_RING - 1);
      if (g % n_ks == n_ks - 1) __builtin_amdgcn_s_barrier();  // the block's row reduction
    }
    return;
  }

  // The wave's column groups: round-robin (cg = wave + DP_CW j), ascending and right-aligned in
  // the DP_MAXCG slots (unused slots first, -1). At k-step ks the groups cg >= 2 ks are active,
  // always the wave's last slots. Every wave runs the same dp_slots(CG0, NCG, ks) slots (the largest
  // active count of any wave, a compile-time number), so the k-step body has no branches; a slot
  // that is idle for this wave multiplies by the all-zero fragment and leaves its sums unchanged.
  // The list is also passed in (cglist) by the host, which builds the same table.
  int cgs[DP_MAXCG];
#pragma unroll
  for (int c = 0; c < DP_MAXCG; c++) cgs[c] = cglist[wave * DP_MAXCG + c];
  float4_t acc[DP_TILES][DP_MAXCG];

  // Two B register sets, alternating by k-step parity (no register copies, so the compiler's
  // wait for the set in use leaves the other set's loads in flight). The last k-step of a block
  // prefetches k-step 0 of the next one into the free set.
  half8_t b0h[DP_MAXCG], b0l[DP_MAXCG], b1h[DP_MAXCG], b1l[DP_MAXCG];
  constexpr int64_t zero_frag = (int64_t)ks_all * n_cg_all;
  // The B fragments are the same for every block. An opaque copy of their base, renewed per
  // block, keeps the compiler from hoisting all n_ks * DP_MAXCG fragment addresses out of the
  // block loop (they would not fit in registers alongside the accumulators). The lambdas are
  // forced inline: their register arrays, passed by pointer, stay in registers only when inlined.
  const half8_t *bfp = bfrag;
  auto load_b = [&](int ks, half8_t *hh, half8_t *ll) __attribute__((always_inline)) {
#pragma unroll
    for (int c = 0; c < DP_MAXCG; c++) {
      if (c < DP_MAXCG - dp_slots(CG0, NCG, ks)) continue;
      const int64_t frag = cgs[c] >= 2 * ks ? (int64_t)ks * n_cg_all + cgs[c] : zero_frag;
      hh[c] = bfp[(frag * 2 + 0) * 64 + lane];
      ll[c] = bfp[(frag * 2 + 1) * 64 + lane];
    }
  };
  // MFMAs with swapped operands: B as the row operand and X as the column operand give the
  // transpose Y^T, so that element e of lane l is column 16 cg + 4 (l / 16) + e of point
  // 16 r + l % 16 (the fragment layouts of both operands serve either role).
  auto step = [&](int slot, int ks, const half8_t *bh,
                  const half8_t *bl) __attribute__((always_inline)) {
    const int c0 = DP_MAXCG - dp_slots(CG0, NCG, ks);
#pragma unroll
    for (int r = 0; r < DP_TILES; r += 2) {
      if (c0 >= DP_MAXCG) break;
      const half8_t ah0 = sa[slot][(r * 2 + 0) * 64 + lane];
      const half8_t al0 = sa[slot][(r * 2 + 1) * 64 + lane];
      const half8_t ah1 = sa[slot][(r * 2 + 2) * 64 + lane];
      const half8_t al1 = sa[slot][(r * 2 + 3) * 64 + lane];
#pragma unroll
      for (int pr = 0; pr < 3; pr++) {
#pragma unroll
        for (int c = 0; c < DP_MAXCG; c++) {
          if (c < c0) continue;
          const half8_t a0 = pr == 2 ? al0 : ah0, a1 = pr == 2 ? al1 : ah1;
          const half8_t b = pr == 1 ? bl[c] : bh[c];
          acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(b, a0, acc[r][c], 0, 0, 0);
          acc[r + 1][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(b, a1, acc[r + 1][c], 0, 0, 0);
        }
      }
    }
  };
  // One block: the k-steps (even ones on set x, odd ones on set y), then the row sums.
  auto block = [&](int64_t i, half8_t *xh, half8_t *xl, half8_t *yh,
                   half8_t *yl) __attribute__((always_inline)) {
#pragma unroll
    for (int r = 0; r < DP_TILES; r++)
#pragma unroll
      for (int c = 0; c < DP_MAXCG; c++) acc[r][c] = (float4_t){0.f, 0.f, 0.f, 0.f};
    const int slot0 = (int)((i * n_ks) & (DP_RING - 1));
    if constexpr (KS_END <= 20) asm volatile("" : "+s"(bfp));
    const bool more = i + 1 < nmine;
#pragma unroll
    for (int ks = 0; ks < n_ks; ks += 2) {
      // Past 20 k-steps (the top column ranges of k = 8) the base is renewed per k-step pair
      // instead, in vector registers: the compiler would otherwise start the B loads of many
      // later k-steps at once, more than the registers left beside the accumulators (and a
      // scalar-register base fails to compile at this size).
      if constexpr (KS_END > 20) asm volatile("" : "+v"(bfp));
      __syncthreads();
      if (ks + 1 < n_ks)
        load_b(ks + 1, yh, yl);
      else if (more)
        load_b(0, yh, yl);
      step((slot0 + ks) & (DP_RING - 1), ks, xh, xl);
      if (ks + 1 >= n_ks) break;
      __syncthreads();
      if (ks + 2 < n_ks)
        load_b(ks + 2, xh, xl);
      else if (more)
        load_b(0, xh, xl);
      step((slot0 + ks + 1) & (DP_RING - 1), ks + 1, yh, yl);
    }
    // Row sums of squares (FP64). Lane l holds, for tile r, point 16 r + l % 16 and columns
    // 4 (l / 16) + e of every slot (inactive slots are still zero): sum them in the lane, then
    // over the four 16-lane rows in two halving exchanges (lanes l ^ 32, then l ^ 16), after
    // which lane l owns the complete sums of tiles 4 (l / 32) + 2 ((l / 16) % 2) + {0, 1}.
    double part[DP_TILES];
#pragma unroll
    for (int r = 0; r < DP_TILES; r++) {
      double sacc = 0.0;
#pragma unroll
      for (int c = 0; c < DP_MAXCG; c++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const double y = (double)acc[r][c][e];
          sacc += y * y;
        }
      part[r] = sacc;
    }
    const bool up32 = (lane & 32) != 0, up16 = (lane & 16) != 0;
    double h4[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const double send = up32 ? part[q] : part[4 + q];
      const double keep = up32 ? part[4 + q] : part[q];
      h4[q] = keep + __shfl_xor(send, 32);
    }
    double h2[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const double send = up16 ? h4[q] : h4[2 + q];
      const double keep = up16 ? h4[2 + q] : h4[q];
      h2[q] = keep + __shfl_xor(send, 16);
    }
    // red is rewritten only after the next block's first barrier, which the reading waves
    // reach after their reads.
    const int rbase = 4 * (lane >> 5) + 2 * ((lane >> 4) & 1);
    red[wave][(rbase + 0) * 16 + (lane & 15)] = h2[0];
    red[wave][(rbase + 1) * 16 + (lane & 15)] = h2[1];
    __syncthreads();
    if (tid < DP_PTS) {
      double sacc = 0.0;
      for (int w = 0; w < DP_CW; w++) sacc += red[w][tid];
      const int64_t p = (blk0 + i * bstride) * DP_PTS + tid;
      if (p < n) dsum[p] = sacc;
    }
  };
  load_b(0, b0h, b0l);
  // With an odd n_ks the sets swap roles from one block to the next.
  for (int64_t i = 0; i < nmine;) {
    block(i, b0h, b0l, b1h, b1l);
    if (++i >= nmine) break;
    if constexpr (KS_END % 2 == 0) continue;
    block(i, b1h, b1l, b0h, b0l);
    ++i;
  }
}

// ---------------------------------------------------------------------------------------------
// Balanced D pass for M = 416 (13 k-steps, 26 column groups). Column group cg is active in
// k-steps ks <= cg / 2 (its level), so the per-k-step work is triangular: 26 - 2 ks groups. With
// one barrier per k-step every wave waits for the busiest, and the round-robin kernel above
// issues sum_ks ceil((26 - 2 ks) / 7) = 32 group-steps per wave for 26 of useful work.
// Here k-steps j and 12 - j (j = 0..5) share one barrier (a super-step; k-step 6 is the last
// super-step alone), and the groups are dealt so that every wave does the same work in every
// super-step: wave m = 0..5 owns {2m, 2m+1} (level m) and {22-2m, 23-2m} (level 11-m), wave 6
// owns {24, 25} (level 12). A level pair (a, 11-a) is active in exactly one k-step of each pair
// (j, 12-j) and in k-step 6, so wave m runs 4 group-steps per super-step (2 in the last one),
// as does wave 6: 26 per wave in 7 barriers, with no zero-operand products.
// Ring position p of a block (the mover's order) holds k-step p / 2 for even p and
// 12 - (p - 1) / 2 for odd p, so each super-step's two k-steps are adjacent in the ring.
// Everything else (ring, mover, B prefetch, row sums) is as in k_dpass_split16.
// ---------------------------------------------------------------------------------------------
namespace b13 {
constexpr int NKS = 13, NCG = 26, NSS = 7;
__host__ __device__ constexpr int pos_ks(int p) { return (p & 1) ? 12 - (p - 1) / 2 : p / 2; }
__host__ __device__ constexpr int ss_first(int s) { return 2 * s; }
__host__ __device__ constexpr int ss_len(int s) { return s < 6 ? 2 : 1; }
__host__ __device__ constexpr int n_groups(int w) { return w < 6 ? 4 : 2; }
__host__ __device__ constexpr int group(int w, int c) {
  return w < 6 ? (c == 0 ? 2 * w : c == 1 ? 2 * w + 1 : c == 2 ? 22 - 2 * w : 23 - 2 * w)
               : 24 + c;
}
__host__ __device__ constexpr bool active(int w, int c, int ks) { return ks <= group(w, c) / 2; }
// Unit u of wave w in super-step s: (position e within the super-step, group slot c), in the
// order e, then c; n_units(w, s) of them.
__host__ __device__ constexpr int n_units(int w, int s) {
  int n = 0;
  for (int e = 0; e < ss_len(s); e++)
    for (int c = 0; c < n_groups(w); c++) n += active(w, c, pos_ks(ss_first(s) + e)) ? 1 : 0;
  return n;
}
__host__ __device__ constexpr int unit_e(int w, int s, int u) {
  for (int e = 0; e < ss_len(s); e++)
    for (int c = 0; c < n_groups(w); c++)
      if (active(w, c, pos_ks(ss_first(s) + e)) && u-- == 0) return e;
  return -1;
}
__host__ __device__ constexpr int unit_c(int w, int s, int u) {
  for (int e = 0; e < ss_len(s); e++)
    for (int c = 0; c < n_groups(w); c++)
      if (active(w, c, pos_ks(ss_first(s) + e)) && u-- == 0) return c;
  return -1;
}
constexpr bool check_balance() {
  int cover[NKS][NCG] = {};
  for (int w = 0; w < DP_CW; w++) {
    int total = 0;
    for (int s = 0; s < NSS; s++) {
      const int nu = n_units(w, s);
      if (nu != (s < 6 ? 4 : 2)) return false;
      total += nu;
      for (int u = 0; u < nu; u++) {
        const int ks = pos_ks(ss_first(s) + unit_e(w, s, u));
        cover[ks][group(w, unit_c(w, s, u))]++;
      }
    }
    if (total != 26) return false;
  }
  // Every active (k-step, group) pair exactly once, no inactive one.
  for (int ks = 0; ks < NKS; ks++)
    for (int cg = 0; cg < NCG; cg++)
      if (cover[ks][cg] != (cg >= 2 * ks ? 1 : 0)) return false;
  return true;
}
static_assert(DP_CW == 7 && check_balance(), "balanced M = 416 assignment");
}  // namespace b13

template <int W>
__device__ __forceinline__ void dpass_b13_wave(half8_t (*sa)[DP_SLOT], double (*red)[DP_PTS],
                                               int64_t n, int64_t nmine, int64_t blk0,
                                               int64_t bstride, const half8_t *__restrict__ bfrag,
                                               double *__restrict__ dsum) {
  using namespace b13;
  constexpr int NG = n_groups(W);
  const int tid = threadIdx.x, lane = tid & 63;
  float4_t acc[DP_TILES][NG];
  half8_t b0h[4], b0l[4], b1h[4], b1l[4];
  const half8_t *bfp = bfrag;
  auto load_b = [&](int s, half8_t *hh, half8_t *ll) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if (u >= n_units(W, s)) continue;
      const int ks = pos_ks(ss_first(s) + unit_e(W, s, u));
      const int64_t frag = (int64_t)ks * NCG + group(W, unit_c(W, s, u));
      hh[u] = bfp[(frag * 2 + 0) * 64 + lane];
      ll[u] = bfp[(frag * 2 + 1) * 64 + lane];
    }
  };
  // Super-step s on ring slots slot0 (position ss_first(s)) and slot0 + 1. The units are
  // ordered by position, so each position's A fragments are loaded once and used up before
  // the next position's.
  auto step = [&](int s, int slot0, const half8_t *bh, const half8_t *bl) {
#pragma unroll
    for (int r = 0; r < DP_TILES; r += 2) {
#pragma unroll
      for (int e = 0; e < 2; e++) {
        if (e >= ss_len(s)) continue;
        const int sl = (slot0 + e) & (DP_RING - 1);
        const half8_t ah0 = sa[sl][(r * 2 + 0) * 64 + lane];
        const half8_t al0 = sa[sl][(r * 2 + 1) * 64 + lane];
        const half8_t ah1 = sa[sl][(r * 2 + 2) * 64 + lane];
        const half8_t al1 = sa[sl][(r * 2 + 3) * 64 + lane];
#pragma unroll
        for (int pr = 0; pr < 3; pr++) {
#pragma unroll
          for (int u = 0; u < 4; u++) {
            if (u >= n_units(W, s) || unit_e(W, s, u) != e) continue;
            const int c = unit_c(W, s, u);
            const half8_t a0 = pr == 2 ? al0 : ah0, a1 = pr == 2 ? al1 : ah1;
            const half8_t b = pr == 1 ? bl[u] : bh[u];
            acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(b, a0, acc[r][c], 0, 0, 0);
            acc[r + 1][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(b, a1, acc[r + 1][c], 0, 0, 0);
          }
        }
      }
    }
  };
  // One block: super-steps 0..6 (even ones on set x, odd ones on set y), then the row sums.
  auto block = [&](int64_t i, half8_t *xh, half8_t *xl, half8_t *yh, half8_t *yl) {
#pragma unroll
    for (int r = 0; r < DP_TILES; r++)
#pragma unroll
      for (int c = 0; c < NG; c++) acc[r][c] = (float4_t){0.f, 0.f, 0.f, 0.f};
    const int pbase = (int)((i * NKS) & (DP_RING - 1));
    asm volatile("" : "+s"(bfp));
    const bool more = i + 1 < nmine;
#pragma unroll
    for (int s = 0; s < NSS; s += 2) {
      __syncthreads();
      if (s + 1 < NSS)
        load_b(s + 1, yh, yl);
      else if (more)
        load_b(0, yh, yl);
      step(s, pbase + ss_first(s), xh, xl);
      if (s + 1 >= NSS) break;
      __syncthreads();
      if (s + 2 < NSS)
        load_b(s + 2, xh, xl);
      else if (more)
        load_b(0, xh, xl);
      step(s + 1, pbase + ss_first(s + 1), yh, yl);
    }
    double part[DP_TILES];
#pragma unroll
    for (int r = 0; r < DP_TILES; r++) {
      double sacc = 0.0;
#pragma unroll
      for (int c = 0; c < NG; c++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const double y = (double)acc[r][c][e];
          sacc += y * y;
        }
      part[r] = sacc;
    }
    const bool up32 = (lane & 32) != 0, up16 = (lane & 16) != 0;
    double h4[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const double send = up32 ? part[q] : part[4 + q];
      const double keep = up32 ? part[4 + q] : part[q];
      h4[q] = keep + __shfl_xor(send, 32);
    }
    double h2[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const double send = up16 ? h4[q] : h4[2 + q];
      const double keep = up16 ? h4[2 + q] : h4[q];
      h2[q] = keep + __shfl_xor(send, 16);
    }
    const int rbase = 4 * (lane >> 5) + 2 * ((lane >> 4) & 1);
    red[W][(rbase + 0) * 16 + (lane & 15)] = h2[0];
    red[W][(rbase + 1) * 16 + (lane & 15)] = h2[1];
    __syncthreads();
    if (tid < DP_PTS) {
      double sacc = 0.0;
      for (int w = 0; w < DP_CW; w++) sacc += red[w][tid];
      const int64_t p = (blk0 + i * bstride) * DP_PTS + tid;
      if (p < n) dsum[p] = sacc;
    }
  };
  load_b(0, b0h, b0l);
  // NSS is odd, so the sets swap roles from one block to the next.
  for (int64_t i = 0; i < nmine;) {
    block(i, b0h, b0l, b1h, b1l);
    if (++i >= nmine) break;
    block(i, b1h, b1l, b0h, b0l);
    ++i;
  }
}

__global__ void __launch_bounds__(64 * DP_WAVES) k_dpass_split16_b13(const half8_t *__restrict__ xd,
                                                                     int64_t n, int64_t nblk,
                                                                     const half8_t *__restrict__ bfrag,
                                                                     double *__restrict__ dsum) {
  using namespace b13;
  __shared__ half8_t sa[DP_RING][DP_SLOT];
  __shared__ double red[DP_CW][DP_PTS];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int64_t blk0 = blockIdx.x, bstride = gridDim.x;
  const int64_t nmine = blk0 < nblk ? (nblk - 1 - blk0) / bstride + 1 : 0;
  switch (wave) {
    case 0: dpass_b13_wave<0>(sa, red, n, nmine, blk0, bstride, bfrag, dsum); return;
    case 1: dpass_b13_wave<1>(sa, red, n, nmine, blk0, bstride, bfrag, dsum); return;
    case 2: dpass_b13_wave<2>(sa, red, n, nmine, blk0, bstride, bfrag, dsum); return;
    case 3: dpass_b13_wave<3>(sa, red, n, nmine, blk0, bstride, bfrag, dsum); return;
    case 4: dpass_b13_wave<4>(sa, red, n, nmine, blk0, bstride, bfrag, dsum); return;
    case 5: dpass_b13_wave<5>(sa, red, n, nmine, blk0, bstride, bfrag, dsum); return;
    case 6: dpass_b13_wave<6>(sa, red, n, nmine, blk0, bstride, bfrag, dsum); return;
    default: break;
  }
  // Data mover: global position g is ring position g % 13 of this workgroup's block g / 13.
  // The barrier that opens a super-step comes after its positions' copies have landed; after
  // it, the slots of the previous super-step are free, so the ring is refilled up to the
  // super-step's first position + DP_RING - 1.
  const int64_t total = nmine * NKS;
  auto issue = [&](int64_t g) {
    const int64_t blk = blk0 + (g / NKS) * bstride;
    const int ks = pos_ks((int)(g % NKS));
    half8_t *slot = sa[(int)(g & (DP_RING - 1))];
    const half8_t *src = xd + (blk * DP_TILES * NKS * 2) * 64 + lane;
#pragma unroll
    for (int b = 0; b < DP_LOADS; b++) {
      const int rg = b >> 1, pl = b & 1;
      __builtin_amdgcn_global_load_lds(
          (const void *)(src + (((int64_t)rg * NKS + ks) * 2 + pl) * 64),
          (__attribute__((address_space(3))) void *)(slot + b * 64), 16, 0, 0);
    }
  };
  int64_t issued = 0;
  for (; issued < DP_RING && issued < total; issued++) issue(issued);
  for (int64_t i = 0; i < nmine; i++) {
    for (int s = 0; s < NSS; s++) {
      const int64_t g0 = i * NKS + ss_first(s), g1 = g0 + ss_len(s) - 1;
      dp_mover_wait(issued - 1 - g1);
      __builtin_amdgcn_s_barrier();
      for (; issued < total && issued <= g0 + DP_RING - 1; issued++) issue(issued);
    }
    __builtin_amdgcn_s_barrier();  // the block's row reduction
  }
}

// Column ranges of the D pass: all M / 16 groups in one range up to M = 448; beyond, ranges of
// DP_CW * DP_MAXCG groups taken from the last group down (the top range needs every k-step, the
// lower ones fewer), so that the k-steps read over all ranges stay few.
std::vector<std::pair<int, int>> dpass_split16_ranges(int64_t m_cols) {
  const int n_cg = (int)(m_cols / 16), cap = DP_CW * DP_MAXCG;
  std::vector<std::pair<int, int>> r;
  if (n_cg <= cap) {
    r.push_back({0, n_cg});
    return r;
  }
  for (int hi = n_cg; hi > 0; hi -= cap) r.push_back({std::max(0, hi - cap), std::min(cap, hi)});
  return r;
}

// dsum[p] = sum over the ranges, in range order, of the range partials (parts[r * n + p]).
__global__ void k_dsum_ranges(const double *__restrict__ parts, int n_ranges, int64_t n,
                              double *__restrict__ dsum) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double s = 0.0;
  for (int r = 0; r < n_ranges; r++) s += parts[r * n + p];
  dsum[p] = s;
}

// n_cu: the device's compute units; the grid is one workgroup per CU (at most one block each).
// cglist: DP_CW x DP_MAXCG group slots per range; dsum_parts: n doubles per range (several
// ranges only).
void launch_dpass_split16(hipStream_t st, const half8_t *xd, int64_t n, int64_t m_cols,
                          const half8_t *bfrag, const int16_t *cglist, int n_cu, double *dsum,
                          double *dsum_parts) {
  if (n <= 0) return;
  const int64_t nblk = (n + DP_PTS - 1) / DP_PTS;
  const int64_t wgs = std::min<int64_t>(nblk, std::max(n_cu, 1));
  const dim3 grid((unsigned)wgs), block(64 * DP_WAVES);
  const int nks = (int)(m_cols / 32);
  if (m_cols <= 32 * 14) {
    switch (nks) {
#define DP_CASE(k)                                                                      \
  case k:                                                                               \
    hipLaunchKernelGGL((k_dpass_split16<k, k, 0, 2 * k>), grid, block, 0, st, xd, n, nblk, \
                       bfrag, cglist, dsum);                                            \
    break;
      DP_CASE(1) DP_CASE(2) DP_CASE(3) DP_CASE(4) DP_CASE(5) DP_CASE(6) DP_CASE(7)
      DP_CASE(8) DP_CASE(9) DP_CASE(10) DP_CASE(11) DP_CASE(12) DP_CASE(14)
      case 13:
        hipLaunchKernelGGL(k_dpass_split16_b13, grid, block, 0, st, xd, n, nblk, bfrag, dsum);
        break;
#undef DP_CASE
      default: break;
    }
    return;
  }
  // Several ranges: the quotient bases of k = 7 and 8 (2N = 630, 920) and the raw k = 8 basis
  // (2N = 990).
  const std::vector<std::pair<int, int>> rr = dpass_split16_ranges(m_cols);
  for (size_t r = 0; r < rr.size(); r++) {
    const int16_t *cl = cglist + r * DP_CW * DP_MAXCG;
    double *out = dsum_parts + r * n;
    const int key = nks * 10000 + rr[r].first * 100 + rr[r].second;
    switch (key) {
#define DP_RANGE(k, c0, nc)                                                                   \
  case k * 10000 + c0 * 100 + nc:                                                             \
    hipLaunchKernelGGL((k_dpass_split16<k, (c0 + nc - 1) / 2 + 1, c0, nc>), grid, block, 0, st, \
                       xd, n, nblk, bfrag, cl, out);                                          \
    break;
      DP_RANGE(20, 12, 28) DP_RANGE(20, 0, 12)
      DP_RANGE(29, 30, 28) DP_RANGE(29, 2, 28) DP_RANGE(29, 0, 2)
      DP_RANGE(31, 34, 28) DP_RANGE(31, 6, 28) DP_RANGE(31, 0, 6)
#undef DP_RANGE
      default: break;  // rejected by the host (dpass_split16_supported)
    }
  }
  hipLaunchKernelGGL(k_dsum_ranges, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     dsum_parts, (int)rr.size(), n, dsum);
}

// The column counts the D pass is built for: every M up to 448, and 640, 928, 992.
bool dpass_split16_supported(int64_t m_cols) {
  return m_cols <= 448 || m_cols == 640 || m_cols == 928 || m_cols == 992;
}

// ---------------------------------------------------------------------------------------------
// Per-point coefficients of the split path. D_p = dsum_p * bscale / rs2_p restores the unscaled
// s_p^H G s_p. The weight on the scaled row is c'_p = w_p / (rs2_p * max(D_p, 1e-30)), for valid
// w (finite, > 0), else 0: the same T as the reference's w_p / max(K_p, 1e-30) on the unscaled
// row (quintic.rs:1761-1775). u_p = sqrt(c'_p) (FP32) scales the T-pass operand z_p = 2^g u_p x_p;
// the largest u sets the global 2^g (per-block maxima, then one workgroup; no atomics).
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_coeffs_split16(const double *__restrict__ w,
                                                        const double *__restrict__ dsum,
                                                        const float *__restrict__ rs2,
                                                        const double *__restrict__ bscale,
                                                        int64_t n, float *__restrict__ u,
                                                        float *__restrict__ block_max) {
  __shared__ float sh[4];
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
  float mv = uv;
  for (int off = 32; off > 0; off >>= 1) mv = fmaxf(mv, __shfl_xor(mv, off));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mv;
  __syncthreads();
  if (threadIdx.x == 0) block_max[blockIdx.x] = fmaxf(fmaxf(sh[0], sh[1]), fmaxf(sh[2], sh[3]));
}

// FP64 T-pass coefficients from the split D pass (the refinement sweep of CY3Q_REFINE_T_FP64):
// c_p = w_p / max(D_p, 1e-30) with D_p = dsum_p * bscale / rs2_p, for valid w, else 0
// (quintic.rs:1757, 1776-1781).
__global__ void k_coeffs_f64_from_split(const double *__restrict__ w,
                                        const double *__restrict__ dsum,
                                        const float *__restrict__ rs2,
                                        const double *__restrict__ bscale, int64_t n,
                                        double *__restrict__ c) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double wp = w[p];
  double cv = 0.0;
  if (isfinite(wp) && wp > 0.0) {
    const double d = dsum[p] * bscale[0] / (double)rs2[p];
    cv = (1.0 / fmax(d, 1e-30)) * wp;
  }
  c[p] = cv;
}

void launch_coeffs_f64_from_split(hipStream_t st, const double *w, const double *dsum,
                                  const float *rs2, const double *bscale, int64_t n, double *c) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_coeffs_f64_from_split, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     w, dsum, rs2, bscale, n, c);
}

// Max over the block maxima -> zscale[0] = 2^g (float, applied to z) and tscale[0] = 4^-g
// (double, applied to T), with g chosen so that max u 2^g < 2^14; the row-scaled |x| < 1, so
// every z stays below 2^14.
__global__ void __launch_bounds__(1024) k_finalize_scale(const float *__restrict__ block_max,
                                                         int64_t nblk, float *__restrict__ zscale,
                                                         double *__restrict__ tscale) {
  __shared__ float sh[1024];
  float mv = 0.f;
  for (int64_t i = threadIdx.x; i < nblk; i += 1024) mv = fmaxf(mv, block_max[i]);
  sh[threadIdx.x] = mv;
  __syncthreads();
  for (int off = 512; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    int e = 0;
    if (sh[0] > 0.f) frexpf(sh[0], &e);  // max u < 2^e
    const int g = 14 - e;
    zscale[0] = ldexpf(1.f, g);
    tscale[0] = ldexp(1.0, -2 * g);
  }
}

// u <- 2^g u, in place (exact: a power of two).
__global__ void __launch_bounds__(256) k_apply_zscale(float *__restrict__ u,
                                                      const float *__restrict__ zscale,
                                                      int64_t n) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) u[p] *= zscale[0];
}

void launch_coeffs_split16(hipStream_t st, const double *w, const double *dsum, const float *rs2,
                           const double *bscale, int64_t n, float *u, float *block_max,
                           float *zscale, double *tscale) {
  const int64_t nblk = (n + 255) / 256;
  if (n > 0)
    hipLaunchKernelGGL(k_coeffs_split16, dim3((unsigned)nblk), dim3(256), 0, st, w, dsum, rs2,
                       bscale, n, u, block_max);
  hipLaunchKernelGGL(k_finalize_scale, dim3(1), dim3(1024), 0, st, block_max, nblk, zscale,
                     tscale);
  if (n > 0)
    hipLaunchKernelGGL(k_apply_zscale, dim3((unsigned)nblk), dim3(256), 0, st, u, zscale, n);
}

}  // namespace cy3q
