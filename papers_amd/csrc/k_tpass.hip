// T pass of the split-precision Donaldson step (reference contraction: quintic.rs:1777-1789):
// T = sum_p c_p x_p x_p^T = Z^T Z over the real-ified section rows, z_p = 2^g sqrt(c'_p) x_p, on
// the fp16 MFMA (v_mfma_f32_32x32x16_f16) with FP32 accumulation within a point chunk.
//
// Staging. One workgroup (8 waves) per CU streams one point chunk, 16 points per stage. Thread c
// owns column c: it loads the column's 16 FP32 values of the next stage (four 16-byte loads,
// each contiguous across the wave's columns) and the stage's 16 coefficients while the current
// stage's MFMAs run, then scales by 2^g sqrt(c'), splits z into
// fp16 hi/lo and writes the transposed tile zT[plane][column][16 points] to the other LDS
// buffer. Two waves share each SIMD, so one wave's conversion overlaps the other's MFMAs.
//
// Tiles. The lower triangle of 32 x 32 tiles (NT = M / 32 indices) is split into per-wave
// programs of sub-blocks (rectangles rows x cols, or lower triangles), fixed at compile time.
// A wave holds a sub-block's row fragments and streams its column fragments, so each fragment
// is read from LDS once per sub-block and stage. For NT = 13 (k = 6, N = 205) the 91 tiles go
// as four 3 x 4 rectangles, one 4-triangle, one 3 x 3 rectangle, a 3 x 2 rectangle plus a
// 3-triangle and two 3-triangles (at most 12 tiles per wave, 24 per SIMD). Other sizes deal
// single tiles round-robin. A product is z_hi z_hi + z_hi z_lo + z_lo z_hi.
//
// Output: FP32 chunk partials [chunk][tile][1024] in MFMA output order; the reduction sums them
// over the chunks in FP64, in chunk order (deterministic).
#include "cy3q_internal.h"

namespace cy3q {

static constexpr int TQ_WAVES = 8;
static constexpr int TQ_PTS = 16;   // points per stage (the MFMA k dimension)
static constexpr int TQ_LD = 24;    // halves per column row in LDS (16 points + 8 padding)
static constexpr int TQ_MAXT = 12;  // accumulator tiles per wave
static constexpr int TQ_MAXNT = 13; // largest tile index count (M <= 416)

struct TqSub {
  int r0, nr, c0, nc, tri;
};
struct TqProg {
  int n;
  TqSub s[TQ_MAXT];
};

__host__ __device__ constexpr int tq_sub_tiles(const TqSub &b) {
  return b.nr <= 0 ? 0 : (b.tri ? b.nr * (b.nr + 1) / 2 : b.nr * b.nc);
}

// Wave w's program for NT tile indices.
__host__ __device__ constexpr TqProg tq_prog(int nt, int w) {
  TqProg p{0, {}};
  if (nt == 13) {
    constexpr TqSub tab[TQ_WAVES][2] = {
        {{4, 3, 0, 4, 0}, {0, 0, 0, 0, 0}},  {{7, 3, 0, 4, 0}, {0, 0, 0, 0, 0}},
        {{10, 3, 0, 4, 0}, {0, 0, 0, 0, 0}}, {{10, 3, 4, 4, 0}, {0, 0, 0, 0, 0}},
        {{0, 4, 0, 4, 1}, {0, 0, 0, 0, 0}},  {{7, 3, 4, 3, 0}, {0, 0, 0, 0, 0}},
        {{10, 3, 8, 2, 0}, {10, 3, 10, 3, 1}}, {{4, 3, 4, 3, 1}, {7, 3, 7, 3, 1}}};
    for (int k = 0; k < 2; k++)
      if (tab[w][k].nr > 0) p.s[p.n++] = tab[w][k];
    return p;
  }
  // Round-robin single tiles in lower-triangle order.
  int t = 0;
  for (int i = 0; i < nt; i++)
    for (int j = 0; j <= i; j++, t++)
      if (t % TQ_WAVES == w && p.n < TQ_MAXT) p.s[p.n++] = TqSub{i, 1, j, 1, i == j ? 1 : 0};
  return p;
}

__host__ __device__ constexpr int tq_prog_tiles(int nt, int w) {
  int s = 0;
  const TqProg p = tq_prog(nt, w);
  for (int k = 0; k < p.n; k++) s += tq_sub_tiles(p.s[k]);
  return s;
}

__host__ __device__ constexpr int tq_sub_off(int nt, int w, int k) {
  int s = 0;
  const TqProg p = tq_prog(nt, w);
  for (int q = 0; q < k; q++) s += tq_sub_tiles(p.s[q]);
  return s;
}

// Every tile of the lower triangle is in exactly one wave program, and no program holds more
// than TQ_MAXT tiles (checked at compile time for each instantiated size).
__host__ __device__ constexpr bool tq_progs_valid(int nt) {
  int cover[TQ_MAXNT * TQ_MAXNT] = {};
  for (int w = 0; w < TQ_WAVES; w++) {
    const TqProg p = tq_prog(nt, w);
    if (tq_prog_tiles(nt, w) > TQ_MAXT) return false;
    for (int k = 0; k < p.n; k++) {
      const TqSub b = p.s[k];
      if (b.tri && (b.r0 != b.c0 || b.nr != b.nc)) return false;
      for (int i = 0; i < b.nr; i++)
        for (int j = 0; j < b.nc; j++) {
          if (b.tri && j > i) continue;
          const int I = b.r0 + i, J = b.c0 + j;
          if (I >= nt || J > I) return false;
          cover[I * TQ_MAXNT + J]++;
        }
    }
  }
  for (int i = 0; i < nt; i++)
    for (int j = 0; j <= i; j++)
      if (cover[i * TQ_MAXNT + j] != 1) return false;
  return true;
}

// One sub-block for one stage: rows R0.. x cols C0.. (TRI: lower triangle, the row fragments
// serve as the column fragments). hi/lo: this lane's position in the two planes of the stage.
template <int R0, int NR, int C0, int NC, int TRI, int OFF>
__device__ __forceinline__ void tq_sub(const _Float16 *hi, const _Float16 *lo,
                                       float16_t (&acc)[TQ_MAXT]) {
  half8_t ah[NR], al[NR];
#pragma unroll
  for (int i = 0; i < NR; i++) {
    ah[i] = *reinterpret_cast<const half8_t *>(hi + (R0 + i) * 32 * TQ_LD);
    al[i] = *reinterpret_cast<const half8_t *>(lo + (R0 + i) * 32 * TQ_LD);
  }
#pragma unroll
  for (int j = 0; j < NC; j++) {
    half8_t bh, bl;
    if constexpr (TRI) {
      bh = ah[j];
      bl = al[j];
    } else {
      bh = *reinterpret_cast<const half8_t *>(hi + (C0 + j) * 32 * TQ_LD);
      bl = *reinterpret_cast<const half8_t *>(lo + (C0 + j) * 32 * TQ_LD);
    }
#pragma unroll
    for (int pr = 0; pr < 3; pr++) {
#pragma unroll
      for (int i = 0; i < NR; i++) {
        if (TRI && i < j) continue;
        const int t = OFF + (TRI ? i * (i + 1) / 2 + j : i * NC + j);
        const half8_t a = pr == 2 ? al[i] : ah[i];
        const half8_t b = pr == 1 ? bl : bh;
        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, acc[t], 0, 0, 0);
      }
    }
  }
}

template <int NT, int W, int K>
__device__ __forceinline__ void tq_run(const _Float16 *hi, const _Float16 *lo,
                                       float16_t (&acc)[TQ_MAXT]) {
  constexpr TqProg P = tq_prog(NT, W);
  if constexpr (K < P.n) {
    constexpr TqSub b = P.s[K];
    tq_sub<b.r0, b.nr, b.c0, b.nc, b.tri, tq_sub_off(NT, W, K)>(hi, lo, acc);
    tq_run<NT, W, K + 1>(hi, lo, acc);
  }
}

template <int NT, int W, int K>
__device__ __forceinline__ void tq_store(float *__restrict__ out, const float16_t (&acc)[TQ_MAXT],
                                         int lane) {
  constexpr TqProg P = tq_prog(NT, W);
  if constexpr (K < P.n) {
    constexpr TqSub b = P.s[K];
    constexpr int OFF = tq_sub_off(NT, W, K);
#pragma unroll
    for (int i = 0; i < b.nr; i++) {
#pragma unroll
      for (int j = 0; j < b.nc; j++) {
        if (b.tri && j > i) continue;
        const int t = OFF + (b.tri ? i * (i + 1) / 2 + j : i * b.nc + j);
        const int I = b.r0 + i, J = b.c0 + j;
        float *o = out + (int64_t)(I * (I + 1) / 2 + J) * 1024;
#pragma unroll
        for (int e = 0; e < 16; e++) o[e * 64 + lane] = acc[t][e];
      }
    }
    tq_store<NT, W, K + 1>(out, acc, lane);
  }
}

// LDS ring of raw stage records: TQ_RING slots, each one stage's FP32 record (4 quads x M
// columns x 4 points, the global layout) and its 16 coefficients (replicated over 64 floats).
static constexpr int TQ_RING = 3;

// s_waitcnt with vmcnt = v (0..63), lgkmcnt = 0 and expcnt at its maximum (gfx9 encoding:
// vmcnt[3:0] in bits 3:0, vmcnt[5:4] in bits 15:14, expcnt in 6:4, lgkmcnt in 11:8).
#define TQ_WAIT_VM_LGKM0(v) __builtin_amdgcn_s_waitcnt(((v)&15) | (((v) >> 4) << 14) | (7 << 4))

template <int NT, int W>
__device__ __forceinline__ void tq_wave(_Float16 *zt, float4_t *ring, float *uring,
                                        const float4_t *__restrict__ xt,
                                        const float *__restrict__ u, int64_t p_begin,
                                        int64_t p_end, float *__restrict__ out) {
  constexpr int M = 32 * NT;
  constexpr int64_t PLANE = (int64_t)M * TQ_LD;
  // Wave TQ_WAVES - 1 holds no column (M <= 416 < 448) and doubles as the data mover: per stage
  // it issues G 1 KB copies of the record and one of the coefficients into the ring.
  constexpr bool MOVER = W == TQ_WAVES - 1;
  constexpr int G = 2 * NT;               // 1 KB granules per stage record (64 M bytes)
  constexpr int COPIES = G + 1;           // copies per stage
  static_assert(M <= 64 * (TQ_WAVES - 1), "the mover wave must hold no column");
  static_assert(2 * COPIES <= 63, "ring waits must fit the vmcnt field");
  const int tid = threadIdx.x, lane = tid & 63;
  float16_t acc[TQ_MAXT];
#pragma unroll
  for (int t = 0; t < TQ_MAXT; t++)
#pragma unroll
    for (int e = 0; e < 16; e++) acc[t][e] = 0.f;

  const int n_stage = (int)((p_end - p_begin + TQ_PTS - 1) / TQ_PTS);
  const int64_t stage0 = p_begin / TQ_PTS;
  const int col = tid;

  // Mover: stage t's record and coefficients into ring slot t % TQ_RING. Each lane copies 16
  // bytes per granule (lane l to byte 16 l of the granule), and 4 bytes of coefficients (lane l
  // gets coefficient l % 16).
  auto issue = [&](int t) {
    const char *src = reinterpret_cast<const char *>(xt + (stage0 + t) * 4 * M) + 16 * lane;
    char *dst = reinterpret_cast<char *>(ring + (t % TQ_RING) * 4 * M);
#pragma unroll
    for (int g = 0; g < G; g++)
      __builtin_amdgcn_global_load_lds((const void *)(src + 1024 * g),
                                       (__attribute__((address_space(3))) void *)(dst + 1024 * g),
                                       16, 0, 0);
    __builtin_amdgcn_global_load_lds(
        (const void *)(u + p_begin + (int64_t)t * TQ_PTS + (lane & (TQ_PTS - 1))),
        (__attribute__((address_space(3))) void *)(uring + (t % TQ_RING) * 64), 4, 0, 0);
  };
  // Mover: wait until the copies of every stage up to `need` have landed, given that stages up
  // to `last` have been issued (copies complete in issue order for the counter), and the
  // wave's own LDS reads too; then a bare barrier (a workgroup fence would wait for every copy
  // in flight). The empty asm statements keep the compiler from moving LDS accesses across.
  auto mover_sync = [&](int need, int last) {
    asm volatile("" ::: "memory");
    const int later = last - need;
    if (later >= 2)
      TQ_WAIT_VM_LGKM0(2 * COPIES);
    else if (later == 1)
      TQ_WAIT_VM_LGKM0(COPIES);
    else
      TQ_WAIT_VM_LGKM0(0);
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
  };
  // Converting waves: stage t from its ring slot, z = u x split into fp16 hi/lo, into zt[buf].
  auto convert = [&](int t, int buf) {
    _Float16 *hi = zt + (int64_t)buf * 2 * PLANE, *lo = hi + PLANE;
    const float4_t *rs = ring + (t % TQ_RING) * 4 * M;
    const float *us = uring + (t % TQ_RING) * 64;
    if (col < M) {
      float4_t rec[4];
#pragma unroll
      for (int q = 0; q < 4; q++) rec[q] = rs[q * M + col];
#pragma unroll
      for (int hf = 0; hf < 2; hf++) {
        half8_t zh, zl;
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const int k = 8 * hf + j;
          const float v = rec[k / 4][k % 4] * us[k];
          const _Float16 h = (_Float16)v;
          zh[j] = h;
          zl[j] = (_Float16)(v - (float)h);
        }
        *reinterpret_cast<half8_t *>(hi + col * TQ_LD + 8 * hf) = zh;
        *reinterpret_cast<half8_t *>(lo + col * TQ_LD + 8 * hf) = zl;
      }
    }
  };

  // Prologue: stages 0 .. TQ_RING - 1 in flight; stage 0 converted into zt[0]; stage 1 landed.
  const int first = n_stage < TQ_RING ? n_stage : TQ_RING;
  if constexpr (MOVER) {
    for (int t = 0; t < first; t++) issue(t);
    mover_sync(0, first - 1);
  } else {
    __syncthreads();
    if (n_stage > 0) convert(0, 0);
  }
  if constexpr (MOVER)
    mover_sync(1, first - 1);
  else
    __syncthreads();
  // Stage s: the mover first refills the slot of stage s (converted in stage s - 1) with stage
  // s + TQ_RING; all waves run their MFMAs on zt[s % 2]; the converting waves fill zt[1 - s % 2]
  // from stage s + 1 (landed before this stage's barrier, and zt[1 - s % 2] was last read
  // before it); the mover then waits for stage s + 2, which the next stage converts.
  // Waves w and w + 4 share a SIMD. Waves 4..6 convert before their MFMAs and waves 0..3
  // after, so one wave's conversion runs under its partner's MFMAs.
  constexpr bool CONVERT_FIRST = W >= TQ_WAVES / 2 && !MOVER;
  // MFMA operand position of this lane: column l % 32 of a tile, points 8 (l / 32) .. + 7.
  const int lofs = (lane & 31) * TQ_LD + 8 * (lane >> 5);
  for (int s = 0; s < n_stage; s++) {
    const int buf = s & 1;
    if constexpr (MOVER) {
      if (s + TQ_RING < n_stage) issue(s + TQ_RING);
    }
    if constexpr (CONVERT_FIRST) {
      if (s + 1 < n_stage) convert(s + 1, buf ^ 1);
    }
    const _Float16 *hi = zt + (int64_t)buf * 2 * PLANE + lofs, *lo = hi + PLANE;
    tq_run<NT, W, 0>(hi, lo, acc);
    if constexpr (MOVER) {
      const int last = s + TQ_RING < n_stage ? s + TQ_RING : n_stage - 1;
      mover_sync(s + 2 < last ? s + 2 : last, last);
    } else {
      if constexpr (!CONVERT_FIRST) {
        if (s + 1 < n_stage) convert(s + 1, buf ^ 1);
      }
      __syncthreads();
    }
  }
  tq_store<NT, W, 0>(out, acc, lane);
}

// u: 2^g sqrt(c'_p), zero-padded to a multiple of 16 points past n.
template <int NT>
__global__ void __launch_bounds__(64 * TQ_WAVES)
k_tpass(const float4_t *__restrict__ xt, const float *__restrict__ u, int64_t n, int64_t chunk,
        float *__restrict__ partials) {
  static_assert(tq_progs_valid(NT), "T-pass wave programs must cover the lower triangle once");
  constexpr int M = 32 * NT;
  __shared__ __attribute__((aligned(16))) _Float16 zt[2 * 2 * M * TQ_LD];
  __shared__ float4_t ring[TQ_RING * 4 * M];
  __shared__ float uring[TQ_RING * 64];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t chunk_id = blockIdx.x;
  const int64_t p_begin = chunk_id * chunk;
  const int64_t p_end = p_begin + chunk < n ? p_begin + chunk : n;
  float *out = partials + chunk_id * (int64_t)(NT * (NT + 1) / 2) * 1024;
  switch (wave) {
    case 0: tq_wave<NT, 0>(zt, ring, uring, xt, u, p_begin, p_end, out); break;
    case 1: tq_wave<NT, 1>(zt, ring, uring, xt, u, p_begin, p_end, out); break;
    case 2: tq_wave<NT, 2>(zt, ring, uring, xt, u, p_begin, p_end, out); break;
    case 3: tq_wave<NT, 3>(zt, ring, uring, xt, u, p_begin, p_end, out); break;
    case 4: tq_wave<NT, 4>(zt, ring, uring, xt, u, p_begin, p_end, out); break;
    case 5: tq_wave<NT, 5>(zt, ring, uring, xt, u, p_begin, p_end, out); break;
    case 6: tq_wave<NT, 6>(zt, ring, uring, xt, u, p_begin, p_end, out); break;
    default: tq_wave<NT, 7>(zt, ring, uring, xt, u, p_begin, p_end, out); break;
  }
}

// FP64 sums of the chunk partials in two passes: pass 1 sums TQ_SEG consecutive chunks per
// segment, pass 2 sums the segments in order, applies the scale of the T pass and scatters to
// the full M x M matrix (the lower element of each diagonal tile, mirrored).
static constexpr int TQ_SEG = 32;

__global__ void k_tpass_seg(const float *__restrict__ partials, int64_t per_chunk,
                            int64_t n_chunks, double *__restrict__ seg) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= per_chunk) return;
  const int64_t c0 = (int64_t)blockIdx.y * TQ_SEG;
  const int64_t c1 = c0 + TQ_SEG < n_chunks ? c0 + TQ_SEG : n_chunks;
  double s = 0.0;
  for (int64_t c = c0; c < c1; c++) s += (double)partials[c * per_chunk + idx];
  seg[(int64_t)blockIdx.y * per_chunk + idx] = s;
}

// 32x32 MFMA output: element e of lane l is (row 8 (e / 4) + 4 (l / 32) + e % 4, column l % 32).
__global__ void k_tpass_final(const double *__restrict__ seg, int64_t per_chunk, int n_seg,
                              int64_t m, double *__restrict__ xx) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= per_chunk) return;
  double s = 0.0;
  for (int g = 0; g < n_seg; g++) s += seg[(int64_t)g * per_chunk + idx];
  const int t = (int)(idx / 1024), el = (int)(idx % 1024);
  int I = 0;
  while ((I + 1) * (I + 2) / 2 <= t) I++;
  const int J = t - I * (I + 1) / 2;
  const int e = el / 64, ln = el % 64;
  const int r = 8 * (e / 4) + 4 * (ln / 32) + (e % 4), c = ln % 32;
  if (I == J && r < c) return;
  const int64_t gi = 32 * (int64_t)I + r, gj = 32 * (int64_t)J + c;
  xx[gi * m + gj] = s;
  xx[gj * m + gi] = s;
}

int tpass_max_cols() { return 32 * TQ_MAXNT; }
size_t tpass_partials(int64_t m_cols, int64_t n_chunks) {
  const int64_t nt = m_cols / 32;
  return (size_t)(nt * (nt + 1) / 2) * 1024 * (size_t)n_chunks;
}
size_t tpass_seg_count(int64_t m_cols, int64_t n_chunks) {
  const int64_t nt = m_cols / 32;
  return (size_t)(nt * (nt + 1) / 2) * 1024 * (size_t)((n_chunks + TQ_SEG - 1) / TQ_SEG);
}

void launch_tpass(hipStream_t st, const float *xt, const float *u, int64_t n, int64_t m_cols,
                  int64_t chunk, int64_t n_chunks, float *partials) {
  const dim3 grid((unsigned)n_chunks), block(64 * TQ_WAVES);
  switch (m_cols / 32) {
#define TQ_CASE(k)                                                                           \
  case k:                                                                                    \
    hipLaunchKernelGGL(k_tpass<k>, grid, block, 0, st, (const float4_t *)xt, u, n, chunk, partials); \
    break;
    TQ_CASE(1) TQ_CASE(2) TQ_CASE(3) TQ_CASE(4) TQ_CASE(5) TQ_CASE(6) TQ_CASE(7)
    TQ_CASE(8) TQ_CASE(9) TQ_CASE(10) TQ_CASE(11) TQ_CASE(12) TQ_CASE(13)
#undef TQ_CASE
    default: break;  // rejected by the host (M <= 416)
  }
}

void launch_tpass_reduce(hipStream_t st, const float *partials, int64_t m_cols, int64_t n_chunks,
                         double *seg, double *xx) {
  const int64_t nt = m_cols / 32, per_chunk = nt * (nt + 1) / 2 * 1024;
  const int n_seg = (int)((n_chunks + TQ_SEG - 1) / TQ_SEG);
  hipLaunchKernelGGL(k_tpass_seg, dim3((unsigned)((per_chunk + 255) / 256), (unsigned)n_seg),
                     dim3(256), 0, st, partials, per_chunk, n_chunks, seg);
  hipLaunchKernelGGL(k_tpass_final, dim3((unsigned)((per_chunk + 255) / 256)), dim3(256), 0, st,
                     seg, per_chunk, n_seg, m_cols, xx);
}

}  // namespace cy3q
