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
#include <algorithm>
#include <vector>

#include "cy3q_internal.h"

namespace cy3q {

static constexpr int TQ_WAVES = 8;
static constexpr int TQ_PTS = 16;   // points per stage (the MFMA k dimension)
static constexpr int TQ_LD = 24;    // halves per column row in LDS (16 points + 8 padding)
static constexpr int TQ_MAXT = 12;  // accumulator tiles per wave
static constexpr int TQ_MAXNT = 13; // largest tile index count of one part (416 columns)

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
__host__ __device__ constexpr TqProg tq_prog_tri(int nt, int w) {
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

// Rectangle part: the NB x NA tiles of T_ba = Z_b^T Z_a for two column segments a < b. The
// staged columns are segment a's (local blocks 0 .. NA - 1, the tile columns) then segment b's
// (local blocks NA .. NA + NB - 1, the tile rows). Sub-blocks of at most 2 rows x 3 columns,
// in row-major order, dealt round-robin to the waves.
__host__ __device__ constexpr TqProg tq_prog_rect(int na, int nb, int w) {
  TqProg p{0, {}};
  int k = 0, tiles = 0;
  for (int r = 0; r < nb; r += 2)
    for (int c = 0; c < na; c += 3, k++) {
      const int nr = nb - r < 2 ? nb - r : 2, nc = na - c < 3 ? na - c : 3;
      if (k % TQ_WAVES == w && tiles + nr * nc <= TQ_MAXT) {
        p.s[p.n++] = TqSub{na + r, nr, c, nc, 0};
        tiles += nr * nc;
      }
    }
  return p;
}

// Part key: KEY = NA + 64 NB; NB = 0 is the lower triangle of NA tile indices.
__host__ __device__ constexpr TqProg tq_prog(int key, int w) {
  return key < 64 ? tq_prog_tri(key, w) : tq_prog_rect(key % 64, key / 64, w);
}

__host__ __device__ constexpr int tq_prog_tiles(int key, int w) {
  int s = 0;
  const TqProg p = tq_prog(key, w);
  for (int k = 0; k < p.n; k++) s += tq_sub_tiles(p.s[k]);
  return s;
}

__host__ __device__ constexpr int tq_sub_off(int key, int w, int k) {
  int s = 0;
  const TqProg p = tq_prog(key, w);
  for (int q = 0; q < k; q++) s += tq_sub_tiles(p.s[q]);
  return s;
}

// Every tile of the part is in exactly one wave program, and no program holds more than TQ_MAXT
// tiles (checked at compile time for each instantiated part).
__host__ __device__ constexpr bool tq_progs_valid(int key) {
  const int na = key % 64, nb = key / 64;
  const int rows = nb > 0 ? nb : na;
  int cover[TQ_MAXNT * TQ_MAXNT] = {};
  for (int w = 0; w < TQ_WAVES; w++) {
    const TqProg p = tq_prog(key, w);
    if (tq_prog_tiles(key, w) > TQ_MAXT) return false;
    for (int k = 0; k < p.n; k++) {
      const TqSub b = p.s[k];
      if (b.tri && (b.r0 != b.c0 || b.nr != b.nc || nb > 0)) return false;
      for (int i = 0; i < b.nr; i++)
        for (int j = 0; j < b.nc; j++) {
          if (b.tri && j > i) continue;
          const int I = b.r0 + i - (nb > 0 ? na : 0), J = b.c0 + j;
          if (I < 0 || I >= rows || J >= na || (nb == 0 && J > I)) return false;
          cover[I * TQ_MAXNT + J]++;
        }
    }
  }
  for (int i = 0; i < rows; i++)
    for (int j = 0; j < (nb > 0 ? na : i + 1); j++)
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

template <int KEY, int W, int K>
__device__ __forceinline__ void tq_run(const _Float16 *hi, const _Float16 *lo,
                                       float16_t (&acc)[TQ_MAXT]) {
  constexpr TqProg P = tq_prog(KEY, W);
  if constexpr (K < P.n) {
    constexpr TqSub b = P.s[K];
    tq_sub<b.r0, b.nr, b.c0, b.nc, b.tri, tq_sub_off(KEY, W, K)>(hi, lo, acc);
    tq_run<KEY, W, K + 1>(hi, lo, acc);
  }
}

// Tile (I, J) of the part goes to out + t * 1024 with t = I (I + 1) / 2 + J (triangle) or
// t = I NA + J (rectangle, I the row block within segment b).
template <int KEY, int W, int K>
__device__ __forceinline__ void tq_store(float *__restrict__ out, const float16_t (&acc)[TQ_MAXT],
                                         int lane) {
  constexpr TqProg P = tq_prog(KEY, W);
  constexpr int NA = KEY % 64, NB = KEY / 64;
  if constexpr (K < P.n) {
    constexpr TqSub b = P.s[K];
    constexpr int OFF = tq_sub_off(KEY, W, K);
#pragma unroll
    for (int i = 0; i < b.nr; i++) {
#pragma unroll
      for (int j = 0; j < b.nc; j++) {
        if (b.tri && j > i) continue;
        const int t = OFF + (b.tri ? i * (i + 1) / 2 + j : i * b.nc + j);
        const int I = b.r0 + i - (NB > 0 ? NA : 0), J = b.c0 + j;
        const int ot = NB > 0 ? I * NA + J : I * (I + 1) / 2 + J;
        float *o = out + (int64_t)ot * 1024;
#pragma unroll
        for (int e = 0; e < 16; e++) o[e * 64 + lane] = acc[t][e];
      }
    }
    tq_store<KEY, W, K + 1>(out, acc, lane);
  }
}

// LDS ring of raw stage records: TQ_RING slots, each one stage's FP32 record (4 quads x M
// columns x 4 points, the global layout) and its 16 coefficients (replicated over 64 floats).
static constexpr int TQ_RING = 3;

// s_waitcnt with vmcnt = v (0..63), lgkmcnt = 0 and expcnt at its maximum (gfx9 encoding:
// vmcnt[3:0] in bits 3:0, vmcnt[5:4] in bits 15:14, expcnt in 6:4, lgkmcnt in 11:8).
#define TQ_WAIT_VM_LGKM0(v) __builtin_amdgcn_s_waitcnt(((v)&15) | (((v) >> 4) << 14) | (7 << 4))

// Part KEY (see tq_prog): the staged columns are one segment record (triangle, xa) or two
// (rectangle: xa then xb, each [stage][quad][its columns] as written by the split planes).
template <int KEY, int W>
__device__ __forceinline__ void tq_wave(_Float16 *zt, float4_t *ring, float *uring,
                                        const float4_t *__restrict__ xa,
                                        const float4_t *__restrict__ xb,
                                        const float *__restrict__ u, int64_t p_begin,
                                        int64_t p_end, float *__restrict__ out) {
  constexpr int NA = KEY % 64, NB = KEY / 64;
  constexpr int MA = 32 * NA, MB = 32 * NB;
  constexpr int M = MA + MB;  // staged columns
  constexpr int64_t PLANE = (int64_t)M * TQ_LD;
  // Wave TQ_WAVES - 1 holds no column (M <= 416 < 448) and doubles as the data mover: per stage
  // it issues G 1 KB copies of the record(s) and one of the coefficients into the ring.
  constexpr bool MOVER = W == TQ_WAVES - 1;
  constexpr int GA = 2 * NA, GB = 2 * NB;  // 1 KB granules per stage record (64 columns bytes)
  constexpr int COPIES = GA + GB + 1;      // copies per stage
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

  // Mover: stage t's record(s) and coefficients into ring slot t % TQ_RING (segment a's record,
  // then segment b's). Each lane copies 16 bytes per granule (lane l to byte 16 l of the
  // granule), and 4 bytes of coefficients (lane l gets coefficient l % 16).
  auto issue = [&](int t) {
    char *dst = reinterpret_cast<char *>(ring + (t % TQ_RING) * 4 * M);
    const char *sa = reinterpret_cast<const char *>(xa + (stage0 + t) * 4 * MA) + 16 * lane;
#pragma unroll
    for (int g = 0; g < GA; g++)
      __builtin_amdgcn_global_load_lds((const void *)(sa + 1024 * g),
                                       (__attribute__((address_space(3))) void *)(dst + 1024 * g),
                                       16, 0, 0);
    if constexpr (NB > 0) {
      const char *sb = reinterpret_cast<const char *>(xb + (stage0 + t) * 4 * MB) + 16 * lane;
#pragma unroll
      for (int g = 0; g < GB; g++)
        __builtin_amdgcn_global_load_lds(
            (const void *)(sb + 1024 * g),
            (__attribute__((address_space(3))) void *)(dst + 64 * MA + 1024 * g), 16, 0, 0);
    }
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
      if (NB == 0 || col < MA) {
#pragma unroll
        for (int q = 0; q < 4; q++) rec[q] = rs[q * MA + col];
      } else {
#pragma unroll
        for (int q = 0; q < 4; q++) rec[q] = rs[4 * MA + q * MB + (col - MA)];
      }
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
    tq_run<KEY, W, 0>(hi, lo, acc);
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
  tq_store<KEY, W, 0>(out, acc, lane);
}

// u: 2^g sqrt(c'_p), zero-padded to a multiple of 16 points past n. KEY < 64: the lower
// triangle of the KEY tile indices of the record xa; KEY = NA + 64 NB: the rectangle of segment
// b (xb) rows by segment a (xa) columns. partials: [chunk][part tiles][1024].
template <int KEY>
__global__ void __launch_bounds__(64 * TQ_WAVES)
k_tpass(const float4_t *__restrict__ xa, const float4_t *__restrict__ xb,
        const float *__restrict__ u, int64_t n, int64_t chunk, float *__restrict__ partials) {
  static_assert(tq_progs_valid(KEY), "T-pass wave programs must cover the part once");
  constexpr int NA = KEY % 64, NB = KEY / 64;
  constexpr int M = 32 * (NA + NB);
  constexpr int TILES = NB > 0 ? NA * NB : NA * (NA + 1) / 2;
  __shared__ __attribute__((aligned(16))) _Float16 zt[2 * 2 * M * TQ_LD];
  __shared__ float4_t ring[TQ_RING * 4 * M];
  __shared__ float uring[TQ_RING * 64];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t chunk_id = blockIdx.x;
  const int64_t p_begin = chunk_id * chunk;
  const int64_t p_end = p_begin + chunk < n ? p_begin + chunk : n;
  float *out = partials + chunk_id * (int64_t)TILES * 1024;
  switch (wave) {
    case 0: tq_wave<KEY, 0>(zt, ring, uring, xa, xb, u, p_begin, p_end, out); break;
    case 1: tq_wave<KEY, 1>(zt, ring, uring, xa, xb, u, p_begin, p_end, out); break;
    case 2: tq_wave<KEY, 2>(zt, ring, uring, xa, xb, u, p_begin, p_end, out); break;
    case 3: tq_wave<KEY, 3>(zt, ring, uring, xa, xb, u, p_begin, p_end, out); break;
    case 4: tq_wave<KEY, 4>(zt, ring, uring, xa, xb, u, p_begin, p_end, out); break;
    case 5: tq_wave<KEY, 5>(zt, ring, uring, xa, xb, u, p_begin, p_end, out); break;
    case 6: tq_wave<KEY, 6>(zt, ring, uring, xa, xb, u, p_begin, p_end, out); break;
    default: tq_wave<KEY, 7>(zt, ring, uring, xa, xb, u, p_begin, p_end, out); break;
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
// Tile t of the part is (I, J) = (row0 + t / na, col0 + t % na) for a rectangle (na > 0), or
// (row0 + I', col0 + J') with t = I' (I' + 1) / 2 + J' for a triangle (na = 0; row0 = col0),
// in 32-column blocks of the full matrix.
__global__ void k_tpass_final(const double *__restrict__ seg, int64_t per_chunk, int n_seg,
                              int64_t m, int row0, int col0, int na, double *__restrict__ xx) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= per_chunk) return;
  double s = 0.0;
  for (int g = 0; g < n_seg; g++) s += seg[(int64_t)g * per_chunk + idx];
  const int t = (int)(idx / 1024), el = (int)(idx % 1024);
  int I, J;
  if (na > 0) {
    I = t / na;
    J = t % na;
  } else {
    I = 0;
    while ((I + 1) * (I + 2) / 2 <= t) I++;
    J = t - I * (I + 1) / 2;
  }
  const int e = el / 64, ln = el % 64;
  const int r = 8 * (e / 4) + 4 * (ln / 32) + (e % 4), c = ln % 32;
  if (na == 0 && I == J && r < c) return;
  const int64_t gi = 32 * (int64_t)(row0 + I) + r, gj = 32 * (int64_t)(col0 + J) + c;
  xx[gi * m + gj] = s;
  xx[gj * m + gi] = s;
}

// Column segments of the T pass: one segment of all M columns up to 416, else segments of
// TQ_SEGB tile indices (192 columns; the last one shorter). The split planes store the stage
// records segment by segment (tpass_segment_cols), and the T pass runs one part per segment
// pair: the lower triangle of each segment and the rectangle of each pair a < b.
static constexpr int TQ_SEGB = 6;

int tpass_max_cols() { return 1024; }
int64_t tpass_segment_cols(int64_t m_cols) {
  return m_cols <= 32 * TQ_MAXNT ? m_cols : 32 * TQ_SEGB;
}

struct TqPart {
  int a, b, na, nb, tiles;  // segments a <= b; na, nb their tile indices (nb = 0: triangle)
  int64_t tile_off;         // tiles of the parts before this one
};

static std::vector<TqPart> tq_parts(int64_t m_cols) {
  const int nt = (int)(m_cols / 32);
  const int sb = (int)(tpass_segment_cols(m_cols) / 32);
  const int n_seg = (nt + sb - 1) / sb;
  std::vector<TqPart> parts;
  int64_t off = 0;
  for (int b = 0; b < n_seg; b++)
    for (int a = 0; a <= b; a++) {
      const int na = std::min(sb, nt - a * sb), nbk = std::min(sb, nt - b * sb);
      TqPart q{a, b, na, a == b ? 0 : nbk, a == b ? na * (na + 1) / 2 : na * nbk, off};
      off += q.tiles;
      parts.push_back(q);
    }
  return parts;
}

size_t tpass_partials(int64_t m_cols, int64_t n_chunks) {
  const int64_t nt = m_cols / 32;
  return (size_t)(nt * (nt + 1) / 2) * 1024 * (size_t)n_chunks;
}
size_t tpass_seg_count(int64_t m_cols, int64_t n_chunks) {
  const int64_t nt = m_cols / 32;
  return (size_t)(nt * (nt + 1) / 2) * 1024 * (size_t)((n_chunks + TQ_SEG - 1) / TQ_SEG);
}

// xt: the stage records of all segments (segment a at n_pad * a * segment columns floats).
void launch_tpass(hipStream_t st, const float *xt, const float *u, int64_t n, int64_t n_pad,
                  int64_t m_cols, int64_t chunk, int64_t n_chunks, float *partials) {
  const dim3 grid((unsigned)n_chunks), block(64 * TQ_WAVES);
  const int64_t sw = tpass_segment_cols(m_cols);
  for (const TqPart &q : tq_parts(m_cols)) {
    const float4_t *xa = reinterpret_cast<const float4_t *>(xt + n_pad * sw * q.a);
    const float4_t *xb = reinterpret_cast<const float4_t *>(xt + n_pad * sw * q.b);
    float *out = partials + q.tile_off * 1024 * n_chunks;
    const int key = q.na + 64 * q.nb;
    switch (key) {
#define TQ_CASE(k)                                                                          \
  case k:                                                                                   \
    hipLaunchKernelGGL(k_tpass<k>, grid, block, 0, st, xa, xb, u, n, chunk, out);           \
    break;
      TQ_CASE(1) TQ_CASE(2) TQ_CASE(3) TQ_CASE(4) TQ_CASE(5) TQ_CASE(6) TQ_CASE(7)
      TQ_CASE(8) TQ_CASE(9) TQ_CASE(10) TQ_CASE(11) TQ_CASE(12) TQ_CASE(13)
      TQ_CASE(6 + 64) TQ_CASE(6 + 128) TQ_CASE(6 + 192) TQ_CASE(6 + 256) TQ_CASE(6 + 320)
      TQ_CASE(6 + 384)
#undef TQ_CASE
      default: break;  // rejected by the host (M <= 1024)
    }
  }
}

void launch_tpass_reduce(hipStream_t st, const float *partials, int64_t m_cols, int64_t n_chunks,
                         double *seg, double *xx) {
  const int n_seg = (int)((n_chunks + TQ_SEG - 1) / TQ_SEG);
  const int sb = (int)(tpass_segment_cols(m_cols) / 32);
  for (const TqPart &q : tq_parts(m_cols)) {
    const int64_t per_chunk = (int64_t)q.tiles * 1024;
    const float *pp = partials + q.tile_off * 1024 * n_chunks;
    hipLaunchKernelGGL(k_tpass_seg, dim3((unsigned)((per_chunk + 255) / 256), (unsigned)n_seg),
                       dim3(256), 0, st, pp, per_chunk, n_chunks, seg);
    hipLaunchKernelGGL(k_tpass_final, dim3((unsigned)((per_chunk + 255) / 256)), dim3(256), 0, st,
                       seg, per_chunk, n_seg, m_cols, sb * q.b, sb * q.a, q.nb > 0 ? q.na : 0, xx);
  }
}

}  // namespace cy3q
