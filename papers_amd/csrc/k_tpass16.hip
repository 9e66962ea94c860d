// T pass of the split-precision Donaldson step (reference contraction: quintic.rs:1777-1789):
// T = sum_p c_p x_p x_p^T over the real-ified section rows, written as X^T (c o X), on the
// fp16 MFMA (v_mfma_f32_16x16x32_f16) with FP32 accumulation within a point chunk.
//
// Operands. A = the static fp16 hi/lo planes of the row-scaled sections (x = hi + lo). B = the
// same rows multiplied by the point coefficient c'_p 2^g in FP32 and re-split into fp16 hi/lo
// in registers. A product is A_hi B_hi + A_hi B_lo + A_lo B_hi.
//
// Work split. One workgroup per CU streams one point chunk, 32 points per stage. The lower
// triangle of 16 x 16 tiles (NT = M / 16 tile indices) is cut into four regular blocks, one per
// wave (one wave per SIMD, so each wave may hold up to 512 registers):
//   wave 0: triangle of indices [0, H);   wave 1: triangle of [H, NT);
//   waves 2 and 3: the rectangle [H, NT) x [0, H), halved in row-major tile order;
// with H = ceil(NT / 2); for NT = 26 that is 91, 91, 85 and 84 tiles. A triangle is worked as
// triangle + rectangle + triangle (at most ceil(H / 2) fragments held at once); a wave reads each
// fragment of a sub-block once per stage and reuses it for all of the sub-block's tiles.
//
// Staging. Each stage (M columns x 32 points x 2 planes, plus the 32 coefficients) is copied
// from HBM straight into a TP_RING-slot LDS ring with global_load_lds_dwordx4, TP_RING - 1
// stages ahead. The 16-byte chunks (8 points of one column and plane) are XOR-swizzled so the
// MFMA operand reads are free of bank conflicts. Every wave waits for its own copies (its
// vector-memory counter tracks them alone), then one barrier per stage publishes the stage.
//
// Output: FP32 chunk partials [chunk][tile][256] in MFMA output order; k_tpass16_reduce sums
// them over the chunks in FP64, in chunk order (deterministic).
#include "cy3q_internal.h"

namespace cy3q {

static constexpr int TP_WAVES = 4;
static constexpr int TP_RING = 2;  // raw stage slots (one more LDS buffer holds scaled B)
static constexpr int TP_PTS = 32;  // points per stage

template <int NT>
__host__ __device__ constexpr int tp_stage_bytes() {
  return 2 * 16 * NT * 64 + TP_PTS * 4;
}

// Byte offset of the 16-byte chunk q (points 8q..8q+7) of column c, plane p, in a stage buffer.
__device__ __forceinline__ int tp_off(int m, int p, int c, int q) {
  return p * m * 64 + c * 64 + 16 * (q ^ ((c >> 1) & 3));
}

// s_waitcnt vmcnt(v) lgkmcnt(0) (gfx9 encoding; v in 0..63).
#define TP_WAIT(v) \
  __builtin_amdgcn_s_waitcnt(((v)&15) | (((v) >> 4) << 14) | (7 << 4) | (0 << 8))

typedef __attribute__((address_space(3))) void *lds_ptr_t;

// Copies of stage s issued by one wave: its NT / 2 blocks of 1 KB (wave 0 also the 128-byte
// coefficient record, with 8 lanes). Block b = plane p, tile f (16 columns); lane l fills LDS
// slot l: column 16 f + l / 4, swizzled chunk (l % 4) ^ swz, i.e. chunk q. In the source the lane
// part of the offset, lane_src, is the same for every block: the swizzle term depends on
// (l / 8) % 4 only.
__device__ __forceinline__ int64_t tp_lane_src(int m, int lane) {
  const int cl = lane >> 2, qs = lane & 3;
  const int q = qs ^ ((cl >> 1) & 3);
  return (int64_t)(q >> 1) * 32 * m + (int64_t)cl * 32 + 8 * (q & 1);  // halves
}

template <int NT>
__device__ __forceinline__ void tp_issue(char *lds, const _Float16 *__restrict__ xt,
                                         const float *__restrict__ cvec, int64_t s_glob, int slot,
                                         int wave, int lane, int64_t lane_src) {
  constexpr int M = 16 * NT, BLK = 2 * NT;
  char *buf = lds + slot * tp_stage_bytes<NT>();
  // 16-point records 2 s_glob and 2 s_glob + 1 of all columns: [rec][col][plane][16 halves].
  const _Float16 *base = xt + 2 * s_glob * 32 * (int64_t)M + lane_src;
#pragma unroll
  for (int k = 0; k < BLK / TP_WAVES; k++) {
    const int b = wave + TP_WAVES * k;
    const int p = b / NT, f = b % NT;
    __builtin_amdgcn_global_load_lds((const void *)(base + f * 512 + p * 16),
                                     (lds_ptr_t)(buf + b * 1024), 16, 0, 0);
  }
  if (wave == 0 && lane < TP_PTS / 4)
    __builtin_amdgcn_global_load_lds((const void *)(cvec + s_glob * TP_PTS + 4 * lane),
                                     (lds_ptr_t)(buf + 2 * M * 64), 16, 0, 0);
}

// One sub-block of a wave for one stage: rows [R0, R0 + NR) x cols [C0, C0 + NC); TRI: the lower
// triangle of the square block (R0 == C0, NR == NC). Accumulators acc[OFF ..]. The NR row
// fragments are held; the column fragments are read (TRI: reused) one at a time, multiplied by
// the lane's 8 point coefficients and re-split into fp16 hi/lo.
template <int NT, int R0, int NR, int C0, int NC, bool TRI>
__host__ __device__ constexpr int tp_sub_tiles() {
  return NR <= 0 ? 0 : (TRI ? NR * (NR + 1) / 2 : NR * NC);
}

// la = raw stage buffer + this lane's offset (column 16 f + l % 16, chunk l / 16, swizzled), lb =
// the same in the scaled buffer: the swizzle term ((16 f + l % 16) / 2) % 4 does not depend on f,
// so every fragment read is la or lb plus a compile-time offset. The NR row fragments (A, raw)
// are held; the column fragments (B, scaled) are read one at a time.
template <int NT, int R0, int NR, int C0, int NC, bool TRI, int OFF, int NACC>
__device__ __forceinline__ void tp_sub(const char *la, const char *lb, float4_t (&acc)[NACC]) {
  if constexpr (NR > 0) {
    constexpr int M = 16 * NT;
    half8_t ah[NR], al[NR];
#pragma unroll
    for (int i = 0; i < NR; i++) {
      ah[i] = *reinterpret_cast<const half8_t *>(la + (R0 + i) * 1024);
      al[i] = *reinterpret_cast<const half8_t *>(la + M * 64 + (R0 + i) * 1024);
    }
#pragma unroll
    for (int j = 0; j < NC; j++) {
      const half8_t bh = *reinterpret_cast<const half8_t *>(lb + (C0 + j) * 1024);
      const half8_t bl = *reinterpret_cast<const half8_t *>(lb + M * 64 + (C0 + j) * 1024);
#pragma unroll
      for (int pr = 0; pr < 3; pr++) {
#pragma unroll
        for (int i = 0; i < NR; i++) {
          if (TRI && i < j) continue;
          const int t = OFF + (TRI ? i * (i + 1) / 2 + j : i * NC + j);
          const half8_t a = pr == 2 ? al[i] : ah[i];
          const half8_t b = pr == 1 ? bl : bh;
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc[t], 0, 0, 0);
        }
      }
    }
  }
}

template <int NT, int R0, int NR, int C0, int NC, bool TRI, int OFF, int NACC>
__device__ __forceinline__ void tp_store(float *__restrict__ out, const float4_t (&acc)[NACC],
                                         int lane) {
  if constexpr (NR > 0) {
#pragma unroll
    for (int i = 0; i < NR; i++) {
#pragma unroll
      for (int j = 0; j < NC; j++) {
        if (TRI && j > i) continue;
        const int t = OFF + (TRI ? i * (i + 1) / 2 + j : i * NC + j);
        const int I = R0 + i, J = C0 + j;
        float *o = out + (int64_t)(I * (I + 1) / 2 + J) * 256;
#pragma unroll
        for (int e = 0; e < 4; e++) o[e * 64 + lane] = acc[t][e];
      }
    }
  }
}

// A wave's work: up to three sub-blocks (row0, rows, col0, cols, triangle?), worked in order
// every stage; their accumulators follow each other. PER = copies this wave issues per stage.
template <int NT, int PER,                                  //
          int AR, int ANR, int AC, int ANC, bool AT,         // sub-block 1
          int BR, int BNR, int BC, int BNC, bool BT,         // sub-block 2
          int CR, int CNR, int CC, int CNC, bool CT>         // sub-block 3
__device__ __forceinline__ void tp_wave(char *lds, const _Float16 *__restrict__ xt,
                                        const float *__restrict__ cvec, int64_t s0, int n_stage,
                                        float *__restrict__ out, int wave, int lane) {
  constexpr int M = 16 * NT;
  constexpr int T1 = tp_sub_tiles<NT, AR, ANR, AC, ANC, AT>();
  constexpr int T2 = tp_sub_tiles<NT, BR, BNR, BC, BNC, BT>();
  constexpr int T3 = tp_sub_tiles<NT, CR, CNR, CC, CNC, CT>();
  constexpr int NACC = (T1 + T2 + T3) > 0 ? (T1 + T2 + T3) : 1;
  float4_t acc[NACC];
#pragma unroll
  for (int t = 0; t < NACC; t++) acc[t] = (float4_t){0.f, 0.f, 0.f, 0.f};
  const int ql = lane >> 4, cl = lane & 15;
  const int64_t lane_src = tp_lane_src(M, lane);
  const int lane_lds = cl * 64 + 16 * (ql ^ ((cl >> 1) & 3));

  char *scaled = lds + TP_RING * tp_stage_bytes<NT>();
  if (n_stage > 0) tp_issue<NT>(lds, xt, cvec, s0, 0, wave, lane, lane_src);
  for (int s = 0; s < n_stage; s++) {
    const char *raw = lds + (s % TP_RING) * tp_stage_bytes<NT>();
    // Barrier A: every wave's copies of stage s have landed (all of this wave's copies are
    // waited for, so no completion order is assumed) and every wave has finished the MFMAs of
    // stage s - 1, so its raw slot and the scaled buffer may be overwritten.
    TP_WAIT(0);
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_sched_barrier(0);
    if (s + 1 < n_stage)
      tp_issue<NT>(lds, xt, cvec, s0 + s + 1, (s + 1) % TP_RING, wave, lane, lane_src);
    // B operand of stage s, once for all waves: x = hi + lo, v = c'' x in FP32, re-split.
    {
      const float *cr = reinterpret_cast<const float *>(raw + 2 * M * 64);
      for (int idx = threadIdx.x; idx < M * 4; idx += 64 * TP_WAVES) {
        const int c = idx >> 2, q = idx & 3;
        const int off = c * 64 + 16 * (q ^ ((c >> 1) & 3));
        const half8_t xh = *reinterpret_cast<const half8_t *>(raw + off);
        const half8_t xl = *reinterpret_cast<const half8_t *>(raw + M * 64 + off);
        const float4_t c0 = *reinterpret_cast<const float4_t *>(cr + 8 * q);
        const float4_t c1 = *reinterpret_cast<const float4_t *>(cr + 8 * q + 4);
        const float cs[8] = {c0[0], c0[1], c0[2], c0[3], c1[0], c1[1], c1[2], c1[3]};
        half8_t bh, bl;
#pragma unroll
        for (int e = 0; e < 8; e++) {
          const float v = ((float)xh[e] + (float)xl[e]) * cs[e];
          const _Float16 h = (_Float16)v;
          bh[e] = h;
          bl[e] = (_Float16)(v - (float)h);
        }
        *reinterpret_cast<half8_t *>(scaled + off) = bh;
        *reinterpret_cast<half8_t *>(scaled + M * 64 + off) = bl;
      }
    }
    // Barrier B: the scaled operand is complete (this wave's LDS writes are waited for first).
    // The copies of stage s + 1 stay in flight.
    __builtin_amdgcn_s_waitcnt((15) | (3 << 14) | (7 << 4) | (0 << 8));  // lgkmcnt(0) only
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (T1 + T2 + T3 > 0) {
      const char *la = raw + lane_lds, *lb = scaled + lane_lds;
      tp_sub<NT, AR, ANR, AC, ANC, AT, 0>(la, lb, acc);
      tp_sub<NT, BR, BNR, BC, BNC, BT, T1>(la, lb, acc);
      tp_sub<NT, CR, CNR, CC, CNC, CT, T1 + T2>(la, lb, acc);
    }
  }
  tp_store<NT, AR, ANR, AC, ANC, AT, 0>(out, acc, lane);
  tp_store<NT, BR, BNR, BC, BNC, BT, T1>(out, acc, lane);
  tp_store<NT, CR, CNR, CC, CNC, CT, T1 + T2>(out, acc, lane);
}

// Wave roles (NT tile indices, H = ceil(NT / 2)):
//   wave 0: triangle [0, H) as triangle [0, S) + rectangle [S, H) x [0, S) + triangle [S, H);
//   wave 1: the same for [H, NT);
//   waves 2, 3: the rectangle [H, NT) x [0, H) in row-major order, first half (K2 full rows and
//   R2 tiles of the next row) to wave 2, the rest to wave 3.
template <int NT>
__global__ void __launch_bounds__(64 * TP_WAVES) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_tpass16(const _Float16 *__restrict__ xt, const float *__restrict__ cvec, int64_t n,
          int64_t chunk, float *__restrict__ partials) {
  constexpr int H = (NT + 1) / 2, L = NT - H;
  constexpr int S0 = (H + 1) / 2, S1 = (L + 1) / 2;
  constexpr int TR = L * H, T2H = (TR + 1) / 2, K2 = T2H / H, R2 = T2H % H;
  constexpr int R2N = R2 > 0 ? 1 : 0;  // rows of the split row
  constexpr int K3 = L - K2 - R2N;
  constexpr int NTILE = NT * (NT + 1) / 2;
  constexpr int PER = NT / 2;  // copies per stage of waves 1..3; wave 0 issues PER + 1
  // TP_RING raw stage slots, then the scaled B operand of the current stage.
  __shared__ __attribute__((aligned(16))) char lds[TP_RING * tp_stage_bytes<NT>() + 2 * 16 * NT * 64];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t chunk_id = blockIdx.x;
  const int64_t p0 = chunk_id * chunk;
  const int64_t p1 = p0 + chunk < n ? p0 + chunk : n;
  const int n_stage = p1 > p0 ? (int)((p1 - p0 + TP_PTS - 1) / TP_PTS) : 0;
  const int64_t s0 = p0 / TP_PTS;
  float *out = partials + chunk_id * (int64_t)NTILE * 256;
  switch (wave) {
    case 0:
      tp_wave<NT, PER + 1, 0, S0, 0, S0, true, S0, H - S0, 0, S0, false, S0, H - S0, S0, H - S0,
              true>(lds, xt, cvec, s0, n_stage, out, wave, lane);
      break;
    case 1:
      tp_wave<NT, PER, H, S1, H, S1, true, H + S1, L - S1, H, S1, false, H + S1, L - S1, H + S1,
              L - S1, true>(lds, xt, cvec, s0, n_stage, out, wave, lane);
      break;
    case 2:
      tp_wave<NT, PER, H, K2, 0, H, false, H + K2, R2N, 0, R2, false, 0, 0, 0, 0,
              false>(lds, xt, cvec, s0, n_stage, out, wave, lane);
      break;
    default:
      tp_wave<NT, PER, H + K2, R2N, R2, H - R2, false, H + K2 + R2N, K3,
              0, H, false, 0, 0, 0, 0, false>(lds, xt, cvec, s0, n_stage, out, wave, lane);
      break;
  }
}

// FP64 sums of the chunk partials in two passes: pass 1 sums TP_SEG consecutive chunks per
// segment, pass 2 sums the segments in order and scatters to the full M x M matrix (lower
// element of each diagonal tile mirrored), times the scale of the T pass.
static constexpr int TP_SEG = 32;

__global__ void k_tpass16_seg(const float *__restrict__ partials, int64_t per_chunk,
                              int64_t n_chunks, double *__restrict__ seg) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= per_chunk) return;
  const int64_t c0 = (int64_t)blockIdx.y * TP_SEG;
  const int64_t c1 = c0 + TP_SEG < n_chunks ? c0 + TP_SEG : n_chunks;
  double s = 0.0;
  for (int64_t c = c0; c < c1; c++) s += (double)partials[c * per_chunk + idx];
  seg[(int64_t)blockIdx.y * per_chunk + idx] = s;
}

__global__ void k_tpass16_final(const double *__restrict__ seg, int64_t per_chunk, int n_seg,
                                int64_t m, double *__restrict__ xx) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= per_chunk) return;
  double s = 0.0;
  for (int g = 0; g < n_seg; g++) s += seg[(int64_t)g * per_chunk + idx];
  const int t = (int)(idx / 256), el = (int)(idx % 256);
  int I = 0;
  while ((I + 1) * (I + 2) / 2 <= t) I++;
  const int J = t - I * (I + 1) / 2;
  const int e = el / 64, ln = el % 64;
  const int r = 4 * (ln / 16) + e, c = ln % 16;
  if (I == J && r < c) return;
  const int64_t gi = 16 * (int64_t)I + r, gj = 16 * (int64_t)J + c;
  xx[gi * m + gj] = s;
  xx[gj * m + gi] = s;
}

int tpass16_max_tiles() { return 26; }
size_t tpass16_partials(int64_t m_cols, int64_t n_chunks) {
  const int64_t nt = m_cols / 16;
  return (size_t)(nt * (nt + 1) / 2) * 256 * (size_t)n_chunks;
}
size_t tpass16_seg_count(int64_t m_cols, int64_t n_chunks) {
  const int64_t nt = m_cols / 16;
  return (size_t)(nt * (nt + 1) / 2) * 256 * (size_t)((n_chunks + TP_SEG - 1) / TP_SEG);
}

void launch_tpass16(hipStream_t st, const _Float16 *xt, const float *cvec, int64_t n,
                    int64_t m_cols, int64_t chunk, int64_t n_chunks, float *partials) {
  const dim3 grid((unsigned)n_chunks), block(64 * TP_WAVES);
  switch (m_cols / 16) {
#define TP_CASE(k)                                                                               \
  case k:                                                                                        \
    hipLaunchKernelGGL(k_tpass16<k>, grid, block, 0, st, xt, cvec, n, chunk, partials); \
    break;
    TP_CASE(2) TP_CASE(4) TP_CASE(6) TP_CASE(8) TP_CASE(10) TP_CASE(12) TP_CASE(14)
    TP_CASE(16) TP_CASE(18) TP_CASE(20) TP_CASE(22) TP_CASE(24) TP_CASE(26)
#undef TP_CASE
    default: break;  // rejected by the host (M <= 416)
  }
}

void launch_tpass16_reduce(hipStream_t st, const float *partials, int64_t m_cols,
                           int64_t n_chunks, double *seg, double *xx) {
  const int64_t nt = m_cols / 16, per_chunk = nt * (nt + 1) / 2 * 256;
  const int n_seg = (int)((n_chunks + TP_SEG - 1) / TP_SEG);
  hipLaunchKernelGGL(k_tpass16_seg, dim3((unsigned)((per_chunk + 255) / 256), (unsigned)n_seg),
                     dim3(256), 0, st, partials, per_chunk, n_chunks, seg);
  hipLaunchKernelGGL(k_tpass16_final, dim3((unsigned)((per_chunk + 255) / 256)), dim3(256), 0, st,
                     seg, per_chunk, n_seg, m_cols, xx);
}

}  // namespace cy3q
