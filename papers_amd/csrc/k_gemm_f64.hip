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
                                                     int64_t n_rowtiles, int n_cb, int64_t ldb,
                                                     int64_t ldy) {
  __shared__ double sa[2][TG_ROWS][TG_KT + 1];
  __shared__ double sb[2][TG_KT][TG_COLS];
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
      rb = *reinterpret_cast<const double2 *>(bm + (k0 + kr) * ldb + c0 + c2);
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
          if (gr < rows) y[gr * ldy + gc] = acc[i][j][e];
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
                     y, mode, full, n_rt, (int)n_cb, m_cols, m_cols);
}

// Y = X B for a dense B (k_cols x n_cols, leading dimension ldb; n_cols a multiple of 32 and
// k_cols of 16): y [rows][n_cols] with leading dimension ldy.
void launch_gemm_nn_f64(hipStream_t st, const double *x, int64_t rows, int64_t k_cols,
                        const double *bmat, int64_t ldb, int64_t n_cols, double *y, int64_t ldy) {
  if (rows <= 0) return;
  int64_t n_rt = (rows + TG_ROWS - 1) / TG_ROWS;
  int64_t n_rt8 = (n_rt + 7) / 8 * 8;
  int64_t n_cb = n_cols / TG_COLS;
  hipLaunchKernelGGL(k_trigemm_f64, dim3((unsigned)(n_rt8 * n_cb)), dim3(256), 0, st, x, rows,
                     k_cols, bmat, y, 0, 1, n_rt, (int)n_cb, ldb, ldy);
}

// ----------------------------------------------------------------------------------------
// D pass in FP64: D_p = |x_p B|^2 = sum over the columns of (X B)_pc^2, B the real-ified W^T
// (block upper triangular: column c needs k < 2 (c / 2) + 2). A workgroup takes a block of
// DG_P = 96 points and a range of at most 448 columns (all of them for M <= 448, so X is read
// from HBM once and D_p is complete in the workgroup). Its 8 waves are 2 point groups x 4
// column owners: wave (pg, cm) holds the 16 x 16 accumulators of point tiles 3 pg .. 3 pg + 2
// and column tiles cg = cm, cm + 4, ... of the range. At stage s the active tiles are those
// with cg >= s, so the owners differ by at most one active tile. Each 16-deep k stage of X (64 rows) and of B
// (16 rows of the range) is copied from global memory into one of two LDS slots with 16-byte
// global_load_lds copies, one stage ahead; copies holding only columns of tiles that are
// already finished are skipped, so the B bytes per stage shrink with the triangle. The
// operands of a 4-deep k step are read from LDS before the previous step's MFMAs issue.
// Epilogue: row sums of squares, reduced over the 16 column lanes and the 2 owners in a fixed
// order; with one column range the coefficient c_p = w_p / max(D_p, 1e-30) (0 for an invalid
// weight) is written too.
// MFMA f64 16x16x4: A[i][k] at lane i + 16k, B[k][j] at lane j + 16k, D[i][j] at lane
// j + 16 (i % 4), element i / 4.
// ----------------------------------------------------------------------------------------
static constexpr int DG_OWN = 4;        // column owners (the other factor of 8 waves: point groups)
static constexpr int DG_PT = 3;         // point tiles of 16 per wave
static constexpr int DG_P = 16 * DG_PT * (8 / DG_OWN);  // points per workgroup
static constexpr int DG_MAXW = 416;     // columns per range
static constexpr int DG_NT = (DG_MAXW / 16 + DG_OWN - 1) / DG_OWN;  // column tiles per wave
static constexpr int DG_KS = 16;        // k per stage
static constexpr int DG_ALD = DG_KS + 2;  // LDS row stride of the X stage (doubles)
static constexpr int DG_ACOPY = 2;      // X-stage copies per wave: 16 KB per stage
static constexpr int DG_ABYTES = 8 * DG_ACOPY * 1024;
static constexpr int DG_ASLOTS = 3;     // X stages in flight: from HBM, two stages ahead
static constexpr int DG_BCPW = 7;       // B-stage copy slots per wave
static_assert(DG_P * DG_ALD * 8 <= DG_ABYTES, "X stage size");
static_assert(8 % DG_OWN == 0, "wave layout");

__host__ __device__ constexpr int dg_bld(int width) { return width + 16; }
__host__ __device__ constexpr int dg_bbytes(int width) {
  return (DG_KS * dg_bld(width) * 8 + 1023) / 1024 * 1024;
}
static_assert(dg_bbytes(DG_MAXW) / 1024 <= 8 * DG_BCPW, "copy slots per wave");
static_assert(DG_ASLOTS * DG_ABYTES + 2 * dg_bbytes(DG_MAXW) <= 160 * 1024, "LDS per CU");

__global__ void __launch_bounds__(512) k_dpass_f64(const double *__restrict__ x, int64_t rows,
                                                   int64_t m, const double *__restrict__ bm,
                                                   const double *__restrict__ w,
                                                   double *__restrict__ dpart,
                                                   double *__restrict__ coef, int n_ranges,
                                                   int range_tiles) {
  extern __shared__ __attribute__((aligned(1024))) char dlds[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  // Waves w and w + 4 share a SIMD; giving them owners cm and cm + 2 (mod 4) pairs an owner with
  // more active tiles at a stage with one with fewer, so the SIMDs stay balanced.
  const int pg = wave / DG_OWN, cm = (wave + 2 * pg) % DG_OWN;
  const int q = (int)(blockIdx.x % n_ranges);
  const int64_t p0 = (int64_t)(blockIdx.x / n_ranges) * DG_P;
  const int ncg = (int)(m / 16);
  const int cg_lo = q * range_tiles;
  const int cg_hi = cg_lo + range_tiles < ncg ? cg_lo + range_tiles : ncg;
  const int ntile_r = cg_hi - cg_lo;
  const int width = 16 * ntile_r;
  const int bld = dg_bld(width);
  const int bbytes = dg_bbytes(width);
  const int bcopies = bbytes / 1024;
  char *aring = dlds, *bring = dlds + DG_ASLOTS * DG_ABYTES;
  const int n_stage = 16 * cg_hi / DG_KS;  // k < 16 cg_hi
  // Copy g of a stage writes LDS bytes [1024 g + 16 lane, +16) of its slot. X copy u of this
  // wave (g = wave + 8 u) reads xb + aofs[u] + KS s, B copy u reads bb + bofs[u] + KS m s
  // (doubles); padding pieces fetch a valid element that is never read. bmax[u]: the last range
  // column a B copy holds (-1: none), so it is skipped once every tile it feeds is finished.
  const double *xb = x + p0 * m;
  const double *bb = bm + 16 * cg_lo;
  const int64_t rows_here = rows - p0 < DG_P ? rows - p0 : DG_P;
  int aofs[DG_ACOPY], bofs[DG_BCPW], bmax[DG_BCPW];
#pragma unroll
  for (int u = 0; u < DG_ACOPY; u++) {
    const int o = (wave + 8 * u) * 1024 + lane * 16;
    const int r = o / (DG_ALD * 8), b = o % (DG_ALD * 8);
    const int row = r < rows_here ? r : (int)rows_here - 1;
    aofs[u] = (r < DG_P && b < DG_KS * 8) ? row * (int)m + b / 8 : 0;
  }
#pragma unroll
  for (int u = 0; u < DG_BCPW; u++) {
    const int g = wave + 8 * u;
    const int o2 = g * 1024 + lane * 16;
    const int kr = o2 / (bld * 8), b = o2 % (bld * 8);
    bofs[u] = (kr < DG_KS && b < width * 8) ? kr * (int)m + b / 8 : 0;
    const int rowb = bld * 8, s0 = g * 1024, e0 = s0 + 1023;
    const int rs = s0 / rowb, re = e0 / rowb;
    const int last = (e0 % rowb) / 8 < width - 1 ? (e0 % rowb) / 8 : width - 1;
    bmax[u] = g >= bcopies || rs >= DG_KS ? -1 : (re > rs ? width - 1 : last);
  }
  auto issue_a = [&](int st) {
    char *dst = aring + (st % DG_ASLOTS) * DG_ABYTES;
#pragma unroll
    for (int u = 0; u < DG_ACOPY; u++)
      __builtin_amdgcn_global_load_lds(
          (const void *)(xb + aofs[u] + (int64_t)st * DG_KS),
          (__attribute__((address_space(3))) void *)(dst + (wave + 8 * u) * 1024), 16, 0, 0);
  };
  auto issue_b = [&](int st) {
    char *dst = bring + (st & 1) * bbytes;
    const int need = st > cg_lo ? 16 * (st - cg_lo) : 0;  // first range column still needed
#pragma unroll
    for (int u = 0; u < DG_BCPW; u++)
      if (bmax[u] >= need)
        __builtin_amdgcn_global_load_lds(
            (const void *)(bb + bofs[u] + (int64_t)st * DG_KS * m),
            (__attribute__((address_space(3))) void *)(dst + (wave + 8 * u) * 1024), 16, 0, 0);
  };
  double4_t acc[DG_PT][DG_NT];
#pragma unroll
  for (int t = 0; t < DG_PT; t++)
#pragma unroll
    for (int v = 0; v < DG_NT; v++) acc[t][v] = (double4_t){0.0, 0.0, 0.0, 0.0};
  // Issue order: X(0), B(0), X(1), then at stage s: B(s + 1), X(s + 2). At the top of stage s
  // the only copies of this wave that may still be in flight are the DG_ACOPY of X(s + 1).
  issue_a(0);
  issue_b(0);
  if (n_stage > 1) issue_a(1);
  // Stages in epochs of this wave's first active column slot: at stage st the active slots are
  // v_lo(st) <= v < v_hi with v_lo(st) = ceil((st - cg_lo - cm) / 4), so epoch E holds the stages
  // st <= cg_lo + cm + 4 E not in an earlier epoch (and epoch 7 the rest, with no slot). E, and v_hi (this wave's number of column
  // slots, dispatched once), are compile-time constants in each epoch's body, so the body has no
  // tile guards and reads no B for finished slots.
  const int v_hi = ntile_r - cm > 0 ? (ntile_r - cm + DG_OWN - 1) / DG_OWN : 0;
  int st = 0;
  auto run_all = [&](auto vhc) {
    constexpr int VHI = decltype(vhc)::value;
    auto run_epoch = [&](auto ec) {
      constexpr int E = decltype(ec)::value;
      // Epoch DG_NT: the stages past every slot of this wave (their B copies of the wave's
      // tiles are skipped, so no slot may read them).
      const int st_end = E == DG_NT ? n_stage
                                    : (cg_lo + cm + 4 * E + 1 < n_stage ? cg_lo + cm + 4 * E + 1
                                                                        : n_stage);
      for (; st < st_end; st++) {
        asm volatile("" ::: "memory");
        if (st + 1 < n_stage)
          __builtin_amdgcn_s_waitcnt(0x72);  // vmcnt(2): X(st) and B(st) have landed
        else
          __builtin_amdgcn_s_waitcnt(0x70);  // vmcnt(0)
        __syncthreads();                     // ... for every wave; B slot (st + 1) & 1 and X
        asm volatile("" ::: "memory");       // slot (st + 2) % 3 are free
        if (st + 1 < n_stage) issue_b(st + 1);
        if (st + 2 < n_stage) issue_a(st + 2);
        if (E >= VHI) continue;  // no active slot: only the copies and the barrier
        const double *sa = reinterpret_cast<const double *>(aring + (st % DG_ASLOTS) * DG_ABYTES);
        const double *sb = reinterpret_cast<const double *>(bring + (st & 1) * bbytes);
        // All operands of a 4-deep k step are read before its MFMAs, and the next step's
        // operands are read before the current step's MFMAs are issued.
        auto load = [&](int kk, double (&a)[DG_PT], double (&b)[DG_NT]) {
          const int kl = 4 * kk + (lane >> 4);
#pragma unroll
          for (int t = 0; t < DG_PT; t++)
            a[t] = sa[(16 * (DG_PT * pg + t) + (lane & 15)) * DG_ALD + kl];
#pragma unroll
          for (int v = E; v < VHI; v++) b[v] = sb[kl * bld + 16 * (cm + DG_OWN * v) + (lane & 15)];
        };
        double a0[DG_PT], b0[DG_NT], a1[DG_PT], b1[DG_NT];
        load(0, a0, b0);
#pragma unroll
        for (int kk = 0; kk < DG_KS / 4; kk++) {
          double(&a)[DG_PT] = (kk & 1) ? a1 : a0;
          double(&b)[DG_NT] = (kk & 1) ? b1 : b0;
          if (kk + 1 < DG_KS / 4) {
            if (kk & 1)
              load(kk + 1, a0, b0);
            else
              load(kk + 1, a1, b1);
          }
#pragma unroll
          for (int v = E; v < VHI; v++)
#pragma unroll
            for (int t = 0; t < DG_PT; t++)
              acc[t][v] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], b[v], acc[t][v], 0, 0, 0);
        }
      }
    };
    run_epoch(std::integral_constant<int, 0>{});
    run_epoch(std::integral_constant<int, 1>{});
    run_epoch(std::integral_constant<int, 2>{});
    run_epoch(std::integral_constant<int, 3>{});
    run_epoch(std::integral_constant<int, 4>{});
    run_epoch(std::integral_constant<int, 5>{});
    run_epoch(std::integral_constant<int, 6>{});
    run_epoch(std::integral_constant<int, 7>{});
  };
  static_assert(DG_NT == 7, "epochs 0 .. 7, v_hi 0 .. 7");
  switch (v_hi) {
    case 0: run_all(std::integral_constant<int, 0>{}); break;
    case 1: run_all(std::integral_constant<int, 1>{}); break;
    case 2: run_all(std::integral_constant<int, 2>{}); break;
    case 3: run_all(std::integral_constant<int, 3>{}); break;
    case 4: run_all(std::integral_constant<int, 4>{}); break;
    case 5: run_all(std::integral_constant<int, 5>{}); break;
    case 6: run_all(std::integral_constant<int, 6>{}); break;
    default: run_all(std::integral_constant<int, 7>{}); break;
  }
  // Row sums of squares: element e of lane l is point 16 (PT pg + t) + 4 e + l / 16.
  __syncthreads();  // the LDS slots are reused for the owner partials
  double *part = reinterpret_cast<double *>(dlds);  // [DG_OWN][DG_P]
#pragma unroll
  for (int t = 0; t < DG_PT; t++)
#pragma unroll
    for (int e = 0; e < 4; e++) {
      double sq = 0.0;
#pragma unroll
      for (int v = 0; v < DG_NT; v++) sq += acc[t][v][e] * acc[t][v][e];
      sq += __shfl_xor(sq, 1);
      sq += __shfl_xor(sq, 2);
      sq += __shfl_xor(sq, 4);
      sq += __shfl_xor(sq, 8);
      if ((lane & 15) == 0) part[cm * DG_P + 16 * (DG_PT * pg + t) + 4 * e + (lane >> 4)] = sq;
    }
  __syncthreads();
  if (tid < DG_P) {
    const int64_t p = p0 + tid;
    if (p < rows) {
      double d = 0.0;
#pragma unroll
      for (int o = 0; o < DG_OWN; o++) d += part[o * DG_P + tid];
      dpart[(int64_t)q * rows + p] = d;
      if (coef) {
        const double wp = w[p];
        coef[p] = (isfinite(wp) && wp > 0.0) ? (1.0 / fmax(d, 1e-30)) * wp : 0.0;
      }
    }
  }
}

// Column ranges of the FP64 D pass for M real columns.
int dpass_f64_ranges(int64_t m_cols) {
  const int64_t ncg = m_cols / 16, per = DG_MAXW / 16;
  return (int)((ncg + per - 1) / per);
}

// D parts (dpart[q][p], q < dpass_f64_ranges(M)) and, with one range, the coefficients c_p;
// with several ranges the coefficients come from launch_point_coeffs_f64 over the parts.
void launch_dpass_f64(hipStream_t st, const double *x, int64_t rows, int64_t m_cols,
                      const double *bmat, const double *w, double *dpart, double *coef) {
  if (rows <= 0) return;
  const int nr = dpass_f64_ranges(m_cols);
  const int ncg = (int)(m_cols / 16);
  const int per = (ncg + nr - 1) / nr;  // tiles per range, balanced
  const int width = 16 * per;
  const size_t lds = (size_t)DG_ASLOTS * DG_ABYTES + 2 * (size_t)dg_bbytes(width);
  const int64_t blocks = (rows + DG_P - 1) / DG_P;
  hipLaunchKernelGGL(k_dpass_f64, dim3((unsigned)(blocks * nr)), dim3(512), lds, st, x, rows,
                     m_cols, bmat, w, dpart, nr == 1 ? coef : nullptr, nr, per);
  if (nr > 1) launch_point_coeffs_f64(st, w, dpart, nr, rows, coef);
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
                                                  const int32_t *__restrict__ pcols,
                                                  int seg_cols,
                                                  double *__restrict__ partials) {
  extern __shared__ double sx[];  // 2 buffers of [SY_KP][staged columns + 1] then c[2][SY_KP]
  // XCD-aware: the parts of one chunk share an XCD (and its L2).
  const int64_t g = blockIdx.x, xcd = g & 7, qq = g >> 3;
  const int part = (int)(qq % n_parts);
  const int64_t chunk_id = (qq / n_parts) * 8 + xcd;
  if (chunk_id >= n_chunks) return;
  // Staged columns: all M, or (segment plans) the part's segment a then segment b, with tile
  // groups renumbered to their staged position.
  int a0 = 0, wa = (int)(m_cols / 16), b0 = 0, wb = 0;
  if (pcols) {
    a0 = pcols[4 * part];
    wa = pcols[4 * part + 1];
    b0 = pcols[4 * part + 2];
    wb = pcols[4 * part + 3];
  }
  const int64_t lcols = 16 * (int64_t)(wa + wb);
  const int64_t ld = (pcols ? seg_cols : m_cols) + 1;
  const int64_t bufsz = SY_KP * ld;
  double *sc = sx + 2 * bufsz;
  auto local_group = [&](int grp) { return grp >= a0 && grp < a0 + wa ? grp - a0 : wa + grp - b0; };
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
    if (trg[t] >= 0) {
      trg[t] = local_group(trg[t]);
      tcg[t] = local_group(tcg[t]);
    }
  }

  double4_t acc[SY_MAXT];
#pragma unroll
  for (int t = 0; t < SY_MAXT; t++) acc[t] = (double4_t){0.0, 0.0, 0.0, 0.0};

  const int64_t p_begin = chunk_id * chunk;
  const int64_t p_end = p_begin + chunk < n ? p_begin + chunk : n;
  const int n_stage = (int)((p_end - p_begin + SY_KP - 1) / SY_KP);
  const int64_t half = lcols / 2;
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
        const int64_t gc = cc < 16 * wa ? 16 * a0 + cc : 16 * b0 + (cc - 16 * wa);
        if (p0 + r < p_end) v = *reinterpret_cast<const double2 *>(x + (p0 + r) * m_cols + gc);
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

// ----------------------------------------------------------------------------------------
// Weighted SYRK for M = 416 (13 column groups of 32: k = 6 with the quotient basis) with
// wave programs fixed at compile time. The 91 lower 32 x 32 tiles go to 16 waves in two parts
// (workgroups of 8 waves per point chunk): part 0 holds tile rows 0..8 (and copies column
// groups 0..8 only), part 1 rows 9..12. A wave's sub-blocks are rectangles or lower triangles
// of tiles; it holds the row fragments of a sub-block and streams its column fragments, 4 f64
// 16 x 16 MFMAs per tile and 4-point k step, the column fragments scaled by c_p. At most 6 tiles
// per wave (192 accumulator VGPRs), and every wave runs its own copy of the stage loop, so no
// accumulators are merged across programs.
// Output: partials[chunk][tile (I (I + 1) / 2 + J)][2][2][256], in MFMA output order.
// ----------------------------------------------------------------------------------------
struct S13Sub {
  int r0, nr, c0, nc, tri;
};
struct S13Prog {
  int n;
  S13Sub s[2];
};
static constexpr int S13_KP = 16;
static constexpr int S13_NT = 13;
static constexpr int S13_M = 32 * S13_NT;
static constexpr int S13_LD = S13_M + 2;

__host__ __device__ constexpr S13Prog s13_prog(int part, int w) {
  constexpr S13Prog tab[2][8] = {
      {{1, {{0, 3, 0, 3, 1}, {}}}, {1, {{3, 3, 3, 3, 1}, {}}}, {1, {{6, 3, 6, 3, 1}, {}}},
       {1, {{3, 3, 0, 2, 0}, {}}}, {1, {{3, 3, 2, 1, 0}, {}}}, {1, {{6, 3, 0, 2, 0}, {}}},
       {1, {{6, 3, 2, 2, 0}, {}}}, {1, {{6, 3, 4, 2, 0}, {}}}},
      {{1, {{9, 3, 9, 3, 1}, {}}}, {1, {{9, 3, 0, 2, 0}, {}}}, {1, {{9, 3, 2, 2, 0}, {}}},
       {1, {{9, 3, 4, 2, 0}, {}}}, {1, {{9, 3, 6, 2, 0}, {}}},
       {2, {{9, 3, 8, 1, 0}, {12, 1, 12, 1, 1}}}, {1, {{12, 1, 0, 6, 0}, {}}},
       {1, {{12, 1, 6, 6, 0}, {}}}}};
  return tab[part][w];
}
__host__ __device__ constexpr int s13_tiles(const S13Sub &b) {
  return b.nr <= 0 ? 0 : (b.tri ? b.nr * (b.nr + 1) / 2 : b.nr * b.nc);
}
__host__ __device__ constexpr int s13_sub_off(int part, int w, int k) {
  int o = 0;
  const S13Prog p = s13_prog(part, w);
  for (int q = 0; q < k; q++) o += s13_tiles(p.s[q]);
  return o;
}
// Every lower tile in exactly one program, at most 6 tiles per wave, part 0 within groups 0..8.
__host__ __device__ constexpr bool s13_valid() {
  int cover[S13_NT * S13_NT] = {};
  for (int part = 0; part < 2; part++)
    for (int w = 0; w < 8; w++) {
      const S13Prog p = s13_prog(part, w);
      int t = 0;
      for (int k = 0; k < p.n; k++) {
        const S13Sub b = p.s[k];
        if (b.tri && (b.r0 != b.c0 || b.nr != b.nc)) return false;
        t += s13_tiles(b);
        for (int i = 0; i < b.nr; i++)
          for (int j = 0; j < b.nc; j++) {
            if (b.tri && j > i) continue;
            const int I = b.r0 + i, J = b.c0 + j;
            if (I >= S13_NT || J > I || (part == 0 && I > 8)) return false;
            cover[I * S13_NT + J]++;
          }
      }
      if (t > 6) return false;
    }
  for (int i = 0; i < S13_NT; i++)
    for (int j = 0; j <= i; j++)
      if (cover[i * S13_NT + j] != 1) return false;
  return true;
}
static_assert(s13_valid(), "FP64 SYRK wave programs must cover the lower triangle once");

// One sub-block for one 4-point k step. rowp: this lane's LDS position in the k step's row
// (point 4 kk + lane / 16, column lane % 16).
template <int R0, int NR, int C0, int NC, int TRI, int OFF>
__device__ __forceinline__ void s13_sub(const double *rowp, double cp, double4_t (&acc)[24]) {
  double a[NR][2];
#pragma unroll
  for (int i = 0; i < NR; i++)
#pragma unroll
    for (int h = 0; h < 2; h++) a[i][h] = rowp[(R0 + i) * 32 + 16 * h];
#pragma unroll
  for (int j = 0; j < NC; j++) {
    double b[2];
#pragma unroll
    for (int h = 0; h < 2; h++) b[h] = cp * rowp[(C0 + j) * 32 + 16 * h];
#pragma unroll
    for (int i = 0; i < NR; i++) {
      if (TRI && i < j) continue;
      const int t = OFF + (TRI ? i * (i + 1) / 2 + j : i * NC + j);
#pragma unroll
      for (int ha = 0; ha < 2; ha++)
#pragma unroll
        for (int hb = 0; hb < 2; hb++)
          acc[4 * t + 2 * ha + hb] = __builtin_amdgcn_mfma_f64_16x16x4f64(
              a[i][ha], b[hb], acc[4 * t + 2 * ha + hb], 0, 0, 0);
    }
  }
}

template <int PART, int W, int K>
__device__ __forceinline__ void s13_run(const double *rowp, double cp, double4_t (&acc)[24]) {
  constexpr S13Prog P = s13_prog(PART, W);
  if constexpr (K < P.n) {
    constexpr S13Sub b = P.s[K];
    s13_sub<b.r0, b.nr, b.c0, b.nc, b.tri, s13_sub_off(PART, W, K)>(rowp, cp, acc);
    s13_run<PART, W, K + 1>(rowp, cp, acc);
  }
}

template <int PART, int W, int K>
__device__ __forceinline__ void s13_store(double *__restrict__ out, const double4_t (&acc)[24],
                                          int lane) {
  constexpr S13Prog P = s13_prog(PART, W);
  if constexpr (K < P.n) {
    constexpr S13Sub b = P.s[K];
    constexpr int OFF = s13_sub_off(PART, W, K);
#pragma unroll
    for (int i = 0; i < b.nr; i++)
#pragma unroll
      for (int j = 0; j < b.nc; j++) {
        if (b.tri && j > i) continue;
        const int t = OFF + (b.tri ? i * (i + 1) / 2 + j : i * b.nc + j);
        const int I = b.r0 + i, J = b.c0 + j;
        double *o = out + (int64_t)(I * (I + 1) / 2 + J) * 1024;
#pragma unroll
        for (int ha = 0; ha < 2; ha++)
#pragma unroll
          for (int hb = 0; hb < 2; hb++)
#pragma unroll
            for (int e = 0; e < 4; e++) o[(ha * 2 + hb) * 256 + e * 64 + lane] = acc[4 * t + 2 * ha + hb][e];
      }
    s13_store<PART, W, K + 1>(out, acc, lane);
  }
}

template <int PART, int W>
__device__ __forceinline__ void s13_wave(const double *sxb, const double *scb, int lane,
                                         double4_t (&acc)[24]) {
#pragma unroll
  for (int kk = 0; kk < S13_KP / 4; kk++) {
    const int kr = 4 * kk + (lane >> 4);
    s13_run<PART, W, 0>(sxb + kr * S13_LD + (lane & 15), scb[kr], acc);
  }
}

// One stage: S13_KP point rows of X (row stride S13_LDB bytes in LDS) copied straight from
// global memory into LDS with 16-byte global_load_lds copies, 1 KB per copy and wave (lane l to
// byte 16 l of the copy), so no staging registers are live across the MFMAs. Copy g of a stage
// belongs to wave g % 8. Bytes of a copy past a row's 416 columns (the row padding) or past the
// stage fetch a valid element and are never read; rows past the chunk's last point fetch that
// point again and get coefficient 0. PART 0 copies only its column groups 0..8.
static constexpr int S13_LDB = S13_LD * 8;                                   // bytes per LDS row
static constexpr int S13_STAGE_B = (S13_KP * S13_LDB + 1023) / 1024 * 1024;  // bytes per buffer
static constexpr int S13_COPIES = S13_STAGE_B / 1024;
static constexpr int S13_CPW = (S13_COPIES + 7) / 8;  // copies per wave and stage

// The whole stage loop of wave W of part PART: each wave has its own copy of the loop, so its
// accumulators never merge with another program's.
template <int PART, int W>
__device__ __forceinline__ void s13_body(const double *__restrict__ x, const double *__restrict__ cw,
                                         int64_t n, int64_t chunk_id, int64_t chunk,
                                         double *__restrict__ partials, char *sx,
                                         double (*sc)[S13_KP]) {
  constexpr int ROWB = PART == 0 ? 9 * 32 * 8 : S13_M * 8;  // bytes copied per row
  const int tid = threadIdx.x, lane = tid & 63;
  constexpr int wave = W;
  const int64_t p_begin = chunk_id * chunk;
  const int64_t p_end = p_begin + chunk < n ? p_begin + chunk : n;
  const int n_stage = (int)((p_end - p_begin + S13_KP - 1) / S13_KP);
  // Per copy slot: the LDS row and the byte offset within the source row (-1: padding).
  int crow[S13_CPW], cbyte[S13_CPW];
#pragma unroll
  for (int u = 0; u < S13_CPW; u++) {
    const int g = wave + 8 * u;
    const int o = g * 1024 + lane * 16;
    const int r = o / S13_LDB, b = o % S13_LDB;
    const bool live = g < S13_COPIES && r < S13_KP && b < ROWB;
    crow[u] = live ? r : 0;
    cbyte[u] = live ? b : (g < S13_COPIES ? 0 : -1);  // -1: no copy for this slot
  }
  auto issue = [&](int st, int buf) {
    const int64_t p0 = p_begin + (int64_t)st * S13_KP;
    const int last = (int)((p_end - p0 < S13_KP ? p_end - p0 : S13_KP) - 1);
    const char *xb = reinterpret_cast<const char *>(x + p0 * S13_M);
    char *dst = sx + buf * S13_STAGE_B;
#pragma unroll
    for (int u = 0; u < S13_CPW; u++) {
      if (wave + 8 * u < S13_COPIES) {
        const int r = crow[u] < last ? crow[u] : last;
        const int b = cbyte[u] < 0 ? 0 : cbyte[u];
        __builtin_amdgcn_global_load_lds(
            (const void *)(xb + (int64_t)r * (S13_M * 8) + b),
            (__attribute__((address_space(3))) void *)(dst + (wave + 8 * u) * 1024), 16, 0, 0);
      }
    }
  };
  double4_t acc[24];
#pragma unroll
  for (int t = 0; t < 24; t++) acc[t] = (double4_t){0.0, 0.0, 0.0, 0.0};
  double rc = 0.0;
  auto fetch_c = [&](int st) {
    const int64_t p0 = p_begin + (int64_t)st * S13_KP;
    if (tid < S13_KP) rc = p0 + tid < p_end ? cw[p0 + tid] : 0.0;
  };
  if (n_stage > 0) {
    issue(0, 0);
    fetch_c(0);
    if (tid < S13_KP) sc[0][tid] = rc;
  }
  // vmcnt(0), lgkmcnt(0): this wave's copies and LDS writes have landed; then the barrier.
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0x70);
  __syncthreads();
  asm volatile("" ::: "memory");
  for (int st = 0; st < n_stage; st++) {
    const int buf = st & 1;
    if (st + 1 < n_stage) {
      issue(st + 1, buf ^ 1);
      fetch_c(st + 1);
    }
    s13_wave<PART, W>(reinterpret_cast<const double *>(sx + buf * S13_STAGE_B), sc[buf], lane,
                      acc);
    if (st + 1 < n_stage && tid < S13_KP) sc[buf ^ 1][tid] = rc;
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0x70);
    __syncthreads();
    asm volatile("" ::: "memory");
  }
  s13_store<PART, W, 0>(partials + chunk_id * (int64_t)(91 * 1024), acc, lane);
}

__global__ void __launch_bounds__(512) k_syrk13_f64(const double *__restrict__ x,
                                                    const double *__restrict__ cw, int64_t n,
                                                    int64_t chunk, int64_t n_chunks,
                                                    double *__restrict__ partials) {
  __shared__ __attribute__((aligned(1024))) char sx[2 * S13_STAGE_B];
  __shared__ double sc[2][S13_KP];
  // XCD-aware: the two parts of a chunk go to one XCD and share its L2.
  const int64_t g = blockIdx.x, xcd = g & 7, qq = g >> 3;
  const int part = (int)(qq & 1);
  const int64_t chunk_id = (qq >> 1) * 8 + xcd;
  if (chunk_id >= n_chunks) return;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  switch (part * 8 + wave) {
    case 0: s13_body<0, 0>(x, cw, n, chunk_id, chunk, partials, sx, sc); break;
    case 1: s13_body<0, 1>(x, cw, n, chunk_id, chunk, partials, sx, sc); break;
    case 2: s13_body<0, 2>(x, cw, n, chunk_id, chunk, partials, sx, sc); break;
    case 3: s13_body<0, 3>(x, cw, n, chunk_id, chunk, partials, sx, sc); break;
    case 4: s13_body<0, 4>(x, cw, n, chunk_id, chunk, partials, sx, sc); break;
    case 5: s13_body<0, 5>(x, cw, n, chunk_id, chunk, partials, sx, sc); break;
    case 6: s13_body<0, 6>(x, cw, n, chunk_id, chunk, partials, sx, sc); break;
    case 7: s13_body<0, 7>(x, cw, n, chunk_id, chunk, partials, sx, sc); break;
    case 8: s13_body<1, 0>(x, cw, n, chunk_id, chunk, partials, sx, sc); break;
    case 9: s13_body<1, 1>(x, cw, n, chunk_id, chunk, partials, sx, sc); break;
    case 10: s13_body<1, 2>(x, cw, n, chunk_id, chunk, partials, sx, sc); break;
    case 11: s13_body<1, 3>(x, cw, n, chunk_id, chunk, partials, sx, sc); break;
    case 12: s13_body<1, 4>(x, cw, n, chunk_id, chunk, partials, sx, sc); break;
    case 13: s13_body<1, 5>(x, cw, n, chunk_id, chunk, partials, sx, sc); break;
    case 14: s13_body<1, 6>(x, cw, n, chunk_id, chunk, partials, sx, sc); break;
    default: s13_body<1, 7>(x, cw, n, chunk_id, chunk, partials, sx, sc); break;
  }
}

// Chunk sums of the 91 tile partials into the full 416 x 416 matrix (a diagonal tile keeps its
// lower elements and mirrors them, so every element has one writer). A workgroup takes 64
// consecutive partial entries; its 16 groups of 64 threads sum the chunks grp, grp + 16, ...
// and the group sums are added in group order, so the result is the same on every run.
static constexpr int RED_GROUPS = 16;

__device__ inline double grouped_chunk_sum(const double *__restrict__ partials, int64_t per_chunk,
                                           int64_t n_chunks, int64_t idx, double (*red)[64]) {
  const int tid = threadIdx.x, grp = tid >> 6;
  double s = 0.0;
  if (idx < per_chunk)
    for (int64_t c = grp; c < n_chunks; c += RED_GROUPS) s += partials[c * per_chunk + idx];
  red[grp][tid & 63] = s;
  __syncthreads();
  double r = 0.0;
  if (tid < 64)
    for (int q = 0; q < RED_GROUPS; q++) r += red[q][tid];
  return r;
}

__global__ void __launch_bounds__(1024) k_syrk13_reduce(const double *__restrict__ partials,
                                                        int64_t n_chunks,
                                                        double *__restrict__ xx) {
  __shared__ double red[RED_GROUPS][64];
  const int64_t per_chunk = 91 * 1024;
  const int64_t idx = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  const double s = grouped_chunk_sum(partials, per_chunk, n_chunks, idx, red);
  if (threadIdx.x >= 64 || idx >= per_chunk) return;
  const int t = (int)(idx / 1024), el = (int)(idx % 1024);
  int I = 0;
  while ((I + 1) * (I + 2) / 2 <= t) I++;
  const int J = t - I * (I + 1) / 2;
  const int q = el / 256, e = (el % 256) / 64, ln = el % 64;
  const int r = 16 * (q >> 1) + 4 * e + ln / 16, c = 16 * (q & 1) + ln % 16;
  if (I == J && r < c) return;
  const int64_t gi = 32 * (int64_t)I + r, gj = 32 * (int64_t)J + c;
  xx[gi * S13_M + gj] = s;
  xx[gj * S13_M + gi] = s;
}

int syrk_f64_max_tiles(int64_t m_cols) { return m_cols <= 448 ? 22 : 12; }

void launch_syrk_f64(hipStream_t st, const SyrkPlan &plan, const double *x, const double *c,
                     int64_t n, double *partials) {
  int64_t n_ch8 = (plan.n_chunks + 7) / 8 * 8;
  int64_t grid = n_ch8 * plan.n_parts;
  if (plan.prog13) {
    hipLaunchKernelGGL(k_syrk13_f64, dim3((unsigned)(n_ch8 * 2)), dim3(512), 0, st, x, c, n,
                       plan.chunk, plan.n_chunks, partials);
  } else if (plan.m_cols <= 448 || plan.d_pcols) {
    // All columns (M <= 448) or a segment pair (at most 416 columns) per 16-point stage.
    const int64_t cols = plan.d_pcols ? plan.seg_cols : plan.m_cols;
    size_t lds = sizeof(double) * 2 * (16 * (cols + 1) + 16);
    hipLaunchKernelGGL((k_syrk_f64<22, 7, 16>), dim3((unsigned)grid), dim3(512), lds, st, x, c, n,
                       plan.m_cols, plan.chunk, plan.n_parts, plan.max_tiles, plan.n_chunks,
                       plan.d_tiles, plan.d_pcols, plan.seg_cols, partials);
  } else {  // m_cols <= 1024: 8-point stages keep the two LDS buffers under 160 KB
    size_t lds = sizeof(double) * 2 * (8 * (plan.m_cols + 1) + 8);
    hipLaunchKernelGGL((k_syrk_f64<12, 8, 8>), dim3((unsigned)grid), dim3(512), lds, st, x, c, n,
                       plan.m_cols, plan.chunk, plan.n_parts, plan.max_tiles, plan.n_chunks,
                       plan.d_tiles, plan.d_pcols, plan.seg_cols, partials);
  }
}

// Sum of the chunk partials (grouped as in k_syrk13_reduce), scattered into the full M x M
// matrix (lower tiles and their mirror). One result per (part, wave, tile slot, element).
// Element (e, lane) of a 16x16 f64 tile is (row 4e + lane/16, col lane%16), the MFMA output layout.
__global__ void __launch_bounds__(1024) k_syrk_reduce(const double *__restrict__ partials,
                                                      const int16_t *__restrict__ tiles,
                                                      int n_parts, int waves, int max_tiles,
                                                      int64_t n_chunks, int tile, int64_t m_cols,
                                                      double *__restrict__ xx) {
  __shared__ double red[RED_GROUPS][64];
  const int64_t per_chunk = (int64_t)n_parts * waves * max_tiles * tile * tile;
  const int64_t idx = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  // Slots padded with -1 are skipped by every thread of the workgroup before the reduction
  // (a slot spans whole workgroups: tile * tile is a multiple of 64).
  const int64_t slot0 = ((int64_t)blockIdx.x * 64) / (tile * tile);
  if ((int64_t)blockIdx.x * 64 >= per_chunk || tiles[2 * slot0] < 0) return;
  const double s = grouped_chunk_sum(partials, per_chunk, n_chunks, idx, red);
  if (threadIdx.x >= 64 || idx >= per_chunk) return;
  const int64_t slot = idx / (tile * tile);
  const int el = (int)(idx % (tile * tile));
  const int rg = tiles[2 * slot], cg = tiles[2 * slot + 1];
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
  if (plan.prog13) {
    hipLaunchKernelGGL(k_syrk13_reduce, dim3((unsigned)((91 * 1024 + 63) / 64)), dim3(1024), 0,
                       st, partials, plan.n_chunks, xx_full);
    return;
  }
  int64_t per_chunk = (int64_t)plan.n_parts * plan.waves * plan.max_tiles * plan.tile * plan.tile;
  hipLaunchKernelGGL(k_syrk_reduce, dim3((unsigned)((per_chunk + 63) / 64)), dim3(1024), 0, st,
                     partials, plan.d_tiles, plan.n_parts, plan.waves, plan.max_tiles,
                     plan.n_chunks, plan.tile, plan.m_cols, xx_full);
}

}  // namespace cy3q
