// Yukawa triple overlaps on the GPU (FP64).
//
// Reference: route34/yukawa_overlap_real.rs. sz_weights (:222-268), triple_overlap_cpu
// (:271-315), bootstrap_one (:318-357) and the CUDA source string of yukawa_overlap_real_gpu.rs
// (:105-158), where one block per ordered triple (i, j, k) re-reads all n points n_modes^3
// times. Here the unordered triples a <= b <= c are one complex contraction over the points,
//   Y[(a, b)][c] = sum_p P_p(a, b) psi_c(p),   P_p(a, b) = w_p conj(Omega_p) psi_a(p) psi_b(p),
// a GEMM of the pair products against the mode values on the FP64 MFMA. The reference's skips
// become zeros: a point with a non-finite weight or Omega contributes nothing (its factor and
// its mode values are zeroed), and a mode value with a non-finite real part is zeroed at its
// point, which removes exactly the point's contributions to the triples containing that mode.
// The sums run over point chunks with per-chunk partials added in chunk order (deterministic).
#include "cy3q_lcg.h"

namespace cy3q {

static constexpr int YK_MAXM = 64;     // max modes
static constexpr int YK_PTS = 16;      // points per stage (the MFMA k dimension, 4 steps)
static constexpr int YK_IMAX = 32;     // row tiles (of 8 pairs) per part
static constexpr int YK_TPW = 8;       // accumulator tiles per wave (8 waves per workgroup)
static constexpr int YK_RGRP = 16;     // chunk groups of the partial reduction

// psi_v[p][2m], [2m + 1] = (re, im) of mode m at point p (element (m, p) of `modes` at
// modes[m * ms + p * ps]), zero padding to vld; g[p] = w_p conj(Omega_p) (omega == nullptr:
// Omega = 1). A point with a non-finite weight or Omega gets g = 0 and zero mode values. With
// check_psi, a mode value with a non-finite real part is zeroed (the reference skips the point
// for every triple holding that mode); without it, it is kept as it is.
__global__ void k_yuk_prep(const double2 *__restrict__ modes, int64_t ms, int64_t ps,
                           const double *__restrict__ w, const double *__restrict__ omega,
                           int64_t n, int check_psi, int nm, int vld, double *__restrict__ psi_v,
                           double2 *__restrict__ g) {
  // One thread per (point, mode slot), point-major: psi_v is written contiguously.
  const int half = vld / 2;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * half) return;
  const int64_t p = e / half;
  const int m = (int)(e - p * half);
  const double wp = w[p];
  const double orr = omega ? omega[2 * p] : 1.0, oi = omega ? omega[2 * p + 1] : 0.0;
  const bool ok = isfinite(wp) && isfinite(orr) && isfinite(oi);
  if (m == 0) g[p] = ok ? make_double2(wp * orr, -(wp * oi)) : make_double2(0.0, 0.0);
  double2 v = make_double2(0.0, 0.0);
  if (m < nm && ok) {
    v = modes[m * ms + p * ps];
    if (check_psi && !isfinite(v.x)) v = make_double2(0.0, 0.0);
  }
  reinterpret_cast<double2 *>(psi_v)[e] = v;
}

// One workgroup per (point chunk, part). A part is a list of 16 x 16 tiles (I, J) of the real
// product U^T V over the points, U = [Re P, Im P] per pair (columns 2 q, 2 q + 1 for pair q),
// V = psi_v, with row tiles I in [ilo, ihi). Per stage of 16 points: the part's pair products
// into LDS (thread tid takes pair slot tid % 256 at points tid / 256, + 2, ...), then 4 f64
// 16 x 16 x 4 MFMAs per tile; wave w holds tiles w, w + 8, ... of the part. The next stage's
// 16 rows of psi_v (one contiguous block) are copied into the other of two LDS slots with
// 16-byte global_load_lds copies while the MFMAs run. Rows past the chunk end are read from
// the last row in range and zeroed in LDS before use, and their g is zero.
// LDS: U [16][uld], V [2][slot of 16 rows of vs doubles], g [2][16] (row strides = 16 mod 32
// doubles, so the four k rows of an operand read fall in different bank halves).
// MFMA f64 16x16x4: A[i][k] at lane i + 16k, B[k][j] at lane j + 16k, D[i][j] at lane
// j + 16 (i % 4), element i / 4.
__global__ void __launch_bounds__(512, 4) k_yuk_pairs(
    const double *__restrict__ psi_v, const double2 *__restrict__ g, int64_t n, int vs, int uld,
    int npair, const short2 *__restrict__ pairs, const short2 *__restrict__ tiles,
    const int4 *__restrict__ parts, int ntiles, int64_t chunk, double *__restrict__ partials) {
  extern __shared__ double2 yk_lds[];
  double *su = reinterpret_cast<double *>(yk_lds);
  const int vbytes = YK_PTS * vs * 8;                // bytes of one stage's V rows
  const int slot = (vbytes + 1023) / 1024 * 1024;   // LDS bytes per V slot
  char *svb = reinterpret_cast<char *>(su + YK_PTS * uld);
  double2 *sg = reinterpret_cast<double2 *>(svb + 2 * slot);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int64_t c = blockIdx.x;
  const int4 part = parts[blockIdx.y];  // tile_begin, tile_end, ilo, ihi
  const int q0 = 8 * part.z;
  const int nslot = 8 * (part.w - part.z);  // pair slots of the part's row tiles (<= 256)
  // Per tile: the lane's U column and V column, packed as (U column) | (V column) << 16.
  int tcol[YK_TPW];
  int nt = 0;
#pragma unroll
  for (int u = 0; u < YK_TPW; u++) {
    const int t = part.x + wave + 8 * u;
    const short2 e = t < part.y ? tiles[t] : make_short2(-1, -1);
    tcol[u] = (16 * (e.x - part.z) + (lane & 15)) | ((16 * e.y + (lane & 15)) << 16);
    if (e.x >= 0) nt = u + 1;
  }
  // This thread's pair slot: modes (pa, pb) of pair q0 + ql, pa < 0 for a slot past the pairs.
  const int ql = tid & 255, kh = tid >> 8;
  int pa = -1, pb = 0;
  if (ql < nslot && q0 + ql < npair) {
    const short2 ab = pairs[q0 + ql];
    pa = ab.x;
    pb = ab.y;
  }
  double4_t acc[YK_TPW];
#pragma unroll
  for (int u = 0; u < YK_TPW; u++) acc[u] = (double4_t){0.0, 0.0, 0.0, 0.0};
  const int64_t p0 = c * chunk, p1 = p0 + chunk < n ? p0 + chunk : n;
  const int nstage = (int)((p1 - p0 + YK_PTS - 1) / YK_PTS);
  const int ncopy = slot / 1024;  // 1 KB copies per stage (<= 24), wave w takes w, w + 8, ...
  const char *vbase = reinterpret_cast<const char *>(psi_v);
  const int64_t last_row = n - 1;
  auto issue = [&](int64_t s0, int buf) {
    for (int u = wave; u < ncopy; u += 8) {
      int o = 1024 * u + 16 * lane;
      o = o < vbytes - 16 ? o : vbytes - 16;
      int64_t r = s0 + o / (8 * vs);
      const int col_b = o % (8 * vs);
      r = r < last_row ? r : last_row;
      __builtin_amdgcn_global_load_lds((const void *)(vbase + r * (int64_t)(8 * vs) + col_b),
                                       (__attribute__((address_space(3))) void *)(
                                           svb + buf * slot + 1024 * u),
                                       16, 0, 0);
    }
  };
  double2 rg = make_double2(0.0, 0.0);
  auto fetch_g = [&](int64_t s0) {
    rg = (tid < YK_PTS && s0 + tid < p1) ? g[s0 + tid] : make_double2(0.0, 0.0);
  };
  issue(p0, 0);
  fetch_g(p0);
  if (tid < YK_PTS) sg[tid] = rg;
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0x70);  // vmcnt(0), lgkmcnt(0)
  __syncthreads();
  asm volatile("" ::: "memory");
  for (int s = 0; s < nstage; s++) {
    const int buf = s & 1;
    const int64_t s0 = p0 + (int64_t)YK_PTS * s;
    double *vb = reinterpret_cast<double *>(svb + buf * slot);
    const int cnt = p1 - s0 < YK_PTS ? (int)(p1 - s0) : YK_PTS;
    if (cnt < YK_PTS) {  // the chunk's last, partial stage: zero the rows past the end
      for (int e = cnt * vs + tid; e < YK_PTS * vs; e += 512) vb[e] = 0.0;
      __syncthreads();
    }
    if (ql < nslot) {
#pragma unroll 2
      for (int j = 0; j < YK_PTS / 2; j++) {
        const int k = kh + 2 * j;
        double2 out = make_double2(0.0, 0.0);
        if (pa >= 0) {
          const double2 x = *reinterpret_cast<const double2 *>(vb + k * vs + 2 * pa);
          const double2 y = *reinterpret_cast<const double2 *>(vb + k * vs + 2 * pb);
          const double2 gk = sg[buf * YK_PTS + k];
          const double xr = x.x * y.x - x.y * y.y, xi = x.x * y.y + x.y * y.x;
          out = make_double2(gk.x * xr - gk.y * xi, gk.x * xi + gk.y * xr);
        }
        *reinterpret_cast<double2 *>(su + k * uld + 2 * ql) = out;
      }
    }
    __syncthreads();
    if (s + 1 < nstage) {
      issue(s0 + YK_PTS, buf ^ 1);
      fetch_g(s0 + YK_PTS);
    }
#pragma unroll 1
    for (int kk = 0; kk < YK_PTS / 4; kk++) {
      const int kr = 4 * kk + (lane >> 4);
      const double *ua = su + kr * uld, *vr = vb + kr * vs;
#pragma unroll
      for (int u = 0; u < YK_TPW; u++) {
        if (u < nt) {
          const double a = ua[tcol[u] & 0xffff];
          const double b = vr[tcol[u] >> 16];
          acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[u], 0, 0, 0);
        }
      }
    }
    if (s + 1 < nstage && tid < YK_PTS) sg[(buf ^ 1) * YK_PTS + tid] = rg;
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0x70);
    __syncthreads();
    asm volatile("" ::: "memory");
  }
#pragma unroll
  for (int u = 0; u < YK_TPW; u++) {
    if (u < nt) {
      double *o = partials + (c * ntiles + part.x + wave + 8 * u) * 256;
#pragma unroll
      for (int e = 0; e < 4; e++) o[e * 64 + lane] = acc[u][e];
    }
  }
}

// cfull[16 I + r][16 J + col] = sum over the chunks of the tile partials. A workgroup takes 64
// consecutive partial entries; its 16 groups of 64 threads sum chunks grp, grp + 16, ... and
// the 16 group sums are added in group order (deterministic).
__global__ void __launch_bounds__(1024) k_yuk_creduce(const double *__restrict__ partials,
                                                      int64_t n_chunks, int ntiles,
                                                      const short2 *__restrict__ tiles, int vld,
                                                      double *__restrict__ cfull) {
  __shared__ double red[YK_RGRP][64];
  const int tid = threadIdx.x, grp = tid >> 6;
  const int64_t idx = (int64_t)blockIdx.x * 64 + (tid & 63);
  const int64_t tot = (int64_t)ntiles * 256;
  double s = 0.0;
  if (idx < tot)
    for (int64_t c = grp; c < n_chunks; c += YK_RGRP) s += partials[c * tot + idx];
  red[grp][tid & 63] = s;
  __syncthreads();
  if (tid >= 64 || idx >= tot) return;
  double r = 0.0;
  for (int q = 0; q < YK_RGRP; q++) r += red[q][tid];
  const int t = (int)(idx / 256), el = (int)(idx % 256);
  const int e = el / 64, ln = el % 64;
  const short2 ij = tiles[t];
  const int row = 16 * ij.x + 4 * e + ln / 16, col = 16 * ij.y + ln % 16;
  cfull[(int64_t)row * vld + col] = r;
}

// Y_abc = sqrt(h_a) sqrt(h_b) sqrt(h_c) (Re, Im of sum_p P_p(a, b) psi_c(p)) from cfull (rows
// 2 q, 2 q + 1 hold Re P, Im P of pair q; columns 2 c, 2 c + 1 Re, Im psi_c); all permutations
// into y_full when y_full != nullptr.
__global__ void k_yuk_assemble(const double *__restrict__ cfull, int vld, int nm,
                               const short4 *__restrict__ trip, int n_trip,
                               const double *__restrict__ hdiag, double2 *__restrict__ y_tri,
                               double2 *__restrict__ y_full) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_trip) return;
  const short4 e = trip[t];
  const int a = e.x, b = e.y, cc = e.z;
  const int q = a * nm - a * (a - 1) / 2 + (b - a);
  const double *r0 = cfull + (int64_t)(2 * q) * vld, *r1 = r0 + vld;
  const double yr = r0[2 * cc] - r1[2 * cc + 1], yi = r0[2 * cc + 1] + r1[2 * cc];
  const double sc = hdiag ? sqrt(hdiag[a]) * sqrt(hdiag[b]) * sqrt(hdiag[cc]) : 1.0;
  const double2 r = make_double2(yr * sc, yi * sc);
  if (y_tri) y_tri[t] = r;
  if (y_full) {
    const int ix[3] = {a, b, cc};
    const int pm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    for (int k = 0; k < 6; k++)
      y_full[((int64_t)ix[pm[k][0]] * nm + ix[pm[k][1]]) * nm + ix[pm[k][2]]] = r;
  }
}

int yukawa_max_modes() { return YK_MAXM; }

static int yk_stride(int cols) { return (cols / 16) % 2 == 1 ? cols : cols + 16; }

YukPlanShape yukawa_plan_shape(int nm, int64_t n) {
  YukPlanShape sh;
  sh.nm = nm;
  sh.npair = nm * (nm + 1) / 2;
  sh.vld = (2 * nm + 15) / 16 * 16;
  const int n_it = (2 * sh.npair + 15) / 16, n_jt = sh.vld / 16;
  for (int i = 0; i < nm; i++)
    for (int j = i; j < nm; j++) sh.pairs.push_back(make_short2((short)i, (short)j));
  for (int it = 0; it < n_it; it++) {
    int bmin = nm;
    for (int q = 8 * it; q < std::min(8 * it + 8, sh.npair); q++)
      bmin = std::min(bmin, (int)sh.pairs[q].y);
    for (int jt = 0; jt < n_jt; jt++)
      if (8 * jt + 7 >= bmin) sh.tiles.push_back(make_short2((short)it, (short)jt));
  }
  sh.ntiles = (int)sh.tiles.size();
  // Parts: consecutive tiles, at most 8 * YK_TPW tiles and YK_IMAX row tiles each.
  int max_rows = 0;
  for (int t0 = 0; t0 < sh.ntiles;) {
    int t1 = t0;
    const int ilo = sh.tiles[t0].x;
    while (t1 < sh.ntiles && t1 - t0 < 8 * YK_TPW && sh.tiles[t1].x - ilo < YK_IMAX) t1++;
    // Keep whole row tiles in a part when the part holds more than one.
    if (t1 < sh.ntiles && sh.tiles[t1].x == sh.tiles[t1 - 1].x && sh.tiles[t0].x != sh.tiles[t1].x)
      while (sh.tiles[t1 - 1].x == sh.tiles[t1].x) t1--;
    const int ihi = sh.tiles[t1 - 1].x + 1;
    sh.parts.push_back(make_int4(t0, t1, ilo, ihi));
    max_rows = std::max(max_rows, ihi - ilo);
    t0 = t1;
  }
  sh.nparts = (int)sh.parts.size();
  sh.uld = yk_stride(16 * max_rows);
  sh.vldp = yk_stride(sh.vld);
  const size_t vslot = ((size_t)YK_PTS * sh.vldp * 8 + 1023) / 1024 * 1024;
  sh.lds_bytes = sizeof(double) * (size_t)YK_PTS * sh.uld + 2 * vslot + sizeof(double2) * 2 * YK_PTS;
  // About 512 workgroups per launch (two per CU), and at least 64 points per chunk.
  const int64_t want = std::max<int64_t>(1, 512 / sh.nparts);
  int64_t chunk = (n + want - 1) / want;
  chunk = std::max<int64_t>(64, (chunk + YK_PTS - 1) / YK_PTS * YK_PTS);
  sh.chunk = chunk;
  sh.n_chunks = std::max<int64_t>(1, (n + chunk - 1) / chunk);
  for (int i = 0; i < nm; i++)
    for (int j = i; j < nm; j++)
      for (int k = j; k < nm; k++) sh.trip.push_back(make_short4((short)i, (short)j, (short)k, 0));
  sh.n_trip = (int)sh.trip.size();
  sh.cfull_rows = 16 * n_it;
  return sh;
}

void launch_yuk_prep(hipStream_t st, const double2 *modes, int64_t mode_stride,
                     int64_t point_stride, const double *w, const double *omega, int64_t n,
                     int check_psi, int nm, int vld, double *psi_v, double2 *g) {
  if (n <= 0) return;
  const int64_t tot = n * (vld / 2);
  hipLaunchKernelGGL(k_yuk_prep, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, modes,
                     mode_stride, point_stride, w, omega, n, check_psi, nm, vld, psi_v, g);
}

void launch_yuk_contract(hipStream_t st, const YukPlan &pl, int64_t n, const double *hdiag,
                         double2 *y_tri, double2 *y_full) {
  const int64_t chunk = pl.chunk;
  const int64_t nc = n > 0 ? (n + chunk - 1) / chunk : 0;
  if (nc > 0)
    hipLaunchKernelGGL(k_yuk_pairs, dim3((unsigned)nc, (unsigned)pl.nparts), dim3(512),
                       pl.lds_bytes, st, pl.psi_v, pl.g, n, pl.vldp, pl.uld, pl.npair, pl.pairs,
                       pl.tiles, pl.parts, pl.ntiles, chunk, pl.partials);
  const int64_t tot = (int64_t)pl.ntiles * 256;
  hipLaunchKernelGGL(k_yuk_creduce, dim3((unsigned)((tot + 63) / 64)), dim3(1024), 0, st,
                     pl.partials, nc, pl.ntiles, pl.tiles, pl.vld, pl.cfull);
  hipLaunchKernelGGL(k_yuk_assemble, dim3((unsigned)((pl.n_trip + 255) / 256)), dim3(256), 0, st,
                     pl.cfull, pl.vld, pl.nm, pl.trip, pl.n_trip, hdiag, y_tri, y_full);
}

// ---------------------------------------------------------------------------------------------
// sz_weights (:222-268) on the device: v_p = w_p / |Omega_p|^2 for a usable point, then the sum,
// the renormalisation and the uniformity score 1 / (1 + sd/mean) over all n entries. One
// workgroup, fixed summation order. out[0] = score.
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_sz_weights(const double *__restrict__ w,
                                                     const double *__restrict__ omega, int64_t n,
                                                     double *__restrict__ sz,
                                                     double *__restrict__ out) {
  __shared__ double s0[1024], s1[1024];
  const int tid = threadIdx.x;
  double tot = 0.0, bad = 0.0;
  for (int64_t a = tid; a < n; a += 1024) {
    const double wa = w[a];
    const double o2 = omega[2 * a] * omega[2 * a] + omega[2 * a + 1] * omega[2 * a + 1];
    double v = 0.0;
    if (!isfinite(wa) || wa <= 0.0 || !isfinite(o2) || o2 <= 0.0)
      bad += 1.0;
    else
      v = wa / o2;
    sz[a] = v;
    tot += v;
  }
  s0[tid] = tot;
  s1[tid] = bad;
  __syncthreads();
  for (int off = 512; off > 0; off >>= 1) {
    if (tid < off) {
      s0[tid] += s0[tid + off];
      s1[tid] += s1[tid + off];
    }
    __syncthreads();
  }
  const double total = s0[0], n_bad = s1[0];
  __syncthreads();
  const double inv = total > 0.0 ? 1.0 / total : 1.0;
  double sm = 0.0;
  for (int64_t a = tid; a < n; a += 1024) {
    const double v = total > 0.0 ? sz[a] * inv : sz[a];
    sz[a] = v;
    sm += v;
  }
  s0[tid] = sm;
  __syncthreads();
  for (int off = 512; off > 0; off >>= 1) {
    if (tid < off) s0[tid] += s0[tid + off];
    __syncthreads();
  }
  const double mean = s0[0] / (double)n;
  __syncthreads();
  double var = 0.0;
  for (int64_t a = tid; a < n; a += 1024) {
    const double d = sz[a] - mean;
    var += d * d;
  }
  s0[tid] = var;
  __syncthreads();
  for (int off = 512; off > 0; off >>= 1) {
    if (tid < off) s0[tid] += s0[tid + off];
    __syncthreads();
  }
  if (tid == 0) {
    const double n_eff = (double)n - n_bad;
    double score = 0.0;
    if (n_eff > 1.0) {
      const double vv = s0[0] / (double)n;
      const double cv = mean > 0.0 ? sqrt(vv) / mean : INFINITY;
      score = fmin(fmax(1.0 / (1.0 + cv), 0.0), 1.0);
    }
    out[0] = score;
  }
}

void launch_sz_weights(hipStream_t st, const double *w, const double *omega, int64_t n, double *sz,
                       double *out) {
  hipLaunchKernelGGL(k_sz_weights, dim3(1), dim3(1024), 0, st, w, omega, n, sz, out);
}

// w_out[p] = count[p] * w[p] (a bootstrap resample as multiplicities), or w[p] / total over the
// first n points when count == nullptr (the half-sample renormalisation of :466-476).
__global__ void k_scale_weights(const double *__restrict__ w, const int32_t *__restrict__ count,
                                const double *__restrict__ total, int64_t n,
                                double *__restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  if (count)
    out[p] = (double)count[p] * w[p];
  else
    out[p] = total[0] > 0.0 ? w[p] / total[0] : w[p];
}

void launch_scale_weights(hipStream_t st, const double *w, const int32_t *count,
                          const double *total, int64_t n, double *out) {
  hipLaunchKernelGGL(k_scale_weights, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, w,
                     count, total, n, out);
}

// Real-ified B operand of psi = S C: b[2 s][2 a] = Re C_sa, b[2 s][2 a + 1] = Im C_sa,
// b[2 s + 1][2 a] = -Im C_sa, b[2 s + 1][2 a + 1] = Re C_sa for sections s < nsec and modes a < nm,
// zero elsewhere ([krows][ncols]).
__global__ void k_modes_bmat(const double2 *__restrict__ coef, int nsec, int nm, int64_t krows,
                             int64_t ncols, double *__restrict__ b) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= krows * ncols) return;
  const int64_t r = e / ncols, col = e % ncols;
  const int64_t sidx = r / 2, a = col / 2;
  double v = 0.0;
  if (sidx < nsec && a < nm) {
    const double2 cf = coef[sidx * nm + a];
    v = (r & 1) == 0 ? ((col & 1) == 0 ? cf.x : cf.y) : ((col & 1) == 0 ? -cf.y : cf.x);
  }
  b[e] = v;
}

void launch_modes_bmat(hipStream_t st, const double2 *coef, int nsec, int nm, int64_t krows,
                       int64_t ncols, double *b) {
  const int64_t tot = krows * ncols;
  hipLaunchKernelGGL(k_modes_bmat, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, coef,
                     nsec, nm, krows, ncols, b);
}

// c_p = w_p |Omega_p|^2 = w_p exp(log|Omega_p|^2) for a usable point, NaN otherwise (so that the
// triple kernel skips it).
__global__ void k_measure_w_omega2(const double *__restrict__ w, const double *__restrict__ log_om,
                                   int64_t n, double *__restrict__ c) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double wp = w[p], lo = log_om[p];
  c[p] = (isfinite(wp) && wp > 0.0 && isfinite(lo)) ? wp * exp(lo) : NAN;
}

void launch_measure_w_omega2(hipStream_t st, const double *w, const double *log_om, int64_t n,
                             double *c) {
  hipLaunchKernelGGL(k_measure_w_omega2, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, w,
                     log_om, n, c);
}

// Bootstrap resample b as multiplicities: draw q (q < n) of the lib.rs LCG seeded with
// seed_b gives index floor(u n) (clamped to n - 1); count[index] += 1. Integer atomics, so the
// counts do not depend on the order of the threads. count must start at zero.
__global__ void k_bootstrap_counts(uint64_t seed_b, int64_t n, int32_t *__restrict__ count) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const uint64_t s = lcg_jump(seed_b * LCG_MUL + 1ULL, (uint64_t)q + 1ULL);
  const uint64_t bits = (s >> 11) & ((1ULL << 53) - 1ULL);
  const double u = (double)bits / (double)(1ULL << 53);
  int64_t idx = (int64_t)(u * (double)n);
  if (idx >= n) idx = n - 1;
  atomicAdd(&count[idx], 1);
}

void launch_bootstrap_counts(hipStream_t st, uint64_t seed_b, int64_t n, int32_t *count) {
  hipLaunchKernelGGL(k_bootstrap_counts, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     seed_b, n, count);
}

// out[j] = (g_2j + i g_2j+1) / sqrt(2), g_k the k-th Box-Muller normal of the lib.rs LCG seeded
// with seed (normal k consumes draws 2k and 2k+1).
__global__ void k_random_coeffs(uint64_t seed, int64_t count, double2 *__restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  Lcg r;
  r.s = lcg_jump(seed * LCG_MUL + 1ULL, 4ULL * (uint64_t)j);
  const double g0 = r.next_normal();
  const double g1 = r.next_normal();
  out[j] = make_double2(g0 * M_SQRT1_2, g1 * M_SQRT1_2);
}

void launch_random_coeffs(hipStream_t st, uint64_t seed, int64_t count, double2 *out) {
  hipLaunchKernelGGL(k_random_coeffs, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, seed,
                     count, out);
}

// acc[e] += |yb[e] - y[e]|^2, then unc[e] = sqrt(acc[e] / nbs).
__global__ void k_sq_dev_acc(const double2 *__restrict__ yb, const double2 *__restrict__ y,
                             int64_t n3, double *__restrict__ acc) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n3) return;
  const double dr = yb[e].x - y[e].x, di = yb[e].y - y[e].y;
  acc[e] += dr * dr + di * di;
}

__global__ void k_sqrt_mean(const double *__restrict__ acc, int64_t n3, double nbs,
                            double *__restrict__ unc) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n3) unc[e] = sqrt(acc[e] / nbs);
}

void launch_bootstrap_acc(hipStream_t st, const double2 *yb, const double2 *y, int64_t n3,
                          double *acc) {
  hipLaunchKernelGGL(k_sq_dev_acc, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, st, yb, y, n3,
                     acc);
}

void launch_bootstrap_final(hipStream_t st, const double *acc, int64_t n3, double nbs,
                            double *unc) {
  hipLaunchKernelGGL(k_sqrt_mean, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, st, acc, n3,
                     nbs, unc);
}

// out[0] = max_e |y - yh| / max_e |y| (0 when y == 0): the convergence ratio of :478-498.
__global__ void __launch_bounds__(256) k_conv_ratio(const double2 *__restrict__ y,
                                                    const double2 *__restrict__ yh, int64_t n3,
                                                    double *__restrict__ out) {
  __shared__ double sd[256], sf[256];
  double md = 0.0, mf = 0.0;
  for (int64_t e = threadIdx.x; e < n3; e += 256) {
    const double dr = y[e].x - yh[e].x, di = y[e].y - yh[e].y;
    md = fmax(md, sqrt(dr * dr + di * di));
    mf = fmax(mf, hypot(y[e].x, y[e].y));
  }
  sd[threadIdx.x] = md;
  sf[threadIdx.x] = mf;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      sd[threadIdx.x] = fmax(sd[threadIdx.x], sd[threadIdx.x + o]);
      sf[threadIdx.x] = fmax(sf[threadIdx.x], sf[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sf[0] > 0.0 ? sd[0] / sf[0] : 0.0;
}

void launch_conv_ratio(hipStream_t st, const double2 *y, const double2 *yh, int64_t n3,
                       double *out) {
  hipLaunchKernelGGL(k_conv_ratio, dim3(1), dim3(256), 0, st, y, yh, n3, out);
}

}  // namespace cy3q
