// Yukawa triple overlaps on the GPU (FP64).
//
// Reference: route34/yukawa_overlap_real.rs. sz_weights (:222-268), triple_overlap_cpu
// (:271-315), bootstrap_one (:318-357) and the CUDA source string of yukawa_overlap_real_gpu.rs
// (:105-158). There, one block per ordered triple (i, j, k) re-reads all n points n_modes^3 times.
// Here every point is read once per triple batch: a workgroup stages a slab of points (all
// modes, point-major) in LDS, and each thread accumulates a few unordered triples i <= j <= k
// over the slab in FP64 registers. The product order is the reference's
// (((conj(Omega) psi_i) psi_j) psi_k) w. Per-chunk partials are summed in chunk order, so the
// result is deterministic, and the symmetric tensor is filled from the unordered triples.
#include "cy3q_lcg.h"

namespace cy3q {

static constexpr int YK_THREADS = 256;
static constexpr int YK_TPB = 8;       // triples per thread per batch
static constexpr int YK_SLAB = 32;     // points per LDS slab
static constexpr int YK_MAXM = 64;     // max modes

struct cplx {
  double re, im;
};
__device__ inline cplx cmul(cplx a, cplx b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// Per-point factor f_p = w_p conj(Omega_p) (zero and flagged for a point the reference skips),
// and the mask of modes whose value has a non-finite real part at the point.
// omega == nullptr means Omega = 1 (a real measure in w).
__global__ void k_yuk_factor(const double *__restrict__ w, const double *__restrict__ omega,
                             const double2 *__restrict__ modes, int64_t ld, int nm, int64_t n,
                             int check_psi,
                             double2 *__restrict__ fac, unsigned long long *__restrict__ bad) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double wp = w[p];
  const double orr = omega ? omega[2 * p] : 1.0, oi = omega ? omega[2 * p + 1] : 0.0;
  const bool ok = isfinite(wp) && isfinite(orr) && isfinite(oi);
  unsigned long long mask = 0ull;
  if (check_psi)
    for (int i = 0; i < nm; i++)
      if (!isfinite(modes[(int64_t)i * ld + p].x)) mask |= 1ull << i;
  fac[p] = ok ? make_double2(orr, -oi) : make_double2(0.0, 0.0);
  // The weight multiplies last, as in the reference, so it stays a separate input. A skipped
  // point gets every mask bit set, which makes every triple skip it.
  bad[p] = ok ? mask : ~0ull;
}

// partials[chunk][t] = sum over the chunk's points of (((f psi_i) psi_j) psi_k) * w.
__global__ void __launch_bounds__(YK_THREADS) k_yuk_triples(
    const double2 *__restrict__ modes, int64_t ld, const double2 *__restrict__ fac,
    const double *__restrict__ w, const unsigned long long *__restrict__ bad, int nm, int64_t n,
    int64_t chunk, const short4 *__restrict__ trip, int n_trip, double2 *__restrict__ partials) {
  __shared__ double2 s_psi[YK_SLAB][YK_MAXM];
  __shared__ double2 s_f[YK_SLAB];
  __shared__ double s_w[YK_SLAB];
  __shared__ unsigned long long s_bad[YK_SLAB];
  const int tid = threadIdx.x;
  const int64_t c = blockIdx.x;
  const int tb = blockIdx.y;
  const int64_t p0 = c * chunk, p1 = min(p0 + chunk, n);
  int ti[YK_TPB], tj[YK_TPB], tk[YK_TPB];
  cplx acc[YK_TPB];
#pragma unroll
  for (int q = 0; q < YK_TPB; q++) {
    const int t = (tb * YK_TPB + q) * YK_THREADS + tid;
    short4 e = t < n_trip ? trip[t] : make_short4(-1, -1, -1, 0);
    ti[q] = e.x;
    tj[q] = e.y;
    tk[q] = e.z;
    acc[q] = {0.0, 0.0};
  }
  for (int64_t s0 = p0; s0 < p1; s0 += YK_SLAB) {
    const int cnt = (int)min((int64_t)YK_SLAB, p1 - s0);
    __syncthreads();
    for (int e = tid; e < YK_SLAB * nm; e += YK_THREADS) {
      const int m = e / YK_SLAB, q = e % YK_SLAB;
      s_psi[q][m] = q < cnt ? modes[(int64_t)m * ld + s0 + q] : make_double2(0.0, 0.0);
    }
    if (tid < YK_SLAB) {
      const bool in = tid < cnt;
      s_f[tid] = in ? fac[s0 + tid] : make_double2(0.0, 0.0);
      s_w[tid] = in ? w[s0 + tid] : 0.0;
      s_bad[tid] = in ? bad[s0 + tid] : ~0ull;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < YK_TPB; q++) {
      if (ti[q] < 0) continue;
      const unsigned long long sel = (1ull << ti[q]) | (1ull << tj[q]) | (1ull << tk[q]);
      for (int pp = 0; pp < cnt; pp++) {
        if (s_bad[pp] & sel) continue;
        const double2 a = s_psi[pp][ti[q]], b = s_psi[pp][tj[q]], d = s_psi[pp][tk[q]];
        const double2 f = s_f[pp];
        cplx v = cmul(cmul(cmul({f.x, f.y}, {a.x, a.y}), {b.x, b.y}), {d.x, d.y});
        const double wp = s_w[pp];
        acc[q].re += v.re * wp;
        acc[q].im += v.im * wp;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < YK_TPB; q++) {
    const int t = (tb * YK_TPB + q) * YK_THREADS + tid;
    if (t < n_trip) partials[c * n_trip + t] = make_double2(acc[q].re, acc[q].im);
  }
}

// y_tri[t] = sqrt(h_i) sqrt(h_j) sqrt(h_k) * sum_c partials[c][t] (chunk order); the full
// tensor (all permutations) when y_full != nullptr.
__global__ void k_yuk_reduce(const double2 *__restrict__ partials, int64_t n_chunks, int n_trip,
                             const short4 *__restrict__ trip, const double *__restrict__ hdiag,
                             int nm, double2 *__restrict__ y_tri, double2 *__restrict__ y_full) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_trip) return;
  double sr = 0.0, si = 0.0;
  for (int64_t c = 0; c < n_chunks; c++) {
    const double2 v = partials[c * n_trip + t];
    sr += v.x;
    si += v.y;
  }
  const short4 e = trip[t];
  const double sc = hdiag ? sqrt(hdiag[e.x]) * sqrt(hdiag[e.y]) * sqrt(hdiag[e.z]) : 1.0;
  const double2 r = make_double2(sr * sc, si * sc);
  if (y_tri) y_tri[t] = r;
  if (y_full) {
    const int ix[3] = {e.x, e.y, e.z};
    const int pm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    for (int q = 0; q < 6; q++)
      y_full[((int64_t)ix[pm[q][0]] * nm + ix[pm[q][1]]) * nm + ix[pm[q][2]]] = r;
  }
}

int yukawa_max_modes() { return YK_MAXM; }

int64_t yukawa_chunks(int64_t n, int n_trip) {
  const int tbatches = (n_trip + YK_THREADS * YK_TPB - 1) / (YK_THREADS * YK_TPB);
  int64_t want = std::max<int64_t>(1, 2048 / tbatches);
  int64_t by_size = std::max<int64_t>(1, (n + 127) / 128);
  return std::min(want, by_size);
}

void launch_yukawa(hipStream_t st, const double2 *modes, int64_t ld, const double *w,
                   const double *omega, int nm, int64_t n, int check_psi, const short4 *trip, int n_trip,
                   const double *hdiag, int64_t n_chunks, double2 *fac, unsigned long long *bad,
                   double2 *partials, double2 *y_tri, double2 *y_full) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_yuk_factor, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, w, omega,
                     modes, ld, nm, n, check_psi, fac, bad);
  const int64_t chunk = (n + n_chunks - 1) / n_chunks;
  const int64_t nc = (n + chunk - 1) / chunk;
  const int tbatches = (n_trip + YK_THREADS * YK_TPB - 1) / (YK_THREADS * YK_TPB);
  hipLaunchKernelGGL(k_yuk_triples, dim3((unsigned)nc, (unsigned)tbatches), dim3(YK_THREADS), 0,
                     st, modes, ld, fac, w, bad, nm, n, chunk, trip, n_trip, partials);
  hipLaunchKernelGGL(k_yuk_reduce, dim3((unsigned)((n_trip + 255) / 256)), dim3(256), 0, st,
                     partials, nc, n_trip, trip, hdiag, nm, y_tri, y_full);
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

// modes[a][p] = sum_b C[b][a] s_b(p) from the real-ified section rows sv [n][m_cols]
// (ψ = S C of the random-mode construction); one thread per (p, a).
__global__ void k_modes_from_sections(const double *__restrict__ sv, int64_t n, int64_t m_cols,
                                      int nsec, const double2 *__restrict__ coef, int nm,
                                      double2 *__restrict__ modes, int64_t ld) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * nm) return;
  const int a = (int)(e / n);
  const int64_t p = e % n;
  const double *row = sv + p * m_cols;
  double sr = 0.0, si = 0.0;
  for (int b = 0; b < nsec; b++) {
    const double xr = row[2 * b], xi = row[2 * b + 1];
    const double2 cb = coef[(int64_t)b * nm + a];
    sr += xr * cb.x - xi * cb.y;
    si += xr * cb.y + xi * cb.x;
  }
  modes[(int64_t)a * ld + p] = make_double2(sr, si);
}

void launch_modes_from_sections(hipStream_t st, const double *sv, int64_t n, int64_t m_cols,
                                int nsec, const double2 *coef, int nm, double2 *modes,
                                int64_t ld) {
  const int64_t tot = n * nm;
  hipLaunchKernelGGL(k_modes_from_sections, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                     sv, n, m_cols, nsec, coef, nm, modes, ld);
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
