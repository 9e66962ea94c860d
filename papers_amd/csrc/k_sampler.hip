// Shiffman–Zelditch line-intersection sampler on the Fermat / Dwork quintic, one line per lane.
//
// Follows sample_quintic_points_sz (cy3/src/quintic.rs:3633-3785) and
// sample_deformed_quintic_points_sz (quintic.rs:3810-3983). Each line consumes exactly 40
// LCG draws (20 Box–Muller normals, cy3/src/lib.rs:222-245), so line l starts from the LCG state
// advanced 40*l steps. The lane reaches that state by affine jump-ahead, which makes every line
// independent and reproduces the reference's sequential draws bit for bit. Accepted candidates
// are compacted in (line, root) order by an exclusive scan, which keeps the reference's point order.
//
// The Newton-projection sampler (sample_quintic_points, quintic.rs:3568-3595, and
// sample_deformed_quintic_points, :3448-3482) runs one try per lane the same way: a try consumes
// 20 draws (10 normals), so try t starts 20*t draws into the stream.
#include "cy3q_lcg.h"
#include "cy3q_point.h"

namespace cy3q {

__device__ inline void quintic_value(const double *z, int dwork, double psi, double &fre,
                                     double &fim) {
  double re = 0.0, im = 0.0;
  for (int j = 0; j < 5; j++) {
    double r = z[2 * j], i = z[2 * j + 1];
    double r2 = r * r, i2 = i * i;
    double z2r = r2 - i2, z2i = 2.0 * r * i;
    double z4r = z2r * z2r - z2i * z2i, z4i = 2.0 * z2r * z2i;
    re += r * z4r - i * z4i;
    im += r * z4i + i * z4r;
  }
  if (dwork && psi != 0.0) {
    double pr = 1.0, pi = 0.0;
    for (int j = 0; j < 5; j++) {
      double r = z[2 * j], i = z[2 * j + 1];
      double nr = pr * r - pi * i, ni = pr * i + pi * r;
      pr = nr;
      pi = ni;
    }
    double sc = -5.0 * psi;
    re += sc * pr;
    im += sc * pi;
  }
  fre = re;
  fim = im;
}

__global__ void __launch_bounds__(256) k_sz_lines(uint64_t seed, int64_t line0, int64_t n_lines,
                                                  int dwork, double psi, int dk_iters,
                                                  int newton_polish, double *__restrict__ cand,
                                                  int32_t *__restrict__ acc) {
  int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n_lines) return;
  int64_t line = line0 + l;
  Lcg rng;
  rng.s = lcg_jump(seed * LCG_MUL + 1ULL, 40ULL * (uint64_t)line);
  double a[5][2], b[5][2];
  for (int i = 0; i < 5; i++) {
    a[i][0] = rng.next_normal();
    a[i][1] = rng.next_normal();
    b[i][0] = rng.next_normal();
    b[i][1] = rng.next_normal();
  }
  const double bin[6] = {1.0, 5.0, 10.0, 10.0, 5.0, 1.0};
  double c[6][2];
  for (int j = 0; j < 6; j++) c[j][0] = c[j][1] = 0.0;
  for (int i = 0; i < 5; i++) {
    double ap[6][2], bp[6][2];
    ap[0][0] = 1.0;
    ap[0][1] = 0.0;
    bp[0][0] = 1.0;
    bp[0][1] = 0.0;
    for (int e = 1; e <= 5; e++) {
      double pr = ap[e - 1][0], pi = ap[e - 1][1];
      ap[e][0] = pr * a[i][0] - pi * a[i][1];
      ap[e][1] = pr * a[i][1] + pi * a[i][0];
      pr = bp[e - 1][0];
      pi = bp[e - 1][1];
      bp[e][0] = pr * b[i][0] - pi * b[i][1];
      bp[e][1] = pr * b[i][1] + pi * b[i][0];
    }
    for (int j = 0; j <= 5; j++) {
      double ar = ap[5 - j][0], ai = ap[5 - j][1], br = bp[j][0], bi = bp[j][1];
      double pr = ar * br - ai * bi, pi = ar * bi + ai * br;
      c[j][0] += bin[j] * pr;
      c[j][1] += bin[j] * pi;
    }
  }
  if (dwork) {
    double prod[6][2];
    for (int j = 0; j < 6; j++) prod[j][0] = prod[j][1] = 0.0;
    prod[0][0] = 1.0;
    for (int i = 0; i < 5; i++) {
      double nx[6][2];
      for (int j = 0; j < 6; j++) nx[j][0] = nx[j][1] = 0.0;
      for (int j = 0; j <= i; j++) {
        double pr = prod[j][0], pi = prod[j][1];
        nx[j][0] += pr * a[i][0] - pi * a[i][1];
        nx[j][1] += pr * a[i][1] + pi * a[i][0];
        nx[j + 1][0] += pr * b[i][0] - pi * b[i][1];
        nx[j + 1][1] += pr * b[i][1] + pi * b[i][0];
      }
      for (int j = 0; j < 6; j++) {
        prod[j][0] = nx[j][0];
        prod[j][1] = nx[j][1];
      }
    }
    double sc = 5.0 * psi;
    for (int j = 0; j <= 5; j++) {
      c[j][0] -= sc * prod[j][0];
      c[j][1] -= sc * prod[j][1];
    }
  }
  int32_t *accl = acc + 5 * l;
  double *cl = cand + 50 * l;
  double denom = c[5][0] * c[5][0] + c[5][1] * c[5][1];
  if (denom < 1e-300) {
    for (int k = 0; k < 5; k++) accl[k] = 0;
    return;
  }
  double ivr = c[5][0] / denom, ivi = -c[5][1] / denom;
  double p[6][2];
  for (int j = 0; j <= 5; j++) {
    p[j][0] = c[j][0] * ivr - c[j][1] * ivi;
    p[j][1] = c[j][0] * ivi + c[j][1] * ivr;
  }
  double radius = 1.0;
  for (int j = 0; j < 5; j++) {
    double mag = sqrt(p[j][0] * p[j][0] + p[j][1] * p[j][1]);
    double r = pow(mag, 1.0 / (double)(5 - j));
    if (r > radius) radius = r;
  }
  double rt[5][2];
  for (int k = 0; k < 5; k++) {
    double th = (2.0 * M_PI / 5.0) * (double)k + 0.4;
    rt[k][0] = radius * cos(th);
    rt[k][1] = radius * sin(th);
  }
  for (int it = 0; it < dk_iters; it++) {
    double pv[5][2];
    for (int k = 0; k < 5; k++) {
      pv[k][0] = rt[k][0];
      pv[k][1] = rt[k][1];
    }
    for (int k = 0; k < 5; k++) {
      double rr = pv[k][0], ri = pv[k][1];
      double vr = 1.0, vi = 0.0;
      for (int j = 4; j >= 0; j--) {
        double nr = vr * rr - vi * ri + p[j][0];
        double ni = vr * ri + vi * rr + p[j][1];
        vr = nr;
        vi = ni;
      }
      double dr = 1.0, di = 0.0;
      for (int j = 0; j < 5; j++) {
        if (j == k) continue;
        double xr = rr - pv[j][0], xi = ri - pv[j][1];
        double nr = dr * xr - di * xi, ni = dr * xi + di * xr;
        dr = nr;
        di = ni;
      }
      double dn = dr * dr + di * di;
      if (dn > 1e-300) {
        double ir = dr / dn, ii = -di / dn;
        rt[k][0] = rr - (vr * ir - vi * ii);
        rt[k][1] = ri - (vr * ii + vi * ir);
      }
    }
  }
  // Optional FP64 Newton polish of each root on the monic line polynomial (an addition to the
  // reference; off by default so the point set matches the reference's).
  for (int pass = 0; pass < newton_polish; pass++) {
    for (int k = 0; k < 5; k++) {
      double rr = rt[k][0], ri = rt[k][1];
      double vr = 1.0, vi = 0.0, dvr = 0.0, dvi = 0.0;
      for (int j = 4; j >= 0; j--) {
        double ndr = dvr * rr - dvi * ri + vr, ndi = dvr * ri + dvi * rr + vi;
        double nr = vr * rr - vi * ri + p[j][0], ni = vr * ri + vi * rr + p[j][1];
        dvr = ndr;
        dvi = ndi;
        vr = nr;
        vi = ni;
      }
      double dn = dvr * dvr + dvi * dvi;
      if (dn > 1e-300) {
        rt[k][0] = rr - (vr * dvr + vi * dvi) / dn;
        rt[k][1] = ri - (vi * dvr - vr * dvi) / dn;
      }
    }
  }
  for (int k = 0; k < 5; k++) {
    double tr = rt[k][0], ti = rt[k][1];
    double z[10];
    accl[k] = 0;
    for (int q = 0; q < 10; q++) z[q] = 0.0;
    if (isfinite(tr) && isfinite(ti)) {
      for (int i = 0; i < 5; i++) {
        z[2 * i] = a[i][0] + tr * b[i][0] - ti * b[i][1];
        z[2 * i + 1] = a[i][1] + tr * b[i][1] + ti * b[i][0];
      }
      double ns = 0.0;
      for (int q = 0; q < 10; q++) ns += z[q] * z[q];
      if (ns >= 1e-30) {
        double nrm = sqrt(ns);
        for (int q = 0; q < 10; q++) z[q] /= nrm;
        double fr, fi;
        quintic_value(z, dwork, psi, fr, fi);
        double fm = sqrt(fr * fr + fi * fi);
        accl[k] = fm > (dwork ? 1e-2 : 1e-3) ? 0 : 1;
      }
    }
    for (int q = 0; q < 10; q++) cl[10 * k + q] = z[q];
  }
}

void launch_sz_lines(hipStream_t st, uint64_t seed, int64_t line0, int64_t n_lines, int dwork,
                     double psi, int dk_iters, int newton_polish, double *cand, int32_t *acc) {
  if (n_lines <= 0) return;
  int64_t blocks = (n_lines + 255) / 256;
  hipLaunchKernelGGL(k_sz_lines, dim3((unsigned)blocks), dim3(256), 0, st, seed, line0, n_lines,
                     dwork, psi, dk_iters, newton_polish, cand, acc);
}

// ---------------------------------------------------------------------------------------
// Newton projection (newton_project_to_quintic, quintic.rs:3334-3389, and its deformed form
// newton_project_to_deformed_quintic, :3395-3443; psi = 0 is the Fermat form exactly): try t
// draws 10 normals, normalises them onto S^9 (a norm below 1e-12 consumes the try with no point)
// and takes up to 50 steps z <- (z - f conj(grad f)/|grad f|^2)/|.|, accepting at |f| < tol
// inside the loop or |f| < 10 tol after it, and rejecting at |grad f|^2 < 1e-14 or a vanishing
// norm. Operation order and divisions as the reference; the file builds with
// -ffp-contract=off.
// ---------------------------------------------------------------------------------------
__device__ inline bool newton_project(double *z, double psi, double tol) {
  const int dwork = psi != 0.0;
  for (int it = 0; it < 50; it++) {
    double fre, fim;
    quintic_value(z, dwork, psi, fre, fim);
    if (fre * fre + fim * fim < tol * tol) return true;
    double g[10];
    grad_quintic(z, psi, g);
    double gn = 0.0;
    for (int k = 0; k < 10; k++) gn += g[k] * g[k];
    if (gn < 1e-14) return false;
    const double scale = -1.0 / gn;
    for (int j = 0; j < 5; j++) {
      const double gr = g[2 * j], gi = g[2 * j + 1];
      const double dr = scale * (fre * gr + fim * gi);
      const double di = scale * (fim * gr - fre * gi);
      z[2 * j] += dr;
      z[2 * j + 1] += di;
    }
    double nsq = 0.0;
    for (int k = 0; k < 10; k++) nsq += z[k] * z[k];
    const double nrm = sqrt(nsq);
    if (nrm < 1e-12) return false;
    for (int k = 0; k < 10; k++) z[k] /= nrm;
  }
  double fre, fim;
  quintic_value(z, dwork, psi, fre, fim);
  const double t10 = 10.0 * tol;
  return fre * fre + fim * fim < t10 * t10;
}

__global__ void __launch_bounds__(256) k_newton_tries(uint64_t seed, int64_t try0, int64_t n_tries,
                                                      double psi, double tol,
                                                      double *__restrict__ cand,
                                                      int32_t *__restrict__ acc) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n_tries) return;
  Lcg rng{lcg_jump(lcg_seed(seed), 20ULL * (uint64_t)(try0 + l))};
  double z[10];
  for (int k = 0; k < 10; k++) z[k] = rng.next_normal();
  double nsq = 0.0;
  for (int k = 0; k < 10; k++) nsq += z[k] * z[k];
  const double nrm = sqrt(nsq);
  bool ok = false;
  if (!(nrm < 1e-12)) {
    for (int k = 0; k < 10; k++) z[k] /= nrm;
    ok = newton_project(z, psi, tol);
  }
  acc[l] = ok ? 1 : 0;
  for (int q = 0; q < 10; q++) cand[10 * l + q] = z[q];
}

void launch_newton_tries(hipStream_t st, uint64_t seed, int64_t try0, int64_t n_tries, double psi,
                         double tol, double *cand, int32_t *acc) {
  if (n_tries <= 0) return;
  hipLaunchKernelGGL(k_newton_tries, dim3((unsigned)((n_tries + 255) / 256)), dim3(256), 0, st,
                     seed, try0, n_tries, psi, tol, cand, acc);
}

// ---------------------------------------------------------------------------------------
// Exclusive scan of accept flags (int32 0/1) into int64 positions, three passes:
// per-1024 block scan, single-block scan of block sums, add of block offsets.
// ---------------------------------------------------------------------------------------
static constexpr int SCAN_BLOCK = 1024;

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_block(const int32_t *__restrict__ flag,
                                                           int64_t n, int64_t *__restrict__ pos,
                                                           int64_t *__restrict__ bsum) {
  __shared__ int64_t sh[SCAN_BLOCK];
  int64_t i = (int64_t)blockIdx.x * SCAN_BLOCK + threadIdx.x;
  int64_t v = (i < n) ? (int64_t)flag[i] : 0;
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < SCAN_BLOCK; off <<= 1) {
    int64_t t = (threadIdx.x >= (unsigned)off) ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) pos[i] = sh[threadIdx.x] - v;
  if (threadIdx.x == SCAN_BLOCK - 1) bsum[blockIdx.x] = sh[threadIdx.x];
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_sums(int64_t *__restrict__ bsum,
                                                          int64_t nb, int64_t *__restrict__ total) {
  __shared__ int64_t sh[SCAN_BLOCK];
  __shared__ int64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < nb; base += SCAN_BLOCK) {
    int64_t i = base + threadIdx.x;
    int64_t v = (i < nb) ? bsum[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_BLOCK; off <<= 1) {
      int64_t t = (threadIdx.x >= (unsigned)off) ? sh[threadIdx.x - off] : 0;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) bsum[i] = carry + sh[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == SCAN_BLOCK - 1) carry += sh[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ void k_scan_add(int64_t *__restrict__ pos, int64_t n,
                           const int64_t *__restrict__ bsum) {
  int64_t i = (int64_t)blockIdx.x * SCAN_BLOCK + threadIdx.x;
  if (i < n) pos[i] += bsum[blockIdx.x];
}

void launch_scan_accept(hipStream_t st, const int32_t *acc, int64_t n_cand, int64_t *pos,
                        int64_t *block_sums, int64_t * /*unused*/, int64_t *total) {
  int64_t nb = (n_cand + SCAN_BLOCK - 1) / SCAN_BLOCK;
  if (nb == 0) nb = 1;
  hipLaunchKernelGGL(k_scan_block, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, st, acc, n_cand, pos,
                     block_sums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_BLOCK), 0, st, block_sums, nb, total);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, st, pos, n_cand,
                     block_sums);
}

__global__ void k_compact(const double *__restrict__ cand, const int32_t *__restrict__ acc,
                          const int64_t *__restrict__ pos, int64_t n_cand, int64_t out_offset,
                          int64_t n_max, double *__restrict__ pts) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_cand || !acc[i]) return;
  int64_t o = out_offset + pos[i];
  if (o >= n_max) return;
  for (int q = 0; q < 10; q++) pts[10 * o + q] = cand[10 * i + q];
}

void launch_compact_points(hipStream_t st, const double *cand, const int32_t *acc,
                           const int64_t *pos, int64_t n_cand, int64_t out_offset,
                           int64_t n_max, double *pts) {
  if (n_cand <= 0) return;
  hipLaunchKernelGGL(k_compact, dim3((unsigned)((n_cand + 255) / 256)), dim3(256), 0, st, cand,
                     acc, pos, n_cand, out_offset, n_max, pts);
}

}  // namespace cy3q
