// sigma-functional on the GPU (FP64).
//
// Reference: compute_sigma_from_workspace (cy3/src/quintic.rs:1301-1499) with
// ambient_metric_5x5 (:4474-4516), quintic_affine_chart_frame (:4541-4584),
// project_to_quintic_tangent (:4739-4776) and det_3x3_complex_hermitian (:4782-4809).
//
// With G = W^H W, the reference's per-point Bergman quantities become inner products of the
// contracted vectors y0 = W s and y_i = W ds/dz_i. Those come from the triangular contraction,
// so the per-point kernel does no N^2 work:
//   K = s^H G s = |y0|^2,  dK_i = (G s)^H ds_i = y0^H y_i,  ddK_ij = ds_j^H G ds_i = y_j^H y_i.
// One wavefront per point: the lanes split the N-length dot products, the results are reduced
// with cross-lane shuffles, and lane 0 does the 5x5 -> 3x3 algebra in FP64 registers.
#include "cy3q_internal.h"

namespace cy3q {

__device__ inline void grad_quintic(const double *z, double psi, double *g) {
  for (int j = 0; j < 5; j++) {
    double r = z[2 * j], i = z[2 * j + 1];
    double r2 = r * r, i2 = i * i;
    double z2r = r2 - i2, z2i = 2.0 * r * i;
    double z4r = z2r * z2r - z2i * z2i, z4i = 2.0 * z2r * z2i;
    g[2 * j] = 5.0 * z4r;
    g[2 * j + 1] = 5.0 * z4i;
  }
  if (psi == 0.0) return;
  double sc = -5.0 * psi;
  for (int k = 0; k < 5; k++) {
    double pr = 1.0, pi = 0.0;
    for (int j = 0; j < 5; j++) {
      if (j == k) continue;
      double r = z[2 * j], i = z[2 * j + 1];
      double nr = pr * r - pi * i, ni = pr * i + pi * r;
      pr = nr;
      pi = ni;
    }
    g[2 * k] += sc * pr;
    g[2 * k + 1] += sc * pi;
  }
}

// Per-point eta from the six contracted vectors (y0 and y_0..y_4). Shared with the adjoint
// kernel, which needs the same intermediates. Returns NaN for points the reference skips.
struct PointMetric {
  double k;             // K (floored at 1e-30)
  double dk[5][2];      // dK_i
  double ga[5][5][2];   // g_amb
  double fr[3][10];     // affine frame
  double gt[3][3][2];   // g_tan
  double det;
};

__device__ inline void ambient_from_dots(double kraw, const double dk[5][2],
                                         const double ddk[5][5][2], PointMetric &pm) {
  pm.k = fmax(kraw, 1e-30);
  const double ik = 1.0 / pm.k, ik2 = ik * ik;
  for (int i = 0; i < 5; i++) {
    pm.dk[i][0] = dk[i][0];
    pm.dk[i][1] = dk[i][1];
  }
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) {
      double t1r = ddk[i][j][0] * ik, t1i = ddk[i][j][1] * ik;
      double pr = dk[i][0] * dk[j][0] + dk[i][1] * dk[j][1];
      double pi = dk[i][1] * dk[j][0] - dk[i][0] * dk[j][1];
      pm.ga[i][j][0] = t1r - pr * ik2;
      pm.ga[i][j][1] = t1i - pi * ik2;
    }
}

__device__ inline void frame_and_det(const double *z, double psi, PointMetric &pm) {
  double g[10];
  grad_quintic(z, psi, g);
  double mz = 0.0;
  int chart = 0;
  for (int k = 0; k < 5; k++) {
    double zs = z[2 * k] * z[2 * k] + z[2 * k + 1] * z[2 * k + 1];
    if (zs > mz) {
      mz = zs;
      chart = k;
    }
  }
  double mg = 0.0;
  int elim = chart == 0 ? 1 : 0;
  for (int k = 0; k < 5; k++) {
    if (k == chart) continue;
    double gq = g[2 * k] * g[2 * k] + g[2 * k + 1] * g[2 * k + 1];
    if (gq > mg) {
      mg = gq;
      elim = k;
    }
  }
  for (int a = 0; a < 3; a++)
    for (int q = 0; q < 10; q++) pm.fr[a][q] = 0.0;
  int fre[3], idx = 0;
  for (int k = 0; k < 5; k++)
    if (k != chart && k != elim) fre[idx++] = k;
  double ger = g[2 * elim], gei = g[2 * elim + 1];
  double gn = ger * ger + gei * gei;
  if (idx == 3 && !(gn < 1e-30)) {
    for (int i = 0; i < 3; i++) {
      int fi = fre[i];
      double gr = g[2 * fi], gim = g[2 * fi + 1];
      double nr = gr * ger + gim * gei, ni = gim * ger - gr * gei;
      pm.fr[i][2 * fi] = 1.0;
      pm.fr[i][2 * fi + 1] = 0.0;
      pm.fr[i][2 * elim] = -nr / gn;
      pm.fr[i][2 * elim + 1] = -ni / gn;
    }
  }
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      double sr = 0.0, si = 0.0;
      for (int i = 0; i < 5; i++) {
        double tar = pm.fr[a][2 * i], tai = pm.fr[a][2 * i + 1];
        double rr = 0.0, ri = 0.0;
        for (int j = 0; j < 5; j++) {
          double gr = pm.ga[i][j][0], gi = pm.ga[i][j][1];
          double tjr = pm.fr[b][2 * j], tji = -pm.fr[b][2 * j + 1];
          rr += gr * tjr - gi * tji;
          ri += gr * tji + gi * tjr;
        }
        sr += tar * rr - tai * ri;
        si += tar * ri + tai * rr;
      }
      pm.gt[a][b][0] = sr;
      pm.gt[a][b][1] = si;
    }
#define CMR(x, y) ((x)[0] * (y)[0] - (x)[1] * (y)[1])
#define CMI(x, y) ((x)[0] * (y)[1] + (x)[1] * (y)[0])
  const double(*gg)[3][2] = pm.gt;
  double c00[2] = {CMR(gg[1][1], gg[2][2]) - CMR(gg[1][2], gg[2][1]),
                   CMI(gg[1][1], gg[2][2]) - CMI(gg[1][2], gg[2][1])};
  double c01[2] = {CMR(gg[1][0], gg[2][2]) - CMR(gg[1][2], gg[2][0]),
                   CMI(gg[1][0], gg[2][2]) - CMI(gg[1][2], gg[2][0])};
  double c02[2] = {CMR(gg[1][0], gg[2][1]) - CMR(gg[1][1], gg[2][0]),
                   CMI(gg[1][0], gg[2][1]) - CMI(gg[1][1], gg[2][0])};
  pm.det = CMR(gg[0][0], c00) - CMR(gg[0][1], c01) + CMR(gg[0][2], c02);
#undef CMR
#undef CMI
}

// y0: [n][m] contracted sections, yd: [n][5][m] contracted derivatives.
__global__ void __launch_bounds__(256) k_sigma_eta(const double *__restrict__ pts,
                                                   const double *__restrict__ y0,
                                                   const double *__restrict__ yd,
                                                   const double *__restrict__ log_om, int64_t n,
                                                   int64_t m, double psi,
                                                   double *__restrict__ eta) {
  const int lane = threadIdx.x & 63;
  const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (p >= n) return;
  // acc layout: [0] K, [1..10] dK (5 complex), [11..60] ddK (25 complex).
  double acc[61];
#pragma unroll
  for (int q = 0; q < 61; q++) acc[q] = 0.0;
  const double *r0 = y0 + p * m;
  const double *rd = yd + p * 5 * m;
  for (int64_t c = 2 * lane; c < m; c += 128) {
    double sr = r0[c], si = r0[c + 1];
    double vr[5], vi[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
      vr[i] = rd[i * m + c];
      vi[i] = rd[i * m + c + 1];
    }
    acc[0] += sr * sr + si * si;
#pragma unroll
    for (int i = 0; i < 5; i++) {  // conj(y0) * y_i
      acc[1 + 2 * i] += sr * vr[i] + si * vi[i];
      acc[2 + 2 * i] += sr * vi[i] - si * vr[i];
    }
#pragma unroll
    for (int i = 0; i < 5; i++)
#pragma unroll
      for (int j = 0; j < 5; j++) {  // ddK_ij = conj(y_j) * y_i
        acc[11 + 2 * (5 * i + j)] += vr[j] * vr[i] + vi[j] * vi[i];
        acc[12 + 2 * (5 * i + j)] += vr[j] * vi[i] - vi[j] * vr[i];
      }
  }
#pragma unroll
  for (int q = 0; q < 61; q++) {
    double v = acc[q];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    acc[q] = v;
  }
  if (lane != 0) return;
  const double lo = log_om[p];
  if (!isfinite(lo)) {
    eta[p] = NAN;
    return;
  }
  double dk[5][2], ddk[5][5][2];
  for (int i = 0; i < 5; i++) {
    dk[i][0] = acc[1 + 2 * i];
    dk[i][1] = acc[2 + 2 * i];
    for (int j = 0; j < 5; j++) {
      ddk[i][j][0] = acc[11 + 2 * (5 * i + j)];
      ddk[i][j][1] = acc[12 + 2 * (5 * i + j)];
    }
  }
  PointMetric pm;
  ambient_from_dots(acc[0], dk, ddk, pm);
  double z[10];
  for (int q = 0; q < 10; q++) z[q] = pts[10 * p + q];
  frame_and_det(z, psi, pm);
  if (!isfinite(pm.det) || fabs(pm.det) < 1e-30) {
    eta[p] = NAN;
    return;
  }
  eta[p] = exp(log(fabs(pm.det)) - lo);
}

void launch_sigma_eta(hipStream_t st, const double *pts, const double *y0, const double *yd,
                      const double *log_om, int64_t n, int64_t m_cols, double psi, double *eta) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_sigma_eta, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, pts, y0, yd,
                     log_om, n, m_cols, psi, eta);
}

// Two deterministic single-workgroup passes (quintic.rs:1469-1498):
//   pass 1: total_w = sum w, weighted = sum w eta, count over valid (eta finite > 0, w finite > 0)
//   pass 2: sum w |eta/kappa - 1|, sum w (eta/kappa - 1)^2
// out[0] = sigma_L1, out[1] = kappa, out[2] = sigma_L2^2 (NaN where the reference returns NaN),
// out[3] = total_w, out[4] = count.
__global__ void __launch_bounds__(1024) k_sigma_reduce(const double *__restrict__ eta,
                                                       const double *__restrict__ w, int64_t n,
                                                       double *__restrict__ out) {
  __shared__ double s0[1024], s1[1024], s2[1024];
  double tw = 0.0, we = 0.0, cnt = 0.0;
  for (int64_t p = threadIdx.x; p < n; p += 1024) {
    double e = eta[p], wp = w[p];
    if (!isfinite(e) || e <= 0.0 || !isfinite(wp) || wp <= 0.0) continue;
    tw += wp;
    we += wp * e;
    cnt += 1.0;
  }
  s0[threadIdx.x] = tw;
  s1[threadIdx.x] = we;
  s2[threadIdx.x] = cnt;
  __syncthreads();
  for (int off = 512; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) {
      s0[threadIdx.x] += s0[threadIdx.x + off];
      s1[threadIdx.x] += s1[threadIdx.x + off];
      s2[threadIdx.x] += s2[threadIdx.x + off];
    }
    __syncthreads();
  }
  const double total_w = s0[0], wsum = s1[0], count = s2[0];
  __syncthreads();
  const bool ok = count > 0.0 && total_w >= 1e-12;
  const double kappa = ok ? wsum / total_w : NAN;
  const bool ok2 = ok && fabs(kappa) >= 1e-30;
  double ad = 0.0, sq = 0.0;
  if (ok2) {
    for (int64_t p = threadIdx.x; p < n; p += 1024) {
      double e = eta[p], wp = w[p];
      if (!isfinite(e) || e <= 0.0 || !isfinite(wp) || wp <= 0.0) continue;
      double r = e / kappa - 1.0;
      ad += wp * fabs(r);
      sq += wp * r * r;
    }
  }
  s0[threadIdx.x] = ad;
  s1[threadIdx.x] = sq;
  __syncthreads();
  for (int off = 512; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) {
      s0[threadIdx.x] += s0[threadIdx.x + off];
      s1[threadIdx.x] += s1[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = ok2 ? s0[0] / total_w : NAN;
    out[1] = ok2 ? kappa : NAN;
    out[2] = ok2 ? s1[0] / total_w : NAN;
    out[3] = total_w;
    out[4] = count;
  }
}

void launch_sigma_reduce(hipStream_t st, const double *eta, const double *w, int64_t n,
                         double *out5) {
  hipLaunchKernelGGL(k_sigma_reduce, dim3(1), dim3(1024), 0, st, eta, w, n, out5);
}

}  // namespace cy3q
