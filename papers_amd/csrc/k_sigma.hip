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
// Several points per wavefront: a group of lanes splits each point's N-length dot products, the
// results are reduced with cross-lane shuffles, and one lane per point does the 5x5 -> 3x3
// algebra in FP64 registers.
#include "cy3q_point.h"

namespace cy3q {

// y0: [n][m] contracted sections, yd: [n][5][m] contracted derivatives. SE_PPW points per wave:
// the SE_L = 64 / SE_PPW lanes of a point split its dot products, then the first lane of each
// group does the 5x5 -> 3x3 algebra, so SE_PPW points' algebra runs side by side.
static constexpr int SE_PPW = 8, SE_L = 64 / SE_PPW;

__global__ void __launch_bounds__(256) k_sigma_eta(const double *__restrict__ pts,
                                                   const double *__restrict__ y0,
                                                   const double *__restrict__ yd,
                                                   const double *__restrict__ log_om, int64_t n,
                                                   int64_t m, double psi,
                                                   double *__restrict__ eta) {
  const int lane = threadIdx.x & 63, sl = lane % SE_L;
  const int64_t p = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * SE_PPW + lane / SE_L;
  const bool in = p < n;
  double acc[61];
  point_dots_l<SE_L>(y0 + (in ? p : 0) * m, yd + (in ? p : 0) * 5 * m, in ? m : 0, sl, acc);
  if (sl != 0 || !in) return;
  const double lo = log_om[p];
  if (!isfinite(lo)) {
    eta[p] = NAN;
    return;
  }
  double dk[5][2], ddk[5][5][2];
  unpack_dots(acc, dk, ddk);
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
  const int64_t per_wg = 4 * SE_PPW;
  hipLaunchKernelGGL(k_sigma_eta, dim3((unsigned)((n + per_wg - 1) / per_wg)), dim3(256), 0, st,
                     pts, y0, yd, log_om, n, m_cols, psi, eta);
}

// Two deterministic single-workgroup passes (quintic.rs:1469-1498):
//   pass 1: total_w = sum w, weighted = sum w eta, count over valid (eta finite > 0, w finite > 0)
//   pass 2: sum w |eta/kappa - 1|, sum w (eta/kappa - 1)^2
// out[0] = sigma_L1, out[1] = kappa, out[2] = sigma_L2^2 (NaN where the reference returns NaN),
// out[3] = total_w, out[4] = count, out[5] = sum w 2 (eta/kappa - 1)(eta/kappa).
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
  double ad = 0.0, sq = 0.0, ga = 0.0;
  if (ok2) {
    for (int64_t p = threadIdx.x; p < n; p += 1024) {
      double e = eta[p], wp = w[p];
      if (!isfinite(e) || e <= 0.0 || !isfinite(wp) || wp <= 0.0) continue;
      double r = e / kappa - 1.0;
      ad += wp * fabs(r);
      sq += wp * r * r;
      ga += wp * 2.0 * r * (e / kappa);
    }
  }
  s0[threadIdx.x] = ad;
  s1[threadIdx.x] = sq;
  s2[threadIdx.x] = ga;
  __syncthreads();
  for (int off = 512; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) {
      s0[threadIdx.x] += s0[threadIdx.x + off];
      s1[threadIdx.x] += s1[threadIdx.x + off];
      s2[threadIdx.x] += s2[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = ok2 ? s0[0] / total_w : NAN;
    out[1] = ok2 ? kappa : NAN;
    out[2] = ok2 ? s1[0] / total_w : NAN;
    out[3] = total_w;
    out[4] = count;
    out[5] = ok2 ? s2[0] : NAN;  // sum w 2 (eta/kappa - 1)(eta/kappa), for the sigma^2 gradient
  }
}

void launch_sigma_reduce(hipStream_t st, const double *eta, const double *w, int64_t n,
                         double *out5) {
  hipLaunchKernelGGL(k_sigma_reduce, dim3(1), dim3(1024), 0, st, eta, w, n, out5);
}

}  // namespace cy3q
