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

// Two deterministic passes (quintic.rs:1469-1498), each over fixed blocks of points followed by
// one workgroup that sums the block partials in block order:
//   pass 1: total_w = sum w, weighted = sum w eta, count over valid (eta finite > 0, w finite > 0)
//   pass 2: sum w |eta/kappa - 1|, sum w (eta/kappa - 1)^2, sum w 2 (eta/kappa - 1)(eta/kappa)
// out[0] = sigma_L1, out[1] = kappa, out[2] = sigma_L2^2 (NaN where the reference returns NaN),
// out[3] = total_w, out[4] = count, out[5] = sum w 2 (eta/kappa - 1)(eta/kappa).
static constexpr int SR_BLOCKS = 512;
static constexpr int SR_THREADS = 256;

__device__ inline void sr_sum3(double &a, double &b, double &c, double (*sh)[SR_THREADS]) {
  const int t = threadIdx.x;
  sh[0][t] = a;
  sh[1][t] = b;
  sh[2][t] = c;
  __syncthreads();
  for (int off = SR_THREADS / 2; off > 0; off >>= 1) {
    if (t < off) {
      sh[0][t] += sh[0][t + off];
      sh[1][t] += sh[1][t + off];
      sh[2][t] += sh[2][t + off];
    }
    __syncthreads();
  }
  a = sh[0][0];
  b = sh[1][0];
  c = sh[2][0];
}

__device__ inline bool sr_valid(double e, double wp) {
  return isfinite(e) && e > 0.0 && isfinite(wp) && wp > 0.0;
}

// Block b covers points [b * chunk, (b + 1) * chunk). stats (pass 2 only): total_w, wsum, count.
template <int PASS>
__global__ void __launch_bounds__(SR_THREADS) k_sigma_part(const double *__restrict__ eta,
                                                           const double *__restrict__ w, int64_t n,
                                                           int64_t chunk, double min_tw,
                                                           const double *__restrict__ stats,
                                                           double *__restrict__ part) {
  __shared__ double sh[3][SR_THREADS];
  const int64_t p0 = (int64_t)blockIdx.x * chunk;
  const int64_t p1 = p0 + chunk < n ? p0 + chunk : n;
  double a = 0.0, b = 0.0, c = 0.0;
  if (PASS == 1) {
    for (int64_t p = p0 + threadIdx.x; p < p1; p += SR_THREADS) {
      const double e = eta[p], wp = w[p];
      if (!sr_valid(e, wp)) continue;
      a += wp;
      b += wp * e;
      c += 1.0;
    }
  } else {
    const double total_w = stats[0], count = stats[2];
    const bool ok = count > 0.0 && total_w >= min_tw;
    const double kappa = ok ? stats[1] / total_w : NAN;
    if (ok && fabs(kappa) >= 1e-30) {
      for (int64_t p = p0 + threadIdx.x; p < p1; p += SR_THREADS) {
        const double e = eta[p], wp = w[p];
        if (!sr_valid(e, wp)) continue;
        const double r = e / kappa - 1.0;
        a += wp * fabs(r);
        b += wp * r * r;
        c += wp * 2.0 * r * (e / kappa);
      }
    }
  }
  sr_sum3(a, b, c, sh);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x + 0] = a;
    part[3 * blockIdx.x + 1] = b;
    part[3 * blockIdx.x + 2] = c;
  }
}

// Block partials in block order. Pass 1 writes stats; pass 2 writes out (with stats).
template <int PASS>
__global__ void __launch_bounds__(SR_THREADS) k_sigma_final(const double *__restrict__ part,
                                                            int nparts, double min_tw,
                                                            double *__restrict__ stats,
                                                            double *__restrict__ out) {
  __shared__ double sh[3][SR_THREADS];
  double a = 0.0, b = 0.0, c = 0.0;
  for (int i = threadIdx.x; i < nparts; i += SR_THREADS) {
    a += part[3 * i];
    b += part[3 * i + 1];
    c += part[3 * i + 2];
  }
  sr_sum3(a, b, c, sh);
  if (threadIdx.x != 0) return;
  if (PASS == 1) {
    stats[0] = a;
    stats[1] = b;
    stats[2] = c;
    return;
  }
  const double total_w = stats[0], count = stats[2];
  const bool ok = count > 0.0 && total_w >= min_tw;
  const double kappa = ok ? stats[1] / total_w : NAN;
  const bool ok2 = ok && fabs(kappa) >= 1e-30;
  out[0] = ok2 ? a / total_w : NAN;
  out[1] = ok2 ? kappa : NAN;
  out[2] = ok2 ? b / total_w : NAN;
  out[3] = total_w;
  out[4] = count;
  out[5] = ok2 ? c : NAN;  // sum w 2 (eta/kappa - 1)(eta/kappa), for the sigma^2 gradient
}

int sigma_reduce_scratch() { return 3 * SR_BLOCKS + 3; }

// scratch: sigma_reduce_scratch() doubles. min_total_w: the weight sum below which sigma is
// NaN (1e-12 in the quintic sigma, quintic.rs:1479; 1e-30 in Cy3SigmaGpu, cy3_sigma_gpu.rs:514).
void launch_sigma_reduce(hipStream_t st, const double *eta, const double *w, int64_t n,
                         double *scratch, double *out6, double min_total_w) {
  const int64_t chunk = (n + SR_BLOCKS - 1) / SR_BLOCKS > 0 ? (n + SR_BLOCKS - 1) / SR_BLOCKS : 1;
  double *part = scratch, *stats = scratch + 3 * SR_BLOCKS;
  hipLaunchKernelGGL(k_sigma_part<1>, dim3(SR_BLOCKS), dim3(SR_THREADS), 0, st, eta, w, n, chunk,
                     min_total_w, stats, part);
  hipLaunchKernelGGL(k_sigma_final<1>, dim3(1), dim3(SR_THREADS), 0, st, part, SR_BLOCKS,
                     min_total_w, stats, out6);
  hipLaunchKernelGGL(k_sigma_part<2>, dim3(SR_BLOCKS), dim3(SR_THREADS), 0, st, eta, w, n, chunk,
                     min_total_w, stats, part);
  hipLaunchKernelGGL(k_sigma_final<2>, dim3(1), dim3(SR_THREADS), 0, st, part, SR_BLOCKS,
                     min_total_w, stats, out6);
}

}  // namespace cy3q
