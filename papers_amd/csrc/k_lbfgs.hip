// L-BFGS on the sigma_L2^2 landscape over (h_re, h_im) = (Re G, Im G): the device side of
// sigma_refine_lbfgs_gpu_legacy (reference: quintic.rs:678-880) and apply_lbfgs_step
// (:2003-2048). The optimiser's vectors (gradients, directions, curvature pairs) are 2 nb^2
// doubles in the gradient layout of sigma_squared_and_gradient (h_re (a, b) for all pairs,
// then h_im (a, b)); they stay on the device, and the host sees only the scalars it branches
// on (g.d, |g|^2, s.y). Every inner product is a fixed-shape block reduction, so runs are
// deterministic.
#include "cy3q_internal.h"

namespace cy3q {

static constexpr int LB_THREADS = 1024;

// Sum over the workgroup of one value per thread, in a fixed tree order. red: LB_THREADS doubles.
__device__ inline double lb_block_sum(double v, double *red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int off = LB_THREADS / 2; off > 0; off >>= 1) {
    if (tid < off) red[tid] += red[tid + off];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__device__ inline double lb_dot(const double *__restrict__ a, const double *__restrict__ b,
                                int64_t n, double *red) {
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += LB_THREADS) acc += a[i] * b[i];
  return lb_block_sum(acc, red);
}

// The two-loop recursion (quintic.rs:746-776): d = -H g from the pairs listed oldest first in
// hist (slot, rho = 1/(s.y)), with the initial scaling gamma = (s.y)/(y.y) of the newest pair
// (1 when y.y <= 1e-30 or rho is not finite, or without pairs). out[0] = g.d, out[1] = g.g.
// Each thread only touches its own strided elements of d, so only the reductions synchronise.
__global__ void __launch_bounds__(LB_THREADS) k_lbfgs_direction(const double *__restrict__ g,
                                                                const double *__restrict__ s_all,
                                                                const double *__restrict__ y_all,
                                                                LbfgsPairs hist, int64_t n2,
                                                                double *__restrict__ d,
                                                                double *__restrict__ out) {
  __shared__ double red[LB_THREADS];
  __shared__ double alpha[LBFGS_MAX_PAIRS];
  const int tid = threadIdx.x;
  for (int64_t i = tid; i < n2; i += LB_THREADS) d[i] = g[i];
  for (int k = hist.count - 1; k >= 0; k--) {
    const double *sk = s_all + (int64_t)hist.slot[k] * n2, *yk = y_all + (int64_t)hist.slot[k] * n2;
    const double a = hist.rho[k] * lb_dot(sk, d, n2, red);
    if (tid == 0) alpha[k] = a;
    for (int64_t i = tid; i < n2; i += LB_THREADS) d[i] -= a * yk[i];
  }
  double gamma = 1.0;
  if (hist.count > 0) {
    const double *yl = y_all + (int64_t)hist.slot[hist.count - 1] * n2;
    const double sy = hist.rho[hist.count - 1];  // the reference's name for rho here
    const double yy = lb_dot(yl, yl, n2, red);
    if (yy > 1e-30 && isfinite(sy)) gamma = 1.0 / (sy * yy);
  }
  for (int64_t i = tid; i < n2; i += LB_THREADS) d[i] *= gamma;
  __syncthreads();  // alpha[] written by thread 0
  for (int k = 0; k < hist.count; k++) {
    const double *sk = s_all + (int64_t)hist.slot[k] * n2, *yk = y_all + (int64_t)hist.slot[k] * n2;
    const double beta = hist.rho[k] * lb_dot(yk, d, n2, red);
    const double c = alpha[k] - beta;
    for (int64_t i = tid; i < n2; i += LB_THREADS) d[i] += sk[i] * c;
  }
  for (int64_t i = tid; i < n2; i += LB_THREADS) d[i] = -d[i];
  const double gd = lb_dot(g, d, n2, red);
  const double gg = lb_dot(g, g, n2, red);
  if (tid == 0) {
    out[0] = gd;
    out[1] = gg;
  }
}

// d = -g with out[0] = g.d and out[1] = g.g: the steepest-descent direction.
__global__ void __launch_bounds__(LB_THREADS) k_lbfgs_steepest(const double *__restrict__ g,
                                                               int64_t n2, double *__restrict__ d,
                                                               double *__restrict__ out) {
  __shared__ double red[LB_THREADS];
  const int tid = threadIdx.x;
  for (int64_t i = tid; i < n2; i += LB_THREADS) d[i] = -g[i];
  const double gd = lb_dot(g, d, n2, red);
  const double gg = lb_dot(g, g, n2, red);
  if (tid == 0) {
    out[0] = gd;
    out[1] = gg;
  }
}

// apply_lbfgs_step (quintic.rs:2003-2048) from the saved point: h_re = g0_re + step d_re for
// every (a, b), then averaged over (a, b), (b, a); h_im(a, b) = g0_im(a, b) + step d_im(a, b)
// and h_im(b, a) = g0_im(b, a) - step d_im(a, b) for a < b; zero h_im diagonal. One thread per
// pair a <= b; the trace renormalisation follows in its own kernel.
__global__ void k_lbfgs_apply(const double *__restrict__ g0, int nb, const double *__restrict__ d,
                              double step, double *__restrict__ g) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nb * nb) return;
  const int a = (int)(e / nb), b = (int)(e % nb);
  if (a > b) return;
  const int64_t n2 = (int64_t)nb * nb;
  const int64_t ab = (int64_t)a * nb + b, ba = (int64_t)b * nb + a;
  if (a == b) {
    g[2 * ab] = g0[2 * ab] + step * d[ab];
    g[2 * ab + 1] = 0.0;
    return;
  }
  const double re_ab = g0[2 * ab] + step * d[ab], re_ba = g0[2 * ba] + step * d[ba];
  const double re = 0.5 * (re_ab + re_ba);
  const double upd = step * d[n2 + ab];
  g[2 * ab] = re;
  g[2 * ba] = re;
  g[2 * ab + 1] = g0[2 * ab + 1] + upd;
  g[2 * ba + 1] = g0[2 * ba + 1] - upd;
}

// A curvature pair (quintic.rs:850-862): s = step d into s_out, y = g_new - g_old into y_out,
// out[0] = s.y.
__global__ void __launch_bounds__(LB_THREADS) k_lbfgs_pair(const double *__restrict__ d,
                                                           double step,
                                                           const double *__restrict__ g_new,
                                                           const double *__restrict__ g_old,
                                                           int64_t n2, double *__restrict__ s_out,
                                                           double *__restrict__ y_out,
                                                           double *__restrict__ out) {
  __shared__ double red[LB_THREADS];
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n2; i += LB_THREADS) {
    const double sv = step * d[i], yv = g_new[i] - g_old[i];
    s_out[i] = sv;
    y_out[i] = yv;
    acc += sv * yv;
  }
  const double sy = lb_block_sum(acc, red);
  if (threadIdx.x == 0) out[0] = sy;
}

void launch_lbfgs_direction(hipStream_t st, const double *g, const double *s_all,
                            const double *y_all, const LbfgsPairs &hist, int64_t n2, double *d,
                            double *out) {
  if (hist.count > 0)
    hipLaunchKernelGGL(k_lbfgs_direction, dim3(1), dim3(LB_THREADS), 0, st, g, s_all, y_all, hist,
                       n2, d, out);
  else
    hipLaunchKernelGGL(k_lbfgs_steepest, dim3(1), dim3(LB_THREADS), 0, st, g, n2, d, out);
}

void launch_lbfgs_apply(hipStream_t st, const double *g0, int nb, const double *d, double step,
                        double *g) {
  const int64_t nn = (int64_t)nb * nb;
  hipLaunchKernelGGL(k_lbfgs_apply, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, g0, nb, d,
                     step, g);
  launch_trace_renorm(st, g, nb);
}

void launch_lbfgs_pair(hipStream_t st, const double *d, double step, const double *g_new,
                       const double *g_old, int64_t n2, double *s_out, double *y_out, double *out) {
  hipLaunchKernelGGL(k_lbfgs_pair, dim3(1), dim3(LB_THREADS), 0, st, d, step, g_new, g_old, n2,
                     s_out, y_out, out);
}

}  // namespace cy3q
