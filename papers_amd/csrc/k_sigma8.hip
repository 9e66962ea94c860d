// Per-point eta for 8-coordinate geometries with caller-supplied tangent frames: the kernel of
// the route34 sigma contract (Cy3SigmaGpu, route34/cy3_sigma_gpu.rs:133-235, the TY / Schoen
// sigma evaluator).
//
// With H = W^H W, the reference's Hermitian forms are inner products of the contracted vectors
// y0 = W s and y_i = W d s / d z_i (i = 0..7), which the triangular contraction produces:
//   K = s^H H s = |y0|^2,   dk_i = s^H H ds_i = y0^H y_i,   M_ij = ds_j^H H ds_i = y_j^H y_i,
// g_amb_ij = M_ij / K - dk_i conj(dk_j) / K^2 (K raised to 1e-30), g_tan = conj(T) g_amb T^T
// with the point's 3 x 8 frame T, and eta = |Re det g_tan| / |Omega|^2; NaN for a point with
// w or |Omega|^2 not finite and positive, or with a non-finite or |det| < 1e-30.
// Eight points per wave: each point's 8 lanes split the N complex entries of the 45 inner
// products, a 3-step butterfly within the 8 lanes gives them the sums (a fixed order), and the
// group's first lane does the 8 x 8 -> 3 x 3 algebra.
#include <algorithm>

#include "cy3q_internal.h"

namespace cy3q {

static constexpr int E8_NC = 8;  // ambient coordinates
static constexpr int E8_NF = 3;  // tangent frame size
static constexpr int E8_NM = E8_NC * (E8_NC + 1) / 2;  // M_ij, i <= j

static constexpr int E8_L = 8;              // lanes per point
static constexpr int E8_PPW = 64 / E8_L;    // points per wave

__device__ inline double group_sum(double v) {
#pragma unroll
  for (int off = E8_L / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// y: [n][9][m] (row 0: y0, rows 1..8: y_i), frames: [n][48] (T_a[i] = (re, im) at
// 2 (8 a + i)), om = |Omega|^2, w: weights.
__global__ void __launch_bounds__(256) k_eta8(const double *__restrict__ y, int64_t m, int nb,
                                              const double *__restrict__ frames,
                                              const double *__restrict__ om,
                                              const double *__restrict__ w, int64_t n,
                                              double *__restrict__ eta) {
  const int lane = threadIdx.x & 63, sl = lane % E8_L;
  const int64_t p = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * E8_PPW + lane / E8_L;
  const bool in = p < n;
  const double *yp = y + (in ? p : 0) * 9 * m;
  double k = 0.0, dkr[E8_NC], dki[E8_NC], mr[E8_NM], mi[E8_NM];
#pragma unroll
  for (int i = 0; i < E8_NC; i++) dkr[i] = dki[i] = 0.0;
#pragma unroll
  for (int t = 0; t < E8_NM; t++) mr[t] = mi[t] = 0.0;
  for (int c = sl; c < (in ? nb : 0); c += E8_L) {
    double vr[9], vi[9];
#pragma unroll
    for (int v = 0; v < 9; v++) {
      const double2 z = *reinterpret_cast<const double2 *>(yp + v * m + 2 * c);
      vr[v] = z.x;
      vi[v] = z.y;
    }
    k += vr[0] * vr[0] + vi[0] * vi[0];
#pragma unroll
    for (int i = 0; i < E8_NC; i++) {  // conj(y0) y_i
      dkr[i] += vr[0] * vr[1 + i] + vi[0] * vi[1 + i];
      dki[i] += vr[0] * vi[1 + i] - vi[0] * vr[1 + i];
    }
    int t = 0;
#pragma unroll
    for (int i = 0; i < E8_NC; i++)
#pragma unroll
      for (int j = i; j < E8_NC; j++, t++) {  // M_ij = conj(y_j) y_i
        mr[t] += vr[1 + j] * vr[1 + i] + vi[1 + j] * vi[1 + i];
        mi[t] += vr[1 + j] * vi[1 + i] - vi[1 + j] * vr[1 + i];
      }
  }
  k = group_sum(k);
#pragma unroll
  for (int i = 0; i < E8_NC; i++) {
    dkr[i] = group_sum(dkr[i]);
    dki[i] = group_sum(dki[i]);
  }
#pragma unroll
  for (int t = 0; t < E8_NM; t++) {
    mr[t] = group_sum(mr[t]);
    mi[t] = group_sum(mi[t]);
  }
  if (sl != 0 || !in) return;
  const double wp = w[p], omp = om[p];
  if (!isfinite(wp) || wp <= 0.0 || !isfinite(omp) || omp <= 0.0) {
    eta[p] = NAN;
    return;
  }
  const double ks = k > 1e-30 ? k : 1e-30;
  const double ik = 1.0 / ks, ik2 = ik * ik;
  double gr[E8_NC][E8_NC], gi[E8_NC][E8_NC];
  {
    int t = 0;
#pragma unroll
    for (int i = 0; i < E8_NC; i++)
#pragma unroll
      for (int j = i; j < E8_NC; j++, t++) {
        // M_ji = conj(M_ij)
        const double a_r = mr[t], a_i = mi[t];
        const double dij_r = dkr[i] * dkr[j] + dki[i] * dki[j];  // dk_i conj(dk_j)
        const double dij_i = dki[i] * dkr[j] - dkr[i] * dki[j];
        gr[i][j] = a_r * ik - dij_r * ik2;
        gi[i][j] = a_i * ik - dij_i * ik2;
        gr[j][i] = a_r * ik - dij_r * ik2;      // dk_j conj(dk_i) = conj(dk_i conj(dk_j))
        gi[j][i] = -a_i * ik + dij_i * ik2;
      }
  }
  const double *fp = frames + p * 48;
  double tr[E8_NF][E8_NC], ti[E8_NF][E8_NC];
#pragma unroll
  for (int a = 0; a < E8_NF; a++)
#pragma unroll
    for (int i = 0; i < E8_NC; i++) {
      tr[a][i] = fp[2 * (a * E8_NC + i)];
      ti[a][i] = fp[2 * (a * E8_NC + i) + 1];
    }
  double hr[E8_NF][E8_NF], hi[E8_NF][E8_NF];  // g_tan = conj(T) g T^T
#pragma unroll
  for (int a = 0; a < E8_NF; a++)
#pragma unroll
    for (int b = 0; b < E8_NF; b++) {
      double sr = 0.0, si = 0.0;
#pragma unroll
      for (int i = 0; i < E8_NC; i++) {
        double rr = 0.0, ri = 0.0;
#pragma unroll
        for (int j = 0; j < E8_NC; j++) {
          rr += gr[i][j] * tr[b][j] - gi[i][j] * ti[b][j];
          ri += gr[i][j] * ti[b][j] + gi[i][j] * tr[b][j];
        }
        const double car = tr[a][i], cai = -ti[a][i];
        sr += car * rr - cai * ri;
        si += car * ri + cai * rr;
      }
      hr[a][b] = sr;
      hi[a][b] = si;
    }
  // Re det of the 3 x 3 complex matrix by the first row's cofactors.
  const double m1r = hr[1][1] * hr[2][2] - hi[1][1] * hi[2][2] - (hr[1][2] * hr[2][1] - hi[1][2] * hi[2][1]);
  const double m1i = hr[1][1] * hi[2][2] + hi[1][1] * hr[2][2] - (hr[1][2] * hi[2][1] + hi[1][2] * hr[2][1]);
  const double m2r = hr[1][0] * hr[2][2] - hi[1][0] * hi[2][2] - (hr[1][2] * hr[2][0] - hi[1][2] * hi[2][0]);
  const double m2i = hr[1][0] * hi[2][2] + hi[1][0] * hr[2][2] - (hr[1][2] * hi[2][0] + hi[1][2] * hr[2][0]);
  const double m3r = hr[1][0] * hr[2][1] - hi[1][0] * hi[2][1] - (hr[1][1] * hr[2][0] - hi[1][1] * hi[2][0]);
  const double m3i = hr[1][0] * hi[2][1] + hi[1][0] * hr[2][1] - (hr[1][1] * hi[2][0] + hi[1][1] * hr[2][0]);
  const double det = hr[0][0] * m1r - hi[0][0] * m1i - (hr[0][1] * m2r - hi[0][1] * m2i) +
                     (hr[0][2] * m3r - hi[0][2] * m3i);
  if (!isfinite(det) || fabs(det) < 1e-30) {
    eta[p] = NAN;
    return;
  }
  eta[p] = fabs(det) / omp;
}

// x9[p][0] = sv[p], x9[p][1 + i] = sd[p][i] (i = 0..7), each row two_n values then zeros to m.
// Grid-stride over the n * 9 * m entries, so any size fits a capped grid.
__global__ void k_pack_rows9(const double *__restrict__ sv, const double *__restrict__ sd,
                             int64_t n, int64_t two_n, int64_t m, double *__restrict__ x9) {
  const int64_t tot = n * 9 * m, step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += step) {
    const int64_t c = e % m, r = e / m, v = r % 9, p = r / 9;
    double x = 0.0;
    if (c < two_n) x = v == 0 ? sv[p * two_n + c] : sd[(p * 8 + (v - 1)) * two_n + c];
    x9[e] = x;
  }
}

void launch_pack_rows9(hipStream_t st, const double *sv, const double *sd, int64_t n,
                       int64_t two_n, int64_t m, double *x9) {
  const int64_t tot = n * 9 * m;
  if (tot <= 0) return;
  const int64_t blocks = std::min<int64_t>((tot + 255) / 256, 1 << 20);
  hipLaunchKernelGGL(k_pack_rows9, dim3((unsigned)blocks), dim3(256), 0, st, sv, sd, n, two_n, m,
                     x9);
}

void launch_eta8(hipStream_t st, const double *y, int64_t m, int nb, const double *frames,
                 const double *om, const double *w, int64_t n, double *eta) {
  if (n <= 0) return;
  const int64_t per_wg = 4 * E8_PPW;
  hipLaunchKernelGGL(k_eta8, dim3((unsigned)((n + per_wg - 1) / per_wg)), dim3(256), 0, st, y, m,
                     nb, frames, om, w, n, eta);
}

}  // namespace cy3q
