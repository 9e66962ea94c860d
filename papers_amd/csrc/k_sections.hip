// Per-point geometry on the quintic: CY-measure weights, log|Omega|^2 with the chart/elimination
// choice, and degree-k section values and their z-derivatives.
//
// Operation order follows the reference so the FP64 values agree to rounding:
//   cy_measure_weights(_psi)      cy3/src/quintic.rs:3507-3563
//   quintic_chart_and_elim        cy3/src/quintic.rs:4870-4916
//   evaluate_quintic_basis        cy3/src/quintic.rs:4021-4071
//   evaluate_quintic_basis_derivs_soa   cy3/src/quintic.rs:1231-1296
// Built with -ffp-contract=off, as the reference (Rust) never fuses multiply-add.
#include "cy3q_internal.h"

namespace cy3q {

__device__ inline void quintic_gradient(const double *z, double psi, double *g) {
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

__global__ void __launch_bounds__(256) k_weights_logomega(const double *__restrict__ pts, int64_t n,
                                                          double psi, double *__restrict__ w,
                                                          double *__restrict__ log_om,
                                                          int32_t *__restrict__ chart_elim) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double z[10], g[10];
  for (int q = 0; q < 10; q++) z[q] = pts[10 * p + q];
  quintic_gradient(z, psi, g);
  double gs = 0.0;
  for (int q = 0; q < 10; q++) gs += g[q] * g[q];
  w[p] = gs < 1e-30 ? 0.0 : 1.0 / gs;
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
  log_om[p] = (mg < 1e-20 || mz < 1e-20) ? (double)NAN : log(mz) - log(mg);
  if (chart_elim) {
    chart_elim[2 * p] = chart;
    chart_elim[2 * p + 1] = elim;
  }
}

void launch_weights_logomega(hipStream_t st, const double *pts, int64_t n, double psi, double *w,
                             double *log_om, int32_t *chart_elim) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_weights_logomega, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, pts,
                     n, psi, w, log_om, chart_elim);
}

// One workgroup per PTS_PER_WG points. The workgroup first builds each point's power table
// z_j^e (e <= kmax) in LDS by repeated multiplication, as the reference does. It then walks the
// (point, monomial) pairs with the monomial index fastest, so the row writes are coalesced.
static constexpr int PTS_PER_WG = 16;
static constexpr int KMAX_LIM = 12;

__global__ void __launch_bounds__(256) k_sections(const double *__restrict__ pts, int64_t n,
                                                  const uint32_t *__restrict__ mono, int nb,
                                                  int kmax, int64_t m_cols,
                                                  double *__restrict__ sv,
                                                  double *__restrict__ sd) {
  __shared__ double pw[PTS_PER_WG][5][KMAX_LIM + 1][2];
  int64_t p0 = (int64_t)blockIdx.x * PTS_PER_WG;
  int tid = threadIdx.x;
  if (tid < PTS_PER_WG * 5) {
    int lp = tid / 5, j = tid % 5;
    int64_t p = p0 + lp;
    if (p < n) {
      double zr = pts[10 * p + 2 * j], zi = pts[10 * p + 2 * j + 1];
      pw[lp][j][0][0] = 1.0;
      pw[lp][j][0][1] = 0.0;
      for (int e = 1; e <= kmax; e++) {
        double pr = pw[lp][j][e - 1][0], pi = pw[lp][j][e - 1][1];
        pw[lp][j][e][0] = pr * zr - pi * zi;
        pw[lp][j][e][1] = pr * zi + pi * zr;
      }
    }
  }
  __syncthreads();
  int64_t pend = p0 + PTS_PER_WG < n ? p0 + PTS_PER_WG : n;
  // Section values: pad columns [2nb, m_cols) with zeros.
  int64_t half_cols = m_cols / 2;
  for (int64_t idx = tid; idx < (pend - p0) * half_cols; idx += blockDim.x) {
    int lp = (int)(idx / half_cols);
    int a = (int)(idx % half_cols);
    int64_t p = p0 + lp;
    double pr = 0.0, pi = 0.0;
    if (a < nb) {
      pr = 1.0;
      for (int j = 0; j < 5; j++) {
        int e = (int)mono[5 * a + j];
        double zr = pw[lp][j][e][0], zi = pw[lp][j][e][1];
        double nr = pr * zr - pi * zi, ni = pr * zi + pi * zr;
        pr = nr;
        pi = ni;
      }
    }
    double2 v;
    v.x = pr;
    v.y = pi;
    reinterpret_cast<double2 *>(sv + p * m_cols)[a] = v;
  }
  if (!sd) return;
  // Derivatives d/dz_i s_a, layout [n][5][m_cols].
  for (int64_t idx = tid; idx < (pend - p0) * 5 * half_cols; idx += blockDim.x) {
    int lp = (int)(idx / (5 * half_cols));
    int rem = (int)(idx % (5 * half_cols));
    int i = rem / (int)half_cols;
    int a = rem % (int)half_cols;
    int64_t p = p0 + lp;
    double pr = 0.0, pi = 0.0;
    if (a < nb) {
      uint32_t mi = mono[5 * a + i];
      if (mi != 0) {
        double fr = pw[lp][i][mi - 1][0], fi = pw[lp][i][mi - 1][1];
        double mf = (double)mi;
        pr = mf * fr;
        pi = mf * fi;
        for (int j = 0; j < 5; j++) {
          if (j == i) continue;
          int e = (int)mono[5 * a + j];
          double zr = pw[lp][j][e][0], zi = pw[lp][j][e][1];
          double nr = pr * zr - pi * zi, ni = pr * zi + pi * zr;
          pr = nr;
          pi = ni;
        }
      }
    }
    double2 v;
    v.x = pr;
    v.y = pi;
    reinterpret_cast<double2 *>(sd + (p * 5 + i) * m_cols)[a] = v;
  }
}

void launch_sections(hipStream_t st, const double *pts, int64_t n, const uint32_t *mono, int nb,
                     int kmax, int64_t m_cols, double *sv, double *sd) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_sections, dim3((unsigned)((n + PTS_PER_WG - 1) / PTS_PER_WG)), dim3(256), 0,
                     st, pts, n, mono, nb, kmax, m_cols, sv, sd);
}

}  // namespace cy3q
