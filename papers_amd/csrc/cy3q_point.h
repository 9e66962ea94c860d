// Per-point FP64 metric algebra shared by the sigma and adjoint kernels.
//
// Reference: ambient_metric_5x5 (cy3/src/quintic.rs:4474-4516), quintic_affine_chart_frame
// (:4541-4584), project_to_quintic_tangent (:4739-4776), det_3x3_complex_hermitian
// (:4782-4809), invert_3x3_complex_hermitian (:2194-2245) and compute_y_5x5 (:2260-2301).
#pragma once
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

// The per-point Bergman quantities and the tangent metric of one point.
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

// Inverse of the 3x3 Hermitian g_tan by cofactors (transposed) over det, then the Hermitian
// average of each off-diagonal pair. False when det is unusable.
__device__ inline bool inv3_herm(const double g[3][3][2], double det, double inv[3][3][2]) {
  if (!isfinite(det) || fabs(det) < 1e-30) return false;
#define CMR(x, y) ((x)[0] * (y)[0] - (x)[1] * (y)[1])
#define CMI(x, y) ((x)[0] * (y)[1] + (x)[1] * (y)[0])
#define COF(o, a, b, c, d, e, f, gg_, h)                                   \
  do {                                                                     \
    (o)[0] = CMR(g[a][b], g[c][d]) - CMR(g[e][f], g[gg_][h]);              \
    (o)[1] = CMI(g[a][b], g[c][d]) - CMI(g[e][f], g[gg_][h]);              \
  } while (0)
  double c[3][3][2];
  COF(c[0][0], 1, 1, 2, 2, 1, 2, 2, 1);
  COF(c[0][1], 1, 2, 2, 0, 1, 0, 2, 2);
  COF(c[0][2], 1, 0, 2, 1, 1, 1, 2, 0);
  COF(c[1][0], 0, 2, 2, 1, 0, 1, 2, 2);
  COF(c[1][1], 0, 0, 2, 2, 0, 2, 2, 0);
  COF(c[1][2], 0, 1, 2, 0, 0, 0, 2, 1);
  COF(c[2][0], 0, 1, 1, 2, 0, 2, 1, 1);
  COF(c[2][1], 0, 2, 1, 0, 0, 0, 1, 2);
  COF(c[2][2], 0, 0, 1, 1, 0, 1, 1, 0);
#undef COF
#undef CMR
#undef CMI
  const double idet = 1.0 / det;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      inv[i][j][0] = c[j][i][0] * idet;
      inv[i][j][1] = c[j][i][1] * idet;
    }
  for (int i = 0; i < 3; i++)
    for (int j = i + 1; j < 3; j++) {
      const double ar = 0.5 * (inv[j][i][0] + inv[i][j][0]);
      const double ai = 0.5 * (inv[j][i][1] - inv[i][j][1]);
      inv[j][i][0] = ar;
      inv[j][i][1] = ai;
      inv[i][j][0] = ar;
      inv[i][j][1] = -ai;
    }
  return true;
}

// Y[i][j] = sum_ab conj(T_a[i]) ginv[a][b] T_b[j] (5 x 5 complex).
__device__ inline void y5_from_frame(const double fr[3][10], const double gi[3][3][2],
                                     double y[5][5][2]) {
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) {
      double sr = 0.0, si = 0.0;
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
          const double gar = gi[a][b][0], gai = gi[a][b][1];
          const double tiar = fr[a][2 * i], tiai = -fr[a][2 * i + 1];
          const double tjbr = fr[b][2 * j], tjbi = fr[b][2 * j + 1];
          const double p1r = tiar * gar - tiai * gai, p1i = tiar * gai + tiai * gar;
          sr += p1r * tjbr - p1i * tjbi;
          si += p1r * tjbi + p1i * tjbr;
        }
      y[i][j][0] = sr;
      y[i][j][1] = si;
    }
}

// The 61 per-point dot products of the contracted vectors y0 = W s and y_i = W ds/dz_i, summed
// over the L lanes that share the point (sl = lane index within the group; every lane of the
// group ends with the totals): [0] K, [1..10] dK_i = conj(y0) y_i, [11..60] ddK_ij =
// conj(y_j) y_i. L is a power of two up to 64.
template <int L>
__device__ inline void point_dots_l(const double *__restrict__ r0, const double *__restrict__ rd,
                                    int64_t m, int sl, double acc[61]) {
#pragma unroll
  for (int q = 0; q < 61; q++) acc[q] = 0.0;
  for (int64_t c = 2 * sl; c < m; c += 2 * L) {
    double sr = r0[c], si = r0[c + 1];
    double vr[5], vi[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
      vr[i] = rd[i * m + c];
      vi[i] = rd[i * m + c + 1];
    }
    acc[0] += sr * sr + si * si;
#pragma unroll
    for (int i = 0; i < 5; i++) {
      acc[1 + 2 * i] += sr * vr[i] + si * vi[i];
      acc[2 + 2 * i] += sr * vi[i] - si * vr[i];
    }
#pragma unroll
    for (int i = 0; i < 5; i++)
#pragma unroll
      for (int j = 0; j < 5; j++) {
        acc[11 + 2 * (5 * i + j)] += vr[j] * vr[i] + vi[j] * vi[i];
        acc[12 + 2 * (5 * i + j)] += vr[j] * vi[i] - vi[j] * vr[i];
      }
  }
#pragma unroll
  for (int q = 0; q < 61; q++) {
    double v = acc[q];
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    acc[q] = v;
  }
}

// One point per 64-lane wave.
__device__ inline void point_dots(const double *__restrict__ r0, const double *__restrict__ rd,
                                  int64_t m, int lane, double acc[61]) {
  point_dots_l<64>(r0, rd, m, lane, acc);
}

__device__ inline void unpack_dots(const double acc[61], double dk[5][2], double ddk[5][5][2]) {
  for (int i = 0; i < 5; i++) {
    dk[i][0] = acc[1 + 2 * i];
    dk[i][1] = acc[2 + 2 * i];
    for (int j = 0; j < 5; j++) {
      ddk[i][j][0] = acc[11 + 2 * (5 * i + j)];
      ddk[i][j][1] = acc[12 + 2 * (5 * i + j)];
    }
  }
}

}  // namespace cy3q
