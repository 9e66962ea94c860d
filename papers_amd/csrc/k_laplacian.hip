// Galerkin assembly of the metric Laplacian on the GPU (FP64 MFMA).
//
// Reference: route34/metric_laplacian.rs. compute_metric_laplacian_spectrum (:296-418),
// build_mass_matrix (:424-471), build_stiffness_matrix (:473-524). The reference loops
// points x basis^2 serially. Here both matrices are Hermitian rank-k updates
//   M = sum_p conj(r0_p) r0_p^T,   K = sum_{p,c} conj(r_pc) r_pc^T
// of per-point rows already scaled by sqrt(w |Omega|^2). The rows are generated per point chunk
// (k_lap_rows_*), real-ified as (re, im) pairs, and contracted by k_herk_f64: lower 128 x 128
// tiles of X^T X on v_mfma_f64_16x16x4f64, accumulated in place (one workgroup per tile and
// buffer, so the sum order is fixed). k_lap_finalize folds the real blocks back into the
// complex matrices.
//
// Two row families:
//  * ambient (the reference's form): phi_alpha = prod z_i^e_i over the 8 ambient coordinates,
//    8 derivative rows per point;
//  * intrinsic on the quintic: phi_ab = s_a conj(s_b) / |z|^(2d) with the raw degree-d monomials
//    s, and 3 rows per point f_c = sum_i R_ci d phi / d conj(z_i), where R = L^T conj(T) factors
//    the pulled-back inverse metric Y = T^H g_tan^-1 T (g_tan^-1 = L L^H), so that
//    sum_c conj(f_c(phi)) f_c(phi') = sum_ij conj(F_j) Y_ij F'_i.
#include "cy3q_point.h"

namespace cy3q {

// ---------------------------------------------------------------------------------------------
// Per-point factor of the intrinsic form: pr[p][0] = sqrt(w |Omega|^2) (0 for a skipped point),
// pr[p][1] = w (0 when skipped), pr[p][2..31] = R (3 x 5 complex); wv[p] = pr[p][1]. A point
// is skipped when w is not finite and positive, log|Omega|^2 is not finite, g_tan is not
// invertible, or g_tan^-1 fails the Cholesky test. LM_PPW points per wave (LM_L lanes each for
// the dot products, one lane each for the algebra).
// ---------------------------------------------------------------------------------------------
static constexpr int LM_PPW = 8, LM_L = 64 / LM_PPW;

__global__ void __launch_bounds__(256) k_lap_metric(const double *__restrict__ pts,
                                                    const double *__restrict__ y0,
                                                    const double *__restrict__ yd,
                                                    const double *__restrict__ w,
                                                    const double *__restrict__ log_om, int64_t n,
                                                    int64_t m, double psi,
                                                    double *__restrict__ pr,
                                                    double *__restrict__ wv) {
  const int lane = threadIdx.x & 63, sl = lane % LM_L;
  const int64_t p = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * LM_PPW + lane / LM_L;
  const bool in = p < n;
  double acc[61];
  point_dots_l<LM_L>(y0 + (in ? p : 0) * m, yd + (in ? p : 0) * 5 * m, in ? m : 0, sl, acc);
  if (sl != 0 || !in) return;
  double *out = pr + 32 * p;
  for (int q = 0; q < 32; q++) out[q] = 0.0;
  wv[p] = 0.0;
  const double wp = w[p], lo = log_om[p];
  if (!isfinite(wp) || wp <= 0.0 || !isfinite(lo)) return;
  double dk[5][2], ddk[5][5][2];
  unpack_dots(acc, dk, ddk);
  PointMetric pm;
  ambient_from_dots(acc[0], dk, ddk, pm);
  double z[10];
  for (int q = 0; q < 10; q++) z[q] = pts[10 * p + q];
  frame_and_det(z, psi, pm);
  double gi[3][3][2];
  if (!inv3_herm(pm.gt, pm.det, gi)) return;
  // Cholesky g_tan^-1 = L L^H (L lower, real diagonal).
  double L[3][3][2] = {};
  for (int j = 0; j < 3; j++) {
    double d = gi[j][j][0];
    for (int k = 0; k < j; k++) d -= L[j][k][0] * L[j][k][0] + L[j][k][1] * L[j][k][1];
    if (!(d > 0.0) || !isfinite(d)) return;
    const double ljj = sqrt(d);
    L[j][j][0] = ljj;
    for (int i = j + 1; i < 3; i++) {
      double sr = gi[i][j][0], si = gi[i][j][1];
      for (int k = 0; k < j; k++) {  // - L[i][k] conj(L[j][k])
        sr -= L[i][k][0] * L[j][k][0] + L[i][k][1] * L[j][k][1];
        si -= L[i][k][1] * L[j][k][0] - L[i][k][0] * L[j][k][1];
      }
      L[i][j][0] = sr / ljj;
      L[i][j][1] = si / ljj;
    }
  }
  // R[c][i] = sum_a L[a][c] conj(T_a[i])
  for (int c = 0; c < 3; c++)
    for (int i = 0; i < 5; i++) {
      double sr = 0.0, si = 0.0;
      for (int a = c; a < 3; a++) {
        const double lr = L[a][c][0], li = L[a][c][1];
        const double tr = pm.fr[a][2 * i], ti = -pm.fr[a][2 * i + 1];
        sr += lr * tr - li * ti;
        si += lr * ti + li * tr;
      }
      out[2 + 2 * (5 * c + i)] = sr;
      out[3 + 2 * (5 * c + i)] = si;
    }
  out[0] = sqrt(wp * exp(lo));
  out[1] = wp;
  wv[p] = wp;
}

void launch_lap_metric(hipStream_t st, const double *pts, const double *y0, const double *yd,
                       const double *w, const double *log_om, int64_t n, int64_t m_cols,
                       double psi, double *pr, double *wv) {
  if (n <= 0) return;
  const int64_t per_wg = 4 * LM_PPW;
  hipLaunchKernelGGL(k_lap_metric, dim3((unsigned)((n + per_wg - 1) / per_wg)), dim3(256), 0, st,
                     pts, y0, yd, w, log_om, n, m_cols, psi, pr, wv);
}

// ---------------------------------------------------------------------------------------------
// Intrinsic rows: X0[p][2 alpha..] = fac u_a conj(s_b), X1[3p + c][2 alpha..] = fac u_a conj(t_cb)
// with u = s / |z|^(2d), G_bi = ds_b/dz_i - d s_b conj(z_i) / |z|^2, t_cb = sum_i conj(R_ci) G_bi.
// Four points per workgroup; columns beyond 2B keep their initial zeros.
// ---------------------------------------------------------------------------------------------
static constexpr int LR_PPB = 4;
static constexpr int LR_MAXS = 70;  // raw monomials of degree <= 4

__global__ void __launch_bounds__(256) k_lap_rows_quintic(const double *__restrict__ pts,
                                                          const double *__restrict__ pr, int64_t n,
                                                          const uint8_t *__restrict__ smono,
                                                          int ns, int sdeg, int64_t mc,
                                                          double *__restrict__ x0,
                                                          double *__restrict__ x1) {
  __shared__ double2 s_u[LR_PPB][LR_MAXS], s_s[LR_PPB][LR_MAXS], s_t[LR_PPB][3][LR_MAXS];
  __shared__ double s_fac[LR_PPB];
  const int tid = threadIdx.x;
  const int q = tid >> 6, l = tid & 63;
  const int64_t pbase = (int64_t)blockIdx.x * LR_PPB;
  {
    const int64_t p = pbase + q;
    if (p < n) {
      double z[10];
      for (int k = 0; k < 10; k++) z[k] = pts[10 * p + k];
      double r2 = 0.0;
      for (int k = 0; k < 5; k++) r2 += z[2 * k] * z[2 * k] + z[2 * k + 1] * z[2 * k + 1];
      double rd = 1.0;
      for (int k = 0; k < sdeg; k++) rd *= r2;
      const double *R = pr + 32 * p + 2;
      if (l == 0) s_fac[q] = pr[32 * p];
      for (int b = l; b < ns; b += 64) {
        const uint8_t *e = smono + 5 * b;
        double vr = 1.0, vi = 0.0;
        for (int k = 0; k < 5; k++)
          for (int r = 0; r < e[k]; r++) {
            const double nr = vr * z[2 * k] - vi * z[2 * k + 1];
            vi = vr * z[2 * k + 1] + vi * z[2 * k];
            vr = nr;
          }
        s_s[q][b] = make_double2(vr, vi);
        s_u[q][b] = make_double2(vr / rd, vi / rd);
        double gr[5], gim[5];
        for (int i = 0; i < 5; i++) {
          double dr = 0.0, di = 0.0;
          if (e[i] > 0) {
            dr = (double)e[i];
            for (int k = 0; k < 5; k++) {
              const int ek = e[k] - (k == i ? 1 : 0);
              for (int r = 0; r < ek; r++) {
                const double nr = dr * z[2 * k] - di * z[2 * k + 1];
                di = dr * z[2 * k + 1] + di * z[2 * k];
                dr = nr;
              }
            }
          }
          // - d s conj(z_i) / |z|^2
          const double cr = vr * z[2 * i] + vi * z[2 * i + 1];
          const double ci = vi * z[2 * i] - vr * z[2 * i + 1];
          gr[i] = dr - (double)sdeg * cr / r2;
          gim[i] = di - (double)sdeg * ci / r2;
        }
        for (int c = 0; c < 3; c++) {
          double tr = 0.0, ti = 0.0;
          for (int i = 0; i < 5; i++) {  // conj(R_ci) G_bi
            const double rr = R[2 * (5 * c + i)], ri = -R[2 * (5 * c + i) + 1];
            tr += rr * gr[i] - ri * gim[i];
            ti += rr * gim[i] + ri * gr[i];
          }
          s_t[q][c][b] = make_double2(tr, ti);
        }
      }
    }
  }
  __syncthreads();
  const int B = ns * ns;
  for (int qq = 0; qq < LR_PPB; qq++) {
    const int64_t p = pbase + qq;
    if (p >= n) break;
    const double f = s_fac[qq];
    double2 *r0 = reinterpret_cast<double2 *>(x0 + p * mc);
    double2 *r1 = reinterpret_cast<double2 *>(x1 + 3 * p * mc);
    const int64_t half = mc / 2;
    for (int al = tid; al < B; al += 256) {
      const int a = al / ns, b = al % ns;
      const double2 u = s_u[qq][a];
      const double ur = u.x * f, ui = u.y * f;
      const double2 sb = s_s[qq][b];
      r0[al] = make_double2(ur * sb.x + ui * sb.y, ui * sb.x - ur * sb.y);
      for (int c = 0; c < 3; c++) {
        const double2 t = s_t[qq][c][b];
        r1[c * half + al] = make_double2(ur * t.x + ui * t.y, ui * t.x - ur * t.y);
      }
    }
  }
}

void launch_lap_rows_quintic(hipStream_t st, const double *pts, const double *pr, int64_t n,
                             const uint8_t *smono, int ns, int sdeg, int64_t mc, double *x0,
                             double *x1) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_lap_rows_quintic, dim3((unsigned)((n + LR_PPB - 1) / LR_PPB)), dim3(256), 0,
                     st, pts, pr, n, smono, ns, sdeg, mc, x0, x1);
}

// ---------------------------------------------------------------------------------------------
// Split-precision form of the intrinsic rows: the same entries, generated 16 points (one stage
// of the split rank-k update) per workgroup and written straight into its fp16 hi/lo planes
// ([stage][column][16 points], the layout of k_split_rows), so the FP64 rows are never stored.
// Buffer 0 holds the mass rows, buffer 1 + c the stiffness rows of f_c. A first pass over the
// same entries finds the chunk's largest |entry| (k_lap_rows_absmax); k_absmax_scale turns it
// into the power-of-two scale that both the planes and the update use.
// LDS (dynamic): u, s and t of the 16 points, 5 ns double2 each, and the 16 factors.
// ---------------------------------------------------------------------------------------------
static constexpr int LS_PTS = 16;

__device__ inline void rows_stage_setup(const double *__restrict__ pts,
                                        const double *__restrict__ pr, int64_t n,
                                        const uint8_t *__restrict__ smono, int ns, int sdeg,
                                        int64_t p0, double2 *su, double2 *ss, double2 *st_,
                                        double *sfac) {
  const int q = threadIdx.x >> 6, l = threadIdx.x & 63;
  for (int j = q; j < LS_PTS; j += 4) {
    const int64_t p = p0 + j;
    double2 *uj = su + j * ns, *sj = ss + j * ns, *tj = st_ + j * 3 * ns;
    if (p >= n) {
      if (l == 0) sfac[j] = 0.0;
      for (int b = l; b < ns; b += 64) {
        uj[b] = sj[b] = make_double2(0.0, 0.0);
        for (int c = 0; c < 3; c++) tj[c * ns + b] = make_double2(0.0, 0.0);
      }
      continue;
    }
    double z[10];
    for (int k = 0; k < 10; k++) z[k] = pts[10 * p + k];
    double r2 = 0.0;
    for (int k = 0; k < 5; k++) r2 += z[2 * k] * z[2 * k] + z[2 * k + 1] * z[2 * k + 1];
    double rd = 1.0;
    for (int k = 0; k < sdeg; k++) rd *= r2;
    const double *R = pr + 32 * p + 2;
    if (l == 0) sfac[j] = pr[32 * p];
    for (int b = l; b < ns; b += 64) {
      const uint8_t *e = smono + 5 * b;
      double vr = 1.0, vi = 0.0;
      for (int k = 0; k < 5; k++)
        for (int r = 0; r < e[k]; r++) {
          const double nr = vr * z[2 * k] - vi * z[2 * k + 1];
          vi = vr * z[2 * k + 1] + vi * z[2 * k];
          vr = nr;
        }
      sj[b] = make_double2(vr, vi);
      uj[b] = make_double2(vr / rd, vi / rd);
      double gr[5], gim[5];
      for (int i = 0; i < 5; i++) {
        double dr = 0.0, di = 0.0;
        if (e[i] > 0) {
          dr = (double)e[i];
          for (int k = 0; k < 5; k++) {
            const int ek = e[k] - (k == i ? 1 : 0);
            for (int r = 0; r < ek; r++) {
              const double nr = dr * z[2 * k] - di * z[2 * k + 1];
              di = dr * z[2 * k + 1] + di * z[2 * k];
              dr = nr;
            }
          }
        }
        const double cr = vr * z[2 * i] + vi * z[2 * i + 1];
        const double ci = vi * z[2 * i] - vr * z[2 * i + 1];
        gr[i] = dr - (double)sdeg * cr / r2;
        gim[i] = di - (double)sdeg * ci / r2;
      }
      for (int c = 0; c < 3; c++) {
        double tr = 0.0, ti = 0.0;
        for (int i = 0; i < 5; i++) {
          const double rr = R[2 * (5 * c + i)], ri = -R[2 * (5 * c + i) + 1];
          tr += rr * gr[i] - ri * gim[i];
          ti += rr * gim[i] + ri * gr[i];
        }
        tj[c * ns + b] = make_double2(tr, ti);
      }
    }
  }
}

// Entry (buffer bf, real column col < 2B) of point j: fac u_a conj(v_b) with v = s (bf = 0) or
// t_{bf-1}, alpha = a ns + b = col / 2, the real (col even) or imaginary part, as the FP64 rows.
__device__ inline double rows_entry(const double2 *su, const double2 *ss, const double2 *st_,
                                    const double *sfac, int ns, int bf, int col, int j) {
  const int al = col >> 1, a = al / ns, b = al - (al / ns) * ns;
  const double f = sfac[j];
  const double2 u = su[j * ns + a];
  const double ur = u.x * f, ui = u.y * f;
  const double2 v = bf == 0 ? ss[j * ns + b] : st_[(j * 3 + (bf - 1)) * ns + b];
  return (col & 1) ? ui * v.x - ur * v.y : ur * v.x + ui * v.y;
}

__global__ void __launch_bounds__(256) k_lap_rows_absmax(const double *__restrict__ pts,
                                                         const double *__restrict__ pr, int64_t n,
                                                         const uint8_t *__restrict__ smono, int ns,
                                                         int sdeg, double *__restrict__ part) {
  extern __shared__ double2 ls[];
  double2 *su = ls, *ss = su + LS_PTS * ns, *st_ = ss + LS_PTS * ns;
  double *sfac = reinterpret_cast<double *>(st_ + 3 * LS_PTS * ns);
  __shared__ double red[256];
  const int64_t p0 = (int64_t)blockIdx.x * LS_PTS;
  rows_stage_setup(pts, pr, n, smono, ns, sdeg, p0, su, ss, st_, sfac);
  __syncthreads();
  const int cols = 2 * ns * ns;
  double mx = 0.0;
  for (int it = threadIdx.x; it < 4 * cols; it += 256) {
    const int bf = it / cols, col = it - bf * cols;
    for (int j = 0; j < LS_PTS; j++) mx = fmax(mx, fabs(rows_entry(su, ss, st_, sfac, ns, bf, col, j)));
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

struct RowPlanes {
  half8_t *h[4], *l[4];
};

__global__ void __launch_bounds__(256) k_lap_rows_planes(const double *__restrict__ pts,
                                                         const double *__restrict__ pr, int64_t n,
                                                         const uint8_t *__restrict__ smono, int ns,
                                                         int sdeg, int64_t mc,
                                                         const double *__restrict__ scale,
                                                         RowPlanes out) {
  extern __shared__ double2 ls[];
  double2 *su = ls, *ss = su + LS_PTS * ns, *st_ = ss + LS_PTS * ns;
  double *sfac = reinterpret_cast<double *>(st_ + 3 * LS_PTS * ns);
  const int64_t s0 = blockIdx.x;
  rows_stage_setup(pts, pr, n, smono, ns, sdeg, s0 * LS_PTS, su, ss, st_, sfac);
  __syncthreads();
  const double sc = scale[0];
  const int cols = 2 * ns * ns;
  for (int64_t it = threadIdx.x; it < 4 * mc; it += 256) {
    const int bf = (int)(it / mc), col = (int)(it - bf * mc);
    half8_t h[2], l[2];
#pragma unroll
    for (int j = 0; j < LS_PTS; j++) {
      const float v =
          col < cols ? (float)(rows_entry(su, ss, st_, sfac, ns, bf, col, j) * sc) : 0.f;
      const _Float16 hv = (_Float16)v;
      h[j >> 3][j & 7] = hv;
      l[j >> 3][j & 7] = (_Float16)(v - (float)hv);
    }
    const int64_t q = s0 * mc + col;
    out.h[bf][q * 2] = h[0];
    out.h[bf][q * 2 + 1] = h[1];
    out.l[bf][q * 2] = l[0];
    out.l[bf][q * 2 + 1] = l[1];
  }
}

static size_t rows_stage_lds(int ns) {
  return sizeof(double2) * 5 * LS_PTS * (size_t)ns + sizeof(double) * LS_PTS;
}

void launch_lap_rows_absmax(hipStream_t st, const double *pts, const double *pr, int64_t n,
                            const uint8_t *smono, int ns, int sdeg, double *part) {
  if (n <= 0) return;
  const int64_t stages = (n + LS_PTS - 1) / LS_PTS;
  hipLaunchKernelGGL(k_lap_rows_absmax, dim3((unsigned)stages), dim3(256), rows_stage_lds(ns), st,
                     pts, pr, n, smono, ns, sdeg, part);
}

void launch_lap_rows_planes(hipStream_t st, const double *pts, const double *pr, int64_t n,
                            const uint8_t *smono, int ns, int sdeg, int64_t mc,
                            const double *scale, half8_t *const *h, half8_t *const *l) {
  if (n <= 0) return;
  RowPlanes o;
  for (int q = 0; q < 4; q++) {
    o.h[q] = h[q];
    o.l[q] = l[q];
  }
  const int64_t stages = (n + LS_PTS - 1) / LS_PTS;
  hipLaunchKernelGGL(k_lap_rows_planes, dim3((unsigned)stages), dim3(256), rows_stage_lds(ns), st,
                     pts, pr, n, smono, ns, sdeg, mc, scale, o);
}

// ---------------------------------------------------------------------------------------------
// Ambient rows (the reference's form): fac = sqrt(w |Omega|^2) for a usable point
// (w finite > 0, Omega finite), 0 otherwise; X0[p] = fac phi, X1[8p + i] = fac d phi / dz_i.
// wv[p] = w for a usable point, 0 otherwise (for the sum of weights).
// ---------------------------------------------------------------------------------------------
__global__ void k_lap_rows_ambient(const double *__restrict__ pts, const double *__restrict__ w,
                                   const double *__restrict__ omega, int64_t n,
                                   const uint8_t *__restrict__ exps, int nb, int64_t mc,
                                   double *__restrict__ x0, double *__restrict__ x1,
                                   double *__restrict__ wv) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * nb) return;
  const int64_t p = e / nb;
  const int al = (int)(e % nb);
  const double wp = w[p], orr = omega[2 * p], oi = omega[2 * p + 1];
  const bool ok = isfinite(wp) && wp > 0.0 && isfinite(orr) && isfinite(oi);
  if (al == 0) wv[p] = ok ? wp : 0.0;
  const double fac = ok ? sqrt((orr * orr + oi * oi) * wp) : 0.0;
  const double *z = pts + 16 * p;
  const uint8_t *ex = exps + 8 * al;
  double vr = 1.0, vi = 0.0;
  for (int k = 0; k < 8; k++)
    for (int r = 0; r < ex[k]; r++) {
      const double nr = vr * z[2 * k] - vi * z[2 * k + 1];
      vi = vr * z[2 * k + 1] + vi * z[2 * k];
      vr = nr;
    }
  double2 *r0 = reinterpret_cast<double2 *>(x0 + p * mc);
  r0[al] = make_double2(fac * vr, fac * vi);
  for (int i = 0; i < 8; i++) {
    double dr = 0.0, di = 0.0;
    if (ex[i] > 0) {
      dr = (double)ex[i];
      for (int k = 0; k < 8; k++) {
        const int ek = ex[k] - (k == i ? 1 : 0);
        for (int r = 0; r < ek; r++) {
          const double nr = dr * z[2 * k] - di * z[2 * k + 1];
          di = dr * z[2 * k + 1] + di * z[2 * k];
          dr = nr;
        }
      }
    }
    double2 *r1 = reinterpret_cast<double2 *>(x1 + (8 * p + i) * mc);
    r1[al] = make_double2(fac * dr, fac * di);
  }
}

void launch_lap_rows_ambient(hipStream_t st, const double *pts, const double *w,
                             const double *omega, int64_t n, const uint8_t *exps, int nb,
                             int64_t mc, double *x0, double *x1, double *wv) {
  if (n <= 0) return;
  const int64_t tot = n * nb;
  hipLaunchKernelGGL(k_lap_rows_ambient, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                     pts, w, omega, n, exps, nb, mc, x0, x1, wv);
}

// ---------------------------------------------------------------------------------------------
// C_b[lower 128 x 128 tiles] += X_b^T X_b over rows [0, rows_b), b = blockIdx.y.
// 4 waves, each a 64 x 64 quadrant as 4 x 4 MFMA tiles; 16-row K stages, LDS double buffer
// filled from registers prefetched one stage ahead.
// MFMA f64 16x16x4 layout: A[i][k] at lane i + 16k, B[k][j] at lane j + 16k; the output element
// e of lane l is (row 4e + l/16, col l%16).
// ---------------------------------------------------------------------------------------------
static constexpr int HK_TB = 128;
static constexpr int HK_KS = 16;
static constexpr int HK_LD = HK_TB + 2;

struct HerkArgs {
  const double *x[12];
  int64_t rows[12];
  double *c[12];
};

__global__ void __launch_bounds__(256) k_herk_f64(HerkArgs args, int64_t mc, int n_tiles) {
  __shared__ double sa[2][HK_KS][HK_LD], sb[2][HK_KS][HK_LD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int b = blockIdx.y;
  const double *__restrict__ x = args.x[b];
  const int64_t rows = args.rows[b];
  double *__restrict__ cm = args.c[b];
  int t = blockIdx.x;
  int I = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= t) I++;
  while (I * (I + 1) / 2 > t) I--;
  const int J = t - I * (I + 1) / 2;
  const bool diag = I == J;
  const int64_t ci = (int64_t)I * HK_TB, cj = (int64_t)J * HK_TB;
  const int wr = 64 * (wave >> 1), wc = 64 * (wave & 1);

  double4_t acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int c = 0; c < 4; c++) acc[a][c] = (double4_t){0.0, 0.0, 0.0, 0.0};

  const int n_stage = (int)((rows + HK_KS - 1) / HK_KS);
  double2 ra[4], rb[4];
  auto fetch = [&](int s) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int e = tid + 256 * u;
      const int r = e >> 6, cc = (e & 63) * 2;
      const int64_t row = (int64_t)s * HK_KS + r;
      const bool in = row < rows;
      ra[u] = in ? *reinterpret_cast<const double2 *>(x + row * mc + ci + cc) : make_double2(0, 0);
      if (!diag)
        rb[u] = in ? *reinterpret_cast<const double2 *>(x + row * mc + cj + cc) : make_double2(0, 0);
    }
  };
  auto put = [&](int buf) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int e = tid + 256 * u;
      const int r = e >> 6, cc = (e & 63) * 2;
      sa[buf][r][cc] = ra[u].x;
      sa[buf][r][cc + 1] = ra[u].y;
      const double2 v = diag ? ra[u] : rb[u];
      sb[buf][r][cc] = v.x;
      sb[buf][r][cc + 1] = v.y;
    }
  };
  if (n_stage > 0) {
    fetch(0);
    put(0);
  }
  __syncthreads();
  for (int s = 0; s < n_stage; s++) {
    const int buf = s & 1;
    if (s + 1 < n_stage) fetch(s + 1);
#pragma unroll
    for (int kk = 0; kk < HK_KS / 4; kk++) {
      const int kr = 4 * kk + (lane >> 4);
      double av[4], bv[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        av[q] = sa[buf][kr][wr + 16 * q + (lane & 15)];
        bv[q] = sb[buf][kr][wc + 16 * q + (lane & 15)];
      }
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int c = 0; c < 4; c++)
          acc[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[c], acc[a][c], 0, 0, 0);
    }
    if (s + 1 < n_stage) put(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int64_t gi = ci + wr + 16 * a + 4 * e + (lane >> 4);
        const int64_t gj = cj + wc + 16 * c + (lane & 15);
        cm[gi * mc + gj] += acc[a][c][e];
      }
}

int herk_max_buffers() { return 12; }

void launch_herk_f64(hipStream_t st, const double *const *x, const int64_t *rows, double *const *c,
                     int nbuf, int64_t mc) {
  HerkArgs a{};
  for (int b = 0; b < nbuf; b++) {
    a.x[b] = x[b];
    a.rows[b] = rows[b];
    a.c[b] = c[b];
  }
  const int nt = (int)(mc / HK_TB);
  const int n_tiles = nt * (nt + 1) / 2;
  hipLaunchKernelGGL(k_herk_f64, dim3((unsigned)n_tiles, (unsigned)nbuf), dim3(256), 0, st, a, mc,
                     n_tiles);
}

// ---------------------------------------------------------------------------------------------
// Complex B x B matrices from the real-ified lower accumulators:
//   M[a][b] = (C[2a][2b] + C[2a+1][2b+1]) + i (C[2a][2b+1] - C[2a+1][2b]), a >= b, / sum w,
// mirrored as conj above the diagonal. cm: the mass accumulator, ck[0..nk): the stiffness
// accumulators (summed in order). Diagonals are real; clamp_k keeps max(re, 0) for K
// (build_stiffness_matrix :517-523).
// ---------------------------------------------------------------------------------------------
struct FinalizeArgs {
  const double *ck[11];
};

__global__ void k_lap_finalize(const double *__restrict__ cmass, FinalizeArgs fk, int nk, int B,
                               int64_t mc, const double *__restrict__ tw, int clamp_k,
                               double2 *__restrict__ M, double2 *__restrict__ K) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)B * B) return;
  const int a = (int)(e / B), b = (int)(e % B);
  if (a < b) return;
  const double inv = tw[0] > 0.0 ? 1.0 / tw[0] : 1.0;
  const int64_t i0 = (int64_t)(2 * a) * mc + 2 * b, i1 = (int64_t)(2 * a + 1) * mc + 2 * b;
  auto fold = [&](const double *c, double &re, double &im) {
    re += c[i0] + c[i1 + 1];
    im += c[i0 + 1] - c[i1];
  };
  double mr = 0.0, mi = 0.0, kr = 0.0, ki = 0.0;
  fold(cmass, mr, mi);
  for (int q = 0; q < nk; q++) fold(fk.ck[q], kr, ki);
  mr *= inv;
  mi *= inv;
  kr *= inv;
  ki *= inv;
  if (a == b) {
    M[(int64_t)a * B + a] = make_double2(mr, 0.0);
    K[(int64_t)a * B + a] = make_double2(clamp_k ? fmax(kr, 0.0) : kr, 0.0);
    return;
  }
  M[(int64_t)a * B + b] = make_double2(mr, mi);
  M[(int64_t)b * B + a] = make_double2(mr, -mi);
  K[(int64_t)a * B + b] = make_double2(kr, ki);
  K[(int64_t)b * B + a] = make_double2(kr, -ki);
}

void launch_lap_finalize(hipStream_t st, const double *cmass, const double *const *ck, int nk,
                         int B, int64_t mc, const double *tw, int clamp_k, double2 *M, double2 *K) {
  FinalizeArgs f{};
  for (int q = 0; q < nk; q++) f.ck[q] = ck[q];
  const int64_t tot = (int64_t)B * B;
  hipLaunchKernelGGL(k_lap_finalize, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, cmass,
                     f, nk, B, mc, tw, clamp_k, M, K);
}

}  // namespace cy3q
