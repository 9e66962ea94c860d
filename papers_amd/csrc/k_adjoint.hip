// Adjoint of the sigma_L2^2 functional with respect to h, and the Adam step on h (FP64).
//
// Reference: per_point_log_det_gradient (cy3/src/quintic.rs:2311-2695),
// sigma_squared_and_gradient (:2714-2860), sigma_refine_analytic (:453-514),
// adam_update_split (:2071-2117), the Hermitian projection (:2122-2146), pack_h_block
// (:2148-2161) and renormalise_h_trace (:3169-3182).
//
// The reference stores the full per-point gradient (2 N^2 numbers per point). Here the per-point
// metric is recomputed and the gradient is formed by contractions over the points instead:
//   d sigma_L2^2 / dh = (1 / sum w) sum_p gamma_p grad L_p,
//   gamma_p = w_p 2 (eta_p/kappa - 1)(eta_p/kappa) - (sum_q w_q 2 (..)(..)) / (kappa sum w) w_p eta_p,
// which folds the d kappa / dh term of the reference into one weight per point. With Y_p, dK_p,
// q_p, q'_p, tr(Y M) and tr(Y dK dK^H) as in the reference, every term of grad L_p is a complex
// outer product, and all of them collapse into one contraction
//   A = sum_p [ sum_j conj(ds_pj) (x) phi'_pj + conj(s_p) (x) v_p ],
//   phi'_pj = (gamma/K) sum_i Y_ji ds_pi,   v_p = (beta_p/2) s_p - (gamma/K^2) (q_p + conj(q'_p)),
//   beta_p = gamma (-2 tr(YM)/K^2 + 4 tr(Y dK dK^H)/K^3),
// whose Hermitian part H = A + A^H gives d/dh_re = Re H (halved on the diagonal) and
// d/dh_im(a < b) = -Im H, over sum w. The contraction is a real GEMM U^T V over the real-ified
// rows (U = the section and derivative rows, V = the rows built here), on the FP64 MFMA.
#include "cy3q_point.h"

namespace cy3q {

// GR_PPW points per wavefront: the GR_L = 64 / GR_PPW lanes of a point split its dot products
// and its output rows, and the first lane of each group does the point's algebra. pts/log_om/
// w/sv/sd: the solver's point data from point p0; y0/yd: the W-contracted rows of the same
// points; sig: the sigma reduction (kappa, sum w, sum of the gamma numerators). Writes
// phi [n][5][m] and vs [n][m] (zero rows for points that do not contribute).
static constexpr int GR_PPW = 8, GR_L = 64 / GR_PPW;

__global__ void __launch_bounds__(256) k_grad_rows(const double *__restrict__ pts,
                                                   const double *__restrict__ y0,
                                                   const double *__restrict__ yd,
                                                   const double *__restrict__ log_om,
                                                   const double *__restrict__ w,
                                                   const double *__restrict__ sv,
                                                   const double *__restrict__ sd, int64_t n,
                                                   int64_t m, double psi,
                                                   const double *__restrict__ sig,
                                                   double *__restrict__ phi,
                                                   double *__restrict__ vs,
                                                   double *__restrict__ eta_out) {
  // Per point: [0] gamma/K, [1] gamma/K^2, [2] beta/2, [3..52] Y (Y[j][i] at 3 + 2 (5 j + i)),
  // [53..62] v_i, [63..72] w_j.
  __shared__ double cf[4 * GR_PPW][73];
  const int lane = threadIdx.x & 63, sl = lane % GR_L;
  const int slot = (threadIdx.x >> 6) * GR_PPW + lane / GR_L;
  const int64_t p = (int64_t)blockIdx.x * 4 * GR_PPW + slot;
  const bool in = p < n;
  double acc[61];
  point_dots_l<GR_L>(y0 + (in ? p : 0) * m, yd + (in ? p : 0) * 5 * m, in ? m : 0, sl, acc);
  if (sl == 0) {
    double *c = cf[slot];
    for (int q = 0; q < 73; q++) c[q] = 0.0;
    // sig == nullptr: the per-point gradient of log eta alone (gamma = 1, no kappa term).
    const bool unit = sig == nullptr;
    const double kappa = unit ? 1.0 : sig[1], tw = unit ? 1.0 : sig[3], sa2 = unit ? 0.0 : sig[5];
    const double wp = in ? w[p] : 0.0, lo = in ? log_om[p] : NAN;
    double eta_p = NAN;
    if (in && isfinite(wp) && wp > 0.0 && isfinite(lo) && isfinite(kappa) && isfinite(sa2) &&
        tw >= 1e-12) {
      double dk[5][2], ddk[5][5][2];
      unpack_dots(acc, dk, ddk);
      PointMetric pm;
      ambient_from_dots(acc[0], dk, ddk, pm);
      double z[10];
      for (int q = 0; q < 10; q++) z[q] = pts[10 * p + q];
      frame_and_det(z, psi, pm);
      double gi[3][3][2];
      if (isfinite(pm.det) && fabs(pm.det) >= 1e-30) {
        const double eta = exp(log(fabs(pm.det)) - lo);
        if (isfinite(eta) && eta > 0.0) eta_p = eta;
        if (isfinite(eta) && eta > 0.0 && inv3_herm(pm.gt, pm.det, gi)) {
          const double en = eta / kappa;
          const double gamma =
              unit ? 1.0 : wp * 2.0 * (en - 1.0) * en - sa2 / (kappa * tw) * (wp * eta);
          double y[5][5][2];
          y5_from_frame(pm.fr, gi, y);
          double tr_ym = 0.0, tr_yd = 0.0;
          for (int i = 0; i < 5; i++)
            for (int j = 0; j < 5; j++) {
              tr_ym += y[j][i][0] * ddk[i][j][0] - y[j][i][1] * ddk[i][j][1];
              const double pr = dk[i][0] * dk[j][0] + dk[i][1] * dk[j][1];
              const double pi = dk[i][1] * dk[j][0] - dk[i][0] * dk[j][1];
              tr_yd += y[j][i][0] * pr - y[j][i][1] * pi;
            }
          const double ik = 1.0 / pm.k, ik2 = ik * ik, ik3 = ik2 * ik;
          c[0] = gamma * ik;
          c[1] = gamma * ik2;
          c[2] = 0.5 * gamma * (-2.0 * tr_ym * ik2 + 4.0 * tr_yd * ik3);
          for (int j = 0; j < 5; j++)
            for (int i = 0; i < 5; i++) {
              c[3 + 2 * (5 * j + i)] = y[j][i][0];
              c[4 + 2 * (5 * j + i)] = y[j][i][1];
            }
          for (int i = 0; i < 5; i++) {  // v_i = sum_j Y_ji conj(dK_j)
            double sr = 0.0, si = 0.0;
            for (int j = 0; j < 5; j++) {
              sr += y[j][i][0] * dk[j][0] + y[j][i][1] * dk[j][1];
              si += y[j][i][1] * dk[j][0] - y[j][i][0] * dk[j][1];
            }
            c[53 + 2 * i] = sr;
            c[54 + 2 * i] = si;
          }
          for (int j = 0; j < 5; j++) {  // w_j = sum_i Y_ji dK_i
            double sr = 0.0, si = 0.0;
            for (int i = 0; i < 5; i++) {
              sr += y[j][i][0] * dk[i][0] - y[j][i][1] * dk[i][1];
              si += y[j][i][0] * dk[i][1] + y[j][i][1] * dk[i][0];
            }
            c[63 + 2 * j] = sr;
            c[64 + 2 * j] = si;
          }
        }
      }
    }
    if (eta_out && in) eta_out[p] = eta_p;
  }
  __syncthreads();
  if (!in) return;
  const double *c = cf[slot];
  const double gk = c[0], gk2 = c[1], hb = c[2];
  const double *sp = sv + p * m, *dp = sd + p * 5 * m;
  double *ph = phi + p * 5 * m, *vp = vs + p * m;
  for (int64_t col = 2 * sl; col < m; col += 2 * GR_L) {
    double dr[5], di[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
      dr[i] = dp[i * m + col];
      di[i] = dp[i * m + col + 1];
    }
#pragma unroll
    for (int j = 0; j < 5; j++) {  // phi'_j = (gamma/K) sum_i Y_ji ds_i
      double sr = 0.0, si = 0.0;
#pragma unroll
      for (int i = 0; i < 5; i++) {
        const double yr = c[3 + 2 * (5 * j + i)], yi = c[4 + 2 * (5 * j + i)];
        sr += yr * dr[i] - yi * di[i];
        si += yr * di[i] + yi * dr[i];
      }
      ph[j * m + col] = gk * sr;
      ph[j * m + col + 1] = gk * si;
    }
    double qr = 0.0, qi = 0.0, pr = 0.0, pi = 0.0;
#pragma unroll
    for (int i = 0; i < 5; i++) {
      const double vr = c[53 + 2 * i], vi = c[54 + 2 * i];
      qr += dr[i] * vr - di[i] * vi;  // q = sum_i ds_i v_i
      qi += dr[i] * vi + di[i] * vr;
      const double wr = c[63 + 2 * i], wi = c[64 + 2 * i];
      pr += dr[i] * wr + di[i] * wi;  // q' = sum_j conj(ds_j) w_j
      pi += dr[i] * wi - di[i] * wr;
    }
    // v = (beta/2) s - (gamma/K^2) (q + conj(q'))
    vp[col] = hb * sp[col] - gk2 * (qr + pr);
    vp[col + 1] = hb * sp[col + 1] - gk2 * (qi - pi);
  }
}

void launch_grad_rows(hipStream_t st, const double *pts, const double *y0, const double *yd,
                      const double *log_om, const double *w, const double *sv, const double *sd,
                      int64_t n, int64_t m_cols, double psi, const double *sig, double *phi,
                      double *vs, double *eta_out) {
  if (n <= 0) return;
  const int64_t per_wg = 4 * GR_PPW;
  hipLaunchKernelGGL(k_grad_rows, dim3((unsigned)((n + per_wg - 1) / per_wg)), dim3(256), 0, st,
                     pts, y0, yd, log_om, w, sv, sd, n, m_cols, psi, sig, phi, vs, eta_out);
}

// ---------------------------------------------------------------------------------------------
// The reference's per-point buffer (cpu_compute_per_point_reference, gpu_adam.rs:2677-2739):
// row p of out ([n][2 + 2 nb^2]) is (eta_p, w_p, grad L_p) with grad L_p = d log eta_p / dh in
// the gradient layout (d/dh_re with the diagonal halved, then d/dh_im for a < b). From the
// gamma = 1 rows of k_grad_rows: A_ab = sum_j conj(ds_j)_a phi'_jb + conj(s_a) v_b over the six
// row pairs of the point; grad_re = Re(A_ab + A_ba), halved on the diagonal, and grad_im =
// -Im(A_ab - A_ba) for a < b, as k_grad_assemble. A point with an invalid weight or eta gets (NaN, 0, 0...); a point whose
// tangent metric does not invert keeps eta and w with a zero gradient, as the reference.
// One workgroup per point; its 12 rows (2 nb doubles each) are staged in LDS.
// ---------------------------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_per_point_grad(const double *__restrict__ sv,
                                                        const double *__restrict__ sd,
                                                        const double *__restrict__ phi,
                                                        const double *__restrict__ vs,
                                                        const double *__restrict__ eta,
                                                        const double *__restrict__ w, int nb,
                                                        int64_t m, double *__restrict__ out) {
  extern __shared__ double rows[];  // u[6][2 nb], v[6][2 nb]: (ds_0..ds_4, s), (phi'_0..4, v)
  const int64_t p = blockIdx.x;
  const int two_n = 2 * nb;
  const int64_t nn = (int64_t)nb * nb, stride = 2 + 2 * nn;
  double *o = out + p * stride;
  const double e = eta[p], wp = w[p];
  const bool valid = isfinite(e) && e > 0.0 && isfinite(wp) && wp > 0.0;
  for (int q = threadIdx.x; q < 6 * two_n; q += blockDim.x) {
    const int r = q / two_n, c = q % two_n;
    rows[q] = r < 5 ? sd[(p * 5 + r) * m + c] : sv[p * m + c];
    rows[6 * two_n + q] = r < 5 ? phi[(p * 5 + r) * m + c] : vs[p * m + c];
  }
  if (threadIdx.x == 0) {
    o[0] = valid ? e : NAN;
    o[1] = valid ? wp : 0.0;
  }
  __syncthreads();
  const double *u = rows, *v = rows + 6 * two_n;
  for (int64_t q = threadIdx.x; q < nn; q += blockDim.x) {
    const int a = (int)(q / nb), b = (int)(q % nb);
    double gre = 0.0, gim = 0.0;
    if (valid) {
      // A_xy = sum_r conj(u_r[x]) v_r[y]
      double abr = 0.0, abi = 0.0, bar = 0.0, bai = 0.0;
#pragma unroll
      for (int r = 0; r < 6; r++) {
        const double *ur = u + r * two_n, *vr = v + r * two_n;
        const double uar = ur[2 * a], uai = ur[2 * a + 1], ubr = ur[2 * b], ubi = ur[2 * b + 1];
        const double var_ = vr[2 * a], vai = vr[2 * a + 1], vbr = vr[2 * b], vbi = vr[2 * b + 1];
        abr += uar * vbr + uai * vbi;
        abi += uar * vbi - uai * vbr;
        bar += ubr * var_ + ubi * vai;
        bai += ubr * vai - ubi * var_;
      }
      gre = abr + bar;
      if (a == b) gre *= 0.5;
      gim = a < b ? -(abi - bai) : 0.0;
    }
    o[2 + q] = gre;
    o[2 + nn + q] = gim;
  }
}

int per_point_grad_max_nb() { return 340; }  // 12 * 2 * 340 * 8 bytes < 64 KB of LDS

void launch_per_point_grad(hipStream_t st, const double *sv, const double *sd, const double *phi,
                           const double *vs, const double *eta, const double *w, int nb,
                           int64_t m_cols, int64_t n, double *out) {
  if (n <= 0) return;
  const size_t lds = sizeof(double) * 12 * 2 * (size_t)nb;
  hipLaunchKernelGGL(k_per_point_grad, dim3((unsigned)n), dim3(256), lds, st, sv, sd, phi, vs, eta,
                     w, nb, m_cols, out);
}

// ---------------------------------------------------------------------------------------------
// C_g = sum over the rows r of group g of U[r]^T V[r] (U, V: [rows][m] row-major, FP64). A
// workgroup computes a 128 x 128 tile of C over its rows: 4 waves, each a 64 x 64 quadrant as
// 4 x 4 tiles of v_mfma_f64_16x16x4f64, 16-row stages in an LDS double buffer filled from
// registers prefetched one stage ahead. Output: part[g][m][m] (columns beyond m are not written).
// v_mfma_f64_16x16x4f64 layout: A[i][k] at lane i + 16 k, B[k][j] at lane j + 16 k,
// D[i][j] at lane j + 16 (i % 4), element i / 4.
// ---------------------------------------------------------------------------------------------
static constexpr int GT = 128;  // output tile edge
static constexpr int GKB = 16;  // rows per stage
static constexpr int GLD = GT + 2;

__global__ void __launch_bounds__(256) k_gemm_tn_f64(const double *__restrict__ u,
                                                     const double *__restrict__ v, int64_t rows,
                                                     int64_t m, int64_t rows_per_group,
                                                     double *__restrict__ part) {
  __shared__ double sa[2][GKB][GLD], sb[2][GKB][GLD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t c0 = (int64_t)blockIdx.x * GT, d0 = (int64_t)blockIdx.y * GT;
  const int64_t g = blockIdx.z;
  const int64_t r0 = g * rows_per_group;
  const int64_t r1 = r0 + rows_per_group < rows ? r0 + rows_per_group : rows;
  const int wr = 64 * (wave >> 1), wc = 64 * (wave & 1);
  double4_t acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int c = 0; c < 4; c++) acc[a][c] = (double4_t){0.0, 0.0, 0.0, 0.0};
  const int64_t n_stage = r1 > r0 ? (r1 - r0 + GKB - 1) / GKB : 0;
  // Thread tid moves pairs e = tid + 256 q: row e >> 6 of the stage, columns 2 (e & 63) + {0, 1}
  // (m is even, so a pair is inside or outside the matrix as a whole).
  double2 ra[4], rb[4];
  auto fetch = [&](int64_t s) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int e = tid + 256 * q;
      const int64_t row = r0 + s * GKB + (e >> 6);
      const int cc = (e & 63) * 2;
      const bool in = row < r1;
      ra[q] = (in && c0 + cc < m) ? *reinterpret_cast<const double2 *>(u + row * m + c0 + cc)
                                  : make_double2(0.0, 0.0);
      rb[q] = (in && d0 + cc < m) ? *reinterpret_cast<const double2 *>(v + row * m + d0 + cc)
                                  : make_double2(0.0, 0.0);
    }
  };
  auto put = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int e = tid + 256 * q;
      const int r = e >> 6, cc = (e & 63) * 2;
      sa[buf][r][cc] = ra[q].x;
      sa[buf][r][cc + 1] = ra[q].y;
      sb[buf][r][cc] = rb[q].x;
      sb[buf][r][cc + 1] = rb[q].y;
    }
  };
  if (n_stage > 0) {
    fetch(0);
    put(0);
  }
  __syncthreads();
  for (int64_t s = 0; s < n_stage; s++) {
    const int buf = (int)(s & 1);
    if (s + 1 < n_stage) fetch(s + 1);
#pragma unroll
    for (int kk = 0; kk < GKB / 4; kk++) {
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
  double *out = part + g * m * m;
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int64_t i = c0 + wr + 16 * a + 4 * e + (lane >> 4);
        const int64_t j = d0 + wc + 16 * c + (lane & 15);
        if (i < m && j < m) out[i * m + j] = acc[a][c][e];
      }
}

// acc[i] += sum_g part[g][i], in group order.
// acc += sum of the group partials. Eight interleaved running sums (group g into sum g % 8,
// then the eight in a fixed tree) keep eight loads in flight per thread; the order is fixed, so
// the result is the same run to run.
__global__ void k_gemm_accumulate(const double *__restrict__ part, int64_t n_groups, int64_t mm,
                                  double *__restrict__ acc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= mm) return;
  double r[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int64_t g = 0;
  for (; g + 8 <= n_groups; g += 8) {
#pragma unroll
    for (int u = 0; u < 8; u++) r[u] += part[(g + u) * mm + i];
  }
  for (int u = 0; g < n_groups; g++, u++) r[u] += part[g * mm + i];
  const double s = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  acc[i] += s;
}

int64_t gemm_tn_groups(int64_t rows, int64_t m_cols) {
  const int64_t nt = (m_cols + GT - 1) / GT;
  int64_t g = std::max<int64_t>(1024 / (nt * nt), 1);
  g = std::min<int64_t>(g, std::max<int64_t>(rows / 256, 1));
  return g;
}

void launch_gemm_tn_acc(hipStream_t st, const double *u, const double *v, int64_t rows,
                        int64_t m_cols, int64_t n_groups, double *part, double *acc) {
  if (rows <= 0) return;
  const int64_t nt = (m_cols + GT - 1) / GT;
  const int64_t rpg = ((rows + n_groups - 1) / n_groups + GKB - 1) / GKB * GKB;
  hipLaunchKernelGGL(k_gemm_tn_f64, dim3((unsigned)nt, (unsigned)nt, (unsigned)n_groups),
                     dim3(256), 0, st, u, v, rows, m_cols, rpg, part);
  const int64_t mm = m_cols * m_cols;
  hipLaunchKernelGGL(k_gemm_accumulate, dim3((unsigned)((mm + 255) / 256)), dim3(256), 0, st, part,
                     n_groups, mm, acc);
}

// grad[a nb + b] = d sigma^2 / d h_re(a, b), grad[nb^2 + a nb + b] = d sigma^2 / d h_im(a, b)
// from the real-ified contraction C (m x m): A_ab = (C[2a][2b] + C[2a+1][2b+1])
// + i (C[2a][2b+1] - C[2a+1][2b]), H = A + A^H.
__global__ void k_grad_assemble(const double *__restrict__ cm, int nb, int64_t m,
                                const double *__restrict__ sig, double *__restrict__ grad) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nb * nb) return;
  const int a = (int)(e / nb), b = (int)(e % nb);
  const double tw = sig[3];
  const bool ok = isfinite(sig[1]) && isfinite(sig[5]) && tw >= 1e-12;
  auto ar = [&](int x, int y) {
    return cm[(2 * (int64_t)x) * m + 2 * y] + cm[(2 * (int64_t)x + 1) * m + 2 * y + 1];
  };
  auto ai = [&](int x, int y) {
    return cm[(2 * (int64_t)x) * m + 2 * y + 1] - cm[(2 * (int64_t)x + 1) * m + 2 * y];
  };
  const double hr = ar(a, b) + ar(b, a), hi = ai(a, b) - ai(b, a);
  double gre = ok ? hr / tw : 0.0, gim = ok ? -hi / tw : 0.0;
  if (a == b) gre *= 0.5;
  if (a >= b) gim = 0.0;
  grad[e] = gre;
  grad[(int64_t)nb * nb + e] = gim;
}

void launch_grad_assemble(hipStream_t st, const double *cm, int nb, int64_t m_cols,
                          const double *sig, double *grad) {
  const int64_t nn = (int64_t)nb * nb;
  hipLaunchKernelGGL(k_grad_assemble, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, cm, nb,
                     m_cols, sig, grad);
}

// ---------------------------------------------------------------------------------------------
// Adam on (h_re, h_im) = (Re G, Im G): the update rule written inline at quintic.rs:5421-5437
// (t <- t + 1, scale = lr sqrt(1 - b2^t) / (1 - b1^t), m, v moments, x -= scale m / (sqrt v + eps)),
// both halves with the same t, then the Hermitian projection (h_re symmetrised, h_im
// antisymmetrised with a zero diagonal), written back into G (complex [nb][nb]). One thread per
// pair a <= b.
// ---------------------------------------------------------------------------------------------
__global__ void k_adam_herm(double *__restrict__ g, int nb, const double *__restrict__ grad,
                            double *__restrict__ mo, double *__restrict__ ve, double scale,
                            double beta1, double beta2, double eps) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nb * nb) return;
  const int a = (int)(e / nb), b = (int)(e % nb);
  if (a > b) return;
  const int64_t n2 = (int64_t)nb * nb;
  const int64_t ab = (int64_t)a * nb + b, ba = (int64_t)b * nb + a;
  auto step = [&](int64_t k, double x) {
    const double gk = grad[k];
    mo[k] = beta1 * mo[k] + (1.0 - beta1) * gk;
    ve[k] = beta2 * ve[k] + (1.0 - beta2) * gk * gk;
    return x - scale * mo[k] / (sqrt(ve[k]) + eps);
  };
  if (a == b) {
    const double re = step(ab, g[2 * ab]);
    step(n2 + ab, g[2 * ab + 1]);  // moments advance; the diagonal of h_im is then zeroed
    g[2 * ab] = re;
    g[2 * ab + 1] = 0.0;
    return;
  }
  const double re_ab = step(ab, g[2 * ab]), re_ba = step(ba, g[2 * ba]);
  const double im_ab = step(n2 + ab, g[2 * ab + 1]), im_ba = step(n2 + ba, g[2 * ba + 1]);
  const double re = 0.5 * (re_ab + re_ba), im = 0.5 * (im_ab - im_ba);
  g[2 * ab] = re;
  g[2 * ab + 1] = im;
  g[2 * ba] = re;
  g[2 * ba + 1] = -im;
}

// Trace renormalisation to tr Re G = nb (when the trace exceeds 1e-10). One workgroup.
__global__ void __launch_bounds__(256) k_trace_renorm(double *__restrict__ g, int nb) {
  __shared__ double sh[256];
  double t = 0.0;
  for (int a = threadIdx.x; a < nb; a += 256) t += g[2 * ((int64_t)a * nb + a)];
  sh[threadIdx.x] = t;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  const double tr = sh[0];
  if (!(tr > 1e-10)) return;
  const double sc = (double)nb / tr;
  for (int64_t e = threadIdx.x; e < 2 * (int64_t)nb * nb; e += 256) g[e] *= sc;
}

void launch_trace_renorm(hipStream_t st, double *g_cplx, int nb) {
  hipLaunchKernelGGL(k_trace_renorm, dim3(1), dim3(256), 0, st, g_cplx, nb);
}

void launch_adam_step(hipStream_t st, double *g_cplx, int nb, const double *grad, double *m,
                      double *v, double scale, double beta1, double beta2, double eps) {
  const int64_t nn = (int64_t)nb * nb;
  hipLaunchKernelGGL(k_adam_herm, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, g_cplx, nb,
                     grad, m, v, scale, beta1, beta2, eps);
  launch_trace_renorm(st, g_cplx, nb);
}

}  // namespace cy3q
