// Sampling, tangent frames and monomial sections on the Tian-Yau and Schoen threefolds, the
// set-up of the reference's 8-coordinate route34 callers.
//
// Reference: CicySampler (cy3/src/cicy_sampler.rs:374-852) and its batch Newton kernel
// (cy3/src/gpu_sampler.rs:60-412); SchoenSampler (route34/schoen_sampler.rs:384-1172);
// ty_affine_chart_frame (route34/ty_metric.rs:1929-2015) and schoen_affine_chart_frame
// (route34/schoen_metric.rs:2202-2273); the section tables (ty_metric.rs:897-1035,
// schoen_metric.rs:1214-1333).
//
// The random draws are a stream of standard normals (the reference's ChaCha8 + ziggurat
// stream, drawn by k_zig_fast / k_zig_gather below). A line that starts at normal s takes its
// origin from normals [s, s + 16), its directions from the next 16 per hypersurface, and
// restart a's initial t from the 2 nh normals after those. Two ways of consuming the stream are reproduced:
//   - restarts (the reference's sequential CPU samplers): a line runs up to 4 Newton restarts
//     and the next line starts after the restarts it used. Its start therefore depends on every
//     line before it. k_cicy_span evaluates a line at every even start in a window at once
//     (each lane one candidate start, all restarts); the host then follows the one chain of
//     starts that the stream takes, and k_cicy_chain recomputes the converged restart of each
//     line on the chain;
//   - lanes (the reference's batch sampler of its GPU build): line l always starts at
//     normal 70 l and runs one damped Newton attempt (k_cicy_lanes).
// k_cicy_post then does the patch rescale, the residue Omega and the pull-back weight of each
// converged line, k_cicy_frames the implicit-function tangent frame of each accepted point.
//
// The normal stream itself is drawn here too: k_zig_fast evaluates, at every position of the
// u64 keystream, the ziggurat's first test (a normal that starts there and takes one word), and
// lists the positions where it fails; the host takes those (the wedge and tail use exp and log,
// whose last bits are the host library's), follows the stream through them, and k_zig_gather
// writes the normals from the resulting runs.
//
// Bit-for-bit agreement with the CPU restatement: every kernel here is compiled with
// -ffp-contract=off and keeps the reference's operation order (num-complex products, naive
// complex division, hypot moduli), so the accept decisions and the points equal the CPU's.
#include "cy3q_internal.h"

namespace cy3q {
namespace {

struct cx {
  double re, im;
};

__device__ inline cx C(double re, double im) { return cx{re, im}; }
__device__ inline cx cadd(cx a, cx b) { return C(a.re + b.re, a.im + b.im); }
__device__ inline cx csub(cx a, cx b) { return C(a.re - b.re, a.im - b.im); }
__device__ inline cx cmul(cx a, cx b) {
  return C(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
__device__ inline cx cdiv(cx a, cx b) {
  const double n = b.re * b.re + b.im * b.im;
  const double re = a.re * b.re + a.im * b.im;
  const double im = a.im * b.re - a.re * b.im;
  return C(re / n, im / n);
}
__device__ inline cx cneg(cx a) { return C(-a.re, -a.im); }
__device__ inline cx cconj(cx a) { return C(a.re, -a.im); }
__device__ inline double cnorm(cx a) { return hypot(a.re, a.im); }
__device__ inline double cnsq(cx a) { return a.re * a.re + a.im * a.im; }

// The defining polynomials, every coefficient 1: [variety][polynomial][term][coordinate].
// TY (BicubicPair::z3_invariant_default, cicy_sampler.rs:244-275): sum z^3, sum w^3, sum z w.
// Schoen (SchoenPoly::z3xz3_invariant_default, schoen_sampler.rs:188-217):
// (x0^3 + x1^3 + x2^3) t0 + x0 x1 x2 t1 and (y0^3 + y1^3 + y2^3) t1 + y0 y1 y2 t0.
__constant__ unsigned char c_exps[2][3][4][8] = {
    {{{3, 0, 0, 0, 0, 0, 0, 0}, {0, 3, 0, 0, 0, 0, 0, 0}, {0, 0, 3, 0, 0, 0, 0, 0}, {0, 0, 0, 3, 0, 0, 0, 0}},
     {{0, 0, 0, 0, 3, 0, 0, 0}, {0, 0, 0, 0, 0, 3, 0, 0}, {0, 0, 0, 0, 0, 0, 3, 0}, {0, 0, 0, 0, 0, 0, 0, 3}},
     {{1, 0, 0, 0, 1, 0, 0, 0}, {0, 1, 0, 0, 0, 1, 0, 0}, {0, 0, 1, 0, 0, 0, 1, 0}, {0, 0, 0, 1, 0, 0, 0, 1}}},
    {{{3, 0, 0, 0, 0, 0, 1, 0}, {0, 3, 0, 0, 0, 0, 1, 0}, {0, 0, 3, 0, 0, 0, 1, 0}, {1, 1, 1, 0, 0, 0, 0, 1}},
     {{0, 0, 0, 3, 0, 0, 0, 1}, {0, 0, 0, 0, 3, 0, 0, 1}, {0, 0, 0, 0, 0, 3, 0, 1}, {0, 0, 0, 1, 1, 1, 1, 0}},
     {{0}}}};

template <int V>
struct Var {
  static constexpr int nh = V == 0 ? 3 : 2;    // hypersurfaces
  static constexpr int nfac = V == 0 ? 2 : 3;  // projective factors
  static constexpr int line_normals = 16 * (1 + nh);
  static constexpr int restart_normals = 2 * nh;
  __device__ static int rs(int f) { return V == 0 ? 4 * f : 3 * f; }
  __device__ static int re(int f) { return V == 0 ? 4 * f + 4 : (f == 2 ? 8 : 3 * f + 3); }
};

// eval_monomial / eval_poly / poly_partial (cicy_sampler.rs:300-350)
__device__ inline cx eval_monomial(const cx *pt, const unsigned char *e) {
  cx acc = C(1.0, 0.0);
  for (int k = 0; k < 8; k++) {
    const int ek = e[k];
    if (ek == 0) continue;
    cx p = pt[k];
    for (int j = 1; j < ek; j++) p = cmul(p, pt[k]);
    acc = cmul(acc, p);
  }
  return acc;
}

template <int V>
__device__ inline cx eval_poly(int i, const cx *pt) {
  cx s = C(0.0, 0.0);
  for (int t = 0; t < 4; t++) s = cadd(s, cmul(C(1.0, 0.0), eval_monomial(pt, c_exps[V][i][t])));
  return s;
}

template <int V>
__device__ inline cx poly_partial(int i, const cx *pt, int k) {
  cx s = C(0.0, 0.0);
  for (int t = 0; t < 4; t++) {
    const int ek = c_exps[V][i][t][k];
    if (ek == 0) continue;
    unsigned char ne[8];
    for (int j = 0; j < 8; j++) ne[j] = c_exps[V][i][t][j];
    ne[k] = (unsigned char)(ek - 1);
    const cx mono = eval_monomial(pt, ne);
    s = cadd(s, cmul(cmul(C(1.0, 0.0), C((double)ek, 0.0)), mono));
  }
  return s;
}

template <int V>
__device__ inline void jacobian(const cx *pt, cx *jac) {
  for (int k = 0; k < 8; k++)
    for (int i = 0; i < Var<V>::nh; i++) jac[i * 8 + k] = poly_partial<V>(i, pt, k);
}

// det_complex_lu (cicy_sampler.rs:959-998)
template <int N>
__device__ inline cx det_lu(const cx *mat) {
  cx a[N * N];
  for (int i = 0; i < N * N; i++) a[i] = mat[i];
  cx det = C(1.0, 0.0);
  for (int k = 0; k < N; k++) {
    int piv = k;
    double best = cnorm(a[k * N + k]);
    for (int p = k + 1; p < N; p++) {
      const double v = cnorm(a[p * N + k]);
      if (v > best) {
        best = v;
        piv = p;
      }
    }
    if (best == 0.0) return C(0.0, 0.0);
    if (piv != k) {
      for (int j = 0; j < N; j++) {
        const cx t = a[k * N + j];
        a[k * N + j] = a[piv * N + j];
        a[piv * N + j] = t;
      }
      det = cneg(det);
    }
    const cx pv = a[k * N + k];
    det = cmul(det, pv);
    const cx ip = cdiv(C(1.0, 0.0), pv);
    for (int i = k + 1; i < N; i++) {
      const cx factor = cmul(a[i * N + k], ip);
      for (int j = k + 1; j < N; j++) a[i * N + j] = csub(a[i * N + j], cmul(factor, a[k * N + j]));
      a[i * N + k] = C(0.0, 0.0);
    }
  }
  return det;
}

// inv_n / invert_n_complex: Gauss-Jordan, pivot floor 1e-30 (cicy_sampler.rs:1000-1064)
template <int N>
__device__ inline bool inv_n(const cx *mat, cx *inv) {
  constexpr int cols = 2 * N;
  cx a[N * cols];
  for (int i = 0; i < N; i++)
    for (int j = 0; j < cols; j++)
      a[i * cols + j] = j < N ? mat[i * N + j] : C(j - N == i ? 1.0 : 0.0, 0.0);
  for (int k = 0; k < N; k++) {
    int piv = k;
    double best = cnorm(a[k * cols + k]);
    for (int p = k + 1; p < N; p++) {
      const double v = cnorm(a[p * cols + k]);
      if (v > best) {
        best = v;
        piv = p;
      }
    }
    if (best < 1.0e-30) return false;
    if (piv != k)
      for (int j = 0; j < cols; j++) {
        const cx t = a[k * cols + j];
        a[k * cols + j] = a[piv * cols + j];
        a[piv * cols + j] = t;
      }
    const cx ip = cdiv(C(1.0, 0.0), a[k * cols + k]);
    for (int j = 0; j < cols; j++) a[k * cols + j] = cmul(a[k * cols + j], ip);
    for (int i = 0; i < N; i++) {
      if (i == k) continue;
      const cx factor = a[i * cols + k];
      if (cnorm(factor) == 0.0) continue;
      for (int j = 0; j < cols; j++) a[i * cols + j] = csub(a[i * cols + j], cmul(a[k * cols + j], factor));
    }
  }
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) inv[i * N + j] = a[i * cols + N + j];
  return true;
}

// gaussian_unit_sphere_c8 (cicy_sampler.rs:883-901) from 16 consecutive normals
__device__ inline void unit_sphere(const double *nrm, cx *v) {
  double sq = 0.0;
  for (int k = 0; k < 8; k++) {
    const double re = nrm[2 * k], im = nrm[2 * k + 1];
    v[k] = C(re, im);
    sq += re * re + im * im;
  }
  const double n = sqrt(sq);
  if (n > 0.0)
    for (int k = 0; k < 8; k++) v[k] = C(v[k].re / n, v[k].im / n);
}

template <int NH>
__device__ inline void line_eval(const cx *p, const cx (*dirs)[8], const cx *t, cx *x) {
  for (int k = 0; k < 8; k++) x[k] = p[k];
  for (int j = 0; j < NH; j++)
    for (int k = 0; k < 8; k++) x[k] = cadd(x[k], cmul(t[j], dirs[j][k]));
}

template <int NH>
__device__ inline double residual_norm(const cx *f) {
  double s = 0.0;
  for (int i = 0; i < NH; i++) s += cnsq(f[i]);
  return sqrt(s);
}

// newton_solve_line with Armijo back-tracking (cicy_sampler.rs:771-853,
// schoen_sampler.rs:643-713). The TY solver gives up on |det M| < 1e-30 before inverting.
template <int V>
__device__ bool newton_restart(const cx *p, const cx (*dirs)[8], cx *t, cx *x) {
  constexpr int NH = Var<V>::nh;
  cx f[NH];
  line_eval<NH>(p, dirs, t, x);
  for (int i = 0; i < NH; i++) f[i] = eval_poly<V>(i, x);
  double res = residual_norm<NH>(f);
  for (int it = 0; it < 64; it++) {
    if (res < 1.0e-8) return true;
    cx jac[NH * 8], m[NH * NH], minv[NH * NH], dt[NH];
    jacobian<V>(x, jac);
    for (int i = 0; i < NH; i++)
      for (int j = 0; j < NH; j++) {
        cx s = C(0.0, 0.0);
        for (int k = 0; k < 8; k++) s = cadd(s, cmul(jac[i * 8 + k], dirs[j][k]));
        m[i * NH + j] = s;
      }
    if (V == 0 && cnorm(det_lu<NH>(m)) < 1.0e-30) return false;
    if (!inv_n<NH>(m, minv)) return false;
    for (int r = 0; r < NH; r++) {
      cx s = C(0.0, 0.0);
      for (int c = 0; c < NH; c++) s = cadd(s, cmul(minv[r * NH + c], f[c]));
      dt[r] = cneg(s);
    }
    double step = 1.0;
    bool accepted = false;
    for (int ls = 0; ls < 32; ls++) {
      cx tt[NH], xt[8], ft[NH];
      for (int r = 0; r < NH; r++) tt[r] = cadd(t[r], cmul(C(step, 0.0), dt[r]));
      line_eval<NH>(p, dirs, tt, xt);
      for (int i = 0; i < NH; i++) ft[i] = eval_poly<V>(i, xt);
      const double rt = residual_norm<NH>(ft);
      if (rt < (1.0 - 1.0e-4 * step) * res) {
        for (int r = 0; r < NH; r++) t[r] = tt[r];
        for (int k = 0; k < 8; k++) x[k] = xt[k];
        for (int i = 0; i < NH; i++) f[i] = ft[i];
        res = rt;
        accepted = true;
        break;
      }
      step *= 0.5;
    }
    if (!accepted) return false;
  }
  return res < 1.0e-8;
}

template <int V>
__device__ inline void load_line(const double *nrm, cx *p, cx (*dirs)[8]) {
  unit_sphere(nrm, p);
  for (int j = 0; j < Var<V>::nh; j++) unit_sphere(nrm + 16 * (1 + j), dirs[j]);
}

// One candidate start per lane: restarts used (1..4) in bits 0-2, converged in bit 3.
template <int V>
__global__ void __launch_bounds__(128) k_cicy_span(const double *__restrict__ nrm, int64_t s0,
                                                   int64_t n_cand, uint8_t *__restrict__ info) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cand) return;
  constexpr int NH = Var<V>::nh;
  const double *base = nrm + s0 + 2 * c;
  cx p[8], dirs[NH][8];
  load_line<V>(base, p, dirs);
  int a = 0;
  bool conv = false;
  for (a = 0; a < 4; a++) {
    const double *tn = base + Var<V>::line_normals + Var<V>::restart_normals * a;
    cx t[NH], x[8];
    for (int j = 0; j < NH; j++) t[j] = C(tn[2 * j], tn[2 * j + 1]);
    if (newton_restart<V>(p, dirs, t, x)) {
      conv = true;
      a++;
      break;
    }
  }
  info[c] = (uint8_t)(a | (conv ? 8 : 0));
}

// The converged restart of each line on the chain, again: x [16] and a flag.
template <int V>
__global__ void __launch_bounds__(128) k_cicy_chain(const double *__restrict__ nrm,
                                                    const int64_t *__restrict__ start,
                                                    const uint8_t *__restrict__ info, int64_t n_lines,
                                                    double *__restrict__ xout, int32_t *__restrict__ flag) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n_lines) return;
  constexpr int NH = Var<V>::nh;
  const uint8_t inf = info[l];
  flag[l] = 0;
  if (!(inf & 8)) return;
  const double *base = nrm + start[l];
  cx p[8], dirs[NH][8], t[NH], x[8];
  load_line<V>(base, p, dirs);
  const double *tn = base + Var<V>::line_normals + Var<V>::restart_normals * ((inf & 7) - 1);
  for (int j = 0; j < NH; j++) t[j] = C(tn[2 * j], tn[2 * j + 1]);
  if (!newton_restart<V>(p, dirs, t, x)) return;
  for (int k = 0; k < 8; k++) {
    xout[16 * l + 2 * k] = x[k].re;
    xout[16 * l + 2 * k + 1] = x[k].im;
  }
  flag[l] = 1;
}

// ---- One damped Newton attempt per line (gpu_sampler.rs:60-412), TY only -------------------

__device__ inline cx cpow_sq(cx b, int n) {  // repeated squaring
  double rr = 1.0, ri = 0.0, br = b.re, bi = b.im;
  while (n > 0) {
    if (n & 1) {
      const double nr = rr * br - ri * bi, ni = rr * bi + ri * br;
      rr = nr;
      ri = ni;
    }
    n >>= 1;
    if (n > 0) {
      const double nbr = br * br - bi * bi, nbi = 2.0 * br * bi;
      br = nbr;
      bi = nbi;
    }
  }
  return C(rr, ri);
}

__device__ inline cx lane_value(int i, const cx *z) {
  double pr = 0.0, pi = 0.0;
  for (int m = 0; m < 4; m++) {
    double tr = 1.0, ti = 0.0;
    for (int j = 0; j < 8; j++) {
      const int e = c_exps[0][i][m][j];
      if (e == 0) continue;
      const cx pw = cpow_sq(z[j], e);
      const double nr = tr * pw.re - ti * pw.im, ni = tr * pw.im + ti * pw.re;
      tr = nr;
      ti = ni;
    }
    pr += tr;
    pi += ti;
  }
  return C(pr, pi);
}

// value and gradient d_k V = e_k V / z_k
__device__ inline cx lane_value_grad(int i, const cx *z, cx *g) {
  double vra = 0.0, via = 0.0;
  for (int k = 0; k < 8; k++) g[k] = C(0.0, 0.0);
  for (int m = 0; m < 4; m++) {
    double vr = 1.0, vi = 0.0;
    for (int j = 0; j < 8; j++) {
      const int e = c_exps[0][i][m][j];
      if (e == 0) continue;
      const cx pw = cpow_sq(z[j], e);
      const double nr = vr * pw.re - vi * pw.im, ni = vr * pw.im + vi * pw.re;
      vr = nr;
      vi = ni;
    }
    vra += vr;
    via += vi;
    for (int k = 0; k < 8; k++) {
      const int e = c_exps[0][i][m][k];
      if (e == 0) continue;
      const double m2 = z[k].re * z[k].re + z[k].im * z[k].im;
      if (m2 < 1.0e-300) continue;
      const double ir = z[k].re / m2, ii = -z[k].im / m2;
      const double dr = vr * ir - vi * ii, di = vr * ii + vi * ir;
      g[k].re += (double)e * dr;
      g[k].im += (double)e * di;
    }
  }
  return C(vra, via);
}

__device__ inline bool lane_solve3(double *mr, double *mi, double *br, double *bi) {
  constexpr int N = 3;
  for (int k = 0; k < N; k++) {
    int p = k;
    double best = mr[k * N + k] * mr[k * N + k] + mi[k * N + k] * mi[k * N + k];
    for (int i = k + 1; i < N; i++) {
      const double mag = mr[i * N + k] * mr[i * N + k] + mi[i * N + k] * mi[i * N + k];
      if (mag > best) {
        best = mag;
        p = i;
      }
    }
    if (best < 1.0e-300) return false;
    if (p != k) {
      for (int j = 0; j < N; j++) {
        const double tr = mr[k * N + j], ti = mi[k * N + j];
        mr[k * N + j] = mr[p * N + j];
        mi[k * N + j] = mi[p * N + j];
        mr[p * N + j] = tr;
        mi[p * N + j] = ti;
      }
      const double tr = br[k], ti = bi[k];
      br[k] = br[p];
      bi[k] = bi[p];
      br[p] = tr;
      bi[p] = ti;
    }
    const double pr = mr[k * N + k], pi = mi[k * N + k], pm = pr * pr + pi * pi;
    for (int i = k + 1; i < N; i++) {
      const double ar = mr[i * N + k], ai = mi[i * N + k];
      const double fr = (ar * pr + ai * pi) / pm, fi = (ai * pr - ar * pi) / pm;
      for (int j = k; j < N; j++) {
        const double tr = fr * mr[k * N + j] - fi * mi[k * N + j];
        const double ti = fr * mi[k * N + j] + fi * mr[k * N + j];
        mr[i * N + j] -= tr;
        mi[i * N + j] -= ti;
      }
      const double tr = fr * br[k] - fi * bi[k], ti = fr * bi[k] + fi * br[k];
      br[i] -= tr;
      bi[i] -= ti;
    }
  }
  for (int i = N - 1; i >= 0; i--) {
    double sr = br[i], si = bi[i];
    for (int j = i + 1; j < N; j++) {
      sr -= mr[i * N + j] * br[j] - mi[i * N + j] * bi[j];
      si -= mr[i * N + j] * bi[j] + mi[i * N + j] * br[j];
    }
    const double dr = mr[i * N + i], di = mi[i * N + i], dm = dr * dr + di * di;
    if (dm < 1.0e-300) return false;
    br[i] = (sr * dr + si * di) / dm;
    bi[i] = (si * dr - sr * di) / dm;
  }
  return true;
}

__device__ inline void lane_x(const cx *p, const cx (*d)[8], const double *tr, const double *ti,
                              cx *x) {
  for (int k = 0; k < 8; k++) {
    double sr = p[k].re, si = p[k].im;
    for (int j = 0; j < 3; j++) {
      sr += tr[j] * d[j][k].re - ti[j] * d[j][k].im;
      si += tr[j] * d[j][k].im + ti[j] * d[j][k].re;
    }
    x[k] = C(sr, si);
  }
}

__global__ void __launch_bounds__(128) k_cicy_lanes(const double *__restrict__ nrm, int64_t n_lanes,
                                                    double *__restrict__ xout,
                                                    int32_t *__restrict__ flag) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n_lanes) return;
  const double *base = nrm + 70 * l;
  cx p[8], d[3][8];
  load_line<0>(base, p, d);
  double tr[3], ti[3];
  for (int j = 0; j < 3; j++) {
    tr[j] = base[64 + 2 * j];
    ti[j] = base[64 + 2 * j + 1];
  }
  bool conv = false;
  for (int it = 0; it < 64; it++) {
    cx x[8], f[3], J[9];
    lane_x(p, d, tr, ti, x);
    for (int i = 0; i < 3; i++) {
      cx g[8];
      f[i] = lane_value_grad(i, x, g);
      for (int j = 0; j < 3; j++) {
        double sr = 0.0, si = 0.0;
        for (int k = 0; k < 8; k++) {
          sr += g[k].re * d[j][k].re - g[k].im * d[j][k].im;
          si += g[k].re * d[j][k].im + g[k].im * d[j][k].re;
        }
        J[i * 3 + j] = C(sr, si);
      }
    }
    double res = 0.0;
    for (int i = 0; i < 3; i++) res += f[i].re * f[i].re + f[i].im * f[i].im;
    res = sqrt(res);
    if (res < 1.0e-8) {
      conv = true;
      break;
    }
    double mr[9], mi[9], br[3], bi[3];
    for (int i = 0; i < 9; i++) {
      mr[i] = J[i].re;
      mi[i] = J[i].im;
    }
    for (int i = 0; i < 3; i++) {
      br[i] = -f[i].re;
      bi[i] = -f[i].im;
    }
    if (!lane_solve3(mr, mi, br, bi)) break;
    double damp = 1.0, nr[3], ni[3], new_res = 0.0;
    for (int trial = 0; trial < 8; trial++) {
      for (int j = 0; j < 3; j++) {
        nr[j] = tr[j] + damp * br[j];
        ni[j] = ti[j] + damp * bi[j];
      }
      cx x2[8];
      lane_x(p, d, nr, ni, x2);
      new_res = 0.0;
      for (int i = 0; i < 3; i++) {
        const cx fv = lane_value(i, x2);
        new_res += fv.re * fv.re + fv.im * fv.im;
      }
      new_res = sqrt(new_res);
      if (new_res < res) break;
      damp *= 0.5;
    }
    if (new_res >= res) break;
    for (int j = 0; j < 3; j++) {
      tr[j] = nr[j];
      ti[j] = ni[j];
    }
    if (new_res < 1.0e-8) {
      conv = true;
      break;
    }
  }
  flag[l] = conv ? 1 : 0;
  if (!conv) return;
  for (int k = 0; k < 8; k++) {  // x = origin + sum_j t_j d_j (cicy_sampler.rs:718-731)
    cx s = p[k];
    for (int j = 0; j < 3; j++) s = cadd(s, cmul(C(tr[j], ti[j]), d[j][k]));
    xout[16 * l + 2 * k] = s.re;
    xout[16 * l + 2 * k + 1] = s.im;
  }
}

// ---- Post-processing and frames -------------------------------------------------------------

__device__ inline int argmax_abs(const cx *c, int s, int e) {
  int best = s;
  double ba = -1.0;
  for (int k = s; k < e; k++) {
    const double a = cnorm(c[k]);
    if (a > ba) {
      ba = a;
      best = k;
    }
  }
  return ba > 0.0 ? best : -1;
}

template <int NH>
__device__ inline bool pick_elim(const cx *jac, bool *taken, int *elim) {
  for (int i = 0; i < NH; i++) {
    int bk = -1;
    double ba = -1.0;
    for (int k = 0; k < 8; k++) {
      if (taken[k]) continue;
      const double a = cnorm(jac[i * 8 + k]);
      if (a > ba) {
        ba = a;
        bk = k;
      }
    }
    if (bk < 0 || ba <= 0.0) return false;
    elim[i] = bk;
    taken[bk] = true;
  }
  return true;
}

// post_process_solved_point (cicy_sampler.rs:457-589; schoen_sampler.rs:744-881), then the
// metric point's patch indices (ty_metric.rs:479-496; schoen_metric.rs:1132-1151).
template <int V>
__global__ void __launch_bounds__(128) k_cicy_post(const double *__restrict__ xin,
                                                   const int32_t *__restrict__ flag, int64_t n,
                                                   double *__restrict__ coords,
                                                   double *__restrict__ omega_sq,
                                                   double *__restrict__ weight,
                                                   int32_t *__restrict__ patch,
                                                   int32_t *__restrict__ ok) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n) return;
  constexpr int NH = Var<V>::nh, NF = Var<V>::nfac;
  ok[l] = 0;
  if (!flag[l]) return;
  cx x[8], xr[8];
  for (int k = 0; k < 8; k++) x[k] = C(xin[16 * l + 2 * k], xin[16 * l + 2 * k + 1]);
  int idx[NF];
  cx scale[NF];
  for (int f = 0; f < NF; f++) {
    idx[f] = argmax_abs(x, Var<V>::rs(f), Var<V>::re(f));
    if (idx[f] < 0) return;
    scale[f] = x[idx[f]];
  }
  for (int f = 0; f < NF; f++)
    if (cnorm(scale[f]) < 2.220446049250313e-16) return;
  for (int f = 0; f < NF; f++)
    for (int k = Var<V>::rs(f); k < Var<V>::re(f); k++) xr[k] = cdiv(x[k], scale[f]);
  int near_one = 0;
  for (int k = 0; k < 8; k++)
    if (cnorm(csub(xr[k], C(1.0, 0.0))) < 1.0e-8) near_one++;
  if (near_one > NF) return;
  cx jac[NH * 8], om[NH * NH], ominv[NH * NH];
  jacobian<V>(xr, jac);
  bool taken[8] = {false, false, false, false, false, false, false, false};
  int elim[NH];
  for (int f = 0; f < NF; f++) taken[idx[f]] = true;
  if (!pick_elim<NH>(jac, taken, elim)) return;
  for (int i = 0; i < NH; i++)
    for (int k = 0; k < NH; k++) om[i * NH + k] = jac[i * 8 + elim[k]];
  const cx det_om = det_lu<NH>(om);
  if (cnorm(det_om) < 2.220446049250313e-16) return;
  const cx omega = cdiv(C(1.0, 0.0), det_om);
  int freec[3], nf = 0;
  for (int k = 0; k < 8; k++)
    if (!taken[k]) {
      if (nf >= 3) return;
      freec[nf++] = k;
    }
  if (nf != 3) return;
  if (!inv_n<NH>(om, ominv)) return;
  // j_pb rows: 1 at the free coordinate, -(Omega^-1 J[:, free])_r at elimination coordinate r.
  cx jpb[3][8];
  for (int a = 0; a < 3; a++) {
    for (int k = 0; k < 8; k++) jpb[a][k] = C(0.0, 0.0);
    jpb[a][freec[a]] = C(1.0, 0.0);
    cx b[NH];
    for (int i = 0; i < NH; i++) b[i] = jac[i * 8 + freec[a]];
    for (int r = 0; r < NH; r++) {
      cx e = C(0.0, 0.0);
      for (int c = 0; c < NH; c++) e = cadd(e, cmul(ominv[r * NH + c], b[c]));
      jpb[a][elim[r]] = cneg(e);
    }
  }
  // g_pb = j_pb g_FS j_pb^H with the block-diagonal Fubini-Study metric; the sums run over all
  // 8 coordinates in order, as the reference's, so structurally zero terms add exact zeros.
  cx tmp[3][8];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 8; b++) tmp[a][b] = C(0.0, 0.0);
  for (int f = 0; f < NF; f++) {
    const int s = Var<V>::rs(f), e = Var<V>::re(f), pt = idx[f];
    double nz = 0.0;
    for (int k = s; k < e; k++)
      if (k != pt) nz += cnsq(xr[k]);
    const double den = (1.0 + nz) * (1.0 + nz);
    for (int bcol = s; bcol < e; bcol++) {
      if (bcol == pt) continue;
      // column bcol of g_FS is non-zero only in rows of the same factor
      for (int a = 0; a < 3; a++) {
        cx acc = C(0.0, 0.0);
        for (int k = 0; k < 8; k++) {
          cx g = C(0.0, 0.0);
          if (k >= s && k < e && k != pt) {
            const double kron = k == bcol ? 1.0 : 0.0;
            const cx num = csub(cmul(C(1.0 + nz, 0.0), C(kron, 0.0)), cmul(xr[k], cconj(xr[bcol])));
            g = cdiv(num, C(den, 0.0));
          }
          acc = cadd(acc, cmul(jpb[a][k], g));
        }
        tmp[a][bcol] = acc;
      }
    }
  }
  cx gpb[9];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      cx s = C(0.0, 0.0);
      for (int k = 0; k < 8; k++) s = cadd(s, cmul(tmp[a][k], cconj(jpb[b][k])));
      gpb[a * 3 + b] = s;
    }
  const double det_g = det_lu<3>(gpb).re;
  if (!isfinite(det_g) || fabs(det_g) < 1.0e-300) return;
  const double w = cnsq(omega) / det_g;
  if (!isfinite(w) || w <= 0.0) return;
  for (int k = 0; k < 8; k++) {
    coords[16 * l + 2 * k] = xr[k].re;
    coords[16 * l + 2 * k + 1] = xr[k].im;
  }
  omega_sq[l] = cnsq(omega);
  weight[l] = w;
  if (V == 0) {
    const int zi = argmax_abs(xr, 0, 4), wi = argmax_abs(xr, 4, 8);
    patch[3 * l] = zi < 0 ? 0 : zi;
    patch[3 * l + 1] = wi < 0 ? 0 : wi - 4;
    patch[3 * l + 2] = -1;
  } else {
    const int xi = argmax_abs(xr, 0, 3), yi = argmax_abs(xr, 3, 6), ti = argmax_abs(xr, 6, 8);
    patch[3 * l] = xi < 0 ? 0 : xi;
    patch[3 * l + 1] = yi < 0 ? 3 : yi;
    patch[3 * l + 2] = ti < 0 ? 6 : ti;
  }
  ok[l] = 1;
}

// ty_affine_chart_frame / schoen_affine_chart_frame: frames [n][3][8][re, im], NaN when the
// chart is degenerate.
template <int V>
__global__ void __launch_bounds__(128) k_cicy_frames(const double *__restrict__ coords,
                                                     const int32_t *__restrict__ patch, int64_t n,
                                                     double *__restrict__ frames) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  constexpr int NH = Var<V>::nh;
  cx pt[8];
  for (int k = 0; k < 8; k++) pt[k] = C(coords[16 * p + 2 * k], coords[16 * p + 2 * k + 1]);
  bool taken[8] = {false, false, false, false, false, false, false, false};
  if (V == 0) {
    taken[patch[3 * p] & 3] = true;
    taken[4 + (patch[3 * p + 1] & 3)] = true;
  } else {
    const int a = patch[3 * p], b = patch[3 * p + 1], c = patch[3 * p + 2];
    taken[min(max(a, 0), 2)] = true;
    taken[min(max(b, 3), 5)] = true;
    taken[min(max(c, 6), 7)] = true;
  }
  cx jac[NH * 8], je[NH * NH], jinv[NH * NH];
  jacobian<V>(pt, jac);
  int elim[NH], freec[3], nf = 0;
  bool ok = pick_elim<NH>(jac, taken, elim);
  if (ok) {
    for (int k = 0; k < 8; k++)
      if (!taken[k]) {
        if (nf < 3) freec[nf] = k;
        nf++;
      }
    ok = nf == 3;
  }
  if (ok) {
    for (int i = 0; i < NH; i++)
      for (int c = 0; c < NH; c++) je[i * NH + c] = jac[i * 8 + elim[c]];
    ok = inv_n<NH>(je, jinv);
  }
  double *out = frames + 48 * p;
  if (!ok) {
    for (int i = 0; i < 48; i++) out[i] = __builtin_nan("");
    return;
  }
  for (int a = 0; a < 3; a++) {
    cx fr[8];
    for (int k = 0; k < 8; k++) fr[k] = C(0.0, 0.0);
    fr[freec[a]] = C(1.0, 0.0);
    cx b[NH];
    for (int i = 0; i < NH; i++) b[i] = jac[i * 8 + freec[a]];
    for (int c = 0; c < NH; c++) {
      cx e = C(0.0, 0.0);
      for (int i = 0; i < NH; i++) e = cadd(e, cmul(jinv[c * NH + i], b[i]));
      fr[elim[c]] = cneg(e);
    }
    for (int k = 0; k < 8; k++) {
      out[(a * 8 + k) * 2] = fr[k].re;
      out[(a * 8 + k) * 2 + 1] = fr[k].im;
    }
  }
}

// Monomial values and their 8 derivatives, one thread per (point, basis function). The powers
// follow the reference's table recurrence pow[k][e] = pow[k][e-1] z_k, so each value is the
// reference's product bit for bit. sv [n][2 nb], sd [n][8][2 nb].
__global__ void __launch_bounds__(256) k_cicy_sections(const double *__restrict__ coords, int64_t n,
                                                       const uint32_t *__restrict__ exps, int nb,
                                                       double *__restrict__ sv,
                                                       double *__restrict__ sd) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n * nb) return;
  const int64_t p = gid / nb;
  const int a = (int)(gid - p * nb);
  uint32_t m[8];
  for (int k = 0; k < 8; k++) m[k] = exps[8 * a + k];
  cx pw[8], pwm1[8];  // z_k^m_k and z_k^(m_k - 1)
  for (int k = 0; k < 8; k++) {
    const cx z = C(coords[16 * p + 2 * k], coords[16 * p + 2 * k + 1]);
    cx prev = C(1.0, 0.0), cur = C(1.0, 0.0);
    for (uint32_t e = 1; e <= m[k]; e++) {
      prev = cur;
      cur = cmul(cur, z);
    }
    pw[k] = cur;
    pwm1[k] = prev;
  }
  cx acc = C(1.0, 0.0);
  for (int k = 0; k < 8; k++) acc = cmul(acc, pw[k]);
  sv[p * 2 * nb + 2 * a] = acc.re;
  sv[p * 2 * nb + 2 * a + 1] = acc.im;
  for (int k = 0; k < 8; k++) {
    cx d = C(0.0, 0.0);
    if (m[k] != 0) {
      d = cmul(C((double)m[k], 0.0), pwm1[k]);
      for (int j = 0; j < 8; j++)
        if (j != k) d = cmul(d, pw[j]);
    }
    double *row = sd + (p * 8 + k) * 2 * nb;
    row[2 * a] = d.re;
    row[2 * a + 1] = d.im;
  }
}

inline unsigned grid(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

void launch_cicy_span(hipStream_t st, int variety, const double *nrm, int64_t s0, int64_t n_cand,
                      uint8_t *info) {
  if (n_cand <= 0) return;
  if (variety == 0)
    hipLaunchKernelGGL(k_cicy_span<0>, dim3(grid(n_cand, 128)), dim3(128), 0, st, nrm, s0, n_cand, info);
  else
    hipLaunchKernelGGL(k_cicy_span<1>, dim3(grid(n_cand, 128)), dim3(128), 0, st, nrm, s0, n_cand, info);
}

void launch_cicy_chain(hipStream_t st, int variety, const double *nrm, const int64_t *start,
                       const uint8_t *info, int64_t n_lines, double *x, int32_t *flag) {
  if (n_lines <= 0) return;
  if (variety == 0)
    hipLaunchKernelGGL(k_cicy_chain<0>, dim3(grid(n_lines, 128)), dim3(128), 0, st, nrm, start, info,
                       n_lines, x, flag);
  else
    hipLaunchKernelGGL(k_cicy_chain<1>, dim3(grid(n_lines, 128)), dim3(128), 0, st, nrm, start, info,
                       n_lines, x, flag);
}

void launch_cicy_lanes(hipStream_t st, const double *nrm, int64_t n_lanes, double *x, int32_t *flag) {
  if (n_lanes <= 0) return;
  hipLaunchKernelGGL(k_cicy_lanes, dim3(grid(n_lanes, 128)), dim3(128), 0, st, nrm, n_lanes, x, flag);
}

void launch_cicy_post(hipStream_t st, int variety, const double *x, const int32_t *flag, int64_t n,
                      double *coords, double *omega_sq, double *weight, int32_t *patch, int32_t *ok) {
  if (n <= 0) return;
  if (variety == 0)
    hipLaunchKernelGGL(k_cicy_post<0>, dim3(grid(n, 128)), dim3(128), 0, st, x, flag, n, coords,
                       omega_sq, weight, patch, ok);
  else
    hipLaunchKernelGGL(k_cicy_post<1>, dim3(grid(n, 128)), dim3(128), 0, st, x, flag, n, coords,
                       omega_sq, weight, patch, ok);
}

void launch_cicy_frames(hipStream_t st, int variety, const double *coords, const int32_t *patch,
                        int64_t n, double *frames) {
  if (n <= 0) return;
  if (variety == 0)
    hipLaunchKernelGGL(k_cicy_frames<0>, dim3(grid(n, 128)), dim3(128), 0, st, coords, patch, n, frames);
  else
    hipLaunchKernelGGL(k_cicy_frames<1>, dim3(grid(n, 128)), dim3(128), 0, st, coords, patch, n, frames);
}

void launch_cicy_sections(hipStream_t st, const double *coords, int64_t n, const uint32_t *exps,
                          int nb, double *sv, double *sd) {
  if (n <= 0 || nb <= 0) return;
  hipLaunchKernelGGL(k_cicy_sections, dim3(grid(n * nb, 256)), dim3(256), 0, st, coords, n, exps, nb,
                     sv, sd);
}


// ---- The normal stream ---------------------------------------------------------------------
// ChaCha8 (rand_chacha 0.9): key words 4-11, a 64-bit block counter in words 12-13, stream 0 in
// words 14-15, 8 rounds; u64 number q of the stream is words 2q and 2q + 1 of the keystream
// (low word first), i.e. words 2 (q % 8), 2 (q % 8) + 1 of block q / 8.
__constant__ double c_zig_x[257];

void upload_zig_x(hipStream_t st, const double *x257) {
  hipMemcpyToSymbolAsync(HIP_SYMBOL(c_zig_x), x257, sizeof(double) * 257, 0,
                         hipMemcpyHostToDevice, st);
}

__device__ inline uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

__device__ inline void chacha8_block(const ZigKey &key, uint64_t counter, uint32_t (&out)[16]) {
  uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
#pragma unroll
  for (int i = 0; i < 8; i++) s[4 + i] = key.k[i];
  s[12] = (uint32_t)counter;
  s[13] = (uint32_t)(counter >> 32);
  s[14] = s[15] = 0;
  uint32_t x[16];
#pragma unroll
  for (int i = 0; i < 16; i++) x[i] = s[i];
#define ZQR(a, b, c, d)                  \
  x[a] += x[b]; x[d] = rotl32(x[d] ^ x[a], 16); \
  x[c] += x[d]; x[b] = rotl32(x[b] ^ x[c], 12); \
  x[a] += x[b]; x[d] = rotl32(x[d] ^ x[a], 8);  \
  x[c] += x[d]; x[b] = rotl32(x[b] ^ x[c], 7);
#pragma unroll
  for (int r = 0; r < 4; r++) {
    ZQR(0, 4, 8, 12) ZQR(1, 5, 9, 13) ZQR(2, 6, 10, 14) ZQR(3, 7, 11, 15)
    ZQR(0, 5, 10, 15) ZQR(1, 6, 11, 12) ZQR(2, 7, 8, 13) ZQR(3, 4, 9, 14)
  }
#undef ZQR
#pragma unroll
  for (int i = 0; i < 16; i++) out[i] = x[i] + s[i];
}

// One thread per keystream block: positions q in [pos0, pos0 + n_pos] of that block; words[q -
// pos0] is the u64 there (one past the range too, for k_zig_exc_words), and for q < pos0 + n_pos
// fast[q - pos0] = u x_i with u = 2 (bits >> 12) / 2^52 - 1 in [-1, 1) and i = bits & 0xff (the
// value if |u x_i| < x_(i+1)); the other positions are appended to exc (in no particular order;
// n_exc counts them all, exc holds the first exc_cap).
__global__ void __launch_bounds__(256) k_zig_fast(ZigKey key, uint64_t pos0, int64_t n_pos,
                                                  double *__restrict__ fast,
                                                  uint64_t *__restrict__ words,
                                                  int64_t *__restrict__ exc, int64_t exc_cap,
                                                  unsigned long long *__restrict__ n_exc) {
  const uint64_t b0 = pos0 / 8;
  const uint64_t blk = b0 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if ((int64_t)(blk * 8) > (int64_t)(pos0 + n_pos)) return;
  uint32_t w[16];
  chacha8_block(key, blk, w);
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const uint64_t q = blk * 8 + j;
    if (q < pos0 || q > pos0 + (uint64_t)n_pos) continue;
    const uint64_t bits = (uint64_t)w[2 * j] | ((uint64_t)w[2 * j + 1] << 32);
    words[q - pos0] = bits;
    if (q == pos0 + (uint64_t)n_pos) continue;  // the word after the range, for exc_words
    const int i = (int)(bits & 0xff);
    const double u = __longlong_as_double((long long)((bits >> 12) | (1024ull << 52))) - 3.0;
    const double x = u * c_zig_x[i];
    fast[q - pos0] = x;
    if (!(fabs(x) < c_zig_x[i + 1])) {
      const unsigned long long slot = atomicAdd(n_exc, 1ull);
      if (slot < (unsigned long long)exc_cap) exc[slot] = (int64_t)q;  // else counted only
    }
  }
}

// out[k] for k < count: in run r (run_k[r] <= k < run_k[r] + run_len[r]) the fast value at
// position run_p[r] + (k - run_k[r]) (relative to the fast array), else the host's value
// exc_v[e] for exc_k[e] == k. Runs and exc_k are sorted by k.
__global__ void __launch_bounds__(256) k_zig_gather(const double *__restrict__ fast,
                                                    const int64_t *__restrict__ run_k,
                                                    const int64_t *__restrict__ run_p,
                                                    const int64_t *__restrict__ run_len,
                                                    int64_t n_runs,
                                                    const int64_t *__restrict__ exc_k,
                                                    const double *__restrict__ exc_v,
                                                    int64_t n_exc, int64_t count,
                                                    double *__restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  int64_t lo = 0, hi = n_runs;  // last run with run_k <= k
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) / 2;
    if (run_k[mid] <= k) lo = mid; else hi = mid;
  }
  if (n_runs > 0 && run_k[lo] <= k && k < run_k[lo] + run_len[lo]) {
    out[k] = fast[run_p[lo] + (k - run_k[lo])];
    return;
  }
  lo = 0;
  hi = n_exc;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) / 2;
    if (exc_k[mid] <= k) lo = mid; else hi = mid;
  }
  out[k] = (n_exc > 0 && exc_k[lo] == k) ? exc_v[lo] : __builtin_nan("");
}

// The two words from each listed position on (the wedge's test reads both): out[2 e] and
// out[2 e + 1] for exc[e] (words holds positions [pos0, pos0 + n_pos]).
__global__ void __launch_bounds__(256) k_zig_exc_words(const uint64_t *__restrict__ words,
                                                       const int64_t *__restrict__ exc, int64_t n,
                                                       uint64_t pos0, uint64_t *__restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int64_t r = exc[e] - (int64_t)pos0;
  out[2 * e] = words[r];
  out[2 * e + 1] = words[r + 1];
}

void launch_zig_fast(hipStream_t st, const ZigKey &key, uint64_t pos0, int64_t n_pos, double *fast,
                     uint64_t *words, int64_t *exc, int64_t exc_cap,
                     unsigned long long *n_exc) {
  if (n_pos <= 0) return;
  const int64_t nblk = (int64_t)((pos0 + n_pos + 8) / 8 - pos0 / 8);  // positions through pos0 + n_pos
  hipLaunchKernelGGL(k_zig_fast, dim3(grid(nblk, 256)), dim3(256), 0, st, key, pos0, n_pos, fast,
                     words, exc, exc_cap, n_exc);
}

void launch_zig_exc_words(hipStream_t st, const uint64_t *words, const int64_t *exc, int64_t n,
                          uint64_t pos0, uint64_t *out) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zig_exc_words, dim3(grid(n, 256)), dim3(256), 0, st, words, exc, n, pos0,
                     out);
}

void launch_zig_gather(hipStream_t st, const double *fast, const int64_t *run_k,
                       const int64_t *run_p, const int64_t *run_len, int64_t n_runs,
                       const int64_t *exc_k, const double *exc_v, int64_t n_exc, int64_t count,
                       double *out) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_zig_gather, dim3(grid(count, 256)), dim3(256), 0, st, fast, run_k, run_p,
                     run_len, n_runs, exc_k, exc_v, n_exc, count, out);
}

}  // namespace cy3q
