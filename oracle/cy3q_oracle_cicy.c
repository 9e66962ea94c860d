/*
 * cy3q oracle, part 2: the 8-coordinate (Tian-Yau and Schoen) sample, section and frame set-up
 * of the reference's route34 callers, restated on the CPU in FP64.
 *
 * TEST INFRASTRUCTURE ONLY. Like cy3q_oracle.c this file is the checker: only tests/, the smoke
 * entry point and the cpu_baseline leg of bench.py load it (through oracle/oracle.py). The
 * product library never links, loads or calls it.
 *
 * Citations are relative to substrate_sm_derivation/cy3_solver/ in the reference. Build with
 * -ffp-contract=off: Rust never fuses a*b+c.
 *
 * Third-party code the reference calls here, restated from its published algorithm:
 *   - rand_chacha 0.9.0 (Cargo.lock:480-483) ChaCha8Rng: the ChaCha stream cipher (Bernstein
 *     2008) with 8 rounds, 64-bit block counter in words 12-13 and a zero 64-bit stream id in
 *     words 14-15, keystream words consumed in order, u64 = lo | hi << 32;
 *   - rand_core 0.9.5 (Cargo.lock:490-493) SeedableRng::seed_from_u64: the 32-byte key is eight
 *     PCG32 outputs (LCG multiplier 6364136223846793005, increment 11634580027462260723, state
 *     advanced before each output, XSH-RR output) in little-endian order;
 *   - rand_distr 0.5.1 (Cargo.lock:499-502) StandardNormal: the 256-layer ziggurat of
 *     Marsaglia & Tsang (2000) with Doornik's sign-from-the-uniform variant: i = low 8 bits of
 *     a u64, u in [-1, 1) from its high 52 bits, x = u X[i]; accept |x| < X[i+1]; layer 0 goes to
 *     the tail (x = ln(U1)/R, y = ln(U2) until -2y >= x^2, result +-(R - x)); otherwise accept
 *     when F[i+1] + (F[i] - F[i+1]) U < exp(-x^2/2). The tables follow the recursion
 *     X[0] = V/f(R), X[1] = R, X[i] = f^-1(V/X[i-1] + f(X[i-1])), X[256] = 0, F = f(X), with
 *     R = 3.654152885361008796 and V = 0.00492867323399, each value rounded to 18 decimals as the
 *     table file stores it.
 * Pins (tests/test_oracle_cicy.py): the ChaCha20 and ChaCha8 all-zero-key keystream vectors,
 * the normal stream's moments, and end to end the reference's determinism golden file
 * (tests/data/p5_10_determinism_golden.json: sigma at h = I and after 5 Donaldson iterations
 * for TY k = 3 and Schoen (3,3,1) on 200 points).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------------------------ */
/* num-complex 0.4 arithmetic, operation for operation (naive division, hypot modulus).        */

typedef struct {
  double re, im;
} cx;

static inline cx C(double re, double im) { cx r = {re, im}; return r; }
static inline cx cadd(cx a, cx b) { return C(a.re + b.re, a.im + b.im); }
static inline cx csub(cx a, cx b) { return C(a.re - b.re, a.im - b.im); }
static inline cx cmul(cx a, cx b) { return C(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline cx cdiv(cx a, cx b) {
  double n = b.re * b.re + b.im * b.im;
  double re = a.re * b.re + a.im * b.im;
  double im = a.im * b.re - a.re * b.im;
  return C(re / n, im / n);
}
static inline cx cdivr(cx a, double r) { return C(a.re / r, a.im / r); }
static inline cx cneg(cx a) { return C(-a.re, -a.im); }
static inline cx cconj(cx a) { return C(a.re, -a.im); }
static inline double cnorm(cx a) { return hypot(a.re, a.im); }
static inline double cnsq(cx a) { return a.re * a.re + a.im * a.im; }

/* ------------------------------------------------------------------------------------------ */
/* ChaCha8Rng::seed_from_u64 and next_u64                                                       */

typedef struct {
  uint32_t key[8];
  uint64_t counter; /* next block */
  uint32_t buf[64]; /* four blocks, as rand_chacha refills them */
  int idx;          /* next word */
  int rounds;
  int64_t normals;  /* standard normals drawn so far */
} orc_chacha;

static inline uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

#define QR(a, b, c, d)                  \
  do {                                  \
    a += b; d ^= a; d = rotl32(d, 16);  \
    c += d; b ^= c; b = rotl32(b, 12);  \
    a += b; d ^= a; d = rotl32(d, 8);   \
    c += d; b ^= c; b = rotl32(b, 7);   \
  } while (0)

static void chacha_block(const uint32_t key[8], uint64_t counter, uint64_t stream, int rounds,
                         uint32_t out[16]) {
  uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
  for (int i = 0; i < 8; i++) s[4 + i] = key[i];
  s[12] = (uint32_t)counter;
  s[13] = (uint32_t)(counter >> 32);
  s[14] = (uint32_t)stream;
  s[15] = (uint32_t)(stream >> 32);
  uint32_t x[16];
  memcpy(x, s, sizeof x);
  for (int r = 0; r < rounds; r += 2) {
    QR(x[0], x[4], x[8], x[12]);
    QR(x[1], x[5], x[9], x[13]);
    QR(x[2], x[6], x[10], x[14]);
    QR(x[3], x[7], x[11], x[15]);
    QR(x[0], x[5], x[10], x[15]);
    QR(x[1], x[6], x[11], x[12]);
    QR(x[2], x[7], x[8], x[13]);
    QR(x[3], x[4], x[9], x[14]);
  }
  for (int i = 0; i < 16; i++) out[i] = x[i] + s[i];
}

static void chacha_seed_u64(orc_chacha *g, uint64_t state, int rounds) {
  const uint64_t MUL = 6364136223846793005ull, INC = 11634580027462260723ull;
  for (int i = 0; i < 8; i++) {
    state = state * MUL + INC;
    uint32_t xorshifted = (uint32_t)(((state >> 18) ^ state) >> 27);
    uint32_t rot = (uint32_t)(state >> 59);
    g->key[i] = (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
  }
  g->counter = 0;
  g->idx = 64;
  g->rounds = rounds;
  g->normals = 0;
}

static uint64_t chacha_next_u64(orc_chacha *g) {
  if (g->idx >= 64) {
    for (int b = 0; b < 4; b++) chacha_block(g->key, g->counter + b, 0, g->rounds, g->buf + 16 * b);
    g->counter += 4;
    g->idx = 0;
  }
  uint64_t lo = g->buf[g->idx], hi = g->buf[g->idx + 1];
  g->idx += 2;
  return lo | (hi << 32);
}

/* ------------------------------------------------------------------------------------------ */
/* rand_distr StandardNormal (ziggurat)                                                         */

static double ZX[257], ZF[257];
static int zig_ready = 0;
static const double ZIG_R = 3.654152885361008796;

static double round18(double v) {
  char b[64];
  snprintf(b, sizeof b, "%.18f", v);
  return strtod(b, NULL);
}

static void zig_init(void) {
  if (zig_ready) return;
  const double r = 3.654152885361008796, v = 0.00492867323399;
  double x[257];
  x[0] = v / exp(-r * r / 2.0);
  x[1] = r;
  for (int i = 2; i < 256; i++) {
    double last = x[i - 1];
    x[i] = sqrt(-2.0 * log(v / last + exp(-last * last / 2.0)));
  }
  x[256] = 0.0;
  for (int i = 0; i < 257; i++) {
    ZX[i] = round18(x[i]);
    ZF[i] = round18(exp(-x[i] * x[i] / 2.0));
  }
  zig_ready = 1;
}

static inline double u64_float_exp(uint64_t frac52, int e) {
  uint64_t bits = frac52 | ((uint64_t)(1023 + e) << 52);
  double d;
  memcpy(&d, &bits, 8);
  return d;
}

static double std_normal(orc_chacha *g) {
  g->normals++;
  for (;;) {
    uint64_t bits = chacha_next_u64(g);
    int i = (int)(bits & 0xff);
    double u = u64_float_exp(bits >> 12, 1) - 3.0;
    double x = u * ZX[i];
    if (fabs(x) < ZX[i + 1]) return x;
    if (i == 0) {
      double xt = 1.0, yt = 0.0;
      while (-2.0 * yt < xt * xt) {
        double x_ = u64_float_exp(chacha_next_u64(g) >> 12, 0) - (1.0 - 2.220446049250313e-16 / 2.0);
        double y_ = u64_float_exp(chacha_next_u64(g) >> 12, 0) - (1.0 - 2.220446049250313e-16 / 2.0);
        xt = log(x_) / ZIG_R;
        yt = log(y_);
      }
      return u < 0.0 ? xt - ZIG_R : ZIG_R - xt;
    }
    double un = (double)(chacha_next_u64(g) >> 11) * (1.0 / 9007199254740992.0);
    if (ZF[i + 1] + (ZF[i] - ZF[i + 1]) * un < exp(-x * x / 2.0)) return x;
  }
}

/* ------------------------------------------------------------------------------------------ */
/* The two varieties: sparse polynomials in 8 ambient coordinates                             */
/* TY: BicubicPair::z3_invariant_default (cicy_sampler.rs:244-275), coords z0..z3 w0..w3.       */
/* Schoen: SchoenPoly::z3xz3_invariant_default (route34/schoen_sampler.rs:188-217), coords      */
/* x0..x2 y0..y2 t0 t1.                                                                         */

typedef struct {
  int nterm;
  unsigned char e[6][8];
} spoly;

typedef struct {
  int nh;          /* hypersurfaces */
  int nfac;        /* projective factors */
  int rs[3], re[3]; /* factor coordinate ranges */
  int ty;          /* 1: TY (det pre-check in Newton, z/w patch logic) */
  spoly f[3];
} variety;

static void make_variety(int which, variety *v) {
  memset(v, 0, sizeof *v);
  if (which == 0) {
    v->nh = 3; v->nfac = 2; v->ty = 1;
    v->rs[0] = 0; v->re[0] = 4; v->rs[1] = 4; v->re[1] = 8;
    for (int i = 0; i < 4; i++) {
      v->f[0].e[i][i] = 3;
      v->f[1].e[i][4 + i] = 3;
      v->f[2].e[i][i] = 1;
      v->f[2].e[i][4 + i] = 1;
    }
    v->f[0].nterm = v->f[1].nterm = v->f[2].nterm = 4;
  } else {
    v->nh = 2; v->nfac = 3; v->ty = 0;
    v->rs[0] = 0; v->re[0] = 3; v->rs[1] = 3; v->re[1] = 6; v->rs[2] = 6; v->re[2] = 8;
    for (int k = 0; k < 3; k++) {
      v->f[0].e[k][k] = 3; v->f[0].e[k][6] = 1;
      v->f[1].e[k][3 + k] = 3; v->f[1].e[k][7] = 1;
    }
    static const unsigned char a[8] = {1, 1, 1, 0, 0, 0, 0, 1}, b[8] = {0, 0, 0, 1, 1, 1, 1, 0};
    memcpy(v->f[0].e[3], a, 8);
    memcpy(v->f[1].e[3], b, 8);
    v->f[0].nterm = v->f[1].nterm = 4;
  }
}

static cx eval_monomial(const cx *pt, const unsigned char *e) { /* cicy_sampler.rs:300-317 */
  cx acc = C(1.0, 0.0);
  for (int k = 0; k < 8; k++) {
    if (e[k] == 0) continue;
    cx p = pt[k];
    for (int j = 1; j < e[k]; j++) p = cmul(p, pt[k]);
    acc = cmul(acc, p);
  }
  return acc;
}

static cx eval_poly(const spoly *f, const cx *pt) { /* :319-326, every coefficient 1 */
  cx s = C(0.0, 0.0);
  for (int t = 0; t < f->nterm; t++) s = cadd(s, cmul(C(1.0, 0.0), eval_monomial(pt, f->e[t])));
  return s;
}

static cx poly_partial(const spoly *f, const cx *pt, int k) { /* :328-350 */
  cx s = C(0.0, 0.0);
  for (int t = 0; t < f->nterm; t++) {
    int ek = f->e[t][k];
    if (ek == 0) continue;
    unsigned char ne[8];
    memcpy(ne, f->e[t], 8);
    ne[k] = (unsigned char)(ek - 1);
    cx mono = eval_monomial(pt, ne);
    s = cadd(s, cmul(cmul(C(1.0, 0.0), C((double)ek, 0.0)), mono));
  }
  return s;
}

static void jacobian(const variety *v, const cx *pt, cx *jac) {
  for (int k = 0; k < 8; k++)
    for (int i = 0; i < v->nh; i++) jac[i * 8 + k] = poly_partial(&v->f[i], pt, k);
}

/* det_complex_lu (cicy_sampler.rs:959-998; schoen_sampler.rs:1000-1036) */
static cx det_lu(const cx *mat, int n) {
  cx a[9];
  memcpy(a, mat, sizeof(cx) * n * n);
  cx det = C(1.0, 0.0);
  for (int k = 0; k < n; k++) {
    int piv = k;
    double best = cnorm(a[k * n + k]);
    for (int p = k + 1; p < n; p++) {
      double v = cnorm(a[p * n + k]);
      if (v > best) { best = v; piv = p; }
    }
    if (best == 0.0) return C(0.0, 0.0);
    if (piv != k) {
      for (int j = 0; j < n; j++) { cx t = a[k * n + j]; a[k * n + j] = a[piv * n + j]; a[piv * n + j] = t; }
      det = cneg(det);
    }
    cx pv = a[k * n + k];
    det = cmul(det, pv);
    cx ip = cdiv(C(1.0, 0.0), pv);
    for (int i = k + 1; i < n; i++) {
      cx factor = cmul(a[i * n + k], ip);
      for (int j = k + 1; j < n; j++) a[i * n + j] = csub(a[i * n + j], cmul(factor, a[k * n + j]));
      a[i * n + k] = C(0.0, 0.0);
    }
  }
  return det;
}

/* inv_n / invert_n_complex: Gauss-Jordan with partial pivoting (cicy_sampler.rs:1000-1064,
 * route34/ty_metric.rs:2032-2083). Returns 0 when a pivot is below 1e-30. */
static int inv_n(const cx *mat, int n, cx *inv) {
  int cols = 2 * n;
  cx a[18];
  for (int i = 0; i < n * cols; i++) a[i] = C(0.0, 0.0);
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < n; j++) a[i * cols + j] = mat[i * n + j];
    a[i * cols + n + i] = C(1.0, 0.0);
  }
  for (int k = 0; k < n; k++) {
    int piv = k;
    double best = cnorm(a[k * cols + k]);
    for (int p = k + 1; p < n; p++) {
      double v = cnorm(a[p * cols + k]);
      if (v > best) { best = v; piv = p; }
    }
    if (best < 1.0e-30) return 0;
    if (piv != k)
      for (int j = 0; j < cols; j++) { cx t = a[k * cols + j]; a[k * cols + j] = a[piv * cols + j]; a[piv * cols + j] = t; }
    cx ip = cdiv(C(1.0, 0.0), a[k * cols + k]);
    for (int j = 0; j < cols; j++) a[k * cols + j] = cmul(a[k * cols + j], ip);
    for (int i = 0; i < n; i++) {
      if (i == k) continue;
      cx factor = a[i * cols + k];
      if (cnorm(factor) == 0.0) continue;
      for (int j = 0; j < cols; j++) a[i * cols + j] = csub(a[i * cols + j], cmul(a[k * cols + j], factor));
    }
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) inv[i * n + j] = a[i * cols + n + j];
  return 1;
}

static void gaussian_unit_sphere(orc_chacha *g, cx *v) { /* cicy_sampler.rs:883-901 */
  double sq = 0.0;
  for (int k = 0; k < 8; k++) {
    double re = std_normal(g);
    double im = std_normal(g);
    v[k] = C(re, im);
    sq += re * re + im * im;
  }
  double n = sqrt(sq);
  if (n > 0.0)
    for (int k = 0; k < 8; k++) v[k] = cdivr(v[k], n);
}

static void line_eval(const cx *p, const cx dirs[][8], const cx *t, int nh, cx *x) {
  for (int k = 0; k < 8; k++) x[k] = p[k];
  for (int j = 0; j < nh; j++)
    for (int k = 0; k < 8; k++) x[k] = cadd(x[k], cmul(t[j], dirs[j][k]));
}

static double residual_norm(const cx *f, int nh) {
  double s = 0.0;
  for (int i = 0; i < nh; i++) s += cnsq(f[i]);
  return sqrt(s);
}

/* newton_solve_line (cicy_sampler.rs:771-853; route34/schoen_sampler.rs:643-713). The TY form
 * bails on |det M| < 1e-30 before inverting; the Schoen form only on a failed inverse. */
static int newton_line(const variety *v, const cx *p, const cx dirs[][8], cx *t, cx *x_out,
                       double *res_out) {
  const int nh = v->nh;
  cx x[8], f[3];
  line_eval(p, dirs, t, nh, x);
  for (int i = 0; i < nh; i++) f[i] = eval_poly(&v->f[i], x);
  double res = residual_norm(f, nh);
  int conv = 0, done = 0;
  for (int it = 0; it < 64; it++) {
    if (res < 1.0e-8) { conv = 1; done = 1; break; }
    cx jac[24], m[9], minv[9], dt[3];
    jacobian(v, x, jac);
    for (int i = 0; i < nh; i++)
      for (int j = 0; j < nh; j++) {
        cx s = C(0.0, 0.0);
        for (int k = 0; k < 8; k++) s = cadd(s, cmul(jac[i * 8 + k], dirs[j][k]));
        m[i * nh + j] = s;
      }
    if (v->ty) {
      cx d = det_lu(m, nh);
      if (cnorm(d) < 1.0e-30) { done = 1; break; }
    }
    if (!inv_n(m, nh, minv)) { done = 1; break; }
    for (int r = 0; r < nh; r++) {
      cx s = C(0.0, 0.0);
      for (int c = 0; c < nh; c++) s = cadd(s, cmul(minv[r * nh + c], f[c]));
      dt[r] = cneg(s);
    }
    double step = 1.0;
    int accepted = 0;
    for (int ls = 0; ls < 32; ls++) {
      cx tt[3], xt[8], ft[3];
      for (int r = 0; r < nh; r++) tt[r] = cadd(t[r], cmul(C(step, 0.0), dt[r]));
      line_eval(p, dirs, tt, nh, xt);
      for (int i = 0; i < nh; i++) ft[i] = eval_poly(&v->f[i], xt);
      double rt = residual_norm(ft, nh);
      if (rt < (1.0 - 1.0e-4 * step) * res) {
        for (int r = 0; r < nh; r++) t[r] = tt[r];
        memcpy(x, xt, sizeof x);
        memcpy(f, ft, sizeof(cx) * nh);
        res = rt;
        accepted = 1;
        break;
      }
      step *= 0.5;
    }
    if (!accepted) { done = 1; break; }
  }
  if (!done) conv = res < 1.0e-8;
  memcpy(x_out, x, sizeof x);
  *res_out = res;
  return conv;
}

/* Greedy elimination columns (cicy_sampler.rs:924-957; schoen_sampler.rs:900-933;
 * ty_metric.rs:1948-1970). taken[] holds the patch coordinates on entry. */
static int pick_elim(const cx *jac, int nh, int *taken, int *elim) {
  for (int i = 0; i < nh; i++) {
    int bk = -1;
    double ba = -1.0;
    for (int k = 0; k < 8; k++) {
      if (taken[k]) continue;
      double a = cnorm(jac[i * 8 + k]);
      if (a > ba) { ba = a; bk = k; }
    }
    if (bk < 0 || ba <= 0.0) return 0;
    elim[i] = bk;
    taken[bk] = 1;
  }
  return 1;
}

/* argmax |c| over [s, e) with the first maximum winning; -1 if every entry is zero */
static int argmax_abs(const cx *c, int s, int e) {
  int best = s;
  double ba = -1.0;
  for (int k = s; k < e; k++) {
    double a = cnorm(c[k]);
    if (a > ba) { ba = a; best = k; }
  }
  return ba > 0.0 ? best : -1;
}

/* FS block in the patch where coordinate `patch` is 1 (cicy_sampler.rs:1066-1127,
 * schoen_sampler.rs:1110-1140) */
static void fs_block(const cx *x, int s, int e, int patch, cx *g) {
  double nz = 0.0;
  for (int k = s; k < e; k++)
    if (k != patch) nz += cnsq(x[k]);
  double den = (1.0 + nz) * (1.0 + nz);
  for (int a = s; a < e; a++) {
    if (a == patch) continue;
    for (int b = s; b < e; b++) {
      if (b == patch) continue;
      double kron = a == b ? 1.0 : 0.0;
      cx num = csub(cmul(C(1.0 + nz, 0.0), C(kron, 0.0)), cmul(x[a], cconj(x[b])));
      g[a * 8 + b] = cdiv(num, C(den, 0.0));
    }
  }
}

/* post_process_solved_point (cicy_sampler.rs:457-589; schoen_sampler.rs:744-881). On success
 * writes the rescaled coordinates, |Omega|^2 and the (unnormalised) weight. */
static int post_process(const variety *v, const cx *x, cx *xr, double *omega_sq, double *weight) {
  int idx[3];
  cx scale[3];
  for (int f = 0; f < v->nfac; f++) {
    idx[f] = argmax_abs(x, v->rs[f], v->re[f]);
    if (idx[f] < 0) return 0;
    scale[f] = x[idx[f]];
  }
  for (int f = 0; f < v->nfac; f++)
    if (cnorm(scale[f]) < 2.220446049250313e-16) return 0;
  for (int f = 0; f < v->nfac; f++)
    for (int k = v->rs[f]; k < v->re[f]; k++) xr[k] = cdiv(x[k], scale[f]);
  int near_one = 0;
  for (int k = 0; k < 8; k++)
    if (cnorm(csub(xr[k], C(1.0, 0.0))) < 1.0e-8) near_one++;
  if (near_one > v->nfac) return 0;
  const int nh = v->nh, nfree = 3;
  cx jac[24], om[9], ominv[9];
  jacobian(v, xr, jac);
  int taken[8] = {0}, elim[3];
  for (int f = 0; f < v->nfac; f++) taken[idx[f]] = 1;
  if (!pick_elim(jac, nh, taken, elim)) return 0;
  for (int i = 0; i < nh; i++)
    for (int k = 0; k < nh; k++) om[i * nh + k] = jac[i * 8 + elim[k]];
  cx det_om = det_lu(om, nh);
  if (cnorm(det_om) < 2.220446049250313e-16) return 0;
  cx omega = cdiv(C(1.0, 0.0), det_om);
  int freec[3], nf = 0;
  for (int k = 0; k < 8; k++)
    if (!taken[k]) {
      if (nf >= nfree) return 0;
      freec[nf++] = k;
    }
  if (nf != nfree) return 0;
  cx jpb[24];
  for (int i = 0; i < 24; i++) jpb[i] = C(0.0, 0.0);
  for (int a = 0; a < nfree; a++) jpb[a * 8 + freec[a]] = C(1.0, 0.0);
  if (!inv_n(om, nh, ominv)) return 0;
  for (int a = 0; a < nfree; a++) {
    cx b[3];
    for (int i = 0; i < nh; i++) b[i] = jac[i * 8 + freec[a]];
    for (int r = 0; r < nh; r++) {
      cx e = C(0.0, 0.0);
      for (int c = 0; c < nh; c++) e = cadd(e, cmul(ominv[r * nh + c], b[c]));
      jpb[a * 8 + elim[r]] = cneg(e);
    }
  }
  cx g[64], tmp[24], gpb[9];
  for (int i = 0; i < 64; i++) g[i] = C(0.0, 0.0);
  for (int f = 0; f < v->nfac; f++) fs_block(xr, v->rs[f], v->re[f], idx[f], g);
  for (int a = 0; a < nfree; a++)
    for (int b = 0; b < 8; b++) {
      cx s = C(0.0, 0.0);
      for (int k = 0; k < 8; k++) s = cadd(s, cmul(jpb[a * 8 + k], g[k * 8 + b]));
      tmp[a * 8 + b] = s;
    }
  for (int a = 0; a < nfree; a++)
    for (int b = 0; b < nfree; b++) {
      cx s = C(0.0, 0.0);
      for (int k = 0; k < 8; k++) s = cadd(s, cmul(tmp[a * 8 + k], cconj(jpb[b * 8 + k])));
      gpb[a * 3 + b] = s;
    }
  double det_g = det_lu(gpb, 3).re;
  if (!isfinite(det_g) || fabs(det_g) < 1.0e-300) return 0;
  double w = cnsq(omega) / det_g;
  if (!isfinite(w) || w <= 0.0) return 0;
  *omega_sq = cnsq(omega);
  *weight = w;
  return 1;
}

/* One line: the draws and Newton restarts of sample_one (cicy_sampler.rs:409-447;
 * schoen_sampler.rs:429-479). Returns 1 and x when a restart converged. *attempts gets the
 * restarts used. */
static int sample_line(const variety *v, orc_chacha *g, cx *x_best, int *attempts) {
  cx p[8], dirs[3][8];
  gaussian_unit_sphere(g, p);
  for (int j = 0; j < v->nh; j++) gaussian_unit_sphere(g, dirs[j]);
  int have = 0;
  double best = INFINITY;
  int a;
  for (a = 0; a < 4; a++) {
    cx t[3], x[8];
    for (int j = 0; j < v->nh; j++) {
      double re = std_normal(g);
      double im = std_normal(g);
      t[j] = C(re, im);
    }
    double res;
    int conv = newton_line(v, p, dirs, t, x, &res);
    if (conv && res < best) {
      best = res;
      memcpy(x_best, x, sizeof x);
      have = 1;
      if (res < 1.0e-8) { a++; break; }
    }
  }
  *attempts = a;
  return have && best < 1.0e-8;
}

/* CicySampler::sample_batch (cicy_sampler.rs:591-617) / SchoenSampler::sample_points
 * (schoen_sampler.rs:482-498): lines until n_target points are accepted or 64 n_target
 * (at least 1024) lines are spent; weights normalised to sum 1 (TY divides by the sum, Schoen
 * multiplies by its reciprocal). No quotient factor. Output per point: coords [16] (re, im of the
 * 8 rescaled coordinates), |Omega|^2, weight, and the patch indices the metric code recomputes
 * (TY: z_idx in 0..4, w_idx in 0..4; Schoen: x_idx, y_idx, t_idx as coordinate indices).
 * line_log (optional, 3 ints per line): index of the line's first standard normal in the stream,
 * restarts used, accepted. */
ORC_API int64_t orc_cicy_sample(int which, uint64_t seed, int64_t n_target, double *coords,
                                double *omega_sq, double *weight, int32_t *idx, int64_t *n_lines,
                                int32_t *line_log, int64_t line_log_cap) {
  zig_init();
  variety v;
  make_variety(which, &v);
  orc_chacha g;
  chacha_seed_u64(&g, seed, 8);
  int64_t max_total = n_target * 64 > 1024 ? n_target * 64 : 1024;
  int64_t n = 0, lines = 0;
  while (n < n_target && lines < max_total) {
    int64_t start_normal = g.normals;
    cx x[8], xr[8];
    int att = 0, ok = sample_line(&v, &g, x, &att);
    double om = 0, w = 0;
    if (ok) ok = post_process(&v, x, xr, &om, &w);
    if (line_log && lines < line_log_cap) {
      line_log[3 * lines] = (int32_t)start_normal;
      line_log[3 * lines + 1] = att;
      line_log[3 * lines + 2] = ok;
    }
    lines++;
    if (!ok) continue;
    for (int k = 0; k < 8; k++) {
      coords[n * 16 + 2 * k] = xr[k].re;
      coords[n * 16 + 2 * k + 1] = xr[k].im;
    }
    omega_sq[n] = om;
    weight[n] = w;
    if (which == 0) { /* ty_metric.rs:479-496 */
      int zi = argmax_abs(xr, 0, 4), wi = argmax_abs(xr, 4, 8);
      idx[3 * n] = zi < 0 ? 0 : zi;
      idx[3 * n + 1] = wi < 0 ? 0 : wi - 4;
      idx[3 * n + 2] = -1;
    } else { /* schoen_metric.rs:1132-1151 */
      int xi = argmax_abs(xr, 0, 3), yi = argmax_abs(xr, 3, 6), ti = argmax_abs(xr, 6, 8);
      idx[3 * n] = xi < 0 ? 0 : xi;
      idx[3 * n + 1] = yi < 0 ? 3 : yi;
      idx[3 * n + 2] = ti < 0 ? 6 : ti;
    }
    n++;
  }
  double sum = 0.0;
  for (int64_t i = 0; i < n; i++) sum += weight[i];
  if (isfinite(sum) && sum > 0.0) {
    if (which == 0) {
      for (int64_t i = 0; i < n; i++) weight[i] /= sum;
    } else {
      double inv = 1.0 / sum;
      for (int64_t i = 0; i < n; i++) weight[i] *= inv;
    }
  }
  if (n_lines) *n_lines = lines;
  return n;
}

/* Affine-chart tangent frames (ty_metric.rs:1929-2015; schoen_metric.rs:2202-2273), packed as
 * frames[p][a][i][re, im] (cy3_sigma_gpu.rs pack_frame); NaN for a degenerate chart. */
ORC_API void orc_cicy_frames(int which, int64_t n, const double *coords, const int32_t *idx,
                             double *frames) {
  variety v;
  make_variety(which, &v);
  for (int64_t p = 0; p < n; p++) {
    cx pt[8];
    for (int k = 0; k < 8; k++) pt[k] = C(coords[p * 16 + 2 * k], coords[p * 16 + 2 * k + 1]);
    int taken[8] = {0}, elim[3], freec[3], nf = 0, ok = 1;
    if (which == 0) {
      taken[idx[3 * p]] = 1;
      taken[4 + idx[3 * p + 1]] = 1;
    } else {
      taken[idx[3 * p]] = taken[idx[3 * p + 1]] = taken[idx[3 * p + 2]] = 1;
    }
    cx jac[24], je[9], jinv[9], fr[3][8];
    jacobian(&v, pt, jac);
    ok = pick_elim(jac, v.nh, taken, elim);
    if (ok) {
      for (int k = 0; k < 8 && ok; k++)
        if (!taken[k]) {
          if (nf >= 3) ok = 0; else freec[nf++] = k;
        }
      if (nf != 3) ok = 0;
    }
    if (ok) {
      for (int i = 0; i < v.nh; i++)
        for (int c = 0; c < v.nh; c++) je[i * v.nh + c] = jac[i * 8 + elim[c]];
      ok = inv_n(je, v.nh, jinv);
    }
    double *out = frames + p * 48;
    if (!ok) {
      for (int i = 0; i < 48; i++) out[i] = NAN;
      continue;
    }
    for (int a = 0; a < 3; a++)
      for (int k = 0; k < 8; k++) fr[a][k] = C(0.0, 0.0);
    for (int a = 0; a < 3; a++) {
      fr[a][freec[a]] = C(1.0, 0.0);
      cx b[3];
      for (int i = 0; i < v.nh; i++) b[i] = jac[i * 8 + freec[a]];
      for (int c = 0; c < v.nh; c++) {
        cx e = C(0.0, 0.0);
        for (int i = 0; i < v.nh; i++) e = cadd(e, cmul(jinv[c * v.nh + i], b[i]));
        fr[a][elim[c]] = cneg(e);
      }
    }
    for (int a = 0; a < 3; a++)
      for (int k = 0; k < 8; k++) {
        out[(a * 8 + k) * 2] = fr[a][k].re;
        out[(a * 8 + k) * 2 + 1] = fr[a][k].im;
      }
  }
}

/* Monomial section values and their 8 holomorphic derivatives from a power table
 * (ty_metric.rs:897-1035; schoen_metric.rs:1214-1333): sv [n][2 nb], sd [n][8][2 nb]. */
ORC_API void orc_cicy_sections(int64_t n, const double *coords, int64_t nb, const uint32_t *exps,
                               double *sv, double *sd) {
  uint32_t kmax = 0;
  for (int64_t i = 0; i < nb * 8; i++)
    if (exps[i] > kmax) kmax = exps[i];
  const int stride = (int)kmax + 1;
  cx *pw = (cx *)malloc(sizeof(cx) * 8 * stride);
  for (int64_t p = 0; p < n; p++) {
    for (int k = 0; k < 8; k++) {
      cx c = C(coords[p * 16 + 2 * k], coords[p * 16 + 2 * k + 1]);
      pw[k * stride] = C(1.0, 0.0);
      for (int e = 1; e < stride; e++) pw[k * stride + e] = cmul(pw[k * stride + e - 1], c);
    }
    for (int64_t a = 0; a < nb; a++) {
      const uint32_t *m = exps + a * 8;
      cx acc = C(1.0, 0.0);
      for (int k = 0; k < 8; k++) acc = cmul(acc, pw[k * stride + m[k]]);
      sv[p * 2 * nb + 2 * a] = acc.re;
      sv[p * 2 * nb + 2 * a + 1] = acc.im;
      for (int k = 0; k < 8; k++) {
        cx d = C(0.0, 0.0);
        if (m[k] != 0) {
          d = C((double)m[k], 0.0);
          d = cmul(d, pw[k * stride + m[k] - 1]);
          for (int j = 0; j < 8; j++)
            if (j != k) d = cmul(d, pw[j * stride + m[j]]);
        }
        double *row = sd + (p * 8 + k) * 2 * nb;
        row[2 * a] = d.re;
        row[2 * a + 1] = d.im;
      }
    }
  }
  free(pw);
}

/* Raw generator access for the known-answer tests. */
ORC_API void orc_chacha_block(const uint32_t *key, uint64_t counter, uint64_t stream, int rounds,
                              uint32_t *out) {
  chacha_block(key, counter, stream, rounds, out);
}

ORC_API void orc_chacha8_u64_stream(uint64_t seed, int64_t n, uint64_t *out) {
  orc_chacha g;
  chacha_seed_u64(&g, seed, 8);
  for (int64_t i = 0; i < n; i++) out[i] = chacha_next_u64(&g);
}

ORC_API void orc_std_normal_stream(uint64_t seed, int64_t n, double *out) {
  zig_init();
  orc_chacha g;
  chacha_seed_u64(&g, seed, 8);
  for (int64_t i = 0; i < n; i++) out[i] = std_normal(&g);
}

ORC_API void orc_zig_tables(double *x, double *f) {
  zig_init();
  memcpy(x, ZX, sizeof ZX);
  memcpy(f, ZF, sizeof ZF);
}

/* post_process_solved_point on a caller's ambient point x [16] (re, im); returns 1 and the
 * rescaled coordinates, |Omega|^2 and the unnormalised weight when accepted. */
ORC_API int orc_cicy_post_process(int which, const double *x16, double *xr16, double *omega_sq,
                                  double *weight) {
  variety v;
  make_variety(which, &v);
  cx x[8], xr[8];
  for (int k = 0; k < 8; k++) x[k] = C(x16[2 * k], x16[2 * k + 1]);
  int ok = post_process(&v, x, xr, omega_sq, weight);
  if (ok)
    for (int k = 0; k < 8; k++) { xr16[2 * k] = xr[k].re; xr16[2 * k + 1] = xr[k].im; }
  return ok;
}

/* ------------------------------------------------------------------------------------------ */
/* The reference's batch sampler with one Newton attempt per line (the `gpu` build of         */
/* CicySampler::sample_batch: sample_batch_gpu, cicy_sampler.rs:637-752, and its kernel       */
/* newton_solve_lines, gpu_sampler.rs:60-412), restated on the CPU. Lines are drawn for       */
/* max(4 n_target, 64) threads up front (origin, three directions, then t_j as (re, im) per   */
/* j), each thread runs one damped Newton solve, and the first n_target converged lines that  */
/* pass post_process_solved_point are kept in thread order.                                   */

static cx lane_cpow(cx b, int n) { /* repeated squaring (gpu_sampler.rs:61-78) */
  double rr = 1.0, ri = 0.0, br = b.re, bi = b.im;
  while (n > 0) {
    if (n & 1) {
      double nr = rr * br - ri * bi;
      double ni = rr * bi + ri * br;
      rr = nr; ri = ni;
    }
    n >>= 1;
    if (n > 0) {
      double nbr = br * br - bi * bi;
      double nbi = 2.0 * br * bi;
      br = nbr; bi = nbi;
    }
  }
  return C(rr, ri);
}

/* value (gpu_sampler.rs:83-106) */
static cx lane_eval(const spoly *f, const cx *z) {
  double pr = 0.0, pi = 0.0;
  for (int m = 0; m < f->nterm; m++) {
    double tr = 1.0, ti = 0.0;
    for (int j = 0; j < 8; j++) {
      int e = f->e[m][j];
      if (e == 0) continue;
      cx pw = lane_cpow(z[j], e);
      double nr = tr * pw.re - ti * pw.im, ni = tr * pw.im + ti * pw.re;
      tr = nr; ti = ni;
    }
    pr += tr; pi += ti;
  }
  return C(pr, pi);
}

/* value and gradient, d_k V = e_k V / z_k (gpu_sampler.rs:120-177) */
static cx lane_eval_grad(const spoly *f, const cx *z, cx *grad) {
  double vra = 0.0, via = 0.0;
  for (int k = 0; k < 8; k++) grad[k] = C(0.0, 0.0);
  for (int m = 0; m < f->nterm; m++) {
    cx pw[8];
    for (int j = 0; j < 8; j++) pw[j] = lane_cpow(z[j], f->e[m][j]);
    double vr = 1.0, vi = 0.0;
    for (int j = 0; j < 8; j++) {
      if (f->e[m][j] == 0) continue;
      double nr = vr * pw[j].re - vi * pw[j].im, ni = vr * pw[j].im + vi * pw[j].re;
      vr = nr; vi = ni;
    }
    vra += vr; via += vi;
    for (int k = 0; k < 8; k++) {
      int e = f->e[m][k];
      if (e == 0) continue;
      double m2 = z[k].re * z[k].re + z[k].im * z[k].im;
      if (m2 < 1.0e-300) continue;
      double ir = z[k].re / m2, ii = -z[k].im / m2;
      double dr = vr * ir - vi * ii, di = vr * ii + vi * ir;
      grad[k].re += (double)e * dr;
      grad[k].im += (double)e * di;
    }
  }
  return C(vra, via);
}

/* LU with partial pivoting on |.|^2, 1e-300 singular floor (gpu_sampler.rs:181-243) */
static int lane_solve3(double *mr, double *mi, double *br, double *bi) {
  const int N = 3;
  for (int k = 0; k < N; k++) {
    int p = k;
    double best = mr[k * N + k] * mr[k * N + k] + mi[k * N + k] * mi[k * N + k];
    for (int i = k + 1; i < N; i++) {
      double mag = mr[i * N + k] * mr[i * N + k] + mi[i * N + k] * mi[i * N + k];
      if (mag > best) { best = mag; p = i; }
    }
    if (best < 1.0e-300) return 0;
    if (p != k) {
      for (int j = 0; j < N; j++) {
        double tr = mr[k * N + j], ti = mi[k * N + j];
        mr[k * N + j] = mr[p * N + j]; mi[k * N + j] = mi[p * N + j];
        mr[p * N + j] = tr; mi[p * N + j] = ti;
      }
      double tr = br[k], ti = bi[k];
      br[k] = br[p]; bi[k] = bi[p];
      br[p] = tr; bi[p] = ti;
    }
    double pr = mr[k * N + k], pi = mi[k * N + k], pm = pr * pr + pi * pi;
    for (int i = k + 1; i < N; i++) {
      double ar = mr[i * N + k], ai = mi[i * N + k];
      double fr = (ar * pr + ai * pi) / pm, fi = (ai * pr - ar * pi) / pm;
      for (int j = k; j < N; j++) {
        double tr = fr * mr[k * N + j] - fi * mi[k * N + j];
        double ti = fr * mi[k * N + j] + fi * mr[k * N + j];
        mr[i * N + j] -= tr; mi[i * N + j] -= ti;
      }
      double tr = fr * br[k] - fi * bi[k], ti = fr * bi[k] + fi * br[k];
      br[i] -= tr; bi[i] -= ti;
    }
  }
  for (int i = N - 1; i >= 0; i--) {
    double sr = br[i], si = bi[i];
    for (int j = i + 1; j < N; j++) {
      sr -= mr[i * N + j] * br[j] - mi[i * N + j] * bi[j];
      si -= mr[i * N + j] * bi[j] + mi[i * N + j] * br[j];
    }
    double dr = mr[i * N + i], di = mi[i * N + i], dm = dr * dr + di * di;
    if (dm < 1.0e-300) return 0;
    br[i] = (sr * dr + si * di) / dm;
    bi[i] = (si * dr - sr * di) / dm;
  }
  return 1;
}

static void lane_x(const cx *p, const cx d[][8], const double *tr, const double *ti, cx *x) {
  for (int k = 0; k < 8; k++) {
    double sr = p[k].re, si = p[k].im;
    for (int j = 0; j < 3; j++) {
      sr += tr[j] * d[j][k].re - ti[j] * d[j][k].im;
      si += tr[j] * d[j][k].im + ti[j] * d[j][k].re;
    }
    x[k] = C(sr, si);
  }
}

/* one damped Newton attempt (gpu_sampler.rs:265-412): max_iter 64, tol 1e-8, up to 8 step
 * halvings accepted on any decrease */
static int lane_newton(const variety *v, const cx *p, const cx d[][8], double *tr, double *ti) {
  int conv = 0;
  for (int it = 0; it < 64; it++) {
    cx x[8], f[3], J[9];
    lane_x(p, d, tr, ti, x);
    for (int i = 0; i < 3; i++) {
      cx g[8];
      f[i] = lane_eval_grad(&v->f[i], x, g);
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
    if (res < 1.0e-8) { conv = 1; break; }
    double mr[9], mi[9], br[3], bi[3];
    for (int i = 0; i < 9; i++) { mr[i] = J[i].re; mi[i] = J[i].im; }
    for (int i = 0; i < 3; i++) { br[i] = -f[i].re; bi[i] = -f[i].im; }
    if (!lane_solve3(mr, mi, br, bi)) break;
    double damp = 1.0, nr_[3], ni_[3], new_res = 0.0;
    for (int trial = 0; trial < 8; trial++) {
      for (int j = 0; j < 3; j++) { nr_[j] = tr[j] + damp * br[j]; ni_[j] = ti[j] + damp * bi[j]; }
      cx x2[8];
      lane_x(p, d, nr_, ni_, x2);
      new_res = 0.0;
      for (int i = 0; i < 3; i++) {
        cx fv = lane_eval(&v->f[i], x2);
        new_res += fv.re * fv.re + fv.im * fv.im;
      }
      new_res = sqrt(new_res);
      if (new_res < res) break;
      damp *= 0.5;
    }
    if (new_res >= res) break;
    for (int j = 0; j < 3; j++) { tr[j] = nr_[j]; ti[j] = ni_[j]; }
    if (new_res < 1.0e-8) { conv = 1; break; }
  }
  return conv;
}

ORC_API int64_t orc_ty_sample_lanes(uint64_t seed, int64_t n_target, double *coords,
                                    double *omega_sq, double *weight, int32_t *idx,
                                    int32_t *lane_ok, int64_t *n_lanes_out) {
  zig_init();
  variety v;
  make_variety(0, &v);
  orc_chacha g;
  chacha_seed_u64(&g, seed, 8);
  int64_t nl = n_target * 4 > 64 ? n_target * 4 : 64;
  cx *P = (cx *)malloc(sizeof(cx) * 8 * nl), *D = (cx *)malloc(sizeof(cx) * 24 * nl);
  double *T = (double *)malloc(sizeof(double) * 6 * nl);
  for (int64_t l = 0; l < nl; l++) {
    gaussian_unit_sphere(&g, P + 8 * l);
    for (int j = 0; j < 3; j++) gaussian_unit_sphere(&g, D + 24 * l + 8 * j);
    for (int j = 0; j < 3; j++) {
      T[6 * l + 2 * j] = std_normal(&g);
      T[6 * l + 2 * j + 1] = std_normal(&g);
    }
  }
  int64_t n = 0;
  for (int64_t l = 0; l < nl; l++) {
    if (lane_ok) lane_ok[l] = 0;
    if (n >= n_target) continue;
    double tr[3], ti[3];
    for (int j = 0; j < 3; j++) { tr[j] = T[6 * l + 2 * j]; ti[j] = T[6 * l + 2 * j + 1]; }
    const cx(*d)[8] = (const cx(*)[8])(D + 24 * l);
    if (!lane_newton(&v, P + 8 * l, d, tr, ti)) continue;
    cx x[8], xr[8];
    for (int k = 0; k < 8; k++) { /* cicy_sampler.rs:718-731 */
      cx s = P[8 * l + k];
      for (int j = 0; j < 3; j++) s = cadd(s, cmul(C(tr[j], ti[j]), d[j][k]));
      x[k] = s;
    }
    double om, w;
    if (!post_process(&v, x, xr, &om, &w)) continue;
    if (lane_ok) lane_ok[l] = 1;
    for (int k = 0; k < 8; k++) { coords[n * 16 + 2 * k] = xr[k].re; coords[n * 16 + 2 * k + 1] = xr[k].im; }
    omega_sq[n] = om;
    weight[n] = w;
    int zi = argmax_abs(xr, 0, 4), wi = argmax_abs(xr, 4, 8);
    idx[3 * n] = zi < 0 ? 0 : zi;
    idx[3 * n + 1] = wi < 0 ? 0 : wi - 4;
    idx[3 * n + 2] = -1;
    n++;
  }
  double sum = 0.0;
  for (int64_t i = 0; i < n; i++) sum += weight[i];
  if (isfinite(sum) && sum > 0.0)
    for (int64_t i = 0; i < n; i++) weight[i] /= sum;
  free(P); free(D); free(T);
  if (n_lanes_out) *n_lanes_out = nl;
  return n;
}
