/*
 * cy3q oracle: a CPU FP64 restatement of the reference's quintic metric hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker. Only tests/, the smoke entry point and
 * the cpu_baseline leg of bench.py load it (through oracle/oracle.py). The product library
 * (papers_amd/libcy3q.so) never links, loads or calls it.
 *
 * Every function follows the reference Rust line for line in operation order. Citations are
 * relative to substrate_sm_derivation/cy3_solver/ in the reference. Build with
 * -ffp-contract=off, because Rust never fuses a*b+c into an FMA. libm (log, cos, sin, pow,
 * exp) is the same glibc the reference binds on Linux.
 *
 * Pins, as checked by tests/test_oracle_*.py:
 *   - LCG stream: closed-form integer recurrence plus Box–Muller. Pinned by integer identities and
 *     brute force.
 *   - Monomial tables: counts from quintic.rs:5472-5484, the closed form C(k+4,4), and the
 *     quotient count C(k+4,4)-C(k-1,4). Bit-exact.
 *   - Sampler, weights, log|Omega|^2, sections, FS-Gram and sigma: pinned end to end by the
 *     published sigma_fs_identity values (output/p5_9_ako_apples_to_apples.json, k=3 and k=4,
 *     200k points). Also by the on-variety residual, the Vieta identities of the DK roots, the
 *     psi=0 Dwork == Fermat identity, and chart invariance of eta.
 *   - Donaldson step (LU inverse): pinned by the published sigma_donaldson values and iteration
 *     counts (same file), and by tr T = N, Hermiticity and G*T = I.
 *   - sigma^2 gradient: pinned by central finite differences of sigma^2.
 *   - Adam: the inline formula of quintic.rs:5421-5437. The bit-level behaviour of
 *     pwos_math::opt::adam is not vendored, so that part is "parity unpinned".
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------------------------ */
/* LCG (cy3/src/lib.rs:222-245)                                                               */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
    uint64_t s;
} orc_lcg;

static void lcg_init(orc_lcg *r, uint64_t seed) { /* lib.rs:226-228 */
    r->s = seed * 6364136223846793005ULL + 1ULL;
}

static double lcg_next_f64(orc_lcg *r) { /* lib.rs:230-237 */
    r->s = r->s * 6364136223846793005ULL + 1442695040888963407ULL;
    uint64_t bits = (r->s >> 11) & ((1ULL << 53) - 1ULL);
    return (double)bits / (double)(1ULL << 53);
}

static double lcg_next_normal(orc_lcg *r) { /* lib.rs:239-244 */
    double u1 = fmax(lcg_next_f64(r) + 1e-12, 1e-12);
    double u2 = lcg_next_f64(r);
    double rr = sqrt(-2.0 * log(u1));
    return rr * cos(2.0 * M_PI * u2);
}

ORC_API void orc_lcg_stream(uint64_t seed, int64_t n, uint64_t *states, double *f64s,
                            double *normals) {
    orc_lcg r;
    lcg_init(&r, seed);
    for (int64_t i = 0; i < n; i++) {
        double u = lcg_next_f64(&r);
        if (states) states[i] = r.s;
        if (f64s) f64s[i] = u;
    }
    if (normals) {
        lcg_init(&r, seed);
        for (int64_t i = 0; i < n; i++) normals[i] = lcg_next_normal(&r);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Monomial tables (quintic.rs:3185-3209; quotient rule route34/ty_metric.rs:873-893)          */
/* ------------------------------------------------------------------------------------------ */
ORC_API int64_t orc_n_sections(uint32_t k) { /* quintic.rs:3185-3188 */
    int64_t kk = k;
    return (kk + 1) * (kk + 2) * (kk + 3) * (kk + 4) / 24;
}

/* Enumerates e0 up, e1 up, e2 up, e3 up, e4 = k - sum (quintic.rs:3192-3209). With quotient != 0,
 * monomials in the leading-monomial ideal <z0^5> of the quintic (e0 >= 5) are dropped. The raw
 * order is kept, as ty_metric.rs:873-893 does for its Groebner-reduced basis. */
ORC_API int64_t orc_monomials(uint32_t k, int quotient, uint32_t *out) {
    int64_t n = 0;
    int kk = (int)k;
    for (int e0 = 0; e0 <= kk; e0++)
        for (int e1 = 0; e1 <= kk - e0; e1++)
            for (int e2 = 0; e2 <= kk - e0 - e1; e2++)
                for (int e3 = 0; e3 <= kk - e0 - e1 - e2; e3++) {
                    int e4 = kk - e0 - e1 - e2 - e3;
                    if (e4 < 0) continue;
                    if (quotient && e0 >= 5) continue;
                    if (out) {
                        out[5 * n + 0] = (uint32_t)e0;
                        out[5 * n + 1] = (uint32_t)e1;
                        out[5 * n + 2] = (uint32_t)e2;
                        out[5 * n + 3] = (uint32_t)e3;
                        out[5 * n + 4] = (uint32_t)e4;
                    }
                    n++;
                }
    return n;
}

/* ------------------------------------------------------------------------------------------ */
/* Quintic polynomial and gradient (quintic.rs:3213-3323, 3985-4001)                          */
/* ------------------------------------------------------------------------------------------ */
static void fermat_quintic(const double *z, double *fre, double *fim) { /* :3213-3235 */
    double re = 0.0, im = 0.0;
    for (int j = 0; j < 5; j++) {
        double r = z[2 * j], i = z[2 * j + 1];
        double r2 = r * r, i2 = i * i;
        double z2_re = r2 - i2;
        double z2_im = 2.0 * r * i;
        double z4_re = z2_re * z2_re - z2_im * z2_im;
        double z4_im = 2.0 * z2_re * z2_im;
        double z5_re = r * z4_re - i * z4_im;
        double z5_im = r * z4_im + i * z4_re;
        re += z5_re;
        im += z5_im;
    }
    *fre = re;
    *fim = im;
}

static void deformed_fermat_quintic(const double *z, double psi, double *fre, double *fim) {
    /* :3248-3268 */
    double re, im;
    fermat_quintic(z, &re, &im);
    if (psi != 0.0) {
        double p_re = 1.0, p_im = 0.0;
        for (int j = 0; j < 5; j++) {
            double r = z[2 * j], i = z[2 * j + 1];
            double nr = p_re * r - p_im * i;
            double ni = p_re * i + p_im * r;
            p_re = nr;
            p_im = ni;
        }
        double scale = -5.0 * psi;
        re += scale * p_re;
        im += scale * p_im;
    }
    *fre = re;
    *fim = im;
}

static void fermat_quintic_gradient(const double *z, double *g) { /* :3305-3323 */
    for (int j = 0; j < 5; j++) {
        double r = z[2 * j], i = z[2 * j + 1];
        double r2 = r * r, i2 = i * i;
        double z2_re = r2 - i2;
        double z2_im = 2.0 * r * i;
        double z4_re = z2_re * z2_re - z2_im * z2_im;
        double z4_im = 2.0 * z2_re * z2_im;
        g[2 * j] = 5.0 * z4_re;
        g[2 * j + 1] = 5.0 * z4_im;
    }
}

static void deformed_gradient(const double *z, double psi, double *g) { /* :3276-3300 */
    fermat_quintic_gradient(z, g);
    if (psi == 0.0) return;
    double scale = -5.0 * psi;
    for (int k = 0; k < 5; k++) {
        double p_re = 1.0, p_im = 0.0;
        for (int j = 0; j < 5; j++) {
            if (j == k) continue;
            double r = z[2 * j], i = z[2 * j + 1];
            double nr = p_re * r - p_im * i;
            double ni = p_re * i + p_im * r;
            p_re = nr;
            p_im = ni;
        }
        g[2 * k] += scale * p_re;
        g[2 * k + 1] += scale * p_im;
    }
}

ORC_API void orc_quintic_eval(const double *pts, int64_t n, double psi, double *f_out,
                              double *grad_out) {
    for (int64_t p = 0; p < n; p++) {
        if (f_out) deformed_fermat_quintic(pts + 10 * p, psi, f_out + 2 * p, f_out + 2 * p + 1);
        if (grad_out) deformed_gradient(pts + 10 * p, psi, grad_out + 10 * p);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Shiffman–Zelditch line sampler (quintic.rs:3633-3785 Fermat; :3810-3983 Dwork)              */
/* ------------------------------------------------------------------------------------------ */
/* One line: draws a, b, builds the monic quintic in t, runs Durand–Kerner and writes the 5
 * candidate z (normalised to |z| = 1) with an accept flag each. Returns 0 when the line is
 * degenerate (|c5|^2 < 1e-300) and the reference 'continue's before emitting anything. */
static int sz_line(orc_lcg *rng, int dwork, double psi, int dk_iters, double zs[5][10],
                   int acc[5]) {
    double a[5][2], b[5][2];
    for (int i = 0; i < 5; i++) {
        a[i][0] = lcg_next_normal(rng);
        a[i][1] = lcg_next_normal(rng);
        b[i][0] = lcg_next_normal(rng);
        b[i][1] = lcg_next_normal(rng);
    }
    static const double bin[6] = {1.0, 5.0, 10.0, 10.0, 5.0, 1.0};
    double c[6][2];
    memset(c, 0, sizeof c);
    for (int i = 0; i < 5; i++) {
        double ap[6][2], bp[6][2];
        ap[0][0] = 1.0;
        ap[0][1] = 0.0;
        bp[0][0] = 1.0;
        bp[0][1] = 0.0;
        for (int e = 1; e <= 5; e++) {
            double pr = ap[e - 1][0], pi = ap[e - 1][1];
            ap[e][0] = pr * a[i][0] - pi * a[i][1];
            ap[e][1] = pr * a[i][1] + pi * a[i][0];
            pr = bp[e - 1][0];
            pi = bp[e - 1][1];
            bp[e][0] = pr * b[i][0] - pi * b[i][1];
            bp[e][1] = pr * b[i][1] + pi * b[i][0];
        }
        for (int j = 0; j <= 5; j++) {
            double ar = ap[5 - j][0], ai = ap[5 - j][1];
            double br = bp[j][0], bi = bp[j][1];
            double pr = ar * br - ai * bi;
            double pi = ar * bi + ai * br;
            c[j][0] += bin[j] * pr;
            c[j][1] += bin[j] * pi;
        }
    }
    if (dwork) { /* :3851-3877 */
        double prod[6][2];
        memset(prod, 0, sizeof prod);
        prod[0][0] = 1.0;
        int deg = 0;
        for (int i = 0; i < 5; i++) {
            double next[6][2];
            memset(next, 0, sizeof next);
            for (int j = 0; j <= deg; j++) {
                double pr = prod[j][0], pi = prod[j][1];
                double ar = a[i][0], ai = a[i][1];
                next[j][0] += pr * ar - pi * ai;
                next[j][1] += pr * ai + pi * ar;
                double br = b[i][0], bi = b[i][1];
                next[j + 1][0] += pr * br - pi * bi;
                next[j + 1][1] += pr * bi + pi * br;
            }
            memcpy(prod, next, sizeof prod);
            deg += 1;
        }
        double scale = 5.0 * psi;
        for (int j = 0; j <= 5; j++) {
            c[j][0] -= scale * prod[j][0];
            c[j][1] -= scale * prod[j][1];
        }
    }
    double cr5 = c[5][0], ci5 = c[5][1];
    double denom = cr5 * cr5 + ci5 * ci5;
    if (denom < 1e-300) return 0;
    double inv_re = cr5 / denom, inv_im = -ci5 / denom;
    double p[6][2];
    for (int j = 0; j <= 5; j++) {
        double cr = c[j][0], ci = c[j][1];
        p[j][0] = cr * inv_re - ci * inv_im;
        p[j][1] = cr * inv_im + ci * inv_re;
    }
    double roots[5][2];
    double radius = 1.0;
    for (int j = 0; j < 5; j++) {
        double mag = sqrt(p[j][0] * p[j][0] + p[j][1] * p[j][1]);
        int ex = 5 - j;
        double r = pow(mag, 1.0 / (double)ex);
        if (r > radius) radius = r;
    }
    for (int k = 0; k < 5; k++) {
        double theta = (2.0 * M_PI / 5.0) * (double)k + 0.4;
        roots[k][0] = radius * cos(theta);
        roots[k][1] = radius * sin(theta);
    }
    for (int it = 0; it < dk_iters; it++) {
        double prev[5][2];
        memcpy(prev, roots, sizeof prev);
        for (int k = 0; k < 5; k++) {
            double rkr = prev[k][0], rki = prev[k][1];
            double val_re = 1.0, val_im = 0.0;
            for (int j = 4; j >= 0; j--) {
                double nr = val_re * rkr - val_im * rki + p[j][0];
                double ni = val_re * rki + val_im * rkr + p[j][1];
                val_re = nr;
                val_im = ni;
            }
            double den_re = 1.0, den_im = 0.0;
            for (int j = 0; j < 5; j++) {
                if (j == k) continue;
                double dr = rkr - prev[j][0];
                double di = rki - prev[j][1];
                double nr = den_re * dr - den_im * di;
                double ni = den_re * di + den_im * dr;
                den_re = nr;
                den_im = ni;
            }
            double den_norm = den_re * den_re + den_im * den_im;
            if (den_norm > 1e-300) {
                double ir = den_re / den_norm, ii = -den_im / den_norm;
                double dr = val_re * ir - val_im * ii;
                double di = val_re * ii + val_im * ir;
                roots[k][0] = rkr - dr;
                roots[k][1] = rki - di;
            }
        }
    }
    for (int k = 0; k < 5; k++) {
        acc[k] = 0;
        double tr = roots[k][0], ti = roots[k][1];
        if (!isfinite(tr) || !isfinite(ti)) continue;
        double *z = zs[k];
        for (int i = 0; i < 5; i++) {
            z[2 * i] = a[i][0] + tr * b[i][0] - ti * b[i][1];
            z[2 * i + 1] = a[i][1] + tr * b[i][1] + ti * b[i][0];
        }
        double nrm_sq = 0.0;
        for (int q = 0; q < 10; q++) nrm_sq += z[q] * z[q];
        if (nrm_sq < 1e-30) continue;
        double nrm = sqrt(nrm_sq);
        for (int q = 0; q < 10; q++) z[q] /= nrm;
        double fr, fi;
        if (dwork)
            deformed_fermat_quintic(z, psi, &fr, &fi);
        else
            fermat_quintic(z, &fr, &fi);
        double fmag = sqrt(fr * fr + fi * fi);
        if (fmag > (dwork ? 1e-2 : 1e-3)) continue;
        acc[k] = 1;
    }
    return 1;
}

/* dwork = 0: sample_quintic_points_sz (Fermat, accept |f| <= 1e-3, max_tries = 4 * lines).
 * dwork = 1: sample_deformed_quintic_points_sz (accept |f| <= 1e-2, max_tries = 8 * lines).
 * out holds n * 10 doubles; returns the number of points written. */
ORC_API int64_t orc_sample_sz(int64_t n, uint64_t seed, int dwork, double psi, int dk_iters,
                              double *out, int64_t *lines_tried) {
    orc_lcg rng;
    lcg_init(&rng, seed);
    int64_t n_lines_target = (n + 4) / 5;
    int64_t max_tries = n_lines_target * (dwork ? 8 : 4);
    int64_t tried = 0, np = 0;
    while (np < n && tried < max_tries) {
        tried++;
        double zs[5][10];
        int acc[5];
        if (!sz_line(&rng, dwork, psi, dk_iters, zs, acc)) continue;
        for (int k = 0; k < 5; k++) {
            if (np >= n) break;
            if (!acc[k]) continue;
            memcpy(out + 10 * np, zs[k], sizeof(double) * 10);
            np++;
        }
    }
    if (lines_tried) *lines_tried = tried;
    return np;
}

/* Per-line candidate generator (test helper): the 5 candidates of each line in order with
 * accept flags; used to check a line-parallel sampler line by line. */
ORC_API void orc_sz_lines(int64_t n_lines, uint64_t seed, int dwork, double psi, int dk_iters,
                          double *zs_out, int32_t *acc_out) {
    orc_lcg rng;
    lcg_init(&rng, seed);
    for (int64_t l = 0; l < n_lines; l++) {
        double zs[5][10];
        int acc[5];
        memset(zs, 0, sizeof zs);
        if (!sz_line(&rng, dwork, psi, dk_iters, zs, acc))
            for (int k = 0; k < 5; k++) acc[k] = 0;
        memcpy(zs_out + 50 * l, zs, sizeof zs);
        for (int k = 0; k < 5; k++) acc_out[5 * l + k] = acc[k];
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Weights and log|Omega|^2 (quintic.rs:3507-3563, 4858-4916)                                  */
/* ------------------------------------------------------------------------------------------ */
static void chart_and_elim(const double *z, const double *g, int *chart_o, int *elim_o,
                           double *log_om) { /* :4870-4916 */
    double max_z_sq = 0.0;
    int chart = 0;
    for (int k = 0; k < 5; k++) {
        double zsq = z[2 * k] * z[2 * k] + z[2 * k + 1] * z[2 * k + 1];
        if (zsq > max_z_sq) {
            max_z_sq = zsq;
            chart = k;
        }
    }
    double max_g_sq = 0.0;
    int elim = chart == 0 ? 1 : 0;
    for (int k = 0; k < 5; k++) {
        if (k == chart) continue;
        double gsq = g[2 * k] * g[2 * k] + g[2 * k + 1] * g[2 * k + 1];
        if (gsq > max_g_sq) {
            max_g_sq = gsq;
            elim = k;
        }
    }
    *chart_o = chart;
    *elim_o = elim;
    if (log_om) *log_om = (max_g_sq < 1e-20 || max_z_sq < 1e-20) ? NAN : log(max_z_sq) - log(max_g_sq);
}

ORC_API void orc_weights_logomega(const double *pts, int64_t n, double psi, double *w,
                                  double *log_om, int32_t *chart_elim) {
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; p++) {
        const double *z = pts + 10 * p;
        double g[10];
        deformed_gradient(z, psi, g);
        double gs = 0.0;
        for (int k = 0; k < 10; k++) gs += g[k] * g[k];
        if (w) w[p] = gs < 1e-30 ? 0.0 : 1.0 / gs;
        int c, e;
        double lo;
        chart_and_elim(z, g, &c, &e, &lo);
        if (log_om) log_om[p] = lo;
        if (chart_elim) {
            chart_elim[2 * p] = c;
            chart_elim[2 * p + 1] = e;
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Section values and derivatives (quintic.rs:4021-4071, 1231-1296)                           */
/* ------------------------------------------------------------------------------------------ */
static void power_table(const double *z, int kmax, double *pw /* [5][kmax+1][2] */) {
    int stride = kmax + 1;
    for (int j = 0; j < 5; j++) {
        pw[j * stride * 2] = 1.0;
        pw[j * stride * 2 + 1] = 0.0;
    }
    for (int j = 0; j < 5; j++) {
        double zr = z[2 * j], zi = z[2 * j + 1];
        for (int e = 1; e <= kmax; e++) {
            double pr = pw[j * stride * 2 + (e - 1) * 2];
            double pi = pw[j * stride * 2 + (e - 1) * 2 + 1];
            pw[j * stride * 2 + e * 2] = pr * zr - pi * zi;
            pw[j * stride * 2 + e * 2 + 1] = pr * zi + pi * zr;
        }
    }
}

static int mono_kmax(const uint32_t *mono, int64_t nb) {
    uint32_t km = 0;
    for (int64_t i = 0; i < 5 * nb; i++)
        if (mono[i] > km) km = mono[i];
    return (int)km;
}

ORC_API void orc_sections(const double *pts, int64_t n, const uint32_t *mono, int64_t nb,
                          double *sv /* [n][2nb] */, double *sd /* [n][5][2nb] or NULL */) {
    int kmax = mono_kmax(mono, nb);
    int stride = kmax + 1;
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; p++) {
        double pw[5 * 64 * 2];
        power_table(pts + 10 * p, kmax, pw);
        double *row = sv + p * 2 * nb;
        for (int64_t a = 0; a < nb; a++) {
            double pr = 1.0, pi = 0.0;
            for (int j = 0; j < 5; j++) {
                int e = (int)mono[5 * a + j];
                double zr = pw[j * stride * 2 + e * 2], zi = pw[j * stride * 2 + e * 2 + 1];
                double nr = pr * zr - pi * zi;
                double ni = pr * zi + pi * zr;
                pr = nr;
                pi = ni;
            }
            row[2 * a] = pr;
            row[2 * a + 1] = pi;
        }
        if (!sd) continue;
        double *slot = sd + p * 10 * nb;
        int64_t two_n = 2 * nb;
        for (int i = 0; i < 5; i++) {
            for (int64_t a = 0; a < nb; a++) {
                const uint32_t *m = mono + 5 * a;
                if (m[i] == 0) {
                    slot[i * two_n + 2 * a] = 0.0;
                    slot[i * two_n + 2 * a + 1] = 0.0;
                    continue;
                }
                double fr = pw[i * stride * 2 + (m[i] - 1) * 2];
                double fi = pw[i * stride * 2 + (m[i] - 1) * 2 + 1];
                double mf = (double)m[i];
                double pr = mf * fr, pi = mf * fi;
                for (int j = 0; j < 5; j++) {
                    if (j == i) continue;
                    int e = (int)m[j];
                    double zr = pw[j * stride * 2 + e * 2], zi = pw[j * stride * 2 + e * 2 + 1];
                    double nr = pr * zr - pi * zi;
                    double ni = pr * zi + pi * zr;
                    pr = nr;
                    pi = ni;
                }
                slot[i * two_n + 2 * a] = pr;
                slot[i * two_n + 2 * a + 1] = pi;
            }
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* FS-Gram orthonormalisation (quintic.rs:2957-3165)                                           */
/* ------------------------------------------------------------------------------------------ */
ORC_API void orc_fs_gram(const double *sv, const double *w, int64_t n, int64_t nb, double *gre,
                         double *gim) { /* :2957-2995, serial in point order */
    int64_t two_n = 2 * nb;
    memset(gre, 0, sizeof(double) * nb * nb);
    memset(gim, 0, sizeof(double) * nb * nb);
    double total_w = 0.0;
    for (int64_t p = 0; p < n; p++) {
        double wp = w[p];
        if (!isfinite(wp) || wp <= 0.0) continue;
        total_w += wp;
        const double *s = sv + p * two_n;
        for (int64_t a = 0; a < nb; a++) {
            double sar = s[2 * a], sai = s[2 * a + 1];
            for (int64_t b = 0; b < nb; b++) {
                double sbr = s[2 * b], sbi = s[2 * b + 1];
                gre[a * nb + b] += wp * (sar * sbr + sai * sbi);
                gim[a * nb + b] += wp * (sar * sbi - sai * sbr);
            }
        }
    }
    double inv_w = total_w > 1e-12 ? 1.0 / total_w : 0.0;
    for (int64_t i = 0; i < nb * nb; i++) {
        gre[i] *= inv_w;
        gim[i] *= inv_w;
    }
}

ORC_API int orc_cholesky(const double *gre, const double *gim, int64_t n, double *lre,
                         double *lim) { /* :3002-3049; returns -1 when not PD */
    memset(lre, 0, sizeof(double) * n * n);
    memset(lim, 0, sizeof(double) * n * n);
    for (int64_t j = 0; j < n; j++) {
        double sdq = 0.0;
        for (int64_t k = 0; k < j; k++) {
            double lr = lre[j * n + k], li = lim[j * n + k];
            sdq += lr * lr + li * li;
        }
        double dv = gre[j * n + j] - sdq;
        if (dv <= 0.0) return -1;
        double ljj = sqrt(dv);
        lre[j * n + j] = ljj;
        lim[j * n + j] = 0.0;
        for (int64_t i = j + 1; i < n; i++) {
            double sr = gre[i * n + j], si = gim[i * n + j];
            for (int64_t k = 0; k < j; k++) {
                double ir = lre[i * n + k], ii = lim[i * n + k];
                double jr = lre[j * n + k], ji = lim[j * n + k];
                sr -= ir * jr + ii * ji;
                si -= ii * jr - ir * ji;
            }
            lre[i * n + j] = sr / ljj;
            lim[i * n + j] = si / ljj;
        }
    }
    return 0;
}

static void forward_solve_conj(const double *lre, const double *lim, double *b, int64_t n,
                               double *y) { /* :3062-3090, y = conj(L)^-1 b */
    for (int64_t i = 0; i < n; i++) {
        double sr = b[2 * i], si = b[2 * i + 1];
        for (int64_t j = 0; j < i; j++) {
            double lr = lre[i * n + j], li = -lim[i * n + j];
            double yr = y[2 * j], yi = y[2 * j + 1];
            sr -= lr * yr - li * yi;
            si -= lr * yi + li * yr;
        }
        double lii = lre[i * n + i];
        y[2 * i] = sr / lii;
        y[2 * i + 1] = si / lii;
    }
}

/* Applies conj(L)^-1 to every section row and, if sd != NULL, every derivative row
 * (apply_inverse_l_to_sections/derivs, :3095-3133). */
ORC_API void orc_apply_inv_l(const double *lre, const double *lim, int64_t n, int64_t nb,
                             double *sv, double *sd) {
    int64_t two_n = 2 * nb;
#pragma omp parallel
    {
        double *y = (double *)malloc(sizeof(double) * two_n);
#pragma omp for schedule(static)
        for (int64_t p = 0; p < n; p++) {
            double *s = sv + p * two_n;
            forward_solve_conj(lre, lim, s, nb, y);
            memcpy(s, y, sizeof(double) * two_n);
            if (sd) {
                for (int i = 0; i < 5; i++) {
                    double *d = sd + p * 5 * two_n + i * two_n;
                    forward_solve_conj(lre, lim, d, nb, y);
                    memcpy(d, y, sizeof(double) * two_n);
                }
            }
        }
        free(y);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Real LU inverse of the 2N x 2N block (stands in for pwos_math::linalg::invert, un-vendored:  */
/* Doolittle LU with partial pivoting, then one forward/back solve per identity column).      */
/* ------------------------------------------------------------------------------------------ */
ORC_API int orc_lu_invert(const double *a_in, int64_t n, double *inv) {
    double *a = (double *)malloc(sizeof(double) * n * n);
    int64_t *perm = (int64_t *)malloc(sizeof(int64_t) * n);
    double *col = (double *)malloc(sizeof(double) * n);
    memcpy(a, a_in, sizeof(double) * n * n);
    for (int64_t i = 0; i < n; i++) perm[i] = i;
    int rc = 0;
    for (int64_t k = 0; k < n && rc == 0; k++) {
        int64_t piv = k;
        double best = fabs(a[k * n + k]);
        for (int64_t i = k + 1; i < n; i++)
            if (fabs(a[i * n + k]) > best) {
                best = fabs(a[i * n + k]);
                piv = i;
            }
        if (best < 1e-300 || !isfinite(best)) {
            rc = -1;
            break;
        }
        if (piv != k) {
            for (int64_t j = 0; j < n; j++) {
                double t = a[k * n + j];
                a[k * n + j] = a[piv * n + j];
                a[piv * n + j] = t;
            }
            int64_t t = perm[k];
            perm[k] = perm[piv];
            perm[piv] = t;
        }
        double akk = a[k * n + k];
        for (int64_t i = k + 1; i < n; i++) {
            double l = a[i * n + k] / akk;
            a[i * n + k] = l;
            for (int64_t j = k + 1; j < n; j++) a[i * n + j] -= l * a[k * n + j];
        }
    }
    if (rc == 0) {
        for (int64_t c = 0; c < n; c++) {
            for (int64_t i = 0; i < n; i++) col[i] = (perm[i] == c) ? 1.0 : 0.0;
            for (int64_t i = 0; i < n; i++) {
                double s = col[i];
                for (int64_t j = 0; j < i; j++) s -= a[i * n + j] * col[j];
                col[i] = s;
            }
            for (int64_t i = n - 1; i >= 0; i--) {
                double s = col[i];
                for (int64_t j = i + 1; j < n; j++) s -= a[i * n + j] * col[j];
                col[i] = s / a[i * n + i];
            }
            for (int64_t i = 0; i < n; i++) inv[i * n + c] = col[i];
        }
    }
    free(a);
    free(perm);
    free(col);
    return rc;
}

/* ------------------------------------------------------------------------------------------ */
/* Donaldson T-operator step (quintic.rs:1734-1894)                                            */
/* ------------------------------------------------------------------------------------------ */
/* h_block [2N x 2N] in, h_new [2N x 2N] out. Optional outputs: t_re/t_im (the trace-normalised
 * lower-index T before inversion), kinv [n] (1/max(K,1e-30)). Returns the Frobenius residual
 * ||h_new - h_block||. */
ORC_API double orc_donaldson_step(const double *sv, const double *w, int64_t n, int64_t nb,
                                  const double *h_block, double *h_new, double *t_re,
                                  double *t_im, double *kinv_out) {
    int64_t two_n = 2 * nb;
    double *kinv = (double *)malloc(sizeof(double) * n);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; p++) { /* :1745-1759 */
        const double *s = sv + p * two_n;
        double k = 0.0;
        for (int64_t i = 0; i < two_n; i++) {
            double rs = 0.0;
            for (int64_t j = 0; j < two_n; j++) rs += h_block[i * two_n + j] * s[j];
            k += s[i] * rs;
        }
        kinv[p] = 1.0 / fmax(k, 1e-30);
    }
    double *hre = (double *)calloc(nb * nb, sizeof(double));
    double *him = (double *)calloc(nb * nb, sizeof(double));
    double total_w = 0.0;
    for (int64_t p = 0; p < n; p++) { /* :1771-1789, sequential over points */
        const double *s = sv + p * two_n;
        double wp = w[p];
        if (!isfinite(wp) || wp <= 0.0) continue;
        total_w += wp;
        double ik = kinv[p] * wp;
        for (int64_t a = 0; a < nb; a++) {
            double sar = s[2 * a], sai = s[2 * a + 1];
            for (int64_t b = 0; b < nb; b++) {
                double sbr = s[2 * b], sbi = s[2 * b + 1];
                hre[a * nb + b] += (sar * sbr + sai * sbi) * ik;
                him[a * nb + b] += (sar * sbi - sai * sbr) * ik;
            }
        }
    }
    double inv_w = total_w > 1e-12 ? 1.0 / total_w : 0.0;
    for (int64_t i = 0; i < nb * nb; i++) {
        hre[i] *= inv_w;
        him[i] *= inv_w;
    }
    double trace = 0.0;
    for (int64_t a = 0; a < nb; a++) trace += hre[a * nb + a];
    if (trace > 1e-10) {
        double scale = (double)nb / trace;
        for (int64_t i = 0; i < nb * nb; i++) {
            hre[i] *= scale;
            him[i] *= scale;
        }
    }
    if (t_re) memcpy(t_re, hre, sizeof(double) * nb * nb);
    if (t_im) memcpy(t_im, him, sizeof(double) * nb * nb);
    if (kinv_out) memcpy(kinv_out, kinv, sizeof(double) * n);
    double *blk = (double *)malloc(sizeof(double) * two_n * two_n);
    double *binv = (double *)malloc(sizeof(double) * two_n * two_n);
    for (int64_t a = 0; a < nb; a++)
        for (int64_t b = 0; b < nb; b++) {
            double r = hre[a * nb + b], i = him[a * nb + b];
            blk[(2 * a) * two_n + 2 * b] = r;
            blk[(2 * a) * two_n + 2 * b + 1] = -i;
            blk[(2 * a + 1) * two_n + 2 * b] = i;
            blk[(2 * a + 1) * two_n + 2 * b + 1] = r;
        }
    if (orc_lu_invert(blk, two_n, binv) == 0) { /* :1859-1872 Hermitian projection */
        for (int64_t a = 0; a < nb; a++)
            for (int64_t b = 0; b < nb; b++) {
                double mrr = binv[(2 * a) * two_n + 2 * b];
                double mri = binv[(2 * a) * two_n + 2 * b + 1];
                double mir = binv[(2 * a + 1) * two_n + 2 * b];
                double mii = binv[(2 * a + 1) * two_n + 2 * b + 1];
                double aa = 0.5 * (mrr + mii), bb = 0.5 * (mir - mri);
                h_new[(2 * a) * two_n + 2 * b] = aa;
                h_new[(2 * a) * two_n + 2 * b + 1] = -bb;
                h_new[(2 * a + 1) * two_n + 2 * b] = bb;
                h_new[(2 * a + 1) * two_n + 2 * b + 1] = aa;
            }
    } else { /* :1874-1883 identity fallback */
        memset(h_new, 0, sizeof(double) * two_n * two_n);
        for (int64_t i = 0; i < two_n; i++) h_new[i * two_n + i] = 1.0;
    }
    double diff = 0.0;
    for (int64_t i = 0; i < two_n * two_n; i++) {
        double d = h_new[i] - h_block[i];
        diff += d * d;
    }
    free(kinv);
    free(hre);
    free(him);
    free(blk);
    free(binv);
    return sqrt(diff);
}

/* ------------------------------------------------------------------------------------------ */
/* sigma-functional (quintic.rs:1301-1499 with helpers :4474-4809)                             */
/* ------------------------------------------------------------------------------------------ */
static void affine_frame(const double *g, int chart, int elim, double fr[3][10]) { /* :4541-4584 */
    memset(fr, 0, sizeof(double) * 30);
    int fre[3], idx = 0;
    for (int k = 0; k < 5; k++)
        if (k != chart && k != elim) fre[idx++] = k;
    if (idx != 3) return;
    double ger = g[2 * elim], gei = g[2 * elim + 1];
    double gn = ger * ger + gei * gei;
    if (gn < 1e-30) return;
    for (int i = 0; i < 3; i++) {
        int fi = fre[i];
        double gr = g[2 * fi], gim = g[2 * fi + 1];
        double nr = gr * ger + gim * gei;
        double ni = gim * ger - gr * gei;
        fr[i][2 * fi] = 1.0;
        fr[i][2 * fi + 1] = 0.0;
        fr[i][2 * elim] = -nr / gn;
        fr[i][2 * elim + 1] = -ni / gn;
    }
}

static void project_tangent(const double ga[5][5][2], const double fr[3][10],
                            double gt[3][3][2]) { /* :4739-4776 */
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            double sr = 0.0, si = 0.0;
            for (int i = 0; i < 5; i++) {
                double tar = fr[a][2 * i], tai = fr[a][2 * i + 1];
                double rr = 0.0, ri = 0.0;
                for (int j = 0; j < 5; j++) {
                    double gr = ga[i][j][0], gi = ga[i][j][1];
                    double tjr = fr[b][2 * j], tji = -fr[b][2 * j + 1];
                    rr += gr * tjr - gi * tji;
                    ri += gr * tji + gi * tjr;
                }
                sr += tar * rr - tai * ri;
                si += tar * ri + tai * rr;
            }
            gt[a][b][0] = sr;
            gt[a][b][1] = si;
        }
}

static double det3(const double g[3][3][2]) { /* :4782-4809 */
#define CMR(x, y) ((x)[0] * (y)[0] - (x)[1] * (y)[1])
#define CMI(x, y) ((x)[0] * (y)[1] + (x)[1] * (y)[0])
    double m1r = CMR(g[1][1], g[2][2]), m1i = CMI(g[1][1], g[2][2]);
    double m2r = CMR(g[1][2], g[2][1]), m2i = CMI(g[1][2], g[2][1]);
    double c00[2] = {m1r - m2r, m1i - m2i};
    double m3r = CMR(g[1][0], g[2][2]), m3i = CMI(g[1][0], g[2][2]);
    double m4r = CMR(g[1][2], g[2][0]), m4i = CMI(g[1][2], g[2][0]);
    double c01[2] = {m3r - m4r, m3i - m4i};
    double m5r = CMR(g[1][0], g[2][1]), m5i = CMI(g[1][0], g[2][1]);
    double m6r = CMR(g[1][1], g[2][0]), m6i = CMI(g[1][1], g[2][0]);
    double c02[2] = {m5r - m6r, m5i - m6i};
    double t1 = CMR(g[0][0], c00), t2 = CMR(g[0][1], c01), t3 = CMR(g[0][2], c02);
    return t1 - t2 + t3;
#undef CMR
#undef CMI
}

/* Per-point eta (r_per_point of :1311-1466); NaN for skipped points. */
static double eta_point(const double *s, const double *d, const double *z, double log_om,
                        const double *h, int64_t nb, double psi, double *scratch) {
    int64_t two_n = 2 * nb;
    if (!isfinite(log_om)) return NAN;
    double k = 0.0;
    for (int64_t i = 0; i < two_n; i++) {
        double rs = 0.0;
        for (int64_t j = 0; j < two_n; j++) rs += h[i * two_n + j] * s[j];
        k += s[i] * rs;
    }
    double ks = fmax(k, 1e-30);
    double *hs = scratch, *hdf = scratch + two_n; /* hdf: [5][two_n] */
    for (int64_t i = 0; i < two_n; i++) {
        double rs = 0.0;
        for (int64_t j = 0; j < two_n; j++) rs += h[i * two_n + j] * s[j];
        hs[i] = rs;
    }
    double dk[5][2];
    for (int i = 0; i < 5; i++) {
        const double *df = d + i * two_n;
        double sr = 0.0, si = 0.0;
        for (int64_t a = 0; a < nb; a++) {
            double dr = df[2 * a], di = df[2 * a + 1];
            double hr = hs[2 * a], hi = hs[2 * a + 1];
            sr += hr * dr + hi * di;
            si += hr * di - hi * dr;
        }
        dk[i][0] = sr;
        dk[i][1] = si;
    }
    for (int i = 0; i < 5; i++) {
        const double *df = d + i * two_n;
        for (int64_t kk = 0; kk < two_n; kk++) {
            double rs = 0.0;
            for (int64_t l = 0; l < two_n; l++) rs += h[kk * two_n + l] * df[l];
            hdf[i * two_n + kk] = rs;
        }
    }
    double ddk[5][5][2];
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) {
            const double *dfj = d + j * two_n;
            double sr = 0.0, si = 0.0;
            for (int64_t a = 0; a < nb; a++) {
                double jr = dfj[2 * a], ji = dfj[2 * a + 1];
                double hr = hdf[i * two_n + 2 * a], hi = hdf[i * two_n + 2 * a + 1];
                sr += jr * hr + ji * hi;
                si += jr * hi - ji * hr;
            }
            ddk[i][j][0] = sr;
            ddk[i][j][1] = si;
        }
    double ik = 1.0 / ks, ik2 = ik * ik;
    double ga[5][5][2];
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) {
            double t1r = ddk[i][j][0] * ik, t1i = ddk[i][j][1] * ik;
            double pr = dk[i][0] * dk[j][0] + dk[i][1] * dk[j][1];
            double pi = dk[i][1] * dk[j][0] - dk[i][0] * dk[j][1];
            ga[i][j][0] = t1r - pr * ik2;
            ga[i][j][1] = t1i - pi * ik2;
        }
    double g[10];
    deformed_gradient(z, psi, g);
    int chart, elim;
    chart_and_elim(z, g, &chart, &elim, NULL);
    double fr[3][10];
    affine_frame(g, chart, elim, fr);
    double gt[3][3][2];
    project_tangent(ga, fr, gt);
    double det = det3(gt);
    if (!isfinite(det) || fabs(det) < 1e-30) return NAN;
    double log_r = log(fabs(det)) - log_om;
    return exp(log_r);
}

ORC_API void orc_eta(const double *pts, const double *sv, const double *sd, const double *log_om,
                     const double *h, int64_t n, int64_t nb, double psi, double *eta) {
#pragma omp parallel
    {
        double *scr = (double *)malloc(sizeof(double) * 12 * nb);
#pragma omp for schedule(static)
        for (int64_t p = 0; p < n; p++)
            eta[p] = eta_point(sv + p * 2 * nb, sd + p * 10 * nb, pts + 10 * p, log_om[p], h, nb,
                               psi, scr);
        free(scr);
    }
}

/* Serial reduction (:1469-1498). Returns sigma_L1; also kappa and sigma_L2^2 (the
 * sigma_squared_and_gradient definition, Sum w (eta/kappa-1)^2 / Sum w). */
ORC_API double orc_sigma_reduce(const double *eta, const double *w, int64_t n, double *kappa_out,
                                double *l2sq_out) {
    double tw = 0.0, ws = 0.0;
    int64_t cnt = 0;
    for (int64_t p = 0; p < n; p++) {
        double e = eta[p], wp = w[p];
        if (!isfinite(e) || e <= 0.0 || !isfinite(wp) || wp <= 0.0) continue;
        tw += wp;
        ws += wp * e;
        cnt++;
    }
    if (kappa_out) *kappa_out = NAN;
    if (l2sq_out) *l2sq_out = NAN;
    if (cnt == 0 || tw < 1e-12) return NAN;
    double kappa = ws / tw;
    if (fabs(kappa) < 1e-30) return NAN;
    double ad = 0.0, sq = 0.0;
    for (int64_t p = 0; p < n; p++) {
        double e = eta[p], wp = w[p];
        if (!isfinite(e) || e <= 0.0 || !isfinite(wp) || wp <= 0.0) continue;
        ad += wp * fabs(e / kappa - 1.0);
        double r = e / kappa - 1.0;
        sq += wp * r * r;
    }
    if (kappa_out) *kappa_out = kappa;
    if (l2sq_out) *l2sq_out = sq / tw;
    return ad / tw;
}

ORC_API double orc_sigma(const double *pts, const double *sv, const double *sd,
                         const double *log_om, const double *w, const double *h, int64_t n,
                         int64_t nb, double psi, double *eta_out) {
    double *eta = eta_out ? eta_out : (double *)malloc(sizeof(double) * n);
    orc_eta(pts, sv, sd, log_om, h, n, nb, psi, eta);
    double s = orc_sigma_reduce(eta, w, n, NULL, NULL);
    if (!eta_out) free(eta);
    return s;
}

/* ------------------------------------------------------------------------------------------ */
/* sigma^2 and its analytic gradient (quintic.rs:2194-2860)                                    */
/* ------------------------------------------------------------------------------------------ */
/* invert_3x3_complex_hermitian (quintic.rs:2194-2245): cofactor transpose / det, then the
 * Hermitian average of each off-diagonal pair. Returns 0 when det is not usable. */
static int inv3_herm(const double g[3][3][2], double inv[3][3][2]) {
#define CM(x, y, o)                                   \
    do {                                              \
        (o)[0] = (x)[0] * (y)[0] - (x)[1] * (y)[1];   \
        (o)[1] = (x)[0] * (y)[1] + (x)[1] * (y)[0];   \
    } while (0)
    double det = det3(g);
    if (!isfinite(det) || fabs(det) < 1e-30) return 0;
    double c[3][3][2], t1[2], t2[2];
    static const int id[9][6] = {
        /* cij = g[a][b]*g[c][d] - g[e][f]*g[g][h]: (a,b,c,d) then (e,f,g,h) */
        {1, 1, 2, 2, 1, 2}, {1, 2, 2, 0, 1, 0}, {1, 0, 2, 1, 1, 1},
        {0, 2, 2, 1, 0, 1}, {0, 0, 2, 2, 0, 2}, {0, 1, 2, 0, 0, 0},
        {0, 1, 1, 2, 0, 2}, {0, 2, 1, 0, 0, 0}, {0, 0, 1, 1, 0, 1}};
    static const int id2[9][2] = {{2, 1}, {2, 2}, {2, 0}, {2, 2}, {2, 0}, {2, 1},
                                  {1, 1}, {1, 2}, {1, 0}};
    for (int q = 0; q < 9; q++) {
        CM(g[id[q][0]][id[q][1]], g[id[q][2]][id[q][3]], t1);
        CM(g[id[q][4]][id[q][5]], g[id2[q][0]][id2[q][1]], t2);
        c[q / 3][q % 3][0] = t1[0] - t2[0];
        c[q / 3][q % 3][1] = t1[1] - t2[1];
    }
    double idet = 1.0 / det;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            inv[i][j][0] = c[j][i][0] * idet;
            inv[i][j][1] = c[j][i][1] * idet;
        }
    for (int i = 0; i < 3; i++)
        for (int j = i + 1; j < 3; j++) {
            double ar = 0.5 * (inv[j][i][0] + inv[i][j][0]);
            double ai = 0.5 * (inv[j][i][1] + (-inv[i][j][1]));
            inv[j][i][0] = ar;
            inv[j][i][1] = ai;
            inv[i][j][0] = ar;
            inv[i][j][1] = -ai;
        }
    return 1;
#undef CM
}

/* per_point_log_det_gradient (quintic.rs:2311-2695), in the reference's operation order.
 * grad: [2 nb^2] (d/dh_re_ab with the diagonal halved, then d/dh_im_ab for a < b). Returns
 * eta_p (exp(log R)) or NaN for a skipped point; *ok = 0 when the point is skipped. */
static double grad_point(const double *s, const double *df, const double *h, const double *z,
                         double log_om, int64_t nb, double psi, double *grad, double *scr,
                         int *ok) {
    const int64_t two_n = 2 * nb;
    double *h_s = scr, *h_dfi = scr + two_n; /* [5][two_n] */
    double *q = h_dfi + 5 * two_n, *qp = q + two_n, *phi = qp + two_n; /* phi [5][nb][2] */
    double *amr = phi + 10 * nb, *ami = amr + nb * nb; /* scratch: 26 nb + 2 nb^2 doubles */
    *ok = 0;
    for (int64_t i = 0; i < two_n; i++) {
        double rs = 0.0;
        for (int64_t j = 0; j < two_n; j++) rs += h[i * two_n + j] * s[j];
        h_s[i] = rs;
    }
    double k_val = 0.0;
    for (int64_t i = 0; i < two_n; i++) k_val += s[i] * h_s[i];
    double k_safe = fmax(k_val, 1e-30);
    double inv_k = 1.0 / k_safe, inv_k2 = inv_k * inv_k, inv_k3 = inv_k * inv_k2;
    for (int i = 0; i < 5; i++) {
        const double *dfi = df + i * two_n;
        for (int64_t kk = 0; kk < two_n; kk++) {
            double rs = 0.0;
            for (int64_t l = 0; l < two_n; l++) rs += h[kk * two_n + l] * dfi[l];
            h_dfi[i * two_n + kk] = rs;
        }
    }
    double dk[5][2];
    for (int i = 0; i < 5; i++) {
        const double *dfi = df + i * two_n;
        double sr = 0.0, si = 0.0;
        for (int64_t a = 0; a < nb; a++) {
            double hdr = h_s[2 * a], hdi = h_s[2 * a + 1], dr = dfi[2 * a], di = dfi[2 * a + 1];
            sr += hdr * dr + hdi * di;
            si += hdr * di - hdi * dr;
        }
        dk[i][0] = sr;
        dk[i][1] = si;
    }
    double m[5][5][2];
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) {
            const double *dfj = df + j * two_n;
            double sr = 0.0, si = 0.0;
            for (int64_t a = 0; a < nb; a++) {
                double jr = dfj[2 * a], ji = dfj[2 * a + 1];
                double hr = h_dfi[i * two_n + 2 * a], hi = h_dfi[i * two_n + 2 * a + 1];
                sr += jr * hr + ji * hi;
                si += jr * hi - ji * hr;
            }
            m[i][j][0] = sr;
            m[i][j][1] = si;
        }
    double ga[5][5][2];
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) {
            double pr = dk[i][0] * dk[j][0] + dk[i][1] * dk[j][1];
            double pi = dk[i][1] * dk[j][0] - dk[i][0] * dk[j][1];
            ga[i][j][0] = m[i][j][0] * inv_k - pr * inv_k2;
            ga[i][j][1] = m[i][j][1] * inv_k - pi * inv_k2;
        }
    double gf[10];
    deformed_gradient(z, psi, gf);
    int chart, elim;
    chart_and_elim(z, gf, &chart, &elim, NULL);
    double fr[3][10];
    affine_frame(gf, chart, elim, fr);
    double gt[3][3][2];
    project_tangent(ga, fr, gt);
    double det = det3(gt);
    for (int64_t t = 0; t < 2 * nb * nb; t++) grad[t] = 0.0;
    if (!isfinite(det) || fabs(det) < 1e-30 || !isfinite(log_om)) return NAN;
    double log_r = log(fabs(det)) - log_om;
    double r_val = exp(log_r);
    double gi[3][3][2];
    if (!inv3_herm(gt, gi)) return r_val; /* zero gradient, as the reference */
    /* compute_y_5x5 (quintic.rs:2260-2301): y[i][j] = sum_ab conj(T_a[i]) ginv[a][b] T_b[j] */
    double y[5][5][2];
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) {
            double sr = 0.0, si = 0.0;
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) {
                    double gar = gi[a][b][0], gai = gi[a][b][1];
                    double tiar = fr[a][2 * i], tiai = -fr[a][2 * i + 1];
                    double tjbr = fr[b][2 * j], tjbi = fr[b][2 * j + 1];
                    double p1r = tiar * gar - tiai * gai, p1i = tiar * gai + tiai * gar;
                    sr += p1r * tjbr - p1i * tjbi;
                    si += p1r * tjbi + p1i * tjbr;
                }
            y[i][j][0] = sr;
            y[i][j][1] = si;
        }
    double tr_ym = 0.0;
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) tr_ym += y[j][i][0] * m[i][j][0] - y[j][i][1] * m[i][j][1];
    double tr_yd = 0.0;
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) {
            double pr = dk[i][0] * dk[j][0] + dk[i][1] * dk[j][1];
            double pi = dk[i][1] * dk[j][0] - dk[i][0] * dk[j][1];
            tr_yd += y[j][i][0] * pr - y[j][i][1] * pi;
        }
    double vi[5][2], wj[5][2];
    for (int i = 0; i < 5; i++) {
        double sr = 0.0, si = 0.0;
        for (int j = 0; j < 5; j++) {
            sr += y[j][i][0] * dk[j][0] + y[j][i][1] * dk[j][1];
            si += y[j][i][1] * dk[j][0] - y[j][i][0] * dk[j][1];
        }
        vi[i][0] = sr;
        vi[i][1] = si;
    }
    for (int64_t b = 0; b < nb; b++) {
        double sr = 0.0, si = 0.0;
        for (int i = 0; i < 5; i++) {
            double dr = df[i * two_n + 2 * b], di = df[i * two_n + 2 * b + 1];
            sr += dr * vi[i][0] - di * vi[i][1];
            si += dr * vi[i][1] + di * vi[i][0];
        }
        q[2 * b] = sr;
        q[2 * b + 1] = si;
    }
    for (int j = 0; j < 5; j++) {
        double sr = 0.0, si = 0.0;
        for (int i = 0; i < 5; i++) {
            sr += y[j][i][0] * dk[i][0] - y[j][i][1] * dk[i][1];
            si += y[j][i][0] * dk[i][1] + y[j][i][1] * dk[i][0];
        }
        wj[j][0] = sr;
        wj[j][1] = si;
    }
    for (int64_t b = 0; b < nb; b++) {
        double sr = 0.0, si = 0.0;
        for (int j = 0; j < 5; j++) {
            double dr = df[j * two_n + 2 * b], di = df[j * two_n + 2 * b + 1];
            sr += dr * wj[j][0] + di * wj[j][1];
            si += dr * wj[j][1] - di * wj[j][0];
        }
        qp[2 * b] = sr;
        qp[2 * b + 1] = si;
    }
    for (int j = 0; j < 5; j++)
        for (int64_t b = 0; b < nb; b++) {
            double sr = 0.0, si = 0.0;
            for (int i = 0; i < 5; i++) {
                double dr = df[i * two_n + 2 * b], di = df[i * two_n + 2 * b + 1];
                sr += y[j][i][0] * dr - y[j][i][1] * di;
                si += y[j][i][0] * di + y[j][i][1] * dr;
            }
            phi[2 * (j * nb + b)] = sr;
            phi[2 * (j * nb + b) + 1] = si;
        }
    for (int64_t a = 0; a < nb; a++)
        for (int64_t b = 0; b < nb; b++) {
            double sr = 0.0, si = 0.0;
            for (int j = 0; j < 5; j++) {
                double dr = df[j * two_n + 2 * a], di = df[j * two_n + 2 * a + 1];
                double pr = phi[2 * (j * nb + b)], pi = phi[2 * (j * nb + b) + 1];
                sr += dr * pr + di * pi;
                si += dr * pi - di * pr;
            }
            amr[a * nb + b] = sr;
            ami[a * nb + b] = si;
        }
    for (int64_t a = 0; a < nb; a++) {
        double sar = s[2 * a], sai = s[2 * a + 1], qar = q[2 * a], qai = q[2 * a + 1];
        double qpar = qp[2 * a], qpai = qp[2 * a + 1];
        for (int64_t b = 0; b < nb; b++) {
            double sbr = s[2 * b], sbi = s[2 * b + 1], qbr = q[2 * b], qbi = q[2 * b + 1];
            double qpbr = qp[2 * b], qpbi = qp[2 * b + 1];
            double re_ss = sar * sbr + sai * sbi, im_ss = sar * sbi - sai * sbr;
            double a_ab_re = amr[a * nb + b], a_ba_re = amr[b * nb + a];
            double a_ab_im = ami[a * nb + b], a_ba_im = ami[b * nb + a];
            double mixed = (sar * qbr + sai * qbi) + (sbr * qar + sbi * qai) +
                           (sar * qpbr - sai * qpbi) + (sbr * qpar - sbi * qpai);
            double im_sqab = sar * qbi - sai * qbr, im_sqba = sbr * qai - sbi * qar;
            double im_sqpab = sar * qpbi + sai * qpbr, im_sqpba = sbr * qpai + sbi * qpar;
            double pair = (a_ab_re + a_ba_re) * inv_k - 2.0 * re_ss * tr_ym * inv_k2 -
                          mixed * inv_k2 + 4.0 * re_ss * tr_yd * inv_k3;
            double dl_re = (a == b) ? 0.5 * pair : pair;
            double dl_im = 0.0;
            if (a < b)
                dl_im = -(a_ab_im - a_ba_im) * inv_k + 2.0 * im_ss * tr_ym * inv_k2 +
                        (im_sqab - im_sqba - im_sqpab + im_sqpba) * inv_k2 -
                        4.0 * im_ss * tr_yd * inv_k3;
            grad[a * nb + b] = dl_re;
            grad[nb * nb + a * nb + b] = dl_im;
        }
    }
    *ok = 1;
    return r_val;
}

/* sigma_squared_and_gradient (quintic.rs:2714-2860). Returns sigma_L2^2; grad [2 nb^2];
 * *sigma_l1 = sigma_L1. Per-point buffer n * (2 + 2 nb^2), as the reference. */
ORC_API double orc_sigma_sq_grad(const double *pts, const double *sv, const double *sd,
                                 const double *log_om, const double *w, const double *h,
                                 int64_t n, int64_t nb, double psi, double *grad,
                                 double *sigma_l1) {
    const int64_t n_dof = 2 * nb * nb, stride = 2 + n_dof;
    double *buf = (double *)malloc(sizeof(double) * (size_t)n * (size_t)stride);
#pragma omp parallel
    {
        double *scr = (double *)malloc(sizeof(double) * (size_t)(26 * nb + 2 * nb * nb + 16));
#pragma omp for schedule(static)
        for (int64_t p = 0; p < n; p++) {
            double *slot = buf + p * stride;
            double wp = w[p];
            if (!isfinite(wp) || wp <= 0.0) {
                slot[0] = NAN;
                slot[1] = 0.0;
                continue;
            }
            int ok = 0;
            double eta = grad_point(sv + p * 2 * nb, sd + p * 10 * nb, h, pts + 10 * p, log_om[p],
                                    nb, psi, slot + 2, scr, &ok);
            if (!isfinite(eta) || eta <= 0.0) {
                slot[0] = NAN;
                slot[1] = 0.0;
                continue;
            }
            slot[0] = eta;
            slot[1] = wp;
        }
        free(scr);
    }
    double tw = 0.0, swe = 0.0;
    double *swg = (double *)calloc((size_t)n_dof, sizeof(double));
    for (int64_t p = 0; p < n; p++) {
        const double *slot = buf + p * stride;
        double e = slot[0], wp = slot[1];
        if (!isfinite(e) || e <= 0.0 || wp == 0.0) continue;
        tw += wp;
        swe += wp * e;
        for (int64_t k = 0; k < n_dof; k++) swg[k] += wp * e * slot[2 + k];
    }
    for (int64_t k = 0; k < n_dof; k++) grad[k] = 0.0;
    if (sigma_l1) *sigma_l1 = NAN;
    if (tw < 1e-12) {
        free(swg);
        free(buf);
        return NAN;
    }
    double kappa = swe / tw;
    if (fabs(kappa) < 1e-30) {
        free(swg);
        free(buf);
        return NAN;
    }
    for (int64_t k = 0; k < n_dof; k++) swg[k] /= tw; /* dkappa */
    double l1 = 0.0, l2 = 0.0;
    for (int64_t p = 0; p < n; p++) {
        const double *slot = buf + p * stride;
        double e = slot[0], wp = slot[1];
        if (!isfinite(e) || e <= 0.0 || wp == 0.0) continue;
        double en = e / kappa, dev = en - 1.0;
        l1 += wp * fabs(dev);
        l2 += wp * dev * dev;
        for (int64_t k = 0; k < n_dof; k++) {
            double d_en = en * (slot[2 + k] - swg[k] / kappa);
            grad[k] += wp * 2.0 * dev * d_en;
        }
    }
    for (int64_t k = 0; k < n_dof; k++) grad[k] /= tw;
    if (sigma_l1) *sigma_l1 = l1 / tw;
    free(swg);
    free(buf);
    return l2 / tw;
}

/* One step of sigma_refine_analytic (quintic.rs:453-514) after the gradient: Adam on
 * (h_re, h_im) = (Re G, Im G) read from the h block, with the update rule written inline at
 * quintic.rs:5421-5437 (scale = lr sqrt(1-b2^t)/(1-b1^t), upd = scale m/(sqrt(v)+eps)); both
 * halves use the same t (adam_update_split, quintic.rs:2071-2117). Then the Hermitian
 * projection (:2122-2146), pack (:2148-2161) and trace renormalisation (:3169-3182).
 * h is updated in place; m, v: [2 nb^2]; returns the new t. */
ORC_API int64_t orc_adam_step(double *h, int64_t nb, const double *grad, double *m, double *v,
                              int64_t t, double lr, double beta1, double beta2, double eps) {
    const int64_t n2 = nb * nb, tn = 2 * nb;
    double *hre = (double *)malloc(sizeof(double) * (size_t)n2);
    double *him = (double *)malloc(sizeof(double) * (size_t)n2);
    for (int64_t a = 0; a < nb; a++)
        for (int64_t b = 0; b < nb; b++) {
            hre[a * nb + b] = h[(2 * a) * tn + 2 * b];
            him[a * nb + b] = -h[(2 * a) * tn + 2 * b + 1];
        }
    int64_t t1 = t + 1;
    double bc1 = 1.0 - pow(beta1, (double)t1), bc2 = 1.0 - pow(beta2, (double)t1);
    double scale = lr * (sqrt(bc2) / bc1);
    for (int half = 0; half < 2; half++) {
        double *x = half == 0 ? hre : him;
        for (int64_t k = 0; k < n2; k++) {
            int64_t g = half * n2 + k;
            m[g] = beta1 * m[g] + (1.0 - beta1) * grad[g];
            v[g] = beta2 * v[g] + (1.0 - beta2) * grad[g] * grad[g];
            x[k] -= scale * m[g] / (sqrt(v[g]) + eps);
        }
    }
    for (int64_t a = 0; a < nb; a++)
        for (int64_t b = a + 1; b < nb; b++) {
            double avg = 0.5 * (hre[a * nb + b] + hre[b * nb + a]);
            hre[a * nb + b] = avg;
            hre[b * nb + a] = avg;
            double avi = 0.5 * (him[a * nb + b] - him[b * nb + a]);
            him[a * nb + b] = avi;
            him[b * nb + a] = -avi;
        }
    for (int64_t a = 0; a < nb; a++) him[a * nb + a] = 0.0;
    for (int64_t a = 0; a < nb; a++)
        for (int64_t b = 0; b < nb; b++) {
            double hr = hre[a * nb + b], hi = him[a * nb + b];
            h[(2 * a) * tn + 2 * b] = hr;
            h[(2 * a) * tn + 2 * b + 1] = -hi;
            h[(2 * a + 1) * tn + 2 * b] = hi;
            h[(2 * a + 1) * tn + 2 * b + 1] = hr;
        }
    double tr = 0.0;
    for (int64_t a = 0; a < nb; a++) tr += h[(2 * a) * tn + 2 * a];
    if (tr > 1e-10) {
        double sc = (double)nb / tr;
        for (int64_t k = 0; k < tn * tn; k++) h[k] *= sc;
    }
    free(hre);
    free(him);
    return t1;
}

ORC_API int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
