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
static int sz_line(orc_lcg *rng, int dwork, double psi, int dk_iters, int polish,
                   double zs[5][10], int acc[5]) {
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
    /* FP64 Newton polish of each root on the monic line polynomial (t <- t - P(t)/P'(t),
     * Horner for both), `polish` passes. Not in the reference (polish = 0 there); the
     * north-star root polish after the Durand-Kerner sweeps. */
    for (int pass = 0; pass < polish; pass++) {
        for (int k = 0; k < 5; k++) {
            double tr = roots[k][0], ti = roots[k][1];
            double vr = 1.0, vi = 0.0, dvr = 0.0, dvi = 0.0;
            for (int j = 4; j >= 0; j--) {
                double ndr = dvr * tr - dvi * ti + vr, ndi = dvr * ti + dvi * tr + vi;
                double nr = vr * tr - vi * ti + p[j][0], ni = vr * ti + vi * tr + p[j][1];
                dvr = ndr;
                dvi = ndi;
                vr = nr;
                vi = ni;
            }
            double dn = dvr * dvr + dvi * dvi;
            if (dn > 1e-300) {
                roots[k][0] = tr - (vr * dvr + vi * dvi) / dn;
                roots[k][1] = ti - (vi * dvr - vr * dvi) / dn;
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
ORC_API int64_t orc_sample_sz_polish(int64_t n, uint64_t seed, int dwork, double psi,
                                     int dk_iters, int polish, double *out,
                                     int64_t *lines_tried);
ORC_API int64_t orc_sample_sz(int64_t n, uint64_t seed, int dwork, double psi, int dk_iters,
                              double *out, int64_t *lines_tried) {
    return orc_sample_sz_polish(n, seed, dwork, psi, dk_iters, 0, out, lines_tried);
}

/* The same sampler with `polish` FP64 Newton passes per root after the DK sweeps. */
ORC_API int64_t orc_sample_sz_polish(int64_t n, uint64_t seed, int dwork, double psi,
                                     int dk_iters, int polish, double *out,
                                     int64_t *lines_tried) {
    orc_lcg rng;
    lcg_init(&rng, seed);
    int64_t n_lines_target = (n + 4) / 5;
    int64_t max_tries = n_lines_target * (dwork ? 8 : 4);
    int64_t tried = 0, np = 0;
    while (np < n && tried < max_tries) {
        tried++;
        double zs[5][10];
        int acc[5];
        if (!sz_line(&rng, dwork, psi, dk_iters, polish, zs, acc)) continue;
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

/* ------------------------------------------------------------------------------------------ */
/* Newton-projection sampler (quintic.rs:3334-3482, 3568-3595)                                 */
/* ------------------------------------------------------------------------------------------ */
/* newton_project_to_quintic (:3334-3389) and newton_project_to_deformed_quintic (:3395-3443):
 * identical at psi = 0, where the deformed forms return the Fermat values. Returns 1 and the
 * projected z in place on acceptance. */
static int newton_project(double *z, double psi, double tol, int max_iter) {
    for (int it = 0; it < max_iter; it++) {
        double f_re, f_im;
        deformed_fermat_quintic(z, psi, &f_re, &f_im);
        double f_mag_sq = f_re * f_re + f_im * f_im;
        if (f_mag_sq < tol * tol) return 1;
        double grad[10];
        deformed_gradient(z, psi, grad);
        double grad_norm_sq = 0.0;
        for (int k = 0; k < 10; k++) grad_norm_sq += grad[k] * grad[k];
        if (grad_norm_sq < 1e-14) return 0; /* singular locus */
        double scale = -1.0 / grad_norm_sq;
        for (int j = 0; j < 5; j++) {
            double gr = grad[2 * j], gi = grad[2 * j + 1];
            double dr = scale * (f_re * gr + f_im * gi);
            double di = scale * (f_im * gr - f_re * gi);
            z[2 * j] += dr;
            z[2 * j + 1] += di;
        }
        double nrm_sq = 0.0;
        for (int k = 0; k < 10; k++) nrm_sq += z[k] * z[k];
        double nrm = sqrt(nrm_sq);
        if (nrm < 1e-12) return 0;
        for (int k = 0; k < 10; k++) z[k] /= nrm;
    }
    double f_re, f_im;
    deformed_fermat_quintic(z, psi, &f_re, &f_im);
    double t10 = 10.0 * tol;
    return f_re * f_re + f_im * f_im < t10 * t10;
}

ORC_API int orc_newton_project(double *z, double psi, double tol, int max_iter) {
    return newton_project(z, psi, tol, max_iter);
}

/* dwork = 0: sample_quintic_points (:3568-3595, max_tries = 4 n); dwork = 1:
 * sample_deformed_quintic_points (:3448-3482, max_tries = 8 n). Each try draws 10 normals,
 * normalises onto S^9 (a norm below 1e-12 skips the try) and keeps the projected point on
 * success. out holds n * 10 doubles; returns the number of points written. acc_out (optional,
 * max_tries ints) records each try's outcome. */
ORC_API int64_t orc_sample_newton(int64_t n, uint64_t seed, int dwork, double psi, double tol,
                                  double *out, int64_t *tries_out, int32_t *acc_out) {
    orc_lcg rng;
    lcg_init(&rng, seed);
    int64_t max_tries = n * (dwork ? 8 : 4);
    int64_t tried = 0, np = 0;
    while (np < n && tried < max_tries) {
        tried++;
        double z[10];
        for (int k = 0; k < 10; k++) z[k] = lcg_next_normal(&rng);
        double nrm_sq = 0.0;
        for (int k = 0; k < 10; k++) nrm_sq += z[k] * z[k];
        double nrm = sqrt(nrm_sq);
        int ok = 0;
        if (!(nrm < 1e-12)) {
            for (int k = 0; k < 10; k++) z[k] /= nrm;
            ok = newton_project(z, psi, tol, 50);
        }
        if (acc_out) acc_out[tried - 1] = ok;
        if (ok) {
            memcpy(out + 10 * np, z, sizeof(double) * 10);
            np++;
        }
    }
    if (tries_out) *tries_out = tried;
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
        if (!sz_line(&rng, dwork, psi, dk_iters, 0, zs, acc))
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

/* The ambient metric g_amb = ddK/K - dK dK^H / K^2 of one point (ambient_metric_5x5,
 * :4474-4516, as inlined in r_per_point :1311-1466). Returns K floored at 1e-30. */
static double ambient_ga(const double *s, const double *d, const double *h, int64_t nb,
                         double *scratch, double ga[5][5][2]) {
    int64_t two_n = 2 * nb;
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
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) {
            double t1r = ddk[i][j][0] * ik, t1i = ddk[i][j][1] * ik;
            double pr = dk[i][0] * dk[j][0] + dk[i][1] * dk[j][1];
            double pi = dk[i][1] * dk[j][0] - dk[i][0] * dk[j][1];
            ga[i][j][0] = t1r - pr * ik2;
            ga[i][j][1] = t1i - pi * ik2;
        }
    return ks;
}

/* Per-point eta (r_per_point of :1311-1466); NaN for skipped points. */
static double eta_point(const double *s, const double *d, const double *z, double log_om,
                        const double *h, int64_t nb, double psi, double *scratch) {
    if (!isfinite(log_om)) return NAN;
    double ga[5][5][2];
    ambient_ga(s, d, h, nb, scratch, ga);
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

/* cpu_compute_per_point_reference (gpu_adam.rs:2677-2739): buf [n][2 + 2 nb^2], row p =
 * (eta_p, w_p, d log eta_p / dh) from per_point_log_det_gradient; (NaN, 0, 0...) where the
 * weight or eta is invalid. */
ORC_API void orc_per_point(const double *pts, const double *sv, const double *sd,
                           const double *log_om, const double *w, const double *h, int64_t n,
                           int64_t nb, double psi, double *buf) {
    const int64_t n_dof = 2 * nb * nb, stride = 2 + n_dof;
#pragma omp parallel
    {
        double *scr = (double *)malloc(sizeof(double) * (size_t)(26 * nb + 2 * nb * nb + 16));
#pragma omp for schedule(static)
        for (int64_t p = 0; p < n; p++) {
            double *slot = buf + p * stride;
            double wp = w[p];
            for (int64_t k = 2; k < stride; k++) slot[k] = 0.0;
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
            } else {
                slot[0] = eta;
                slot[1] = wp;
            }
        }
        free(scr);
    }
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

/* ------------------------------------------------------------------------------------------ */
/* Complex helpers with num-complex 0.4's operation order (Mul, Div, norm = hypot, arg = atan2). */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
    double re, im;
} orc_c;
static inline orc_c oc(double re, double im) { orc_c z = {re, im}; return z; }
static inline orc_c c_mul(orc_c a, orc_c b) {
    return oc(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline orc_c c_add(orc_c a, orc_c b) { return oc(a.re + b.re, a.im + b.im); }
static inline orc_c c_sub(orc_c a, orc_c b) { return oc(a.re - b.re, a.im - b.im); }
static inline orc_c c_conj(orc_c a) { return oc(a.re, -a.im); }
static inline double c_nsq(orc_c a) { return a.re * a.re + a.im * a.im; }
static inline double c_norm(orc_c a) { return hypot(a.re, a.im); }
static inline orc_c c_div(orc_c a, orc_c b) {
    double ns = b.re * b.re + b.im * b.im;
    double re = a.re * b.re + a.im * b.im;
    double im = a.im * b.re - a.re * b.im;
    return oc(re / ns, im / ns);
}
static inline orc_c c_scale(orc_c a, double s) { return oc(a.re * s, a.im * s); }

/* ------------------------------------------------------------------------------------------ */
/* Metric Laplacian, ambient (Bochner) form: route34/metric_laplacian.rs                       */
/* ------------------------------------------------------------------------------------------ */
/* monomials_4 (:118-130) and build_test_basis (:87-115): every bigraded monomial of total
 * degree 1..max_deg in the 8 ambient coordinates, the constant excluded. exps: [nb][8]. */
ORC_API int64_t orc_lap_basis(uint32_t max_deg, uint32_t *exps) {
    int64_t cnt = 0;
    for (uint32_t dz = 0; dz <= max_deg; dz++)
        for (uint32_t dw = 0; dw <= max_deg - dz; dw++) {
            if (dz == 0 && dw == 0) continue;
            for (uint32_t a = 0; a <= dz; a++)
                for (uint32_t b = 0; b <= dz - a; b++)
                    for (uint32_t c = 0; c <= dz - a - b; c++) {
                        uint32_t e = dz - a - b - c;
                        for (uint32_t a2 = 0; a2 <= dw; a2++)
                            for (uint32_t b2 = 0; b2 <= dw - a2; b2++)
                                for (uint32_t c2 = 0; c2 <= dw - a2 - b2; c2++) {
                                    uint32_t e2 = dw - a2 - b2 - c2;
                                    if (exps) {
                                        uint32_t *o = exps + 8 * cnt;
                                        o[0] = a; o[1] = b; o[2] = c; o[3] = e;
                                        o[4] = a2; o[5] = b2; o[6] = c2; o[7] = e2;
                                    }
                                    cnt++;
                                }
                    }
        }
    return cnt;
}

/* eval_monomial (:132-153): square-and-multiply per coordinate. */
static orc_c eval_mono8(const double *pt, const uint32_t *e) {
    orc_c acc = oc(1.0, 0.0);
    for (int k = 0; k < 8; k++) {
        uint32_t ee = e[k];
        if (ee == 0) continue;
        orc_c base = oc(pt[2 * k], pt[2 * k + 1]);
        orc_c term = oc(1.0, 0.0);
        while (ee > 0) {
            if (ee & 1) term = c_mul(term, base);
            ee >>= 1;
            if (ee > 0) base = c_mul(base, base);
        }
        acc = c_mul(acc, term);
    }
    return acc;
}

/* build_mass_matrix (:424-471) and build_stiffness_matrix (:473-524) over the basis values
 * evaluated as in compute_metric_laplacian_spectrum (:318-340). pts: [n][8][2], omega: [n][2].
 * M, K: [nb][nb][2]. */
ORC_API void orc_lap_ambient_mk(const double *pts, const double *w, const double *omega, int64_t n,
                                const uint32_t *exps, int64_t nb, double *M, double *K) {
    orc_c *bat = (orc_c *)calloc((size_t)(nb * n), sizeof(orc_c));
    orc_c *dbat = (orc_c *)calloc((size_t)(nb * 8 * n), sizeof(orc_c));
#pragma omp parallel for schedule(static)
    for (int64_t al = 0; al < nb; al++) {
        const uint32_t *ex = exps + 8 * al;
        for (int64_t p = 0; p < n; p++) bat[al * n + p] = eval_mono8(pts + 16 * p, ex);
        for (int i = 0; i < 8; i++) {
            if (ex[i] == 0) continue; /* d_dz (:159-168) returns coefficient 0 */
            uint32_t de[8];
            memcpy(de, ex, sizeof(de));
            double coeff = (double)de[i];
            de[i] -= 1;
            orc_c cz = oc(coeff, 0.0);
            for (int64_t p = 0; p < n; p++)
                dbat[(al * 8 + i) * n + p] = c_mul(cz, eval_mono8(pts + 16 * p, de));
        }
    }
    orc_c *m = (orc_c *)calloc((size_t)(nb * nb), sizeof(orc_c));
    orc_c *k = (orc_c *)calloc((size_t)(nb * nb), sizeof(orc_c));
    double tw_m = 0.0, tw_k = 0.0;
    /* The reference loops points outermost; rows alpha are independent, so each thread owns a
     * row range and keeps the point order, giving identical sums. */
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t al = 0; al < nb; al++) {
        for (int64_t p = 0; p < n; p++) {
            double wp = w[p];
            if (!isfinite(wp) || wp <= 0.0) continue;
            orc_c om = oc(omega[2 * p], omega[2 * p + 1]);
            if (!isfinite(om.re) || !isfinite(om.im)) continue;
            orc_c pw = oc(c_nsq(om) * wp, 0.0);
            orc_c sa = bat[al * n + p];
            if (c_nsq(sa) != 0.0) {
                orc_c sac = c_conj(sa);
                for (int64_t be = al; be < nb; be++) {
                    orc_c inc = c_mul(c_mul(sac, bat[be * n + p]), pw);
                    m[al * nb + be] = c_add(m[al * nb + be], inc);
                }
            }
        }
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t al = 0; al < nb; al++) {
        for (int64_t p = 0; p < n; p++) {
            double wp = w[p];
            if (!isfinite(wp) || wp <= 0.0) continue;
            orc_c om = oc(omega[2 * p], omega[2 * p + 1]);
            if (!isfinite(om.re) || !isfinite(om.im)) continue;
            orc_c pw = oc(c_nsq(om) * wp, 0.0);
            for (int i = 0; i < 8; i++) {
                orc_c da = dbat[(al * 8 + i) * n + p];
                if (c_nsq(da) == 0.0) continue;
                orc_c dac = c_conj(da);
                for (int64_t be = al; be < nb; be++) {
                    orc_c db = dbat[(be * 8 + i) * n + p];
                    if (c_nsq(db) == 0.0) continue;
                    orc_c inc = c_mul(c_mul(dac, db), pw);
                    k[al * nb + be] = c_add(k[al * nb + be], inc);
                }
            }
        }
    }
    for (int64_t p = 0; p < n; p++) {
        double wp = w[p];
        if (!isfinite(wp) || wp <= 0.0) continue;
        if (!isfinite(omega[2 * p]) || !isfinite(omega[2 * p + 1])) continue;
        tw_m += wp;
        tw_k += wp;
    }
    orc_c *mats[2] = {m, k};
    double tws[2] = {tw_m, tw_k};
    for (int q = 0; q < 2; q++) {
        orc_c *x = mats[q];
        if (tws[q] > 0.0) {
            orc_c inv = oc(1.0 / tws[q], 0.0);
            for (int64_t e = 0; e < nb * nb; e++) x[e] = c_mul(x[e], inv);
        }
        for (int64_t i = 0; i < nb; i++) {
            orc_c z = x[i * nb + i];
            x[i * nb + i] = oc(q == 0 ? z.re : fmax(z.re, 0.0), 0.0);
            for (int64_t j = i + 1; j < nb; j++) x[j * nb + i] = c_conj(x[i * nb + j]);
        }
    }
    memcpy(M, m, sizeof(orc_c) * nb * nb);
    memcpy(K, k, sizeof(orc_c) * nb * nb);
    free(bat);
    free(dbat);
    free(m);
    free(k);
}

/* matmul_complex (:531-545). */
static void c_matmul(const orc_c *a, const orc_c *b, int64_t n, orc_c *out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = 0; j < n; j++) {
            orc_c acc = oc(0.0, 0.0);
            for (int64_t q = 0; q < n; q++) acc = c_add(acc, c_mul(a[i * n + q], b[q * n + j]));
            out[i * n + j] = acc;
        }
}

/* invert_hermitian_regularised (:547-613): Gauss-Jordan on [H + lambda I | I] with partial
 * pivoting on |.|, lambda = max(eps ||H||_F, 1e-30). */
ORC_API void orc_herm_inv_reg(const double *h_in, int64_t n, double eps, double *out_d) {
    const orc_c *h = (const orc_c *)h_in;
    orc_c *out = (orc_c *)out_d;
    if (n == 0) return;
    double fs = 0.0;
    for (int64_t e = 0; e < n * n; e++) fs += c_nsq(h[e]);
    double lambda = fmax(sqrt(fs) * eps, 1.0e-30);
    int64_t tn = 2 * n;
    orc_c *m = (orc_c *)calloc((size_t)(n * tn), sizeof(orc_c));
    for (int64_t i = 0; i < n; i++) {
        for (int64_t j = 0; j < n; j++) m[i * tn + j] = h[i * n + j];
        m[i * tn + i] = c_add(m[i * tn + i], oc(lambda, 0.0));
        m[i * tn + n + i] = oc(1.0, 0.0);
    }
    for (int64_t kk = 0; kk < n; kk++) {
        int64_t piv = kk;
        double best = c_norm(m[kk * tn + kk]);
        for (int64_t r = kk + 1; r < n; r++) {
            double v = c_norm(m[r * tn + kk]);
            if (v > best) {
                best = v;
                piv = r;
            }
        }
        if (best < 1.0e-300) {
            double inv = 1.0 / fmax(lambda, 1.0e-30);
            for (int64_t e = 0; e < n * n; e++) out[e] = oc(0.0, 0.0);
            for (int64_t i = 0; i < n; i++) out[i * n + i] = oc(inv, 0.0);
            free(m);
            return;
        }
        if (piv != kk)
            for (int64_t j = 0; j < tn; j++) {
                orc_c t = m[kk * tn + j];
                m[kk * tn + j] = m[piv * tn + j];
                m[piv * tn + j] = t;
            }
        orc_c pinv = c_div(oc(1.0, 0.0), m[kk * tn + kk]);
        for (int64_t j = 0; j < tn; j++) m[kk * tn + j] = c_mul(m[kk * tn + j], pinv);
#pragma omp parallel for schedule(static)
        for (int64_t r = 0; r < n; r++) {
            if (r == kk) continue;
            orc_c f = m[r * tn + kk];
            if (c_norm(f) == 0.0) continue;
            for (int64_t j = 0; j < tn; j++)
                m[r * tn + j] = c_sub(m[r * tn + j], c_mul(f, m[kk * tn + j]));
        }
    }
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = 0; j < n; j++) out[i * n + j] = m[i * tn + n + j];
    free(m);
}

/* hermitian_jacobi_n (:616-693): cyclic complex Jacobi; eigenvalues ascending, eigenvector
 * columns aligned. evecs may be NULL. Returns the number of sweeps run. */
ORC_API int64_t orc_herm_jacobi(const double *a_in, int64_t n, int64_t max_sweeps, double tol,
                                double *eig, double *evecs_d) {
    orc_c *a = (orc_c *)malloc(sizeof(orc_c) * n * n);
    orc_c *v = (orc_c *)calloc((size_t)(n * n), sizeof(orc_c));
    memcpy(a, a_in, sizeof(orc_c) * n * n);
    for (int64_t i = 0; i < n; i++) v[i * n + i] = oc(1.0, 0.0);
    int64_t sweeps = 0;
    for (int64_t sw = 0; sw < max_sweeps; sw++) {
        double off = 0.0;
        for (int64_t p = 0; p < n; p++)
            for (int64_t q = 0; q < n; q++)
                if (p != q) off += c_nsq(a[p * n + q]);
        if (sqrt(off) < tol) break;
        sweeps++;
        for (int64_t p = 0; p + 1 < n; p++)
            for (int64_t q = p + 1; q < n; q++) {
                orc_c apq = a[p * n + q];
                if (c_norm(apq) < tol) continue;
                double app = a[p * n + p].re, aqq = a[q * n + q].re;
                double abs_pq = c_norm(apq);
                double phi = abs_pq < 1.0e-300 ? 0.0 : atan2(apq.im, apq.re);
                double th = fabs(app - aqq) < 1.0e-300 ? 0.78539816339744830962
                                                        : 0.5 * atan(2.0 * abs_pq / (app - aqq));
                double c = cos(th), s = sin(th);
                orc_c csp = c_mul(oc(cos(phi), sin(phi)), oc(s, 0.0));
                orc_c cc = oc(c, 0.0);
                for (int64_t k2 = 0; k2 < n; k2++) {
                    orc_c akp = a[k2 * n + p], akq = a[k2 * n + q];
                    a[k2 * n + p] = c_add(c_mul(akp, cc), c_mul(akq, c_conj(csp)));
                    a[k2 * n + q] = c_sub(c_mul(akq, cc), c_mul(akp, csp));
                }
                for (int64_t k2 = 0; k2 < n; k2++) {
                    orc_c apk = a[p * n + k2], aqk = a[q * n + k2];
                    a[p * n + k2] = c_add(c_mul(apk, cc), c_mul(aqk, csp));
                    a[q * n + k2] = c_sub(c_mul(aqk, cc), c_mul(apk, c_conj(csp)));
                }
                a[p * n + q] = oc(0.0, 0.0);
                a[q * n + p] = oc(0.0, 0.0);
                a[p * n + p] = oc(a[p * n + p].re, 0.0);
                a[q * n + q] = oc(a[q * n + q].re, 0.0);
                for (int64_t k2 = 0; k2 < n; k2++) {
                    orc_c vkp = v[k2 * n + p], vkq = v[k2 * n + q];
                    v[k2 * n + p] = c_add(c_mul(vkp, cc), c_mul(vkq, c_conj(csp)));
                    v[k2 * n + q] = c_sub(c_mul(vkq, cc), c_mul(vkp, csp));
                }
            }
    }
    /* Stable ascending sort of (index, value), as Rust's sort_by. */
    int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * n);
    for (int64_t i = 0; i < n; i++) idx[i] = i;
    for (int64_t i = 1; i < n; i++) {
        int64_t t = idx[i];
        double tv = a[t * n + t].re;
        int64_t j = i;
        while (j > 0 && a[idx[j - 1] * n + idx[j - 1]].re > tv) {
            idx[j] = idx[j - 1];
            j--;
        }
        idx[j] = t;
    }
    for (int64_t k2 = 0; k2 < n; k2++) {
        eig[k2] = a[idx[k2] * n + idx[k2]].re;
        if (evecs_d) {
            orc_c *ev = (orc_c *)evecs_d;
            for (int64_t i = 0; i < n; i++) ev[i * n + k2] = v[i * n + idx[k2]];
        }
    }
    free(idx);
    free(a);
    free(v);
    return sweeps;
}

/* compute_metric_laplacian_spectrum (:296-418) after the assembly: M^-1 K with the regularised
 * inverse, Hermitian projection, Jacobi, then |lambda| sorted ascending. full: [nb]. */
ORC_API void orc_lap_spectrum_from_mk(const double *M, const double *K, int64_t nb,
                                      int64_t max_sweeps, double tol, double reg, double *full) {
    orc_c *minv = (orc_c *)malloc(sizeof(orc_c) * nb * nb);
    orc_c *l = (orc_c *)malloc(sizeof(orc_c) * nb * nb);
    orc_herm_inv_reg(M, nb, reg, (double *)minv);
    c_matmul(minv, (const orc_c *)K, nb, l);
    for (int64_t i = 0; i < nb; i++) {
        l[i * nb + i] = oc(l[i * nb + i].re, 0.0);
        for (int64_t j = i + 1; j < nb; j++) {
            orc_c avg = c_scale(c_add(l[i * nb + j], c_conj(l[j * nb + i])), 0.5);
            l[i * nb + j] = avg;
            l[j * nb + i] = c_conj(avg);
        }
    }
    double *ev = (double *)malloc(sizeof(double) * nb);
    orc_herm_jacobi((const double *)l, nb, max_sweeps, tol, ev, NULL);
    for (int64_t i = 0; i < nb; i++) ev[i] = fabs(ev[i]);
    /* stable sort by |lambda| */
    for (int64_t i = 1; i < nb; i++) {
        double t = ev[i];
        int64_t j = i;
        while (j > 0 && ev[j - 1] > t) {
            ev[j] = ev[j - 1];
            j--;
        }
        ev[j] = t;
    }
    memcpy(full, ev, sizeof(double) * nb);
    free(ev);
    free(minv);
    free(l);
}

/* ------------------------------------------------------------------------------------------ */
/* Metric Laplacian on the quintic with the intrinsic inverse metric g^{a b-bar} of the         */
/* balanced h, on phi_ab = s_a conj(s_b) / |z|^(2d) with s the raw degree-d monomials.          */
/* The reference states the ambient form only (metric_laplacian.rs:28-41 names the intrinsic    */
/* form as the missing refinement); the measure (w |Omega|^2 / sum w) and the Hermitian pencil  */
/* K v = lambda M v follow metric_laplacian.rs:424-524.                                          */
/*   M[ab][a'b'] = sum_p c_p conj(phi_ab) phi_a'b' / sum w                                       */
/*   K[ab][a'b'] = sum_p c_p sum_ij conj(F^ab_j) Y_ij F^a'b'_i / sum w                            */
/* with c_p = w_p |Omega_p|^2, F_i = d phi / d conj(z_i), Y = T^H g_tan^-1 T the pulled-back      */
/* inverse metric (compute_y_5x5, quintic.rs:2260-2301). Index alpha = a ns + b.                */
/* ------------------------------------------------------------------------------------------ */
ORC_API double orc_lap_quintic_mk(const double *pts, const double *w, const double *log_om,
                                  const double *sv, const double *sd, const double *h, int64_t n,
                                  int64_t nb, double psi, const uint32_t *smono, int64_t ns,
                                  int sdeg, double *M, double *K) {
    const int64_t B = ns * ns;
    int nth = 1;
#ifdef _OPENMP
    nth = omp_get_max_threads();
#endif
    orc_c *pm = (orc_c *)calloc((size_t)(nth * B * B), sizeof(orc_c));
    orc_c *pk = (orc_c *)calloc((size_t)(nth * B * B), sizeof(orc_c));
    double *ptw = (double *)calloc((size_t)nth, sizeof(double));
#pragma omp parallel
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        double *scr = (double *)malloc(sizeof(double) * 12 * nb);
        orc_c *phi = (orc_c *)malloc(sizeof(orc_c) * B);
        orc_c *F = (orc_c *)malloc(sizeof(orc_c) * B * 5);
        orc_c *YF = (orc_c *)malloc(sizeof(orc_c) * B * 5);
        orc_c *sm = (orc_c *)malloc(sizeof(orc_c) * ns);
        orc_c *G = (orc_c *)malloc(sizeof(orc_c) * ns * 5);
        orc_c *mm = pm + (int64_t)t * B * B, *kk = pk + (int64_t)t * B * B;
#pragma omp for schedule(static)
        for (int64_t p = 0; p < n; p++) {
            double wp = w[p], lo = log_om[p];
            if (!isfinite(wp) || wp <= 0.0 || !isfinite(lo)) continue;
            const double *z = pts + 10 * p;
            double ga[5][5][2];
            ambient_ga(sv + p * 2 * nb, sd + p * 10 * nb, h, nb, scr, ga);
            double g[10];
            deformed_gradient(z, psi, g);
            int chart, elim;
            chart_and_elim(z, g, &chart, &elim, NULL);
            double fr[3][10];
            affine_frame(g, chart, elim, fr);
            double gt[3][3][2], gi[3][3][2];
            project_tangent(ga, fr, gt);
            if (!inv3_herm(gt, gi)) continue;
            orc_c Y[5][5];
            for (int i = 0; i < 5; i++)
                for (int j = 0; j < 5; j++) {
                    orc_c acc = oc(0.0, 0.0);
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) {
                            orc_c ta = oc(fr[a][2 * i], -fr[a][2 * i + 1]);
                            orc_c tb = oc(fr[b][2 * j], fr[b][2 * j + 1]);
                            acc = c_add(acc, c_mul(c_mul(ta, oc(gi[a][b][0], gi[a][b][1])), tb));
                        }
                    Y[i][j] = acc;
                }
            double c = wp * exp(lo);
            double r2 = 0.0;
            for (int i = 0; i < 5; i++) r2 += z[2 * i] * z[2 * i] + z[2 * i + 1] * z[2 * i + 1];
            double rd = pow(r2, (double)sdeg);
            for (int64_t b = 0; b < ns; b++) {
                const uint32_t *e = smono + 5 * b;
                orc_c v = oc(1.0, 0.0);
                for (int i = 0; i < 5; i++)
                    for (uint32_t q = 0; q < e[i]; q++) v = c_mul(v, oc(z[2 * i], z[2 * i + 1]));
                sm[b] = v;
                for (int i = 0; i < 5; i++) {
                    orc_c dv = oc(0.0, 0.0);
                    if (e[i] > 0) {
                        dv = oc((double)e[i], 0.0);
                        for (int j = 0; j < 5; j++) {
                            uint32_t ej = e[j] - (j == i ? 1u : 0u);
                            for (uint32_t q = 0; q < ej; q++)
                                dv = c_mul(dv, oc(z[2 * j], z[2 * j + 1]));
                        }
                    }
                    orc_c corr = c_scale(c_mul(v, oc(z[2 * i], -z[2 * i + 1])), (double)sdeg / r2);
                    G[b * 5 + i] = c_sub(dv, corr);
                }
            }
            for (int64_t a = 0; a < ns; a++) {
                orc_c u = c_scale(sm[a], 1.0 / rd);
                for (int64_t b = 0; b < ns; b++) {
                    int64_t al = a * ns + b;
                    phi[al] = c_mul(u, c_conj(sm[b]));
                    for (int i = 0; i < 5; i++) F[al * 5 + i] = c_mul(u, c_conj(G[b * 5 + i]));
                }
            }
            for (int64_t be = 0; be < B; be++)
                for (int j = 0; j < 5; j++) {
                    orc_c acc = oc(0.0, 0.0);
                    for (int i = 0; i < 5; i++) acc = c_add(acc, c_mul(Y[i][j], F[be * 5 + i]));
                    YF[be * 5 + j] = acc;
                }
            for (int64_t al = 0; al < B; al++) {
                orc_c pa = c_scale(c_conj(phi[al]), c);
                orc_c fa[5];
                for (int j = 0; j < 5; j++) fa[j] = c_scale(c_conj(F[al * 5 + j]), c);
                orc_c *mrow = mm + al * B, *krow = kk + al * B;
                for (int64_t be = 0; be < B; be++) {
                    mrow[be] = c_add(mrow[be], c_mul(pa, phi[be]));
                    orc_c acc = oc(0.0, 0.0);
                    for (int j = 0; j < 5; j++) acc = c_add(acc, c_mul(fa[j], YF[be * 5 + j]));
                    krow[be] = c_add(krow[be], acc);
                }
            }
            ptw[t] += wp;
        }
        free(scr);
        free(phi);
        free(F);
        free(YF);
        free(sm);
        free(G);
    }
    double tw = 0.0;
    for (int t = 0; t < nth; t++) tw += ptw[t];
    orc_c *Mo = (orc_c *)M, *Ko = (orc_c *)K;
    for (int64_t e = 0; e < B * B; e++) {
        orc_c sm_ = oc(0.0, 0.0), sk = oc(0.0, 0.0);
        for (int t = 0; t < nth; t++) {
            sm_ = c_add(sm_, pm[(int64_t)t * B * B + e]);
            sk = c_add(sk, pk[(int64_t)t * B * B + e]);
        }
        Mo[e] = tw > 0.0 ? c_scale(sm_, 1.0 / tw) : sm_;
        Ko[e] = tw > 0.0 ? c_scale(sk, 1.0 / tw) : sk;
    }
    free(pm);
    free(pk);
    free(ptw);
    return tw;
}

/* ------------------------------------------------------------------------------------------ */
/* Yukawa triple overlaps: route34/yukawa_overlap_real.rs                                      */
/* ------------------------------------------------------------------------------------------ */
/* sz_weights (:222-268): w / |Omega|^2 renormalised to sum one; returns the uniformity score. */
ORC_API double orc_yuk_sz_weights(const double *w, const double *omega, int64_t n, double *sz) {
    double total = 0.0;
    int64_t bad = 0;
    for (int64_t a = 0; a < n; a++) sz[a] = 0.0;
    for (int64_t a = 0; a < n; a++) {
        double wa = w[a];
        if (!isfinite(wa) || wa <= 0.0) {
            bad++;
            continue;
        }
        double o2 = c_nsq(oc(omega[2 * a], omega[2 * a + 1]));
        if (!isfinite(o2) || o2 <= 0.0) {
            bad++;
            continue;
        }
        double v = wa / o2;
        sz[a] = v;
        total += v;
    }
    if (total > 0.0) {
        double inv = 1.0 / total;
        for (int64_t a = 0; a < n; a++) sz[a] *= inv;
    }
    double n_eff = (double)(n - bad);
    if (n_eff <= 1.0) return 0.0;
    double mean = 0.0;
    for (int64_t a = 0; a < n; a++) mean += sz[a];
    mean /= (double)n;
    double var = 0.0;
    for (int64_t a = 0; a < n; a++) var += (sz[a] - mean) * (sz[a] - mean);
    var /= (double)n;
    double cv = mean > 0.0 ? sqrt(var) / mean : INFINITY;
    double score = 1.0 / (1.0 + cv);
    return score < 0.0 ? 0.0 : (score > 1.0 ? 1.0 : score);
}

/* triple_overlap_cpu (:271-315) when idx == NULL (check_psi = 1), bootstrap_one (:318-357) over
 * the given resample indices otherwise (no psi finiteness check, as the reference).
 * modes: [nm][n][2], y: [nm][nm][nm][2]. */
ORC_API void orc_yuk_triple(const double *modes, const double *omega, const double *weights,
                            const double *hdiag, int64_t nm, int64_t n, const int64_t *idx,
                            int64_t n_idx, double *y) {
    const orc_c *md = (const orc_c *)modes;
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
    for (int64_t i = 0; i < nm; i++)
        for (int64_t j = 0; j < nm; j++)
            for (int64_t k = 0; k < nm; k++) {
                orc_c acc = oc(0.0, 0.0);
                int64_t cnt = idx ? n_idx : n;
                for (int64_t q = 0; q < cnt; q++) {
                    int64_t p = idx ? idx[q] : q;
                    double wv = weights[p];
                    if (!isfinite(wv)) continue;
                    orc_c om = oc(omega[2 * p], omega[2 * p + 1]);
                    if (!isfinite(om.re) || !isfinite(om.im)) continue;
                    orc_c pi_ = md[i * n + p], pj = md[j * n + p], pk = md[k * n + p];
                    if (!idx && (!isfinite(pi_.re) || !isfinite(pj.re) || !isfinite(pk.re)))
                        continue;
                    acc = c_add(acc, c_mul(c_mul(c_mul(c_mul(c_conj(om), pi_), pj), pk),
                                           oc(wv, 0.0)));
                }
                double sc = sqrt(hdiag ? hdiag[i] : 1.0) * sqrt(hdiag ? hdiag[j] : 1.0) *
                            sqrt(hdiag ? hdiag[k] : 1.0);
                orc_c r = c_mul(acc, oc(sc, 0.0));
                int64_t o = (i * nm + j) * nm + k;
                y[2 * o] = r.re;
                y[2 * o + 1] = r.im;
            }
}

/* Bootstrap resample indices: n draws of floor(u n) from the lib.rs LCG seeded with
 * seed + b * 2654435761 (wrapping). The reference draws them from ChaCha20Rng::random_range
 * (yukawa_overlap_real.rs:327-331), a third-party stream that is not restated here. */
ORC_API void orc_yuk_bootstrap_indices(uint64_t seed, int64_t b, int64_t n, int64_t *idx) {
    orc_lcg r;
    lcg_init(&r, seed + (uint64_t)b * 2654435761ull);
    for (int64_t q = 0; q < n; q++) {
        int64_t p = (int64_t)(lcg_next_f64(&r) * (double)n);
        idx[q] = p >= n ? n - 1 : p;
    }
}

/* compute_yukawa_couplings (:372-525) without the HYM projection (h_v_diag given): SZ weights,
 * central tensor, bootstrap standard deviation and the N vs N/2 convergence ratio. */
ORC_API void orc_yuk_couplings(const double *modes, const double *omega, const double *w,
                               const double *hdiag, int64_t nm, int64_t n, int64_t n_bs,
                               uint64_t seed, double *y, double *unc, double *score_out,
                               double *conv_out) {
    double *sz = (double *)malloc(sizeof(double) * n);
    double score = orc_yuk_sz_weights(w, omega, n, sz);
    orc_yuk_triple(modes, omega, sz, hdiag, nm, n, NULL, 0, y);
    int64_t n3 = nm * nm * nm;
    int64_t nb_ = n_bs < 1 ? 1 : n_bs;
    double *acc = (double *)calloc((size_t)n3, sizeof(double));
    double *yb = (double *)malloc(sizeof(double) * 2 * n3);
    int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * n);
    for (int64_t b = 0; b < nb_; b++) {
        orc_yuk_bootstrap_indices(seed, b, n, idx);
        orc_yuk_triple(modes, omega, sz, hdiag, nm, n, idx, n, yb);
        for (int64_t e = 0; e < n3; e++) {
            double dr = yb[2 * e] - y[2 * e], di = yb[2 * e + 1] - y[2 * e + 1];
            acc[e] += dr * dr + di * di;
        }
    }
    for (int64_t e = 0; e < n3; e++) unc[e] = sqrt(acc[e] / (double)nb_);
    int64_t half = n / 2;
    double conv = NAN;
    if (half >= 8 && nm >= 1) {
        double *hm = (double *)malloc(sizeof(double) * 2 * nm * half);
        for (int64_t i = 0; i < nm; i++)
            memcpy(hm + 2 * i * half, modes + 2 * i * n, sizeof(double) * 2 * half);
        double tot = 0.0;
        for (int64_t p = 0; p < half; p++) tot += sz[p];
        double *hw = (double *)malloc(sizeof(double) * half);
        for (int64_t p = 0; p < half; p++) hw[p] = tot > 0.0 ? sz[p] / tot : sz[p];
        orc_yuk_triple(hm, omega, hw, hdiag, nm, half, NULL, 0, yb);
        double md = 0.0, mf = 0.0;
        for (int64_t e = 0; e < n3; e++) {
            double dr = y[2 * e] - yb[2 * e], di = y[2 * e + 1] - yb[2 * e + 1];
            double d = sqrt(dr * dr + di * di);
            if (d > md) md = d;
            double f = hypot(y[2 * e], y[2 * e + 1]);
            if (f > mf) mf = f;
        }
        conv = mf > 0.0 ? md / mf : 0.0;
        free(hm);
        free(hw);
    }
    *score_out = score;
    *conv_out = conv;
    free(sz);
    free(acc);
    free(yb);
    free(idx);
}

/* Per-point ambient metric and affine frame (for the independent numpy checks of the
 * Laplacian stiffness). ga: [n][5][5][2], fr: [n][3][10], ok: [n] (1 when the tangent metric
 * is invertible and the point is kept). */
ORC_API void orc_point_metric(const double *pts, const double *sv, const double *sd,
                              const double *h, int64_t n, int64_t nb, double psi, double *ga_out,
                              double *fr_out, int32_t *ok) {
#pragma omp parallel
    {
        double *scr = (double *)malloc(sizeof(double) * 12 * nb);
#pragma omp for schedule(static)
        for (int64_t p = 0; p < n; p++) {
            const double *z = pts + 10 * p;
            double ga[5][5][2];
            ambient_ga(sv + p * 2 * nb, sd + p * 10 * nb, h, nb, scr, ga);
            double g[10];
            deformed_gradient(z, psi, g);
            int chart, elim;
            chart_and_elim(z, g, &chart, &elim, NULL);
            double fr[3][10];
            affine_frame(g, chart, elim, fr);
            double gt[3][3][2], gi[3][3][2];
            project_tangent(ga, fr, gt);
            ok[p] = inv3_herm(gt, gi);
            memcpy(ga_out + 50 * p, ga, sizeof(ga));
            memcpy(fr_out + 30 * p, fr, sizeof(fr));
        }
        free(scr);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Legacy un-inverted Donaldson iteration: donaldson_iteration_quintic (quintic.rs:4140-4222)   */
/* with w == NULL, donaldson_iteration_quintic_weighted (:4256-4331) otherwise. h_new: [2N][2N]. */
/* ------------------------------------------------------------------------------------------ */
ORC_API void orc_legacy_step(const double *sv, const double *w, const double *h, int64_t n,
                             int64_t nb, double *h_new) {
    int64_t two_n = 2 * nb;
    double *inv_k = (double *)malloc(sizeof(double) * n);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; p++) {
        const double *s = sv + p * two_n;
        double k = 0.0;
        for (int64_t i = 0; i < two_n; i++) {
            double rs = 0.0;
            for (int64_t j = 0; j < two_n; j++) rs += h[i * two_n + j] * s[j];
            k += s[i] * rs;
        }
        inv_k[p] = 1.0 / fmax(k, 1e-30);
    }
    double *hre = (double *)calloc((size_t)(nb * nb), sizeof(double));
    double *him = (double *)calloc((size_t)(nb * nb), sizeof(double));
    double total_w = 0.0;
    for (int64_t p = 0; p < n; p++) {
        const double *s = sv + p * two_n;
        double ik = inv_k[p];
        if (w) {
            double wp = w[p];
            if (!isfinite(wp) || wp <= 0.0) continue;
            total_w += wp;
            ik = inv_k[p] * wp;
        }
        for (int64_t a = 0; a < nb; a++) {
            double sar = s[2 * a], sai = s[2 * a + 1];
            for (int64_t b = 0; b < nb; b++) {
                double sbr = s[2 * b], sbi = s[2 * b + 1];
                hre[a * nb + b] += (sar * sbr + sai * sbi) * ik;
                him[a * nb + b] += (sar * sbi - sai * sbr) * ik;
            }
        }
    }
    double sc0 = w ? (total_w > 1e-12 ? 1.0 / total_w : 0.0) : 1.0 / (double)n;
    for (int64_t e = 0; e < nb * nb; e++) {
        hre[e] *= sc0;
        him[e] *= sc0;
    }
    double tr = 0.0;
    for (int64_t a = 0; a < nb; a++) tr += hre[a * nb + a];
    if (tr > 1e-10) {
        double sc = (double)nb / tr;
        for (int64_t e = 0; e < nb * nb; e++) {
            hre[e] *= sc;
            him[e] *= sc;
        }
    }
    for (int64_t a = 0; a < nb; a++)
        for (int64_t b = 0; b < nb; b++) {
            double r = hre[a * nb + b], i = him[a * nb + b];
            h_new[(2 * a) * two_n + 2 * b] = r;
            h_new[(2 * a) * two_n + 2 * b + 1] = -i;
            h_new[(2 * a + 1) * two_n + 2 * b] = i;
            h_new[(2 * a + 1) * two_n + 2 * b + 1] = r;
        }
    free(inv_k);
    free(hre);
    free(him);
}
