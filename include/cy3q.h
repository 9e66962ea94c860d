/*
 * cy3q: MI355X-native quintic Calabi–Yau metric hot path. C ABI.
 *
 * The library (papers_amd/libcy3q.so) replaces the reference crate's QuinticSolver path
 * (gable-digital/papers, substrate_sm_derivation/cy3_solver, cited below as cy3/). That path
 * covers the SZ sampler, the section basis, FS-Gram orthonormalisation, the Donaldson
 * T-operator, the sigma-functional and its analytic gradient with Adam. Every compute stage
 * runs in hand-written HIP kernels for gfx950 on one GPU. The host side only orchestrates.
 *
 * Conventions (the same as the reference):
 *   - points: [n][10] doubles, (re, im) of z_0..z_4, normalised to |z| = 1;
 *   - section values: [n][2N], interleaved (re, im) per basis element;
 *   - section derivatives: [n][5][2N];
 *   - h_block: the 2N x 2N real block of the upper-index Hermitian G, with
 *     H[2a][2b] = Re G_ab, H[2a][2b+1] = -Im G_ab, H[2a+1][2b] = Im G_ab,
 *     H[2a+1][2b+1] = Re G_ab (cy3/src/quintic.rs:2148-2161).
 *
 * Every function returns CY3Q_OK (0) or a negative status. cy3q_last_error() gives the
 * message of the calling thread's last failure. Handles are thread-compatible: use one handle
 * from one thread at a time. Each handle owns one HIP stream on its device.
 */
#ifndef CY3Q_H
#define CY3Q_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CY3Q_OK 0
#define CY3Q_ERR_ARG (-1)    /* bad argument or unsupported option */
#define CY3Q_ERR_HIP (-2)    /* HIP runtime failure (allocation, launch) */
#define CY3Q_ERR_NOT_PD (-3) /* FS Gram not positive definite (quintic.rs:3148-3149) */
#define CY3Q_ERR_EMPTY (-4)  /* no point sampled: the reference returns None (quintic.rs:203-205) */
#define CY3Q_ERR_STATE (-5)  /* call out of order */

/* Sampler kinds (cy3/src/quintic.rs:3613-3616). */
#define CY3Q_SAMPLER_NEWTON 0       /* sample_quintic_points, quintic.rs:3568-3595 */
#define CY3Q_SAMPLER_SZ 1           /* sample_quintic_points_sz, quintic.rs:3633-3785 */
#define CY3Q_SAMPLER_SZ_DWORK 2     /* sample_deformed_quintic_points_sz, quintic.rs:3810-3983 */
#define CY3Q_SAMPLER_NEWTON_DWORK 3 /* sample_deformed_quintic_points, quintic.rs:3448-3482 */

/* Arithmetic of the T-operator contractions. */
#define CY3Q_PREC_FP64 0    /* plain FP64 MFMA */
#define CY3Q_PREC_SPLIT16 1 /* fp16 hi/lo split MFMA, FP32 accumulate, FP64 across chunks */

typedef struct cy3q_options {
  int quotient;      /* 0: raw C(k+4,4) basis (quintic.rs:3192-3209); 1: drop e0 >= 5, the
                        leading-monomial ideal <z0^5> of the quintic (route34/ty_metric.rs:873-893) */
  int precision;     /* CY3Q_PREC_* for the Donaldson contractions */
  int device;        /* HIP device ordinal */
  int dk_iters;      /* Durand–Kerner sweeps per line (reference: 50) */
  int newton_polish; /* FP64 Newton steps per root after DK (reference: 0) */
  double newton_tol; /* tolerance of the Newton-projection samplers (the reference's newton_tol
                        argument; default 1e-12) */
} cy3q_options;

typedef struct cy3q_solver cy3q_solver;

/* Fills the defaults: raw basis, FP64, device 0, 50 DK sweeps, no polish, newton_tol 1e-12. */
void cy3q_options_default(cy3q_options *opt);
/* Library version, the clang/hipcc version that built it and the GPU target (recorded in the
 * run manifest; the reference's build.rs records RUSTC_VERSION for it, cy3/build.rs). */
const char *cy3q_version(void);
const char *cy3q_last_error(void);

/* Index tables. n_sections_quintic (quintic.rs:3185-3188) and
 * build_degree_k_quintic_monomials (quintic.rs:3192-3209). out holds cap rows of 5 exponents. */
int cy3q_n_sections(uint32_t k, int quotient, int64_t *n);
int cy3q_monomials(uint32_t k, int quotient, uint32_t *out, int64_t cap, int64_t *n);

/* The QuinticSolver constructors (quintic.rs:174-322), by sampler:
 *   CY3Q_SAMPLER_NEWTON        QuinticSolver::new(k, n, seed, tol), equivalently
 *                              new_with_sampler(.., NewtonProjection) (psi must be 0);
 *   CY3Q_SAMPLER_SZ            new_with_sampler(k, n, seed, tol, ShiffmanZelditch) (psi 0);
 *   CY3Q_SAMPLER_SZ_DWORK      new_with_psi_sz(k, n, seed, psi);
 *   CY3Q_SAMPLER_NEWTON_DWORK  new_with_psi(k, n, seed, tol, psi).
 * tol is opt->newton_tol. Samples, weights, log|Omega|^2 and sections are computed on the GPU.
 * h starts at the identity. */
int cy3q_solver_new(uint32_t k, int64_t n_points, uint64_t seed, double psi, int sampler,
                    const cy3q_options *opt, cy3q_solver **out);
/* Same workspace built from caller-supplied points ([n][10], host memory). */
int cy3q_solver_new_from_points(uint32_t k, const double *points, int64_t n, double psi,
                                const cy3q_options *opt, cy3q_solver **out);
void cy3q_solver_free(cy3q_solver *s);

int cy3q_solver_info(const cy3q_solver *s, int64_t *n_points, int64_t *n_basis, uint32_t *k,
                     double *psi);
int cy3q_solver_set_precision(cy3q_solver *s, int precision);

/* QuinticSolver::orthonormalise_basis_fs_gram (quintic.rs:3140-3165). */
int cy3q_solver_orthonormalise_fs_gram(cy3q_solver *s);
/* QuinticSolver::donaldson_step (quintic.rs:427-430, 1734-1894): one T-operator iteration.
 * residual = ||H_new - H_old||_F. */
int cy3q_solver_donaldson_step(cy3q_solver *s, double *residual);
/* QuinticSolver::donaldson_solve (quintic.rs:433-443). With tol <= 0 no early stop is taken
 * and the iterations run back to back with no host synchronisation. residuals: max_iter
 * entries or NULL. In CY3Q_PREC_SPLIT16 the last min(refine, max_iter) iterations are
 * refinement sweeps (cy3q_solver_set_refine, _set_refine_kind); with tol > 0 a fast iteration below tol
 * ends the fast phase early, the refinement sweeps follow, and the solve stops at the first
 * sweep below tol or after `refine` sweeps. */
int cy3q_solver_donaldson_solve(cy3q_solver *s, int64_t max_iter, double tol, int64_t *iters,
                                double *residuals);
/* The legacy un-inverted iteration h <- T(h), trace-normalised, in FP64 on the solver's current
 * sections: donaldson_iteration_quintic (quintic.rs:4140-4222) when weighted = 0 (all points,
 * 1 / n_points), donaldson_iteration_quintic_weighted (:4256-4331) when weighted = 1 (valid
 * weights, 1 / sum w). residual = ||H_new - H_old||_F. The solve starts from h = I, as
 * donaldson_solve_quintic(_weighted) (:4224-4248, :4333-4357), and stops at residual < tol. */
int cy3q_solver_donaldson_step_legacy(cy3q_solver *s, int weighted, double *residual);
int cy3q_solver_donaldson_solve_legacy(cy3q_solver *s, int weighted, int64_t max_iter, double tol,
                                       int64_t *iters, double *residuals);
/* Number of refinement sweeps that end every split-precision donaldson_solve (default 1; 0
 * turns refinement off). No counterpart in the reference, which has only the FP64 path. A sweep
 * re-applies the T-operator to the fast path's metric with the T pass in FP64 (and, for
 * CY3Q_REFINE_STEP_FP64, the D pass too), so the returned h and residual are those of that
 * iteration. */
int cy3q_solver_set_refine(cy3q_solver *s, int64_t sweeps);

/* What one refinement sweep is. CY3Q_REFINE_T_FP64 (the default): the split D pass, then the
 * FP64 T pass and the FP64 N x N solve (a final FP64 T-iteration correction: the D pass's
 * rounding is independent per point and averages out in T). CY3Q_REFINE_STEP_FP64: a whole
 * FP64 Donaldson step. No counterpart in the reference. */
#define CY3Q_REFINE_STEP_FP64 0
#define CY3Q_REFINE_T_FP64 1
int cy3q_solver_set_refine_kind(cy3q_solver *s, int kind);
/* Tile parameters (the autotuner's knobs; no counterpart in the reference): point chunks per
 * CU of the split-precision T pass (1..16) and the target workgroup count of the FP64 rank-k
 * T pass (64..8192). 0 keeps a value. Results change only by summation order. */
int cy3q_solver_set_tuning(cy3q_solver *s, int64_t tpass_wgs_per_cu, int64_t syrk_target_wgs);
int cy3q_solver_get_tuning(cy3q_solver *s, int64_t *tpass_wgs_per_cu, int64_t *syrk_target_wgs);
/* QuinticSolver::sigma (quintic.rs:423-425, 1301-1499). */
int cy3q_solver_sigma(cy3q_solver *s, double *sigma);
/* sigma with its side outputs: kappa, sigma_L2^2 = sum w (eta/kappa-1)^2 / sum w, and the
 * per-point eta ([n], NaN where skipped) when eta != NULL. */
int cy3q_solver_sigma_ext(cy3q_solver *s, double *sigma_l1, double *kappa, double *sigma_l2sq,
                          double *eta);

/* sigma_at_metric_closure and sigma_and_etas_at_metric_closure (quintic.rs:1520-1700) with the
 * solver's Bergman metric g = ddbar log(s^H G s) of its current G as the metric callback: a fresh
 * sample of n_points points (sampler and seed as the reference's sample_quintic_points_with,
 * quintic.rs:4006-4016; the Dwork samplers use the solver's psi, the Fermat samplers need
 * psi = 0), the CY weights, the sections of the solver's basis at those points carried through
 * the FS-Gram basis changes already applied to the solver, then sigma as
 * compute_sigma_from_workspace (quintic.rs:1301-1499). The solver's own points and state are
 * not touched. *n_actual: points sampled; sigma_l1, kappa and sigma_l2sq are NaN for an empty
 * sample. eta ([n_points], NaN where skipped) and points ([n_points][10]) may be NULL. */
int cy3q_solver_sigma_at_sample(cy3q_solver *s, int64_t n_points, uint64_t seed, int sampler,
                                double *sigma_l1, double *kappa, double *sigma_l2sq,
                                int64_t *n_actual, double *eta, double *points);

/* sigma_squared_and_gradient (quintic.rs:2714-2860): sigma_L2^2 = sum w (eta/kappa-1)^2 / sum w,
 * its gradient with respect to h and sigma_L1. grad: 2 N^2 doubles or NULL; grad[a N + b] =
 * d/dh_re(a, b) (the diagonal halved, as the reference), grad[N^2 + a N + b] = d/dh_im(a, b) for
 * a < b (0 otherwise), where h_re = Re G and h_im = Im G. The per-point metric is recomputed on
 * the GPU (no per-point gradient is stored). */
int cy3q_solver_sigma_sq_grad(cy3q_solver *s, double *sigma_l2sq, double *grad,
                              double *sigma_l1);
/* QuinticSolver::sigma_refine_analytic (quintic.rs:453-514): n_iter Adam steps on (h_re, h_im)
 * driven by the analytic gradient, each followed by the Hermitian projection and the trace
 * renormalisation; the moments start at zero. history: n_iter doubles (sigma_L1 before each
 * step) or NULL; *n_done: steps taken (the loop stops at a non-finite sigma). Default Adam
 * parameters are the reference's (quintic.rs:386-388); cy3q_solver_set_adam changes them. */
int cy3q_solver_sigma_refine_analytic(cy3q_solver *s, int64_t n_iter, double lr,
                                      double *history, int64_t *n_done);
int cy3q_solver_set_adam(cy3q_solver *s, double beta1, double beta2, double eps);
/* QuinticSolver::sigma_refine_lbfgs_gpu (quintic.rs:678-880, 901-1094): L-BFGS on sigma_L2^2
 * over (h_re, h_im) with the m_history newest curvature pairs (1..63), a backtracking Armijo
 * line search (c1 = 1e-4, at most 30 halvings, first trial step 1, or 1/|g| for a steepest-
 * descent step), a steepest-descent retry when the line search fails, and the Hermitian
 * projection plus trace renormalisation of apply_lbfgs_step (:2003-2048) at every trial point.
 * The point with the lowest sigma_L1 is kept at the end. history: n_iter + 1 doubles (sigma_L1
 * at the start and after each accepted step) or NULL; *n_done: entries written. A trial point
 * whose G is not positive definite is rejected by the line search. */
int cy3q_solver_sigma_refine_lbfgs(cy3q_solver *s, int64_t n_iter, int64_t m_history,
                                   double *history, int64_t *n_done);
/* QuinticSolver::sigma_refine_analytic_with_restarts (quintic.rs:545-582) and its GPU form
 * (:1145-1182): Adam from the current G, then n_restarts more runs, each from the best G so far
 * plus a Hermitian perturbation (perturb_h_block_hermitian, :1195-1223: uniform in
 * [-scale, scale) from the lib.rs LCG with `seed`, real on the diagonal, re and im for a < b)
 * and the trace renormalisation. The best G (lowest sigma_L1) is kept. history: n_restarts + 1
 * doubles (sigma_L1 after each run) or NULL. */
int cy3q_solver_sigma_refine_analytic_restarts(cy3q_solver *s, int64_t n_restarts,
                                               int64_t n_iter_per_restart, double lr,
                                               double perturb_scale, uint64_t seed,
                                               double *best_sigma, double *history);
/* QuinticSolver::sigma_refine_lbfgs_with_restarts_gpu (quintic.rs:1100-1139): the restart loop
 * above around cy3q_solver_sigma_refine_lbfgs. */
int cy3q_solver_sigma_refine_lbfgs_restarts(cy3q_solver *s, int64_t n_restarts,
                                            int64_t n_iter_per_restart, int64_t m_history,
                                            double perturb_scale, uint64_t seed,
                                            double *best_sigma, double *history);

/* h_block access (4 N^2 doubles). set_h_block takes the Hermitian part. */
int cy3q_solver_get_h_block(cy3q_solver *s, double *h_block);
int cy3q_solver_set_h_block(cy3q_solver *s, const double *h_block);
/* Last trace-normalised T (complex, [N][N][2]) of the previous donaldson_step. */
int cy3q_solver_get_t(cy3q_solver *s, double *t_cplx);

/* Copies of the workspace's static buffers (host memory). */
int cy3q_solver_get_points(cy3q_solver *s, double *points);
int cy3q_solver_get_weights(cy3q_solver *s, double *weights, double *log_omega);
int cy3q_solver_get_sections(cy3q_solver *s, double *section_values, double *section_derivs);
/* Replaces the solver's weights (n doubles, host memory), as the reference's callers do by
 * writing QuinticSolver::weights (e.g. quintic.rs:6631-6642). Later steps, sigma and gradients
 * use them; the FS Gram already applied is not redone. */
int cy3q_solver_set_weights(cy3q_solver *s, const double *weights);
/* Arithmetic of the Galerkin rank-k updates of cy3q_solver_laplacian_spectrum: CY3Q_PREC_FP64
 * (default; the reference's FP64 sums, compute_metric_laplacian_spectrum, metric_laplacian.rs:
 * 296-418) or CY3Q_PREC_SPLIT16 (fp16 hi/lo split MFMA with FP32 partials over at most 4096
 * points, summed in FP64). The eigensolve is FP64 either way. No reference counterpart. */
int cy3q_solver_set_galerkin_precision(cy3q_solver *s, int precision);

/* monge_ampere_residual_quintic(_weighted) (quintic.rs:4961-5049): sigma_L1 of the Fermat
 * quintic metric h_block (2N x 2N block over the raw degree-k monomials, the legacy pipeline's
 * basis) at the given points ([n][10], host), with the CY weights 1/|grad f|^2, or the given
 * weights (n doubles) when weights is not NULL. NaN where the reference returns NaN. */
int cy3q_monge_ampere_residual(uint32_t k, const double *points, int64_t n, const double *h_block,
                               const double *weights, int device, double *sigma);

/* Per-stage device time (milliseconds, from HIP events on the solver's stream) accumulated
 * since profiling was switched on: out[i] for stage i, count[i] launches. Stage ids are listed in
 * cy3q_stage_name(). */
int cy3q_solver_set_profiling(cy3q_solver *s, int on);
int cy3q_solver_stage_times(cy3q_solver *s, double *ms, int64_t *count, int cap, int *n_stages);
const char *cy3q_stage_name(int stage);
int cy3q_solver_synchronize(cy3q_solver *s);

/* ============================================================================================
 * Galerkin assembly: metric Laplacian and Yukawa triple overlaps.
 * Host buffers in, host buffers out; every stage runs on `device` in FP64 (MFMA for the
 * assembly contractions). Complex arrays are interleaved (re, im) doubles.
 * ========================================================================================== */

/* MetricLaplacianConfig (route34/metric_laplacian.rs:224-271) and its Default (:277-290). */
typedef struct cy3q_laplacian_config {
  uint32_t max_total_degree; /* test monomials of total degree 1..max (default 4) */
  int64_t n_low_eigenvalues; /* default 15 */
  int64_t jacobi_max_sweeps; /* default 128 */
  double jacobi_tol;         /* default 1e-12 */
  double mass_regularisation; /* eps of the (M + eps ||M||_F I)^-1 inverse, default 1e-10 */
} cy3q_laplacian_config;
void cy3q_laplacian_config_default(cy3q_laplacian_config *cfg);

/* Number of test monomials of build_test_basis(max_total_degree) (metric_laplacian.rs:87-115)
 * and, when exps != NULL, their exponents ([n][8]). */
int cy3q_laplacian_basis(uint32_t max_total_degree, uint32_t *exps, int64_t cap, int64_t *n);

/* compute_metric_laplacian_spectrum (route34/metric_laplacian.rs:296-418), the ambient form
 * over a MetricBackground (route34/hym_hermitian.rs:257-294): points [n][8] complex, weights
 * [n], omega [n] complex. M, K as build_mass_matrix / build_stiffness_matrix (:424-524), then the
 * spectrum of the Hermitian part of (M + eps ||M||_F I)^-1 K, |lambda| ascending (:342-389).
 * eig_low: min(n_low, basis_dim) values; eig_full ([basis_dim]), mass and stiffness
 * ([basis_dim][basis_dim] complex) may be NULL. The Jacobi stops on the relative off-diagonal
 * norm (jacobi_tol relative to ||A||_F) or after jacobi_max_sweeps sweeps. */
int cy3q_metric_laplacian_spectrum(const double *points, const double *weights,
                                   const double *omega, int64_t n,
                                   const cy3q_laplacian_config *cfg, int device, double *eig_low,
                                   double *eig_full, double *mass, double *stiffness,
                                   int64_t *basis_dim);

/* The metric Laplacian of the solver's current h on the quintic, with the intrinsic inverse
 * metric g^{a b-bar} (the refinement metric_laplacian.rs:28-41 describes) on the test functions
 * phi_ab = s_a conj(s_b) / |z|^(2 d), s the raw degree-d monomials (d = section_degree <= 4,
 * alpha = a * C(d+4,4) + b). Measure w |Omega|^2 / sum w as build_mass_matrix. The generalized
 * problem K v = lambda M v is solved in FP64 on the GPU: C = W K W^H with W = chol(M)^-1, or,
 * when M is not positive definite, with M^(-1/2) over the eigenvectors of M with eigenvalue >
 * deflation_tol * max (the null-space deflation of metric_laplacian.rs:250-263; max_sweeps
 * bounds that Jacobi eigensolve of M). The n_low lowest eigenvalues of C come from Householder
 * tridiagonalisation and bisection. eig_low: the n_low smallest eigenvalues, ascending; *rank:
 * dimension kept. mass/stiffness ([B][B] complex) may be NULL. */
int cy3q_solver_laplacian_spectrum(cy3q_solver *s, uint32_t section_degree, int64_t n_low,
                                   double deflation_tol, int64_t max_sweeps, double *eig_low,
                                   double *mass, double *stiffness, int64_t *basis_dim,
                                   int64_t *rank);

/* Hermitian eigensolver used by both (hermitian_jacobi_n, metric_laplacian.rs:616-693):
 * a: [n][n] complex Hermitian; eig: n values ascending; vecs ([n][n] complex, column j for
 * eig[j]) may be NULL. *sweeps: sweeps run. */
int cy3q_hermitian_eigen(const double *a, int64_t n, int64_t max_sweeps, double tol, int device,
                         double *eig, double *vecs, int64_t *sweeps);

/* (T + shift I)^-1 for the 8-coordinate Donaldson loop, where the reference inverts the 2N x 2N
 * real block of T (with its Tikhonov shift on the diagonal) by LU (pwos_math::linalg::invert
 * at route34/ty_metric.rs:1241 and route34/schoen_metric.rs:1534). t, out: [n][n] complex,
 * host memory; t Hermitian. Computed on `device` through the inverse Cholesky factor (W^H W
 * with W = L^-1); CY3Q_ERR_NOT_PD (and out unspecified) when T + shift I is not numerically
 * positive definite, where an LU may still invert it: callers keep the LU for that case. */
int cy3q_hermitian_inverse(const double *t, int64_t n, double shift, int device, double *out);

/* No counterpart in the reference. cy3q_hermitian_eigen without any device allocation, for
 * callers that cannot allocate (code under graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): d_a ([n][n] complex, read only), d_eig (n values, ascending) and d_vecs
 * ([n][n] complex or NULL) are device pointers on `device`; *sweeps is host memory. The working
 * matrices live in a fixed 16.8 MB device scratch array of the library (one per device, kept
 * until the process ends), so n <= 512 (CY3Q_ERR_ARG beyond). Work runs on the null stream and
 * the call returns when the results are written. Calls that overlap on one device share the
 * scratch array and race; serialise them. */
int cy3q_hermitian_eigen_noalloc(const double *d_a, int64_t n, int64_t max_sweeps, double tol,
                                 int device, double *d_eig, double *d_vecs, int64_t *sweeps);

/* sz_weights (route34/yukawa_overlap_real.rs:222-268): w / |Omega|^2 renormalised to sum one,
 * zero for an unusable point; *uniformity = 1 / (1 + sd/mean). */
int cy3q_yukawa_sz_weights(const double *weights, const double *omega, int64_t n, int device,
                           double *sz, double *uniformity);

/* triple_overlap_cpu (route34/yukawa_overlap_real.rs:271-315): modes [n_modes][n] complex,
 * omega [n] complex, weights [n], h_v_diag [n_modes] or NULL (= 1). y: [n_modes]^3 complex,
 * Y_ijk = sqrt(h_i h_j h_k) sum_p w_p conj(Omega_p) psi_i psi_j psi_k over usable points.
 * n_modes <= 64. */
int cy3q_yukawa_triple_overlap(const double *modes, const double *omega, const double *weights,
                               const double *h_v_diag, int64_t n_modes, int64_t n, int device,
                               double *y);

/* compute_yukawa_couplings (route34/yukawa_overlap_real.rs:372-525) given h_v_diag: the SZ
 * weights, the central tensor, the bootstrap standard deviation over n_bootstrap resamples and
 * the N vs N/2 convergence ratio. Resample b draws n indices floor(u n) from the lib.rs LCG
 * seeded with seed + b * 2654435761; the reference's ChaCha20 resampling stream is not
 * reproduced. y: [n_modes]^3 complex, y_unc: [n_modes]^3. */
int cy3q_yukawa_couplings(const double *modes, const double *omega, const double *weights,
                          const double *h_v_diag, int64_t n_modes, int64_t n,
                          int64_t n_bootstrap, uint64_t seed, int device, double *y,
                          double *y_unc, double *uniformity, double *convergence_ratio);

/* Triple overlaps of n_modes random holomorphic modes psi_a = sum_b C_ba s_b over the raw
 * degree-d monomials s at the solver's points, with measure c_p = w_p |Omega_p|^2:
 * Y_abc = sum_p c_p psi_a psi_b psi_c for the C(n_modes + 2, 3) triples a <= b <= c (in the
 * order a, then b, then c ascending). C_ba = (g1 + i g2) / sqrt(2) with g Box-Muller normals of
 * the lib.rs LCG seeded with coeff_seed, drawn b-major. y_triples: [n_triples] complex;
 * coeffs ([nmono][n_modes] complex) may be NULL. */
int cy3q_solver_yukawa_random_modes(cy3q_solver *s, uint32_t degree, int64_t n_modes,
                                    uint64_t coeff_seed, double *y_triples, double *coeffs,
                                    int64_t *n_triples);

/* ============================================================================================
 * Batched instances (no counterpart in the reference, whose instances run one at a time; the
 * nearest analogue is the multi-stream task orchestrator, cy3/src/orchestrator.rs:1-12).
 * ========================================================================================== */
typedef struct cy3q_instance_spec {
  uint32_t k;       /* section degree */
  int64_t n_points; /* sample size */
  uint64_t seed;    /* sampler seed */
  double psi;       /* Dwork modulus (0: Fermat sampler) */
} cy3q_instance_spec;
typedef struct cy3q_batch cy3q_batch;

/* n instance specs from the lib.rs LCG seeded with `seed`, three draws u per instance:
 * psi = 0.9 u1, k = 2 + floor(5 u2) (2..6), N_p = round(exp(ln 2e4 + u3 (ln 5e5 - ln 2e4)));
 * sampler seed 42 + i. */
int cy3q_batch_draw_specs(uint64_t seed, int n, cy3q_instance_spec *out);
/* One solver per spec (QuinticSolver::new_with_psi_sz / new_with_sampler, then
 * orthonormalise_basis_fs_gram), built concurrently; the options apply to every instance. */
int cy3q_batch_new(const cy3q_instance_spec *specs, int n, const cy3q_options *opt,
                   cy3q_batch **out);
void cy3q_batch_free(cy3q_batch *b);
int cy3q_batch_size(cy3q_batch *b, int *n);
/* Borrowed handle of instance i (owned by the batch): sigma, h access, etc. */
int cy3q_batch_solver(cy3q_batch *b, int i, cy3q_solver **s);
/* h = I for every instance, then donaldson_solve(iters, 0) for each, one after another. */
int cy3q_batch_solve_sequential(cy3q_batch *b, int64_t iters);
/* The same solves, all instances captured into one hipGraph (h = I reset included) on first
 * use for this `iters`, then one graph launch. Per-instance results equal the sequential ones
 * bit for bit. */
int cy3q_batch_solve_graph(cy3q_batch *b, int64_t iters);
/* Times the tuning candidates per (k, quotient, round(log2 N_p)) and applies the fastest to
 * every instance, reusing and updating the text cache at cache_path (may be NULL).
 * *n_measured: shapes timed (not found in the cache). Leaves h = I. */
int cy3q_batch_autotune(cy3q_batch *b, const char *cache_path, int64_t *n_measured);

/* Device time per stage of the calling thread's last Galerkin call (HIP events on its stream):
 * ms[i], count[i] for stage i < *n_stages, named by cy3q_galerkin_stage_name(i). */
int cy3q_galerkin_stage_times(double *ms, int64_t *count, int cap, int *n_stages);
const char *cy3q_galerkin_stage_name(int stage);

/* ---- The route34 T-operator contract: Cy3DonaldsonGpu (route34/cy3_donaldson_gpu.rs:164-333),
 * the GPU T-operator the TY / Schoen metric solvers call. A workspace without points: the
 * caller's section values and weights are uploaded once, and each call maps h (route34
 * convention: K_p = sum_ab h_re Re(conj(s_a) s_b) + h_im Im(conj(s_a) s_b), i.e. h = conj(G) of
 * the quintic solver) to the unnormalised h_pair[a][b] = sum_p (w_p / max(K_p, 1e-30))
 * conj(s_a) s_b over points with finite w_p > 0 (no division by sum w, no trace).
 * Differences: h must be Hermitian positive definite (K is formed as |W s|^2 with W^H W =
 * conj(h); CY3Q_ERR_NOT_PD otherwise), and the sums run in a fixed order, not the reference's
 * 256-lane tree (results agree to rounding). */
typedef struct cy3q_top cy3q_top;
/* Cy3DonaldsonGpu::new (:182-209); 2 n_basis <= 1024. */
int cy3q_top_new(int64_t n_points, int64_t n_basis, int device, cy3q_top **out);
void cy3q_top_free(cy3q_top *t);
/* Cy3DonaldsonGpu::upload_static (:211-247): section_values[n_points][2 n_basis] ((re, im) per
 * basis element), weights[n_points]. */
int cy3q_top_upload_static(cy3q_top *t, const double *section_values, const double *weights);
/* Cy3DonaldsonGpu::t_operator_raw (:249-320): h_re, h_im, out_re, out_im are n_basis x
 * n_basis row-major. */
int cy3q_top_t_operator_raw(cy3q_top *t, const double *h_re, const double *h_im, double *out_re,
                            double *out_im);

/* ---- The route34 sigma contract: Cy3SigmaGpu (route34/cy3_sigma_gpu.rs:314-560), the sigma
 * evaluator of the TY / Schoen solvers on 8 ambient coordinates with caller-supplied frames.
 * section_values[n][2 n_basis], section_derivs[n][8][2 n_basis] (d/dz_i, i = 0..7), frames[n][48]
 * (T_a[i] = (re, im) at 2 (8 a + i), a = 0..2), omega_sq[n] = |Omega|^2, weights[n]. h is used
 * as given (K = s^H h s, :80-107) and must be Hermitian positive definite. sigma =
 * sum w (eta / kappa - 1)^2 / sum w over points with finite eta > 0 and w > 0, NaN when no
 * point is valid or the total weight is below 1e-30 (cy3_sigma_gpu.rs:514-519); kappa =
 * sum w eta / sum w. Difference: the sums run in a fixed order. */
typedef struct cy3q_sig8 cy3q_sig8;
int cy3q_sig8_new(int64_t n_points, int64_t n_basis, int device, cy3q_sig8 **out);  /* :336-381 */
void cy3q_sig8_free(cy3q_sig8 *g);
int cy3q_sig8_upload_static(cy3q_sig8 *g, const double *section_values,               /* :383-443 */
                            const double *section_derivs, const double *frames,
                            const double *omega_sq, const double *weights);
int cy3q_sig8_compute_sigma(cy3q_sig8 *g, const double *h_re, const double *h_im,   /* :445-547 */
                            double *sigma);

/* ---- Tian-Yau and Schoen set-up for the 8-coordinate callers of the two contracts above:
 * samples, Groebner-reduced section bases, tangent frames and section tables
 * (solve_ty_metric, route34/ty_metric.rs:461-548; solve_schoen_metric, route34/schoen_metric.rs:
 * 782-830). Coordinates are [n][8] complex as (re, im) pairs: TY z0..z3 w0..w3, Schoen
 * x0..x2 y0..y2 t0 t1. */
#define CY3Q_CICY_TY 0     /* BicubicPair::z3_invariant_default (cicy_sampler.rs:244-275) */
#define CY3Q_CICY_SCHOEN 1 /* SchoenPoly::z3xz3_invariant_default (schoen_sampler.rs:188-217) */
#define CY3Q_CICY_RESTARTS 0 /* the sequential samplers: up to 4 Newton restarts per line */
#define CY3Q_CICY_LANES 1    /* TY only: the batch sampler of the reference's GPU build */

/* The samplers' normal stream: rand_distr 0.5 StandardNormal over rand_chacha 0.9 ChaCha8Rng
 * seeded with seed_from_u64(seed) (cicy_sampler.rs:136-138, route34/schoen_sampler.rs:980-998),
 * normals 0 .. count - 1 into out (host memory), drawn on `device` as the samplers draw them. */
int cy3q_cicy_normals(uint64_t seed, int64_t count, int device, double *out);

/* Sample n_target points with the reference's ChaCha8 + StandardNormal stream for `seed`.
 *   CY3Q_CICY_RESTARTS: CicySampler::sample_batch (cicy_sampler.rs:591-617, the CPU build) or
 *     SchoenSampler::sample_points (route34/schoen_sampler.rs:482-498);
 *   CY3Q_CICY_LANES: CicySampler::sample_batch_gpu (cicy_sampler.rs:637-752, the GPU build):
 *     max(4 n_target, 64) lines drawn up front, one damped Newton attempt each
 *     (gpu_sampler.rs:265-412), the first n_target accepted kept.
 * Outputs: rescaled coordinates [n][16], |Omega|^2, weights normalised to sum one (no quotient
 * factor: apply_z3_quotient divides by 3, apply_z3xz3_quotient by 9), the metric point's patch
 * indices [n][3] (TY: z_idx in 0..3, w_idx in 0..3, -1; Schoen: x_idx, y_idx, t_idx as
 * coordinate indices, ty_metric.rs:479-496 and schoen_metric.rs:1132-1151), *n_out accepted
 * points (fewer than n_target when the line budget runs out) and *n_lines lines used. The
 * output arrays hold n_target rows. */
int cy3q_cicy_sample(int variety, int mode, uint64_t seed, int64_t n_target, int device,
                     double *coords, double *omega_sq, double *weight, int32_t *patch,
                     int64_t *n_out, int64_t *n_lines);
/* build_ty_invariant_reduced_basis (ty_metric.rs:873-895; degrees[0] = k) and
 * build_schoen_invariant_reduced_basis (schoen_metric.rs:1185-1212; degrees = d_x, d_y, d_t):
 * *n monomials, written as [n][8] exponents to out when out is not NULL (cap rows). */
int cy3q_cicy_basis(int variety, const uint32_t *degrees, uint32_t *out, int64_t cap, int64_t *n);
/* ty_affine_chart_frame (ty_metric.rs:1929-2015) / schoen_affine_chart_frame
 * (schoen_metric.rs:2202-2273) at each point: frames [n][3][8][re, im] in the layout of
 * build_ty_frames / build_schoen_frames (ty_metric.rs:1472-1494, schoen_metric.rs:1758-1783),
 * NaN rows where the chart is degenerate. */
int cy3q_cicy_frames(int variety, int64_t n, const double *coords, const int32_t *patch,
                     int device, double *frames);
/* evaluate_section_basis_at_points and _derivs_at_points (ty_metric.rs:949-1035,
 * schoen_metric.rs:1259-1333): section_values [n][2 n_basis], section_derivs
 * [n][8][2 n_basis]. */
int cy3q_cicy_sections(int64_t n, const double *coords, int64_t n_basis, const uint32_t *exps,
                       int device, double *section_values, double *section_derivs);

/* Sections, derivatives and frames of caller points (as cy3q_cicy_sections and cy3q_cicy_frames
 * compute them) built on the device straight into the static buffers of a T-operator and a sigma
 * contract, with the weights and |Omega|^2: the effect of build_ty_frames + the two
 * upload_static calls of solve_ty_metric (ty_metric.rs:521-545) without passing the tables
 * through host memory. No counterpart in the reference: its tables are host vectors. Both
 * contracts must have been created for (n, n_basis) on the same device. */
int cy3q_cicy_upload_contracts(int variety, int64_t n, const double *coords, const int32_t *patch,
                               int64_t n_basis, const uint32_t *exps, const double *omega_sq,
                               const double *weights, cy3q_top *top, cy3q_sig8 *sig);

/* ---- The sigma-gradient kernel contract: AdamGradientKernel (gpu_adam.rs:1533-2194), the
 * object the reference's callers build once per (n_points, n_basis) and pass to
 * sigma_refine_analytic_gpu / sigma_refine_lbfgs_gpu (quintic.rs:6514, 6650, 7637). Static
 * inputs are the solver's arrays in the reference layouts: section_values[n][2 n_basis],
 * section_derivs[n][5][2 n_basis], points[n][10], weights[n], log_omega_sq[n] (Fermat quintic,
 * as the reference kernel). h_block is the 2N x 2N real block; its Hermitian part is used and
 * must be positive definite (CY3Q_ERR_NOT_PD otherwise; the reference forms s^T H s
 * directly). Nothing per point is stored for the gradient: the per-point metric is recomputed. */
typedef struct cy3q_adamgrad cy3q_adamgrad;
/* AdamGradientKernel::new (gpu_adam.rs:1563-1614). */
int cy3q_adamgrad_new(int64_t n_points, int64_t n_basis, int device, cy3q_adamgrad **out);
void cy3q_adamgrad_free(cy3q_adamgrad *g);
/* AdamGradientKernel::upload_static_inputs (gpu_adam.rs:2062-2113). */
int cy3q_adamgrad_upload_static_inputs(cy3q_adamgrad *g, const double *section_values,
                                       const double *section_derivs, const double *points,
                                       const double *weights, const double *log_omega_sq);
/* AdamGradientKernel::compute_sigma_grad_h_only (gpu_adam.rs:1769-1938): sigma_L2^2, its
 * gradient (2 n_basis^2 doubles, the layout of cy3q_solver_sigma_sq_grad) and sigma_L1. */
int cy3q_adamgrad_compute_sigma_grad_h_only(cy3q_adamgrad *g, const double *h_block,
                                            double *sigma_l2sq, double *grad, double *sigma_l1);
/* AdamGradientKernel::compute_per_point_h_only (gpu_adam.rs:2119-2176): per_point_buf
 * [n][2 + 2 n_basis^2], row p = (eta_p, w_p, d log eta_p / dh) in the reference's layout
 * (cpu_compute_per_point_reference, gpu_adam.rs:2677-2739): (NaN, 0, 0, ...) for a point with an
 * invalid weight or eta, a zero gradient where the tangent metric does not invert.
 * n_basis <= 340. */
int cy3q_adamgrad_compute_per_point_h_only(cy3q_adamgrad *g, const double *h_block,
                                           double *per_point_buf);
/* AdamGradientKernel::compute_per_point (gpu_adam.rs:1940-2060): upload_static_inputs, then
 * compute_per_point_h_only. */
int cy3q_adamgrad_compute_per_point(cy3q_adamgrad *g, const double *section_values,
                                    const double *section_derivs, const double *h_block,
                                    const double *points, const double *weights,
                                    const double *log_omega_sq, double *per_point_buf);
/* AdamGradientKernel::total_device_bytes (gpu_adam.rs:2180-2194). */
int cy3q_adamgrad_total_device_bytes(const cy3q_adamgrad *g, int64_t *bytes);

#ifdef __cplusplus
}
#endif
#endif /* CY3Q_H */
