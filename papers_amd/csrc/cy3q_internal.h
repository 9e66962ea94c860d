// Internal declarations shared by the cy3q HIP kernels and the C++ host solver.
// Written for gfx950 (CDNA4, 64-wide wavefronts, MFMA); no other target is built.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstddef>
#include <string>
#include <utility>
#include <vector>

namespace cy3q {

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));
typedef float float16_t __attribute__((ext_vector_type(16)));
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------------------
// Layout constants
// ---------------------------------------------------------------------------------------
// A section row is stored real-ified and interleaved, exactly as the reference's
// section_values row (re_0, im_0, re_1, im_1, ...), zero-padded from 2N to M columns.
// M is a multiple of 32 so every MFMA column group of 16 and every 32-wide tile is whole.
inline int64_t padded_cols(int64_t n_basis) { return (2 * n_basis + 31) / 32 * 32; }

// ---------------------------------------------------------------------------------------
// Status codes (mirrored in include/cy3q.h)
// ---------------------------------------------------------------------------------------
enum {
  ST_OK = 0,
  ST_ERR_ARG = -1,
  ST_ERR_HIP = -2,
  ST_ERR_NOT_PD = -3,
  ST_ERR_EMPTY = -4,
  ST_ERR_STATE = -5,
};

// ---------------------------------------------------------------------------------------
// Kernel launchers (defined in the .hip files; all asynchronous on `st`)
// ---------------------------------------------------------------------------------------

// Sampler (k_sampler.hip). Lines [line0, line0 + n_lines) of the SZ stream for `seed`.
// cand: [n_lines][5][10] doubles, acc: [n_lines][5] int32 accept flags.
void launch_sz_lines(hipStream_t st, uint64_t seed, int64_t line0, int64_t n_lines, int dwork,
                     double psi, int dk_iters, int newton_polish, double *cand, int32_t *acc);
void launch_newton_tries(hipStream_t st, uint64_t seed, int64_t try0, int64_t n_tries, double psi,
                         double tol, double *cand, int32_t *acc);
// Counts accepted candidates per block of `block` lines; writes block sums.
void launch_scan_accept(hipStream_t st, const int32_t *acc, int64_t n_cand, int64_t *pos,
                        int64_t *block_sums, int64_t *scratch_sums, int64_t *total);
void launch_compact_points(hipStream_t st, const double *cand, const int32_t *acc,
                           const int64_t *pos, int64_t n_cand, int64_t out_offset,
                           int64_t n_max, double *pts);

// Tian-Yau / Schoen set-up (k_cicy.hip). variety 0 = TY, 1 = Schoen. nrm: standard normals.
void launch_cicy_span(hipStream_t st, int variety, const double *nrm, int64_t s0, int64_t n_cand,
                      uint8_t *info);
void launch_cicy_chain(hipStream_t st, int variety, const double *nrm, const int64_t *start,
                       const uint8_t *info, int64_t n_lines, double *x, int32_t *flag);
void launch_cicy_lanes(hipStream_t st, const double *nrm, int64_t n_lanes, double *x, int32_t *flag);
// The ChaCha8 + ziggurat normal stream (k_cicy.hip): key words of the keystream, the fast-path
// values and failed positions of a range of u64 positions, and the gather of normals from runs.
struct ZigKey {
  uint32_t k[8];
};
void upload_zig_x(hipStream_t st, const double *x257);
void launch_zig_fast(hipStream_t st, const ZigKey &key, uint64_t pos0, int64_t n_pos, double *fast,
                     uint64_t *words, int64_t *exc, int64_t exc_cap,
                     unsigned long long *n_exc);
void launch_zig_exc_words(hipStream_t st, const uint64_t *words, const int64_t *exc, int64_t n,
                          uint64_t pos0, uint64_t *out);
void launch_zig_gather(hipStream_t st, const double *fast, const int64_t *run_k,
                       const int64_t *run_p, const int64_t *run_len, int64_t n_runs,
                       const int64_t *exc_k, const double *exc_v, int64_t n_exc, int64_t count,
                       double *out);
void launch_cicy_post(hipStream_t st, int variety, const double *x, const int32_t *flag, int64_t n,
                      double *coords, double *omega_sq, double *weight, int32_t *patch, int32_t *ok);
void launch_cicy_frames(hipStream_t st, int variety, const double *coords, const int32_t *patch,
                        int64_t n, double *frames);
void launch_cicy_sections(hipStream_t st, const double *coords, int64_t n, const uint32_t *exps,
                          int nb, double *sv, double *sd);

// Per-point geometry (k_sections.hip).
void launch_weights_logomega(hipStream_t st, const double *pts, int64_t n, double psi,
                             double *w, double *log_om, int32_t *chart_elim);
void launch_sections(hipStream_t st, const double *pts, int64_t n, const uint32_t *mono,
                     int nb, int kmax, int64_t m_cols, double *sv, double *sd);

// FP64 MFMA contractions (k_gemm_f64.hip).
// y[r][:] = x[r][:] * B for r < rows (B: [M][M] row-major, block-upper-triangular), with
// mode 0: store y (ld = m_cols); mode 1: rowsumsq partial per 32-col part into dpart.
void launch_trigemm_f64(hipStream_t st, const double *x, int64_t rows, int64_t m_cols,
                        const double *bmat, double *y, int mode, int full);
// FP64 D pass: D parts dpart[q][p] (q < dpass_f64_ranges(M)) of D_p = |x_p B|^2 and the
// coefficients coef[p] = w_p / max(D_p, 1e-30) (0 for an invalid weight).
int dpass_f64_ranges(int64_t m_cols);
void launch_dpass_f64(hipStream_t st, const double *x, int64_t rows, int64_t m_cols,
                      const double *bmat, const double *w, double *dpart, double *coef);
// Weighted symmetric rank-k: partial lower tiles of X^T diag(c) X over point chunks.
struct SyrkPlan;
int syrk_f64_max_tiles(int64_t m_cols);  // accumulator tiles per wave of k_syrk_f64
void launch_syrk_f64(hipStream_t st, const SyrkPlan &plan, const double *x, const double *c,
                     int64_t n, double *partials);
void launch_syrk_reduce(hipStream_t st, const SyrkPlan &plan, const double *partials,
                        double *xx_full);

// Split-fp16 MFMA contractions (k_split16.hip).
void launch_split_planes(hipStream_t st, const double *sv, int64_t n, int64_t n_pad, int64_t m,
                         int *ex, float *rs2, _Float16 *xd, float *xt);
void launch_build_bfrag16(hipStream_t st, const double *v_cplx, int nb, int64_t m_cols,
                          half8_t *bfrag, double *b_scale);
int dpass_split16_waves();
int dpass_split16_max_cg();
std::vector<std::pair<int, int>> dpass_split16_ranges(int64_t m_cols);
bool dpass_split16_supported(int64_t m_cols);
void launch_dpass_split16(hipStream_t st, const half8_t *xd, int64_t n, int64_t m_cols,
                          const half8_t *bfrag, const int16_t *cglist, int n_cu, double *dsum,
                          double *dsum_parts);
void launch_coeffs_f64_from_split(hipStream_t st, const double *w, const double *dsum,
                                  const float *rs2, const double *bscale, int64_t n, double *c);
// Per-point eta of the 8-coordinate sigma contract (k_sigma8.hip): y [n][9][m], frames [n][48].
void launch_pack_rows9(hipStream_t st, const double *sv, const double *sd, int64_t n,
                       int64_t two_n, int64_t m, double *x9);
void launch_eta8(hipStream_t st, const double *y, int64_t m, int nb, const double *frames,
                 const double *om, const double *w, int64_t n, double *eta);
void launch_coeffs_split16(hipStream_t st, const double *w, const double *dsum, const float *rs2,
                           const double *bscale, int64_t n, float *u, float *block_max,
                           float *zscale, double *tscale);
// Split-fp16 T pass (k_tpass.hip).
int tpass_max_cols();
int64_t tpass_segment_cols(int64_t m_cols);
size_t tpass_partials(int64_t m_cols, int64_t n_chunks);
size_t tpass_seg_count(int64_t m_cols, int64_t n_chunks);
void launch_tpass(hipStream_t st, const float *xt, const float *u, int64_t n, int64_t n_pad,
                  int64_t m_cols, int64_t chunk, int64_t n_chunks, float *partials);
void launch_tpass_reduce(hipStream_t st, const float *partials, int64_t m_cols, int64_t n_chunks,
                         double *seg, double *xx);

// Per-point coefficients, reductions, B operand, layout conversions (k_dense.hip).
void launch_point_coeffs_f64(hipStream_t st, const double *w, const double *dpart, int n_parts,
                             int64_t n, double *c);
void launch_valid_weights(hipStream_t st, const double *w, int64_t n, double *c);
void launch_sum_f64(hipStream_t st, const double *v, int64_t n, double *out);
void launch_build_bmat(hipStream_t st, const double *v_cplx, int nb, int64_t m_cols,
                       int conj_v, double *bmat);
void launch_cplx_to_block(hipStream_t st, const double *g_cplx, int nb, double *h_block);
void launch_block_to_cplx(hipStream_t st, const double *h_block, int nb, double *g_cplx);
void launch_identity_cplx(hipStream_t st, int nb, double *g_cplx);
void launch_hermitian_part(hipStream_t st, double *g_cplx, int nb);
void launch_legacy_replace(hipStream_t st, const double *t, int nb, double *g, double *resid);
void launch_fill_f64(hipStream_t st, double *x, int64_t n, double v);

// N x N solve (k_hsolve.hip): T from xx; inverse Cholesky factor W = L^-1; G = W^H W.
void launch_t_from_xx(hipStream_t st, const double *xx, int64_t m_cols, int nb,
                      const double *sum_w, double hscale, const double *dscale,
                      int trace_normalise, double *scale_buf, double *t, double *a, double *r);
void launch_elim_init(hipStream_t st, const double *src, int nb, double *a, double *r);
void launch_inverse_cholesky(hipStream_t st, double *a, double *r, double *einv, double *w, int nb,
                             int32_t *status);
// With counter (a zero-initialised unsigned that the kernel leaves at zero), the last tile
// workgroup forms the residual; otherwise a second launch does.
void launch_gram_residual(hipStream_t st, double *w, int nb, double *g, const int32_t *status,
                          double *partial, double *residual, unsigned *counter = nullptr);

// sigma (k_sigma.hip).
void launch_sigma_eta(hipStream_t st, const double *pts, const double *y0, const double *yd,
                      const double *log_om, int64_t n, int64_t m_cols, double psi,
                      double *eta);
// sigma reduction into out6 (see k_sigma.hip); scratch holds sigma_reduce_scratch() doubles.
int sigma_reduce_scratch();
void launch_sigma_reduce(hipStream_t st, const double *eta, const double *w, int64_t n,
                         double *scratch, double *out6, double min_total_w = 1e-12);
// Adjoint of sigma_L2^2 and Adam (k_adjoint.hip).
void launch_grad_rows(hipStream_t st, const double *pts, const double *y0, const double *yd,
                      const double *log_om, const double *w, const double *sv, const double *sd,
                      int64_t n, int64_t m_cols, double psi, const double *sig, double *phi,
                      double *vs, double *eta_out = nullptr);
// Rows of the reference's per-point buffer ([n][2 + 2 nb^2]: eta, w, d log eta / dh) from the
// sig == nullptr rows of launch_grad_rows and their eta. 12 * 2 nb doubles of LDS per point.
int per_point_grad_max_nb();
void launch_per_point_grad(hipStream_t st, const double *sv, const double *sd, const double *phi,
                           const double *vs, const double *eta, const double *w, int nb,
                           int64_t m_cols, int64_t n, double *out);
int64_t gemm_tn_groups(int64_t rows, int64_t m_cols);
void launch_gemm_tn_acc(hipStream_t st, const double *u, const double *v, int64_t rows,
                        int64_t m_cols, int64_t n_groups, double *part, double *acc);
void launch_grad_assemble(hipStream_t st, const double *cm, int nb, int64_t m_cols,
                          const double *sig, double *grad);
void launch_adam_step(hipStream_t st, double *g_cplx, int nb, const double *grad, double *m,
                      double *v, double scale, double beta1, double beta2, double eps);
// Trace renormalisation of G to tr Re G = nb (when the trace exceeds 1e-10).
void launch_trace_renorm(hipStream_t st, double *g_cplx, int nb);

// L-BFGS vectors (k_lbfgs.hip). The curvature pairs in use, oldest first: slot[k] indexes the
// pair's rows of the s and y stores, rho[k] = 1 / (s.y).
constexpr int LBFGS_MAX_PAIRS = 64;
struct LbfgsPairs {
  int count;
  int slot[LBFGS_MAX_PAIRS];
  double rho[LBFGS_MAX_PAIRS];
};
void launch_lbfgs_direction(hipStream_t st, const double *g, const double *s_all,
                            const double *y_all, const LbfgsPairs &hist, int64_t n2, double *d,
                            double *out);
void launch_lbfgs_apply(hipStream_t st, const double *g0, int nb, const double *d, double step,
                        double *g);
void launch_lbfgs_pair(hipStream_t st, const double *d, double step, const double *g_new,
                       const double *g_old, int64_t n2, double *s_out, double *y_out, double *out);

// Metric Laplacian assembly (k_laplacian.hip).
void launch_lap_metric(hipStream_t st, const double *pts, const double *y0, const double *yd,
                       const double *w, const double *log_om, int64_t n, int64_t m_cols,
                       double psi, double *pr, double *wv);
void launch_lap_rows_quintic(hipStream_t st, const double *pts, const double *pr, int64_t n,
                             const uint8_t *smono, int ns, int sdeg, int64_t mc, double *x0,
                             double *x1);
// The intrinsic rows of the split assembly straight into the fp16 hi/lo planes of its four
// buffers (16 points per stage), after a pass for the chunk's largest |entry| (one per stage).
void launch_lap_rows_absmax(hipStream_t st, const double *pts, const double *pr, int64_t n,
                            const uint8_t *smono, int ns, int sdeg, double *part);
void launch_lap_rows_planes(hipStream_t st, const double *pts, const double *pr, int64_t n,
                            const uint8_t *smono, int ns, int sdeg, int64_t mc,
                            const double *scale, half8_t *const *h, half8_t *const *l);
void launch_lap_rows_ambient(hipStream_t st, const double *pts, const double *w,
                             const double *omega, int64_t n, const uint8_t *exps, int nb,
                             int64_t mc, double *x0, double *x1, double *wv);
int herk_max_buffers();
// Split-precision rank-k update of the Galerkin assembly (k_herk_split.hip).
// One power-of-two range for the rows of two arrays: scale[0] applied, scale[1] restores products.
void launch_rows_scale_pair(hipStream_t st, const double *a, int64_t na, const double *b,
                            int64_t nb, double *part, double *scale);
int rows_scale_parts();
void launch_rows_scale_from_parts(hipStream_t st, const double *part, int nparts, double *scale);
// ldx: the leading dimension of x (columns past it are zero); 0 means mc.
void launch_split_rows(hipStream_t st, const double *x, int64_t rows, int64_t stages, int64_t mc,
                       const double *scale, half8_t *xh, half8_t *xl, int64_t ldx = 0);
// Y = X B on the fp16 split MFMA (FP32 accumulation, FP64 output). X [rows][m] FP64 is first
// written as row-scaled hi/lo planes of (m / 16) * rowgemm_rows_pad(rows) * 2 units each, with
// the row exponents ex; B is the real-ified W^T as hi/lo planes of (m / 16) *
// rowgemm_cols_pad(m) * 2 units from launch_split_rows(B, m, m / 16, cols_pad, bscale, .., m).
// y [rows][ldy]. Requires rowgemm_supported(m) (m a multiple of 32, at most 1024).
int64_t rowgemm_rows_pad(int64_t rows);
int64_t rowgemm_cols_pad(int64_t m);
inline bool rowgemm_supported(int64_t m) { return m > 0 && m % 32 == 0 && m <= 1024; }
void launch_rows_split_planes(hipStream_t st, const double *x, int64_t rows, int64_t m, int *ex,
                              half8_t *xh, half8_t *xl);
void launch_rowgemm_split16(hipStream_t st, const half8_t *xh, const half8_t *xl, int64_t rows,
                            int64_t m, const int *ex, const half8_t *bh, const half8_t *bl,
                            const double *bscale, double *y, int64_t ldy);
int herk_split_parts(int nbuf, int n_tiles, int n_cu);
void launch_herk_split16(hipStream_t st, const half8_t *const *xh, const half8_t *const *xl,
                         double *const *c, int nbuf, int64_t stages, int64_t mc, int nparts,
                         const double *scale);
void launch_sum_parts(hipStream_t st, double *c, int64_t len, int nparts);
int herk_split_tile();
void launch_herk_f64(hipStream_t st, const double *const *x, const int64_t *rows, double *const *c,
                     int nbuf, int64_t mc);
void launch_lap_finalize(hipStream_t st, const double *cmass, const double *const *ck, int nk,
                         int B, int64_t mc, const double *tw, int clamp_k, double2 *M, double2 *K);

// Hermitian Jacobi eigensolver and complex GEMM (k_eigen.hip).
void launch_jacobi_round(hipStream_t st, const double2 *A, double2 *A2, const double2 *V,
                         double2 *V2, const int *tag, int *tag2, int np, int want_v);
int offnorm_parts();
void launch_offnorm(hipStream_t st, const double2 *A, int np, double *part, double *out2);
void launch_jacobi_init(hipStream_t st, const double2 *src, int n, int np, double shift,
                        double2 *A, double2 *V, int *tag);
void launch_diag_re(hipStream_t st, const double2 *A, int np, double *d);
void launch_gather_cols(hipStream_t st, const double2 *V, int np, int n, const int *col,
                        const double *scale, int r, double2 *out);
void launch_herm_project(hipStream_t st, double2 *A, int n);
// Householder tridiagonalisation (in place; d: n, e: n - 1, v: n double2, p: 2 n double2, scal: 3) and
// bisection of the lowest n_want eigenvalues of the tridiagonal (k_tridiag.hip).
int tridiag_smem_ok(int n);
void launch_hh_tridiag(hipStream_t st, double2 *A, int n, double2 *v, double2 *p, double *scal,
                       double *d, double *e);
void launch_tridiag_bisect(hipStream_t st, const double *d, const double *e, int n, int n_want,
                           double *ev);
void launch_add_diag(hipStream_t st, double2 *a, int n, double shift);
void launch_zgemm(hipStream_t st, const double2 *A, const double2 *B, double2 *C, int M, int N,
                  int K, int lda, int ldb, int ldc, int opa, int opb);

// Yukawa triple overlaps (k_yukawa.hip): a plan of pair products x mode values on the FP64
// MFMA for nm modes over n points.
struct YukDims {
  int nm = 0, npair = 0, vld = 0, vldp = 0, uld = 0, ntiles = 0, nparts = 0, n_trip = 0;
  int cfull_rows = 0;
  int64_t chunk = 0, n_chunks = 0;
  size_t lds_bytes = 0;
};
struct YukPlanShape : YukDims {
  std::vector<short2> pairs, tiles;
  std::vector<int4> parts;
  std::vector<short4> trip;
};
struct YukPlan : YukDims {
  double *psi_v = nullptr, *partials = nullptr, *cfull = nullptr;
  double2 *g = nullptr;
  const short2 *pairs = nullptr, *tiles = nullptr;
  const int4 *parts = nullptr;
  const short4 *trip = nullptr;
};
int yukawa_max_modes();
YukPlanShape yukawa_plan_shape(int nm, int64_t n);
void launch_yuk_prep(hipStream_t st, const double2 *modes, int64_t mode_stride,
                     int64_t point_stride, const double *w, const double *omega, int64_t n,
                     int check_psi, int nm, int vld, double *psi_v, double2 *g);
void launch_yuk_contract(hipStream_t st, const YukPlan &pl, int64_t n, const double *hdiag,
                         double2 *y_tri, double2 *y_full);
void launch_modes_bmat(hipStream_t st, const double2 *coef, int nsec, int nm, int64_t krows,
                       int64_t ncols, double *b);
void launch_gemm_nn_f64(hipStream_t st, const double *x, int64_t rows, int64_t k_cols,
                        const double *bmat, int64_t ldb, int64_t n_cols, double *y, int64_t ldy);
void launch_sz_weights(hipStream_t st, const double *w, const double *omega, int64_t n, double *sz,
                       double *out);
void launch_scale_weights(hipStream_t st, const double *w, const int32_t *count,
                          const double *total, int64_t n, double *out);
void launch_measure_w_omega2(hipStream_t st, const double *w, const double *log_om, int64_t n,
                             double *c);
void launch_bootstrap_counts(hipStream_t st, uint64_t seed_b, int64_t n, int32_t *count);
void launch_random_coeffs(hipStream_t st, uint64_t seed, int64_t count, double2 *out);
void launch_bootstrap_acc(hipStream_t st, const double2 *yb, const double2 *y, int64_t n3,
                          double *acc);
void launch_bootstrap_final(hipStream_t st, const double *acc, int64_t n3, double nbs,
                            double *unc);
void launch_conv_ratio(hipStream_t st, const double2 *y, const double2 *yh, int64_t n3,
                       double *out);

// ---------------------------------------------------------------------------------------
// SYRK work partition: lower 16x16 (FP64) or 32x32 (split16) output sub-tiles of the
// M x M symmetric result, grouped into "parts" (one workgroup each per point chunk), each
// part split over the waves of the workgroup. Built on the host (see solver.cpp).
// ---------------------------------------------------------------------------------------
struct SyrkPlan {
  int64_t m_cols = 0;     // M
  int prog13 = 0;         // 1: k_syrk13_f64 (M = 416, fixed wave programs)
  int tile = 16;          // sub-tile edge (16 for FP64 MFMA, 32 for split16)
  int n_groups = 0;       // M / tile
  int n_parts = 0;        // workgroups per point chunk
  int waves = 8;          // waves per workgroup
  int max_tiles = 0;      // max sub-tiles per wave (accumulator budget)
  int64_t chunk = 0;      // points per workgroup
  int64_t n_chunks = 0;
  // Per (part, wave): list of sub-tile (row group, col group) pairs, padded with -1.
  // Layout: [n_parts][waves][max_tiles][2] int16.
  int16_t *d_tiles = nullptr;
  // Row-group order of the per-wave lists lets a wave reuse the row fragment.
  // Segment plans (M > 448): per part, (first column group, groups) of the two staged column
  // segments, [n_parts][4] int32; seg_cols = the most columns a part stages. Otherwise null / 0
  // (every part stages all M columns).
  int32_t *d_pcols = nullptr;
  int seg_cols = 0;
};

// ---------------------------------------------------------------------------------------
// Host-side access to a solver's device workspace for the Galerkin entry points
// (solver.cpp). The pointers stay owned by the solver.
// ---------------------------------------------------------------------------------------
struct SolverView {
  hipStream_t st = nullptr;
  int device = 0;
  int64_t n = 0, m = 0;  // points, padded real columns of the section rows
  int nb = 0;
  double psi = 0.0;
  const double *pts = nullptr, *w = nullptr, *log_om = nullptr, *sv = nullptr, *sd = nullptr;
  const double *bmat = nullptr;  // real-ified W = L^-1 of the current h (W s gives K = |W s|^2)
  int galerkin_precision = 0;    // CY3Q_PREC_* of the Galerkin rank-k updates
};

}  // namespace cy3q

struct cy3q_solver;
namespace cy3q {
// Records msg as the calling thread's last error and returns code.
int set_error(int code, const std::string &msg);
// Fills v for s (building the contraction operand of the current h first).
int solver_view(cy3q_solver *s, SolverView *v);
// Batched and captured solves (batch.cpp): the solver's stream; set-up that must precede a
// capture; donaldson_solve(iters, tol <= 0) as launches only; h = I as launches; the last
// residual of the previous enqueued solve.
hipStream_t solver_stream(cy3q_solver *s);
int solver_prepare_enqueue(cy3q_solver *s, int64_t iters);
int solver_enqueue_solve(cy3q_solver *s, int64_t iters);
int solver_reset_identity(cy3q_solver *s);
int solver_last_residual(cy3q_solver *s, int64_t iters, double *out);
}  // namespace cy3q
