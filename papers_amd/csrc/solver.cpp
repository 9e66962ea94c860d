// Host side of cy3q: the QuinticSolver workspace and the C ABI of include/cy3q.h.
//
// Mirrors the reference crate's QuinticSolver (cy3/src/quintic.rs:117-165, constructors
// :174-393, methods :423-514). The host keeps the same workspace: points, weights, log|Omega|^2,
// section values and derivatives, and h. Every buffer lives on one GPU, and every stage is a
// HIP kernel on the solver's stream. The host allocates, orders the launches and copies out
// the few scalars the API returns.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cy3q.h"
#include "cy3q_internal.h"
#include "cy3q_lcg.h"

using namespace cy3q;

namespace {

thread_local std::string g_last_error;

int poison_allocs() {
  static const int mode = [] {
    const char *v = std::getenv("CY3Q_POISON");
    return (v && (v[0] == '1' || v[0] == '2')) ? v[0] - '0' : 0;
  }();
  return mode;
}

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail(CY3Q_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));         \
  } while (0)

#define ST_TRY(expr)          \
  do {                        \
    int rc_ = (expr);         \
    if (rc_ != CY3Q_OK) return rc_; \
  } while (0)

// Stage ids for the profiling counters.
enum Stage {
  STG_SAMPLE = 0,
  STG_SECTIONS,
  STG_FSGRAM,
  STG_DPASS,
  STG_COEFF,
  STG_TPASS,
  STG_TREDUCE,
  STG_NXN,
  STG_SIGMA,
  STG_COUNT
};
const char *kStageNames[STG_COUNT] = {"sample",  "sections", "fs_gram",  "d_pass", "coeffs",
                                      "t_pass",  "t_reduce", "nxn_fp64", "sigma"};

int64_t n_sections(uint32_t k) {
  int64_t kk = k;
  return (kk + 1) * (kk + 2) * (kk + 3) * (kk + 4) / 24;
}

// build_degree_k_quintic_monomials (quintic.rs:3192-3209): e0 up, e1 up, e2 up, e3 up,
// e4 = k - sum. The quotient basis drops e0 >= 5 in the same order.
std::vector<uint32_t> monomials(uint32_t k, int quotient) {
  std::vector<uint32_t> out;
  int kk = (int)k;
  for (int e0 = 0; e0 <= kk; e0++)
    for (int e1 = 0; e1 <= kk - e0; e1++)
      for (int e2 = 0; e2 <= kk - e0 - e1; e2++)
        for (int e3 = 0; e3 <= kk - e0 - e1 - e2; e3++) {
          int e4 = kk - e0 - e1 - e2 - e3;
          if (e4 < 0) continue;
          if (quotient && e0 >= 5) continue;
          out.insert(out.end(), {(uint32_t)e0, (uint32_t)e1, (uint32_t)e2, (uint32_t)e3,
                                 (uint32_t)e4});
        }
  return out;
}

}  // namespace

struct cy3q_solver {
  int device = 0;
  hipStream_t st = nullptr;
  cy3q_options opt{};
  uint32_t k = 0;
  int nb = 0;
  int64_t n = 0;
  int64_t m = 0;  // padded real columns
  double psi = 0.0;
  int precision = CY3Q_PREC_FP64;
  int64_t refine = 1;  // FP64 sweeps that end a split-precision donaldson_solve
  int refine_kind = CY3Q_REFINE_T_FP64;  // what one refinement sweep is
  std::vector<uint32_t> mono;
  std::vector<void *> allocs;
  std::vector<size_t> alloc_bytes;  // bytes of each entry of allocs

  // static workspace
  double *pts = nullptr, *w = nullptr, *log_om = nullptr, *sv = nullptr, *sd = nullptr;
  int32_t *chart_elim = nullptr;
  uint32_t *d_mono = nullptr;
  double *wvalid = nullptr, *sum_w = nullptr;
  // h state: G = W^H W, W lower triangular
  double *g = nullptr, *wmat = nullptr, *t = nullptr, *wnew = nullptr, *gnew = nullptr;
  double *bmat = nullptr, *dpart = nullptr, *coef = nullptr, *xx = nullptr, *partials = nullptr;
  // Accumulated FS-Gram basis change (complex N x N, lower triangular): the solver's sections
  // are conj(wfs) times the raw monomials. nullptr while no FS-Gram has been applied.
  double *wfs = nullptr;
  double *resid = nullptr, *sig_out = nullptr, *sig_scratch = nullptr;
  // elimination workspace of the N x N solve
  double *abuf = nullptr, *rbuf = nullptr, *einv = nullptr, *partial = nullptr,
         *scale_buf = nullptr;
  int32_t *status = nullptr;
  bool bmat_valid = false;
  SyrkPlan plan64;
  int64_t partials_cap = 0;
  // split16 path (allocated on first use)
  bool x32_valid = false, bfrag_valid = false;
  float *rs2 = nullptr, *u = nullptr, *zscale = nullptr, *partials32 = nullptr,
        *block_max = nullptr;
  double *tseg = nullptr;            // T-pass segment sums
  int64_t tp_chunk = 0, tp_chunks = 0;
  int n_cu = 1;                // compute units of the device (persistent D-pass grid)
  int galerkin_precision = CY3Q_PREC_FP64;  // Galerkin rank-k updates
  int64_t tp_mult = 1;         // T-pass point chunks per CU (a tuning knob)
  int64_t syrk_target = 1024;  // workgroups of the FP64 rank-k pass (a tuning knob)
  double *batch_res = nullptr;  // residuals of captured solves
  double *ones = nullptr, *n_as_w = nullptr;  // unit weights and their sum (legacy iteration)
  int64_t batch_res_cap = 0;
  int *row_ex = nullptr;
  _Float16 *xd = nullptr;
  float *xt = nullptr;
  double *dsum = nullptr, *dsum_parts = nullptr, *bscale = nullptr, *tscale = nullptr;
  half8_t *bfrag = nullptr;
  int16_t *cglist = nullptr;

  // sigma^2 gradient and Adam (allocated on first use)
  double *grad_dev = nullptr, *adam_m = nullptr, *adam_v = nullptr;
  // sigma / gradient workspaces (allocated on first use and kept): the contracted rows of every
  // point when they fit in kRowsResidentBytes (else per chunk), eta, and the gradient rows.
  double *y0all = nullptr, *ydall = nullptr, *eta_buf = nullptr;
  bool rows_resident = false;
  double *g_phi = nullptr, *g_vs = nullptr, *g_part = nullptr, *g_cm = nullptr;
  int64_t g_chunk = 0, g_groups = 0;
  double adam_beta1 = 0.9, adam_beta2 = 0.999, adam_eps = 1e-8;

  // profiling
  bool profiling = false;
  double stage_ms[STG_COUNT] = {0};
  int64_t stage_cnt[STG_COUNT] = {0};
  hipEvent_t ev_a = nullptr, ev_b = nullptr;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;

  ~cy3q_solver() {
    if (device >= 0) hipSetDevice(device);
    if (st) hipStreamSynchronize(st);
    for (auto &pe : pending) {
      hipEventDestroy(pe.second.first);
      hipEventDestroy(pe.second.second);
    }
    for (void *p : allocs) hipFree(p);
    if (plan64.d_tiles) hipFree(plan64.d_tiles);
    if (plan64.d_pcols) hipFree(plan64.d_pcols);
    if (st) hipStreamDestroy(st);
  }

  template <class T>
  int alloc(T **p, size_t count) {
    void *q = nullptr;
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess)
      return fail(CY3Q_ERR_HIP, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    // CY3Q_POISON=1 fills every new buffer with 0xFF bytes (NaN in f64, f32 and f16), =2 with
    // 0x7F bytes (huge finite f64 and f32), so that a read before the first write shows up in
    // the results, also through max-reductions that skip NaN (a checking aid).
    if (const int pv = poison_allocs()) {
      e = hipMemsetAsync(q, pv == 2 ? 0x7F : 0xFF, bytes, st);
      if (e != hipSuccess) return fail(CY3Q_ERR_HIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
    }
    allocs.push_back(q);
    alloc_bytes.push_back(bytes);
    *p = (T *)q;
    return CY3Q_OK;
  }
  void release(void *p) {
    auto it = std::find(allocs.begin(), allocs.end(), p);
    if (it != allocs.end()) {
      hipFree(p);
      alloc_bytes.erase(alloc_bytes.begin() + (it - allocs.begin()));
      allocs.erase(it);
    }
  }
  size_t device_bytes() const {
    size_t b = 0;
    for (size_t x : alloc_bytes) b += x;
    return b;
  }

  // --- profiling helpers: events bracket a stage on the solver's stream ---
  void stage_begin(int) {
    if (!profiling) return;
    hipEvent_t a;
    hipEventCreate(&a);
    hipEventRecord(a, st);
    pending.push_back({-1, {a, nullptr}});
  }
  void stage_end(int id) {
    if (!profiling || pending.empty()) return;
    hipEvent_t b;
    hipEventCreate(&b);
    hipEventRecord(b, st);
    pending.back().first = id;
    pending.back().second.second = b;
  }
  void collect() {
    for (auto &pe : pending) {
      if (pe.first < 0 || !pe.second.second) continue;
      hipEventSynchronize(pe.second.second);
      float ms = 0.f;
      hipEventElapsedTime(&ms, pe.second.first, pe.second.second);
      stage_ms[pe.first] += ms;
      stage_cnt[pe.first] += 1;
      hipEventDestroy(pe.second.first);
      hipEventDestroy(pe.second.second);
    }
    pending.clear();
  }
};

namespace {

// Plan of the FP64 rank-k pass past 448 columns: the column groups are cut into segments of
// SY_SEGG groups (208 columns), and each part is a segment pair a <= b: the lower triangle of
// segment a's tiles (a = b) or all of segment b's rows by segment a's columns. A part's
// workgroups stage only its two segments (at most 416 columns) per 16-point stage, where the
// all-column plan re-read every point's whole row once per part and staged 8 points at a time.
// (At k = 8, 500k points: 22.4 ms per pass against 23.8 ms; 144-column segments, at two
// workgroups per CU, took 29 ms.) Per part, pcols holds (first group, groups) of segment a and
// of segment b; the tiles keep their global (row group, column group) for the reduction.
constexpr int SY_SEGG = 13;

int build_syrk_segment_plan(cy3q_solver *s, SyrkPlan &pl, int64_t target_wgs, int64_t min_chunk,
                            int64_t kp) {
  const int ng = pl.n_groups, waves = pl.waves;
  const int n_seg = (ng + SY_SEGG - 1) / SY_SEGG;
  std::vector<std::vector<std::pair<int, int>>> part_tiles;
  std::vector<int32_t> pc;
  for (int b = 0; b < n_seg; b++)
    for (int a = 0; a <= b; a++) {
      const int a0 = a * SY_SEGG, wa = std::min(SY_SEGG, ng - a0);
      const int b0 = b * SY_SEGG, wb = std::min(SY_SEGG, ng - b0);
      std::vector<std::pair<int, int>> tl;
      for (int r = 0; r < (a == b ? wa : wb); r++)
        for (int c = 0; c < (a == b ? r + 1 : wa); c++)
          tl.push_back({(a == b ? a0 : b0) + r, a0 + c});
      part_tiles.push_back(tl);
      pc.insert(pc.end(), {a0, wa, a == b ? 0 : b0, a == b ? 0 : wb});
    }
  pl.n_parts = (int)part_tiles.size();
  int max_t = 0;
  for (const auto &tl : part_tiles) max_t = std::max<int>(max_t, ((int)tl.size() + waves - 1) / waves);
  pl.max_tiles = max_t;
  pl.seg_cols = 16 * 2 * SY_SEGG;
  std::vector<int16_t> h((size_t)pl.n_parts * waves * pl.max_tiles * 2, (int16_t)-1);
  for (int q = 0; q < pl.n_parts; q++) {
    const auto &tl = part_tiles[q];
    const int64_t total = (int64_t)tl.size(), base = total / waves, extra = total % waves;
    int64_t idx = 0;
    for (int w = 0; w < waves; w++) {
      const int64_t take = base + (w < extra ? 1 : 0);
      for (int64_t t = 0; t < take; t++, idx++) {
        const size_t o = (((size_t)q * waves + w) * pl.max_tiles + t) * 2;
        h[o] = (int16_t)tl[idx].first;
        h[o + 1] = (int16_t)tl[idx].second;
      }
    }
  }
  int64_t want = std::max<int64_t>(1, target_wgs / pl.n_parts);
  int64_t by_size = std::max<int64_t>(1, (s->n + min_chunk - 1) / min_chunk);
  pl.n_chunks = std::max<int64_t>(1, std::min(want, by_size));
  pl.chunk = (s->n + pl.n_chunks - 1) / pl.n_chunks;
  pl.chunk = (pl.chunk + kp - 1) / kp * kp;
  pl.n_chunks = (s->n + pl.chunk - 1) / pl.chunk;
  if (pl.d_tiles) hipFree(pl.d_tiles);
  if (pl.d_pcols) hipFree(pl.d_pcols);
  pl.d_pcols = nullptr;
  HIP_TRY(hipMalloc(&pl.d_tiles, h.size() * sizeof(int16_t)));
  HIP_TRY(hipMemcpy(pl.d_tiles, h.data(), h.size() * sizeof(int16_t), hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc(&pl.d_pcols, pc.size() * sizeof(int32_t)));
  HIP_TRY(hipMemcpy(pl.d_pcols, pc.data(), pc.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  return CY3Q_OK;
}

// Lower sub-tiles of the M x M symmetric product, in row-group-major order, dealt evenly in
// contiguous runs to (part, wave) slots of at most max_tiles each. With pad_repeat, a short
// list is padded with repeats of its first tile (for kernels that read a static number of
// slots; the reduction skips the repeats).
// Host rows of `cols` doubles to device rows of pitch `ld` (zero padding after `cols`). A
// pitched copy straight from pageable host memory goes row by row (about 9 us a row), so the
// rows cross in one contiguous copy to a staging buffer and are padded on the device.
int upload_rows_padded(cy3q_solver *s, double *dst, const double *src, int64_t rows,
                       int64_t cols, int64_t ld) {
  double *raw = nullptr;
  ST_TRY(s->alloc(&raw, rows * cols));
  HIP_TRY(hipMemcpyAsync(raw, src, sizeof(double) * rows * cols, hipMemcpyHostToDevice, s->st));
  HIP_TRY(hipMemsetAsync(dst, 0, sizeof(double) * rows * ld, s->st));
  HIP_TRY(hipMemcpy2DAsync(dst, sizeof(double) * ld, raw, sizeof(double) * cols,
                           sizeof(double) * cols, rows, hipMemcpyDeviceToDevice, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  s->release(raw);
  return CY3Q_OK;
}

// The reverse: device rows of pitch `ld` to packed host rows of `cols` doubles.
int download_rows_packed(cy3q_solver *s, double *dst, const double *src, int64_t rows,
                         int64_t cols, int64_t ld) {
  double *raw = nullptr;
  ST_TRY(s->alloc(&raw, rows * cols));
  HIP_TRY(hipMemcpy2DAsync(raw, sizeof(double) * cols, src, sizeof(double) * ld,
                           sizeof(double) * cols, rows, hipMemcpyDeviceToDevice, s->st));
  HIP_TRY(hipMemcpyAsync(dst, raw, sizeof(double) * rows * cols, hipMemcpyDeviceToHost, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  s->release(raw);
  return CY3Q_OK;
}

int build_syrk_plan(cy3q_solver *s, SyrkPlan &pl, int tile, int waves, int cap_tiles,
                    int64_t target_wgs, int64_t min_chunk, int64_t kp, bool pad_repeat = false,
                    int fixed_tiles = 0) {
  pl.m_cols = s->m;
  if (tile == 16 && !pad_repeat && fixed_tiles == 0 && s->m == 416) {
    // k_syrk13_f64: two workgroups of 8 waves per point chunk, 91 tiles of 32 x 32 per chunk
    // partial; target_wgs / 8 chunks (one 107 KB LDS workgroup per CU at a time).
    pl.prog13 = 1;
    pl.tile = 32;
    pl.waves = 8;
    pl.n_groups = 13;
    pl.n_parts = 2;
    pl.max_tiles = 6;
    const int64_t kp13 = 16;
    int64_t want = std::max<int64_t>(1, target_wgs / 8);
    int64_t by_size = std::max<int64_t>(1, (s->n + min_chunk - 1) / min_chunk);
    pl.n_chunks = std::max<int64_t>(1, std::min(want, by_size));
    pl.chunk = (s->n + pl.n_chunks - 1) / pl.n_chunks;
    pl.chunk = (pl.chunk + kp13 - 1) / kp13 * kp13;
    pl.n_chunks = (s->n + pl.chunk - 1) / pl.chunk;
    return CY3Q_OK;
  }
  pl.prog13 = 0;
  pl.tile = tile;
  pl.waves = waves;
  pl.n_groups = (int)(s->m / tile);
  if (tile == 16 && !pad_repeat && fixed_tiles == 0 && s->m > 448)
    return build_syrk_segment_plan(s, pl, target_wgs, min_chunk, kp);
  pl.seg_cols = 0;
  std::vector<std::pair<int, int>> tl;
  for (int r = 0; r < pl.n_groups; r++)
    for (int c = 0; c <= r; c++) tl.push_back({r, c});
  const int64_t total = (int64_t)tl.size();
  pl.n_parts = (int)((total + (int64_t)waves * cap_tiles - 1) / ((int64_t)waves * cap_tiles));
  const int64_t slots = (int64_t)pl.n_parts * waves;
  pl.max_tiles = fixed_tiles > 0 ? fixed_tiles : (int)((total + slots - 1) / slots);
  std::vector<int16_t> h((size_t)slots * pl.max_tiles * 2, (int16_t)-1);
  const int64_t base = total / slots, extra = total % slots;
  int64_t idx = 0;
  for (int64_t sl = 0; sl < slots; sl++) {
    const int64_t take = base + (sl < extra ? 1 : 0);
    for (int64_t t = 0; t < take; t++, idx++) {
      h[(sl * pl.max_tiles + t) * 2] = (int16_t)tl[idx].first;
      h[(sl * pl.max_tiles + t) * 2 + 1] = (int16_t)tl[idx].second;
    }
    if (pad_repeat && take > 0)
      for (int64_t t = take; t < pl.max_tiles; t++) {
        h[(sl * pl.max_tiles + t) * 2] = h[(sl * pl.max_tiles) * 2];
        h[(sl * pl.max_tiles + t) * 2 + 1] = h[(sl * pl.max_tiles) * 2 + 1];
      }
  }
  int64_t want = std::max<int64_t>(1, target_wgs / pl.n_parts);
  int64_t by_size = std::max<int64_t>(1, (s->n + min_chunk - 1) / min_chunk);
  pl.n_chunks = std::max<int64_t>(1, std::min(want, by_size));
  pl.chunk = (s->n + pl.n_chunks - 1) / pl.n_chunks;
  pl.chunk = (pl.chunk + kp - 1) / kp * kp;
  pl.n_chunks = (s->n + pl.chunk - 1) / pl.chunk;
  if (pl.d_tiles) hipFree(pl.d_tiles);
  HIP_TRY(hipMalloc(&pl.d_tiles, h.size() * sizeof(int16_t)));
  HIP_TRY(hipMemcpy(pl.d_tiles, h.data(), h.size() * sizeof(int16_t), hipMemcpyHostToDevice));
  return CY3Q_OK;
}

int64_t syrk_partials_count(const SyrkPlan &pl) {
  return pl.n_chunks * pl.n_parts * pl.waves * (int64_t)pl.max_tiles * pl.tile * pl.tile;
}

// Sections (+ derivatives), weights, log|Omega|^2 and the state buffers for n points in s->pts.
int finish_workspace(cy3q_solver *s) {
  const int64_t n = s->n, m = s->m;
  const int nb = s->nb;
  ST_TRY(s->alloc(&s->w, n));
  ST_TRY(s->alloc(&s->log_om, n));
  ST_TRY(s->alloc(&s->chart_elim, 2 * n));
  ST_TRY(s->alloc(&s->wvalid, n));
  ST_TRY(s->alloc(&s->sum_w, 1));
  ST_TRY(s->alloc(&s->sv, n * m));
  ST_TRY(s->alloc(&s->sd, n * 5 * m));
  ST_TRY(s->alloc(&s->d_mono, s->mono.size()));
  HIP_TRY(hipMemcpyAsync(s->d_mono, s->mono.data(), s->mono.size() * sizeof(uint32_t),
                         hipMemcpyHostToDevice, s->st));
  uint32_t kmax = 0;
  for (uint32_t e : s->mono) kmax = std::max(kmax, e);
  if (kmax > 12) return fail(CY3Q_ERR_ARG, "degree k > 12 is not supported");
  s->stage_begin(STG_SECTIONS);
  launch_weights_logomega(s->st, s->pts, n, s->psi, s->w, s->log_om, s->chart_elim);
  launch_sections(s->st, s->pts, n, s->d_mono, nb, (int)kmax, m, s->sv, s->sd);
  launch_valid_weights(s->st, s->w, n, s->wvalid);
  launch_sum_f64(s->st, s->wvalid, n, s->sum_w);
  s->stage_end(STG_SECTIONS);
  HIP_TRY(hipGetLastError());
  const int64_t nn = (int64_t)nb * nb;
  ST_TRY(s->alloc(&s->g, 2 * nn));
  ST_TRY(s->alloc(&s->wmat, 2 * nn));
  ST_TRY(s->alloc(&s->t, 2 * nn));
  ST_TRY(s->alloc(&s->wnew, 2 * nn));
  ST_TRY(s->alloc(&s->gnew, 2 * nn));
  ST_TRY(s->alloc(&s->bmat, m * m));
  ST_TRY(s->alloc(&s->xx, m * m));
  ST_TRY(s->alloc(&s->coef, n));
  ST_TRY(s->alloc(&s->dpart, (m / 32) * n));
  ST_TRY(s->alloc(&s->resid, 1));
  ST_TRY(s->alloc(&s->sig_out, 8));
  ST_TRY(s->alloc(&s->sig_scratch, sigma_reduce_scratch()));
  ST_TRY(s->alloc(&s->status, 4));
  // status[3] is the completion counter of the Gram kernel's fused residual: it starts at zero
  // and every launch leaves it at zero.
  HIP_TRY(hipMemsetAsync(s->status, 0, 4 * sizeof(int32_t), s->st));
  const int nblk = (nb + 31) / 32;
  ST_TRY(s->alloc(&s->abuf, 2 * nn));
  ST_TRY(s->alloc(&s->rbuf, 2 * nn));
  ST_TRY(s->alloc(&s->einv, (size_t)nblk * 32 * 32 * 2));
  ST_TRY(s->alloc(&s->partial, (size_t)nblk * nblk));
  ST_TRY(s->alloc(&s->scale_buf, 2));
  launch_identity_cplx(s->st, nb, s->g);
  launch_identity_cplx(s->st, nb, s->wmat);
  launch_identity_cplx(s->st, nb, s->t);
  s->bmat_valid = false;
  if (s->m > 1024) return fail(CY3Q_ERR_ARG, "2N > 1024 (k > 9) is not supported");
  ST_TRY(build_syrk_plan(s, s->plan64, 16, 8, syrk_f64_max_tiles(s->m), s->syrk_target, 256, 16));
  s->partials_cap = syrk_partials_count(s->plan64);
  ST_TRY(s->alloc(&s->partials, s->partials_cap));
  HIP_TRY(hipStreamSynchronize(s->st));
  return CY3Q_OK;
}

int make_solver(uint32_t k, double psi, const cy3q_options *opt, cy3q_solver **out,
                cy3q_solver **sp) {
  if (!out) return fail(CY3Q_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (k < 1 || k > 12) return fail(CY3Q_ERR_ARG, "k must be in 1..12");
  cy3q_solver *s = new cy3q_solver();
  if (opt)
    s->opt = *opt;
  else
    cy3q_options_default(&s->opt);
  s->device = s->opt.device;
  hipError_t e = hipSetDevice(s->device);
  if (e != hipSuccess) {
    delete s;
    return fail(CY3Q_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
  }
  e = hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking);
  if (e != hipSuccess) {
    s->st = nullptr;
    delete s;
    return fail(CY3Q_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  s->k = k;
  s->psi = psi;
  s->precision = s->opt.precision;
  s->mono = monomials(k, s->opt.quotient);
  s->nb = (int)(s->mono.size() / 5);
  s->m = padded_cols(s->nb);
  *sp = s;
  return CY3Q_OK;
}

// SZ sampling on the GPU in batches of lines until n points are accepted or the reference's
// max_tries lines have been drawn (4x the target lines for Fermat, 8x for Dwork).
int sample_points(cy3q_solver *s, int64_t n_target, uint64_t seed, int dwork) {
  const int64_t lines_target = (n_target + 4) / 5;
  const int64_t max_tries = lines_target * (dwork ? 8 : 4);
  int64_t batch = std::max<int64_t>(1024, lines_target + lines_target / 32 + 64);
  batch = std::min(batch, max_tries);
  double *cand = nullptr;
  int32_t *acc = nullptr;
  int64_t *pos = nullptr, *bsum = nullptr, *total = nullptr;
  ST_TRY(s->alloc(&s->pts, n_target * 10));
  ST_TRY(s->alloc(&cand, batch * 50));
  ST_TRY(s->alloc(&acc, batch * 5));
  ST_TRY(s->alloc(&pos, batch * 5));
  ST_TRY(s->alloc(&bsum, (batch * 5 + 1023) / 1024 + 1));
  ST_TRY(s->alloc(&total, 1));
  int64_t have = 0, line0 = 0;
  s->stage_begin(STG_SAMPLE);
  while (have < n_target && line0 < max_tries) {
    int64_t nl = std::min(batch, max_tries - line0);
    launch_sz_lines(s->st, seed, line0, nl, dwork, s->psi, s->opt.dk_iters, s->opt.newton_polish,
                    cand, acc);
    launch_scan_accept(s->st, acc, nl * 5, pos, bsum, nullptr, total);
    launch_compact_points(s->st, cand, acc, pos, nl * 5, have, n_target, s->pts);
    HIP_TRY(hipGetLastError());
    int64_t got = 0;
    HIP_TRY(hipMemcpyAsync(&got, total, sizeof(int64_t), hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    have = std::min(n_target, have + got);
    line0 += nl;
  }
  s->stage_end(STG_SAMPLE);
  s->release(cand);
  s->release(acc);
  s->release(pos);
  s->release(bsum);
  s->release(total);
  s->n = have;
  return CY3Q_OK;
}

// sample_quintic_points / sample_deformed_quintic_points (quintic.rs:3448-3482, 3568-3595):
// tries in order until n_target points are accepted or the reference's try cap (4 n, 8 n for the
// deformed form) is reached. Each batch of tries is projected in parallel, then the accepted
// points are compacted in try order.
int sample_points_newton(cy3q_solver *s, int64_t n_target, uint64_t seed, double tol, int dwork) {
  const int64_t max_tries = n_target * (dwork ? 8 : 4);
  int64_t batch = std::max<int64_t>(1024, n_target + n_target / 16 + 64);
  batch = std::min(batch, max_tries);
  double *cand = nullptr;
  int32_t *acc = nullptr;
  int64_t *pos = nullptr, *bsum = nullptr, *total = nullptr;
  ST_TRY(s->alloc(&s->pts, n_target * 10));
  ST_TRY(s->alloc(&cand, batch * 10));
  ST_TRY(s->alloc(&acc, batch));
  ST_TRY(s->alloc(&pos, batch));
  ST_TRY(s->alloc(&bsum, (batch + 1023) / 1024 + 1));
  ST_TRY(s->alloc(&total, 1));
  int64_t have = 0, try0 = 0;
  s->stage_begin(STG_SAMPLE);
  while (have < n_target && try0 < max_tries) {
    const int64_t nt = std::min(batch, max_tries - try0);
    launch_newton_tries(s->st, seed, try0, nt, s->psi, tol, cand, acc);
    launch_scan_accept(s->st, acc, nt, pos, bsum, nullptr, total);
    launch_compact_points(s->st, cand, acc, pos, nt, have, n_target, s->pts);
    HIP_TRY(hipGetLastError());
    int64_t got = 0;
    HIP_TRY(hipMemcpyAsync(&got, total, sizeof(int64_t), hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    have = std::min(n_target, have + got);
    try0 += nt;
  }
  s->stage_end(STG_SAMPLE);
  s->release(cand);
  s->release(acc);
  s->release(pos);
  s->release(bsum);
  s->release(total);
  s->n = have;
  return CY3Q_OK;
}

// Section rows and derivative rows of s replaced by conj(V) times them, V lower triangular
// (complex N x N on the device): the basis change of the FS-Gram orthonormalisation.
int rebase_sections(cy3q_solver *s, const double *v_cplx) {
  const int64_t n = s->n, m = s->m;
  launch_build_bmat(s->st, v_cplx, s->nb, m, 1, s->bmat);
  s->bmat_valid = false;
  double *sv2 = nullptr, *sd2 = nullptr;
  ST_TRY(s->alloc(&sv2, n * m));
  launch_trigemm_f64(s->st, s->sv, n, m, s->bmat, sv2, 0, 0);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->st));
  s->release(s->sv);
  s->sv = sv2;
  ST_TRY(s->alloc(&sd2, n * 5 * m));
  launch_trigemm_f64(s->st, s->sd, n * 5, m, s->bmat, sd2, 0, 0);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->st));
  s->release(s->sd);
  s->sd = sd2;
  return CY3Q_OK;
}

// T-operator pieces -------------------------------------------------------------------------
// B(W) for the triangular contraction, rebuilt lazily after W changes.
void ensure_bmat(cy3q_solver *s) {
  if (s->bmat_valid) return;
  launch_build_bmat(s->st, s->wmat, s->nb, s->m, 0, s->bmat);
  s->bmat_valid = true;
}

// One Donaldson step in FP64; residual left on the device in s->resid.
int t_step_fp64(cy3q_solver *s) {
  const int64_t n = s->n, m = s->m;
  ensure_bmat(s);
  s->stage_begin(STG_DPASS);
  launch_dpass_f64(s->st, s->sv, n, m, s->bmat, s->w, s->dpart, s->coef);
  s->stage_end(STG_DPASS);
  s->stage_begin(STG_TPASS);
  launch_syrk_f64(s->st, s->plan64, s->sv, s->coef, n, s->partials);
  s->stage_end(STG_TPASS);
  s->stage_begin(STG_TREDUCE);
  launch_syrk_reduce(s->st, s->plan64, s->partials, s->xx);
  s->stage_end(STG_TREDUCE);
  return CY3Q_OK;
}

// T from xx, then W = L^-1 (T = L L^H), G = W^H W and the residual; W becomes the state.
int finish_step(cy3q_solver *s, double scale, const double *dscale) {
  const int nb = s->nb;
  s->stage_begin(STG_NXN);
  launch_t_from_xx(s->st, s->xx, s->m, nb, s->sum_w, scale, dscale, 1, s->scale_buf, s->t,
                   s->abuf, s->rbuf);
  launch_inverse_cholesky(s->st, s->abuf, s->rbuf, s->einv, s->wnew, nb, s->status);
  launch_gram_residual(s->st, s->wnew, nb, s->g, s->status, s->partial, s->resid,
                       reinterpret_cast<unsigned *>(s->status + 3));
  std::swap(s->wmat, s->wnew);
  s->bmat_valid = false;
  s->stage_end(STG_NXN);
  HIP_TRY(hipGetLastError());
  return CY3Q_OK;
}

// Split path set-up: the static fp16 hi/lo operand planes of the (orthonormalised) sections,
// the D-pass column-group lists and the T-pass plan.
int ensure_split16(cy3q_solver *s) {
  const int64_t n = s->n, m = s->m;
  const int n_cg = (int)(m / 16);
  const int dw = dpass_split16_waves(), dmc = dpass_split16_max_cg();
  if (!dpass_split16_supported(m) || m > tpass_max_cols())
    return fail(CY3Q_ERR_ARG,
                "split16 precision supports 2N <= 448, the quotient bases of k = 7, 8 (2N = 630, "
                "920) and the raw k = 8 basis (2N = 990); use FP64");
  if (!s->xd) {
    const int64_t n_pad = (n + 127) / 128 * 128;
    ST_TRY(s->alloc(&s->xd, n_pad * m * 2));
    ST_TRY(s->alloc(&s->xt, n_pad * m));
    ST_TRY(s->alloc(&s->row_ex, n));
    ST_TRY(s->alloc(&s->rs2, n));
    // T-pass coefficients, zero past n: the last stage of the last chunk reads up to a
    // multiple of 16 points (<= n_pad).
    ST_TRY(s->alloc(&s->u, n_pad));
    HIP_TRY(hipMemsetAsync(s->u, 0, n_pad * sizeof(float), s->st));
    ST_TRY(s->alloc(&s->dsum, n));
    ST_TRY(s->alloc(&s->block_max, (n + 255) / 256 + 1));
    ST_TRY(s->alloc(&s->bscale, 2));
    ST_TRY(s->alloc(&s->tscale, 1));
    ST_TRY(s->alloc(&s->zscale, 1));
    ST_TRY(s->alloc(&s->bfrag, ((size_t)(m / 32) * n_cg + 1) * 2 * 64));
    // Column groups to waves round-robin: at k-step ks the active groups are cg >= 2 ks, a
    // contiguous range, which round-robin spreads evenly over the waves (each k-step ends at a
    // barrier, so the per-step balance is what counts). Each wave's list is ascending and
    // right-aligned, unused slots first (-1), so its active groups are always its last slots.
    // Past 448 columns the groups are split into column ranges (one D-pass launch each, whose
    // partial sums are then added): one such table per range, over the range's groups.
    const std::vector<std::pair<int, int>> ranges = dpass_split16_ranges(m);
    std::vector<int16_t> cl(ranges.size() * dw * dmc, (int16_t)-1);
    for (size_t r = 0; r < ranges.size(); r++)
      for (int w = 0; w < dw; w++) {
        const int cnt = (ranges[r].second - w + dw - 1) / dw;
        for (int j = 0; j < cnt; j++)
          cl[(r * dw + w) * dmc + (dmc - cnt) + j] = (int16_t)(ranges[r].first + w + j * dw);
      }
    if (ranges.size() > 1) ST_TRY(s->alloc(&s->dsum_parts, (int64_t)ranges.size() * n));
    ST_TRY(s->alloc(&s->cglist, cl.size()));
    HIP_TRY(hipMemcpyAsync(s->cglist, cl.data(), cl.size() * sizeof(int16_t),
                           hipMemcpyHostToDevice, s->st));
    // T pass: one workgroup per CU, each streaming one chunk of a multiple of 16 points.
    int n_cu = 0;
    HIP_TRY(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, s->device));
    n_cu = std::max(n_cu, 1);
    s->n_cu = n_cu;
    const int64_t slots = (int64_t)n_cu * s->tp_mult;
    s->tp_chunk = std::max<int64_t>((n + slots * 16 - 1) / (slots * 16), 1) * 16;
    s->tp_chunks = (n + s->tp_chunk - 1) / s->tp_chunk;
    ST_TRY(s->alloc(&s->partials32, tpass_partials(m, s->tp_chunks)));
    ST_TRY(s->alloc(&s->tseg, tpass_seg_count(m, s->tp_chunks)));
    HIP_TRY(hipStreamSynchronize(s->st));
  }
  if (!s->x32_valid) {
    const int64_t n_pad = (n + 127) / 128 * 128;
    launch_split_planes(s->st, s->sv, n, n_pad, m, s->row_ex, s->rs2, s->xd, s->xt);
    s->x32_valid = true;
  }
  return CY3Q_OK;
}

int t_step_split16(cy3q_solver *s) {
  const int64_t n = s->n, m = s->m;
  ST_TRY(ensure_split16(s));
  if (!s->bfrag_valid) {
    launch_build_bfrag16(s->st, s->wmat, s->nb, m, s->bfrag, s->bscale);
    s->bfrag_valid = true;
  }
  s->stage_begin(STG_DPASS);
  launch_dpass_split16(s->st, (const half8_t *)s->xd, n, m, s->bfrag, s->cglist, s->n_cu,
                       s->dsum, s->dsum_parts);
  s->stage_end(STG_DPASS);
  s->stage_begin(STG_COEFF);
  launch_coeffs_split16(s->st, s->w, s->dsum, s->rs2, s->bscale, n, s->u, s->block_max,
                        s->zscale, s->tscale);
  s->stage_end(STG_COEFF);
  s->stage_begin(STG_TPASS);
  launch_tpass(s->st, s->xt, s->u, n, (n + 127) / 128 * 128, m, s->tp_chunk, s->tp_chunks,
               s->partials32);
  s->stage_end(STG_TPASS);
  s->stage_begin(STG_TREDUCE);
  launch_tpass_reduce(s->st, s->partials32, m, s->tp_chunks, s->tseg, s->xx);
  s->stage_end(STG_TREDUCE);
  return CY3Q_OK;
}

// Refinement sweep of kind CY3Q_REFINE_T_FP64: the split D pass, then the FP64 T pass on
// FP64 coefficients. The D pass's rounding is independent from point to point and averages
// out in T; the T pass's does not (it sets the split path's fixed point).
int t_step_refine_t64(cy3q_solver *s) {
  const int64_t n = s->n, m = s->m;
  ST_TRY(ensure_split16(s));
  if (!s->bfrag_valid) {
    launch_build_bfrag16(s->st, s->wmat, s->nb, m, s->bfrag, s->bscale);
    s->bfrag_valid = true;
  }
  s->stage_begin(STG_DPASS);
  launch_dpass_split16(s->st, (const half8_t *)s->xd, n, m, s->bfrag, s->cglist, s->n_cu,
                       s->dsum, s->dsum_parts);
  s->stage_end(STG_DPASS);
  s->stage_begin(STG_COEFF);
  launch_coeffs_f64_from_split(s->st, s->w, s->dsum, s->rs2, s->bscale, n, s->coef);
  s->stage_end(STG_COEFF);
  s->stage_begin(STG_TPASS);
  launch_syrk_f64(s->st, s->plan64, s->sv, s->coef, n, s->partials);
  s->stage_end(STG_TPASS);
  s->stage_begin(STG_TREDUCE);
  launch_syrk_reduce(s->st, s->plan64, s->partials, s->xx);
  s->stage_end(STG_TREDUCE);
  return CY3Q_OK;
}

// Step codes: CY3Q_PREC_FP64, CY3Q_PREC_SPLIT16, or PREC_REFINE_T64 (the sweep above).
constexpr int PREC_REFINE_T64 = 100;

// The step code of one refinement sweep of s.
int refine_step_code(const cy3q_solver *s) {
  return s->refine_kind == CY3Q_REFINE_T_FP64 ? PREC_REFINE_T64 : CY3Q_PREC_FP64;
}

int step_in(cy3q_solver *s, int precision) {
  if (precision == CY3Q_PREC_FP64) {
    ST_TRY(t_step_fp64(s));
    ST_TRY(finish_step(s, 1.0, nullptr));
  } else if (precision == PREC_REFINE_T64) {
    ST_TRY(t_step_refine_t64(s));
    ST_TRY(finish_step(s, 1.0, nullptr));
  } else {
    ST_TRY(t_step_split16(s));
    ST_TRY(finish_step(s, 1.0, s->tscale));
  }
  s->bfrag_valid = false;
  return CY3Q_OK;
}

int one_step(cy3q_solver *s) { return step_in(s, s->precision); }

// W from an arbitrary Hermitian PD G: W1 = chol(G)^-1, G^-1 = W1^H W1, W = chol(G^-1)^-1,
// so that W^H W = G. *pd = false when either factorisation finds G not positive definite.
int w_from_g_pd(cy3q_solver *s, bool *pd) {
  const int nb = s->nb;
  launch_elim_init(s->st, s->g, nb, s->abuf, s->rbuf);
  launch_inverse_cholesky(s->st, s->abuf, s->rbuf, s->einv, s->wnew, nb, s->status);
  launch_gram_residual(s->st, s->wnew, nb, s->gnew, s->status, nullptr, nullptr);  // G^-1
  launch_elim_init(s->st, s->gnew, nb, s->abuf, s->rbuf);
  launch_inverse_cholesky(s->st, s->abuf, s->rbuf, s->einv, s->wmat, nb, s->status + 1);
  s->bmat_valid = false;
  s->bfrag_valid = false;
  int32_t st2[2] = {0, 0};
  HIP_TRY(hipMemcpyAsync(st2, s->status, sizeof(st2), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  *pd = st2[0] == 0 && st2[1] == 0;
  return CY3Q_OK;
}

int w_from_g(cy3q_solver *s) {
  bool pd = false;
  ST_TRY(w_from_g_pd(s, &pd));
  if (!pd) return fail(CY3Q_ERR_ARG, "h_block is not positive definite");
  return CY3Q_OK;
}

// Contracted rows y0 = W s and y_i = W ds_i of every point stay resident when they fit in this
// many bytes; otherwise they are recomputed per chunk of the sigma and gradient passes.
constexpr double kRowsResidentBytes = 48e9;
constexpr int64_t kRowChunk = 65536;

int ensure_sigma_ws(cy3q_solver *s) {
  if (!s->eta_buf) ST_TRY(s->alloc(&s->eta_buf, s->n));
  if (!s->y0all) {
    const double bytes = 8.0 * 6.0 * (double)s->n * (double)s->m;
    s->rows_resident = bytes <= kRowsResidentBytes;
    const int64_t rows = s->rows_resident ? s->n : std::min<int64_t>(s->n, kRowChunk);
    ST_TRY(s->alloc(&s->y0all, rows * s->m));
    ST_TRY(s->alloc(&s->ydall, rows * 5 * s->m));
  }
  return CY3Q_OK;
}

// eta for every point (device, [n]) and the two-pass reduction into s->sig_out. With resident
// rows, y0 and yd hold every point's contracted rows afterwards.
int sigma_device(cy3q_solver *s, double *eta) {
  const int64_t n = s->n, m = s->m;
  if (!eta) return fail(CY3Q_ERR_STATE, "sigma: no eta buffer");
  ensure_bmat(s);
  ST_TRY(ensure_sigma_ws(s));
  if (s->rows_resident) {
    launch_trigemm_f64(s->st, s->sv, n, m, s->bmat, s->y0all, 0, 0);
    launch_trigemm_f64(s->st, s->sd, n * 5, m, s->bmat, s->ydall, 0, 0);
    launch_sigma_eta(s->st, s->pts, s->y0all, s->ydall, s->log_om, n, m, s->psi, eta);
  } else {
    for (int64_t p0 = 0; p0 < n; p0 += kRowChunk) {
      const int64_t cn = std::min(kRowChunk, n - p0);
      launch_trigemm_f64(s->st, s->sv + p0 * m, cn, m, s->bmat, s->y0all, 0, 0);
      launch_trigemm_f64(s->st, s->sd + p0 * 5 * m, cn * 5, m, s->bmat, s->ydall, 0, 0);
      launch_sigma_eta(s->st, s->pts + 10 * p0, s->y0all, s->ydall, s->log_om + p0, cn, m, s->psi,
                       eta + p0);
    }
  }
  launch_sigma_reduce(s->st, eta, s->w, n, s->sig_scratch, s->sig_out);
  HIP_TRY(hipGetLastError());
  return CY3Q_OK;
}

int sigma_impl(cy3q_solver *s, double *out5, double *eta_host) {
  const int64_t n = s->n;
  s->stage_begin(STG_SIGMA);
  ST_TRY(ensure_sigma_ws(s));
  double *eta = s->eta_buf;
  ST_TRY(sigma_device(s, eta));
  s->stage_end(STG_SIGMA);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out5, s->sig_out, 5 * sizeof(double), hipMemcpyDeviceToHost, s->st));
  if (eta_host)
    HIP_TRY(hipMemcpyAsync(eta_host, eta, n * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  s->collect();
  return CY3Q_OK;
}

// sigma_L2^2 and its gradient (sigma_squared_and_gradient, quintic.rs:2714-2860) into
// s->grad_dev [2 nb^2]; out6 = the sigma reduction (sigma_L1, kappa, sigma_L2^2, sum w, count,
// gamma numerator sum). Pass 1 is the sigma pass; pass 2 recomputes the contracted rows per
// chunk, builds the adjoint rows and contracts them with the section and derivative rows.
int sigma_grad_impl(cy3q_solver *s, double *out6) {
  const int64_t n = s->n, m = s->m;
  const int nb = s->nb;
  if (!s->grad_dev) {
    ST_TRY(s->alloc(&s->grad_dev, 2 * (int64_t)nb * nb));
    ST_TRY(s->alloc(&s->adam_m, 2 * (int64_t)nb * nb));
    ST_TRY(s->alloc(&s->adam_v, 2 * (int64_t)nb * nb));
  }
  s->stage_begin(STG_SIGMA);
  ST_TRY(ensure_sigma_ws(s));  // allocates eta_buf before it is passed on
  ST_TRY(sigma_device(s, s->eta_buf));
  const int64_t chunk = std::min<int64_t>(n, kRowChunk);
  if (!s->g_phi) {
    s->g_chunk = chunk;
    s->g_groups = gemm_tn_groups(5 * chunk, m);
    ST_TRY(s->alloc(&s->g_phi, chunk * 5 * m));
    ST_TRY(s->alloc(&s->g_vs, chunk * m));
    ST_TRY(s->alloc(&s->g_part, s->g_groups * m * m));
    ST_TRY(s->alloc(&s->g_cm, m * m));
  }
  double *phi = s->g_phi, *vs = s->g_vs, *part = s->g_part, *cm = s->g_cm;
  const int64_t groups = s->g_groups;
  HIP_TRY(hipMemsetAsync(cm, 0, sizeof(double) * m * m, s->st));
  for (int64_t p0 = 0; p0 < n; p0 += chunk) {
    const int64_t cn = std::min(chunk, n - p0);
    const double *y0 = s->y0all, *yd = s->ydall;
    if (s->rows_resident) {
      y0 += p0 * m;
      yd += p0 * 5 * m;
    } else {
      launch_trigemm_f64(s->st, s->sv + p0 * m, cn, m, s->bmat, s->y0all, 0, 0);
      launch_trigemm_f64(s->st, s->sd + p0 * 5 * m, cn * 5, m, s->bmat, s->ydall, 0, 0);
    }
    launch_grad_rows(s->st, s->pts + 10 * p0, y0, yd, s->log_om + p0, s->w + p0, s->sv + p0 * m,
                     s->sd + p0 * 5 * m, cn, m, s->psi, s->sig_out, phi, vs);
    launch_gemm_tn_acc(s->st, s->sd + p0 * 5 * m, phi, cn * 5, m,
                       std::min(groups, gemm_tn_groups(5 * cn, m)), part, cm);
    launch_gemm_tn_acc(s->st, s->sv + p0 * m, vs, cn, m, std::min(groups, gemm_tn_groups(cn, m)),
                       part, cm);
  }
  launch_grad_assemble(s->st, cm, nb, m, s->sig_out, s->grad_dev);
  s->stage_end(STG_SIGMA);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out6, s->sig_out, 6 * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  s->collect();
  return CY3Q_OK;
}

}  // namespace

namespace {
// donaldson_iteration_quintic(_weighted) (quintic.rs:4140-4331): h <- T(h) with no inversion.
// The FP64 D pass and T pass give T = sum c_p conj(s_a) s_b with c_p = w_p / K_p over the valid
// weights (weighted), or c_p = 1 / K_p over all points with sum 1 / n_points (unweighted); then
// the trace normalisation, G = T, the block residual and W from the new G.
int legacy_step(cy3q_solver *s, int weighted, double *residual) {
  const int64_t n = s->n, m = s->m;
  if (!weighted && !s->ones) {
    ST_TRY(s->alloc(&s->ones, n));
    ST_TRY(s->alloc(&s->n_as_w, 1));
    launch_fill_f64(s->st, s->ones, n, 1.0);
    launch_sum_f64(s->st, s->ones, n, s->n_as_w);
  }
  const double *wv = weighted ? s->w : s->ones;
  ensure_bmat(s);
  launch_dpass_f64(s->st, s->sv, n, m, s->bmat, wv, s->dpart, s->coef);
  launch_syrk_f64(s->st, s->plan64, s->sv, s->coef, n, s->partials);
  launch_syrk_reduce(s->st, s->plan64, s->partials, s->xx);
  launch_t_from_xx(s->st, s->xx, m, s->nb, weighted ? s->sum_w : s->n_as_w, 1.0, nullptr, 1,
                   s->scale_buf, s->t, s->abuf, s->rbuf);
  launch_legacy_replace(s->st, s->t, s->nb, s->g, s->resid);
  HIP_TRY(hipGetLastError());
  ST_TRY(w_from_g(s));
  if (residual) {
    HIP_TRY(hipMemcpyAsync(residual, s->resid, sizeof(double), hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
  }
  return CY3Q_OK;
}
}  // namespace

namespace cy3q {
int set_error(int code, const std::string &msg) { return fail(code, msg); }

int solver_view(cy3q_solver *s, SolverView *v) {
  if (!s || !v) return fail(CY3Q_ERR_ARG, "null solver");
  HIP_TRY(hipSetDevice(s->device));
  ensure_bmat(s);
  v->st = s->st;
  v->device = s->device;
  v->n = s->n;
  v->m = s->m;
  v->nb = s->nb;
  v->psi = s->psi;
  v->pts = s->pts;
  v->w = s->w;
  v->log_om = s->log_om;
  v->sv = s->sv;
  v->sd = s->sd;
  v->bmat = s->bmat;
  v->galerkin_precision = s->galerkin_precision;
  return CY3Q_OK;
}

hipStream_t solver_stream(cy3q_solver *s) { return s->st; }

// Everything a captured solve needs, allocated and built now: the split path's static operand
// planes and plan, and a residual slot per iteration.
int solver_prepare_enqueue(cy3q_solver *s, int64_t iters) {
  HIP_TRY(hipSetDevice(s->device));
  if (s->precision == CY3Q_PREC_SPLIT16) ST_TRY(ensure_split16(s));
  if (s->batch_res_cap < iters) {
    if (s->batch_res) s->release(s->batch_res);
    s->batch_res = nullptr;
    ST_TRY(s->alloc(&s->batch_res, std::max<int64_t>(iters, 1)));
    s->batch_res_cap = std::max<int64_t>(iters, 1);
  }
  HIP_TRY(hipStreamSynchronize(s->st));
  return CY3Q_OK;
}

// The iterations of donaldson_solve(iters, tol <= 0) as kernel launches only: no allocation,
// copy or synchronisation, so the sequence can be captured into a graph.
int solver_enqueue_solve(cy3q_solver *s, int64_t iters) {
  if (s->batch_res_cap < iters) return fail(CY3Q_ERR_STATE, "solver_prepare_enqueue first");
  const int64_t n_ref = s->precision == CY3Q_PREC_FP64 ? 0 : std::min(s->refine, iters);
  for (int64_t it = 0; it < iters; it++) {
    double *keep = s->resid;
    s->resid = s->batch_res + it;
    const int prec = it >= iters - n_ref ? refine_step_code(s) : s->precision;
    int rc = step_in(s, prec);
    s->resid = keep;
    if (rc != CY3Q_OK) return rc;
  }
  return CY3Q_OK;
}

// h = I as kernel launches (G, W and T to the identity).
int solver_reset_identity(cy3q_solver *s) {
  launch_identity_cplx(s->st, s->nb, s->g);
  launch_identity_cplx(s->st, s->nb, s->wmat);
  launch_identity_cplx(s->st, s->nb, s->t);
  s->bmat_valid = false;
  s->bfrag_valid = false;
  HIP_TRY(hipGetLastError());
  return CY3Q_OK;
}

int solver_last_residual(cy3q_solver *s, int64_t iters, double *out) {
  if (iters < 1 || s->batch_res_cap < iters) return fail(CY3Q_ERR_STATE, "no captured solve");
  HIP_TRY(hipMemcpyAsync(out, s->batch_res + iters - 1, sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  return CY3Q_OK;
}
}  // namespace cy3q

// ============================================================================================
// C ABI
// ============================================================================================
extern "C" {

void cy3q_options_default(cy3q_options *opt) {
  if (!opt) return;
  opt->quotient = 0;
  opt->precision = CY3Q_PREC_FP64;
  opt->device = 0;
  opt->dk_iters = 50;
  opt->newton_polish = 0;
  opt->newton_tol = 1e-12;
}

const char *cy3q_version(void) { return "cy3q 0.1 (gfx950; clang " __clang_version__ ")"; }
const char *cy3q_last_error(void) { return g_last_error.c_str(); }

int cy3q_n_sections(uint32_t k, int quotient, int64_t *n) {
  if (!n) return fail(CY3Q_ERR_ARG, "n is NULL");
  *n = (int64_t)(monomials(k, quotient).size() / 5);
  (void)n_sections;
  return CY3Q_OK;
}

int cy3q_monomials(uint32_t k, int quotient, uint32_t *out, int64_t cap, int64_t *n) {
  std::vector<uint32_t> mo = monomials(k, quotient);
  int64_t cnt = (int64_t)(mo.size() / 5);
  if (n) *n = cnt;
  if (out) {
    if (cap < cnt) return fail(CY3Q_ERR_ARG, "output capacity too small");
    std::memcpy(out, mo.data(), mo.size() * sizeof(uint32_t));
  }
  return CY3Q_OK;
}

int cy3q_solver_new(uint32_t k, int64_t n_points, uint64_t seed, double psi, int sampler,
                    const cy3q_options *opt, cy3q_solver **out) {
  if (n_points <= 0) return fail(CY3Q_ERR_ARG, "n_points must be > 0");
  if (sampler != CY3Q_SAMPLER_SZ && sampler != CY3Q_SAMPLER_SZ_DWORK &&
      sampler != CY3Q_SAMPLER_NEWTON && sampler != CY3Q_SAMPLER_NEWTON_DWORK)
    return fail(CY3Q_ERR_ARG, "unknown sampler");
  if ((sampler == CY3Q_SAMPLER_SZ || sampler == CY3Q_SAMPLER_NEWTON) && psi != 0.0)
    return fail(CY3Q_ERR_ARG, "the Fermat samplers take psi = 0; use a Dwork sampler");
  const bool newton = sampler == CY3Q_SAMPLER_NEWTON || sampler == CY3Q_SAMPLER_NEWTON_DWORK;
  const double tol = opt ? opt->newton_tol : 1e-12;
  if (newton && !(tol > 0.0 && std::isfinite(tol)))
    return fail(CY3Q_ERR_ARG, "newton_tol must be finite and > 0");
  cy3q_solver *s = nullptr;
  ST_TRY(make_solver(k, psi, opt, out, &s));
  const int dwork = sampler == CY3Q_SAMPLER_SZ_DWORK || sampler == CY3Q_SAMPLER_NEWTON_DWORK;
  int rc = newton ? sample_points_newton(s, n_points, seed, tol, dwork)
                  : sample_points(s, n_points, seed, dwork);
  if (rc == CY3Q_OK && s->n == 0) rc = fail(CY3Q_ERR_EMPTY, "no point was accepted");
  if (rc == CY3Q_OK) rc = finish_workspace(s);
  if (rc != CY3Q_OK) {
    delete s;
    return rc;
  }
  *out = s;
  return CY3Q_OK;
}

int cy3q_solver_new_from_points(uint32_t k, const double *points, int64_t n, double psi,
                                const cy3q_options *opt, cy3q_solver **out) {
  if (!points || n <= 0) return fail(CY3Q_ERR_EMPTY, "no points");
  cy3q_solver *s = nullptr;
  ST_TRY(make_solver(k, psi, opt, out, &s));
  s->n = n;
  int rc = s->alloc(&s->pts, n * 10);
  if (rc == CY3Q_OK) {
    hipError_t e = hipMemcpy(s->pts, points, sizeof(double) * 10 * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = fail(CY3Q_ERR_HIP, hipGetErrorString(e));
  }
  if (rc == CY3Q_OK) rc = finish_workspace(s);
  if (rc != CY3Q_OK) {
    delete s;
    return rc;
  }
  *out = s;
  return CY3Q_OK;
}

void cy3q_solver_free(cy3q_solver *s) { delete s; }

int cy3q_solver_info(const cy3q_solver *s, int64_t *n_points, int64_t *n_basis, uint32_t *k,
                     double *psi) {
  if (!s) return fail(CY3Q_ERR_ARG, "NULL solver");
  if (n_points) *n_points = s->n;
  if (n_basis) *n_basis = s->nb;
  if (k) *k = s->k;
  if (psi) *psi = s->psi;
  return CY3Q_OK;
}

int cy3q_solver_set_precision(cy3q_solver *s, int precision) {
  if (!s) return fail(CY3Q_ERR_ARG, "NULL solver");
  if (precision != CY3Q_PREC_FP64 && precision != CY3Q_PREC_SPLIT16)
    return fail(CY3Q_ERR_ARG, "unknown precision");
  s->precision = precision;
  return CY3Q_OK;
}

int cy3q_solver_orthonormalise_fs_gram(cy3q_solver *s) {
  if (!s) return fail(CY3Q_ERR_ARG, "NULL solver");
  HIP_TRY(hipSetDevice(s->device));
  const int64_t n = s->n, m = s->m;
  const int nb = s->nb;
  s->stage_begin(STG_FSGRAM);
  // Gram = sum_p w_p conj(s_p) s_p^T / sum w: the weighted SYRK with c = valid w, no trace
  // normalisation (compute_fs_gram_matrix, quintic.rs:2957-2995).
  launch_syrk_f64(s->st, s->plan64, s->sv, s->wvalid, n, s->partials);
  launch_syrk_reduce(s->st, s->plan64, s->partials, s->xx);
  launch_t_from_xx(s->st, s->xx, m, nb, s->sum_w, 1.0, nullptr, 0, s->scale_buf, nullptr, s->abuf,
                   s->rbuf);
  // Gram = L L^H; W = L^-1 comes out of the same elimination.
  launch_inverse_cholesky(s->st, s->abuf, s->rbuf, s->einv, s->wnew, nb, s->status);
  int32_t stt = 0;
  HIP_TRY(hipMemcpyAsync(&stt, s->status, sizeof(int32_t), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  if (stt != 0) return fail(CY3Q_ERR_NOT_PD, "FS Gram matrix not positive-definite");
  // s' = conj(L)^-1 s = conj(W) s for the section rows and all derivative rows
  // (apply_inverse_l_to_sections/derivs, quintic.rs:3095-3133), as triangular contractions.
  ST_TRY(rebase_sections(s, s->wnew));
  // Keep the accumulated basis change for evaluations at other points: a second FS-Gram
  // composes, conj(W2) conj(W1) = conj(W2 W1).
  const int64_t nn = (int64_t)nb * nb;
  if (!s->wfs) {
    ST_TRY(s->alloc(&s->wfs, 2 * nn));
    HIP_TRY(hipMemcpyAsync(s->wfs, s->wnew, sizeof(double) * 2 * nn, hipMemcpyDeviceToDevice,
                           s->st));
  } else {
    launch_zgemm(s->st, (const double2 *)s->wnew, (const double2 *)s->wfs, (double2 *)s->gnew, nb,
                 nb, nb, nb, nb, nb, 0, 0);
    HIP_TRY(hipMemcpyAsync(s->wfs, s->gnew, sizeof(double) * 2 * nn, hipMemcpyDeviceToDevice,
                           s->st));
  }
  // h <- I in the new basis.
  launch_identity_cplx(s->st, nb, s->g);
  launch_identity_cplx(s->st, nb, s->wmat);
  s->bmat_valid = false;
  s->bfrag_valid = false;
  s->x32_valid = false;
  s->stage_end(STG_FSGRAM);
  HIP_TRY(hipStreamSynchronize(s->st));
  s->collect();
  return CY3Q_OK;
}

int cy3q_solver_donaldson_step(cy3q_solver *s, double *residual) {
  if (!s) return fail(CY3Q_ERR_ARG, "NULL solver");
  HIP_TRY(hipSetDevice(s->device));
  ST_TRY(one_step(s));
  double r = 0.0;
  HIP_TRY(hipMemcpyAsync(&r, s->resid, sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  s->collect();
  if (residual) *residual = r;
  return CY3Q_OK;
}

int cy3q_solver_set_refine(cy3q_solver *s, int64_t sweeps) {
  if (!s) return fail(CY3Q_ERR_ARG, "NULL solver");
  if (sweeps < 0) return fail(CY3Q_ERR_ARG, "refinement sweeps must be >= 0");
  s->refine = sweeps;
  return CY3Q_OK;
}

int cy3q_solver_set_refine_kind(cy3q_solver *s, int kind) {
  if (!s) return fail(CY3Q_ERR_ARG, "NULL solver");
  if (kind != CY3Q_REFINE_STEP_FP64 && kind != CY3Q_REFINE_T_FP64)
    return fail(CY3Q_ERR_ARG, "kind must be CY3Q_REFINE_STEP_FP64 or CY3Q_REFINE_T_FP64");
  s->refine_kind = kind;
  return CY3Q_OK;
}

int cy3q_solver_donaldson_step_legacy(cy3q_solver *s, int weighted, double *residual) {
  if (!s) return fail(CY3Q_ERR_ARG, "NULL solver");
  HIP_TRY(hipSetDevice(s->device));
  return legacy_step(s, weighted, residual);
}

int cy3q_solver_donaldson_solve_legacy(cy3q_solver *s, int weighted, int64_t max_iter, double tol,
                                       int64_t *iters, double *residuals) {
  if (!s) return fail(CY3Q_ERR_ARG, "NULL solver");
  if (max_iter < 0) return fail(CY3Q_ERR_ARG, "max_iter must be >= 0");
  HIP_TRY(hipSetDevice(s->device));
  launch_identity_cplx(s->st, s->nb, s->g);  // init_h_block_identity
  ST_TRY(w_from_g(s));
  int64_t done = 0;
  for (int64_t it = 0; it < max_iter; it++) {
    double r = 0.0;
    ST_TRY(legacy_step(s, weighted, &r));
    if (residuals) residuals[it] = r;
    done = it + 1;
    if (r < tol) break;
  }
  if (iters) *iters = done;
  return CY3Q_OK;
}

int cy3q_solver_set_tuning(cy3q_solver *s, int64_t tpass_wgs_per_cu, int64_t syrk_target_wgs) {
  if (!s) return fail(CY3Q_ERR_ARG, "NULL solver");
  if (tpass_wgs_per_cu < 0 || tpass_wgs_per_cu > 16 || syrk_target_wgs < 0 ||
      (syrk_target_wgs > 0 && (syrk_target_wgs < 64 || syrk_target_wgs > 8192)))
    return fail(CY3Q_ERR_ARG, "tuning out of range (1..16 chunks per CU, 64..8192 workgroups)");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(hipStreamSynchronize(s->st));
  if (tpass_wgs_per_cu > 0 && tpass_wgs_per_cu != s->tp_mult) {
    s->tp_mult = tpass_wgs_per_cu;
    if (s->xd) {  // the split path is set up: re-plan its T-pass chunks
      int n_cu = 0;
      HIP_TRY(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, s->device));
      const int64_t slots = (int64_t)std::max(n_cu, 1) * s->tp_mult;
      s->tp_chunk = std::max<int64_t>((s->n + slots * 16 - 1) / (slots * 16), 1) * 16;
      s->tp_chunks = (s->n + s->tp_chunk - 1) / s->tp_chunk;
      s->release(s->partials32);
      s->release(s->tseg);
      s->partials32 = nullptr;
      s->tseg = nullptr;
      ST_TRY(s->alloc(&s->partials32, tpass_partials(s->m, s->tp_chunks)));
      ST_TRY(s->alloc(&s->tseg, tpass_seg_count(s->m, s->tp_chunks)));
    }
  }
  if (syrk_target_wgs > 0 && syrk_target_wgs != s->syrk_target) {
    s->syrk_target = syrk_target_wgs;
    ST_TRY(build_syrk_plan(s, s->plan64, 16, 8, syrk_f64_max_tiles(s->m), s->syrk_target, 256, 16));
    const int64_t need = syrk_partials_count(s->plan64);
    if (need > s->partials_cap) {
      s->release(s->partials);
      s->partials = nullptr;
      ST_TRY(s->alloc(&s->partials, need));
      s->partials_cap = need;
    }
  }
  HIP_TRY(hipStreamSynchronize(s->st));
  return CY3Q_OK;
}

int cy3q_solver_get_tuning(cy3q_solver *s, int64_t *tpass_wgs_per_cu, int64_t *syrk_target_wgs) {
  if (!s) return fail(CY3Q_ERR_ARG, "NULL solver");
  if (tpass_wgs_per_cu) *tpass_wgs_per_cu = s->tp_mult;
  if (syrk_target_wgs) *syrk_target_wgs = s->syrk_target;
  return CY3Q_OK;
}

int cy3q_solver_donaldson_solve(cy3q_solver *s, int64_t max_iter, double tol, int64_t *iters,
                                double *residuals) {
  if (!s) return fail(CY3Q_ERR_ARG, "NULL solver");
  if (max_iter < 0) return fail(CY3Q_ERR_ARG, "max_iter must be >= 0");
  HIP_TRY(hipSetDevice(s->device));
  // Split precision: the fast iterations run first; the last `refine` iterations (at most
  // max_iter) are FP64 T-operator sweeps from the fast path's metric.
  const int64_t n_ref = s->precision == CY3Q_PREC_FP64 ? 0 : std::min(s->refine, max_iter);
  const int64_t n_fast = max_iter - n_ref;
  auto prec_of = [&](int64_t it, bool fast_done) {
    return (it >= n_fast || fast_done) ? refine_step_code(s) : s->precision;
  };
  int64_t done = 0;
  if (tol > 0.0) {
    // Early stop: a fast iteration below tol ends the fast phase; the refinement sweeps follow
    // and the solve stops at the first of them below tol.
    bool fast_done = false;
    int64_t ref_done = 0;
    for (int64_t it = 0; it < max_iter; it++) {
      const int prec = prec_of(it, fast_done);
      ST_TRY(step_in(s, prec));
      double r = 0.0;
      HIP_TRY(hipMemcpyAsync(&r, s->resid, sizeof(double), hipMemcpyDeviceToHost, s->st));
      HIP_TRY(hipStreamSynchronize(s->st));
      if (residuals) residuals[it] = r;
      done = it + 1;
      if (prec != s->precision || s->precision == CY3Q_PREC_FP64) {
        ref_done++;
        if (r < tol || (s->precision != CY3Q_PREC_FP64 && ref_done >= n_ref)) break;
      } else if (r < tol) {
        fast_done = true;
        if (n_ref == 0) break;
      }
    }
    s->collect();
  } else {
    double *dres = nullptr;
    ST_TRY(s->alloc(&dres, std::max<int64_t>(max_iter, 1)));
    for (int64_t it = 0; it < max_iter; it++) {
      double *keep = s->resid;
      s->resid = dres + it;
      int rc = step_in(s, prec_of(it, false));
      s->resid = keep;
      if (rc != CY3Q_OK) return rc;
      done = it + 1;
    }
    if (residuals && max_iter > 0)
      HIP_TRY(hipMemcpyAsync(residuals, dres, sizeof(double) * max_iter, hipMemcpyDeviceToHost,
                             s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    s->release(dres);
    s->collect();
  }
  if (iters) *iters = done;
  return CY3Q_OK;
}

int cy3q_solver_sigma_sq_grad(cy3q_solver *s, double *sigma_l2sq, double *grad,
                              double *sigma_l1) {
  if (!s) return fail(CY3Q_ERR_ARG, "NULL solver");
  HIP_TRY(hipSetDevice(s->device));
  double o[6];
  ST_TRY(sigma_grad_impl(s, o));
  if (sigma_l2sq) *sigma_l2sq = o[2];
  if (sigma_l1) *sigma_l1 = o[0];
  if (grad)
    HIP_TRY(hipMemcpy(grad, s->grad_dev, sizeof(double) * 2 * s->nb * s->nb,
                      hipMemcpyDeviceToHost));
  return CY3Q_OK;
}

int cy3q_solver_set_adam(cy3q_solver *s, double beta1, double beta2, double eps) {
  if (!s) return fail(CY3Q_ERR_ARG, "NULL solver");
  if (!(beta1 >= 0.0 && beta1 < 1.0 && beta2 >= 0.0 && beta2 < 1.0 && eps > 0.0))
    return fail(CY3Q_ERR_ARG, "Adam needs 0 <= beta < 1 and eps > 0");
  s->adam_beta1 = beta1;
  s->adam_beta2 = beta2;
  s->adam_eps = eps;
  return CY3Q_OK;
}

int cy3q_solver_sigma_refine_analytic(cy3q_solver *s, int64_t n_iter, double lr,
                                      double *history, int64_t *n_done) {
  if (!s) return fail(CY3Q_ERR_ARG, "NULL solver");
  if (n_iter < 0) return fail(CY3Q_ERR_ARG, "n_iter must be >= 0");
  HIP_TRY(hipSetDevice(s->device));
  const int nb = s->nb;
  const int64_t n2 = 2 * (int64_t)nb * nb;
  int64_t done = 0;
  if (n_iter > 0) {
    double o[6];
    ST_TRY(sigma_grad_impl(s, o));  // allocates the Adam state on first use
    HIP_TRY(hipMemsetAsync(s->adam_m, 0, sizeof(double) * n2, s->st));
    HIP_TRY(hipMemsetAsync(s->adam_v, 0, sizeof(double) * n2, s->st));
    for (int64_t it = 0; it < n_iter; it++) {
      if (it > 0) ST_TRY(sigma_grad_impl(s, o));
      if (!std::isfinite(o[0])) break;
      if (history) history[it] = o[0];
      done = it + 1;
      const double t1 = (double)(it + 1);
      const double bc1 = 1.0 - std::pow(s->adam_beta1, t1), bc2 = 1.0 - std::pow(s->adam_beta2, t1);
      launch_adam_step(s->st, s->g, nb, s->grad_dev, s->adam_m, s->adam_v,
                       lr * (std::sqrt(bc2) / bc1), s->adam_beta1, s->adam_beta2, s->adam_eps);
      ST_TRY(w_from_g(s));
    }
  }
  if (n_done) *n_done = done;
  return CY3Q_OK;
}

}  // extern "C"

namespace {

// sigma_L2^2 (*f), its gradient (into s->grad_dev) and sigma_L1 at the current G. A G with no
// W (not positive definite) gives f = +inf and sigma_L1 = NaN, which every line search rejects.
int refine_eval(cy3q_solver *s, double *f, double *sigma_l1) {
  bool pd = false;
  ST_TRY(w_from_g_pd(s, &pd));
  if (!pd) {
    *f = INFINITY;
    *sigma_l1 = NAN;
    return CY3Q_OK;
  }
  double o[6];
  ST_TRY(sigma_grad_impl(s, o));
  *f = o[2];
  *sigma_l1 = o[0];
  return CY3Q_OK;
}

// Device scalars of the L-BFGS kernels, read back after each launch.
int read_scalars(cy3q_solver *s, const double *dev, int count, double *host) {
  HIP_TRY(hipMemcpyAsync(host, dev, sizeof(double) * count, hipMemcpyDeviceToHost, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  return CY3Q_OK;
}

int copy_dev(cy3q_solver *s, double *dst, const double *src, int64_t count) {
  HIP_TRY(hipMemcpyAsync(dst, src, sizeof(double) * count, hipMemcpyDeviceToDevice, s->st));
  return CY3Q_OK;
}

// sigma_refine_lbfgs_gpu_legacy (quintic.rs:678-880) with the vectors on the device. The pair
// stores hold m_history + 1 slots, so a candidate pair can be formed before it is known whether
// it enters the history (s.y > 1e-12), without overwriting the oldest pair.
int lbfgs_run(cy3q_solver *s, int64_t n_iter, int m_history, double *history, int64_t *n_done) {
  const int nb = s->nb;
  const int64_t n2 = 2 * (int64_t)nb * nb;  // gradient entries = doubles of the complex G
  const int n_slots = m_history + 1;
  double *s_all = nullptr, *y_all = nullptr, *d = nullptr, *g_cur = nullptr, *g_save = nullptr,
         *g_best = nullptr, *scal = nullptr;
  auto cleanup = [&] {
    for (double *p : {s_all, y_all, d, g_cur, g_save, g_best, scal})
      if (p) s->release(p);
  };
  auto run = [&]() -> int {
    ST_TRY(s->alloc(&s_all, (size_t)n_slots * n2));
    ST_TRY(s->alloc(&y_all, (size_t)n_slots * n2));
    ST_TRY(s->alloc(&d, n2));
    ST_TRY(s->alloc(&g_cur, n2));
    ST_TRY(s->alloc(&g_save, n2));
    ST_TRY(s->alloc(&g_best, n2));
    ST_TRY(s->alloc(&scal, 4));
    int64_t written = 0;
    double f_cur = 0.0, sig = 0.0;
    ST_TRY(refine_eval(s, &f_cur, &sig));
    if (!std::isfinite(sig)) {
      *n_done = 0;
      return CY3Q_OK;
    }
    if (history) history[written] = sig;
    written++;
    double best = sig;
    ST_TRY(copy_dev(s, g_best, s->g, n2));
    ST_TRY(copy_dev(s, g_cur, s->grad_dev, n2));
    LbfgsPairs hist{};
    hist.count = 0;
    const double c1 = 1e-4;
    for (int64_t iter = 0; iter < n_iter;) {
      double sc[2];
      launch_lbfgs_direction(s->st, g_cur, s_all, y_all, hist, n2, d, scal);
      ST_TRY(read_scalars(s, scal, 2, sc));
      double gd = sc[0], gg = sc[1];
      bool used_lbfgs = hist.count > 0;
      if (!std::isfinite(gd) || gd >= 0.0) {
        hist.count = 0;
        launch_lbfgs_direction(s->st, g_cur, s_all, y_all, hist, n2, d, scal);
        ST_TRY(read_scalars(s, scal, 2, sc));
        gd = sc[0];
        gg = sc[1];
        used_lbfgs = false;
        if (gd >= 0.0) break;  // zero gradient: converged
      }
      const double g_norm = std::sqrt(gg);
      double step = used_lbfgs ? 1.0 : (g_norm > 1e-30 ? 1.0 / g_norm : 1.0);
      ST_TRY(copy_dev(s, g_save, s->g, n2));
      bool accepted = false;
      double f_new = 0.0, sig_new = 0.0;
      for (int ls = 0; ls < 30; ls++) {
        launch_lbfgs_apply(s->st, g_save, nb, d, step, s->g);
        ST_TRY(refine_eval(s, &f_new, &sig_new));
        if (std::isfinite(f_new) && f_new <= f_cur + c1 * step * gd) {
          accepted = true;
          break;
        }
        step *= 0.5;
      }
      if (!accepted && used_lbfgs) {
        // The line search along the L-BFGS direction failed: drop the history and retry this
        // iteration along -g (quintic.rs:810-841).
        hist.count = 0;
        launch_lbfgs_direction(s->st, g_cur, s_all, y_all, hist, n2, d, scal);
        ST_TRY(read_scalars(s, scal, 2, sc));
        gd = sc[0];
        const double gn = std::sqrt(sc[1]);
        double sd_step = gn > 1e-30 ? 1.0 / gn : 1.0;
        for (int ls = 0; ls < 30; ls++) {
          launch_lbfgs_apply(s->st, g_save, nb, d, sd_step, s->g);
          ST_TRY(refine_eval(s, &f_new, &sig_new));
          if (std::isfinite(f_new) && f_new <= f_cur + c1 * sd_step * gd) {
            accepted = true;
            step = sd_step;
            break;
          }
          sd_step *= 0.5;
        }
      }
      if (!accepted) {  // neither direction gives a decrease: converged
        ST_TRY(copy_dev(s, s->g, g_save, n2));
        break;
      }
      // Candidate pair in a free slot; it enters the history when s.y > 1e-12, evicting the
      // oldest pair beyond m_history.
      int slot = 0;
      while (std::find(hist.slot, hist.slot + hist.count, slot) != hist.slot + hist.count) slot++;
      launch_lbfgs_pair(s->st, d, step, s->grad_dev, g_cur, n2, s_all + (int64_t)slot * n2,
                        y_all + (int64_t)slot * n2, scal + 2);
      double sy = 0.0;
      ST_TRY(read_scalars(s, scal + 2, 1, &sy));
      if (sy > 1e-12) {
        if (hist.count == m_history) {
          for (int k = 1; k < hist.count; k++) {
            hist.slot[k - 1] = hist.slot[k];
            hist.rho[k - 1] = hist.rho[k];
          }
          hist.count--;
        }
        hist.slot[hist.count] = slot;
        hist.rho[hist.count] = 1.0 / sy;
        hist.count++;
      }
      if (std::isfinite(sig_new) && sig_new < best) {
        best = sig_new;
        ST_TRY(copy_dev(s, g_best, s->g, n2));
      }
      f_cur = f_new;
      ST_TRY(copy_dev(s, g_cur, s->grad_dev, n2));
      if (history) history[written] = sig_new;
      written++;
      iter++;
    }
    // The best iterate (lowest sigma_L1) with its W.
    ST_TRY(copy_dev(s, s->g, g_best, n2));
    ST_TRY(w_from_g(s));
    *n_done = written;
    return CY3Q_OK;
  };
  const int rc = run();
  HIP_TRY(hipStreamSynchronize(s->st));
  cleanup();
  return rc;
}

// The restart loop of sigma_refine_analytic_with_restarts (quintic.rs:545-582) and
// sigma_refine_lbfgs_with_restarts_gpu (:1100-1139) around `refine`, with
// perturb_h_block_hermitian (:1195-1223) on G (h_re = Re G, h_im = Im G) in the LCG's draw
// order: for each a the diagonal, then re and im of each (a, b), b > a.
template <typename F>
int restarts_run(cy3q_solver *s, int64_t n_restarts, double scale, uint64_t seed, F refine,
                 double *best_sigma, double *history) {
  const int nb = s->nb;
  const int64_t n2 = 2 * (int64_t)nb * nb;
  double *g_best = nullptr;
  ST_TRY(s->alloc(&g_best, n2));
  auto run = [&]() -> int {
    double o[5];
    ST_TRY(refine());
    ST_TRY(sigma_impl(s, o, nullptr));
    double best = o[0];
    ST_TRY(copy_dev(s, g_best, s->g, n2));
    if (history) history[0] = best;
    Lcg rng{lcg_seed(seed)};
    std::vector<double> hg(n2);
    for (int64_t r = 0; r < n_restarts; r++) {
      HIP_TRY(hipMemcpyAsync(hg.data(), g_best, sizeof(double) * n2, hipMemcpyDeviceToHost,
                             s->st));
      HIP_TRY(hipStreamSynchronize(s->st));
      for (int a = 0; a < nb; a++) {
        const int64_t aa = (int64_t)a * nb + a;
        hg[2 * aa] += scale * (rng.next_f64() * 2.0 - 1.0);
        for (int b = a + 1; b < nb; b++) {
          const double dr = scale * (rng.next_f64() * 2.0 - 1.0);
          const double di = scale * (rng.next_f64() * 2.0 - 1.0);
          const int64_t ab = (int64_t)a * nb + b, ba = (int64_t)b * nb + a;
          hg[2 * ab] += dr;
          hg[2 * ba] += dr;
          hg[2 * ab + 1] += di;
          hg[2 * ba + 1] -= di;
        }
      }
      HIP_TRY(hipMemcpyAsync(s->g, hg.data(), sizeof(double) * n2, hipMemcpyHostToDevice, s->st));
      launch_trace_renorm(s->st, s->g, nb);
      bool pd = false;
      ST_TRY(w_from_g_pd(s, &pd));
      double trial = NAN;
      if (pd) {  // a perturbation that leaves G without a W counts as a failed restart
        ST_TRY(refine());
        ST_TRY(sigma_impl(s, o, nullptr));
        trial = o[0];
      }
      if (history) history[r + 1] = trial;
      if (std::isfinite(trial) && trial < best) {
        best = trial;
        ST_TRY(copy_dev(s, g_best, s->g, n2));
      }
    }
    ST_TRY(copy_dev(s, s->g, g_best, n2));
    ST_TRY(w_from_g(s));
    if (best_sigma) *best_sigma = best;
    return CY3Q_OK;
  };
  const int rc = run();
  HIP_TRY(hipStreamSynchronize(s->st));
  s->release(g_best);
  return rc;
}

}  // namespace

extern "C" {

int cy3q_solver_sigma_refine_lbfgs(cy3q_solver *s, int64_t n_iter, int64_t m_history,
                                   double *history, int64_t *n_done) {
  if (!s) return fail(CY3Q_ERR_ARG, "NULL solver");
  if (n_iter < 0) return fail(CY3Q_ERR_ARG, "n_iter must be >= 0");
  if (m_history < 1 || m_history > LBFGS_MAX_PAIRS - 1)
    return fail(CY3Q_ERR_ARG, "m_history must be in 1..63");
  HIP_TRY(hipSetDevice(s->device));
  int64_t done = 0;
  ST_TRY(lbfgs_run(s, n_iter, (int)m_history, history, &done));
  if (n_done) *n_done = done;
  return CY3Q_OK;
}

int cy3q_solver_sigma_refine_analytic_restarts(cy3q_solver *s, int64_t n_restarts,
                                               int64_t n_iter_per_restart, double lr,
                                               double perturb_scale, uint64_t seed,
                                               double *best_sigma, double *history) {
  if (!s) return fail(CY3Q_ERR_ARG, "NULL solver");
  if (n_restarts < 0 || n_iter_per_restart < 0)
    return fail(CY3Q_ERR_ARG, "restart and iteration counts must be >= 0");
  HIP_TRY(hipSetDevice(s->device));
  return restarts_run(
      s, n_restarts, perturb_scale, seed,
      [&] { return cy3q_solver_sigma_refine_analytic(s, n_iter_per_restart, lr, nullptr, nullptr); },
      best_sigma, history);
}

int cy3q_solver_sigma_refine_lbfgs_restarts(cy3q_solver *s, int64_t n_restarts,
                                            int64_t n_iter_per_restart, int64_t m_history,
                                            double perturb_scale, uint64_t seed,
                                            double *best_sigma, double *history) {
  if (!s) return fail(CY3Q_ERR_ARG, "NULL solver");
  if (n_restarts < 0 || n_iter_per_restart < 0)
    return fail(CY3Q_ERR_ARG, "restart and iteration counts must be >= 0");
  if (m_history < 1 || m_history > LBFGS_MAX_PAIRS - 1)
    return fail(CY3Q_ERR_ARG, "m_history must be in 1..63");
  HIP_TRY(hipSetDevice(s->device));
  return restarts_run(
      s, n_restarts, perturb_scale, seed,
      [&] {
        int64_t done = 0;
        return lbfgs_run(s, n_iter_per_restart, (int)m_history, nullptr, &done);
      },
      best_sigma, history);
}

int cy3q_solver_sigma_ext(cy3q_solver *s, double *sigma_l1, double *kappa, double *sigma_l2sq,
                          double *eta) {
  if (!s) return fail(CY3Q_ERR_ARG, "NULL solver");
  HIP_TRY(hipSetDevice(s->device));
  double o[5];
  ST_TRY(sigma_impl(s, o, eta));
  if (sigma_l1) *sigma_l1 = o[0];
  if (kappa) *kappa = o[1];
  if (sigma_l2sq) *sigma_l2sq = o[2];
  return CY3Q_OK;
}

int cy3q_solver_sigma(cy3q_solver *s, double *sigma) {
  return cy3q_solver_sigma_ext(s, sigma, nullptr, nullptr, nullptr);
}

int cy3q_solver_sigma_at_sample(cy3q_solver *s, int64_t n_points, uint64_t seed, int sampler,
                                double *sigma_l1, double *kappa, double *sigma_l2sq,
                                int64_t *n_actual, double *eta, double *points) {
  if (!s) return fail(CY3Q_ERR_ARG, "NULL solver");
  if (n_points <= 0) return fail(CY3Q_ERR_ARG, "n_points must be > 0");
  if (sampler != CY3Q_SAMPLER_SZ && sampler != CY3Q_SAMPLER_SZ_DWORK &&
      sampler != CY3Q_SAMPLER_NEWTON && sampler != CY3Q_SAMPLER_NEWTON_DWORK)
    return fail(CY3Q_ERR_ARG, "unknown sampler");
  const bool dwork = sampler == CY3Q_SAMPLER_SZ_DWORK || sampler == CY3Q_SAMPLER_NEWTON_DWORK;
  if (!dwork && s->psi != 0.0)
    return fail(CY3Q_ERR_ARG, "the solver's metric lives on a Dwork quintic; use a Dwork sampler");
  const bool newton = sampler == CY3Q_SAMPLER_NEWTON || sampler == CY3Q_SAMPLER_NEWTON_DWORK;
  HIP_TRY(hipSetDevice(s->device));
  // The metric's state must be complete before the second workspace reads it on its stream.
  HIP_TRY(hipStreamSynchronize(s->st));
  cy3q_solver *t = nullptr, *tout = nullptr;
  ST_TRY(make_solver(s->k, dwork ? s->psi : 0.0, &s->opt, &tout, &t));
  auto done = [&](int rc) {
    delete t;
    return rc;
  };
  int rc = newton ? sample_points_newton(t, n_points, seed, s->opt.newton_tol, dwork ? 1 : 0)
                  : sample_points(t, n_points, seed, dwork ? 1 : 0);
  if (rc != CY3Q_OK) return done(rc);
  if (n_actual) *n_actual = t->n;
  if (t->n == 0) {
    // The reference returns NaN for an empty sample (quintic.rs:1531-1533).
    if (sigma_l1) *sigma_l1 = NAN;
    if (kappa) *kappa = NAN;
    if (sigma_l2sq) *sigma_l2sq = NAN;
    return done(CY3Q_OK);
  }
  if ((rc = finish_workspace(t)) != CY3Q_OK) return done(rc);
  if (s->wfs && (rc = rebase_sections(t, s->wfs)) != CY3Q_OK) return done(rc);
  const int64_t nn = (int64_t)s->nb * s->nb;
  hipError_t e = hipMemcpyAsync(t->wmat, s->wmat, sizeof(double) * 2 * nn,
                                hipMemcpyDeviceToDevice, t->st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(t->g, s->g, sizeof(double) * 2 * nn, hipMemcpyDeviceToDevice, t->st);
  if (e != hipSuccess) return done(fail(CY3Q_ERR_HIP, hipGetErrorString(e)));
  t->bmat_valid = false;
  double o[5];
  if ((rc = sigma_impl(t, o, eta)) != CY3Q_OK) return done(rc);
  if (points) {
    e = hipMemcpy(points, t->pts, sizeof(double) * 10 * t->n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return done(fail(CY3Q_ERR_HIP, hipGetErrorString(e)));
  }
  if (sigma_l1) *sigma_l1 = o[0];
  if (kappa) *kappa = o[1];
  if (sigma_l2sq) *sigma_l2sq = o[2];
  return done(CY3Q_OK);
}

int cy3q_solver_get_h_block(cy3q_solver *s, double *h_block) {
  if (!s || !h_block) return fail(CY3Q_ERR_ARG, "NULL argument");
  HIP_TRY(hipSetDevice(s->device));
  const int64_t tn = 2 * (int64_t)s->nb;
  double *d = nullptr;
  ST_TRY(s->alloc(&d, tn * tn));
  launch_cplx_to_block(s->st, s->g, s->nb, d);
  HIP_TRY(hipMemcpyAsync(h_block, d, sizeof(double) * tn * tn, hipMemcpyDeviceToHost, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  s->release(d);
  return CY3Q_OK;
}

int cy3q_solver_set_h_block(cy3q_solver *s, const double *h_block) {
  if (!s || !h_block) return fail(CY3Q_ERR_ARG, "NULL argument");
  HIP_TRY(hipSetDevice(s->device));
  const int64_t tn = 2 * (int64_t)s->nb;
  double *d = nullptr;
  ST_TRY(s->alloc(&d, tn * tn));
  HIP_TRY(hipMemcpyAsync(d, h_block, sizeof(double) * tn * tn, hipMemcpyHostToDevice, s->st));
  launch_block_to_cplx(s->st, d, s->nb, s->g);
  launch_hermitian_part(s->st, s->g, s->nb);
  HIP_TRY(hipStreamSynchronize(s->st));
  s->release(d);
  return w_from_g(s);
}

int cy3q_solver_get_t(cy3q_solver *s, double *t_cplx) {
  if (!s || !t_cplx) return fail(CY3Q_ERR_ARG, "NULL argument");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(hipMemcpyAsync(t_cplx, s->t, sizeof(double) * 2 * s->nb * s->nb, hipMemcpyDeviceToHost,
                         s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  return CY3Q_OK;
}

int cy3q_solver_get_points(cy3q_solver *s, double *points) {
  if (!s || !points) return fail(CY3Q_ERR_ARG, "NULL argument");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(hipMemcpyAsync(points, s->pts, sizeof(double) * 10 * s->n, hipMemcpyDeviceToHost, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  return CY3Q_OK;
}

int cy3q_solver_get_weights(cy3q_solver *s, double *weights, double *log_omega) {
  if (!s) return fail(CY3Q_ERR_ARG, "NULL solver");
  HIP_TRY(hipSetDevice(s->device));
  if (weights)
    HIP_TRY(hipMemcpyAsync(weights, s->w, sizeof(double) * s->n, hipMemcpyDeviceToHost, s->st));
  if (log_omega)
    HIP_TRY(hipMemcpyAsync(log_omega, s->log_om, sizeof(double) * s->n, hipMemcpyDeviceToHost,
                           s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  return CY3Q_OK;
}

int cy3q_solver_set_galerkin_precision(cy3q_solver *s, int precision) {
  if (!s) return fail(CY3Q_ERR_ARG, "NULL solver");
  if (precision != CY3Q_PREC_FP64 && precision != CY3Q_PREC_SPLIT16)
    return fail(CY3Q_ERR_ARG, "precision must be CY3Q_PREC_FP64 or CY3Q_PREC_SPLIT16");
  s->galerkin_precision = precision;
  return CY3Q_OK;
}

int cy3q_solver_set_weights(cy3q_solver *s, const double *weights) {
  if (!s || !weights) return fail(CY3Q_ERR_ARG, "NULL argument");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(hipMemcpyAsync(s->w, weights, sizeof(double) * s->n, hipMemcpyHostToDevice, s->st));
  launch_valid_weights(s->st, s->w, s->n, s->wvalid);
  launch_sum_f64(s->st, s->wvalid, s->n, s->sum_w);
  HIP_TRY(hipStreamSynchronize(s->st));
  return CY3Q_OK;
}

int cy3q_monge_ampere_residual(uint32_t k, const double *points, int64_t n, const double *h_block,
                               const double *weights, int device, double *sigma) {
  if (!points || !h_block || !sigma) return fail(CY3Q_ERR_ARG, "NULL argument");
  if (n <= 0) return fail(CY3Q_ERR_EMPTY, "no points");
  cy3q_options opt;
  cy3q_options_default(&opt);
  opt.device = device;
  cy3q_solver *s = nullptr;
  ST_TRY(cy3q_solver_new_from_points(k, points, n, 0.0, &opt, &s));
  int rc = cy3q_solver_set_h_block(s, h_block);
  if (rc == CY3Q_OK && weights) rc = cy3q_solver_set_weights(s, weights);
  if (rc == CY3Q_OK) rc = cy3q_solver_sigma(s, sigma);
  cy3q_solver_free(s);
  return rc;
}

int cy3q_solver_get_sections(cy3q_solver *s, double *section_values, double *section_derivs) {
  if (!s) return fail(CY3Q_ERR_ARG, "NULL solver");
  HIP_TRY(hipSetDevice(s->device));
  const int64_t two_n = 2 * (int64_t)s->nb;
  if (section_values) ST_TRY(download_rows_packed(s, section_values, s->sv, s->n, two_n, s->m));
  if (section_derivs)
    ST_TRY(download_rows_packed(s, section_derivs, s->sd, s->n * 5, two_n, s->m));
  return CY3Q_OK;
}

int cy3q_solver_set_profiling(cy3q_solver *s, int on) {
  if (!s) return fail(CY3Q_ERR_ARG, "NULL solver");
  s->collect();
  s->profiling = on != 0;
  for (int i = 0; i < STG_COUNT; i++) {
    s->stage_ms[i] = 0.0;
    s->stage_cnt[i] = 0;
  }
  return CY3Q_OK;
}

int cy3q_solver_stage_times(cy3q_solver *s, double *ms, int64_t *count, int cap, int *n_stages) {
  if (!s) return fail(CY3Q_ERR_ARG, "NULL solver");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(hipStreamSynchronize(s->st));
  s->collect();
  if (n_stages) *n_stages = STG_COUNT;
  for (int i = 0; i < STG_COUNT && i < cap; i++) {
    if (ms) ms[i] = s->stage_ms[i];
    if (count) count[i] = s->stage_cnt[i];
  }
  return CY3Q_OK;
}

const char *cy3q_stage_name(int stage) {
  return (stage >= 0 && stage < STG_COUNT) ? kStageNames[stage] : "";
}

int cy3q_solver_synchronize(cy3q_solver *s) {
  if (!s) return fail(CY3Q_ERR_ARG, "NULL solver");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(hipStreamSynchronize(s->st));
  s->collect();
  return CY3Q_OK;
}

}  // extern "C"

// ---- The route34 T-operator contract (Cy3DonaldsonGpu, route34/cy3_donaldson_gpu.rs:164-333) --
// A solver workspace without points: the caller's section values (n_points x n_basis complex,
// (re, im) interleaved) and weights are the static inputs, and each call maps a route34 h
// (h_route34 = conj(G) of this library's G, so K_p = s^H G s is the route34
// sum h_re Re(conj(s_a) s_b) + h_im Im(conj(s_a) s_b)) to the unnormalised
// h_pair = sum_p (w_p / max(K_p, 1e-30)) conj(s_a) s_b over points with finite w_p > 0.
struct cy3q_top {
  cy3q_solver *s = nullptr;
  bool uploaded = false;
  std::vector<double> xx_host;
};

extern "C" {

int cy3q_top_new(int64_t n_points, int64_t n_basis, int device, cy3q_top **out) {
  if (!out) return fail(CY3Q_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (n_points <= 0 || n_basis <= 0)
    return fail(CY3Q_ERR_ARG, "n_points and n_basis must be > 0");
  if (padded_cols(n_basis) > 1024) return fail(CY3Q_ERR_ARG, "2 n_basis > 1024 is not supported");
  cy3q_solver *s = new cy3q_solver();
  cy3q_options_default(&s->opt);
  s->opt.device = device;
  s->device = device;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) {
    delete s;
    return fail(CY3Q_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
  }
  e = hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking);
  if (e != hipSuccess) {
    s->st = nullptr;
    delete s;
    return fail(CY3Q_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  auto bail = [&](int rc) {
    delete s;
    return rc;
  };
  s->nb = (int)n_basis;
  s->m = padded_cols(n_basis);
  s->n = n_points;
  const int64_t n = s->n, m = s->m, nn = n_basis * n_basis;
  const int nblk = (s->nb + 31) / 32;
  int rc = CY3Q_OK;
  if ((rc = s->alloc(&s->w, n)) || (rc = s->alloc(&s->wvalid, n)) ||
      (rc = s->alloc(&s->sum_w, 1)) || (rc = s->alloc(&s->sv, n * m)) ||
      (rc = s->alloc(&s->g, 2 * nn)) || (rc = s->alloc(&s->wmat, 2 * nn)) ||
      (rc = s->alloc(&s->t, 2 * nn)) || (rc = s->alloc(&s->wnew, 2 * nn)) ||
      (rc = s->alloc(&s->gnew, 2 * nn)) || (rc = s->alloc(&s->bmat, m * m)) ||
      (rc = s->alloc(&s->xx, m * m)) || (rc = s->alloc(&s->coef, n)) ||
      (rc = s->alloc(&s->dpart, (m / 32) * n)) || (rc = s->alloc(&s->resid, 1)) ||
      (rc = s->alloc(&s->status, 4)) || (rc = s->alloc(&s->abuf, 2 * nn)) ||
      (rc = s->alloc(&s->rbuf, 2 * nn)) ||
      (rc = s->alloc(&s->einv, (size_t)nblk * 32 * 32 * 2)) ||
      (rc = s->alloc(&s->partial, (size_t)nblk * nblk)) || (rc = s->alloc(&s->scale_buf, 2)))
    return bail(rc);
  if ((rc = build_syrk_plan(s, s->plan64, 16, 8, syrk_f64_max_tiles(s->m), s->syrk_target, 256,
                            16)))
    return bail(rc);
  s->partials_cap = syrk_partials_count(s->plan64);
  if ((rc = s->alloc(&s->partials, s->partials_cap))) return bail(rc);
  cy3q_top *t = new cy3q_top();
  t->s = s;
  t->xx_host.resize((size_t)(m * m));
  *out = t;
  return CY3Q_OK;
}

void cy3q_top_free(cy3q_top *t) {
  if (!t) return;
  delete t->s;
  delete t;
}

int cy3q_top_upload_static(cy3q_top *t, const double *section_values, const double *weights) {
  if (!t || !section_values || !weights) return fail(CY3Q_ERR_ARG, "NULL argument");
  cy3q_solver *s = t->s;
  HIP_TRY(hipSetDevice(s->device));
  const int64_t n = s->n, m = s->m, two_n = 2 * (int64_t)s->nb;
  // Row p of sv: the caller's 2 n_basis values, then zero padding to m columns.
  ST_TRY(upload_rows_padded(s, s->sv, section_values, n, two_n, m));
  HIP_TRY(hipMemcpyAsync(s->w, weights, sizeof(double) * n, hipMemcpyHostToDevice, s->st));
  launch_valid_weights(s->st, s->w, n, s->wvalid);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->st));
  t->uploaded = true;
  return CY3Q_OK;
}

int cy3q_top_t_operator_raw(cy3q_top *t, const double *h_re, const double *h_im, double *out_re,
                            double *out_im) {
  if (!t || !h_re || !h_im || !out_re || !out_im) return fail(CY3Q_ERR_ARG, "NULL argument");
  if (!t->uploaded) return fail(CY3Q_ERR_STATE, "cy3q_top_upload_static first");
  cy3q_solver *s = t->s;
  HIP_TRY(hipSetDevice(s->device));
  const int nb = s->nb;
  const int64_t n = s->n, m = s->m, nn = (int64_t)nb * nb;
  // G = conj(h_route34), then W with W^H W = G for the D pass.
  std::vector<double> gh((size_t)(2 * nn));
  for (int64_t e = 0; e < nn; e++) {
    gh[2 * e] = h_re[e];
    gh[2 * e + 1] = -h_im[e];
  }
  HIP_TRY(hipMemcpyAsync(s->g, gh.data(), sizeof(double) * 2 * nn, hipMemcpyHostToDevice, s->st));
  bool pd = false;
  ST_TRY(w_from_g_pd(s, &pd));
  if (!pd) return fail(CY3Q_ERR_NOT_PD, "h is not Hermitian positive definite");
  ensure_bmat(s);
  launch_dpass_f64(s->st, s->sv, n, m, s->bmat, s->w, s->dpart, s->coef);
  launch_syrk_f64(s->st, s->plan64, s->sv, s->coef, n, s->partials);
  launch_syrk_reduce(s->st, s->plan64, s->partials, s->xx);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(t->xx_host.data(), s->xx, sizeof(double) * m * m, hipMemcpyDeviceToHost,
                         s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  const double *xx = t->xx_host.data();
  for (int a = 0; a < nb; a++)
    for (int b = 0; b < nb; b++) {
      const int64_t i = 2 * a, j = 2 * b;
      out_re[(int64_t)a * nb + b] = xx[i * m + j] + xx[(i + 1) * m + j + 1];
      out_im[(int64_t)a * nb + b] = xx[i * m + j + 1] - xx[(i + 1) * m + j];
    }
  return CY3Q_OK;
}

}  // extern "C"

// ---- The route34 sigma contract (Cy3SigmaGpu, route34/cy3_sigma_gpu.rs:314-560) ------------
// Static inputs: section values [n][2N], their 8 ambient derivatives [n][8][2N], the 3 x 8
// complex tangent frames [n][48], |Omega|^2 and weights. Per call, h (used as given:
// K = s^H h s, route34/cy3_sigma_gpu.rs:80-107) gives W with W^H W = h, the contracted rows of
// each point chunk (k_trigemm_f64 over 9 rows per point), eta (k_eta8) and sigma =
// sum w (eta / kappa - 1)^2 / sum w over valid points (the two-pass reduction).
struct cy3q_sig8 {
  cy3q_solver *s = nullptr;
  bool uploaded = false;
  double *x9 = nullptr, *y9 = nullptr, *frames = nullptr, *om = nullptr, *eta = nullptr;
  int64_t chunk = 0;
};

extern "C" {

int cy3q_sig8_new(int64_t n_points, int64_t n_basis, int device, cy3q_sig8 **out) {
  if (!out) return fail(CY3Q_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (n_points <= 0 || n_basis <= 0)
    return fail(CY3Q_ERR_ARG, "n_points and n_basis must be > 0");
  if (padded_cols(n_basis) > 1024) return fail(CY3Q_ERR_ARG, "2 n_basis > 1024 is not supported");
  cy3q_solver *s = new cy3q_solver();
  cy3q_options_default(&s->opt);
  s->opt.device = device;
  s->device = device;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) {
    delete s;
    return fail(CY3Q_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
  }
  e = hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking);
  if (e != hipSuccess) {
    s->st = nullptr;
    delete s;
    return fail(CY3Q_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  cy3q_sig8 *g = new cy3q_sig8();
  g->s = s;
  auto bail = [&](int rc) {
    delete g->s;
    delete g;
    return rc;
  };
  s->nb = (int)n_basis;
  s->m = padded_cols(n_basis);
  s->n = n_points;
  const int64_t n = s->n, m = s->m, nn = n_basis * n_basis;
  const int nblk = (s->nb + 31) / 32;
  g->chunk = std::min<int64_t>(n, 32768);
  int rc = CY3Q_OK;
  if ((rc = s->alloc(&s->w, n)) || (rc = s->alloc(&g->om, n)) || (rc = s->alloc(&g->eta, n)) ||
      (rc = s->alloc(&g->frames, n * 48)) || (rc = s->alloc(&g->x9, n * 9 * m)) ||
      (rc = s->alloc(&g->y9, g->chunk * 9 * m)) || (rc = s->alloc(&s->g, 2 * nn)) ||
      (rc = s->alloc(&s->wmat, 2 * nn)) || (rc = s->alloc(&s->wnew, 2 * nn)) ||
      (rc = s->alloc(&s->gnew, 2 * nn)) || (rc = s->alloc(&s->bmat, m * m)) ||
      (rc = s->alloc(&s->status, 4)) || (rc = s->alloc(&s->abuf, 2 * nn)) ||
      (rc = s->alloc(&s->rbuf, 2 * nn)) ||
      (rc = s->alloc(&s->einv, (size_t)nblk * 32 * 32 * 2)) ||
      (rc = s->alloc(&s->partial, (size_t)nblk * nblk)) || (rc = s->alloc(&s->sig_out, 8)) ||
      (rc = s->alloc(&s->sig_scratch, sigma_reduce_scratch())))
    return bail(rc);
  *out = g;
  return CY3Q_OK;
}

void cy3q_sig8_free(cy3q_sig8 *g) {
  if (!g) return;
  delete g->s;
  delete g;
}

int cy3q_sig8_upload_static(cy3q_sig8 *g, const double *section_values,
                            const double *section_derivs, const double *frames,
                            const double *omega_sq, const double *weights) {
  if (!g || !section_values || !section_derivs || !frames || !omega_sq || !weights)
    return fail(CY3Q_ERR_ARG, "NULL argument");
  cy3q_solver *s = g->s;
  HIP_TRY(hipSetDevice(s->device));
  const int64_t n = s->n, m = s->m, two_n = 2 * (int64_t)s->nb;
  // x9 row 9 p: the section values of point p, rows 9 p + 1 + i: d/dz_i; zero padding to m.
  // The host arrays go to staging buffers, then one kernel packs the rows.
  double *sv_raw = nullptr, *sd_raw = nullptr;
  ST_TRY(s->alloc(&sv_raw, n * two_n));
  ST_TRY(s->alloc(&sd_raw, n * 8 * two_n));
  HIP_TRY(hipMemcpyAsync(sv_raw, section_values, sizeof(double) * n * two_n, hipMemcpyHostToDevice,
                         s->st));
  HIP_TRY(hipMemcpyAsync(sd_raw, section_derivs, sizeof(double) * n * 8 * two_n,
                         hipMemcpyHostToDevice, s->st));
  launch_pack_rows9(s->st, sv_raw, sd_raw, n, two_n, m, g->x9);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->st));
  s->release(sv_raw);
  s->release(sd_raw);
  HIP_TRY(hipMemcpyAsync(g->frames, frames, sizeof(double) * n * 48, hipMemcpyHostToDevice, s->st));
  HIP_TRY(hipMemcpyAsync(g->om, omega_sq, sizeof(double) * n, hipMemcpyHostToDevice, s->st));
  HIP_TRY(hipMemcpyAsync(s->w, weights, sizeof(double) * n, hipMemcpyHostToDevice, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  g->uploaded = true;
  return CY3Q_OK;
}

int cy3q_sig8_compute_sigma(cy3q_sig8 *g, const double *h_re, const double *h_im, double *sigma) {
  if (!g || !h_re || !h_im || !sigma) return fail(CY3Q_ERR_ARG, "NULL argument");
  if (!g->uploaded) return fail(CY3Q_ERR_STATE, "cy3q_sig8_upload_static first");
  cy3q_solver *s = g->s;
  HIP_TRY(hipSetDevice(s->device));
  const int nb = s->nb;
  const int64_t n = s->n, m = s->m, nn = (int64_t)nb * nb;
  std::vector<double> gh((size_t)(2 * nn));
  for (int64_t e = 0; e < nn; e++) {
    gh[2 * e] = h_re[e];
    gh[2 * e + 1] = h_im[e];
  }
  HIP_TRY(hipMemcpyAsync(s->g, gh.data(), sizeof(double) * 2 * nn, hipMemcpyHostToDevice, s->st));
  bool pd = false;
  ST_TRY(w_from_g_pd(s, &pd));
  if (!pd) return fail(CY3Q_ERR_NOT_PD, "h is not Hermitian positive definite");
  ensure_bmat(s);
  for (int64_t p0 = 0; p0 < n; p0 += g->chunk) {
    const int64_t cn = std::min(g->chunk, n - p0);
    launch_trigemm_f64(s->st, g->x9 + p0 * 9 * m, cn * 9, m, s->bmat, g->y9, 0, 0);
    launch_eta8(s->st, g->y9, m, nb, g->frames + p0 * 48, g->om + p0, s->w + p0, cn, g->eta + p0);
  }
  launch_sigma_reduce(s->st, g->eta, s->w, n, s->sig_scratch, s->sig_out, 1e-30);
  HIP_TRY(hipGetLastError());
  double o[6];
  HIP_TRY(hipMemcpyAsync(o, s->sig_out, sizeof(o), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  *sigma = o[2];
  return CY3Q_OK;
}

// Sections, derivatives and frames of caller points evaluated on the device straight into the
// static buffers of a T-operator and a sigma contract: the same arrays cy3q_cicy_sections and
// cy3q_cicy_frames return, without the round trip through host memory.
int cy3q_cicy_upload_contracts(int variety, int64_t n, const double *coords, const int32_t *patch,
                               int64_t n_basis, const uint32_t *exps, const double *omega_sq,
                               const double *weights, cy3q_top *top, cy3q_sig8 *sig) {
  if (!coords || !patch || !exps || !omega_sq || !weights || !top || !sig)
    return fail(CY3Q_ERR_ARG, "NULL argument");
  if (variety != CY3Q_CICY_TY && variety != CY3Q_CICY_SCHOEN)
    return fail(CY3Q_ERR_ARG, "variety must be CY3Q_CICY_TY or CY3Q_CICY_SCHOEN");
  cy3q_solver *s = top->s, *s2 = sig->s;
  if (s->n != n || s2->n != n || s->nb != n_basis || s2->nb != n_basis)
    return fail(CY3Q_ERR_ARG, "contract sizes differ from n / n_basis");
  if (s->device != s2->device) return fail(CY3Q_ERR_ARG, "contracts on different devices");
  for (int64_t i = 0; i < 8 * n_basis; i++)
    if (exps[i] > 64) return fail(CY3Q_ERR_ARG, "exponent above 64");
  for (int64_t p = 0; p < n; p++) {
    const int32_t *q = patch + 3 * p;
    const bool good = variety == CY3Q_CICY_TY
                          ? (q[0] >= 0 && q[0] < 4 && q[1] >= 0 && q[1] < 4)
                          : (q[0] >= 0 && q[0] < 3 && q[1] >= 3 && q[1] < 6 && q[2] >= 6 && q[2] < 8);
    if (!good) return fail(CY3Q_ERR_ARG, "patch index out of range at point " + std::to_string(p));
  }
  HIP_TRY(hipSetDevice(s->device));
  const int64_t m = s->m, two_n = 2 * n_basis;
  double *d_c = nullptr, *sv_raw = nullptr, *sd_raw = nullptr;
  int32_t *d_p = nullptr;
  uint32_t *d_e = nullptr;
  ST_TRY(s->alloc(&d_c, n * 16));
  ST_TRY(s->alloc(&d_p, n * 3));
  ST_TRY(s->alloc(&d_e, n_basis * 8));
  ST_TRY(s->alloc(&sv_raw, n * two_n));
  ST_TRY(s->alloc(&sd_raw, n * 8 * two_n));
  HIP_TRY(hipMemcpyAsync(d_c, coords, sizeof(double) * 16 * n, hipMemcpyHostToDevice, s->st));
  HIP_TRY(hipMemcpyAsync(d_p, patch, sizeof(int32_t) * 3 * n, hipMemcpyHostToDevice, s->st));
  HIP_TRY(hipMemcpyAsync(d_e, exps, sizeof(uint32_t) * 8 * n_basis, hipMemcpyHostToDevice, s->st));
  launch_cicy_sections(s->st, d_c, n, d_e, (int)n_basis, sv_raw, sd_raw);
  // T-operator: rows padded to m columns, weights and their validity.
  HIP_TRY(hipMemsetAsync(s->sv, 0, sizeof(double) * n * m, s->st));
  HIP_TRY(hipMemcpy2DAsync(s->sv, sizeof(double) * m, sv_raw, sizeof(double) * two_n,
                           sizeof(double) * two_n, n, hipMemcpyDeviceToDevice, s->st));
  HIP_TRY(hipMemcpyAsync(s->w, weights, sizeof(double) * n, hipMemcpyHostToDevice, s->st));
  launch_valid_weights(s->st, s->w, n, s->wvalid);
  // sigma: the nine packed rows per point, frames, |Omega|^2, weights.
  launch_pack_rows9(s->st, sv_raw, sd_raw, n, two_n, s2->m, sig->x9);
  launch_cicy_frames(s->st, variety, d_c, d_p, n, sig->frames);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(sig->om, omega_sq, sizeof(double) * n, hipMemcpyHostToDevice, s->st));
  HIP_TRY(hipMemcpyAsync(s2->w, weights, sizeof(double) * n, hipMemcpyHostToDevice, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  for (void *q : {(void *)d_c, (void *)d_p, (void *)d_e, (void *)sv_raw, (void *)sd_raw}) s->release(q);
  top->uploaded = true;
  sig->uploaded = true;
  return CY3Q_OK;
}

}  // extern "C"

// ---- The sigma-gradient kernel contract: AdamGradientKernel (gpu_adam.rs:1533-2194) ----------
// A points-free solver workspace holds the caller's static inputs (sections, derivatives,
// points, weights, log|Omega|^2; the Fermat quintic, as the reference kernel). Per call, h gives
// W with W^H W = G (its Hermitian part, as cy3q_solver_set_h_block), then either the sigma^2
// gradient of the solver (sigma_grad_impl) or the reference's per-point buffer in chunks.
struct cy3q_adamgrad {
  cy3q_solver *s = nullptr;
  bool uploaded = false;
  double *pp = nullptr, *eta = nullptr;  // per-point output chunk and eta of the chunk
  int64_t pp_chunk = 0;
  std::vector<double> grad_host;
};

namespace {
constexpr double kPerPointChunkBytes = 1 << 30;

int adamgrad_set_h(cy3q_adamgrad *g, const double *h_block) {
  cy3q_solver *s = g->s;
  const int64_t tn = 2 * (int64_t)s->nb;
  double *d = nullptr;
  ST_TRY(s->alloc(&d, tn * tn));
  HIP_TRY(hipMemcpyAsync(d, h_block, sizeof(double) * tn * tn, hipMemcpyHostToDevice, s->st));
  launch_block_to_cplx(s->st, d, s->nb, s->g);
  launch_hermitian_part(s->st, s->g, s->nb);
  HIP_TRY(hipStreamSynchronize(s->st));
  s->release(d);
  bool pd = false;
  ST_TRY(w_from_g_pd(s, &pd));
  if (!pd) return fail(CY3Q_ERR_NOT_PD, "h_block is not Hermitian positive definite");
  return CY3Q_OK;
}
}  // namespace

extern "C" {

int cy3q_adamgrad_new(int64_t n_points, int64_t n_basis, int device, cy3q_adamgrad **out) {
  if (!out) return fail(CY3Q_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (n_points <= 0 || n_basis <= 0)
    return fail(CY3Q_ERR_ARG, "n_points and n_basis must be > 0");
  if (padded_cols(n_basis) > 1024) return fail(CY3Q_ERR_ARG, "2 n_basis > 1024 is not supported");
  cy3q_solver *s = new cy3q_solver();
  cy3q_options_default(&s->opt);
  s->opt.device = device;
  s->device = device;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) {
    delete s;
    return fail(CY3Q_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
  }
  e = hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking);
  if (e != hipSuccess) {
    s->st = nullptr;
    delete s;
    return fail(CY3Q_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  s->nb = (int)n_basis;
  s->m = padded_cols(n_basis);
  s->n = n_points;
  const int64_t n = s->n, m = s->m, nn = n_basis * n_basis;
  const int nblk = (s->nb + 31) / 32;
  cy3q_adamgrad *g = new cy3q_adamgrad();
  g->s = s;
  auto bail = [&](int rc) {
    delete s;
    delete g;
    return rc;
  };
  int rc = CY3Q_OK;
  if ((rc = s->alloc(&s->pts, n * 10)) || (rc = s->alloc(&s->w, n)) ||
      (rc = s->alloc(&s->log_om, n)) || (rc = s->alloc(&s->wvalid, n)) ||
      (rc = s->alloc(&s->sum_w, 1)) || (rc = s->alloc(&s->sv, n * m)) ||
      (rc = s->alloc(&s->sd, n * 5 * m)) || (rc = s->alloc(&s->g, 2 * nn)) ||
      (rc = s->alloc(&s->wmat, 2 * nn)) || (rc = s->alloc(&s->wnew, 2 * nn)) ||
      (rc = s->alloc(&s->gnew, 2 * nn)) || (rc = s->alloc(&s->bmat, m * m)) ||
      (rc = s->alloc(&s->status, 4)) || (rc = s->alloc(&s->abuf, 2 * nn)) ||
      (rc = s->alloc(&s->rbuf, 2 * nn)) ||
      (rc = s->alloc(&s->einv, (size_t)nblk * 32 * 32 * 2)) ||
      (rc = s->alloc(&s->partial, (size_t)nblk * nblk)) || (rc = s->alloc(&s->sig_out, 8)) ||
      (rc = s->alloc(&s->sig_scratch, sigma_reduce_scratch())))
    return bail(rc);
  launch_identity_cplx(s->st, s->nb, s->g);
  launch_identity_cplx(s->st, s->nb, s->wmat);
  e = hipStreamSynchronize(s->st);
  if (e != hipSuccess) return bail(fail(CY3Q_ERR_HIP, hipGetErrorString(e)));
  *out = g;
  return CY3Q_OK;
}

void cy3q_adamgrad_free(cy3q_adamgrad *g) {
  if (!g) return;
  delete g->s;
  delete g;
}

int cy3q_adamgrad_upload_static_inputs(cy3q_adamgrad *g, const double *section_values,
                                       const double *section_derivs, const double *points,
                                       const double *weights, const double *log_omega_sq) {
  if (!g || !section_values || !section_derivs || !points || !weights || !log_omega_sq)
    return fail(CY3Q_ERR_ARG, "NULL argument");
  cy3q_solver *s = g->s;
  HIP_TRY(hipSetDevice(s->device));
  const int64_t n = s->n, m = s->m, two_n = 2 * (int64_t)s->nb;
  // Rows padded from 2 N to m columns with zeros.
  ST_TRY(upload_rows_padded(s, s->sv, section_values, n, two_n, m));
  ST_TRY(upload_rows_padded(s, s->sd, section_derivs, n * 5, two_n, m));
  HIP_TRY(hipMemcpyAsync(s->pts, points, sizeof(double) * 10 * n, hipMemcpyHostToDevice, s->st));
  HIP_TRY(hipMemcpyAsync(s->w, weights, sizeof(double) * n, hipMemcpyHostToDevice, s->st));
  HIP_TRY(hipMemcpyAsync(s->log_om, log_omega_sq, sizeof(double) * n, hipMemcpyHostToDevice,
                         s->st));
  launch_valid_weights(s->st, s->w, n, s->wvalid);
  launch_sum_f64(s->st, s->wvalid, n, s->sum_w);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->st));
  // Contracted rows computed for the previous inputs are stale.
  s->bmat_valid = false;
  g->uploaded = true;
  return CY3Q_OK;
}

int cy3q_adamgrad_compute_sigma_grad_h_only(cy3q_adamgrad *g, const double *h_block,
                                            double *sigma_l2sq, double *grad, double *sigma_l1) {
  if (!g || !h_block) return fail(CY3Q_ERR_ARG, "NULL argument");
  if (!g->uploaded) return fail(CY3Q_ERR_STATE, "cy3q_adamgrad_upload_static_inputs first");
  cy3q_solver *s = g->s;
  HIP_TRY(hipSetDevice(s->device));
  ST_TRY(adamgrad_set_h(g, h_block));
  double o[6];
  ST_TRY(sigma_grad_impl(s, o));
  if (grad)
    HIP_TRY(hipMemcpy(grad, s->grad_dev, sizeof(double) * 2 * (int64_t)s->nb * s->nb,
                      hipMemcpyDeviceToHost));
  if (sigma_l2sq) *sigma_l2sq = o[2];
  if (sigma_l1) *sigma_l1 = o[0];
  return CY3Q_OK;
}

int cy3q_adamgrad_compute_per_point_h_only(cy3q_adamgrad *g, const double *h_block,
                                           double *per_point_buf) {
  if (!g || !h_block || !per_point_buf) return fail(CY3Q_ERR_ARG, "NULL argument");
  if (!g->uploaded) return fail(CY3Q_ERR_STATE, "cy3q_adamgrad_upload_static_inputs first");
  cy3q_solver *s = g->s;
  if (s->nb > per_point_grad_max_nb())
    return fail(CY3Q_ERR_ARG, "the per-point buffer supports n_basis <= " +
                                  std::to_string(per_point_grad_max_nb()));
  HIP_TRY(hipSetDevice(s->device));
  ST_TRY(adamgrad_set_h(g, h_block));
  const int64_t n = s->n, m = s->m, nn = (int64_t)s->nb * s->nb, stride = 2 + 2 * nn;
  if (!g->pp) {
    int64_t chunk = (int64_t)(kPerPointChunkBytes / (8.0 * (double)stride));
    chunk = std::max<int64_t>(1, std::min<int64_t>({chunk, n, kRowChunk}));
    g->pp_chunk = chunk;
    ST_TRY(s->alloc(&g->pp, chunk * stride));
    ST_TRY(s->alloc(&g->eta, chunk));
    if (!s->g_phi) {
      s->g_chunk = std::min<int64_t>(n, kRowChunk);
      s->g_groups = gemm_tn_groups(5 * s->g_chunk, m);
      ST_TRY(s->alloc(&s->g_phi, s->g_chunk * 5 * m));
      ST_TRY(s->alloc(&s->g_vs, s->g_chunk * m));
      ST_TRY(s->alloc(&s->g_part, s->g_groups * m * m));
      ST_TRY(s->alloc(&s->g_cm, m * m));
    }
    if (!s->y0all) {
      ST_TRY(s->alloc(&s->y0all, s->g_chunk * m));
      ST_TRY(s->alloc(&s->ydall, s->g_chunk * 5 * m));
    }
  }
  ensure_bmat(s);
  const int64_t chunk = std::min(g->pp_chunk, s->g_chunk);
  for (int64_t p0 = 0; p0 < n; p0 += chunk) {
    const int64_t cn = std::min(chunk, n - p0);
    // With rows resident from a gradient call, the chunk's rows are recomputed here all the same
    // (they may belong to an older h).
    launch_trigemm_f64(s->st, s->sv + p0 * m, cn, m, s->bmat, s->y0all, 0, 0);
    launch_trigemm_f64(s->st, s->sd + p0 * 5 * m, cn * 5, m, s->bmat, s->ydall, 0, 0);
    launch_grad_rows(s->st, s->pts + 10 * p0, s->y0all, s->ydall, s->log_om + p0, s->w + p0,
                     s->sv + p0 * m, s->sd + p0 * 5 * m, cn, m, s->psi, nullptr, s->g_phi,
                     s->g_vs, g->eta);
    launch_per_point_grad(s->st, s->sv + p0 * m, s->sd + p0 * 5 * m, s->g_phi, s->g_vs, g->eta,
                          s->w + p0, s->nb, m, cn, g->pp);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(per_point_buf + p0 * stride, g->pp, sizeof(double) * cn * stride,
                           hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
  }
  return CY3Q_OK;
}

int cy3q_adamgrad_compute_per_point(cy3q_adamgrad *g, const double *section_values,
                                    const double *section_derivs, const double *h_block,
                                    const double *points, const double *weights,
                                    const double *log_omega_sq, double *per_point_buf) {
  ST_TRY(cy3q_adamgrad_upload_static_inputs(g, section_values, section_derivs, points, weights,
                                            log_omega_sq));
  return cy3q_adamgrad_compute_per_point_h_only(g, h_block, per_point_buf);
}

int cy3q_adamgrad_total_device_bytes(const cy3q_adamgrad *g, int64_t *bytes) {
  if (!g || !bytes) return fail(CY3Q_ERR_ARG, "NULL argument");
  *bytes = (int64_t)g->s->device_bytes();
  return CY3Q_OK;
}

}  // extern "C"
