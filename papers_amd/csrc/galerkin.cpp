// Host side of the Galerkin entry points: metric Laplacian spectra, the Hermitian eigensolver
// and the Yukawa triple overlaps (include/cy3q.h). The host orders the launches, allocates
// device buffers and moves inputs and results. Every arithmetic stage is a HIP kernel
// (k_laplacian.hip, k_eigen.hip, k_yukawa.hip). The host only sorts the final eigenvalues
// and picks the columns that the deflation keeps.
//
// References: route34/metric_laplacian.rs (:296-693) and route34/yukawa_overlap_real.rs
// (:222-525).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/cy3q.h"
#include "cy3q_internal.h"

using namespace cy3q;

namespace {

#define G_HIP(expr)                                                                          \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return set_error(CY3Q_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
  } while (0)

#define G_TRY(expr)                  \
  do {                               \
    int rc_ = (expr);                \
    if (rc_ != CY3Q_OK) return rc_;  \
  } while (0)

// Device allocations of one call, freed when the call returns. The stream is either borrowed
// (a solver's) or owned (created on `device` for a stand-alone call).
struct Ctx {
  int device = 0;
  hipStream_t st = nullptr;
  bool own = false;
  std::vector<void *> mem;
  ~Ctx() {
    if (st) hipStreamSynchronize(st);
    for (void *p : mem) hipFree(p);
    if (own && st) hipStreamDestroy(st);
  }
  int open(int dev) {
    device = dev;
    G_HIP(hipSetDevice(dev));
    G_HIP(hipStreamCreate(&st));
    own = true;
    return CY3Q_OK;
  }
  template <class T>
  int alloc(T **p, size_t count) {
    void *q = nullptr;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess)
      return set_error(CY3Q_ERR_HIP,
                       "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    mem.push_back(q);
    *p = (T *)q;
    return CY3Q_OK;
  }
  template <class T>
  int zeros(T **p, size_t count) {
    G_TRY(alloc(p, count));
    G_HIP(hipMemsetAsync(*p, 0, std::max<size_t>(count, 1) * sizeof(T), st));
    return CY3Q_OK;
  }
  template <class T>
  int upload(T **p, const T *src, size_t count) {
    G_TRY(alloc(p, count));
    G_HIP(hipMemcpyAsync(*p, src, count * sizeof(T), hipMemcpyHostToDevice, st));
    return CY3Q_OK;
  }
  void release(void *p) {
    auto it = std::find(mem.begin(), mem.end(), p);
    if (it != mem.end()) {
      hipStreamSynchronize(st);
      hipFree(p);
      mem.erase(it);
    }
  }
};

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// Per-stage device time of the calling thread's last Galerkin call (HIP events on its stream).
enum GStage {
  GS_ROWS = 0,   // per-point metric factor and row generation
  GS_HERK,       // Hermitian rank-k accumulation (FP64 MFMA, or the fp16 split)
  GS_FINALIZE,   // complex M, K from the real accumulators
  GS_EIGEN_M,    // factorisation of M: inverse Cholesky, or eigenvectors when M is singular
  GS_REDUCE,     // M^-1/2 reduction GEMMs (or M^-1 K)
  GS_EIGEN_C,    // eigensolve of the reduced pencil
  GS_YUK_MODES,  // mode values psi = S C
  GS_YUK_TRIPLES,
  GS_SPLIT,      // row range and fp16 hi/lo planes of the split rank-k update
  GS_COUNT
};
const char *kGStageNames[GS_COUNT] = {"rows", "herk", "finalize", "factor_mass",
                                      "reduce", "eigen_reduced", "yukawa_modes",
                                      "yukawa_triples", "split_rows"};
thread_local double g_stage_ms[GS_COUNT];
thread_local int64_t g_stage_cnt[GS_COUNT];

struct StageTimer {
  hipStream_t st;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev;
  explicit StageTimer(hipStream_t s) : st(s) {
    for (int i = 0; i < GS_COUNT; i++) {
      g_stage_ms[i] = 0.0;
      g_stage_cnt[i] = 0;
    }
  }
  void begin() {
    hipEvent_t a = nullptr;
    hipEventCreate(&a);
    hipEventRecord(a, st);
    ev.push_back({-1, {a, nullptr}});
  }
  void end(int id) {
    hipEvent_t b = nullptr;
    hipEventCreate(&b);
    hipEventRecord(b, st);
    ev.back().first = id;
    ev.back().second.second = b;
  }
  ~StageTimer() {
    for (auto &e : ev) {
      if (e.second.second) {
        hipEventSynchronize(e.second.second);
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e.second.first, e.second.second) == hipSuccess && e.first >= 0) {
          g_stage_ms[e.first] += ms;
          g_stage_cnt[e.first] += 1;
        }
        hipEventDestroy(e.second.second);
      }
      hipEventDestroy(e.second.first);
    }
  }
};

// ------------------------------------------------------------------------------------------
// Hermitian eigensolver: parallel Jacobi sweeps (k_jacobi_round) until the off-diagonal
// Frobenius norm is at most tol * ||A||_F (or is zero), or max_sweeps sweeps. a: device
// [n][n] complex; shift is added to the diagonal first. eig: ascending; vecs (device
// [n][n], column j for eig[j]) when not null.
// ------------------------------------------------------------------------------------------
// Device buffers of one Jacobi eigensolve of an n x n matrix (np = n rounded up to even):
// A0, A1 (and V0, V1 for vectors) np x np, t0, t1 np tags, part 2 offnorm_parts(), nrm 2,
// diag np, cols np.
struct HeevBufs {
  double2 *A0 = nullptr, *A1 = nullptr, *V0 = nullptr, *V1 = nullptr;
  int *t0 = nullptr, *t1 = nullptr, *cols = nullptr;
  double *part = nullptr, *nrm = nullptr, *diag = nullptr;
};

// The parallel Jacobi sweeps on b (hermitian_jacobi_n, metric_laplacian.rs:616-693) until the
// relative off-diagonal norm is below tol or max_sweeps have run; eig (host, n values) ascending,
// vecs (device, [n][n], column j for eig[j]) when b holds V0/V1. d and tag are host arrays of
// np entries. No allocation.
int heev_core(hipStream_t st, const double2 *a, int n, int64_t max_sweeps, double tol,
              double shift, HeevBufs b, double *eig, double2 *vecs, int64_t *sweeps_out,
              double *d, int *tag) {
  const int np = n + (n & 1);
  launch_jacobi_init(st, a, n, np, shift, b.A0, vecs ? b.V0 : nullptr, b.t0);
  int64_t sw = 0;
  for (;;) {
    double h2[2];
    launch_offnorm(st, b.A0, np, b.part, b.nrm);
    G_HIP(hipMemcpyAsync(h2, b.nrm, sizeof(h2), hipMemcpyDeviceToHost, st));
    G_HIP(hipStreamSynchronize(st));
    if (!std::isfinite(h2[0]) || !std::isfinite(h2[1]))
      return set_error(CY3Q_ERR_ARG, "non-finite entries in the eigenproblem");
    if (h2[0] == 0.0 || std::sqrt(h2[0]) <= tol * std::sqrt(h2[1]) || sw >= max_sweeps) break;
    for (int r = 0; r < np - 1; r++) {
      launch_jacobi_round(st, b.A0, b.A1, b.V0, b.V1, b.t0, b.t1, np, vecs ? 1 : 0);
      std::swap(b.A0, b.A1);
      std::swap(b.V0, b.V1);
      std::swap(b.t0, b.t1);
    }
    G_HIP(hipGetLastError());
    sw++;
  }
  launch_diag_re(st, b.A0, np, b.diag);
  G_HIP(hipMemcpyAsync(d, b.diag, sizeof(double) * np, hipMemcpyDeviceToHost, st));
  G_HIP(hipMemcpyAsync(tag, b.t0, sizeof(int) * np, hipMemcpyDeviceToHost, st));
  G_HIP(hipStreamSynchronize(st));
  // The slots holding the n eigenvalues (the padding slot of an odd n has tag n), ascending by
  // value; equal values keep slot order (an in-place sort with the slot as the tie-break).
  int m = 0;
  for (int x = 0; x < np; x++)
    if (tag[x] < n) tag[m++] = x;
  std::sort(tag, tag + m, [&](int a_, int b_) { return d[a_] < d[b_] || (d[a_] == d[b_] && a_ < b_); });
  for (int j = 0; j < n; j++) eig[j] = d[tag[j]];
  if (vecs) {
    G_HIP(hipMemcpyAsync(b.cols, tag, sizeof(int) * n, hipMemcpyHostToDevice, st));
    launch_gather_cols(st, b.V0, np, n, b.cols, nullptr, n, vecs);
    G_HIP(hipStreamSynchronize(st));
  }
  if (sweeps_out) *sweeps_out = sw;
  G_HIP(hipGetLastError());
  return CY3Q_OK;
}

int heev(Ctx &c, const double2 *a, int n, int64_t max_sweeps, double tol, double shift,
         std::vector<double> &eig, double2 *vecs, int64_t *sweeps_out) {
  const int np = n + (n & 1);
  HeevBufs b;
  G_TRY(c.alloc(&b.A0, (size_t)np * np));
  G_TRY(c.alloc(&b.A1, (size_t)np * np));
  if (vecs) {
    G_TRY(c.alloc(&b.V0, (size_t)np * np));
    G_TRY(c.alloc(&b.V1, (size_t)np * np));
  }
  G_TRY(c.alloc(&b.t0, np));
  G_TRY(c.alloc(&b.t1, np));
  G_TRY(c.alloc(&b.cols, np));
  G_TRY(c.alloc(&b.part, 2 * offnorm_parts()));
  G_TRY(c.alloc(&b.nrm, 2));
  G_TRY(c.alloc(&b.diag, np));
  std::vector<double> d(np);
  std::vector<int> tag(np);
  eig.resize(n);
  return heev_core(c.st, a, n, max_sweeps, tol, shift, b, eig.data(), vecs, sweeps_out, d.data(),
                   tag.data());
}

}  // namespace

// Scratch of cy3q_hermitian_eigen_noalloc: the eigensolve buffers of HeevBufs for matrices up to
// kNoAllocMaxN, carved in order A0, A1, V0, V1 (kNoAllocMaxN^2 complex each), part
// (2 kNoAllocParts), nrm (2), diag (kNoAllocMaxN doubles), t0, t1, cols (kNoAllocMaxN ints
// each): 16.8 MB per device, held until the process ends. Every call on a device shares it.
constexpr int kNoAllocMaxN = 512;
constexpr int kNoAllocParts = 256;
constexpr size_t kNoAllocBytes =
    4 * sizeof(double2) * (size_t)kNoAllocMaxN * kNoAllocMaxN +
    sizeof(double) * (2 * kNoAllocParts + 2 + kNoAllocMaxN) + sizeof(int) * 3 * kNoAllocMaxN;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

// The device address of `scratch` on `device` (each device holds its own copy), looked up once
// per device with that device current. The host-side name of the array is not a device
// address and is never passed to a kernel or a copy.
int noalloc_scratch(int device, unsigned char **out) {
  static std::mutex mu;
  static unsigned char *addr[64] = {};
  int count = 0;
  G_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count || device >= 64)
    return set_error(CY3Q_ERR_ARG, "bad device ordinal");
  G_HIP(hipSetDevice(device));
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(CY3Q_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                         hipGetErrorString(e));
    addr[device] = static_cast<unsigned char *>(p);
  }
  *out = addr[device];
  return CY3Q_OK;
}

// Frobenius norm of a device [n][n] complex matrix.
int frob(Ctx &c, const double2 *a, int n, double *out) {
  double *part, *nrm;
  G_TRY(c.alloc(&part, 2 * offnorm_parts()));
  G_TRY(c.alloc(&nrm, 2));
  launch_offnorm(c.st, a, n, part, nrm);
  double h2[2];
  G_HIP(hipMemcpyAsync(h2, nrm, sizeof(h2), hipMemcpyDeviceToHost, c.st));
  G_HIP(hipStreamSynchronize(c.st));
  *out = std::sqrt(h2[1]);
  return CY3Q_OK;
}

// W = L^-1 with M + shift I = L L^H (FP64 tiled elimination of k_hsolve.hip). *ok = false when
// the matrix is not positive definite (then W is not usable).
int inverse_cholesky(Ctx &c, const double2 *M, int n, double shift, double2 *W, bool *ok) {
  const int nblk = (n + 31) / 32;
  double *a, *r, *einv;
  int32_t *status;
  G_TRY(c.alloc(&a, (size_t)2 * n * n));
  G_TRY(c.alloc(&r, (size_t)2 * n * n));
  G_TRY(c.alloc(&einv, (size_t)nblk * 32 * 32 * 2));
  G_TRY(c.alloc(&status, 4));
  G_HIP(hipMemsetAsync(W, 0, sizeof(double2) * n * n, c.st));  // the upper triangle stays zero
  launch_elim_init(c.st, reinterpret_cast<const double *>(M), n, a, r);
  if (shift != 0.0) launch_add_diag(c.st, reinterpret_cast<double2 *>(a), n, shift);
  launch_inverse_cholesky(c.st, a, r, einv, reinterpret_cast<double *>(W), n, status);
  int32_t h = -1;
  G_HIP(hipMemcpyAsync(&h, status, sizeof(h), hipMemcpyDeviceToHost, c.st));
  G_HIP(hipStreamSynchronize(c.st));
  *ok = h == 0;
  for (void *p : {(void *)a, (void *)r, (void *)einv, (void *)status}) c.release(p);
  return CY3Q_OK;
}

// Spectrum step of compute_metric_laplacian_spectrum (metric_laplacian.rs:342-389):
// (M + lambda I)^-1 K with lambda = max(reg ||M||_F, 1e-30), through the eigenvectors of M, then
// the Hermitian part, then its eigenvalues as |lambda| ascending.
int reference_spectrum(Ctx &c, const double2 *M, const double2 *K, int nb, int64_t max_sweeps,
                       double tol, double reg, std::vector<double> &full, StageTimer &tm) {
  double fm;
  G_TRY(frob(c, M, nb, &fm));
  const double lambda = std::max(fm * reg, 1.0e-30);
  double2 *U, *Y, *Minv, *L;
  G_TRY(c.alloc(&U, (size_t)nb * nb));
  G_TRY(c.alloc(&Y, (size_t)nb * nb));
  G_TRY(c.alloc(&Minv, (size_t)nb * nb));
  G_TRY(c.alloc(&L, (size_t)nb * nb));
  // (M + lambda I)^-1 = W^H W from the inverse Cholesky factor; the eigenvectors of M are the
  // fallback when the factorisation meets a non-positive pivot.
  bool chol_ok = false;
  tm.begin();
  G_TRY(inverse_cholesky(c, M, nb, lambda, U, &chol_ok));
  tm.end(GS_EIGEN_M);
  if (chol_ok) {
    tm.begin();
    launch_zgemm(c.st, U, U, Minv, nb, nb, nb, nb, nb, nb, 1, 0);  // W^H W
    launch_zgemm(c.st, Minv, K, L, nb, nb, nb, nb, nb, nb, 0, 0);  // M^-1 K
    launch_herm_project(c.st, L, nb);
    tm.end(GS_REDUCE);
    std::vector<double> ev;
    tm.begin();
    G_TRY(heev(c, L, nb, max_sweeps, tol, 0.0, ev, nullptr, nullptr));
    tm.end(GS_EIGEN_C);
    full.resize(nb);
    for (int j = 0; j < nb; j++) full[j] = std::fabs(ev[j]);
    std::stable_sort(full.begin(), full.end());
    return CY3Q_OK;
  }
  std::vector<double> mu;
  tm.begin();
  G_TRY(heev(c, M, nb, max_sweeps, tol, lambda, mu, U, nullptr));
  tm.end(GS_EIGEN_M);
  std::vector<double> inv(nb);
  for (int j = 0; j < nb; j++) inv[j] = 1.0 / mu[j];
  std::vector<int> ident(nb);
  std::iota(ident.begin(), ident.end(), 0);
  double *d_inv;
  int *d_id;
  G_TRY(c.upload(&d_inv, inv.data(), nb));
  G_TRY(c.upload(&d_id, ident.data(), nb));
  tm.begin();
  launch_gather_cols(c.st, U, nb, nb, d_id, d_inv, nb, Y);           // U diag(1/mu)
  launch_zgemm(c.st, Y, U, Minv, nb, nb, nb, nb, nb, nb, 0, 1);       // ... U^H
  launch_zgemm(c.st, Minv, K, L, nb, nb, nb, nb, nb, nb, 0, 0);       // M^-1 K
  launch_herm_project(c.st, L, nb);
  tm.end(GS_REDUCE);
  std::vector<double> ev;
  tm.begin();
  G_TRY(heev(c, L, nb, max_sweeps, tol, 0.0, ev, nullptr, nullptr));
  tm.end(GS_EIGEN_C);
  full.resize(nb);
  for (int j = 0; j < nb; j++) full[j] = std::fabs(ev[j]);
  std::stable_sort(full.begin(), full.end());
  return CY3Q_OK;
}

// Generalized K v = lambda M v by the canonical reduction X = U_keep diag(mu^-1/2):
// C = X^H K X, the eigenvalues of its Hermitian part ascending.
// The lowest n_want eigenvalues (ascending) of a Hermitian matrix: Householder
// tridiagonalisation of a copy, then bisection.
int tridiag_eigvals(Ctx &c, const double2 *a, int n, int n_want, std::vector<double> &ev) {
  if (!tridiag_smem_ok(n)) return set_error(CY3Q_ERR_ARG, "eigenproblem too large");
  n_want = std::max(0, std::min(n_want, n));
  double2 *A, *v, *p;
  double *scal, *d, *e, *dev;
  G_TRY(c.alloc(&A, (size_t)n * n));
  G_TRY(c.alloc(&v, (size_t)n));
  G_TRY(c.alloc(&p, (size_t)(2 * n)));  // p, then the contiguous Householder column
  G_TRY(c.alloc(&scal, 3));
  G_TRY(c.alloc(&d, (size_t)n));
  G_TRY(c.alloc(&e, (size_t)n));
  G_TRY(c.alloc(&dev, (size_t)std::max(n_want, 1)));
  G_HIP(hipMemcpyAsync(A, a, sizeof(double2) * n * n, hipMemcpyDeviceToDevice, c.st));
  launch_hh_tridiag(c.st, A, n, v, p, scal, d, e);
  launch_tridiag_bisect(c.st, d, e, n, n_want, dev);
  G_HIP(hipGetLastError());
  ev.assign(n_want, 0.0);
  if (n_want > 0)
    G_HIP(hipMemcpyAsync(ev.data(), dev, sizeof(double) * n_want, hipMemcpyDeviceToHost, c.st));
  G_HIP(hipStreamSynchronize(c.st));
  for (double x : ev)
    if (!std::isfinite(x)) return set_error(CY3Q_ERR_ARG, "non-finite entries in the eigenproblem");
  for (void *q : {(void *)A, (void *)v, (void *)p, (void *)scal, (void *)d, (void *)e, (void *)dev})
    c.release(q);
  return CY3Q_OK;
}

// The lowest n_want eigenvalues of the pencil (K, M).
int generalized_spectrum(Ctx &c, const double2 *M, const double2 *K, int nb, int64_t max_sweeps,
                         double tol, double defl, int n_want, std::vector<double> &ev,
                         int64_t *rank, StageTimer &tm) {
  double2 *U;
  G_TRY(c.alloc(&U, (size_t)nb * nb));
  // Positive definite M: C = W K W^H with W = chol(M)^-1, whose eigenvalues are those of the
  // pencil. Otherwise the canonical reduction below deflates M's null space.
  {
    bool chol_ok = false;
    tm.begin();
    G_TRY(inverse_cholesky(c, M, nb, 0.0, U, &chol_ok));
    tm.end(GS_EIGEN_M);
    if (chol_ok) {
      double2 *T, *C;
      G_TRY(c.alloc(&T, (size_t)nb * nb));
      G_TRY(c.alloc(&C, (size_t)nb * nb));
      tm.begin();
      launch_zgemm(c.st, U, K, T, nb, nb, nb, nb, nb, nb, 0, 0);  // W K
      launch_zgemm(c.st, T, U, C, nb, nb, nb, nb, nb, nb, 0, 1);  // W K W^H
      launch_herm_project(c.st, C, nb);
      tm.end(GS_REDUCE);
      if (rank) *rank = nb;
      tm.begin();
      G_TRY(tridiag_eigvals(c, C, nb, n_want, ev));
      tm.end(GS_EIGEN_C);
      return CY3Q_OK;
    }
  }
  std::vector<double> mu;
  tm.begin();
  G_TRY(heev(c, M, nb, max_sweeps, tol, 0.0, mu, U, nullptr));
  tm.end(GS_EIGEN_M);
  const double mmax = mu.empty() ? 0.0 : mu.back();
  std::vector<int> keep;
  std::vector<double> sc;
  for (int j = 0; j < nb; j++)
    if (mu[j] > defl * mmax && mu[j] > 0.0) {
      keep.push_back(j);
      sc.push_back(1.0 / std::sqrt(mu[j]));
    }
  const int r = (int)keep.size();
  if (rank) *rank = r;
  if (r == 0) return set_error(CY3Q_ERR_ARG, "mass matrix has no positive eigenvalue");
  int *d_keep;
  double *d_sc;
  double2 *X, *T, *C;
  G_TRY(c.upload(&d_keep, keep.data(), r));
  G_TRY(c.upload(&d_sc, sc.data(), r));
  G_TRY(c.alloc(&X, (size_t)nb * r));
  G_TRY(c.alloc(&T, (size_t)nb * r));
  G_TRY(c.alloc(&C, (size_t)r * r));
  tm.begin();
  launch_gather_cols(c.st, U, nb, nb, d_keep, d_sc, r, X);
  launch_zgemm(c.st, K, X, T, nb, r, nb, nb, r, r, 0, 0);  // K X
  launch_zgemm(c.st, X, T, C, r, r, nb, r, r, r, 1, 0);    // X^H K X
  launch_herm_project(c.st, C, r);
  tm.end(GS_REDUCE);
  tm.begin();
  G_TRY(tridiag_eigvals(c, C, r, n_want, ev));
  tm.end(GS_EIGEN_C);
  return CY3Q_OK;
}

// build_test_basis (metric_laplacian.rs:87-115): (dz, dw) with dz + dw <= max, (0, 0) excluded,
// then monomials_4(dz) x monomials_4(dw), each in a, b, c ascending order (:118-130).
std::vector<uint8_t> lap_basis(uint32_t max_deg) {
  std::vector<uint8_t> out;
  auto m4 = [](uint32_t d) {
    std::vector<std::array<uint8_t, 4>> v;
    for (uint32_t a = 0; a <= d; a++)
      for (uint32_t b = 0; b <= d - a; b++)
        for (uint32_t cc = 0; cc <= d - a - b; cc++)
          v.push_back({(uint8_t)a, (uint8_t)b, (uint8_t)cc, (uint8_t)(d - a - b - cc)});
    return v;
  };
  for (uint32_t dz = 0; dz <= max_deg; dz++)
    for (uint32_t dw = 0; dw <= max_deg - dz; dw++) {
      if (dz == 0 && dw == 0) continue;
      auto zm = m4(dz), wm = m4(dw);
      for (auto &a : zm)
        for (auto &b : wm) {
          out.insert(out.end(), a.begin(), a.end());
          out.insert(out.end(), b.begin(), b.end());
        }
    }
  return out;
}

// Raw degree-d monomials in 5 variables in the order of build_degree_k_quintic_monomials
// (quintic.rs:3192-3209): e0 up, e1 up, e2 up, e3 up, e4 = d - sum.
std::vector<uint8_t> raw_monomials(uint32_t d) {
  std::vector<uint8_t> out;
  const int k = (int)d;
  for (int e0 = 0; e0 <= k; e0++)
    for (int e1 = 0; e1 <= k - e0; e1++)
      for (int e2 = 0; e2 <= k - e0 - e1; e2++)
        for (int e3 = 0; e3 <= k - e0 - e1 - e2; e3++)
          out.insert(out.end(), {(uint8_t)e0, (uint8_t)e1, (uint8_t)e2, (uint8_t)e3,
                                 (uint8_t)(k - e0 - e1 - e2 - e3)});
  return out;
}

// Device Yukawa plan for n points and nm modes.
struct YukWs {
  YukPlan pl;
  int init(Ctx &c, int nm, int64_t n) {
    YukPlanShape sh = yukawa_plan_shape(nm, n);
    static_cast<YukDims &>(pl) = sh;
    short2 *pairs, *tiles;
    int4 *parts;
    short4 *trip;
    G_TRY(c.upload(&pairs, sh.pairs.data(), sh.pairs.size()));
    G_TRY(c.upload(&tiles, sh.tiles.data(), sh.tiles.size()));
    G_TRY(c.upload(&parts, sh.parts.data(), sh.parts.size()));
    G_TRY(c.upload(&trip, sh.trip.data(), sh.trip.size()));
    pl.pairs = pairs;
    pl.tiles = tiles;
    pl.parts = parts;
    pl.trip = trip;
    G_TRY(c.alloc(&pl.psi_v, (size_t)n * sh.vldp));
    G_TRY(c.alloc(&pl.g, (size_t)n));
    G_TRY(c.alloc(&pl.partials, (size_t)sh.n_chunks * sh.ntiles * 256));
    G_TRY(c.alloc(&pl.cfull, (size_t)sh.cfull_rows * sh.vld));
    return CY3Q_OK;
  }
};

}  // namespace

extern "C" {

int cy3q_galerkin_stage_times(double *ms, int64_t *count, int cap, int *n_stages) {
  if (n_stages) *n_stages = GS_COUNT;
  for (int i = 0; i < GS_COUNT && i < cap; i++) {
    if (ms) ms[i] = g_stage_ms[i];
    if (count) count[i] = g_stage_cnt[i];
  }
  return CY3Q_OK;
}

const char *cy3q_galerkin_stage_name(int stage) {
  return (stage >= 0 && stage < GS_COUNT) ? kGStageNames[stage] : "";
}

void cy3q_laplacian_config_default(cy3q_laplacian_config *cfg) {
  if (!cfg) return;
  cfg->max_total_degree = 4;
  cfg->n_low_eigenvalues = 15;
  cfg->jacobi_max_sweeps = 128;
  cfg->jacobi_tol = 1.0e-12;
  cfg->mass_regularisation = 1.0e-10;
}

int cy3q_laplacian_basis(uint32_t max_total_degree, uint32_t *exps, int64_t cap, int64_t *n) {
  if (max_total_degree > 8) return set_error(CY3Q_ERR_ARG, "max_total_degree > 8");
  std::vector<uint8_t> b = lap_basis(max_total_degree);
  const int64_t cnt = (int64_t)(b.size() / 8);
  if (n) *n = cnt;
  if (exps) {
    if (cap < cnt) return set_error(CY3Q_ERR_ARG, "output capacity too small");
    for (size_t i = 0; i < b.size(); i++) exps[i] = b[i];
  }
  return CY3Q_OK;
}

int cy3q_hermitian_eigen(const double *a, int64_t n, int64_t max_sweeps, double tol, int device,
                         double *eig, double *vecs, int64_t *sweeps) {
  if (!a || !eig || n <= 0 || n > 8192) return set_error(CY3Q_ERR_ARG, "bad matrix arguments");
  Ctx c;
  G_TRY(c.open(device));
  double2 *d_a, *d_v = nullptr;
  G_TRY(c.upload(&d_a, reinterpret_cast<const double2 *>(a), (size_t)(n * n)));
  if (vecs) G_TRY(c.alloc(&d_v, (size_t)(n * n)));
  std::vector<double> ev;
  G_TRY(heev(c, d_a, (int)n, max_sweeps, tol, 0.0, ev, d_v, sweeps));
  std::memcpy(eig, ev.data(), sizeof(double) * n);
  if (vecs)
    G_HIP(hipMemcpyAsync(vecs, d_v, sizeof(double2) * n * n, hipMemcpyDeviceToHost, c.st));
  G_HIP(hipStreamSynchronize(c.st));
  return CY3Q_OK;
}

// (T + shift I)^-1 of a Hermitian positive definite T through the inverse Cholesky factor
// (k_hsolve.hip): W = L^-1, T + shift I = L L^H, inverse = W^H W. One workspace and stream per
// device, grown to the largest n seen and kept until the process ends; calls on one device
// are serialised by its mutex.
namespace {
struct InvWorkspace {
  std::mutex mu;
  hipStream_t st = nullptr;
  int64_t cap = 0;
  double *a = nullptr, *r = nullptr, *einv = nullptr, *w = nullptr, *g = nullptr;
  int32_t *status = nullptr;
};
constexpr int kInvDevices = 64;
InvWorkspace g_inv_ws[kInvDevices];
}  // namespace

int cy3q_hermitian_inverse(const double *t, int64_t n, double shift, int device, double *out) {
  if (!t || !out || n <= 0 || n > 4096) return set_error(CY3Q_ERR_ARG, "bad matrix arguments");
  if (device < 0 || device >= kInvDevices) return set_error(CY3Q_ERR_ARG, "bad device index");
  if (!std::isfinite(shift)) return set_error(CY3Q_ERR_ARG, "shift must be finite");
  InvWorkspace &ws = g_inv_ws[device];
  std::lock_guard<std::mutex> lock(ws.mu);
  G_HIP(hipSetDevice(device));
  if (!ws.st) G_HIP(hipStreamCreate(&ws.st));
  const int64_t nblk = (n + 31) / 32;
  if (ws.cap < n) {
    for (void *p : {(void *)ws.a, (void *)ws.r, (void *)ws.einv, (void *)ws.w, (void *)ws.g,
                    (void *)ws.status})
      if (p) G_HIP(hipFree(p));
    ws.a = ws.r = ws.einv = ws.w = ws.g = nullptr;
    ws.status = nullptr;
    ws.cap = 0;
    const size_t mat = sizeof(double) * 2 * n * n;
    G_HIP(hipMalloc(&ws.a, mat));
    G_HIP(hipMalloc(&ws.r, mat));
    G_HIP(hipMalloc(&ws.w, mat));
    G_HIP(hipMalloc(&ws.g, mat));
    G_HIP(hipMalloc(&ws.einv, sizeof(double) * 2 * 32 * 32 * nblk));
    G_HIP(hipMalloc(&ws.status, sizeof(int32_t)));
    ws.cap = n;
  }
  const size_t bytes = sizeof(double) * 2 * n * n;
  G_HIP(hipMemcpyAsync(ws.g, t, bytes, hipMemcpyHostToDevice, ws.st));
  launch_elim_init(ws.st, ws.g, (int)n, ws.a, ws.r);
  if (shift != 0.0) launch_add_diag(ws.st, reinterpret_cast<double2 *>(ws.a), (int)n, shift);
  launch_fill_f64(ws.st, ws.w, 2 * n * n, 0.0);
  launch_inverse_cholesky(ws.st, ws.a, ws.r, ws.einv, ws.w, (int)n, ws.status);
  launch_gram_residual(ws.st, ws.w, (int)n, ws.g, ws.status, nullptr, nullptr, nullptr);
  G_HIP(hipGetLastError());
  int32_t st_h = -1;
  G_HIP(hipMemcpyAsync(&st_h, ws.status, sizeof(st_h), hipMemcpyDeviceToHost, ws.st));
  G_HIP(hipMemcpyAsync(out, ws.g, bytes, hipMemcpyDeviceToHost, ws.st));
  G_HIP(hipStreamSynchronize(ws.st));
  if (st_h != 0)
    return set_error(CY3Q_ERR_NOT_PD, "cy3q_hermitian_inverse: the matrix (with its shift) is "
                                      "not numerically positive definite");
  return CY3Q_OK;
}

int cy3q_hermitian_eigen_noalloc(const double *d_a, int64_t n, int64_t max_sweeps, double tol,
                                 int device, double *d_eig, double *d_vecs, int64_t *sweeps) {
  if (!d_a || !d_eig || n <= 0) return set_error(CY3Q_ERR_ARG, "bad matrix arguments");
  if (n > kNoAllocMaxN)
    return set_error(CY3Q_ERR_ARG, "cy3q_hermitian_eigen_noalloc: n = " + std::to_string(n) +
                                       " exceeds its scratch limit of " +
                                       std::to_string(kNoAllocMaxN));
  if (offnorm_parts() > kNoAllocParts)
    return set_error(CY3Q_ERR_STATE, "offnorm partials exceed the scratch layout");
  unsigned char *base = nullptr;
  G_TRY(noalloc_scratch(device, &base));
  const size_t mat = (size_t)kNoAllocMaxN * kNoAllocMaxN;
  HeevBufs b;
  b.A0 = reinterpret_cast<double2 *>(base);
  b.A1 = b.A0 + mat;
  b.V0 = b.A1 + mat;
  b.V1 = b.V0 + mat;
  b.part = reinterpret_cast<double *>(b.V1 + mat);
  b.nrm = b.part + 2 * kNoAllocParts;
  b.diag = b.nrm + 2;
  b.t0 = reinterpret_cast<int *>(b.diag + kNoAllocMaxN);
  b.t1 = b.t0 + kNoAllocMaxN;
  b.cols = b.t1 + kNoAllocMaxN;
  double d[kNoAllocMaxN], eig[kNoAllocMaxN];
  int tag[kNoAllocMaxN];
  // The null stream: no stream is created.
  const hipStream_t st = nullptr;
  G_TRY(heev_core(st, reinterpret_cast<const double2 *>(d_a), (int)n, max_sweeps, tol, 0.0, b, eig,
                  reinterpret_cast<double2 *>(d_vecs), sweeps, d, tag));
  G_HIP(hipMemcpyAsync(d_eig, eig, sizeof(double) * n, hipMemcpyHostToDevice, st));
  G_HIP(hipStreamSynchronize(st));
  return CY3Q_OK;
}

int cy3q_metric_laplacian_spectrum(const double *points, const double *weights,
                                   const double *omega, int64_t n,
                                   const cy3q_laplacian_config *cfg_in, int device,
                                   double *eig_low, double *eig_full, double *mass,
                                   double *stiffness, int64_t *basis_dim) {
  cy3q_laplacian_config cfg;
  cy3q_laplacian_config_default(&cfg);
  if (cfg_in) cfg = *cfg_in;
  if (cfg.max_total_degree > 6) return set_error(CY3Q_ERR_ARG, "max_total_degree > 6");
  std::vector<uint8_t> ex = lap_basis(cfg.max_total_degree);
  const int nb = (int)(ex.size() / 8);
  if (basis_dim) *basis_dim = nb;
  if (nb == 0 || n <= 0) return CY3Q_OK;  // the reference returns an empty spectrum (:305-316)
  if (!points || !weights || !omega) return set_error(CY3Q_ERR_ARG, "null input");
  Ctx c;
  G_TRY(c.open(device));
  StageTimer tm(c.st);
  const int64_t mc = round_up(2 * (int64_t)nb, 128);
  const int nder = 8;
  const int64_t chunk = std::min<int64_t>(n, std::max<int64_t>(1024, (int64_t)(3e8 / (9.0 * mc))));
  double *d_pts, *d_w, *d_om, *x0, *x1, *wv, *tw, *cmass, *ck[8];
  uint8_t *d_ex;
  G_TRY(c.upload(&d_pts, points, (size_t)(16 * n)));
  G_TRY(c.upload(&d_w, weights, (size_t)n));
  G_TRY(c.upload(&d_om, omega, (size_t)(2 * n)));
  G_TRY(c.upload(&d_ex, ex.data(), ex.size()));
  G_TRY(c.zeros(&x0, (size_t)(chunk * mc)));
  G_TRY(c.zeros(&x1, (size_t)(nder * chunk * mc)));
  G_TRY(c.alloc(&wv, (size_t)n));
  G_TRY(c.alloc(&tw, 1));
  G_TRY(c.zeros(&cmass, (size_t)(mc * mc)));
  for (int q = 0; q < nder; q++) G_TRY(c.zeros(&ck[q], (size_t)(mc * mc)));
  for (int64_t p0 = 0; p0 < n; p0 += chunk) {
    const int64_t cn = std::min(chunk, n - p0);
    tm.begin();
    launch_lap_rows_ambient(c.st, d_pts + 16 * p0, d_w + p0, d_om + 2 * p0, cn, d_ex, nb, mc, x0,
                            x1, wv + p0);
    tm.end(GS_ROWS);
    const double *xs[9];
    int64_t rows[9];
    double *cs[9];
    xs[0] = x0;
    rows[0] = cn;
    cs[0] = cmass;
    for (int q = 0; q < nder; q++) {
      xs[1 + q] = x1 + (int64_t)q * cn * mc;
      rows[1 + q] = cn;
      cs[1 + q] = ck[q];
    }
    tm.begin();
    launch_herk_f64(c.st, xs, rows, cs, 1 + nder, mc);
    tm.end(GS_HERK);
  }
  launch_sum_f64(c.st, wv, n, tw);
  double2 *M, *K;
  G_TRY(c.alloc(&M, (size_t)nb * nb));
  G_TRY(c.alloc(&K, (size_t)nb * nb));
  tm.begin();
  launch_lap_finalize(c.st, cmass, ck, nder, nb, mc, tw, 1, M, K);
  tm.end(GS_FINALIZE);
  G_HIP(hipGetLastError());
  c.release(x0);
  c.release(x1);
  if (mass) G_HIP(hipMemcpyAsync(mass, M, sizeof(double2) * nb * nb, hipMemcpyDeviceToHost, c.st));
  if (stiffness)
    G_HIP(hipMemcpyAsync(stiffness, K, sizeof(double2) * nb * nb, hipMemcpyDeviceToHost, c.st));
  std::vector<double> full;
  G_TRY(reference_spectrum(c, M, K, nb, cfg.jacobi_max_sweeps, cfg.jacobi_tol,
                           cfg.mass_regularisation, full, tm));
  const int64_t n_low = std::min<int64_t>(std::max<int64_t>(cfg.n_low_eigenvalues, 0), nb);
  if (eig_low) std::memcpy(eig_low, full.data(), sizeof(double) * n_low);
  if (eig_full) std::memcpy(eig_full, full.data(), sizeof(double) * nb);
  G_HIP(hipStreamSynchronize(c.st));
  return CY3Q_OK;
}

int cy3q_solver_laplacian_spectrum(cy3q_solver *s, uint32_t section_degree, int64_t n_low,
                                   double deflation_tol, int64_t max_sweeps, double *eig_low,
                                   double *mass, double *stiffness, int64_t *basis_dim,
                                   int64_t *rank) {
  if (section_degree < 1 || section_degree > 4)
    return set_error(CY3Q_ERR_ARG, "section_degree must be in 1..4");
  SolverView v;
  G_TRY(solver_view(s, &v));
  Ctx c;
  c.device = v.device;
  c.st = v.st;
  StageTimer tm(c.st);
  std::vector<uint8_t> sm = raw_monomials(section_degree);
  const int ns = (int)(sm.size() / 5);
  const int B = ns * ns;
  if (basis_dim) *basis_dim = B;
  const int64_t n = v.n, m = v.m;
  const bool split = v.galerkin_precision == CY3Q_PREC_SPLIT16;
  // The split update tiles the columns by 256; the FP64 one by 128.
  const int64_t mc = round_up(2 * (int64_t)B, split ? herk_split_tile() : 128);
  const int64_t chunk = std::min<int64_t>(n, std::max<int64_t>(256, (int64_t)(6e8 / (4.0 * mc))));
  // Split path: the point parts of each tile's sum (one accumulator copy each), as many as fill
  // the XCDs' rounds of workgroups best.
  int nparts = 1;
  if (split) {
    int n_cu = 0;
    G_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, c.device));
    const int64_t nt = mc / herk_split_tile(), tiles = 4 * nt * (nt + 1) / 2;
    nparts = herk_split_parts(4, (int)(tiles / 4), n_cu);
  }
  const int64_t stages = (chunk + 15) / 16;
  uint8_t *d_sm;
  double *y0, *yd, *pr, *wv, *tw, *x0, *x1, *cmass, *ck[3];
  double *part = nullptr, *scale = nullptr;
  half8_t *planes = nullptr;
  G_TRY(c.upload(&d_sm, sm.data(), sm.size()));
  G_TRY(c.alloc(&y0, (size_t)(chunk * m)));
  G_TRY(c.alloc(&yd, (size_t)(chunk * 5 * m)));
  G_TRY(c.alloc(&pr, (size_t)(32 * chunk)));
  G_TRY(c.alloc(&wv, (size_t)n));
  G_TRY(c.alloc(&tw, 1));
  x0 = x1 = nullptr;
  if (!split) {  // the split path writes its fp16 planes straight from the row kernel
    G_TRY(c.zeros(&x0, (size_t)(chunk * mc)));
    G_TRY(c.zeros(&x1, (size_t)(3 * chunk * mc)));
  }
  G_TRY(c.zeros(&cmass, (size_t)(nparts * mc * mc)));
  for (int q = 0; q < 3; q++) G_TRY(c.zeros(&ck[q], (size_t)(nparts * mc * mc)));
  // Split path: the metric contraction y = W x of the section and derivative rows also runs on
  // the fp16 split MFMA, with B = W^T as hi/lo planes (one power-of-two scale) and the rows as
  // hi/lo planes with one exponent per row.
  const bool split_rows = split && rowgemm_supported(m);
  double *bscale = nullptr;
  half8_t *wph = nullptr, *wpl = nullptr, *xph = nullptr, *xpl = nullptr;
  int *rex = nullptr;
  if (split) {
    G_TRY(c.alloc(&part, (size_t)std::max<int64_t>(rows_scale_parts(), stages)));
    G_TRY(c.alloc(&scale, 2));
    G_TRY(c.alloc(&planes, (size_t)(4 * 2 * stages * mc * 2)));  // 4 buffers x hi/lo planes
  }
  if (split_rows) {
    const int64_t mcp = rowgemm_cols_pad(m), xunits = (m / 16) * rowgemm_rows_pad(5 * chunk) * 2;
    G_TRY(c.alloc(&bscale, 2));
    G_TRY(c.alloc(&wph, (size_t)((m / 16) * mcp * 2)));
    G_TRY(c.alloc(&wpl, (size_t)((m / 16) * mcp * 2)));
    G_TRY(c.alloc(&xph, (size_t)xunits));
    G_TRY(c.alloc(&xpl, (size_t)xunits));
    G_TRY(c.alloc(&rex, (size_t)(5 * chunk)));
    tm.begin();
    launch_rows_scale_pair(c.st, v.bmat, m * m, nullptr, 0, part, bscale);
    launch_split_rows(c.st, v.bmat, m, m / 16, mcp, bscale, wph, wpl, m);
    tm.end(GS_ROWS);
  }
  for (int64_t p0 = 0; p0 < n; p0 += chunk) {
    const int64_t cn = std::min(chunk, n - p0);
    tm.begin();
    if (split_rows) {
      launch_rows_split_planes(c.st, v.sv + p0 * m, cn, m, rex, xph, xpl);
      launch_rowgemm_split16(c.st, xph, xpl, cn, m, rex, wph, wpl, bscale, y0, m);
      launch_rows_split_planes(c.st, v.sd + p0 * 5 * m, cn * 5, m, rex, xph, xpl);
      launch_rowgemm_split16(c.st, xph, xpl, cn * 5, m, rex, wph, wpl, bscale, yd, m);
    } else {
      launch_trigemm_f64(c.st, v.sv + p0 * m, cn, m, v.bmat, y0, 0, 0);
      launch_trigemm_f64(c.st, v.sd + p0 * 5 * m, cn * 5, m, v.bmat, yd, 0, 0);
    }
    launch_lap_metric(c.st, v.pts + 10 * p0, y0, yd, v.w + p0, v.log_om + p0, cn, m, v.psi, pr,
                      wv + p0);
    if (!split)
      launch_lap_rows_quintic(c.st, v.pts + 10 * p0, pr, cn, d_sm, ns, (int)section_degree, mc,
                              x0, x1);
    tm.end(GS_ROWS);
    const double *xs[4] = {x0, x1, x1 + cn * mc, x1 + 2 * cn * mc};
    int64_t rows[4] = {cn, cn, cn, cn};
    double *cs[4] = {cmass, ck[0], ck[1], ck[2]};
    if (split) {
      // One power-of-two range for the chunk's rows (the largest |entry| over the mass and
      // stiffness rows, from a pass that generates them without storing them), then the rows
      // straight into the fp16 hi/lo planes: buffer 0 the mass rows, buffer 1 + c the rows f_c.
      const int64_t cst = (cn + 15) / 16;
      tm.begin();
      launch_lap_rows_absmax(c.st, v.pts + 10 * p0, pr, cn, d_sm, ns, (int)section_degree, part);
      launch_rows_scale_from_parts(c.st, part, (int)cst, scale);
      half8_t *ph[4], *pl[4];
      for (int q = 0; q < 4; q++) {
        ph[q] = planes + (int64_t)q * 2 * stages * mc * 2;
        pl[q] = ph[q] + stages * mc * 2;
      }
      launch_lap_rows_planes(c.st, v.pts + 10 * p0, pr, cn, d_sm, ns, (int)section_degree, mc,
                             scale, ph, pl);
      tm.end(GS_SPLIT);
      tm.begin();
      const half8_t *cph[4] = {ph[0], ph[1], ph[2], ph[3]}, *cpl[4] = {pl[0], pl[1], pl[2], pl[3]};
      launch_herk_split16(c.st, cph, cpl, cs, 4, cst, mc, nparts, scale);
      tm.end(GS_HERK);
    } else {
      tm.begin();
      launch_herk_f64(c.st, xs, rows, cs, 4, mc);
      tm.end(GS_HERK);
    }
  }
  if (split) {
    tm.begin();
    launch_sum_parts(c.st, cmass, mc * mc, nparts);
    for (int q = 0; q < 3; q++) launch_sum_parts(c.st, ck[q], mc * mc, nparts);
    tm.end(GS_HERK);
    for (void *p : {(void *)part, (void *)scale, (void *)planes, (void *)bscale, (void *)wph,
                    (void *)wpl, (void *)xph, (void *)xpl, (void *)rex})
      c.release(p);
  }
  launch_sum_f64(c.st, wv, n, tw);
  double2 *M, *K;
  G_TRY(c.alloc(&M, (size_t)B * B));
  G_TRY(c.alloc(&K, (size_t)B * B));
  tm.begin();
  launch_lap_finalize(c.st, cmass, ck, 3, B, mc, tw, 0, M, K);
  tm.end(GS_FINALIZE);
  G_HIP(hipGetLastError());
  for (void *p : {(void *)x0, (void *)x1, (void *)y0, (void *)yd}) c.release(p);
  if (mass) G_HIP(hipMemcpyAsync(mass, M, sizeof(double2) * B * B, hipMemcpyDeviceToHost, c.st));
  if (stiffness)
    G_HIP(hipMemcpyAsync(stiffness, K, sizeof(double2) * B * B, hipMemcpyDeviceToHost, c.st));
  std::vector<double> ev;
  G_TRY(generalized_spectrum(c, M, K, B, max_sweeps > 0 ? max_sweeps : 64, 1e-14,
                             deflation_tol > 0.0 ? deflation_tol : 1e-10,
                             (int)std::min<int64_t>(std::max<int64_t>(n_low, 0), B), ev, rank,
                             tm));
  const int64_t nl = std::min<int64_t>(std::max<int64_t>(n_low, 0), (int64_t)ev.size());
  if (eig_low) std::memcpy(eig_low, ev.data(), sizeof(double) * nl);
  G_HIP(hipStreamSynchronize(c.st));
  return CY3Q_OK;
}

int cy3q_yukawa_sz_weights(const double *weights, const double *omega, int64_t n, int device,
                           double *sz, double *uniformity) {
  if (!weights || !omega || n <= 0) return set_error(CY3Q_ERR_ARG, "bad input");
  Ctx c;
  G_TRY(c.open(device));
  double *d_w, *d_om, *d_sz, *d_out;
  G_TRY(c.upload(&d_w, weights, (size_t)n));
  G_TRY(c.upload(&d_om, omega, (size_t)(2 * n)));
  G_TRY(c.alloc(&d_sz, (size_t)n));
  G_TRY(c.alloc(&d_out, 1));
  launch_sz_weights(c.st, d_w, d_om, n, d_sz, d_out);
  if (sz) G_HIP(hipMemcpyAsync(sz, d_sz, sizeof(double) * n, hipMemcpyDeviceToHost, c.st));
  if (uniformity) G_HIP(hipMemcpyAsync(uniformity, d_out, sizeof(double), hipMemcpyDeviceToHost, c.st));
  G_HIP(hipStreamSynchronize(c.st));
  return CY3Q_OK;
}

int cy3q_yukawa_triple_overlap(const double *modes, const double *omega, const double *weights,
                               const double *h_v_diag, int64_t n_modes, int64_t n, int device,
                               double *y) {
  if (n_modes < 1 || n_modes > yukawa_max_modes())
    return set_error(CY3Q_ERR_ARG, "n_modes must be in 1..64");
  if (!modes || !omega || !weights || !y || n <= 0) return set_error(CY3Q_ERR_ARG, "bad input");
  Ctx c;
  G_TRY(c.open(device));
  const int nm = (int)n_modes;
  double2 *d_modes, *d_y;
  double *d_om, *d_w, *d_h = nullptr;
  G_TRY(c.upload(&d_modes, reinterpret_cast<const double2 *>(modes), (size_t)(nm * n)));
  G_TRY(c.upload(&d_om, omega, (size_t)(2 * n)));
  G_TRY(c.upload(&d_w, weights, (size_t)n));
  if (h_v_diag) G_TRY(c.upload(&d_h, h_v_diag, (size_t)nm));
  YukWs ws;
  G_TRY(ws.init(c, nm, n));
  G_TRY(c.alloc(&d_y, (size_t)nm * nm * nm));
  StageTimer tm(c.st);
  tm.begin();
  launch_yuk_prep(c.st, d_modes, n, 1, d_w, d_om, n, 1, nm, ws.pl.vldp, ws.pl.psi_v, ws.pl.g);
  launch_yuk_contract(c.st, ws.pl, n, d_h, nullptr, d_y);
  tm.end(GS_YUK_TRIPLES);
  G_HIP(hipGetLastError());
  G_HIP(hipMemcpyAsync(y, d_y, sizeof(double2) * nm * nm * nm, hipMemcpyDeviceToHost, c.st));
  G_HIP(hipStreamSynchronize(c.st));
  return CY3Q_OK;
}

int cy3q_yukawa_couplings(const double *modes, const double *omega, const double *weights,
                          const double *h_v_diag, int64_t n_modes, int64_t n,
                          int64_t n_bootstrap, uint64_t seed, int device, double *y,
                          double *y_unc, double *uniformity, double *convergence_ratio) {
  if (n_modes < 1 || n_modes > yukawa_max_modes())
    return set_error(CY3Q_ERR_ARG, "n_modes must be in 1..64");
  if (!modes || !omega || !weights || n <= 0) return set_error(CY3Q_ERR_ARG, "bad input");
  Ctx c;
  G_TRY(c.open(device));
  const int nm = (int)n_modes;
  const int64_t n3 = (int64_t)nm * nm * nm;
  double2 *d_modes, *d_y, *d_yb;
  double *d_om, *d_w, *d_sz, *d_wb, *d_h = nullptr, *d_score, *d_acc, *d_unc, *d_tot, *d_conv;
  int32_t *d_cnt;
  G_TRY(c.upload(&d_modes, reinterpret_cast<const double2 *>(modes), (size_t)(nm * n)));
  G_TRY(c.upload(&d_om, omega, (size_t)(2 * n)));
  G_TRY(c.upload(&d_w, weights, (size_t)n));
  if (h_v_diag) G_TRY(c.upload(&d_h, h_v_diag, (size_t)nm));
  G_TRY(c.alloc(&d_sz, (size_t)n));
  G_TRY(c.alloc(&d_wb, (size_t)n));
  G_TRY(c.alloc(&d_score, 1));
  G_TRY(c.alloc(&d_tot, 1));
  G_TRY(c.alloc(&d_conv, 1));
  G_TRY(c.alloc(&d_cnt, (size_t)n));
  G_TRY(c.alloc(&d_y, (size_t)n3));
  G_TRY(c.alloc(&d_yb, (size_t)n3));
  G_TRY(c.zeros(&d_acc, (size_t)n3));
  G_TRY(c.alloc(&d_unc, (size_t)n3));
  YukWs ws;
  G_TRY(ws.init(c, nm, n));
  launch_sz_weights(c.st, d_w, d_om, n, d_sz, d_score);
  launch_yuk_prep(c.st, d_modes, n, 1, d_sz, d_om, n, 1, nm, ws.pl.vldp, ws.pl.psi_v, ws.pl.g);
  launch_yuk_contract(c.st, ws.pl, n, d_h, nullptr, d_y);
  const int64_t nbs = std::max<int64_t>(n_bootstrap, 1);
  for (int64_t b = 0; b < nbs; b++) {
    G_HIP(hipMemsetAsync(d_cnt, 0, sizeof(int32_t) * n, c.st));
    launch_bootstrap_counts(c.st, seed + (uint64_t)b * 2654435761ull, n, d_cnt);
    launch_scale_weights(c.st, d_sz, d_cnt, nullptr, n, d_wb);
    launch_yuk_prep(c.st, d_modes, n, 1, d_wb, d_om, n, 0, nm, ws.pl.vldp, ws.pl.psi_v, ws.pl.g);
    launch_yuk_contract(c.st, ws.pl, n, d_h, nullptr, d_yb);
    launch_bootstrap_acc(c.st, d_yb, d_y, n3, d_acc);
  }
  launch_bootstrap_final(c.st, d_acc, n3, (double)nbs, d_unc);
  double conv = NAN;
  const int64_t half = n / 2;
  if (half >= 8) {
    launch_sum_f64(c.st, d_sz, half, d_tot);
    launch_scale_weights(c.st, d_sz, nullptr, d_tot, half, d_wb);
    launch_yuk_prep(c.st, d_modes, n, 1, d_wb, d_om, half, 1, nm, ws.pl.vldp, ws.pl.psi_v,
                    ws.pl.g);
    launch_yuk_contract(c.st, ws.pl, half, d_h, nullptr, d_yb);
    launch_conv_ratio(c.st, d_y, d_yb, n3, d_conv);
    G_HIP(hipMemcpyAsync(&conv, d_conv, sizeof(double), hipMemcpyDeviceToHost, c.st));
  }
  G_HIP(hipGetLastError());
  if (y) G_HIP(hipMemcpyAsync(y, d_y, sizeof(double2) * n3, hipMemcpyDeviceToHost, c.st));
  if (y_unc) G_HIP(hipMemcpyAsync(y_unc, d_unc, sizeof(double) * n3, hipMemcpyDeviceToHost, c.st));
  if (uniformity)
    G_HIP(hipMemcpyAsync(uniformity, d_score, sizeof(double), hipMemcpyDeviceToHost, c.st));
  G_HIP(hipStreamSynchronize(c.st));
  if (convergence_ratio) *convergence_ratio = conv;
  return CY3Q_OK;
}

int cy3q_solver_yukawa_random_modes(cy3q_solver *s, uint32_t degree, int64_t n_modes,
                                    uint64_t coeff_seed, double *y_triples, double *coeffs,
                                    int64_t *n_triples) {
  if (n_modes < 1 || n_modes > yukawa_max_modes())
    return set_error(CY3Q_ERR_ARG, "n_modes must be in 1..64");
  if (degree < 1 || degree > 12) return set_error(CY3Q_ERR_ARG, "degree must be in 1..12");
  SolverView v;
  G_TRY(solver_view(s, &v));
  Ctx c;
  c.device = v.device;
  c.st = v.st;
  const int nm = (int)n_modes;
  std::vector<uint8_t> sm8 = raw_monomials(degree);
  std::vector<uint32_t> sm(sm8.begin(), sm8.end());
  const int nsec = (int)(sm.size() / 5);
  const int64_t n = v.n;
  const int64_t m5 = padded_cols(nsec);
  const int64_t chunk = std::min<int64_t>(n, 65536);
  // psi = S C as one real GEMM per point chunk: the section rows [Re s, Im s] times the
  // real-ified coefficients give the rows [Re psi_a, Im psi_a] of the modes.
  const int64_t ncols = (2 * nm + 31) / 32 * 32;
  StageTimer tm(c.st);
  uint32_t *d_mono;
  double *sv5, *cw, *bmat, *mrows;
  double2 *coef, *d_y;
  G_TRY(c.upload(&d_mono, sm.data(), sm.size()));
  G_TRY(c.alloc(&sv5, (size_t)(chunk * m5)));
  G_TRY(c.alloc(&mrows, (size_t)(chunk * ncols)));
  G_TRY(c.alloc(&cw, (size_t)n));
  G_TRY(c.alloc(&coef, (size_t)nsec * nm));
  G_TRY(c.alloc(&bmat, (size_t)(m5 * ncols)));
  YukWs ws;
  G_TRY(ws.init(c, nm, n));
  G_TRY(c.alloc(&d_y, (size_t)ws.pl.n_trip));
  tm.begin();
  launch_random_coeffs(c.st, coeff_seed, (int64_t)nsec * nm, coef);
  launch_modes_bmat(c.st, coef, nsec, nm, m5, ncols, bmat);
  launch_measure_w_omega2(c.st, v.w, v.log_om, n, cw);
  for (int64_t p0 = 0; p0 < n; p0 += chunk) {
    const int64_t cn = std::min(chunk, n - p0);
    launch_sections(c.st, v.pts + 10 * p0, cn, d_mono, nsec, (int)degree, m5, sv5, nullptr);
    launch_gemm_nn_f64(c.st, sv5, cn, m5, bmat, ncols, ncols, mrows, ncols);
    launch_yuk_prep(c.st, reinterpret_cast<const double2 *>(mrows), 1, ncols / 2, cw + p0,
                    nullptr, cn, 0, nm, ws.pl.vldp, ws.pl.psi_v + p0 * ws.pl.vldp, ws.pl.g + p0);
  }
  tm.end(GS_YUK_MODES);
  tm.begin();
  launch_yuk_contract(c.st, ws.pl, n, nullptr, d_y, nullptr);
  tm.end(GS_YUK_TRIPLES);
  G_HIP(hipGetLastError());
  if (n_triples) *n_triples = ws.pl.n_trip;
  if (y_triples)
    G_HIP(hipMemcpyAsync(y_triples, d_y, sizeof(double2) * ws.pl.n_trip, hipMemcpyDeviceToHost, c.st));
  if (coeffs)
    G_HIP(hipMemcpyAsync(coeffs, coef, sizeof(double2) * nsec * nm, hipMemcpyDeviceToHost, c.st));
  G_HIP(hipStreamSynchronize(c.st));
  return CY3Q_OK;
}

}  // extern "C"
