// Batched multi-instance Donaldson balancing: many (psi, k, N_p) quintic instances on one GPU.
//
// The reference has no batch mode. Its instances run one after another, each with its own
// QuinticSolver; the nearest analogue is the multi-stream task orchestrator of
// cy3/src/orchestrator.rs:1-12. Here every instance keeps its own solver
// and HIP stream. Construction (sampling, weights, sections, FS-Gram) runs the instances
// concurrently from host threads, because the sampler reads its accepted counts on the host.
// The balancing iterations of all instances are then captured once into a single hipGraph:
// one fork event, each instance's reset and iteration chain on its own stream, and one join.
// Each solve is then one graph launch, and the GPU overlaps the independent instances, most of
// which are far too small to fill 256 CUs alone. The autotuner times the candidate T-pass
// point chunkings and FP64 rank-k workgroup counts per (k, N_p bucket), keeps the fastest, and
// caches the choice in a text file.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cy3q.h"
#include "cy3q_internal.h"
#include "cy3q_lcg.h"

using namespace cy3q;

struct cy3q_batch {
  int device = 0;
  int precision = CY3Q_PREC_FP64;  // of every instance (from the options)
  std::vector<cy3q_instance_spec> specs;
  std::vector<cy3q_solver *> inst;
  hipStream_t origin = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int64_t graph_iters = -1;

  void drop_graph() {
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    graph_iters = -1;
  }
  ~cy3q_batch() {
    hipSetDevice(device);
    if (origin) hipStreamSynchronize(origin);
    drop_graph();
    for (hipEvent_t e : join) hipEventDestroy(e);
    if (fork) hipEventDestroy(fork);
    for (cy3q_solver *s : inst) cy3q_solver_free(s);
    if (origin) hipStreamDestroy(origin);
  }
};

namespace {

#define B_HIP(expr)                                                                          \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return set_error(CY3Q_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
  } while (0)

#define B_TRY(expr)                  \
  do {                               \
    int rc_ = (expr);                \
    if (rc_ != CY3Q_OK) return rc_;  \
  } while (0)


int solve_timed(cy3q_solver *s, int64_t iters, double *ms) {
  B_TRY(cy3q_solver_synchronize(s));
  const auto t0 = std::chrono::steady_clock::now();
  B_TRY(cy3q_solver_donaldson_solve(s, iters, 0.0, nullptr, nullptr));
  const auto t1 = std::chrono::steady_clock::now();
  *ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  return CY3Q_OK;
}

std::string tune_key(const cy3q_instance_spec &sp, int quotient) {
  const int bucket = (int)std::lround(std::log2((double)std::max<int64_t>(sp.n_points, 1)));
  return "k" + std::to_string(sp.k) + "_q" + std::to_string(quotient) + "_n2e" +
         std::to_string(bucket);
}

}  // namespace

extern "C" {

int cy3q_batch_draw_specs(uint64_t seed, int n, cy3q_instance_spec *out) {
  if (n < 0 || (n > 0 && !out)) return set_error(CY3Q_ERR_ARG, "bad arguments");
  Lcg r{lcg_seed(seed)};
  const double l0 = std::log(2.0e4), l1 = std::log(5.0e5);
  for (int i = 0; i < n; i++) {
    const double u_psi = r.next_f64(), u_k = r.next_f64(), u_n = r.next_f64();
    out[i].psi = 0.9 * u_psi;
    out[i].k = 2u + (uint32_t)std::min(4.0, std::floor(5.0 * u_k));
    out[i].n_points = (int64_t)std::llround(std::exp(l0 + u_n * (l1 - l0)));
    out[i].seed = 42ull + (uint64_t)i;
  }
  return CY3Q_OK;
}

int cy3q_batch_new(const cy3q_instance_spec *specs, int n, const cy3q_options *opt,
                   cy3q_batch **out) {
  if (!out || !specs || n <= 0 || n > 1024) return set_error(CY3Q_ERR_ARG, "bad arguments");
  *out = nullptr;
  cy3q_options o;
  cy3q_options_default(&o);
  if (opt) o = *opt;
  for (int i = 0; i < n; i++)
    if (specs[i].k < 1 || specs[i].k > 12 || specs[i].n_points <= 0 || !(specs[i].psi >= 0.0))
      return set_error(CY3Q_ERR_ARG, "instance " + std::to_string(i) + ": bad spec");
  cy3q_batch *b = new cy3q_batch();
  b->device = o.device;
  b->precision = o.precision;
  b->specs.assign(specs, specs + n);
  b->inst.assign(n, nullptr);
  std::vector<int> rc(n, CY3Q_OK);
  std::vector<std::string> msg(n);
  {
    std::vector<std::thread> th;
    const int n_threads = std::min(n, 16);
    for (int t = 0; t < n_threads; t++)
      th.emplace_back([&, t] {
        hipSetDevice(o.device);
        for (int i = t; i < n; i += n_threads) {
          const cy3q_instance_spec &sp = specs[i];
          const int sampler = sp.psi == 0.0 ? CY3Q_SAMPLER_SZ : CY3Q_SAMPLER_SZ_DWORK;
          cy3q_solver *s = nullptr;
          int r = cy3q_solver_new(sp.k, sp.n_points, sp.seed, sp.psi, sampler, &o, &s);
          if (r == CY3Q_OK) r = cy3q_solver_orthonormalise_fs_gram(s);
          if (r == CY3Q_OK) r = cy3q_solver_synchronize(s);
          rc[i] = r;
          b->inst[i] = s;
          if (r != CY3Q_OK) msg[i] = cy3q_last_error();
        }
      });
    for (auto &t : th) t.join();
  }
  for (int i = 0; i < n; i++)
    if (rc[i] != CY3Q_OK) {
      const int r = rc[i];
      const std::string m = "instance " + std::to_string(i) + ": " + msg[i];
      delete b;
      return set_error(r, m);
    }
  B_HIP(hipSetDevice(o.device));
  B_HIP(hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking));
  B_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
  b->join.assign(n, nullptr);
  for (int i = 0; i < n; i++) B_HIP(hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming));
  *out = b;
  return CY3Q_OK;
}

void cy3q_batch_free(cy3q_batch *b) { delete b; }

int cy3q_batch_size(cy3q_batch *b, int *n) {
  if (!b || !n) return set_error(CY3Q_ERR_ARG, "bad arguments");
  *n = (int)b->inst.size();
  return CY3Q_OK;
}

int cy3q_batch_solver(cy3q_batch *b, int i, cy3q_solver **s) {
  if (!b || !s || i < 0 || i >= (int)b->inst.size()) return set_error(CY3Q_ERR_ARG, "bad index");
  *s = b->inst[i];
  return CY3Q_OK;
}

int cy3q_batch_solve_sequential(cy3q_batch *b, int64_t iters) {
  if (!b || iters < 0) return set_error(CY3Q_ERR_ARG, "bad arguments");
  B_HIP(hipSetDevice(b->device));
  for (cy3q_solver *s : b->inst) {
    B_TRY(solver_reset_identity(s));
    B_TRY(cy3q_solver_donaldson_solve(s, iters, 0.0, nullptr, nullptr));
  }
  return CY3Q_OK;
}

int cy3q_batch_solve_graph(cy3q_batch *b, int64_t iters) {
  if (!b || iters < 0) return set_error(CY3Q_ERR_ARG, "bad arguments");
  B_HIP(hipSetDevice(b->device));
  if (b->graph_iters != iters || !b->exec) {
    b->drop_graph();
    for (cy3q_solver *s : b->inst) B_TRY(solver_prepare_enqueue(s, iters));
    B_HIP(hipStreamBeginCapture(b->origin, hipStreamCaptureModeRelaxed));
    int rc = CY3Q_OK;
    hipError_t e = hipEventRecord(b->fork, b->origin);
    for (size_t i = 0; i < b->inst.size() && rc == CY3Q_OK && e == hipSuccess; i++) {
      hipStream_t st = solver_stream(b->inst[i]);
      e = hipStreamWaitEvent(st, b->fork, 0);
      if (e != hipSuccess) break;
      rc = solver_reset_identity(b->inst[i]);
      if (rc == CY3Q_OK) rc = solver_enqueue_solve(b->inst[i], iters);
      if (rc == CY3Q_OK) e = hipEventRecord(b->join[i], st);
      if (rc == CY3Q_OK && e == hipSuccess) e = hipStreamWaitEvent(b->origin, b->join[i], 0);
    }
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(b->origin, &g);
    if (rc != CY3Q_OK) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess || e2 != hipSuccess) {
      if (g) hipGraphDestroy(g);
      return set_error(CY3Q_ERR_HIP, std::string("graph capture: ") +
                                         hipGetErrorString(e != hipSuccess ? e : e2));
    }
    b->graph = g;
    B_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->graph_iters = iters;
  }
  B_HIP(hipGraphLaunch(b->exec, b->origin));
  B_HIP(hipStreamSynchronize(b->origin));
  return CY3Q_OK;
}

int cy3q_batch_autotune(cy3q_batch *b, const char *cache_path, int64_t *n_measured) {
  if (!b) return set_error(CY3Q_ERR_ARG, "NULL batch");
  B_HIP(hipSetDevice(b->device));
  std::map<std::string, std::pair<int64_t, int64_t>> table;
  if (cache_path) {
    if (FILE *f = std::fopen(cache_path, "r")) {
      char key[128];
      long long tp, sy;
      while (std::fscanf(f, "%127s %lld %lld", key, &tp, &sy) == 3) table[key] = {tp, sy};
      std::fclose(f);
    }
  }
  static const int64_t kTp[] = {1, 2, 4, 8};
  static const int64_t kSy[] = {256, 512, 1024, 2048};
  int64_t measured = 0;
  for (size_t i = 0; i < b->inst.size(); i++) {
    cy3q_solver *s = b->inst[i];
    int64_t n_pts = 0, nb = 0;
    uint32_t k = 0;
    double psi = 0;
    B_TRY(cy3q_solver_info(s, &n_pts, &nb, &k, &psi));
    const int quotient = (int)(nb != (int64_t)((k + 1) * (k + 2) * (k + 3) * (k + 4) / 24));
    const std::string key = tune_key(b->specs[i], quotient);
    auto it = table.find(key);
    if (it == table.end()) {
      int64_t prec_tp = 1, prec_sy = 1024;
      double best = 1e300;
      // T-pass chunking, timed on split-precision iterations with the refinement sweep off.
      for (int64_t tp : kTp) {
        if (b->precision != CY3Q_PREC_SPLIT16) break;
        B_TRY(cy3q_solver_set_tuning(s, tp, 0));
        B_TRY(cy3q_solver_set_precision(s, CY3Q_PREC_SPLIT16));
        B_TRY(cy3q_solver_set_refine(s, 0));
        double ms = 0.0, m2 = 0.0;
        B_TRY(solve_timed(s, 1, &ms));  // warm-up
        B_TRY(solve_timed(s, 3, &ms));
        B_TRY(solve_timed(s, 3, &m2));
        ms = std::min(ms, m2);
        if (ms < best) {
          best = ms;
          prec_tp = tp;
        }
      }
      best = 1e300;
      for (int64_t sy : kSy) {
        B_TRY(cy3q_solver_set_tuning(s, 0, sy));
        B_TRY(cy3q_solver_set_precision(s, CY3Q_PREC_FP64));
        double ms = 0.0, m2 = 0.0;
        B_TRY(solve_timed(s, 1, &ms));
        B_TRY(solve_timed(s, 2, &ms));
        B_TRY(solve_timed(s, 2, &m2));
        ms = std::min(ms, m2);
        if (ms < best) {
          best = ms;
          prec_sy = sy;
        }
      }
      table[key] = {prec_tp, prec_sy};
      it = table.find(key);
      measured++;
    }
    B_TRY(cy3q_solver_set_tuning(s, it->second.first, it->second.second));
    B_TRY(cy3q_solver_set_precision(s, b->precision));
    B_TRY(cy3q_solver_set_refine(s, 1));
    B_TRY(solver_reset_identity(s));
    B_TRY(cy3q_solver_synchronize(s));
  }
  b->drop_graph();  // buffers may have moved
  if (cache_path) {
    if (FILE *f = std::fopen(cache_path, "w")) {
      for (auto &kv : table)
        std::fprintf(f, "%s %lld %lld\n", kv.first.c_str(), (long long)kv.second.first,
                     (long long)kv.second.second);
      std::fclose(f);
    }
  }
  if (n_measured) *n_measured = measured;
  return CY3Q_OK;
}

}  // extern "C"
