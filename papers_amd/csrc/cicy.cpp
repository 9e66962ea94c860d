// Host side of the Tian-Yau and Schoen set-up entry points (include/cy3q.h, cy3q_cicy_*):
// the random stream, the Groebner-reduced section bases and the sequencing of the sampler,
// frame and section kernels of k_cicy.hip.
//
// The random stream is the reference's: rand_chacha 0.9 ChaCha8Rng seeded through rand_core
// 0.9's seed_from_u64 (eight PCG32 outputs as the key), read as u64 = lo | hi << 32, and
// rand_distr 0.5's StandardNormal ziggurat on it (cy3/src/cicy_sampler.rs:136-138,
// route34/schoen_sampler.rs:980-998). The GPU evaluates the ziggurat's first test at every
// position of the keystream; the host follows the stream through the positions where it fails
// (about 1 in 80, with the wedge and tail in its own exp and log) and the GPU gathers the
// normals (DeviceNormals). The GPU does every line's Newton solves.
//
// The section bases keep the (k, k) / (d_x, d_y, d_t) monomials that are invariant under the
// quotient group and lie outside the ideal of leading monomials of the reduced Groebner basis
// of the defining ideal under degree-lex (route34/ty_metric.rs:839-895,
// route34/schoen_metric.rs:1172-1212, route34/groebner.rs). That ideal of leading monomials
// depends only on the ideal and the order; it is computed here by Buchberger's algorithm over
// exact rationals.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <set>
#include <string>
#include <vector>

#include "../../include/cy3q.h"
#include "cy3q_internal.h"

using namespace cy3q;

namespace {

#define C_HIP(expr)                                                                          \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return set_error(CY3Q_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
  } while (0)

#define C_TRY(expr)                  \
  do {                               \
    int rc_ = (expr);                \
    if (rc_ != CY3Q_OK) return rc_;  \
  } while (0)

// ---- Random stream ----------------------------------------------------------------------

struct ZigTables {
  double x[257], f[257];
};

// The 256-layer normal ziggurat of Marsaglia & Tsang with R = 3.654152885361008796 and
// V = 0.00492867323399: x0 = V / f(R), x1 = R, x_i = f^-1(V / x_{i-1} + f(x_{i-1})), x256 = 0,
// f = exp(-x^2 / 2). Each entry is rounded to 18 decimals, as the table rand_distr compiles in.
const ZigTables &zig_tables() {
  static ZigTables t;
  static std::once_flag once;
  std::call_once(once, [] {
    auto r18 = [](double v) {
      char b[64];
      std::snprintf(b, sizeof b, "%.18f", v);
      return std::strtod(b, nullptr);
    };
    const double r = 3.654152885361008796, v = 0.00492867323399;
    double x[257];
    x[0] = v / std::exp(-r * r / 2.0);
    x[1] = r;
    for (int i = 2; i < 256; i++) {
      const double last = x[i - 1];
      x[i] = std::sqrt(-2.0 * std::log(v / last + std::exp(-last * last / 2.0)));
    }
    x[256] = 0.0;
    for (int i = 0; i < 257; i++) {
      t.x[i] = r18(x[i]);
      t.f[i] = r18(std::exp(-x[i] * x[i] / 2.0));
    }
  });
  return t;
}

class NormalStream {
 public:
  explicit NormalStream(uint64_t seed) {
    const uint64_t mul = 6364136223846793005ull, inc = 11634580027462260723ull;
    for (int i = 0; i < 8; i++) {  // SeedableRng::seed_from_u64: PCG32 words
      seed = seed * mul + inc;
      const uint32_t xs = (uint32_t)(((seed >> 18) ^ seed) >> 27);
      const uint32_t rot = (uint32_t)(seed >> 59);
      key_[i] = (xs >> rot) | (xs << ((32 - rot) & 31));
    }
  }
  const uint32_t *key() const { return key_; }

  // The normal that starts at u64 position e of the stream, given that its first test fails
  // there (|u x_i| >= x_(i+1), u = 2 (bits >> 12) / 2^52 - 1, i = bits & 0xff): for i = 0 the
  // tail beyond R (pairs of uniforms in (0, 1] until -2 log b >= (log a / R)^2), else the wedge
  // test f_(i+1) + (f_i - f_(i+1)) v < exp(-x^2 / 2) with v from the next word. *words: the u64s
  // it takes; a rejected wedge (*accepted false) takes 2, and the stream's next attempt starts
  // after them.
  // bits, next: the words at e and e + 1 (only a tail reads further words).
  void slow_at(uint64_t e, uint64_t bits, uint64_t next, bool *accepted, double *value,
               int64_t *words) {
    const ZigTables &z = zig_tables();
    const int i = (int)(bits & 0xff);
    const double u = with_exponent(bits >> 12, 1) - 3.0;
    const double x = u * z.x[i];
    if (i == 0) {
      const double R = 3.654152885361008796;
      double xt = 1.0, yt = 0.0;
      uint64_t q = e + 1;
      while (-2.0 * yt < xt * xt) {
        const double a = with_exponent((q == e + 1 ? next : u64_at(q)) >> 12, 0) - (1.0 - 0x1p-53);
        const double b = with_exponent(u64_at(q + 1) >> 12, 0) - (1.0 - 0x1p-53);
        q += 2;
        xt = std::log(a) / R;
        yt = std::log(b);
      }
      *accepted = true;
      *value = u < 0.0 ? xt - R : R - xt;
      *words = (int64_t)(q - e);
      return;
    }
    const double un = (double)(next >> 11) * 0x1p-53;
    *accepted = z.f[i + 1] + (z.f[i] - z.f[i + 1]) * un < std::exp(-x * x / 2.0);
    *value = x;
    *words = 2;
  }

 private:
  static double with_exponent(uint64_t frac52, int e) {
    const uint64_t b = frac52 | ((uint64_t)(1023 + e) << 52);
    double d;
    std::memcpy(&d, &b, 8);
    return d;
  }
  static uint32_t rotl(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

  // u64 number q of the stream (block q / 8), with the last block kept.
  uint64_t u64_at(uint64_t q) {
    const uint64_t blk = q / 8;
    if (blk != rblk_) {
      uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
      for (int i = 0; i < 8; i++) s[4 + i] = key_[i];
      s[12] = (uint32_t)blk;
      s[13] = (uint32_t)(blk >> 32);
      s[14] = s[15] = 0;
      uint32_t x[16];
      std::memcpy(x, s, sizeof x);
      auto qr = [&](int a, int bb, int cc, int d) {
        x[a] += x[bb]; x[d] ^= x[a]; x[d] = rotl(x[d], 16);
        x[cc] += x[d]; x[bb] ^= x[cc]; x[bb] = rotl(x[bb], 12);
        x[a] += x[bb]; x[d] ^= x[a]; x[d] = rotl(x[d], 8);
        x[cc] += x[d]; x[bb] ^= x[cc]; x[bb] = rotl(x[bb], 7);
      };
      for (int r = 0; r < 4; r++) {
        qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15);
        qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14);
      }
      for (int i = 0; i < 16; i++) rbuf_[i] = x[i] + s[i];
      rblk_ = blk;
    }
    const int j = (int)(q % 8);
    return (uint64_t)rbuf_[2 * j] | ((uint64_t)rbuf_[2 * j + 1] << 32);
  }

  uint32_t key_[8];
  uint64_t rblk_ = ~0ull;
  uint32_t rbuf_[16];
};


// ---- Groebner bases over Q ------------------------------------------------------------------

using Exp = std::array<uint8_t, 8>;

struct Rat {
  int64_t n = 0, d = 1;
};

bool rat_make(__int128 n, __int128 d, Rat *out) {
  if (d < 0) { n = -n; d = -d; }
  __int128 a = n < 0 ? -n : n, b = d;
  while (b) { __int128 t = a % b; a = b; b = t; }
  if (a > 1) { n /= a; d /= a; }
  const __int128 lim = (__int128)1 << 62;
  if (n > lim || n < -lim || d > lim) return false;
  out->n = (int64_t)n;
  out->d = (int64_t)d;
  return true;
}

int deg(const Exp &e) { int s = 0; for (uint8_t v : e) s += v; return s; }

// Degree-lex with x0 > x1 > ... > x7 (groebner.rs:125-157): true when a > b.
bool deglex_gt(const Exp &a, const Exp &b) {
  const int da = deg(a), db = deg(b);
  if (da != db) return da > db;
  for (int i = 0; i < 8; i++)
    if (a[i] != b[i]) return a[i] > b[i];
  return false;
}

struct ExpGt {
  bool operator()(const Exp &a, const Exp &b) const { return deglex_gt(a, b); }
};

using Poly = std::map<Exp, Rat, ExpGt>;  // first entry = leading term

bool divides(const Exp &a, const Exp &b) {
  for (int i = 0; i < 8; i++)
    if (a[i] > b[i]) return false;
  return true;
}

// p -= c x^shift q
bool sub_scaled(Poly &p, const Poly &q, Rat c, const Exp &shift) {
  for (const auto &t : q) {
    Exp e;
    for (int i = 0; i < 8; i++) e[i] = (uint8_t)(t.first[i] + shift[i]);
    Rat prod;
    if (!rat_make((__int128)c.n * t.second.n, (__int128)c.d * t.second.d, &prod)) return false;
    auto it = p.find(e);
    if (it == p.end()) {
      p[e] = Rat{-prod.n, prod.d};
    } else {
      Rat r;
      if (!rat_make((__int128)it->second.n * prod.d - (__int128)prod.n * it->second.d,
                    (__int128)it->second.d * prod.d, &r))
        return false;
      if (r.n == 0) p.erase(it); else it->second = r;
    }
  }
  return true;
}

// Full normal form of f modulo g.
bool normal_form(Poly f, const std::vector<Poly> &g, Poly *out) {
  Poly r;
  while (!f.empty()) {
    const auto lt = *f.begin();
    bool reduced = false;
    for (const Poly &q : g) {
      const auto &lq = *q.begin();
      if (!divides(lq.first, lt.first)) continue;
      Exp shift;
      for (int i = 0; i < 8; i++) shift[i] = (uint8_t)(lt.first[i] - lq.first[i]);
      Rat c;
      if (!rat_make((__int128)lt.second.n * lq.second.d, (__int128)lt.second.d * lq.second.n, &c))
        return false;
      if (!sub_scaled(f, q, c, shift)) return false;
      reduced = true;
      break;
    }
    if (!reduced) {
      r[lt.first] = lt.second;
      f.erase(f.begin());
    }
  }
  *out = std::move(r);
  return true;
}

// Leading monomials of a minimal Groebner basis of <gens> (equal to those of the reduced
// basis). Normal selection strategy; pairs with coprime leading monomials are skipped.
bool groebner_lms(const std::vector<Poly> &gens, std::vector<Exp> *lms_out) {
  std::vector<Poly> g = gens;
  std::set<std::pair<int, int>> pairs;
  for (int i = 0; i < (int)g.size(); i++)
    for (int j = i + 1; j < (int)g.size(); j++) pairs.insert({i, j});
  auto lcm_deg = [&](const std::pair<int, int> &q) {
    const Exp &a = g[q.first].begin()->first, &b = g[q.second].begin()->first;
    int s = 0;
    for (int i = 0; i < 8; i++) s += std::max(a[i], b[i]);
    return s;
  };
  int guard = 0;
  while (!pairs.empty()) {
    if (++guard > 100000) return false;
    auto best = pairs.begin();
    int bd = lcm_deg(*best);
    for (auto it = std::next(pairs.begin()); it != pairs.end(); ++it) {
      const int d = lcm_deg(*it);
      if (d < bd) { bd = d; best = it; }
    }
    const auto q = *best;
    pairs.erase(best);
    const auto &fi = *g[q.first].begin(), &fj = *g[q.second].begin();
    bool coprime = true;
    for (int k = 0; k < 8; k++)
      if (fi.first[k] && fj.first[k]) coprime = false;
    if (coprime) continue;
    Exp l, si, sj;
    for (int k = 0; k < 8; k++) {
      l[k] = std::max(fi.first[k], fj.first[k]);
      si[k] = (uint8_t)(l[k] - fi.first[k]);
      sj[k] = (uint8_t)(l[k] - fj.first[k]);
    }
    // S = x^si f_i / lc_i - x^sj f_j / lc_j
    Poly s;
    if (!sub_scaled(s, g[q.first], Rat{-fi.second.d, fi.second.n}, si)) return false;
    Rat cj;
    if (!rat_make(fj.second.d, fj.second.n, &cj)) return false;
    if (!sub_scaled(s, g[q.second], cj, sj)) return false;
    Poly r;
    if (!normal_form(s, g, &r)) return false;
    if (r.empty()) continue;
    g.push_back(std::move(r));
    for (int k = 0; k + 1 < (int)g.size(); k++) pairs.insert({k, (int)g.size() - 1});
  }
  std::vector<Exp> lms;
  for (const Poly &p : g) lms.push_back(p.begin()->first);
  std::vector<Exp> keep;
  for (size_t a = 0; a < lms.size(); a++) {
    bool redundant = false;
    for (size_t b = 0; b < lms.size() && !redundant; b++)
      if (b != a && divides(lms[b], lms[a]) && (lms[b] != lms[a] || b < a)) redundant = true;
    if (!redundant) keep.push_back(lms[a]);
  }
  *lms_out = keep;
  return true;
}

Poly poly_of(const std::vector<std::pair<int, Exp>> &terms) {
  Poly p;
  for (const auto &t : terms) p[t.second] = Rat{t.first, 1};
  return p;
}

// ty_generators / schoen_generators (route34/groebner.rs:665-744)
std::vector<Poly> generators(int variety) {
  std::vector<Poly> out;
  if (variety == CY3Q_CICY_TY) {
    std::vector<std::pair<int, Exp>> f1, f2, f3;
    for (int k = 0; k < 4; k++) {
      Exp a{}, b{}, c{};
      a[k] = 3;
      b[4 + k] = 3;
      c[k] = 1;
      c[4 + k] = 1;
      f1.push_back({1, a});
      f2.push_back({1, b});
      f3.push_back({1, c});
    }
    out = {poly_of(f1), poly_of(f2), poly_of(f3)};
  } else {
    std::vector<std::pair<int, Exp>> f1, f2;
    for (int k = 0; k < 3; k++) {
      Exp a{}, a2{}, b{}, b2{};
      a[k] = 3; a[6] = 1;
      a2[k] = 3; a2[7] = 1;
      b[3 + k] = 3; b[6] = 1;
      b2[3 + k] = 3; b2[7] = 1;
      f1.push_back({1, a});
      f1.push_back({k == 0 ? 1 : -1, a2});
      f2.push_back({1, b});
      f2.push_back({k == 0 ? 1 : -1, b2});
    }
    out = {poly_of(f1), poly_of(f2)};
  }
  return out;
}

int cached_lms(int variety, const std::vector<Exp> **out) {
  static std::vector<Exp> lms[2];
  static bool ready[2] = {false, false};
  static std::mutex mu;
  std::lock_guard<std::mutex> lock(mu);
  if (!ready[variety]) {
    if (!groebner_lms(generators(variety), &lms[variety]))
      return set_error(CY3Q_ERR_STATE, "Groebner basis computation did not finish");
    ready[variety] = true;
  }
  *out = &lms[variety];
  return CY3Q_OK;
}

// ---- Device buffers of one call ---------------------------------------------------------------

// Device blocks of the set-up calls, cached per device across calls: a sampler runs once per
// solve, and allocating its buffers afresh after a solve has freed its own costs a few
// milliseconds per call. Blocks come in power-of-two size classes (64 KiB up); at most 1 GiB
// per device is kept, until the process ends.
constexpr int kCacheDevices = 64;
constexpr size_t kCacheLimit = (size_t)1 << 30;
struct BlockCache {
  std::mutex mu;
  std::multimap<size_t, void *> free_blocks[kCacheDevices];
  size_t cached[kCacheDevices] = {};
};
BlockCache &block_cache() {
  static BlockCache c;
  return c;
}
size_t size_class(size_t bytes) {
  size_t c = (size_t)1 << 16;
  while (c < bytes) c <<= 1;
  return c;
}

struct Call {
  hipStream_t st = nullptr;
  int device = 0;
  std::vector<std::pair<void *, size_t>> mem;  // live blocks and their size classes
  ~Call() {
    if (st) hipStreamSynchronize(st);
    for (auto &b : mem) give_back(b.first, b.second);
    if (st) hipStreamDestroy(st);
  }
  int open(int dev) {
    if (dev < 0 || dev >= kCacheDevices) return set_error(CY3Q_ERR_ARG, "bad device index");
    device = dev;
    C_HIP(hipSetDevice(dev));
    C_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    return CY3Q_OK;
  }
  template <class T>
  int alloc(T **p, size_t count) {
    const size_t cls = size_class(std::max<size_t>(count, 1) * sizeof(T));
    void *q = nullptr;
    {
      BlockCache &bc = block_cache();
      std::lock_guard<std::mutex> lock(bc.mu);
      auto it = bc.free_blocks[device].find(cls);
      if (it != bc.free_blocks[device].end()) {
        q = it->second;
        bc.free_blocks[device].erase(it);
        bc.cached[device] -= cls;
      }
    }
    if (!q) {
      hipError_t e = hipMalloc(&q, cls);
      if (e != hipSuccess)
        return set_error(CY3Q_ERR_HIP,
                         "hipMalloc(" + std::to_string(cls) + "): " + hipGetErrorString(e));
    }
    mem.emplace_back(q, cls);
    *p = (T *)q;
    return CY3Q_OK;
  }
  void release(void *p) {
    auto it = std::find_if(mem.begin(), mem.end(),
                           [p](const std::pair<void *, size_t> &b) { return b.first == p; });
    if (it != mem.end()) {
      hipStreamSynchronize(st);
      give_back(it->first, it->second);
      mem.erase(it);
    }
  }

 private:
  void give_back(void *p, size_t cls) {
    BlockCache &bc = block_cache();
    std::lock_guard<std::mutex> lock(bc.mu);
    if (bc.cached[device] + cls <= kCacheLimit) {
      bc.free_blocks[device].emplace(cls, p);
      bc.cached[device] += cls;
    } else {
      hipFree(p);
    }
  }
};

constexpr int kLineNormals[2] = {64, 48};  // origin + one direction per hypersurface, 16 each
constexpr int kRestartNormals[2] = {6, 4};  // one complex t per hypersurface

struct Accepted {
  std::vector<double> coords, om, w;
  std::vector<int32_t> patch;
  int64_t n = 0;
};

// The normal stream drawn on the device, in order: generate(count, d_out) writes the stream's
// next `count` normals. Per pass over a range of u64 positions, k_zig_fast evaluates the
// ziggurat's first test at every position (exact: a subtraction and a product of doubles) and
// lists the positions where it fails (about 1.2%); the host follows the stream from the
// current position: a position whose first test passes is a normal of one word, so the starts
// run on until the next listed position, which the host evaluates itself (wedge or tail,
// NormalStream::slow_at, with its exp and log); k_zig_gather then writes the normals.
class DeviceNormals {
 public:
  explicit DeviceNormals(uint64_t seed) : host_(seed) {
    std::memcpy(key_.k, host_.key(), sizeof key_.k);
  }
  int64_t generated() const { return generated_; }

  int generate(Call &c, int64_t count, double *d_out) {
    if (!tables_) {
      upload_zig_x(c.st, zig_tables().x);
      tables_ = true;
    }
    int64_t produced = 0;
    while (produced < count) {
      const int64_t need = count - produced;
      const int64_t n_pos = need + need / 16 + 256;
      // The failed positions are about 1.2% of a pass; a pass whose list overflows its
      // capacity runs again with room for every position.
      int64_t exc_cap = n_pos / 16 + 1024;
      unsigned long long n_exc = 0;
      for (;;) {
        C_TRY(grow(c, n_pos, exc_cap));
        C_HIP(hipMemsetAsync(d_nexc_, 0, sizeof(unsigned long long), c.st));
        launch_zig_fast(c.st, key_, pos_, n_pos, d_fast_, d_words_, d_exc_, exc_cap_, d_nexc_);
        C_HIP(hipGetLastError());
        C_HIP(hipMemcpyAsync(&n_exc, d_nexc_, sizeof n_exc, hipMemcpyDeviceToHost, c.st));
        C_HIP(hipStreamSynchronize(c.st));
        if ((int64_t)n_exc <= exc_cap_) break;
        exc_cap = n_pos;
      }
      // The listed positions with their two words, in stream order.
      exc_.resize((size_t)n_exc);
      exw_.resize(2 * (size_t)n_exc);
      if (n_exc) {
        launch_zig_exc_words(c.st, d_words_, d_exc_, (int64_t)n_exc, pos_, d_exw_);
        C_HIP(hipGetLastError());
        C_HIP(hipMemcpyAsync(exc_.data(), d_exc_, sizeof(int64_t) * n_exc, hipMemcpyDeviceToHost,
                             c.st));
        C_HIP(hipMemcpyAsync(exw_.data(), d_exw_, sizeof(uint64_t) * 2 * n_exc,
                             hipMemcpyDeviceToHost, c.st));
      }
      C_HIP(hipStreamSynchronize(c.st));
      order_.resize((size_t)n_exc);
      std::iota(order_.begin(), order_.end(), (size_t)0);
      std::sort(order_.begin(), order_.end(), [&](size_t a, size_t b) { return exc_[a] < exc_[b]; });
      // Follow the stream through [pos_, pos_ + n_pos).
      run_k_.clear();
      run_p_.clear();
      run_len_.clear();
      ek_.clear();
      ev_.clear();
      const int64_t base = produced;
      const uint64_t end = pos_ + (uint64_t)n_pos;
      uint64_t p = pos_;
      size_t ei = 0;
      while (produced < count && p < end) {
        while (ei < order_.size() && (uint64_t)exc_[order_[ei]] < p) ei++;
        const uint64_t e = ei < order_.size() ? (uint64_t)exc_[order_[ei]] : end;
        const int64_t r = std::min<int64_t>((int64_t)(e - p), count - produced);
        if (r > 0) {
          run_k_.push_back(produced - base);
          run_p_.push_back((int64_t)(p - pos_));
          run_len_.push_back(r);
          produced += r;
          p += (uint64_t)r;
        }
        if (produced == count || p >= end) break;
        bool accepted;
        double v;
        int64_t words;
        const size_t oe = order_[ei];
        host_.slow_at(e, exw_[2 * oe], exw_[2 * oe + 1], &accepted, &v, &words);
        if (accepted) {
          ek_.push_back(produced - base);
          ev_.push_back(v);
          produced++;
        }
        p = e + (uint64_t)words;
      }
      // Runs of one pass number at most its listed positions + 1.
      const int64_t nr = (int64_t)run_k_.size(), ne = (int64_t)ek_.size();
      int64_t *d_rk, *d_rp, *d_rl, *d_ek;
      double *d_ev;
      C_TRY(c.alloc(&d_rk, (size_t)nr));
      C_TRY(c.alloc(&d_rp, (size_t)nr));
      C_TRY(c.alloc(&d_rl, (size_t)nr));
      C_TRY(c.alloc(&d_ek, (size_t)ne));
      C_TRY(c.alloc(&d_ev, (size_t)ne));
      if (nr) {
        C_HIP(hipMemcpyAsync(d_rk, run_k_.data(), sizeof(int64_t) * nr, hipMemcpyHostToDevice, c.st));
        C_HIP(hipMemcpyAsync(d_rp, run_p_.data(), sizeof(int64_t) * nr, hipMemcpyHostToDevice, c.st));
        C_HIP(hipMemcpyAsync(d_rl, run_len_.data(), sizeof(int64_t) * nr, hipMemcpyHostToDevice, c.st));
      }
      if (ne) {
        C_HIP(hipMemcpyAsync(d_ek, ek_.data(), sizeof(int64_t) * ne, hipMemcpyHostToDevice, c.st));
        C_HIP(hipMemcpyAsync(d_ev, ev_.data(), sizeof(double) * ne, hipMemcpyHostToDevice, c.st));
      }
      launch_zig_gather(c.st, d_fast_, d_rk, d_rp, d_rl, nr, d_ek, d_ev, ne, produced - base,
                        d_out + base);
      C_HIP(hipGetLastError());
      for (void *q : {(void *)d_rk, (void *)d_rp, (void *)d_rl, (void *)d_ek, (void *)d_ev})
        c.release(q);  // synchronises the stream, so the host vectors may change after this
      pos_ = p;
    }
    generated_ += count;
    return CY3Q_OK;
  }

 private:
  int grow(Call &c, int64_t n_pos, int64_t exc_cap) {
    if (n_pos > cap_) {
      if (d_fast_) c.release(d_fast_);
      if (d_words_) c.release(d_words_);
      C_TRY(c.alloc(&d_fast_, (size_t)n_pos));
      C_TRY(c.alloc(&d_words_, (size_t)n_pos + 1));
      cap_ = n_pos;
    }
    if (exc_cap > exc_cap_) {
      if (d_exc_) c.release(d_exc_);
      if (d_exw_) c.release(d_exw_);
      C_TRY(c.alloc(&d_exc_, (size_t)exc_cap));
      C_TRY(c.alloc(&d_exw_, 2 * (size_t)exc_cap));
      exc_cap_ = exc_cap;
    }
    if (!d_nexc_) C_TRY(c.alloc(&d_nexc_, 1));
    return CY3Q_OK;
  }

  NormalStream host_;
  ZigKey key_;
  uint64_t pos_ = 0;        // u64 position where the next normal starts
  int64_t generated_ = 0;   // normals written so far
  bool tables_ = false;
  int64_t cap_ = 0, exc_cap_ = 0;
  double *d_fast_ = nullptr;
  uint64_t *d_words_ = nullptr, *d_exw_ = nullptr;
  int64_t *d_exc_ = nullptr;
  unsigned long long *d_nexc_ = nullptr;
  std::vector<int64_t> exc_, run_k_, run_p_, run_len_, ek_;
  std::vector<uint64_t> exw_;
  std::vector<size_t> order_;
  std::vector<double> ev_;
};


// Post-process `n` solved lines already on the device (x, flag) and append the accepted ones
// (in line order) until `n_target`. Returns the number of lines consumed (up to the line that
// produced the last point kept, or all n).
int post_and_collect(Call &c, int variety, const double *d_x, const int32_t *d_flag, int64_t n,
                     int64_t n_target, Accepted &acc, int64_t *lines_used) {
  double *d_coords, *d_om, *d_w;
  int32_t *d_patch, *d_ok;
  C_TRY(c.alloc(&d_coords, (size_t)n * 16));
  C_TRY(c.alloc(&d_om, (size_t)n));
  C_TRY(c.alloc(&d_w, (size_t)n));
  C_TRY(c.alloc(&d_patch, (size_t)n * 3));
  C_TRY(c.alloc(&d_ok, (size_t)n));
  launch_cicy_post(c.st, variety, d_x, d_flag, n, d_coords, d_om, d_w, d_patch, d_ok);
  C_HIP(hipGetLastError());
  std::vector<int32_t> ok((size_t)n), patch((size_t)n * 3);
  std::vector<double> coords((size_t)n * 16), om((size_t)n), w((size_t)n);
  C_HIP(hipMemcpyAsync(ok.data(), d_ok, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c.st));
  C_HIP(hipMemcpyAsync(coords.data(), d_coords, sizeof(double) * n * 16, hipMemcpyDeviceToHost, c.st));
  C_HIP(hipMemcpyAsync(om.data(), d_om, sizeof(double) * n, hipMemcpyDeviceToHost, c.st));
  C_HIP(hipMemcpyAsync(w.data(), d_w, sizeof(double) * n, hipMemcpyDeviceToHost, c.st));
  C_HIP(hipMemcpyAsync(patch.data(), d_patch, sizeof(int32_t) * n * 3, hipMemcpyDeviceToHost, c.st));
  C_HIP(hipStreamSynchronize(c.st));
  int64_t used = n;
  for (int64_t l = 0; l < n; l++) {
    if (acc.n >= n_target) {
      used = l;
      break;
    }
    if (!ok[l]) continue;
    acc.coords.insert(acc.coords.end(), coords.begin() + 16 * l, coords.begin() + 16 * l + 16);
    acc.om.push_back(om[l]);
    acc.w.push_back(w[l]);
    acc.patch.insert(acc.patch.end(), patch.begin() + 3 * l, patch.begin() + 3 * l + 3);
    acc.n++;
  }
  c.release(d_coords);
  c.release(d_om);
  c.release(d_w);
  c.release(d_patch);
  c.release(d_ok);
  *lines_used = used;
  return CY3Q_OK;
}

// The sequential samplers (restarts): windows of candidate starts, the chain through them, and
// the converged restart of each chained line.
int sample_restarts(Call &c, int variety, uint64_t seed, int64_t n_target, Accepted &acc,
                    int64_t *n_lines) {
  const int ln = kLineNormals[variety], rn = kRestartNormals[variety];
  const int64_t tail = ln + 4 * rn;
  const int64_t max_lines = std::max<int64_t>(n_target * 64, 1024);
  DeviceNormals dn(seed);
  // The previous window's normals [prev_lo, prev_hi) stay on the device: the next window starts
  // inside it (where the chain stopped), so its overlap is copied and only the rest is drawn.
  double *d_prev = nullptr;
  int64_t prev_lo = 0, prev_hi = 0;
  int64_t s = 0, lines = 0;
  while (acc.n < n_target && lines < max_lines) {
    // Lines per window: the expected remainder plus slack, at most 2^20 (the window then holds
    // up to about 2^26 normals).
    const int64_t want = std::min<int64_t>(
        (int64_t)1 << 20, std::max<int64_t>(256, (int64_t)((n_target - acc.n) * 1.15) + 64));
    const int64_t n_cand = want * (ln + rn) / 2;  // even starts in [s, s + 2 n_cand)
    const int64_t end = s + 2 * n_cand;
    // The window holds normals [s, hi), at least [s, end + tail) and everything drawn so far:
    // [s, prev_hi) comes from the previous window (prev_hi is always the stream's end), the
    // rest is drawn now.
    const int64_t hi = std::max<int64_t>(end + tail, dn.generated());
    double *d_nrm;
    uint8_t *d_info;
    C_TRY(c.alloc(&d_nrm, (size_t)(hi - s)));
    C_TRY(c.alloc(&d_info, (size_t)n_cand));
    if (s < prev_lo || (d_prev && prev_hi != dn.generated()) || s > dn.generated())
      return set_error(CY3Q_ERR_STATE, "sampler window does not continue the normal stream");
    if (d_prev && prev_hi > s)
      C_HIP(hipMemcpyAsync(d_nrm, d_prev + (s - prev_lo), sizeof(double) * (prev_hi - s),
                           hipMemcpyDeviceToDevice, c.st));
    if (hi > dn.generated())
      C_TRY(dn.generate(c, hi - dn.generated(), d_nrm + (dn.generated() - s)));
    if (d_prev) c.release(d_prev);
    d_prev = d_nrm;
    prev_lo = s;
    prev_hi = hi;
    launch_cicy_span(c.st, variety, d_nrm, 0, n_cand, d_info);
    C_HIP(hipGetLastError());
    std::vector<uint8_t> info((size_t)n_cand);
    C_HIP(hipMemcpyAsync(info.data(), d_info, n_cand, hipMemcpyDeviceToHost, c.st));
    C_HIP(hipStreamSynchronize(c.st));
    // The chain: each line starts where the previous line's restarts ended.
    std::vector<int64_t> starts;
    std::vector<uint8_t> cinfo;
    int64_t t = s;
    while (t < end && lines + (int64_t)starts.size() < max_lines) {
      const uint8_t inf = info[(size_t)((t - s) / 2)];
      starts.push_back(t - s);
      cinfo.push_back(inf);
      t += ln + (int64_t)rn * (inf & 7);
    }
    const int64_t nl = (int64_t)starts.size();
    int64_t *d_start;
    uint8_t *d_cinfo;
    double *d_x;
    int32_t *d_flag;
    C_TRY(c.alloc(&d_start, (size_t)nl));
    C_TRY(c.alloc(&d_cinfo, (size_t)nl));
    C_TRY(c.alloc(&d_x, (size_t)nl * 16));
    C_TRY(c.alloc(&d_flag, (size_t)nl));
    C_HIP(hipMemcpyAsync(d_start, starts.data(), sizeof(int64_t) * nl, hipMemcpyHostToDevice, c.st));
    C_HIP(hipMemcpyAsync(d_cinfo, cinfo.data(), nl, hipMemcpyHostToDevice, c.st));
    launch_cicy_chain(c.st, variety, d_nrm, d_start, d_cinfo, nl, d_x, d_flag);
    C_HIP(hipGetLastError());
    int64_t used = 0;
    C_TRY(post_and_collect(c, variety, d_x, d_flag, nl, n_target, acc, &used));
    lines += used;
    s = t;
    for (void *p : {(void *)d_info, (void *)d_start, (void *)d_cinfo, (void *)d_x,
                    (void *)d_flag})
      c.release(p);
  }
  *n_lines = lines;
  return CY3Q_OK;
}

// The batch sampler of the reference's GPU build: max(4 n_target, 64) lines drawn up front,
// 70 normals each, one damped Newton attempt per line. Lines are independent, so they run in
// chunks (in line order) until n_target points are kept; the draws of lines past the last one
// needed are never made (they only advance a generator the reference then drops).
int sample_lanes(Call &c, uint64_t seed, int64_t n_target, Accepted &acc, int64_t *n_lines) {
  const int64_t nl = std::max<int64_t>(4 * n_target, 64);
  // Chunks shrink as points accumulate, so the first one sizes the buffers.
  const int64_t cap = std::min<int64_t>(
      std::min<int64_t>(nl, (int64_t)1 << 20), (int64_t)(n_target * 1.08) + 64);
  DeviceNormals dn(seed);
  double *d_nrm, *d_x;
  int32_t *d_flag;
  C_TRY(c.alloc(&d_nrm, (size_t)cap * 70));
  C_TRY(c.alloc(&d_x, (size_t)cap * 16));
  C_TRY(c.alloc(&d_flag, (size_t)cap));
  for (int64_t l0 = 0; l0 < nl && acc.n < n_target;) {
    // the remaining points plus slack (about 97% of lines give a point)
    const int64_t want = (int64_t)((n_target - acc.n) * 1.08) + 64;
    const int64_t m = std::min(std::min(cap, want), nl - l0);
    C_TRY(dn.generate(c, 70 * m, d_nrm));  // normals [70 l0, 70 (l0 + m))
    launch_cicy_lanes(c.st, d_nrm, m, d_x, d_flag);
    C_HIP(hipGetLastError());
    int64_t used = 0;
    C_TRY(post_and_collect(c, CY3Q_CICY_TY, d_x, d_flag, m, n_target, acc, &used));
    l0 += m;
  }
  *n_lines = nl;
  return CY3Q_OK;
}

}  // namespace

extern "C" {

int cy3q_cicy_basis(int variety, const uint32_t *degrees, uint32_t *out, int64_t cap, int64_t *n) {
  if (!degrees || !n) return set_error(CY3Q_ERR_ARG, "NULL argument");
  if (variety != CY3Q_CICY_TY && variety != CY3Q_CICY_SCHOEN)
    return set_error(CY3Q_ERR_ARG, "variety must be CY3Q_CICY_TY or CY3Q_CICY_SCHOEN");
  for (int i = 0; i < (variety == CY3Q_CICY_TY ? 1 : 3); i++)
    if (degrees[i] > 40) return set_error(CY3Q_ERR_ARG, "degree above 40");
  const std::vector<Exp> *lms = nullptr;
  C_TRY(cached_lms(variety, &lms));
  auto keep = [&](const Exp &m) {
    for (const Exp &l : *lms)
      if (divides(l, m)) return false;
    return true;
  };
  std::vector<Exp> basis;
  if (variety == CY3Q_CICY_TY) {
    // enumerate_bigraded_kk_monomials (ty_metric.rs:792-820), z3_character (quotient.rs:55-57)
    const int k = (int)degrees[0];
    std::vector<std::array<uint8_t, 4>> zm;
    for (int a0 = 0; a0 <= k; a0++)
      for (int a1 = 0; a1 <= k - a0; a1++)
        for (int a2 = 0; a2 <= k - a0 - a1; a2++)
          zm.push_back({(uint8_t)a0, (uint8_t)a1, (uint8_t)a2, (uint8_t)(k - a0 - a1 - a2)});
    for (const auto &z : zm)
      for (const auto &w : zm) {
        Exp m{z[0], z[1], z[2], z[3], w[0], w[1], w[2], w[3]};
        if ((m[1] + 2 * m[2] + m[5] + 2 * m[6]) % 3 != 0) continue;
        if (keep(m)) basis.push_back(m);
      }
  } else {
    // enumerate_bidegree_monomials (z3xz3_projector.rs:298-315); alpha and beta characters
    // (:105-113)
    const int dx = (int)degrees[0], dy = (int)degrees[1], dt = (int)degrees[2];
    for (int a0 = 0; a0 <= dx; a0++)
      for (int a1 = 0; a1 <= dx - a0; a1++)
        for (int b0 = 0; b0 <= dy; b0++)
          for (int b1 = 0; b1 <= dy - b0; b1++)
            for (int c0 = 0; c0 <= dt; c0++) {
              Exp m{(uint8_t)a0, (uint8_t)a1, (uint8_t)(dx - a0 - a1), (uint8_t)b0, (uint8_t)b1,
                    (uint8_t)(dy - b0 - b1), (uint8_t)c0, (uint8_t)(dt - c0)};
              if ((m[1] + 2 * m[2] + m[4] + 2 * m[5]) % 3 != 0 || m[7] % 3 != 0) continue;
              if (keep(m)) basis.push_back(m);
            }
  }
  *n = (int64_t)basis.size();
  if (!out) return CY3Q_OK;
  if (cap < (int64_t)basis.size()) return set_error(CY3Q_ERR_ARG, "cap below the basis size");
  for (size_t a = 0; a < basis.size(); a++)
    for (int k = 0; k < 8; k++) out[8 * a + k] = basis[a][k];
  return CY3Q_OK;
}

int cy3q_cicy_normals(uint64_t seed, int64_t count, int device, double *out) {
  if (!out || count < 0) return set_error(CY3Q_ERR_ARG, "cy3q_cicy_normals: bad arguments");
  if (count == 0) return CY3Q_OK;
  Call c;
  C_TRY(c.open(device));
  double *d_out;
  C_TRY(c.alloc(&d_out, (size_t)count));
  DeviceNormals dn(seed);
  C_TRY(dn.generate(c, count, d_out));
  C_HIP(hipMemcpyAsync(out, d_out, sizeof(double) * count, hipMemcpyDeviceToHost, c.st));
  C_HIP(hipStreamSynchronize(c.st));
  return CY3Q_OK;
}

int cy3q_cicy_sample(int variety, int mode, uint64_t seed, int64_t n_target, int device,
                     double *coords, double *omega_sq, double *weight, int32_t *patch,
                     int64_t *n_out, int64_t *n_lines) {
  if (!coords || !omega_sq || !weight || !patch || !n_out)
    return set_error(CY3Q_ERR_ARG, "NULL argument");
  *n_out = 0;
  if (n_lines) *n_lines = 0;
  if (variety != CY3Q_CICY_TY && variety != CY3Q_CICY_SCHOEN)
    return set_error(CY3Q_ERR_ARG, "variety must be CY3Q_CICY_TY or CY3Q_CICY_SCHOEN");
  if (mode != CY3Q_CICY_RESTARTS && mode != CY3Q_CICY_LANES)
    return set_error(CY3Q_ERR_ARG, "mode must be CY3Q_CICY_RESTARTS or CY3Q_CICY_LANES");
  if (mode == CY3Q_CICY_LANES && variety != CY3Q_CICY_TY)
    return set_error(CY3Q_ERR_ARG, "the one-attempt batch sampler exists for Tian-Yau only");
  if (n_target <= 0 || n_target > ((int64_t)1 << 26))
    return set_error(CY3Q_ERR_ARG, "n_target must be in [1, 2^26]");
  Call c;
  C_TRY(c.open(device));
  Accepted acc;
  int64_t lines = 0;
  if (mode == CY3Q_CICY_LANES)
    C_TRY(sample_lanes(c, seed, n_target, acc, &lines));
  else
    C_TRY(sample_restarts(c, variety, seed, n_target, acc, &lines));
  // Weights to sum one: TY divides by the sum (cicy_sampler.rs:607-615), Schoen multiplies by its
  // reciprocal (schoen_sampler.rs:1174-1182).
  double sum = 0.0;
  for (int64_t i = 0; i < acc.n; i++) sum += acc.w[i];
  if (std::isfinite(sum) && sum > 0.0) {
    if (variety == CY3Q_CICY_TY) {
      for (int64_t i = 0; i < acc.n; i++) acc.w[i] /= sum;
    } else {
      const double inv = 1.0 / sum;
      for (int64_t i = 0; i < acc.n; i++) acc.w[i] *= inv;
    }
  }
  std::memcpy(coords, acc.coords.data(), sizeof(double) * 16 * acc.n);
  std::memcpy(omega_sq, acc.om.data(), sizeof(double) * acc.n);
  std::memcpy(weight, acc.w.data(), sizeof(double) * acc.n);
  std::memcpy(patch, acc.patch.data(), sizeof(int32_t) * 3 * acc.n);
  *n_out = acc.n;
  if (n_lines) *n_lines = lines;
  return CY3Q_OK;
}

int cy3q_cicy_frames(int variety, int64_t n, const double *coords, const int32_t *patch,
                     int device, double *frames) {
  if (!coords || !patch || !frames) return set_error(CY3Q_ERR_ARG, "NULL argument");
  if (variety != CY3Q_CICY_TY && variety != CY3Q_CICY_SCHOEN)
    return set_error(CY3Q_ERR_ARG, "variety must be CY3Q_CICY_TY or CY3Q_CICY_SCHOEN");
  if (n <= 0) return set_error(CY3Q_ERR_ARG, "n must be > 0");
  for (int64_t p = 0; p < n; p++) {
    const int32_t *q = patch + 3 * p;
    const bool good = variety == CY3Q_CICY_TY
                          ? (q[0] >= 0 && q[0] < 4 && q[1] >= 0 && q[1] < 4)
                          : (q[0] >= 0 && q[0] < 3 && q[1] >= 3 && q[1] < 6 && q[2] >= 6 && q[2] < 8);
    if (!good) return set_error(CY3Q_ERR_ARG, "patch index out of range at point " + std::to_string(p));
  }
  Call c;
  C_TRY(c.open(device));
  double *d_c, *d_f;
  int32_t *d_p;
  C_TRY(c.alloc(&d_c, (size_t)n * 16));
  C_TRY(c.alloc(&d_p, (size_t)n * 3));
  C_TRY(c.alloc(&d_f, (size_t)n * 48));
  C_HIP(hipMemcpyAsync(d_c, coords, sizeof(double) * 16 * n, hipMemcpyHostToDevice, c.st));
  C_HIP(hipMemcpyAsync(d_p, patch, sizeof(int32_t) * 3 * n, hipMemcpyHostToDevice, c.st));
  launch_cicy_frames(c.st, variety, d_c, d_p, n, d_f);
  C_HIP(hipGetLastError());
  C_HIP(hipMemcpyAsync(frames, d_f, sizeof(double) * 48 * n, hipMemcpyDeviceToHost, c.st));
  C_HIP(hipStreamSynchronize(c.st));
  return CY3Q_OK;
}

int cy3q_cicy_sections(int64_t n, const double *coords, int64_t n_basis, const uint32_t *exps,
                       int device, double *section_values, double *section_derivs) {
  if (!coords || !exps || !section_values || !section_derivs)
    return set_error(CY3Q_ERR_ARG, "NULL argument");
  if (n <= 0 || n_basis <= 0 || n_basis > 4096)
    return set_error(CY3Q_ERR_ARG, "n must be > 0 and n_basis in [1, 4096]");
  for (int64_t i = 0; i < 8 * n_basis; i++)
    if (exps[i] > 64) return set_error(CY3Q_ERR_ARG, "exponent above 64");
  Call c;
  C_TRY(c.open(device));
  double *d_c, *d_sv, *d_sd;
  uint32_t *d_e;
  C_TRY(c.alloc(&d_c, (size_t)n * 16));
  C_TRY(c.alloc(&d_e, (size_t)n_basis * 8));
  C_TRY(c.alloc(&d_sv, (size_t)n * 2 * n_basis));
  C_TRY(c.alloc(&d_sd, (size_t)n * 16 * n_basis));
  C_HIP(hipMemcpyAsync(d_c, coords, sizeof(double) * 16 * n, hipMemcpyHostToDevice, c.st));
  C_HIP(hipMemcpyAsync(d_e, exps, sizeof(uint32_t) * 8 * n_basis, hipMemcpyHostToDevice, c.st));
  launch_cicy_sections(c.st, d_c, n, d_e, (int)n_basis, d_sv, d_sd);
  C_HIP(hipGetLastError());
  C_HIP(hipMemcpyAsync(section_values, d_sv, sizeof(double) * n * 2 * n_basis, hipMemcpyDeviceToHost, c.st));
  C_HIP(hipMemcpyAsync(section_derivs, d_sd, sizeof(double) * n * 16 * n_basis, hipMemcpyDeviceToHost, c.st));
  C_HIP(hipStreamSynchronize(c.st));
  return CY3Q_OK;
}

}  // extern "C"
