// The reference crate's LCG (cy3/src/lib.rs:222-245) for the device and the host: the 64-bit
// multiplier and increment of lib.rs:230-237, the seeding of LCG::new, 53-bit doubles, and
// Box-Muller normals (:239-244), plus an O(log n) jump-ahead so that draw q of a stream can be
// taken by thread q.
#pragma once
#include "cy3q_internal.h"

namespace cy3q {

static constexpr uint64_t LCG_MUL = 6364136223846793005ULL;
static constexpr uint64_t LCG_INC = 1442695040888963407ULL;

// LCG::new(seed): the state after seeding (lib.rs:230-233).
__host__ __device__ inline uint64_t lcg_seed(uint64_t seed) { return seed * LCG_MUL + 1ULL; }

struct Lcg {
  uint64_t s;
  __host__ __device__ double next_f64() {
    s = s * LCG_MUL + LCG_INC;
    uint64_t bits = (s >> 11) & ((1ULL << 53) - 1ULL);
    return (double)bits / (double)(1ULL << 53);
  }
  __device__ double next_normal() {
    double u1 = fmax(next_f64() + 1e-12, 1e-12);
    double u2 = next_f64();
    double r = sqrt(-2.0 * log(u1));
    return r * cos(2.0 * M_PI * u2);
  }
};

// State after `steps` applications of s -> MUL*s + INC, starting from s0.
__device__ inline uint64_t lcg_jump(uint64_t s0, uint64_t steps) {
  uint64_t acc_a = 1, acc_c = 0, a = LCG_MUL, c = LCG_INC;
  while (steps) {
    if (steps & 1ULL) {
      acc_a = a * acc_a;
      acc_c = a * acc_c + c;
    }
    c = a * c + c;
    a = a * a;
    steps >>= 1;
  }
  return acc_a * s0 + acc_c;
}

}  // namespace cy3q
