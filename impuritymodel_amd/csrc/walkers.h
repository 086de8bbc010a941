// Metropolis walkers of a surrogate fitting loop, stepped on the host next to
// the ensemble's device prediction.
#pragma once

#include <cstdint>
#include <vector>

#include "common.h"
#include "mlp.h"

namespace imped {

// SplitMix64 (Steele, Lea and Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014): the state advances by the golden-ratio gamma and
// each output is the state through the paper's 64-bit finalizer.
struct SplitMix64 {
  uint64_t state;
  uint64_t next() {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // the top 53 bits as a double in [0, 1)
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
  // Box-Muller: two uniforms give two normals, the cosine one first, then
  // the sine one
  double normal();
  double spare = 0.0;
  bool has_spare = false;
};

struct Walkers {
  Mlp* net = nullptr;
  int W = 0, P = 0;
  std::vector<double> theta, cur, prop;  // [W][P], [W], [W][P]
  std::vector<double> lo, hi, span;      // box and step * (hi - lo), [P]
  std::vector<uint8_t> waiting;          // [W]
  double temperature = 0.0, threshold = 0.0, best = 0.0;
  SplitMix64 rng{0};   // normals (proposals)
  SplitMix64 urng{0};  // uniforms (decisions)
  // normals drawn ahead, first in first out: the pool is refilled while the
  // GPU predicts, and the k-th normal used is always the stream's k-th
  std::vector<double> pool;
  size_t pool_head = 0;
  // parameters of the surrogate-served proposals, [n][P]: the first
  // record_cap of them since the last counter reset (a bounded record: a long
  // run does not grow host memory without limit)
  std::vector<double> served;
  int64_t record_cap = int64_t(1) << 20;
  int64_t n_predicted = 0, n_served = 0, n_accepted = 0;
  int n_waiting = 0;
  // per-step scratch
  std::vector<int32_t> free_idx;
  std::vector<float> x;
  std::vector<double> dis, mis;
};

// One step: every walker that is not waiting proposes, the ensemble answers
// all proposals in one prediction (misfit to target on the device), and each
// trusted proposal (disagreement < threshold) takes its Metropolis decision;
// the others wait.  Returns the number of waiting walkers.
int walkers_step(Walkers& w, const double* weights, double scale, const double* target, hipStream_t s);
// Metropolis decisions of count waiting walkers idx[] (ascending) with the
// given misfits (from exact evaluation); they stop waiting.
void walkers_decide(Walkers& w, int count, const int32_t* idx, const double* misfit);

}  // namespace imped
