// Metropolis walkers of a surrogate fitting loop (host side).  One step per
// call: proposals, one ensemble prediction with device misfits, decisions.
// The arithmetic is the Python fitting loop's, operation for operation and
// without fused multiply-adds, so the same random numbers give the same walk.
#include <cmath>

#include "walkers.h"

#pragma clang fp contract(off)

namespace imped {

double SplitMix64::normal() {
  if (has_spare) {
    has_spare = false;
    return spare;
  }
  const double u1 = 1.0 - uniform();  // (0, 1]
  const double u2 = uniform();
  const double r = std::sqrt(-2.0 * std::log(u1)), a = 2.0 * M_PI * u2;
  spare = r * std::sin(a);
  has_spare = true;
  return r * std::cos(a);
}

namespace {

void decide_one(Walkers& w, int i, double misfit) {
  const double u = w.urng.uniform();
  if (u < std::exp(-(misfit - w.cur[i]) / w.temperature)) {
    for (int j = 0; j < w.P; ++j) w.theta[(size_t)i * w.P + j] = w.prop[(size_t)i * w.P + j];
    w.cur[i] = misfit;
    w.best = std::min(w.best, misfit);
    ++w.n_accepted;
  }
}

// Tops the pool up to n unused normals.
void refill(Walkers& w, size_t n) {
  if (w.pool_head > 0 && w.pool_head >= w.pool.size() / 2) {
    w.pool.erase(w.pool.begin(), w.pool.begin() + w.pool_head);
    w.pool_head = 0;
  }
  while (w.pool.size() - w.pool_head < n) w.pool.push_back(w.rng.normal());
}

}  // namespace

int walkers_step(Walkers& w, const double* weights, double scale, const double* target, hipStream_t s) {
  const int P = w.P;
  w.free_idx.clear();
  for (int i = 0; i < w.W; ++i)
    if (!w.waiting[i]) w.free_idx.push_back(i);
  const int nf = (int)w.free_idx.size();
  if (nf == 0) return w.n_waiting;
  w.x.resize((size_t)nf * P);
  refill(w, (size_t)nf * P);
  // proposals theta + step (hi - lo) z, reflected into the box and clipped;
  // network inputs 2 (p - lo) / (hi - lo) - 1
  for (int k = 0; k < nf; ++k) {
    const int i = w.free_idx[k];
    for (int j = 0; j < P; ++j) {
      const double lo = w.lo[j], hi = w.hi[j];
      const double z = w.pool[w.pool_head++];
      double p = w.theta[(size_t)i * P + j] + w.span[j] * z;
      if (p < lo) p = 2.0 * lo - p;
      if (p > hi) p = 2.0 * hi - p;
      p = std::min(std::max(p, lo), hi);
      w.prop[(size_t)i * P + j] = p;
      w.x[(size_t)k * P + j] = (float)(2.0 * (p - lo) / (hi - lo) - 1.0);
    }
  }
  w.dis.resize(nf);
  w.mis.resize(nf);
  mlp_predict_begin(*w.net, w.x.data(), nf, weights, scale, target, nullptr, s);
  refill(w, (size_t)w.W * P);  // the next step's normals, while the GPU predicts
  mlp_predict_end(*w.net, nf, true, w.dis.data(), w.mis.data(), s);
  w.n_predicted += nf;
  for (int k = 0; k < nf; ++k) {
    const int i = w.free_idx[k];
    if (w.dis[k] < w.threshold) {
      if (w.n_served < w.record_cap)
        w.served.insert(w.served.end(), w.prop.begin() + (size_t)i * P, w.prop.begin() + (size_t)(i + 1) * P);
      ++w.n_served;
    }
  }
  // decisions in walker order, one uniform per served proposal
  for (int k = 0; k < nf; ++k) {
    const int i = w.free_idx[k];
    if (w.dis[k] < w.threshold) {
      decide_one(w, i, w.mis[k]);
    } else {
      w.waiting[i] = 1;
      ++w.n_waiting;
    }
  }
  return w.n_waiting;
}

void walkers_decide(Walkers& w, int count, const int32_t* idx, const double* misfit) {
  for (int k = 0; k < count; ++k) {
    const int i = idx[k];
    decide_one(w, i, misfit[k]);
    if (w.waiting[i]) {
      w.waiting[i] = 0;
      --w.n_waiting;
    }
  }
}

}  // namespace imped
