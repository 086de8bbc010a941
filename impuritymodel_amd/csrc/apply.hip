// Operator application kernels: one lane per source product state.
//
// Replaces the inner loops of applyOp (impurityModel/ed/finite.py:1357-1388,
// 1415-1458) used by the Hamiltonian build (get_hamiltonian_matrix,
// finite.py:1585-1609), the BFS closure (expand_basis, finite.py:1737-1798)
// and transition operators (getSpectra, spectra.py:786).
//
// Each process (term) of an operator is compiled on the host to bit masks:
// it is non-zero on state s iff (s & care) == val, it maps s to s ^ flip, and
// its fermion sign is (-1)^(popcount(s & sgn) + parity) -- the product of the
// reference's sequential create/remove signs (create.py:142-148,
// remove.py:141-147).  Terms are grouped by flip mask: a group maps a source
// to exactly one target, so per-source duplicates merge inside the group in
// the operator dict's order, the order in which applyOp adds them.
//
// A wave handles 64 sources of one model, so the group/term loops are
// wave-uniform.  The wave copies its model's group and term tables into its
// own LDS slice in chunks (coalesced vector loads, TERM_CHUNK terms and
// GROUP_CHUNK groups at a time), and the loops read them from there: a term
// then costs an LDS broadcast instead of a dependent scalar-cache miss.
#include "kernels.h"

namespace imped {

__device__ __forceinline__ bool key_eq(const uint64_t* a, const uint64_t* b, int W) {
  for (int w = 0; w < W; ++w)
    if (a[w] != b[w]) return false;
  return true;
}

__device__ __forceinline__ uint64_t load_coherent(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int32_t load_coherent(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Look up (chain, key); returns pool id, CAND_FLAG|cand id or HASH_EMPTY.
template <int W>
__device__ int32_t hash_find(const HashView& h, const CandView* c, int32_t chain, const uint64_t* key) {
  uint32_t slot = hash_key((uint32_t)chain, key, W) & h.mask;
  for (;;) {
    int32_t cur = load_coherent(&h.table[slot]);
    if (cur == HASH_EMPTY) return HASH_EMPTY;
    if (cur & CAND_FLAG) {
      if (c != nullptr) {
        int32_t ci = cur & ~CAND_FLAG;
        uint64_t k[W];
        for (int w = 0; w < W; ++w) k[w] = load_coherent(&c->keys[(size_t)ci * W + w]);
        if (load_coherent(&c->chain[ci]) == chain && key_eq(k, key, W)) return cur;
      }
    } else {
      if (h.pool_chain[cur] == chain && key_eq(h.pool_keys + (size_t)cur * W, key, W)) return cur;
    }
    slot = (slot + 1) & h.mask;
  }
}

// Insert (chain, key) as a new candidate unless present.  The candidate's
// key is written and fenced before its id is published by the CAS, and
// readers load candidate keys with device-coherent loads.
template <int W>
__device__ void hash_discover(const HashView& h, const CandView& c, int32_t chain, const uint64_t* key) {
  uint32_t slot = hash_key((uint32_t)chain, key, W) & h.mask;
  int32_t mine = -1;
  for (;;) {
    int32_t cur = load_coherent(&h.table[slot]);
    if (cur == HASH_EMPTY) {
      if (mine < 0) {
        mine = atomicAdd(c.count, 1);
        if (mine >= c.cap) return;  // host detects the overflow from the count
        for (int w = 0; w < W; ++w) c.keys[(size_t)mine * W + w] = key[w];
        c.chain[mine] = chain;
        __threadfence();
      }
      int32_t prev = atomicCAS(&h.table[slot], HASH_EMPTY, CAND_FLAG | mine);
      if (prev == HASH_EMPTY) {
        c.slot[mine] = (int32_t)slot;
        int32_t a = atomicAdd(c.alive_count, 1);
        c.alive_list[a] = mine;
        return;
      }
      cur = prev;
    }
    bool same;
    if (cur & CAND_FLAG) {
      int32_t ci = cur & ~CAND_FLAG;
      uint64_t k[W];
      for (int w = 0; w < W; ++w) k[w] = load_coherent(&c.keys[(size_t)ci * W + w]);
      same = load_coherent(&c.chain[ci]) == chain && key_eq(k, key, W);
    } else {
      same = h.pool_chain[cur] == chain && key_eq(h.pool_keys + (size_t)cur * W, key, W);
    }
    if (same) {
      if (mine >= 0) c.chain[mine] = -1;  // claimed slot stays unused
      return;
    }
    slot = (slot + 1) & h.mask;
  }
}

constexpr int TERM_CHUNK = 128;  // terms per staged chunk (per wave)
constexpr int GROUP_CHUNK = 64;  // groups per staged chunk (per wave)
constexpr int RESTR_LDS = 16;    // restrictions kept in LDS (more are read from memory)

// LDS writes by some lanes of a wave, read by every lane of the same wave.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int W, int MODE, bool FILL>
__global__ __launch_bounds__(256) void apply_kernel(ApplyArgs a) {
  __shared__ uint64_t l_mask[4][TERM_CHUNK * 3 * W];
  __shared__ double2 l_coef[4][TERM_CHUNK];
  __shared__ uint32_t l_meta[4][TERM_CHUNK];
  __shared__ int32_t l_gto[4][GROUP_CHUNK + 1];
  __shared__ uint64_t l_flip[4][GROUP_CHUNK * W];
  __shared__ uint64_t l_rmask[4][RESTR_LDS * W];
  __shared__ int32_t l_rlohi[4][RESTR_LDS * 2];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int64_t task_id = (int64_t)blockIdx.x * (blockDim.x >> 6) + wv;
  if (task_id >= a.n_tasks) return;
  const WaveTask task = a.tasks[task_id];
  const int model = __builtin_amdgcn_readfirstlane(task.model);
  const int chain = __builtin_amdgcn_readfirstlane(task.chain);
  const int64_t src = task.begin + lane;
  const bool active = src < task.end;

  uint64_t s[W];
  for (int w = 0; w < W; ++w) s[w] = active ? a.src_keys[src * W + w] : 0ull;
  double2 amp = make_double2(1.0, 0.0);
  int32_t spos = 0;
  if (MODE == MODE_VECTOR && active) {
    amp = a.src_amps[src];
    spos = a.src_pos[src];
  }
  int64_t out = 0;
  int32_t cnt = 0;
  if (FILL && active) out = a.offsets[src];

  const OpsView& o = a.ops;
  const int g0 = o.model_group_off[model];
  const int g1 = o.model_group_off[model + 1];
  const int r0 = o.model_restr_off[model];
  const int r1 = o.model_restr_off[model + 1];
  const int t_end = o.group_term_off[g1];  // the model's terms end here
  uint64_t* lm = l_mask[wv];
  double2* lc = l_coef[wv];
  uint32_t* lt = l_meta[wv];
  int32_t* lg = l_gto[wv];
  uint64_t* lf = l_flip[wv];
  int gc0 = g0, gc1 = g0;  // staged groups [gc0, gc1)
  int tc0 = 0, tc1 = 0;    // staged terms [tc0, tc1)
  const int nr = r1 - r0;
  const int nrl = nr < RESTR_LDS ? nr : RESTR_LDS;
  uint64_t* lrm = l_rmask[wv];
  int32_t* lrl = l_rlohi[wv];
  if (MODE != MODE_FIXED) {
    for (int i = lane; i < nrl * W; i += 64) lrm[i] = o.restr_mask[(size_t)r0 * W + i];
    for (int i = lane; i < nrl * 2; i += 64) lrl[i] = o.restr_lohi[2 * r0 + i];
  }

  for (int g = g0; g < g1; ++g) {
    if (g >= gc1) {
      wave_lds_sync();  // every lane is done with the previous chunk
      gc0 = g;
      gc1 = g + GROUP_CHUNK < g1 ? g + GROUP_CHUNK : g1;
      for (int i = lane; i <= gc1 - gc0; i += 64) lg[i] = o.group_term_off[gc0 + i];
      for (int i = lane; i < (gc1 - gc0) * W; i += 64) lf[i] = o.group_flip[(size_t)gc0 * W + i];
      wave_lds_sync();
    }
    const int tb = lg[g - gc0];
    const int te = lg[g - gc0 + 1];
    double2 acc = make_double2(0.0, 0.0);
    bool any = false;
    int first = 0;
    // one matching term: add its signed value (in term order)
    auto take = [&](int t) {
      const uint64_t* m = lm + (size_t)(t - tc0) * 3 * W;
      const uint32_t meta = lt[t - tc0];
      uint32_t par = meta & 1u;
#pragma unroll
      for (int w = 0; w < W; ++w) par ^= (uint32_t)__popcll(s[w] & m[2 * W + w]) & 1u;
      const double2 c = lc[t - tc0];
      if (par) {
        acc.x -= c.x;
        acc.y -= c.y;
      } else {
        acc.x += c.x;
        acc.y += c.y;
      }
      if (!any) first = (int)(meta >> 1);
      any = true;
    };
    auto matches = [&](int t) {
      const uint64_t* m = lm + (size_t)(t - tc0) * 3 * W;
      bool ok = true;
#pragma unroll
      for (int w = 0; w < W; ++w) ok = ok && ((s[w] & m[w]) == m[W + w]);
      return ok;
    };
    for (int t = tb; t < te;) {
      if (t >= tc1 || te > tc1) {
        // stage from this term on (a group is staged whole when it fits a chunk)
        if (t >= tc1 || (t == tb && te - tb <= TERM_CHUNK)) {
          wave_lds_sync();
          tc0 = t;
          tc1 = t + TERM_CHUNK < t_end ? t + TERM_CHUNK : t_end;
          for (int i = lane; i < (tc1 - tc0) * 3 * W; i += 64) lm[i] = o.term_masks[(size_t)tc0 * 3 * W + i];
          for (int i = lane; i < tc1 - tc0; i += 64) {
            lc[i] = o.term_coef[tc0 + i];
            lt[i] = o.term_meta[tc0 + i];
          }
          wave_lds_sync();
        }
      }
      const int tend = te < tc1 ? te : tc1;
      // four mask tests per round trip, the matches taken in term order
      for (; t + 4 <= tend; t += 4) {
        const bool k0 = matches(t), k1 = matches(t + 1), k2 = matches(t + 2), k3 = matches(t + 3);
        if (k0) take(t);
        if (k1) take(t + 1);
        if (k2) take(t + 2);
        if (k3) take(t + 3);
      }
      for (; t < tend; ++t)
        if (matches(t)) take(t);
    }
    if (!(any && active)) continue;
    uint64_t tgt[W];
#pragma unroll
    for (int w = 0; w < W; ++w) tgt[w] = s[w] ^ lf[(g - gc0) * W + w];
    bool keep = true;
    if (MODE != MODE_VECTOR) keep = abs2_ref(acc) >= a.weight_min;
    if (MODE != MODE_FIXED) {
      for (int r = 0; r < nrl && keep; ++r) {
        int n = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) n += __popcll(tgt[w] & lrm[r * W + w]);
        keep = n >= lrl[2 * r] && n <= lrl[2 * r + 1];
      }
      for (int r = r0 + nrl; r < r1 && keep; ++r) {
        int n = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) n += __popcll(tgt[w] & o.restr_mask[(size_t)r * W + w]);
        keep = n >= o.restr_lohi[2 * r] && n <= o.restr_lohi[2 * r + 1];
      }
    }
    int32_t local = 0;
    if (MODE == MODE_FIXED && keep) {
      int32_t id = hash_find<W>(a.hash, nullptr, chain, tgt);
      keep = id >= 0;
      if (keep) local = a.pool_local[id];
    }
    if (!keep) continue;
    if (!FILL) {
      ++cnt;
      continue;
    }
    if (MODE == MODE_EXPAND) {
      // new targets are entered into the hash by discover_kernel afterwards
      for (int w = 0; w < W; ++w) a.out_keys[out * W + w] = tgt[w];
      a.out_val[out] = acc;
      a.out_chain[out] = chain;
    } else if (MODE == MODE_FIXED) {
      a.out_col[out] = local;
      a.out_val[out] = cconj(acc);
    } else {
      for (int w = 0; w < W; ++w) a.out_keys[out * W + w] = tgt[w];
      a.out_val[out] = cmul(amp, acc);
      a.out_vec[out] = chain;
      a.out_src[out] = spos;
      a.out_first[out] = first;
    }
    ++out;
  }
  if (!FILL && active) a.counts[src] = cnt;
}

template <int W>
static void launch_w(int mode, bool fill, const ApplyArgs& a, hipStream_t s) {
  const int threads = 256;
  const unsigned blocks = blocks_for(a.n_tasks, threads / 64);
  if (a.n_tasks == 0) return;
  TimedLaunch tl(fill ? "apply_fill" : "apply_count", s);
#define IMPED_LAUNCH(M, F) hipLaunchKernelGGL((apply_kernel<W, M, F>), dim3(blocks), dim3(threads), 0, s, a)
  if (mode == MODE_EXPAND) {
    if (fill) IMPED_LAUNCH(MODE_EXPAND, true); else IMPED_LAUNCH(MODE_EXPAND, false);
  } else if (mode == MODE_FIXED) {
    if (fill) IMPED_LAUNCH(MODE_FIXED, true); else IMPED_LAUNCH(MODE_FIXED, false);
  } else {
    if (fill) IMPED_LAUNCH(MODE_VECTOR, true); else IMPED_LAUNCH(MODE_VECTOR, false);
  }
#undef IMPED_LAUNCH
  HIP_CHECK(hipGetLastError());
}

// Every expand-fill entry (chain[e], keys[e]) becomes a candidate unless it
// is a pool state or already a candidate: one thread per entry, so the
// dependent hash probes of many entries overlap.
template <int W>
__global__ __launch_bounds__(256) void discover_kernel(const uint64_t* keys, const int32_t* chain, int64_t E, HashView h,
                                                       CandView c) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  uint64_t k[W];
#pragma unroll
  for (int w = 0; w < W; ++w) k[w] = keys[e * W + w];
  hash_discover<W>(h, c, chain[e], k);
}

void launch_discover(int W, const uint64_t* keys, const int32_t* chain, int64_t E, HashView h, CandView c,
                     hipStream_t s) {
  if (E <= 0) return;
  TimedLaunch tl("discover", s);
  if (W == 1)
    hipLaunchKernelGGL(discover_kernel<1>, dim3(blocks_for(E, 256)), dim3(256), 0, s, keys, chain, E, h, c);
  else
    hipLaunchKernelGGL(discover_kernel<2>, dim3(blocks_for(E, 256)), dim3(256), 0, s, keys, chain, E, h, c);
  HIP_CHECK(hipGetLastError());
}

void launch_apply(int W, int mode, bool fill, const ApplyArgs& a, hipStream_t s) {
  if (W == 1)
    launch_w<1>(mode, fill, a, s);
  else
    launch_w<2>(mode, fill, a, s);
}

// ------------------------------------------------------------------ hash maintenance

__global__ void hash_clear_kernel(int32_t* t, uint32_t cap) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cap) t[i] = HASH_EMPTY;
}

void launch_hash_clear(int32_t* table, uint32_t cap, hipStream_t s) {
  hipLaunchKernelGGL(hash_clear_kernel, dim3(blocks_for(cap, 256)), dim3(256), 0, s, table, cap);
  HIP_CHECK(hipGetLastError());
}

// Insert pool ids [first, first+count); ids are unique per (chain, key).
template <int W>
__global__ void hash_insert_pool_kernel(HashView h, int64_t first, int64_t count) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int32_t id = (int32_t)(first + i);
  const uint64_t* key = h.pool_keys + (size_t)id * W;
  uint32_t slot = hash_key((uint32_t)h.pool_chain[id], key, W) & h.mask;
  for (;;) {
    int32_t prev = atomicCAS(&h.table[slot], HASH_EMPTY, id);
    if (prev == HASH_EMPTY) return;
    slot = (slot + 1) & h.mask;
  }
}

void launch_hash_insert_pool(int W, HashView h, int64_t first, int64_t count, hipStream_t s) {
  if (count == 0) return;
  if (W == 1)
    hipLaunchKernelGGL(hash_insert_pool_kernel<1>, dim3(blocks_for(count, 256)), dim3(256), 0, s, h, first, count);
  else
    hipLaunchKernelGGL(hash_insert_pool_kernel<2>, dim3(blocks_for(count, 256)), dim3(256), 0, s, h, first, count);
  HIP_CHECK(hipGetLastError());
}

}  // namespace imped
