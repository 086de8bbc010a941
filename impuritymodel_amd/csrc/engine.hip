// Host orchestration and the C ABI of the engine (libimped.so).
//
// Handles own device memory.  A *batch* is a set of chains, each the
// Hilbert-space sector of one Green's function (or one model's ground-state
// sector); every step below runs for all chains of a batch in the same
// launches.
//
//   ops    compiled operator tables of several models        (imped_ops_*)
//   vec    sparse multi-configurational states on the device (imped_vec_*)
//   batch  bases + CSR Hamiltonians + Lanczos/eigen results  (imped_batch_*)
//
// Reference counterparts: applyOp (impurityModel/ed/finite.py:1301),
// get_hamiltonian_matrix (finite.py:1514), expand_basis_and_hamiltonian
// (finite.py:1806), get_tridiagonal_krylov_vectors (finite.py:1897),
// eigensystem (finite.py:58), the continued fraction of getGreen
// (spectra.py:677-680) and thermal_average (average.py:11).
#include <algorithm>
#include <array>
#include <stdexcept>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <unordered_map>

#include "../../include/imped.h"
#include "kernels.h"
#include "krylov.h"
#include "mlp.h"
#include "walkers.h"

namespace imped {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

Tuning g_tuning;

// ------------------------------------------------------------------ staged small reads

namespace {
struct PendingRead {
  void* dst;
  const char* slot;
  size_t bytes;
  hipStream_t s;
};
struct Staging {
  std::mutex mu;
  char* ring = nullptr;
  size_t cap = 0, pos = 0;
  std::vector<PendingRead> pending;
};
Staging& staging() {
  static Staging* st = new Staging;
  return *st;
}
constexpr size_t kStageMax = 4 << 20;    // reads up to this size are staged
constexpr size_t kRingBytes = 32 << 20;  // pinned ring
}  // namespace

void d2h_bytes(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (!bytes) return;
  if (bytes > kStageMax) {
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    return;
  }
  Staging& st = staging();
  {
    std::lock_guard<std::mutex> lk(st.mu);
    if (!st.ring) {
      HIP_CHECK(hipHostMalloc((void**)&st.ring, kRingBytes, hipHostMallocDefault));
      st.cap = kRingBytes;
    }
  }
  const size_t need = (bytes + 63) & ~(size_t)63;
  if (st.pos + need > st.cap) stream_sync(s);
  std::lock_guard<std::mutex> lk(st.mu);
  if (st.pos + need > st.cap) {
    // reads of other streams still hold the ring: read this one directly
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    return;
  }
  char* slot = st.ring + st.pos;
  st.pos += need;
  HIP_CHECK(hipMemcpyAsync(slot, src, bytes, hipMemcpyDeviceToHost, s));
  st.pending.push_back(PendingRead{dst, slot, bytes, s});
}

// ------------------------------------------------------------------ staged small uploads

namespace {
struct Uploads {
  std::mutex mu;
  char* ring = nullptr;
  size_t cap = 0, pos = 0;
  std::vector<hipStream_t> users;  // streams with queued copies from the ring
};
Uploads& uploads() {
  static Uploads* u = new Uploads;
  return *u;
}
constexpr size_t kUpMax = 1 << 20;    // uploads up to this size are staged
constexpr size_t kUpRing = 16 << 20;  // pinned ring
}  // namespace

void h2d_bytes(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (!bytes) return;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  HIP_CHECK(hipStreamIsCapturing(s, &cs));
  if (bytes <= kUpMax && cs == hipStreamCaptureStatusNone) {
    Uploads& u = uploads();
    std::lock_guard<std::mutex> lk(u.mu);
    if (!u.ring) {
      HIP_CHECK(hipHostMalloc((void**)&u.ring, kUpRing, hipHostMallocDefault));
      u.cap = kUpRing;
    }
    const size_t need = (bytes + 255) & ~(size_t)255;
    if (u.pos + need <= u.cap) {
      char* slot = u.ring + u.pos;
      u.pos += need;
      std::memcpy(slot, src, bytes);
      if (std::find(u.users.begin(), u.users.end(), s) == u.users.end()) u.users.push_back(s);
      HIP_CHECK(hipMemcpyAsync(dst, slot, bytes, hipMemcpyHostToDevice, s));
      return;
    }
  }
  HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
}

void stream_sync(hipStream_t s) {
  HIP_CHECK(hipStreamSynchronize(s));
  {
    // the ring's copies on s are done; the ring is free once no stream holds it
    Uploads& u = uploads();
    std::lock_guard<std::mutex> lk(u.mu);
    u.users.erase(std::remove(u.users.begin(), u.users.end(), s), u.users.end());
    if (u.users.empty()) u.pos = 0;
  }
  Staging& st = staging();
  std::lock_guard<std::mutex> lk(st.mu);
  size_t keep = 0;
  for (const PendingRead& r : st.pending) {
    if (r.s == s) std::memcpy(r.dst, r.slot, r.bytes);
    else st.pending[keep++] = r;
  }
  st.pending.resize(keep);
  if (st.pending.empty()) st.pos = 0;
}

// ------------------------------------------------------------------ device memory cache

namespace {
struct MemCache {
  std::mutex mu;
  std::unordered_map<size_t, std::vector<void*>> idle;
  size_t idle_bytes = 0;
};
// never destroyed: handles released during process exit still find it
MemCache& mem_cache() {
  static MemCache* c = new MemCache;
  return *c;
}
constexpr size_t kIdleCap = size_t(16) << 30;  // idle blocks kept at most
}  // namespace

size_t devmem_class(size_t n) {
  if (n <= 256) return 256;
  if (n <= (size_t(64) << 20)) {
    size_t c = 256;
    while (c < n) c <<= 1;
    return c;
  }
  const size_t g = size_t(2) << 20;
  return (n + g - 1) / g * g;
}

void* devmem_alloc(size_t bytes) {
  MemCache& mc = mem_cache();
  {
    std::lock_guard<std::mutex> lk(mc.mu);
    auto it = mc.idle.find(bytes);
    if (it != mc.idle.end() && !it->second.empty()) {
      void* p = it->second.back();
      it->second.pop_back();
      mc.idle_bytes -= bytes;
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e == hipErrorOutOfMemory) {
    // give the idle blocks back to the runtime and try once more
    (void)hipGetLastError();
    std::lock_guard<std::mutex> lk(mc.mu);
    (void)hipDeviceSynchronize();
    for (auto& kv : mc.idle)
      for (void* q : kv.second) (void)hipFree(q);
    mc.idle.clear();
    mc.idle_bytes = 0;
    e = hipMalloc(&p, bytes);
  }
  HIP_CHECK(e);
  return p;
}

void devmem_free(void* p, size_t bytes) {
  if (!p) return;
  MemCache& mc = mem_cache();
  std::lock_guard<std::mutex> lk(mc.mu);
  if (mc.idle_bytes + bytes > kIdleCap) {
    (void)hipFree(p);
    return;
  }
  mc.idle[bytes].push_back(p);
  mc.idle_bytes += bytes;
}

// ------------------------------------------------------------------ launch timing

struct TimingRecord {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
};
static bool g_timing = false;
static std::vector<std::pair<std::string, TimingRecord>> g_timings;

bool timing_enabled() { return g_timing; }

static TimingRecord& timing_record(const char* name) {
  for (auto& kv : g_timings)
    if (kv.first == name) return kv.second;
  g_timings.emplace_back(name, TimingRecord{});
  return g_timings.back().second;
}

TimedLaunch::TimedLaunch(const char* n, hipStream_t st) : name(n), s(st) {
  if (!g_timing) return;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  HIP_CHECK(hipEventRecord(e0, s));
}

TimedLaunch::~TimedLaunch() {
  if (!e0) return;
  (void)hipEventRecord(e1, s);
  timing_record(name).events.emplace_back(e0, e1);
}

static void timing_clear() {
  for (auto& kv : g_timings)
    for (auto& ev : kv.second.events) {
      (void)hipEventDestroy(ev.first);
      (void)hipEventDestroy(ev.second);
    }
  g_timings.clear();
}

static hipStream_t g_stream = nullptr;

struct Ops {
  int W = 1, M = 0, n_orb = 0;
  int64_t T = 0;      // terms over all models
  bool real = false;  // every term value is real, so H is real in the product basis
  DevBuf model_group_off, group_term_off, group_flip, term_masks, term_meta, term_coef, model_restr_off, restr_mask,
      restr_lohi;
  OpsView view() const {
    return OpsView{model_group_off.as<int32_t>(), group_term_off.as<int32_t>(), group_flip.as<uint64_t>(),
                   term_masks.as<uint64_t>(),     term_meta.as<uint32_t>(),     term_coef.as<double2>(),
                   model_restr_off.as<int32_t>(), restr_mask.as<uint64_t>(),    restr_lohi.as<int32_t>()};
  }
};

// Sparse vectors: vector v holds entries [off[v], off[v+1]).
struct Vec {
  int W = 1, V = 0;
  std::vector<int64_t> h_off;  // V+1
  DevBuf keys, amps, off, vid, pos, norm2;
  int64_t total() const { return h_off.empty() ? 0 : h_off.back(); }
};

struct Batch {
  Ops* ops = nullptr;
  int C = 0, W = 1;
  std::vector<int32_t> h_chain_model;
  DevBuf chain_model;
  bool chain_model_dev = false;
  // seeds (host offsets; device keys/amps)
  std::vector<int64_t> h_seed_off;
  DevBuf seed_keys, seed_amps, seed_off_dev;
  bool have_seed_amps = false;
  // state pool (global ids in discovery order)
  int64_t P = 0, Pcap = 0;
  DevBuf pool_keys, pool_chain, pool_local, pool_coloff, pool_colcnt;
  // hash
  uint32_t hcap = 0;
  DevBuf table;
  // per-chain sizes
  DevBuf chain_size, seg_start;
  std::vector<int64_t> h_n, h_row_base, h_nnz;
  // column storage (expansion)
  int64_t E = 0;
  DevBuf col_keys, col_val;
  // CSR in chain-local order, rows of chain c at [row_base[c], row_base[c]+n[c])
  int64_t R = 0, NNZ = 0;
  DevBuf rowptr, col, val, row_chain, row_base, chain_n;
  // Lanczos
  int K = 0;
  DevBuf alpha, beta, kout;  // kout holds kout[C] then status[C]
  // ground states
  int nev = 0;
  DevBuf evals, evecs;
  bool built = false;
  bool row_keys = false;  // pool_keys holds the basis in row order (built, not set_csr)
  bool real = false;      // all CSR values real
  // start vectors formed on the device (imped_batch_seed_affine), one entry per row
  DevBuf v0dev;
  bool have_v0dev = false;
  DevBuf fill_meta;  // chain table of imped_batch_set_affine
  DevBuf seed_meta, norm2dev;  // chain table and |T psi|^2 of imped_batch_seed_affine
  bool have_norm2dev = false;
  DevBuf val_re;  // real parts of val for the grid-wide kernels (real CSR with a chain above flat_min_n)
  bool have_val_re = false;
};

// Seed amplitudes are placed on the first rows of their chain: every chain
// must hold at least as many rows as it has seeds.
static bool seeds_fit(const Batch& b) {
  if ((int)b.h_seed_off.size() != b.C + 1) return false;
  for (int c = 0; c < b.C; ++c)
    if (b.h_seed_off[c + 1] - b.h_seed_off[c] > b.h_n[c]) return false;
  return true;
}

// ------------------------------------------------------------------ small kernels

__global__ void set_local_seed_kernel(const int64_t* seed_off, int C, int32_t* pool_local, int32_t* pool_chain,
                                      int64_t* chain_size) {
  int c = blockIdx.x;
  if (c >= C) return;
  int64_t a = seed_off[c], b = seed_off[c + 1];
  for (int64_t i = a + threadIdx.x; i < b; i += blockDim.x) {
    pool_local[i] = (int32_t)(i - a);
    pool_chain[i] = c;
  }
  if (threadIdx.x == 0) chain_size[c] = b - a;
}

__global__ void append_layer_kernel(int W, const int32_t* perm, int64_t n, const uint64_t* cand_keys,
                                    const int32_t* cand_chain, const int32_t* cand_slot, int64_t P0, uint64_t* pool_keys,
                                    int32_t* pool_chain, int32_t* table) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t c = perm[i];
  int64_t g = P0 + i;
  for (int w = 0; w < W; ++w) pool_keys[g * W + w] = cand_keys[(size_t)c * W + w];
  pool_chain[g] = cand_chain[c];
  table[cand_slot[c]] = (int32_t)g;
}

__global__ void layer_heads_kernel(const int32_t* pool_chain, int64_t P0, int64_t n, int64_t* seg_start) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t c = pool_chain[P0 + i];
  if (i == 0 || pool_chain[P0 + i - 1] != c) seg_start[c] = i;
}

// Local index of each new state: chain size before the layer + rank in its
// chain's segment.  chain_size is only read here ...
__global__ void layer_local_kernel(const int32_t* pool_chain, int64_t P0, int64_t n, const int64_t* seg_start,
                                   const int64_t* chain_size, int32_t* pool_local) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t c = pool_chain[P0 + i];
  pool_local[P0 + i] = (int32_t)(chain_size[c] + (i - seg_start[c]));
}

// ... and grown afterwards by the last state of each segment.
__global__ void layer_grow_kernel(const int32_t* pool_chain, int64_t P0, int64_t n, const int64_t* seg_start,
                                  int64_t* chain_size) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t c = pool_chain[P0 + i];
  if (i == n - 1 || pool_chain[P0 + i + 1] != c) chain_size[c] += i - seg_start[c] + 1;
}

__global__ void record_cols_kernel(const int64_t* offsets, const int32_t* counts, int64_t n, int64_t base,
                                   int64_t* pool_coloff, int32_t* pool_colcnt, int64_t P0) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  pool_coloff[P0 + i] = base + offsets[i];
  pool_colcnt[P0 + i] = counts[i];
}

__global__ void row_order_kernel(const int32_t* pool_chain, const int32_t* pool_local, const int64_t* row_base,
                                 int64_t P, int32_t* row_gid, int32_t* row_chain) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= P) return;
  int32_t c = pool_chain[g];
  int64_t r = row_base[c] + pool_local[g];
  row_gid[r] = (int32_t)g;
  row_chain[r] = c;
}

__global__ void row_counts_kernel(const int32_t* row_gid, const int32_t* pool_colcnt, int64_t R, int32_t* cnt) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < R) cnt[r] = pool_colcnt[row_gid[r]];
}

// Row r of H is the conjugate of column r: <r|H|t> = conj(<t|H|r>).
template <int W>
__global__ void csr_fill_kernel(HashView h, const int32_t* row_gid, const int32_t* row_chain, int64_t R,
                                const int64_t* pool_coloff, const uint64_t* col_keys, const double2* col_vals,
                                const int32_t* pool_local, const int64_t* rowptr, int32_t* col, double2* val,
                                int32_t* err) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  int32_t g = row_gid[r];
  int32_t c = row_chain[r];
  int64_t src = pool_coloff[g];
  int64_t a = rowptr[r], b = rowptr[r + 1];
  for (int64_t e = a; e < b; ++e) {
    const uint64_t* k = col_keys + (size_t)(src + (e - a)) * W;
    uint32_t slot = hash_key((uint32_t)c, k, W) & h.mask;
    int32_t id = HASH_EMPTY;
    for (;;) {
      int32_t cur = h.table[slot];
      if (cur == HASH_EMPTY) break;
      if (!(cur & CAND_FLAG) && h.pool_chain[cur] == c) {
        bool eq = true;
        for (int w = 0; w < W; ++w) eq = eq && h.pool_keys[(size_t)cur * W + w] == k[w];
        if (eq) {
          id = cur;
          break;
        }
      }
      slot = (slot + 1) & h.mask;
    }
    if (id < 0) {
      atomicExch(err, 1);
      col[e] = 0;
      val[e] = make_double2(0.0, 0.0);
      continue;
    }
    col[e] = pool_local[id];
    val[e] = cconj(col_vals[src + (e - a)]);
  }
}

__global__ void keys_by_rows_kernel(int W, const int32_t* row_gid, int64_t R, const uint64_t* pool_keys,
                                    uint64_t* out) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  for (int w = 0; w < W; ++w) out[r * W + w] = pool_keys[(size_t)row_gid[r] * W + w];
}

__global__ void chain_of_rows_kernel(const int64_t* row_base, const int64_t* chain_n, int C, int32_t* row_chain) {
  int c = blockIdx.x;
  if (c >= C) return;
  for (int64_t i = threadIdx.x; i < chain_n[c]; i += blockDim.x) row_chain[row_base[c] + i] = c;
}

// ------------------------------------------------------------------ helpers

static hipStream_t stream() { return g_stream; }

static int bits_for(int64_t n) {
  int b = 1;
  while ((int64_t(1) << b) < n) ++b;
  return b;
}

static void ensure_hash(Batch& b, int64_t need_entries) {
  uint64_t want = 1024;
  while (want < (uint64_t)need_entries * 2) want <<= 1;
  if (want <= b.hcap) return;
  if (want > (1ull << 31)) throw HipError{"hash table too large"};
  b.hcap = (uint32_t)want;
  b.table.reserve((size_t)b.hcap * sizeof(int32_t));
  launch_hash_clear(b.table.as<int32_t>(), b.hcap, stream());
  HashView hv{b.table.as<int32_t>(), b.hcap - 1, b.pool_keys.as<uint64_t>(), b.pool_chain.as<int32_t>()};
  launch_hash_insert_pool(b.W, hv, 0, b.P, stream());
}

static void ensure_pool(Batch& b, int64_t need) {
  if (need <= b.Pcap) return;
  int64_t cap = std::max<int64_t>(need, b.Pcap * 2);
  const size_t W = b.W;
  b.pool_keys.grow_keep(cap * W * 8, b.P * W * 8, stream());
  b.pool_chain.grow_keep(cap * 4, b.P * 4, stream());
  b.pool_local.grow_keep(cap * 4, b.P * 4, stream());
  b.pool_coloff.grow_keep(cap * 8, b.P * 8, stream());
  b.pool_colcnt.grow_keep(cap * 4, b.P * 4, stream());
  b.Pcap = cap;
}

static HashView hash_view(Batch& b) {
  return HashView{b.table.as<int32_t>(), b.hcap - 1, b.pool_keys.as<uint64_t>(), b.pool_chain.as<int32_t>()};
}

// Scratch shared by the build steps.
struct BuildScratch {
  DevBuf tasks, counts, offsets, cand_keys, cand_chain, cand_slot, cand_cnt, alive, perm, rowgid, err, col_chain;
};
static BuildScratch& bscratch() {
  static BuildScratch s;
  return s;
}

static int64_t make_tasks(Batch& b, int64_t begin, int64_t end) {
  BuildScratch& sc = bscratch();
  int64_t cap = (end - begin) / 64 + b.C + 1;
  sc.tasks.reserve(cap * sizeof(WaveTask));
  return build_wave_tasks(b.pool_chain.as<int32_t>(), begin, end, b.chain_model.as<int32_t>(), b.C, sc.tasks.p,
                          cap, stream(), nullptr);
}

// Apply the Hamiltonian to pool states [f0, f1): count, scan, fill.  In
// expand mode new targets become candidates; returns the entry count.
static int64_t apply_layer(Batch& b, int64_t f0, int64_t f1, int mode, double weight_min, int64_t* n_alive) {
  BuildScratch& sc = bscratch();
  const int64_t n = f1 - f0;
  int64_t nt = make_tasks(b, f0, f1);
  ApplyArgs a{};
  a.ops = b.ops->view();
  a.tasks = sc.tasks.as<WaveTask>();
  a.n_tasks = nt;
  a.src_keys = b.pool_keys.as<uint64_t>();
  a.weight_min = weight_min;
  sc.counts.reserve(n * sizeof(int32_t));
  // the wave tasks carry absolute source ids; counts/offsets are indexed the
  // same way through a shifted base pointer
  a.counts = sc.counts.as<int32_t>() - f0;
  a.hash = hash_view(b);
  a.pool_local = b.pool_local.as<int32_t>();
  launch_apply(b.W, mode, false, a, stream());
  sc.offsets.reserve((n + 1) * sizeof(int64_t));
  int64_t E = scan_counts(sc.counts.as<int32_t>(), n, sc.offsets.as<int64_t>(), stream());
  a.offsets = sc.offsets.as<int64_t>() - f0;
  if (mode == MODE_EXPAND) {
    int64_t base = b.E;
    b.col_keys.grow_keep((base + E) * b.W * 8, base * b.W * 8, stream());
    b.col_val.grow_keep((base + E) * 16, base * 16, stream());
    hipLaunchKernelGGL(record_cols_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), sc.offsets.as<int64_t>(),
                       sc.counts.as<int32_t>(), n, base, b.pool_coloff.as<int64_t>(), b.pool_colcnt.as<int32_t>(), f0);
    ensure_hash(b, b.P + E);
    a.hash = hash_view(b);
    sc.cand_keys.reserve(std::max<int64_t>(E, 1) * b.W * 8);
    sc.cand_chain.reserve(std::max<int64_t>(E, 1) * 4);
    sc.cand_slot.reserve(std::max<int64_t>(E, 1) * 4);
    sc.alive.reserve(std::max<int64_t>(E, 1) * 4);
    sc.cand_cnt.reserve(2 * sizeof(int32_t));
    HIP_CHECK(hipMemsetAsync(sc.cand_cnt.p, 0, 2 * sizeof(int32_t), stream()));
    a.out_keys = b.col_keys.as<uint64_t>() + base * b.W;
    a.out_val = b.col_val.as<double2>() + base;
    sc.col_chain.reserve(std::max<int64_t>(E, 1) * 4);
    a.out_chain = sc.col_chain.as<int32_t>();
    a.offsets = sc.offsets.as<int64_t>() - f0;
    a.cand = CandView{sc.cand_keys.as<uint64_t>(), sc.cand_chain.as<int32_t>(), sc.cand_slot.as<int32_t>(),
                      sc.cand_cnt.as<int32_t>(),   sc.alive.as<int32_t>(),      sc.cand_cnt.as<int32_t>() + 1,
                      (int32_t)std::min<int64_t>(E, INT32_MAX)};
    if (E > 0) {
      launch_apply(b.W, mode, true, a, stream());
      launch_discover(b.W, a.out_keys, a.out_chain, E, a.hash, a.cand, stream());
    }
    int32_t cc[2] = {0, 0};
    d2h_bytes(cc, sc.cand_cnt.p, sizeof(cc), stream());
    stream_sync(stream());
    b.E = base + E;
    *n_alive = cc[1];
  } else {
    // fixed basis: write the CSR directly (rows are pool ids = local ids)
    b.rowptr.reserve((n + 1) * sizeof(int64_t));
    HIP_CHECK(hipMemcpyAsync(b.rowptr.p, sc.offsets.p, n * sizeof(int64_t), hipMemcpyDeviceToDevice, stream()));
    HIP_CHECK(hipMemcpyAsync(b.rowptr.as<int64_t>() + n, &E, sizeof(int64_t), hipMemcpyHostToDevice, stream()));
    stream_sync(stream());
    b.col.reserve(std::max<int64_t>(E, 1) * 4);
    b.val.reserve(std::max<int64_t>(E, 1) * 16);
    a.out_col = b.col.as<int32_t>();
    a.out_val = b.val.as<double2>();
    if (E > 0) launch_apply(b.W, mode, true, a, stream());
    stream_sync(stream());
    *n_alive = 0;
  }
  return E;
}

static void finalize_sizes(Batch& b) {
  b.h_n.assign(b.C, 0);
  d2h_bytes(b.h_n.data(), b.chain_size.p, b.C * sizeof(int64_t), stream());
  stream_sync(stream());
  b.h_row_base.assign(b.C + 1, 0);
  for (int c = 0; c < b.C; ++c) b.h_row_base[c + 1] = b.h_row_base[c] + b.h_n[c];
  b.R = b.h_row_base[b.C];
  h2d(b.row_base, b.h_row_base.data(), b.C + 1, stream());
  h2d(b.chain_n, b.h_n.data(), b.C, stream());
}

__global__ void gather_rowptr_kernel(const int64_t* rowptr, const int64_t* row_base, int C, int64_t* out) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c <= C) out[c] = rowptr[row_base[c]];
}

static void compute_chain_nnz(Batch& b) {
  std::vector<int64_t> rp(b.C + 1);
  b.h_nnz.assign(b.C, 0);
  DevBuf tmp;
  tmp.reserve((b.C + 1) * sizeof(int64_t));
  hipLaunchKernelGGL(gather_rowptr_kernel, dim3(blocks_for(b.C + 1, 256)), dim3(256), 0, stream(),
                     b.rowptr.as<int64_t>(), b.row_base.as<int64_t>(), b.C, tmp.as<int64_t>());
  HIP_CHECK(hipGetLastError());
  d2h(rp.data(), tmp.p, b.C + 1, stream());
  stream_sync(stream());
  for (int c = 0; c < b.C; ++c) b.h_nnz[c] = rp[c + 1] - rp[c];
  b.NNZ = rp[b.C];
}

// Start vectors: the seed amplitudes of chain c fill its first local rows.
__global__ void place_seed_amps_kernel(const int64_t* seed_off, const int64_t* row_base, int C, const double2* amps,
                                       double2* v0) {
  int c = blockIdx.x;
  if (c >= C) return;
  const int64_t a = seed_off[c], b = seed_off[c + 1], r = row_base[c];
  for (int64_t i = a + threadIdx.x; i < b; i += blockDim.x) v0[r + (i - a)] = amps[i];
}

static void build(Batch& b, int mode, double weight_min, int max_layers) {
  const hipStream_t s = stream();
  BuildScratch& sc = bscratch();
  const int64_t S = b.h_seed_off.back();
  b.P = 0;
  b.E = 0;
  ensure_pool(b, std::max<int64_t>(S, 1024));
  if (S) HIP_CHECK(hipMemcpyAsync(b.pool_keys.p, b.seed_keys.p, S * b.W * 8, hipMemcpyDeviceToDevice, s));
  DevBuf& seed_off = b.seed_off_dev;
  h2d(seed_off, b.h_seed_off.data(), b.C + 1, s);
  b.chain_size.reserve(b.C * sizeof(int64_t));
  b.seg_start.reserve(b.C * sizeof(int64_t));
  if (b.C) {
    hipLaunchKernelGGL(set_local_seed_kernel, dim3(b.C), dim3(256), 0, s, seed_off.as<int64_t>(), b.C,
                       b.pool_local.as<int32_t>(), b.pool_chain.as<int32_t>(), b.chain_size.as<int64_t>());
    HIP_CHECK(hipGetLastError());
  }
  b.P = S;
  b.hcap = 0;
  ensure_hash(b, std::max<int64_t>(S, 512));

  if (mode == 1) {
    // Fixed basis (get_hamiltonian_matrix): rows are seeds in order.
    int64_t na = 0;
    int64_t E = apply_layer(b, 0, S, MODE_FIXED, weight_min, &na);
    finalize_sizes(b);
    b.NNZ = E;
    b.row_chain.reserve(std::max<int64_t>(b.R, 1) * 4);
    if (b.C)
      hipLaunchKernelGGL(chain_of_rows_kernel, dim3(b.C), dim3(256), 0, s, b.row_base.as<int64_t>(),
                         b.chain_n.as<int64_t>(), b.C, b.row_chain.as<int32_t>());
    compute_chain_nnz(b);
    b.built = true;
    b.row_keys = true;
    b.real = b.ops->real;
    return;
  }

  int64_t f0 = 0, f1 = S;
  int layers = 0;
  while (f1 > f0) {
    if (max_layers > 0 && layers >= max_layers) break;
    ++layers;
    int64_t n_alive = 0;
    apply_layer(b, f0, f1, MODE_EXPAND, weight_min, &n_alive);
    if (n_alive == 0) {
      f0 = f1;
      break;
    }
    // sort the new states by (chain, key) and append them
    sc.perm.reserve(n_alive * sizeof(int32_t));
    HIP_CHECK(hipMemcpyAsync(sc.perm.p, sc.alive.p, n_alive * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    sort_chain_keys(b.W, b.ops->n_orb, bits_for(b.C), sc.cand_keys.as<uint64_t>(), sc.cand_chain.as<int32_t>(),
                    sc.perm.as<int32_t>(), n_alive, s);
    ensure_pool(b, b.P + n_alive);
    const int64_t P0 = b.P;
    hipLaunchKernelGGL(append_layer_kernel, dim3(blocks_for(n_alive, 256)), dim3(256), 0, s, b.W,
                       sc.perm.as<int32_t>(), n_alive, sc.cand_keys.as<uint64_t>(), sc.cand_chain.as<int32_t>(),
                       sc.cand_slot.as<int32_t>(), P0, b.pool_keys.as<uint64_t>(), b.pool_chain.as<int32_t>(),
                       b.table.as<int32_t>());
    hipLaunchKernelGGL(layer_heads_kernel, dim3(blocks_for(n_alive, 256)), dim3(256), 0, s,
                       b.pool_chain.as<int32_t>(), P0, n_alive, b.seg_start.as<int64_t>());
    hipLaunchKernelGGL(layer_local_kernel, dim3(blocks_for(n_alive, 256)), dim3(256), 0, s,
                       b.pool_chain.as<int32_t>(), P0, n_alive, b.seg_start.as<int64_t>(),
                       b.chain_size.as<int64_t>(), b.pool_local.as<int32_t>());
    hipLaunchKernelGGL(layer_grow_kernel, dim3(blocks_for(n_alive, 256)), dim3(256), 0, s,
                       b.pool_chain.as<int32_t>(), P0, n_alive, b.seg_start.as<int64_t>(),
                       b.chain_size.as<int64_t>());
    HIP_CHECK(hipGetLastError());
    b.P = P0 + n_alive;
    f0 = f1;
    f1 = b.P;
  }
  // States of an unfinished last layer keep empty columns.
  if (f1 > f0) {
    HIP_CHECK(hipMemsetAsync(b.pool_colcnt.as<int32_t>() + f0, 0, (f1 - f0) * 4, s));
  }
  finalize_sizes(b);
  // CSR rows in chain-local order
  sc.rowgid.reserve(std::max<int64_t>(b.R, 1) * 4);
  b.row_chain.reserve(std::max<int64_t>(b.R, 1) * 4);
  hipLaunchKernelGGL(row_order_kernel, dim3(blocks_for(b.P, 256)), dim3(256), 0, s, b.pool_chain.as<int32_t>(),
                     b.pool_local.as<int32_t>(), b.row_base.as<int64_t>(), b.P, sc.rowgid.as<int32_t>(),
                     b.row_chain.as<int32_t>());
  sc.counts.reserve(std::max<int64_t>(b.R, 1) * 4);
  hipLaunchKernelGGL(row_counts_kernel, dim3(blocks_for(b.R, 256)), dim3(256), 0, s, sc.rowgid.as<int32_t>(),
                     b.pool_colcnt.as<int32_t>(), b.R, sc.counts.as<int32_t>());
  b.rowptr.reserve((b.R + 1) * sizeof(int64_t));
  int64_t nnz = scan_counts(sc.counts.as<int32_t>(), b.R, b.rowptr.as<int64_t>(), s);
  HIP_CHECK(hipMemcpyAsync(b.rowptr.as<int64_t>() + b.R, &nnz, sizeof(int64_t), hipMemcpyHostToDevice, s));
  stream_sync(s);
  b.col.reserve(std::max<int64_t>(nnz, 1) * 4);
  b.val.reserve(std::max<int64_t>(nnz, 1) * 16);
  sc.err.reserve(4);
  HIP_CHECK(hipMemsetAsync(sc.err.p, 0, 4, s));
  if (b.R) {
    TimedLaunch tl("csr_fill", s);
    if (b.W == 1)
      hipLaunchKernelGGL(csr_fill_kernel<1>, dim3(blocks_for(b.R, 128)), dim3(128), 0, s, hash_view(b),
                         sc.rowgid.as<int32_t>(), b.row_chain.as<int32_t>(), b.R, b.pool_coloff.as<int64_t>(),
                         b.col_keys.as<uint64_t>(), b.col_val.as<double2>(), b.pool_local.as<int32_t>(),
                         b.rowptr.as<int64_t>(), b.col.as<int32_t>(), b.val.as<double2>(), sc.err.as<int32_t>());
    else
      hipLaunchKernelGGL(csr_fill_kernel<2>, dim3(blocks_for(b.R, 128)), dim3(128), 0, s, hash_view(b),
                         sc.rowgid.as<int32_t>(), b.row_chain.as<int32_t>(), b.R, b.pool_coloff.as<int64_t>(),
                         b.col_keys.as<uint64_t>(), b.col_val.as<double2>(), b.pool_local.as<int32_t>(),
                         b.rowptr.as<int64_t>(), b.col.as<int32_t>(), b.val.as<double2>(), sc.err.as<int32_t>());
    HIP_CHECK(hipGetLastError());
  }
  // basis keys in row order (reuse col_keys tail is not safe; keep a copy in seed_keys)
  b.seed_keys.reserve(std::max<int64_t>(b.R, 1) * b.W * 8);
  DevBuf tmp;
  tmp.reserve(std::max<int64_t>(b.R, 1) * b.W * 8);
  hipLaunchKernelGGL(keys_by_rows_kernel, dim3(blocks_for(b.R, 256)), dim3(256), 0, s, b.W, sc.rowgid.as<int32_t>(),
                     b.R, b.pool_keys.as<uint64_t>(), tmp.as<uint64_t>());
  HIP_CHECK(hipMemcpyAsync(b.pool_keys.p, tmp.p, b.R * b.W * 8, hipMemcpyDeviceToDevice, s));
  int32_t err = 0;
  d2h_bytes(&err, sc.err.p, 4, s);
  stream_sync(s);
  if (err) throw HipError{"internal error: CSR target missing from basis"};
  compute_chain_nnz(b);
  b.built = true;
  b.row_keys = true;
  b.real = b.ops->real;
}

// ------------------------------------------------------------------ affine CSR structures
//
// A set of fixed sparsity patterns (structures) whose values are affine in
// a parameter vector: value = base + sum_p theta_p dirs_p, entry by entry.  A
// model family whose Hamiltonian is affine in its parameters has, on a fixed
// sector, a fixed pattern and affine values; a batch of parameter sets then
// needs one fill kernel instead of operator application, sorting and
// hashing.  Structures with no parameters hold fixed matrices (transition
// operators from a ground sector into an excited one).

struct AffineCsr {
  int S = 0, P = 0;
  std::vector<int64_t> h_n, h_nnz, h_row_off, h_nnz_off, h_maxcol;  // per structure (h_maxcol = -1: empty)
  int64_t total_rows = 0, total_nnz = 0;
  DevBuf srowptr, scol, base, dirs, meta;  // meta: n, row_off, nnz_off per structure (int64)
  bool real = true;
};

// Rows, columns and values of every chain of a batch from its structure and
// model: chain c = blockIdx.x, gridDim.y blocks per chain.  meta: row_base [C+1], chain_n [C], nnz_base
// [C+1], chain_struct [C], chain_model [C] (int64).
__global__ __launch_bounds__(256) void affine_fill_kernel(const int64_t* smeta, int S, const int64_t* srowptr,
                                                          const int32_t* scol, const double2* base,
                                                          const double2* dirs, int P, int64_t total_nnz,
                                                          const int64_t* meta, int C, const double* theta,
                                                          double weight_min, int64_t* rowptr, int32_t* col,
                                                          double2* val, int64_t* row_base_out, int64_t* chain_n_out) {
  const int c = blockIdx.x;
  const int64_t* row_base = meta;
  const int64_t* chain_n = meta + C + 1;
  const int64_t* nnz_base = chain_n + C;
  const int64_t* chain_struct = nnz_base + C + 1;
  const int64_t* chain_model = chain_struct + C;
  const int st = (int)chain_struct[c];
  const int64_t n = smeta[st], roff = smeta[S + st], eoff = smeta[2 * S + st];
  const int64_t nnz = srowptr[roff + n];
  const double* th = theta + (size_t)chain_model[c] * P;
  const int64_t first = (int64_t)blockIdx.y * blockDim.x + threadIdx.x, stride = (int64_t)gridDim.y * blockDim.x;
  for (int64_t i = first; i < n; i += stride) rowptr[row_base[c] + i] = nnz_base[c] + srowptr[roff + i];
  for (int64_t k = first; k < nnz; k += stride) {
    const int64_t e = eoff + k;
    double vx = base[e].x, vy = base[e].y;
    for (int p = 0; p < P; ++p) {
      const double2 d = dirs[(size_t)p * total_nnz + e];
      vx += th[p] * d.x;
      vy += th[p] * d.y;
    }
    // the pruning of the Hamiltonian build, |h|^2 < weight_min
    if (vx * vx + vy * vy < weight_min) vx = vy = 0.0;
    col[nnz_base[c] + k] = scol[e];
    val[nnz_base[c] + k] = make_double2(vx, vy);
  }
  if (blockIdx.y == 0 && threadIdx.x == 0) {
    row_base_out[c] = row_base[c];
    chain_n_out[c] = chain_n[c];
    if (c == C - 1) {
      rowptr[row_base[C]] = nnz_base[C];
      row_base_out[C] = row_base[C];
    }
  }
}

// Start vectors of chain c = blockIdx.x (one wave): v = T psi over the
// chain's rows, T = structure chain_struct[c] (rows: the chain's rows;
// columns: rows of ground chain src[c]), psi = eigenvector vec[c] of that
// ground chain with components |psi_i|^2 < psi_wmin dropped (the eigenvector
// dict of eigensystem), entries |v_r|^2 < wmin dropped (applyOp's pruning),
// then normalized; norm2[c] = |v|^2 before normalization.  Each row sums its
// T entries in stored (source) order.
__global__ __launch_bounds__(64) void affine_seed_kernel(const int64_t* smeta, int S, const int64_t* srowptr,
                                                         const int32_t* scol, const double2* tval,
                                                         const int64_t* xmeta, int C, const int64_t* g_row_base,
                                                         const int64_t* g_chain_n, const double2* evecs, int nev,
                                                         double psi_wmin, double wmin, double2* v0,
                                                         double* norm2) {
  const int c = blockIdx.x;
  const int64_t* x_row_base = xmeta;
  const int64_t* src = xmeta + C + 1;
  const int64_t* vec = src + C;
  const int64_t* tstruct = vec + C;
  const int st = (int)tstruct[c];
  const int64_t n = smeta[st], roff = smeta[S + st], eoff = smeta[2 * S + st];
  const int g = (int)src[c];
  const int64_t ng = g_chain_n[g];
  const double2* psi = evecs + (size_t)g_row_base[g] * nev + (size_t)vec[c] * ng;
  double2* v = v0 + x_row_base[c];
  double p = 0.0;
  for (int64_t r = threadIdx.x; r < n; r += 64) {
    double ax = 0.0, ay = 0.0;
    for (int64_t e = srowptr[roff + r]; e < srowptr[roff + r + 1]; ++e) {
      const double2 t = tval[eoff + e];
      const double2 y = psi[scol[eoff + e]];
      if (y.x * y.x + y.y * y.y < psi_wmin) continue;
      ax += t.x * y.x - t.y * y.y;
      ay += t.x * y.y + t.y * y.x;
    }
    if (ax * ax + ay * ay < wmin) ax = ay = 0.0;
    v[r] = make_double2(ax, ay);
    p += ax * ax + ay * ay;
  }
  p = wave_sum(p);
  const double inv = p > 0.0 ? 1.0 / sqrt(p) : 0.0;
  for (int64_t r = threadIdx.x; r < n; r += 64) v[r] = make_double2(v[r].x * inv, v[r].y * inv);
  if (threadIdx.x == 0) norm2[c] = p;
}

__global__ void real_part_kernel(const double2* v, int64_t n, double* out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = v[i].x;
}

// The Krylov view of a batch; for a real CSR with a chain run by the
// grid-wide kernels, the real parts of the values (made once per CSR) ride
// along so those kernels read 8 bytes per value.
KrylovBatch krylov_view(Batch* b) {
  KrylovBatch kb{b->C,
                 b->row_base.as<int64_t>(),
                 b->chain_n.as<int64_t>(),
                 b->rowptr.as<int64_t>(),
                 b->col.as<int32_t>(),
                 b->val.as<double2>(),
                 b->h_n,
                 b->h_nnz};
  bool flat = false;
  for (int c = 0; c < b->C; ++c) flat = flat || b->h_n[c] > g_tuning.flat_min_n;
  if (b->real && flat && g_tuning.real_values) {
    if (!b->have_val_re) {
      b->val_re.reserve(std::max<int64_t>(b->NNZ, 1) * sizeof(double));
      if (b->NNZ > 0) {
        const unsigned nb = (unsigned)std::min<int64_t>((b->NNZ + 255) / 256, 8192);
        hipLaunchKernelGGL(real_part_kernel, dim3(nb), dim3(256), 0, stream(), b->val.as<double2>(), b->NNZ,
                           b->val_re.as<double>());
        HIP_CHECK(hipGetLastError());
      }
      b->have_val_re = true;
    }
    kb.val_re = b->val_re.as<double>();
  }
  return kb;
}

}  // namespace imped

using namespace imped;

#define API_BEGIN try {
#define API_END                                  \
  }                                              \
  catch (const HipError& e) {                    \
    set_error(e.msg);                            \
    return IMPED_ERR_HIP;                        \
  }                                              \
  catch (const std::exception& e) {              \
    set_error(e.what());                         \
    return IMPED_ERR_INVALID;                    \
  }                                              \
  return IMPED_OK;

extern "C" {

const char* imped_last_error(void) { return g_last_error.c_str(); }

int imped_abi_version(void) { return IMPED_ABI_VERSION; }

int imped_set_stream(void* stream) {
  g_stream = reinterpret_cast<hipStream_t>(stream);
  return IMPED_OK;
}

int imped_device_synchronize(void) {
  API_BEGIN
  stream_sync(stream());
  API_END
}

int imped_timing_enable(int on) {
  API_BEGIN
  g_timing = on != 0;
  if (!g_timing) timing_clear();
  API_END
}

int imped_timing_reset(void) {
  API_BEGIN
  stream_sync(stream());
  timing_clear();
  API_END
}

int imped_timing_query(const char* name, double* total_ms, int64_t* launches) {
  API_BEGIN
  stream_sync(stream());
  double t = 0.0;
  int64_t k = 0;
  for (auto& kv : g_timings) {
    if (kv.first != name) continue;
    for (auto& ev : kv.second.events) {
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, ev.first, ev.second));
      t += ms;
      ++k;
    }
  }
  *total_ms = t;
  *launches = k;
  API_END
}

// ------------------------------------------------------------------ restricted basis (host)

namespace {

using Mask = std::array<uint64_t, IMPED_MAXW>;

void set_bit(Mask& m, int i) { m[i / 64] |= 1ull << (63 - i % 64); }

// All k-subsets of orbitals [first, first+n) in itertools.combinations order.
void combinations(int first, int n, int k, std::vector<Mask>& out) {
  out.clear();
  if (k < 0 || k > n) {
    if (k < 0) throw std::invalid_argument("negative occupation in basis enumeration");
    return;
  }
  std::vector<int> idx(k);
  for (int i = 0; i < k; ++i) idx[i] = i;
  for (;;) {
    Mask m{};
    for (int i = 0; i < k; ++i) set_bit(m, first + idx[i]);
    out.push_back(m);
    int i = k - 1;
    while (i >= 0 && idx[i] == n - k + i) --i;
    if (i < 0) break;
    ++idx[i];
    for (int j = i + 1; j < k; ++j) idx[j] = idx[j - 1] + 1;
  }
}

}  // namespace

int imped_enumerate_basis(int n_orb, int W, int n_shells, const int32_t* shell, int64_t cap, int64_t* n_out,
                          uint64_t* keys) {
  API_BEGIN
  if (W < 1 || W > IMPED_MAXW || n_orb > 64 * W || n_shells < 1) {
    set_error("imped_enumerate_basis: invalid W / n_orb / n_shells");
    return IMPED_ERR_INVALID;
  }
  std::vector<std::vector<Mask>> per_shell(n_shells);
  std::vector<Mask> ci, cv, cc;
  for (int l = 0; l < n_shells; ++l) {
    const int32_t* p = shell + 10 * l;
    const int imp_first = p[0], n_imp = p[1], val_first = p[2], n_val = p[3], con_first = p[4], n_con = p[5];
    const int n0 = p[6], dn_tol = p[7], dn_val_max = p[8], dn_con_max = p[9];
    for (int dv = 0; dv <= dn_val_max; ++dv)
      for (int dc = 0; dc <= dn_con_max; ++dc) {
        const int dn = dv - dc;
        if (std::abs(dn) > dn_tol) continue;
        const int nimp = n0 + dn, nval = n_val - dv, ncon = dc;
        if (nimp > n_imp || ncon > n_con) throw std::invalid_argument("over-occupied shell in basis enumeration");
        combinations(imp_first, n_imp, nimp, ci);
        if (n_val == 0) cv.assign(1, Mask{});
        else combinations(val_first, n_val, nval, cv);
        if (n_con == 0) cc.assign(1, Mask{});
        else combinations(con_first, n_con, ncon, cc);
        for (auto& a : ci)
          for (auto& b : cv)
            for (auto& c : cc) {
              Mask m;
              for (int w = 0; w < IMPED_MAXW; ++w) m[w] = a[w] | b[w] | c[w];
              per_shell[l].push_back(m);
            }
      }
  }
  int64_t total = 1;
  for (auto& v : per_shell) total *= (int64_t)v.size();
  *n_out = total;
  if (!keys) return IMPED_OK;
  if (total > cap) {
    set_error("imped_enumerate_basis: output capacity too small");
    return IMPED_ERR_CAPACITY;
  }
  // itertools.product order: the first shell varies slowest
  std::vector<size_t> idx(n_shells, 0);
  for (int64_t k = 0; k < total; ++k) {
    Mask m{};
    for (int l = 0; l < n_shells; ++l)
      for (int w = 0; w < IMPED_MAXW; ++w) m[w] |= per_shell[l][idx[l]][w];
    for (int w = 0; w < W; ++w) keys[k * W + w] = m[w];
    for (int l = n_shells - 1; l >= 0; --l) {
      if (++idx[l] < per_shell[l].size()) break;
      idx[l] = 0;
    }
  }
  API_END
}

int imped_sort_u64(const uint64_t* keys, int64_t n, int begin_bit, int end_bit, int32_t* perm) {
  API_BEGIN
  if (n < 0 || begin_bit < 0 || end_bit > 64 || begin_bit > end_bit || n > INT32_MAX) {
    set_error("imped_sort_u64: invalid arguments");
    return IMPED_ERR_INVALID;
  }
  DevBuf dk, dp;
  h2d(dk, keys, std::max<int64_t>(n, 1), stream());
  h2d(dp, perm, std::max<int64_t>(n, 1), stream());
  sort_u64_perm(dk.as<uint64_t>(), dp.as<int32_t>(), n, begin_bit, end_bit, stream());
  d2h(perm, dp.p, n, stream());
  stream_sync(stream());
  API_END
}

int imped_set_tuning(const char* name, int64_t value) {
  API_BEGIN
  const std::string n(name);
  if (n == "lanczos_wave_max_n")
    g_tuning.lanczos_wave_max_n = value;
  else if (n == "lanczos_wg256_max_n")
    g_tuning.lanczos_wg256_max_n = value;
  else if (n == "eigen_bt") {
    if (value != 0 && value != 256 && value != 512 && value != 1024) {
      set_error("imped_set_tuning: eigen_bt must be 0 (automatic), 256, 512 or 1024");
      return IMPED_ERR_INVALID;
    }
    g_tuning.eigen_bt = value;
  } else if (n == "eigen_wave_max_n")
    g_tuning.eigen_wave_max_n = value;
  else if (n == "eigen_dense_max_n")
    g_tuning.eigen_dense_max_n = value;
  else if (n == "lanczos_dense_max_n")
    g_tuning.lanczos_dense_max_n = value;
  else if (n == "eigen_real")
    g_tuning.eigen_real = value != 0;
  else if (n == "eigen_dgks")
    g_tuning.eigen_dgks = value != 0;
  else if (n == "flat_min_n")
    g_tuning.flat_min_n = value;
  else if (n == "spmv_lpr") {
    if (value != 4 && value != 8 && value != 16 && value != 32 && value != 64) {
      set_error("imped_set_tuning: spmv_lpr must be 4, 8, 16, 32 or 64");
      return IMPED_ERR_INVALID;
    }
    g_tuning.spmv_lpr = value;
  } else if (n == "small_sort") {
    g_tuning.small_sort = value != 0;
  } else if (n == "mlp_graph") {
    g_tuning.mlp_graph = value != 0;
  } else if (n == "lanczos_lds") {
    g_tuning.lanczos_lds = value != 0;
  } else if (n == "lanczos_lds_waves") {
    if (value != 1 && value != 2 && value != 4) {
      set_error("imped_set_tuning: lanczos_lds_waves must be 1, 2 or 4");
      return IMPED_ERR_INVALID;
    }
    g_tuning.lanczos_lds_waves = value;
  } else if (n == "eigen_lds_vectors") {
    g_tuning.eigen_lds_vectors = value < 0 ? -1 : value;
  } else if (n == "lanczos_graph_steps") {
    if (value < 0 || value > 256) {
      set_error("imped_set_tuning: lanczos_graph_steps must be in [0, 256]");
      return IMPED_ERR_INVALID;
    }
    g_tuning.lanczos_graph_steps = value;
  } else if (n == "real_values") {
    g_tuning.real_values = value != 0;
  } else if (n == "eigen_check_start") {
    if (value < 0) {
      set_error("imped_set_tuning: eigen_check_start must be >= 0");
      return IMPED_ERR_INVALID;
    }
    g_tuning.eigen_check_start = value;
  } else if (n == "eigen_check_every") {
    if (value < 1) {
      set_error("imped_set_tuning: eigen_check_every must be >= 1");
      return IMPED_ERR_INVALID;
    }
    g_tuning.eigen_check_every = value;
  } else {
    set_error("imped_set_tuning: unknown parameter " + n);
    return IMPED_ERR_INVALID;
  }
  API_END
}

int imped_ops_create(int W, int n_orb, int n_models, const int32_t* model_group_off, const int32_t* group_term_off,
                     const uint64_t* group_flip, const uint64_t* term_masks, const uint32_t* term_meta,
                     const double* term_coef, const int32_t* model_restr_off, const uint64_t* restr_mask,
                     const int32_t* restr_lohi, imped_ops* out) {
  API_BEGIN
  if (W < 1 || W > IMPED_MAXW || n_models < 1 || n_orb < 1 || n_orb > 64 * W) {
    set_error("imped_ops_create: invalid W / n_orb / n_models");
    return IMPED_ERR_INVALID;
  }
  auto o = std::make_unique<Ops>();
  o->W = W;
  o->M = n_models;
  o->n_orb = n_orb;
  const int G = model_group_off[n_models];
  const int T = group_term_off[G];
  const int R = model_restr_off[n_models];
  o->T = T;
  o->real = true;
  for (int t = 0; t < T; ++t)
    if (term_coef[2 * t + 1] != 0.0) o->real = false;
  hipStream_t s = stream();
  h2d(o->model_group_off, model_group_off, n_models + 1, s);
  h2d(o->group_term_off, group_term_off, G + 1, s);
  h2d(o->group_flip, group_flip, (size_t)G * W, s);
  h2d(o->term_masks, term_masks, (size_t)T * 3 * W, s);
  h2d(o->term_meta, term_meta, T, s);
  h2d(o->term_coef, term_coef, (size_t)T * 2, s);
  h2d(o->model_restr_off, model_restr_off, n_models + 1, s);
  h2d(o->restr_mask, restr_mask, (size_t)R * W, s);
  h2d(o->restr_lohi, restr_lohi, (size_t)R * 2, s);
  stream_sync(s);
  *out = reinterpret_cast<imped_ops>(o.release());
  API_END
}

int imped_ops_set_coef(imped_ops ops, const double* term_coef, int64_t n_terms) {
  API_BEGIN
  Ops* o = reinterpret_cast<Ops*>(ops);
  if (!o || n_terms != o->T || (n_terms > 0 && !term_coef)) {
    set_error("imped_ops_set_coef: the term count differs from the operator's");
    return IMPED_ERR_INVALID;
  }
  bool real = true;
  for (int64_t t = 0; t < n_terms; ++t)
    if (term_coef[2 * t + 1] != 0.0) real = false;
  hipStream_t s = stream();
  // stream order: launches already queued finish with the old values
  if (n_terms) HIP_CHECK(hipMemcpyAsync(o->term_coef.p, term_coef, n_terms * 16, hipMemcpyHostToDevice, s));
  stream_sync(s);
  o->real = real;
  API_END
}

// coef[m][t] = base[t] + sum_p theta[m][p] dirs[p][t] for m < M, t < T1.
__global__ void affine_coef_kernel(const double2* base, const double2* dirs, const double* theta, int P, int64_t T1,
                                   int M, double2* coef) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)M * T1) return;
  const int64_t m = i / T1, t = i % T1;
  double ax = 0.0, ay = 0.0;
  for (int p = 0; p < P; ++p) {
    const double th = theta[m * P + p];
    const double2 d = dirs[(int64_t)p * T1 + t];
    ax += th * d.x;
    ay += th * d.y;
  }
  coef[i] = make_double2(base[t].x + ax, base[t].y + ay);
}

int imped_ops_set_affine(imped_ops ops, int n_params, int64_t terms_per_model, const double* base, const double* dirs,
                         int n_models, const double* theta) {
  API_BEGIN
  Ops* o = reinterpret_cast<Ops*>(ops);
  if (!o || n_params < 0 || n_models < 1 || n_models > o->M || terms_per_model * o->M != o->T || !base ||
      (n_params > 0 && (!dirs || !theta))) {
    set_error("imped_ops_set_affine: every model must hold terms_per_model terms and n_models <= the model count");
    return IMPED_ERR_INVALID;
  }
  const int64_t T1 = terms_per_model;
  bool real = true;
  for (int64_t i = 0; i < T1 * (n_params + 1); ++i)
    if ((i < T1 ? base[2 * i + 1] : dirs[2 * (i - T1) + 1]) != 0.0) real = false;
  static DevBuf d_base, d_dirs, d_theta;
  hipStream_t s = stream();
  h2d(d_base, base, (size_t)T1 * 2, s);
  if (n_params > 0) {
    h2d(d_dirs, dirs, (size_t)T1 * 2 * n_params, s);
    h2d(d_theta, theta, (size_t)n_models * n_params, s);
  }
  if (T1 > 0) {
    hipLaunchKernelGGL(affine_coef_kernel, dim3(blocks_for((int64_t)n_models * T1, 256)), dim3(256), 0, s,
                       d_base.as<double2>(), d_dirs.as<double2>(), d_theta.as<double>(), n_params, T1, n_models,
                       o->term_coef.as<double2>());
    HIP_CHECK(hipGetLastError());
  }
  stream_sync(s);  // the host arrays may be released on return
  // models past n_models keep their values: real only if they were real too
  o->real = real && (n_models == o->M || o->real);
  API_END
}

int imped_ops_destroy(imped_ops ops) {
  delete reinterpret_cast<Ops*>(ops);
  return IMPED_OK;
}

int imped_batch_create(imped_ops ops, int n_chains, const int32_t* chain_model, imped_batch* out) {
  API_BEGIN
  Ops* o = reinterpret_cast<Ops*>(ops);
  if (!o || n_chains < 0) {
    set_error("imped_batch_create: invalid arguments");
    return IMPED_ERR_INVALID;
  }
  for (int c = 0; c < n_chains; ++c) {
    if (chain_model[c] < 0 || chain_model[c] >= o->M) {
      set_error("imped_batch_create: chain model out of range");
      return IMPED_ERR_INVALID;
    }
  }
  auto b = std::make_unique<Batch>();
  b->ops = o;
  b->C = n_chains;
  b->W = o->W;
  b->h_chain_model.assign(chain_model, chain_model + n_chains);  // sent to the device by the first build
  b->h_seed_off.assign(n_chains + 1, 0);
  *out = reinterpret_cast<imped_batch>(b.release());
  API_END
}

int imped_batch_destroy(imped_batch batch) {
  delete reinterpret_cast<Batch*>(batch);
  return IMPED_OK;
}

int imped_batch_set_seeds(imped_batch batch, const int64_t* seed_off, const uint64_t* seed_keys,
                          const double* seed_amps) {
  API_BEGIN
  Batch* b = reinterpret_cast<Batch*>(batch);
  for (int c = 0; c < b->C; ++c)
    if (seed_off[c + 1] < seed_off[c] || seed_off[0] != 0) {
      set_error("imped_batch_set_seeds: offsets must start at 0 and be non-decreasing");
      return IMPED_ERR_INVALID;
    }
  b->h_seed_off.assign(seed_off, seed_off + b->C + 1);
  // device offsets are needed by every seed consumer, built or not (set_csr)
  h2d(b->seed_off_dev, b->h_seed_off.data(), b->C + 1, stream());
  const int64_t S = b->h_seed_off.back();
  h2d(b->seed_keys, seed_keys, std::max<int64_t>(S, 1) * b->W, stream());
  b->have_seed_amps = seed_amps != nullptr;
  if (seed_amps) h2d(b->seed_amps, seed_amps, std::max<int64_t>(S, 1) * 2, stream());
  b->built = false;
  API_END
}

int imped_batch_seed_from_vec(imped_batch batch, imped_vec vec) {
  API_BEGIN
  Batch* b = reinterpret_cast<Batch*>(batch);
  Vec* v = reinterpret_cast<Vec*>(vec);
  if (v->V != b->C || v->W != b->W) {
    set_error("imped_batch_seed_from_vec: vector count or word count mismatch");
    return IMPED_ERR_INVALID;
  }
  b->h_seed_off = v->h_off;
  h2d(b->seed_off_dev, b->h_seed_off.data(), b->C + 1, stream());
  const int64_t S = v->total();
  b->seed_keys.reserve(std::max<int64_t>(S, 1) * b->W * 8);
  b->seed_amps.reserve(std::max<int64_t>(S, 1) * 16);
  if (S) {
    HIP_CHECK(hipMemcpyAsync(b->seed_keys.p, v->keys.p, S * b->W * 8, hipMemcpyDeviceToDevice, stream()));
    HIP_CHECK(hipMemcpyAsync(b->seed_amps.p, v->amps.p, S * 16, hipMemcpyDeviceToDevice, stream()));
  }
  b->have_seed_amps = true;
  b->built = false;
  API_END
}

int imped_batch_build(imped_batch batch, int fixed_basis, double weight_min, int max_layers) {
  API_BEGIN
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (!b->chain_model_dev) {
    h2d(b->chain_model, b->h_chain_model.data(), std::max<size_t>(b->h_chain_model.size(), 1), stream());
    b->chain_model_dev = true;
  }
  b->have_v0dev = false;
  b->have_norm2dev = false;
  b->have_val_re = false;
  build(*b, fixed_basis ? 1 : 0, weight_min, max_layers);
  API_END
}

int imped_batch_set_csr(imped_batch batch, const int64_t* chain_n, const int64_t* rowptr, const int32_t* col,
                        const double* val) {
  API_BEGIN
  Batch* b = reinterpret_cast<Batch*>(batch);
  b->have_v0dev = false;
  b->have_norm2dev = false;
  b->have_val_re = false;
  b->h_n.assign(chain_n, chain_n + b->C);
  b->h_row_base.assign(b->C + 1, 0);
  for (int c = 0; c < b->C; ++c) {
    if (chain_n[c] < 0) {
      set_error("imped_batch_set_csr: negative chain size");
      return IMPED_ERR_INVALID;
    }
    b->h_row_base[c + 1] = b->h_row_base[c] + chain_n[c];
  }
  b->R = b->h_row_base[b->C];
  b->NNZ = rowptr[b->R];
  if (rowptr[0] != 0 || b->NNZ < 0) {
    set_error("imped_batch_set_csr: rowptr must start at 0");
    return IMPED_ERR_INVALID;
  }
  for (int64_t r = 0; r < b->R; ++r)
    if (rowptr[r + 1] < rowptr[r]) {
      set_error("imped_batch_set_csr: rowptr must be non-decreasing");
      return IMPED_ERR_INVALID;
    }
  for (int c = 0; c < b->C; ++c)
    for (int64_t e = rowptr[b->h_row_base[c]]; e < rowptr[b->h_row_base[c + 1]]; ++e)
      if (col[e] < 0 || col[e] >= chain_n[c]) {
        set_error("imped_batch_set_csr: column index out of range");
        return IMPED_ERR_INVALID;
      }
  hipStream_t s = stream();
  h2d(b->rowptr, rowptr, b->R + 1, s);
  h2d(b->col, col, std::max<int64_t>(b->NNZ, 1), s);
  h2d(b->val, val, std::max<int64_t>(b->NNZ, 1) * 2, s);
  h2d(b->row_base, b->h_row_base.data(), b->C + 1, s);
  h2d(b->chain_n, b->h_n.data(), std::max(b->C, 1), s);
  b->h_nnz.assign(b->C, 0);
  for (int c = 0; c < b->C; ++c) b->h_nnz[c] = rowptr[b->h_row_base[c + 1]] - rowptr[b->h_row_base[c]];
  stream_sync(s);
  b->real = true;
  for (int64_t e = 0; e < b->NNZ; ++e)
    if (val[2 * e + 1] != 0.0) b->real = false;
  b->built = true;
  b->row_keys = false;
  API_END
}

int imped_batch_sizes(imped_batch batch, int64_t* chain_n, int64_t* chain_nnz, int64_t* totals) {
  API_BEGIN
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (!b->built) {
    set_error("imped_batch_sizes: batch not built");
    return IMPED_ERR_INVALID;
  }
  if (chain_n) std::copy(b->h_n.begin(), b->h_n.end(), chain_n);
  if (chain_nnz) std::copy(b->h_nnz.begin(), b->h_nnz.end(), chain_nnz);
  if (totals) {
    totals[0] = b->R;
    totals[1] = b->NNZ;
  }
  API_END
}

int imped_batch_get_basis(imped_batch batch, uint64_t* keys) {
  API_BEGIN
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (!b->row_keys) {
    set_error("imped_batch_get_basis: the batch has no basis keys (not built by imped_batch_build)");
    return IMPED_ERR_INVALID;
  }
  d2h(keys, b->pool_keys.p, (size_t)b->R * b->W, stream());
  stream_sync(stream());
  API_END
}

int imped_batch_get_csr(imped_batch batch, int64_t* rowptr, int32_t* col, double* val) {
  API_BEGIN
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (rowptr) d2h(rowptr, b->rowptr.p, b->R + 1, stream());
  if (col) d2h(col, b->col.p, b->NNZ, stream());
  if (val) d2h(val, b->val.p, b->NNZ * 2, stream());
  stream_sync(stream());
  API_END
}

/* ------------------------------------------------------------------ affine CSR structures */

int imped_affine_create(int n_struct, const int64_t* struct_n, const int64_t* rowptr, const int32_t* col,
                        const double* base, int n_params, const double* dirs, imped_affine* out) {
  API_BEGIN
  if (n_struct < 1 || !struct_n || !rowptr || !col || !base || n_params < 0 || (n_params > 0 && !dirs) || !out) {
    set_error("imped_affine_create: n_struct >= 1, sizes, rowptr, col, base (and dirs when n_params > 0)");
    return IMPED_ERR_INVALID;
  }
  auto a = std::make_unique<AffineCsr>();
  a->S = n_struct;
  a->P = n_params;
  a->h_n.assign(struct_n, struct_n + n_struct);
  a->h_row_off.assign(n_struct, 0);
  a->h_nnz_off.assign(n_struct, 0);
  a->h_nnz.assign(n_struct, 0);
  a->h_maxcol.assign(n_struct, -1);
  int64_t ro = 0, eo = 0;
  for (int st = 0; st < n_struct; ++st) {
    if (struct_n[st] < 0) {
      set_error("imped_affine_create: negative structure size");
      return IMPED_ERR_INVALID;
    }
    a->h_row_off[st] = ro;
    a->h_nnz_off[st] = eo;
    const int64_t* rp = rowptr + ro;
    if (rp[0] != 0) {
      set_error("imped_affine_create: every structure's rowptr starts at 0");
      return IMPED_ERR_INVALID;
    }
    for (int64_t i = 0; i < struct_n[st]; ++i)
      if (rp[i + 1] < rp[i]) {
        set_error("imped_affine_create: rowptr must be non-decreasing");
        return IMPED_ERR_INVALID;
      }
    a->h_nnz[st] = rp[struct_n[st]];
    ro += struct_n[st] + 1;
    eo += a->h_nnz[st];
  }
  a->total_rows = ro;
  a->total_nnz = eo;
  for (int st = 0; st < n_struct; ++st)
    for (int64_t e = a->h_nnz_off[st]; e < a->h_nnz_off[st] + a->h_nnz[st]; ++e) {
      if (col[e] < 0) {
        set_error("imped_affine_create: negative column index");
        return IMPED_ERR_INVALID;
      }
      a->h_maxcol[st] = std::max<int64_t>(a->h_maxcol[st], col[e]);
    }
  for (int64_t e = 0; e < eo && a->real; ++e) a->real = base[2 * e + 1] == 0.0;
  for (int64_t e = 0; e < eo * n_params && a->real; ++e) a->real = dirs[2 * e + 1] == 0.0;
  std::vector<int64_t> sm(3 * (size_t)n_struct);
  for (int st = 0; st < n_struct; ++st) {
    sm[st] = a->h_n[st];
    sm[n_struct + st] = a->h_row_off[st];
    sm[2 * n_struct + st] = a->h_nnz_off[st];
  }
  hipStream_t s = stream();
  h2d(a->meta, sm.data(), sm.size(), s);
  h2d(a->srowptr, rowptr, ro, s);
  h2d(a->scol, col, std::max<int64_t>(eo, 1), s);
  h2d(a->base, base, std::max<int64_t>(eo, 1) * 2, s);
  if (n_params > 0) h2d(a->dirs, dirs, std::max<int64_t>(eo * n_params, 1) * 2, s);
  stream_sync(s);
  *out = reinterpret_cast<imped_affine>(a.release());
  API_END
}

int imped_affine_destroy(imped_affine aff) {
  delete reinterpret_cast<AffineCsr*>(aff);
  return IMPED_OK;
}

int imped_batch_set_affine(imped_batch batch, imped_affine aff, const int32_t* chain_struct,
                           const int32_t* chain_model, int n_models, const double* theta, double weight_min) {
  API_BEGIN
  Batch* b = reinterpret_cast<Batch*>(batch);
  const AffineCsr* a = reinterpret_cast<const AffineCsr*>(aff);
  if (!b || !a || (b->C > 0 && (!chain_struct || !chain_model)) || n_models < 0 ||
      (a->P > 0 && n_models > 0 && !theta)) {
    set_error("imped_batch_set_affine: batch, structures, chain_struct and chain_model (and theta) required");
    return IMPED_ERR_INVALID;
  }
  const int C = b->C;
  for (int c = 0; c < C; ++c) {
    const int st = chain_struct[c];
    if (st < 0 || st >= a->S || chain_model[c] < 0 || (a->P > 0 && chain_model[c] >= n_models)) {
      set_error("imped_batch_set_affine: chain structure or model index out of range");
      return IMPED_ERR_INVALID;
    }
    if (a->h_maxcol[st] >= a->h_n[st]) {
      set_error("imped_batch_set_affine: a structure used as a Hamiltonian must be square");
      return IMPED_ERR_INVALID;
    }
  }
  std::vector<int64_t> meta(3 * (size_t)C + 2 + 2 * (size_t)C);
  int64_t* row_base = meta.data();
  int64_t* chain_n = row_base + C + 1;
  int64_t* nnz_base = chain_n + C;
  int64_t* cs = nnz_base + C + 1;
  int64_t* cm = cs + C;
  b->h_n.assign(C, 0);
  b->h_nnz.assign(C, 0);
  b->h_row_base.assign(C + 1, 0);
  int64_t maxw = 1;
  for (int c = 0; c < C; ++c) {
    const int st = chain_struct[c];
    b->h_n[c] = a->h_n[st];
    b->h_nnz[c] = a->h_nnz[st];
    row_base[c + 1] = row_base[c] + a->h_n[st];
    nnz_base[c + 1] = nnz_base[c] + a->h_nnz[st];
    chain_n[c] = a->h_n[st];
    cs[c] = st;
    cm[c] = chain_model[c];
    maxw = std::max(maxw, std::max(a->h_n[st], a->h_nnz[st]));
  }
  for (int c = 0; c <= C; ++c) b->h_row_base[c] = row_base[c];
  b->R = row_base[C];
  b->NNZ = nnz_base[C];
  hipStream_t s = stream();
  // the chain table and theta in one copy, kept with the batch
  HostPack pk;
  const size_t o_meta = pk.add(meta.data(), meta.size());
  const size_t o_th = pk.add(theta, a->P > 0 ? (size_t)n_models * a->P : 0);
  pk.upload(b->fill_meta, s);
  const int64_t* dmeta = HostPack::at<int64_t>(b->fill_meta, o_meta);
  const double* dth = HostPack::at<double>(b->fill_meta, o_th);
  b->rowptr.reserve((b->R + 1) * sizeof(int64_t));
  b->col.reserve(std::max<int64_t>(b->NNZ, 1) * 4);
  b->val.reserve(std::max<int64_t>(b->NNZ, 1) * 16);
  b->row_base.reserve((C + 1) * sizeof(int64_t));
  b->chain_n.reserve(std::max(C, 1) * sizeof(int64_t));
  if (C == 0) {
    const int64_t zero = 0;
    h2d(b->rowptr, &zero, 1, s);
    h2d(b->row_base, &zero, 1, s);
  } else {
    const unsigned gy = (unsigned)std::min<int64_t>((maxw + 255) / 256, 64);
    TimedLaunch tl("affine_fill", s);
    hipLaunchKernelGGL(affine_fill_kernel, dim3(C, gy), dim3(256), 0, s, a->meta.as<int64_t>(), a->S,
                       a->srowptr.as<int64_t>(), a->scol.as<int32_t>(), a->base.as<double2>(),
                       a->P > 0 ? a->dirs.as<double2>() : nullptr, a->P, a->total_nnz, dmeta, C,
                       a->P > 0 ? dth : nullptr, weight_min, b->rowptr.as<int64_t>(),
                       b->col.as<int32_t>(), b->val.as<double2>(), b->row_base.as<int64_t>(),
                       b->chain_n.as<int64_t>());
    HIP_CHECK(hipGetLastError());
  }
  b->built = true;
  b->row_keys = false;
  b->real = a->real;
  b->have_v0dev = false;
  b->have_norm2dev = false;
  b->have_val_re = false;
  b->have_seed_amps = false;
  API_END
}

int imped_batch_seed_affine(imped_batch batch, imped_batch ground, imped_affine trans, const int32_t* src_chain,
                            const int32_t* src_vec, const int32_t* chain_struct, double psi_weight_min,
                            double weight_min, double* norm2) {
  API_BEGIN
  Batch* b = reinterpret_cast<Batch*>(batch);
  const Batch* g = reinterpret_cast<const Batch*>(ground);
  const AffineCsr* a = reinterpret_cast<const AffineCsr*>(trans);
  if (!b || !g || !a || !b->built || !g->built || g->nev < 1 || (b->C > 0 && (!src_chain || !src_vec ||
      !chain_struct))) {
    set_error("imped_batch_seed_affine: built batches (the ground one after imped_batch_eigen), structures and "
              "chain maps required");
    return IMPED_ERR_INVALID;
  }
  const int C = b->C;
  std::vector<int64_t> meta(4 * (size_t)C + 1);
  for (int c = 0; c <= C; ++c) meta[c] = b->h_row_base[c];
  for (int c = 0; c < C; ++c) {
    const int st = chain_struct[c];
    if (src_chain[c] < 0 || src_chain[c] >= g->C || src_vec[c] < 0 || src_vec[c] >= g->nev || st < 0 ||
        st >= a->S || a->h_n[st] != b->h_n[c] || a->h_maxcol[st] >= g->h_n[src_chain[c]]) {
      set_error("imped_batch_seed_affine: chain maps out of range, or a structure's rows (columns) differ from "
                "the chain's (exceed the ground chain's)");
      return IMPED_ERR_INVALID;
    }
    meta[C + 1 + c] = src_chain[c];
    meta[2 * C + 1 + c] = src_vec[c];
    meta[3 * C + 1 + c] = st;
  }
  // structure columns index the ground chain's rows
  hipStream_t s = stream();
  h2d(b->seed_meta, meta.data(), meta.size(), s);
  b->v0dev.reserve(std::max<int64_t>(b->R, 1) * 16);
  b->norm2dev.reserve(std::max(C, 1) * sizeof(double));
  if (C > 0) {
    TimedLaunch tl("affine_seed", s);
    hipLaunchKernelGGL(affine_seed_kernel, dim3(C), dim3(64), 0, s, a->meta.as<int64_t>(), a->S,
                       a->srowptr.as<int64_t>(), a->scol.as<int32_t>(), a->base.as<double2>(),
                       b->seed_meta.as<int64_t>(), C, g->row_base.as<int64_t>(), g->chain_n.as<int64_t>(),
                       g->evecs.as<double2>(), g->nev, psi_weight_min, weight_min, b->v0dev.as<double2>(),
                       b->norm2dev.as<double>());
    HIP_CHECK(hipGetLastError());
    if (norm2) d2h(norm2, b->norm2dev.p, C, s);
  }
  // without norm2 nothing is read back and nothing waits
  if (norm2) stream_sync(s);
  b->have_v0dev = true;
  b->have_norm2dev = true;
  API_END
}

/* ------------------------------------------------------------------ family XAS in one call */

int imped_family_xas(imped_affine ground, imped_affine sector, imped_affine trans, int n_comp, int n_pol,
                     const int32_t* sector_index, int n_models, const double* theta, const double* kT, int nev,
                     int n_psi_max, double energy_cut, int krylov_size, int reorth, double eig_tol,
                     int eig_max_steps, uint64_t seed, double h_weight_min, double psi_weight_min,
                     double weight_min, const double* omega, int n_w, double delta, double* spectra,
                     double* energies, int32_t* n_kept, int32_t* eig_steps, int32_t* chain_struct, int32_t* kout,
                     int32_t* info) {
  API_BEGIN
  const AffineCsr* ag = reinterpret_cast<const AffineCsr*>(ground);
  const AffineCsr* as = reinterpret_cast<const AffineCsr*>(sector);
  const AffineCsr* at = reinterpret_cast<const AffineCsr*>(trans);
  if (!ag || !as || !at || n_comp < 1 || n_comp > ag->S || n_pol < 1 || !sector_index || n_models < 0 ||
      (n_models > 0 && (!theta || !kT || !spectra || !energies || !n_kept)) || nev < 1 || n_psi_max < 1 ||
      krylov_size < 1 || n_w < 0 || (n_w > 0 && !omega)) {
    set_error("imped_family_xas: structures, n_comp <= ground structures, sector_index, theta, kT, outputs, "
              "nev >= 1, n_psi_max >= 1 and krylov_size >= 1 required");
    return IMPED_ERR_INVALID;
  }
  for (int i = 0; i < n_comp * n_pol; ++i)
    if (sector_index[i] < 0 || sector_index[i] >= as->S || sector_index[i] >= at->S) {
      set_error("imped_family_xas: sector index out of range");
      return IMPED_ERR_INVALID;
    }
  for (int m = 0; m < n_models; ++m)
    if (!(kT[m] > 0.0)) {
      set_error("imped_family_xas: kT must be positive");
      return IMPED_ERR_INVALID;
    }
  if (n_models == 0) return IMPED_OK;
  auto make_batch = [](int C) {
    auto b = std::make_unique<Batch>();
    b->C = C;
    b->h_seed_off.assign(C + 1, 0);
    return b;
  };
  // ground chains (model, component)
  const int G = n_models * n_comp;
  auto gb = make_batch(G);
  std::vector<int32_t> g_struct(G), g_model(G);
  for (int c = 0; c < G; ++c) {
    g_struct[c] = c % n_comp;
    g_model[c] = c / n_comp;
  }
  int rc = imped_batch_set_affine(reinterpret_cast<imped_batch>(gb.get()), ground, g_struct.data(), g_model.data(),
                                  n_models, theta, h_weight_min);
  if (rc != IMPED_OK) return rc;
  std::vector<double> ev((size_t)G * nev);
  std::vector<int32_t> st(G);
  rc = imped_batch_eigen(reinterpret_cast<imped_batch>(gb.get()), nev, eig_tol, eig_max_steps, seed, ev.data(),
                         nullptr, st.data());
  if (rc != IMPED_OK) return rc;
  if (eig_steps) std::copy(st.begin(), st.end(), eig_steps);
  // per model: the lowest n_psi_max of the components' eigenvalues (stable
  // in component-major order), kept within energy_cut kT of the lowest
  std::vector<int32_t> x_struct, x_model, x_src, x_vec, x_sign, x_z, x_spec;
  std::vector<double> x_e;
  std::vector<int> ord(n_comp * nev);
  for (int m = 0; m < n_models; ++m) {
    const double* e = ev.data() + (size_t)m * n_comp * nev;
    for (int i = 0; i < n_comp * nev; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return e[x] < e[y]; });
    const int top = std::min(n_psi_max, n_comp * nev);
    int kept = 0;
    while (kept < top && e[ord[kept]] - e[ord[0]] < energy_cut * kT[m]) ++kept;
    n_kept[m] = kept;
    for (int i = 0; i < n_psi_max; ++i) energies[(size_t)m * n_psi_max + i] = i < kept ? e[ord[i]] : NAN;
    for (int t = 0; t < n_pol; ++t)
      for (int i = 0; i < kept; ++i) {
        const int k = ord[i] / nev, v = ord[i] % nev;
        x_struct.push_back(sector_index[k * n_pol + t]);
        x_model.push_back(m);
        x_src.push_back(m * n_comp + k);
        x_vec.push_back(v);
        x_e.push_back(e[ord[i]]);
        x_sign.push_back(1);
        x_z.push_back(t == 0);
        x_spec.push_back(m);
      }
  }
  const int X = (int)x_struct.size();
  auto xb = make_batch(X);
  rc = imped_batch_set_affine(reinterpret_cast<imped_batch>(xb.get()), sector, x_struct.data(), x_model.data(),
                              n_models, theta, h_weight_min);
  if (rc != IMPED_OK) return rc;
  rc = imped_batch_seed_affine(reinterpret_cast<imped_batch>(xb.get()), reinterpret_cast<imped_batch>(gb.get()),
                               trans, x_src.data(), x_vec.data(), x_struct.data(), psi_weight_min, weight_min, nullptr);
  if (rc != IMPED_OK) return rc;
  int64_t nmax = 1;
  for (int c = 0; c < X; ++c) nmax = std::max(nmax, xb->h_n[c]);
  const int K = (int)std::min<int64_t>(krylov_size, nmax);
  rc = imped_batch_lanczos(reinterpret_cast<imped_batch>(xb.get()), nullptr, K, reorth, nullptr, nullptr, nullptr,
                           nullptr);
  if (rc != IMPED_OK) return rc;
  // kout and status come back with the spectra (one wait, in run_cf)
  std::vector<int32_t> ks(2 * (size_t)X);
  d2h(ks.data(), xb->kout.p, ks.size(), stream());
  run_cf(X, K, xb->alpha.as<double>(), xb->beta.as<double>(), xb->kout.as<int32_t>(), x_e.data(), nullptr,
         x_sign.data(), x_z.data(), kT, x_spec.data(), n_models, omega, n_w, delta, nullptr, spectra, stream(),
         xb->norm2dev.as<double>());
  if (chain_struct) std::copy(x_struct.begin(), x_struct.end(), chain_struct);
  if (kout) std::copy(ks.begin(), ks.begin() + X, kout);
  if (info) {
    info[0] = X;
    info[1] = K;
    info[2] = info[3] = 0;
  }
  for (int c = 0; c < X; ++c)
    if (ks[X + c] != IMPED_OK) {
      set_error("imped_family_xas: Lanczos breakdown (beta == 0) in an excited chain");
      return IMPED_ERR_BREAKDOWN;
    }
  API_END
}

int imped_batch_lanczos(imped_batch batch, const double* v0_host, int K, int reorth, double* alpha, double* beta,
                        int32_t* kout, int32_t* status) {
  API_BEGIN
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (!b->built || K < 1) {
    set_error("imped_batch_lanczos: batch not built or K < 1");
    return IMPED_ERR_INVALID;
  }
  DevBuf v0;
  const double2* v0d = nullptr;
  if (v0_host) {
    h2d(v0, v0_host, std::max<int64_t>(b->R, 1) * 2, stream());
    v0d = v0.as<double2>();
  } else if (b->have_v0dev) {
    v0d = b->v0dev.as<double2>();
  } else {
    if (!b->have_seed_amps || !seeds_fit(*b)) {
      set_error("imped_batch_lanczos: no start vector (or more seeds than rows in a chain)");
      return IMPED_ERR_INVALID;
    }
    // seed amplitudes occupy the first local rows of every chain
    v0.reserve(std::max<int64_t>(b->R, 1) * 16);
    HIP_CHECK(hipMemsetAsync(v0.p, 0, std::max<int64_t>(b->R, 1) * 16, stream()));
    if (b->C) {
      hipLaunchKernelGGL(place_seed_amps_kernel, dim3(b->C), dim3(64), 0, stream(), b->seed_off_dev.as<int64_t>(),
                         b->row_base.as<int64_t>(), b->C, b->seed_amps.as<double2>(), v0.as<double2>());
      HIP_CHECK(hipGetLastError());
    }
    v0d = v0.as<double2>();
  }
  b->K = K;
  b->alpha.reserve((size_t)std::max(b->C, 1) * K * 8);
  b->beta.reserve((size_t)std::max(b->C, 1) * K * 8);
  b->kout.reserve((size_t)std::max(b->C, 1) * 8);  // kout[C], then status[C]
  int32_t* status_d = b->kout.as<int32_t>() + std::max(b->C, 1);
  const KrylovBatch kb = krylov_view(b);
  run_lanczos(kb, v0d, K, reorth, b->alpha.as<double>(), b->beta.as<double>(), b->kout.as<int32_t>(), status_d,
              stream());
  if (alpha) d2h(alpha, b->alpha.p, (size_t)b->C * K, stream());
  if (beta) d2h(beta, b->beta.p, (size_t)b->C * K, stream());
  if (!alpha && !beta && !kout && !status) return IMPED_OK;  // results stay on the device; nothing waits
  std::vector<int32_t> ks;
  if (kout && status && b->C) {
    ks.resize(2 * (size_t)b->C);
    d2h(ks.data(), b->kout.p, ks.size(), stream());  // kout and status in one read
  } else {
    if (kout) d2h(kout, b->kout.p, b->C, stream());
    if (status) d2h(status, status_d, b->C, stream());
  }
  stream_sync(stream());
  if (!ks.empty()) {
    std::copy(ks.begin(), ks.begin() + b->C, kout);
    std::copy(ks.begin() + b->C, ks.end(), status);
  }
  API_END
}

int imped_batch_eigen(imped_batch batch, int nev, double tol, int max_steps, uint64_t seed, double* evals,
                      double* evecs_host, int32_t* steps_out) {
  API_BEGIN
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (!b->built || nev < 1) {
    set_error("imped_batch_eigen: batch not built or nev < 1");
    return IMPED_ERR_INVALID;
  }
  b->nev = nev;
  // evals[C][nev], then steps[C] and conv[C] (int32): one read brings all back
  const int C1 = std::max(b->C, 1);
  const size_t ev_bytes = (size_t)C1 * nev * 8;
  b->evals.reserve(ev_bytes + (size_t)C1 * 8);
  int32_t* steps_d = reinterpret_cast<int32_t*>(static_cast<char*>(b->evals.p) + ev_bytes);
  int32_t* conv_d = steps_d + C1;
  b->evecs.reserve(std::max<int64_t>(b->R, 1) * nev * 16);
  const KrylovBatch kb = krylov_view(b);
  int rc = run_eigen(kb, nev, tol, max_steps, seed, b->evals.as<double>(), b->evecs.as<double2>(), steps_d, conv_d,
                     b->real, stream());
  if (rc != IMPED_OK) return rc;
  std::vector<char> es(ev_bytes + (size_t)C1 * 8);
  if (b->C) d2h(es.data(), b->evals.p, es.size(), stream());
  if (evecs_host) d2h(evecs_host, b->evecs.p, (size_t)b->R * nev * 2, stream());
  stream_sync(stream());
  if (evals) std::memcpy(evals, es.data(), (size_t)b->C * nev * 8);
  if (steps_out) std::memcpy(steps_out, es.data() + ev_bytes, (size_t)b->C * 4);
  const int32_t* conv = reinterpret_cast<const int32_t*>(es.data() + ev_bytes + (size_t)C1 * 4);
  for (int c = 0; c < b->C; ++c)
    if (!conv[c]) rc = IMPED_ERR_NOCONV;
  if (rc != IMPED_OK) {
    set_error("imped_batch_eigen: not converged within max_steps");
    return rc;
  }
  API_END
}

int imped_batch_solve(imped_batch batch, int n_shift, const double* shifts, double tol, int max_steps,
                      int32_t* steps_out, double* resid_out, double* x_out) {
  API_BEGIN
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (!b->built || n_shift < 1) {
    set_error("imped_batch_solve: batch not built or no shifts");
    return IMPED_ERR_INVALID;
  }
  if (!b->have_seed_amps || !seeds_fit(*b)) {
    set_error("imped_batch_solve: no right-hand side (or more seeds than rows in a chain)");
    return IMPED_ERR_INVALID;
  }
  DevBuf y;
  y.reserve(std::max<int64_t>(b->R, 1) * 16);
  HIP_CHECK(hipMemsetAsync(y.p, 0, std::max<int64_t>(b->R, 1) * 16, stream()));
  if (b->C) {
    hipLaunchKernelGGL(place_seed_amps_kernel, dim3(b->C), dim3(64), 0, stream(), b->seed_off_dev.as<int64_t>(),
                       b->row_base.as<int64_t>(), b->C, b->seed_amps.as<double2>(), y.as<double2>());
    HIP_CHECK(hipGetLastError());
  }
  b->nev = 0;
  b->evecs.reserve(std::max<int64_t>(b->R, 1) * n_shift * 16);
  KrylovBatch kb{b->C,
                 b->row_base.as<int64_t>(),
                 b->chain_n.as<int64_t>(),
                 b->rowptr.as<int64_t>(),
                 b->col.as<int32_t>(),
                 b->val.as<double2>(),
                 b->h_n,
                 b->h_nnz};
  const int rc = run_solve(kb, y.as<double2>(), n_shift, shifts, tol, max_steps, b->evecs.as<double2>(), steps_out,
                           resid_out, stream());
  if (rc == IMPED_ERR_INVALID) return rc;
  b->nev = n_shift;
  if (x_out) {
    d2h(x_out, b->evecs.p, (size_t)b->R * n_shift * 2, stream());
    stream_sync(stream());
  }
  if (rc != IMPED_OK) {
    set_error("imped_batch_solve: residual above tolerance after max_steps");
    return rc;
  }
  API_END
}

int imped_batch_cf(imped_batch batch, const double* energy, const double* norm2, const int32_t* sign,
                   const int32_t* zflag, const double* kT, const int32_t* spec_id, int n_spec, const double* omega, int n_w,
                   double delta, double* g_out, double* spec_out) {
  API_BEGIN
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (b->K < 1) {
    set_error("imped_batch_cf: run imped_batch_lanczos first");
    return IMPED_ERR_INVALID;
  }
  for (int q = 0; q < n_spec; ++q)
    if (!(kT[q] > 0.0)) {
      set_error("imped_batch_cf: kT must be positive");
      return IMPED_ERR_INVALID;
    }
  if (!norm2 && !b->have_norm2dev) {
    set_error("imped_batch_cf: norm2 is NULL and the batch has no device norms (imped_batch_seed_affine)");
    return IMPED_ERR_INVALID;
  }
  run_cf(b->C, b->K, b->alpha.as<double>(), b->beta.as<double>(), b->kout.as<int32_t>(), energy, norm2, sign, zflag,
         kT, spec_id, n_spec, omega, n_w, delta, g_out, spec_out, stream(), b->norm2dev.as<double>());
  API_END
}

// ------------------------------------------------------------------ sparse vectors

int imped_vec_create(int W, int n_vec, const int64_t* off, const uint64_t* keys, const double* amps, imped_vec* out) {
  API_BEGIN
  auto v = std::make_unique<Vec>();
  v->W = W;
  v->V = n_vec;
  v->h_off.assign(off, off + n_vec + 1);
  const int64_t n = v->total();
  h2d(v->keys, keys, std::max<int64_t>(n, 1) * W, stream());
  h2d(v->amps, amps, std::max<int64_t>(n, 1) * 2, stream());
  stream_sync(stream());
  *out = reinterpret_cast<imped_vec>(v.release());
  API_END
}

int imped_vec_destroy(imped_vec vec) {
  delete reinterpret_cast<Vec*>(vec);
  return IMPED_OK;
}

int imped_vec_offsets(imped_vec vec, int64_t* off) {
  Vec* v = reinterpret_cast<Vec*>(vec);
  std::copy(v->h_off.begin(), v->h_off.end(), off);
  return IMPED_OK;
}

int imped_vec_get(imped_vec vec, uint64_t* keys, double* amps, double* norm2) {
  API_BEGIN
  Vec* v = reinterpret_cast<Vec*>(vec);
  const int64_t n = v->total();
  if (keys) d2h(keys, v->keys.p, (size_t)n * v->W, stream());
  if (amps) d2h(amps, v->amps.p, (size_t)n * 2, stream());
  if (norm2 && v->norm2.p) d2h(norm2, v->norm2.p, v->V, stream());
  stream_sync(stream());
  API_END
}

int imped_vec_apply(imped_ops ops, imped_vec src, int n_out, const int32_t* out_src, const int32_t* out_model,
                    double weight_min, int normalize, imped_vec* out) {
  API_BEGIN
  Ops* o = reinterpret_cast<Ops*>(ops);
  Vec* v = reinterpret_cast<Vec*>(src);
  if (v->W != o->W) {
    set_error("imped_vec_apply: word count mismatch");
    return IMPED_ERR_INVALID;
  }
  for (int i = 0; i < n_out; ++i)
    if (out_src[i] < 0 || out_src[i] >= v->V || out_model[i] < 0 || out_model[i] >= o->M) {
      set_error("imped_vec_apply: output vector mapping out of range");
      return IMPED_ERR_INVALID;
    }
  auto r = std::make_unique<Vec>();
  r->W = o->W;
  r->V = n_out;
  vector_apply(o->view(), v->W, v->h_off, v->keys.as<uint64_t>(), v->amps.as<double2>(), n_out, out_src,
               out_model, weight_min, normalize != 0, r->h_off, r->keys, r->amps, r->norm2, stream(), o->n_orb);
  *out = reinterpret_cast<imped_vec>(r.release());
  API_END
}

int imped_vec_prune(imped_vec src, double weight_min, imped_vec* out) {
  API_BEGIN
  Vec* v = reinterpret_cast<Vec*>(src);
  auto r = std::make_unique<Vec>();
  r->W = v->W;
  r->V = v->V;
  prune_vec(v->W, v->h_off, v->keys.as<uint64_t>(), v->amps.as<double2>(), weight_min, r->h_off, r->keys, r->amps,
            stream());
  if (v->norm2.p && v->V) {
    r->norm2.reserve((size_t)v->V * 8);
    HIP_CHECK(hipMemcpyAsync(r->norm2.p, v->norm2.p, (size_t)v->V * 8, hipMemcpyDeviceToDevice, stream()));
    stream_sync(stream());
  }
  *out = reinterpret_cast<imped_vec>(r.release());
  API_END
}

int imped_vec_from_eigen(imped_batch batch, int n_out, const int32_t* chain, const int32_t* which, double weight_min,
                         imped_vec* out) {
  API_BEGIN
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (b->nev < 1) {
    set_error("imped_vec_from_eigen: run imped_batch_eigen or imped_batch_solve first");
    return IMPED_ERR_INVALID;
  }
  if (!b->row_keys) {
    set_error("imped_vec_from_eigen: the batch has no basis keys (not built by imped_batch_build)");
    return IMPED_ERR_INVALID;
  }
  for (int i = 0; i < n_out; ++i)
    if (chain[i] < 0 || chain[i] >= b->C || which[i] < 0 || which[i] >= b->nev) {
      set_error("imped_vec_from_eigen: index out of range");
      return IMPED_ERR_INVALID;
    }
  auto r = std::make_unique<Vec>();
  r->W = b->W;
  r->V = n_out;
  truncate_eigvecs(b->W, b->pool_keys.as<uint64_t>(), b->evecs.as<double2>(), b->nev, b->h_row_base, b->h_n, n_out,
                   chain, which, weight_min, r->h_off, r->keys, r->amps, stream());
  *out = reinterpret_cast<imped_vec>(r.release());
  API_END
}

// ------------------------------------------------------------------ surrogate network

int imped_mlp_create(int n_members, int n_layers, const int32_t* dims, int max_batch, imped_mlp* out) {
  API_BEGIN
  if (n_members < 1 || n_members > 64 || n_layers < 1 || n_layers > 16 || max_batch < 1) {
    set_error("imped_mlp_create: need 1..64 members, 1..16 layers and max_batch >= 1");
    return IMPED_ERR_INVALID;
  }
  for (int l = 0; l <= n_layers; ++l)
    if (dims[l] < 1 || dims[l] > 65536) {
      set_error("imped_mlp_create: layer widths must be in [1, 65536]");
      return IMPED_ERR_INVALID;
    }
  *out = reinterpret_cast<imped_mlp>(mlp_create(n_members, n_layers, dims, max_batch));
  API_END
}

int imped_mlp_destroy(imped_mlp net) {
  mlp_destroy(reinterpret_cast<Mlp*>(net));
  return IMPED_OK;
}

int imped_mlp_param_count(imped_mlp net, int64_t* n) {
  *n = mlp_param_count(*reinterpret_cast<Mlp*>(net));
  return IMPED_OK;
}

int imped_mlp_set_params(imped_mlp net, const float* params) {
  API_BEGIN
  mlp_set_params(*reinterpret_cast<Mlp*>(net), params, stream());
  API_END
}

int imped_mlp_get_params(imped_mlp net, float* params) {
  API_BEGIN
  mlp_get_params(*reinterpret_cast<Mlp*>(net), params, stream());
  API_END
}

int imped_mlp_forward(imped_mlp net, const float* x, int batch, float* y) {
  API_BEGIN
  if (batch < 1) {
    set_error("imped_mlp_forward: batch must be >= 1");
    return IMPED_ERR_INVALID;
  }
  mlp_forward(*reinterpret_cast<Mlp*>(net), x, batch, y, stream());
  API_END
}

int imped_pinned_alloc(int64_t bytes, void** out) {
  API_BEGIN
  if (bytes < 1 || !out) {
    set_error("imped_pinned_alloc: bytes must be >= 1");
    return IMPED_ERR_INVALID;
  }
  HIP_CHECK(hipHostMalloc(out, (size_t)bytes, hipHostMallocDefault));
  API_END
}

int imped_pinned_free(void* p) {
  API_BEGIN
  if (p) HIP_CHECK(hipHostFree(p));
  API_END
}

int imped_mlp_predict(imped_mlp net, const float* x, int batch, const double* weights, double scale, double* mean,
                      double* disagreement) {
  API_BEGIN
  if (!net || !x || !disagreement || batch < 1 || !weights || !(scale > 0.0)) {
    set_error("imped_mlp_predict: network, inputs, disagreement and weights required, batch >= 1, scale > 0");
    return IMPED_ERR_INVALID;
  }
  mlp_predict(*reinterpret_cast<Mlp*>(net), x, batch, weights, scale, nullptr, mean, disagreement, nullptr, stream());
  API_END
}

int imped_walkers_create(imped_mlp net, int n_walkers, int n_params, const double* theta, const double* cur,
                         const double* lo, const double* hi, double step, double temperature, double threshold,
                         uint64_t seed, imped_walkers* out) {
  API_BEGIN
  if (!net || n_walkers < 1 || n_params < 1 || !theta || !cur || !lo || !hi || !(temperature > 0.0) || !out) {
    set_error("imped_walkers_create: need a network, n_walkers, n_params >= 1, states, a box and temperature > 0");
    return IMPED_ERR_INVALID;
  }
  Mlp* m = reinterpret_cast<Mlp*>(net);
  int32_t dims[17];  // at most 16 layers
  mlp_layers(*m, dims);
  if (dims[0] != n_params) {
    set_error("imped_walkers_create: n_params must equal the network's input width");
    return IMPED_ERR_INVALID;
  }
  for (int j = 0; j < n_params; ++j)
    if (!(hi[j] > lo[j])) {
      set_error("imped_walkers_create: every box needs hi > lo");
      return IMPED_ERR_INVALID;
    }
  auto w = std::make_unique<Walkers>();
  w->net = m;
  w->W = n_walkers;
  w->P = n_params;
  w->theta.assign(theta, theta + (size_t)n_walkers * n_params);
  w->prop = w->theta;
  w->cur.assign(cur, cur + n_walkers);
  w->lo.assign(lo, lo + n_params);
  w->hi.assign(hi, hi + n_params);
  w->span.resize(n_params);
  for (int j = 0; j < n_params; ++j) w->span[j] = step * (hi[j] - lo[j]);
  w->waiting.assign(n_walkers, 0);
  w->temperature = temperature;
  w->threshold = threshold;
  w->best = *std::min_element(w->cur.begin(), w->cur.end());
  w->rng.state = seed;
  w->urng.state = ~seed;
  *out = reinterpret_cast<imped_walkers>(w.release());
  API_END
}

int imped_walkers_destroy(imped_walkers w) {
  delete reinterpret_cast<Walkers*>(w);
  return IMPED_OK;
}

int imped_walkers_step(imped_walkers wk, const double* weights, double scale, const double* target,
                       int32_t* n_waiting) {
  API_BEGIN
  if (!wk || !weights || !target || !(scale > 0.0)) {
    set_error("imped_walkers_step: walkers, weights, target and scale > 0 required");
    return IMPED_ERR_INVALID;
  }
  Walkers& w = *reinterpret_cast<Walkers*>(wk);
  const int nw = walkers_step(w, weights, scale, target, stream());
  if (n_waiting) *n_waiting = nw;
  API_END
}

int imped_walkers_waiting(imped_walkers wk, int32_t* idx, double* proposals, int32_t* count) {
  API_BEGIN
  if (!wk || !count) {
    set_error("imped_walkers_waiting: walkers and count required");
    return IMPED_ERR_INVALID;
  }
  const Walkers& w = *reinterpret_cast<const Walkers*>(wk);
  int k = 0;
  for (int i = 0; i < w.W; ++i)
    if (w.waiting[i]) {
      if (idx) idx[k] = i;
      if (proposals)
        std::copy(w.prop.begin() + (size_t)i * w.P, w.prop.begin() + (size_t)(i + 1) * w.P,
                  proposals + (size_t)k * w.P);
      ++k;
    }
  *count = k;
  API_END
}

int imped_walkers_decide(imped_walkers wk, int32_t count, const int32_t* idx, const double* misfit) {
  API_BEGIN
  if (!wk || count < 0 || (count > 0 && (!idx || !misfit))) {
    set_error("imped_walkers_decide: walkers required; indices and misfits for count > 0");
    return IMPED_ERR_INVALID;
  }
  Walkers& w = *reinterpret_cast<Walkers*>(wk);
  for (int k = 0; k < count; ++k)
    if (idx[k] < 0 || idx[k] >= w.W || !w.waiting[idx[k]] || (k > 0 && idx[k] <= idx[k - 1])) {
      set_error("imped_walkers_decide: indices must be waiting walkers in ascending order");
      return IMPED_ERR_INVALID;
    }
  walkers_decide(w, count, idx, misfit);
  API_END
}

int imped_walkers_state(imped_walkers wk, double* theta, double* cur, double* best, int64_t* counters) {
  API_BEGIN
  if (!wk) {
    set_error("imped_walkers_state: walkers required");
    return IMPED_ERR_INVALID;
  }
  const Walkers& w = *reinterpret_cast<const Walkers*>(wk);
  if (theta) std::copy(w.theta.begin(), w.theta.end(), theta);
  if (cur) std::copy(w.cur.begin(), w.cur.end(), cur);
  if (best) *best = w.best;
  if (counters) {
    counters[0] = w.n_predicted;
    counters[1] = w.n_served;
    counters[2] = w.n_accepted;
    counters[3] = w.n_waiting;
  }
  API_END
}

int imped_walkers_served(imped_walkers wk, int64_t start, int64_t count, double* params) {
  API_BEGIN
  if (!wk || (count > 0 && !params)) {
    set_error("imped_walkers_served: walkers and an output array required");
    return IMPED_ERR_INVALID;
  }
  const Walkers& w = *reinterpret_cast<const Walkers*>(wk);
  const int64_t recorded = (int64_t)(w.served.size() / (size_t)w.P);
  if (start < 0 || count < 0 || start + count > recorded) {
    set_error("imped_walkers_served: rows out of range (only the recorded rows can be read)");
    return IMPED_ERR_INVALID;
  }
  std::copy(w.served.begin() + start * w.P, w.served.begin() + (start + count) * w.P, params);
  API_END
}

int imped_walkers_set_record_cap(imped_walkers wk, int64_t max_rows) {
  if (!wk || max_rows < 0) {
    set_error("imped_walkers_set_record_cap: walkers required and max_rows >= 0");
    return IMPED_ERR_INVALID;
  }
  Walkers& w = *reinterpret_cast<Walkers*>(wk);
  w.record_cap = max_rows;
  if ((int64_t)(w.served.size() / (size_t)w.P) > max_rows) w.served.resize((size_t)max_rows * w.P);
  return IMPED_OK;
}

int imped_walkers_reset_counters(imped_walkers wk) {
  if (!wk) {
    set_error("imped_walkers_reset_counters: walkers required");
    return IMPED_ERR_INVALID;
  }
  Walkers& w = *reinterpret_cast<Walkers*>(wk);
  w.n_predicted = w.n_served = w.n_accepted = 0;
  w.served.clear();
  return IMPED_OK;
}

int imped_mlp_misfit(imped_mlp net, const float* x, int batch, const double* weights, double scale,
                     const double* target, double* disagreement, double* misfit, double* mean) {
  API_BEGIN
  if (batch < 1 || !weights || !target || !disagreement || !misfit || !(scale > 0.0)) {
    set_error("imped_mlp_misfit: batch must be >= 1, weights, target and outputs given and scale > 0");
    return IMPED_ERR_INVALID;
  }
  mlp_predict(*reinterpret_cast<Mlp*>(net), x, batch, weights, scale, target, mean, disagreement, misfit, stream());
  API_END
}

int imped_mlp_gradients(imped_mlp net, const float* x, const float* y, int batch, float* grads, double* loss) {
  API_BEGIN
  if (batch < 1) {
    set_error("imped_mlp_gradients: batch must be >= 1");
    return IMPED_ERR_INVALID;
  }
  mlp_gradients(*reinterpret_cast<Mlp*>(net), x, y, batch, grads, loss, stream());
  API_END
}

int imped_mlp_set_data(imped_mlp net, const float* x, const float* y, int64_t n) {
  API_BEGIN
  if (n < 1) {
    set_error("imped_mlp_set_data: empty data set");
    return IMPED_ERR_INVALID;
  }
  mlp_set_data(*reinterpret_cast<Mlp*>(net), x, y, n, stream());
  API_END
}

int imped_mlp_train(imped_mlp net, const int32_t* batches, int steps, int batch, float lr, float beta1, float beta2,
                    float eps, double* losses) {
  API_BEGIN
  Mlp* m = reinterpret_cast<Mlp*>(net);
  const int64_t n = mlp_data_count(*m);
  if (steps < 0 || batch < 1 || n < 1) {
    set_error("imped_mlp_train: set the data first; steps >= 0, batch >= 1");
    return IMPED_ERR_INVALID;
  }
  for (int64_t i = 0; i < (int64_t)steps * batch; ++i)
    if (batches[i] < 0 || batches[i] >= n) {
      set_error("imped_mlp_train: minibatch row index out of range");
      return IMPED_ERR_INVALID;
    }
  mlp_train(*m, batches, steps, batch, lr, beta1, beta2, eps, losses, stream());
  API_END
}

}  // extern "C"
