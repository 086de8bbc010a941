// Scans, sorts and wave-task construction for the basis builder.
//
// Sorting is what fixes the reference's basis order: every BFS layer is
// appended in ascending key order (``basis += sorted(basis_new)``,
// impurityModel/ed/finite.py:1749), and applyOp's insertion order decides the
// order of a transition operator's output (finite.py:1375-1388).  Records are
// sorted by this file's stable LSD radix passes over the key words (least
// significant first) and then the chain id; counts become offsets by this
// file's exclusive scan.
#include <cstring>

#include "kernels.h"

namespace imped {

namespace {

struct Scratch {
  DevBuf temp, wide, k0, k1, v1, hist, hoff, gather;
};
Scratch& scratch() {
  static Scratch s;
  return s;
}

// ---- exclusive scan of int32 counts into int64 offsets
//
// A tile is SCAN_T threads x SCAN_ITEMS consecutive counts per thread: each
// thread sums its items, each wave scans its threads' sums with lane
// shuffles, the wave totals are scanned through LDS, and every thread writes
// its items' offsets.  An input of one tile takes one launch; a larger one
// takes its tile totals, a scan of those totals (the same way, one level
// down) and the tile scans.

constexpr int SCAN_T = 1024;
constexpr int SCAN_ITEMS = 8;
constexpr int64_t SCAN_TILE = (int64_t)SCAN_T * SCAN_ITEMS;

__device__ __forceinline__ int64_t wave_inclusive_scan(int64_t x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int64_t y = __shfl_up(x, d);
    if (lane >= d) x += y;
  }
  return x;
}

// Exclusive prefix of v over the workgroup (thread order); *total gets the sum.
__device__ __forceinline__ int64_t block_exclusive_scan(int64_t v, int64_t* total) {
  __shared__ int64_t wsum[SCAN_T / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t inc = wave_inclusive_scan(v);
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  if (wave == 0) {
    const int64_t t = lane < SCAN_T / 64 ? wsum[lane] : 0;
    const int64_t ti = wave_inclusive_scan(t);
    if (lane < SCAN_T / 64) wsum[lane] = ti - t;
    if (lane == SCAN_T / 64 - 1) *total = ti;
  }
  __syncthreads();
  const int64_t r = wsum[wave] + inc - v;
  __syncthreads();
  return r;
}

template <class In>
__device__ __forceinline__ void load_items(const In* in, int64_t n, int64_t first, int64_t* it) {
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const int64_t i = first + k;
    it[k] = i < n ? (int64_t)in[i] : 0;
  }
}

// Sum of each tile's values.
template <class In>
__global__ __launch_bounds__(SCAN_T) void scan_tile_sums_kernel(const In* in, int64_t n, int64_t* sums) {
  __shared__ int64_t total;
  int64_t it[SCAN_ITEMS];
  load_items(in, n, (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS, it);
  int64_t t = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) t += it[k];
  block_exclusive_scan(t, &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Exclusive offsets of every value: the tile's prefix (sums, or 0) plus the
// in-tile scan.  A thread reads its own items before writing them, so out
// may be in.
template <class In>
__global__ __launch_bounds__(SCAN_T) void scan_tiles_kernel(const In* in, int64_t n, const int64_t* sums,
                                                            int64_t* out) {
  __shared__ int64_t total;
  int64_t it[SCAN_ITEMS];
  const int64_t first = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  load_items(in, n, first, it);
  int64_t t = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) t += it[k];
  int64_t run = block_exclusive_scan(t, &total) + (sums ? sums[blockIdx.x] : 0);
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    if (first + k < n) out[first + k] = run;
    run += it[k];
  }
}

// Exclusive scan of n values of in into out, one level per SCAN_TILE factor:
// tile sums, their scan (recursively), tile scans.
template <class In>
void scan_level(const In* in, int64_t n, int64_t* out, int level, hipStream_t s) {
  const int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (tiles == 1) {
    hipLaunchKernelGGL(scan_tiles_kernel<In>, dim3(1), dim3(SCAN_T), 0, s, in, n, (const int64_t*)nullptr, out);
    HIP_CHECK(hipGetLastError());
    return;
  }
  if (tiles > (int64_t)UINT32_MAX) throw HipError{"scan: too many values"};
  static DevBuf levels[4];  // tile sums of each level (4 levels cover 2^52 values)
  if (level >= 4) throw HipError{"scan: too many levels"};
  DevBuf& sb = levels[level];
  sb.reserve(tiles * sizeof(int64_t));
  int64_t* sums = sb.as<int64_t>();
  hipLaunchKernelGGL(scan_tile_sums_kernel<In>, dim3((unsigned)tiles), dim3(SCAN_T), 0, s, in, n, sums);
  HIP_CHECK(hipGetLastError());
  scan_level<int64_t>(sums, tiles, sums, level + 1, s);
  hipLaunchKernelGGL(scan_tiles_kernel<In>, dim3((unsigned)tiles), dim3(SCAN_T), 0, s, in, n, (const int64_t*)sums,
                     out);
  HIP_CHECK(hipGetLastError());
}

// Exclusive scan of int32 counts into int64 offsets, stream-ordered.
void scan_counts_async(const int32_t* counts, int64_t n, int64_t* offsets, hipStream_t s) {
  if (n == 0) return;
  scan_level<int32_t>(counts, n, offsets, 0, s);
}

__global__ void gather_word_kernel(const uint64_t* keys, int W, int w, const int32_t* perm, uint64_t* out,
                                   int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = keys[(size_t)perm[i] * W + w];
}

__global__ void gather_u64_kernel(const uint64_t* src, const int32_t* perm, uint64_t* out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = src[perm[i]];
}

// ---- stable LSD radix sort of (uint64 key, int32 payload) pairs, 8-bit digits
//
// Pass over one digit: (1) per-tile digit histograms, (2) one exclusive scan
// over the digit-major histogram matrix (digit d of tile b lands after all
// smaller digits and after digit d of tiles < b), (3) a stable scatter: each
// tile walks its elements in index order, 256 at a time; inside a wave the
// lanes holding the same digit are found with 8 ballots and ranked by lane,
// waves are ordered through LDS counters.

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;

__global__ __launch_bounds__(RS_THREADS) void rs_hist_kernel(const uint64_t* keys, int64_t n, int shift,
                                                             int n_tiles, int32_t* hist) {
  __shared__ int32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  for (int k = 0; k < RS_ITEMS; ++k) {
    const int64_t i = base + (int64_t)k * RS_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * n_tiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(RS_THREADS) void rs_scatter_kernel(const uint64_t* keys, const int32_t* vals, int64_t n,
                                                                int shift, int n_tiles, const int64_t* offs,
                                                                uint64_t* keys_out, int32_t* vals_out) {
  __shared__ int64_t run[256];
  __shared__ int32_t wcnt[RS_THREADS / 64][256];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  run[threadIdx.x] = offs[(int64_t)threadIdx.x * n_tiles + blockIdx.x];
  const uint64_t lt = (1ull << lane) - 1ull;
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  for (int k = 0; k < RS_ITEMS; ++k) {
    for (int j = threadIdx.x; j < (RS_THREADS / 64) * 256; j += RS_THREADS) (&wcnt[0][0])[j] = 0;
    __syncthreads();
    const int64_t i = base + (int64_t)k * RS_THREADS + threadIdx.x;
    const bool valid = i < n;
    uint64_t key = 0;
    int32_t val = 0;
    uint32_t d = 0;
    if (valid) {
      key = keys[i];
      val = vals[i];
      d = (uint32_t)(key >> shift) & 255u;
    }
    uint64_t eq = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = __ballot(bit);
      eq &= bit ? m : ~m;
    }
    const int rank = __popcll(eq & lt);
    if (valid && rank == 0) wcnt[wave][d] = __popcll(eq);
    __syncthreads();
    if (valid) {
      int64_t pos = run[d] + rank;
      for (int w = 0; w < wave; ++w) pos += wcnt[w][d];
      keys_out[pos] = key;
      vals_out[pos] = val;
    }
    __syncthreads();
    int32_t add = 0;
    for (int w = 0; w < RS_THREADS / 64; ++w) add += wcnt[w][threadIdx.x];
    run[threadIdx.x] += add;
    __syncthreads();
  }
}

// Stable sort of perm by keys_in[i] (already gathered in perm order) on bits
// [begin_bit, end_bit).  perm is updated in place.
void sort_pass(const uint64_t* keys_in, int32_t* perm, int64_t n, int begin_bit, int end_bit, hipStream_t s) {
  if (n <= 1 || begin_bit >= end_bit) return;
  if (n > INT32_MAX) throw HipError{"sort: too many records"};
  Scratch& sc = scratch();
  const int n_tiles = (int)((n + RS_TILE - 1) / RS_TILE);
  sc.k0.reserve(n * 8);
  sc.k1.reserve(n * 8);
  sc.v1.reserve(n * 4);
  sc.hist.reserve((size_t)256 * n_tiles * 4);
  sc.hoff.reserve(((size_t)256 * n_tiles + 1) * 8);
  if (keys_in != sc.k0.as<uint64_t>())
    HIP_CHECK(hipMemcpyAsync(sc.k0.p, keys_in, n * 8, hipMemcpyDeviceToDevice, s));
  uint64_t* ka = sc.k0.as<uint64_t>();
  uint64_t* kb = sc.k1.as<uint64_t>();
  int32_t* va = perm;
  int32_t* vb = sc.v1.as<int32_t>();
  for (int shift = begin_bit; shift < end_bit; shift += 8) {
    hipLaunchKernelGGL(rs_hist_kernel, dim3(n_tiles), dim3(RS_THREADS), 0, s, ka, n, shift, n_tiles,
                       sc.hist.as<int32_t>());
    HIP_CHECK(hipGetLastError());
    scan_counts_async(sc.hist.as<int32_t>(), (int64_t)256 * n_tiles, sc.hoff.as<int64_t>(), s);
    hipLaunchKernelGGL(rs_scatter_kernel, dim3(n_tiles), dim3(RS_THREADS), 0, s, ka, va, n, shift, n_tiles,
                       sc.hoff.as<int64_t>(), kb, vb);
    HIP_CHECK(hipGetLastError());
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (va != perm) HIP_CHECK(hipMemcpyAsync(perm, va, n * 4, hipMemcpyDeviceToDevice, s));
}

int bits_for(int64_t n) {
  int b = 1;
  while ((int64_t(1) << b) < n) ++b;
  return b;
}

// Digit passes of a word whose used bits are [64 - used, 64), rounded to
// whole 8-bit digits from the bottom of the used range.
void sort_word(const uint64_t* gathered_in_k0, int32_t* perm, int64_t n, int used_bits, hipStream_t s) {
  int begin = 64 - used_bits;
  begin -= begin % 8;
  sort_pass(gathered_in_k0, perm, n, begin, 64, s);
}

__global__ void gather_i32_as_u64_kernel(const int32_t* src, const int32_t* perm, uint64_t* out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (uint64_t)(uint32_t)src[perm[i]];
}

// One-workgroup sort of up to SMALL_SORT_MAX records whose (prim, key,
// position) fits one 63-bit code: code = prim << (n_orb + ib) | key << ib | i
// with i the record's position in perm, so every code is distinct and the
// order is the stable (prim, key) order of the digit passes.  Bitonic network
// in LDS (padding codes all ones sort last), then perm[i] = perm[code_i & mask].
constexpr int SMALL_SORT_MAX = 16384;
constexpr int SMALL_SORT_THREADS = 1024;

__global__ __launch_bounds__(SMALL_SORT_THREADS) void small_sort_kernel(const uint64_t* keys, const int32_t* prim,
                                                                        int32_t* perm, int n, int npow2, int n_orb,
                                                                        int ib) {
  extern __shared__ uint64_t code[];
  const int shift_prim = n_orb + ib;
  for (int i = threadIdx.x; i < npow2; i += SMALL_SORT_THREADS) {
    uint64_t c = ~0ull;
    if (i < n) {
      const int32_t r = perm[i];
      const uint64_t key = n_orb > 0 ? keys[r] >> (64 - n_orb) : 0ull;
      const uint64_t pr = shift_prim < 64 ? (uint64_t)(uint32_t)prim[r] << shift_prim : 0ull;
      c = pr | (key << ib) | (uint64_t)i;
    }
    code[i] = c;
  }
  __syncthreads();
  // each stage's npow2/2 compare-exchange pairs are disjoint: a thread loads
  // all of its pairs, then compares, then stores (one LDS round trip per
  // stage).  Pair q of a stage: i = its index with a zero inserted at bit
  // log2(j), partner i + j, ascending when (i & k) == 0.
  constexpr int PPT = SMALL_SORT_MAX / 2 / SMALL_SORT_THREADS;  // pairs per thread at most
  const int half = npow2 >> 1;
  for (int k = 2; k <= npow2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      uint64_t av[PPT], bv[PPT];
      int iv[PPT];
#pragma unroll
      for (int u = 0; u < PPT; ++u) {
        const int q = threadIdx.x + u * SMALL_SORT_THREADS;
        const int i = ((q & ~(j - 1)) << 1) | (q & (j - 1));
        iv[u] = i;
        if (q < half) {
          av[u] = code[i];
          bv[u] = code[i + j];
        }
      }
#pragma unroll
      for (int u = 0; u < PPT; ++u) {
        const int q = threadIdx.x + u * SMALL_SORT_THREADS;
        if (q < half && (av[u] > bv[u]) == ((iv[u] & k) == 0)) {
          code[iv[u]] = bv[u];
          code[iv[u] + j] = av[u];
        }
      }
      __syncthreads();
    }
  }
  const uint64_t mask = (1ull << ib) - 1;
  int32_t v[SMALL_SORT_MAX / SMALL_SORT_THREADS];
  int q = 0;
  for (int i = threadIdx.x; i < n; i += SMALL_SORT_THREADS) v[q++] = perm[(int)(code[i] & mask)];
  __syncthreads();  // every read of the old perm precedes the writes
  q = 0;
  for (int i = threadIdx.x; i < n; i += SMALL_SORT_THREADS) perm[i] = v[q++];
}

bool small_sort(int W, int n_orb, int n_prim_bits, const uint64_t* keys, const int32_t* prim, int32_t* perm,
                int64_t n, hipStream_t s) {
  if (W != 1 || n > SMALL_SORT_MAX || n_orb > 64) return false;
  const int ib = bits_for(n);
  if (n_prim_bits + n_orb + ib > 63) return false;
  int npow2 = 1;
  while (npow2 < n) npow2 <<= 1;
  static bool attr = false;
  if (!attr) {
    HIP_CHECK(hipFuncSetAttribute((const void*)small_sort_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  SMALL_SORT_MAX * (int)sizeof(uint64_t)));
    attr = true;
  }
  hipLaunchKernelGGL(small_sort_kernel, dim3(1), dim3(SMALL_SORT_THREADS), (size_t)npow2 * sizeof(uint64_t), s, keys,
                     prim, perm, (int)n, npow2, n_orb, ib);
  HIP_CHECK(hipGetLastError());
  return true;
}

void sort_words_then(int W, int n_orb, int n_prim_bits, const uint64_t* keys, const int32_t* prim, int32_t* perm,
                     int64_t n, hipStream_t s) {
  if (n <= 1) return;
  if (g_tuning.small_sort && small_sort(W, n_orb, n_prim_bits, keys, prim, perm, n, s)) return;
  Scratch& sc = scratch();
  sc.k0.reserve(n * sizeof(uint64_t));
  for (int w = W - 1; w >= 0; --w) {
    // Orbital bits of word w occupy its top bits; lower bits are zero.
    int used = n_orb - 64 * w;
    if (used <= 0) continue;
    if (used > 64) used = 64;
    hipLaunchKernelGGL(gather_word_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, keys, W, w, perm,
                       sc.k0.as<uint64_t>(), n);
    sort_word(sc.k0.as<uint64_t>(), perm, n, used, s);
  }
  sc.k0.reserve(n * sizeof(uint64_t));
  hipLaunchKernelGGL(gather_i32_as_u64_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, prim, perm,
                     sc.k0.as<uint64_t>(), n);
  sort_pass(sc.k0.as<uint64_t>(), perm, n, 0, n_prim_bits, s);
}

}  // namespace

int64_t scan_counts(const int32_t* counts, int64_t n, int64_t* offsets, hipStream_t s) {
  if (n == 0) return 0;
  scan_counts_async(counts, n, offsets, s);
  int64_t last_off = 0;
  int32_t last_cnt = 0;
  d2h_bytes(&last_off, offsets + n - 1, sizeof(int64_t), s);
  d2h_bytes(&last_cnt, counts + n - 1, sizeof(int32_t), s);
  stream_sync(s);
  return last_off + last_cnt;
}

void sort_chain_keys(int W, int n_orb, int n_chain_bits, const uint64_t* keys, const int32_t* chain, int32_t* perm,
                     int64_t n, hipStream_t s) {
  TimedLaunch tl("sort", s);  // the span of all passes of one BFS layer's sort
  sort_words_then(W, n_orb, n_chain_bits, keys, chain, perm, n, s);
}

void sort_vec_keys_stable(int W, int n_orb, int n_vec_bits, const uint64_t* keys, const int32_t* vec, int32_t* perm,
                          int64_t n, hipStream_t s) {
  sort_words_then(W, n_orb, n_vec_bits, keys, vec, perm, n, s);
}

void sort_pairs_u64_u32(const uint64_t* sec, const int32_t* prim, int n_prim_bits, int32_t* perm, int64_t n,
                        hipStream_t s) {
  if (n <= 1) return;
  Scratch& sc = scratch();
  sc.k0.reserve(n * sizeof(uint64_t));
  hipLaunchKernelGGL(gather_u64_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, sec, perm, sc.k0.as<uint64_t>(),
                     n);
  sort_pass(sc.k0.as<uint64_t>(), perm, n, 0, 64, s);
  hipLaunchKernelGGL(gather_i32_as_u64_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, prim, perm,
                     sc.k0.as<uint64_t>(), n);
  sort_pass(sc.k0.as<uint64_t>(), perm, n, 0, n_prim_bits, s);
}

// Test hook: stable sort of perm by 64-bit keys (bits [begin, end)).
void sort_u64_perm(const uint64_t* keys, int32_t* perm, int64_t n, int begin_bit, int end_bit, hipStream_t s) {
  if (n <= 1) return;
  Scratch& sc = scratch();
  sc.k0.reserve(n * sizeof(uint64_t));
  hipLaunchKernelGGL(gather_u64_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, keys, perm, sc.k0.as<uint64_t>(),
                     n);
  sort_pass(sc.k0.as<uint64_t>(), perm, n, begin_bit, end_bit, s);
}

// ------------------------------------------------------------------ wave tasks

namespace {

// Segment [lo, hi) of chain c within a chain-sorted array, by binary search.
__global__ void chain_segments_kernel(const int32_t* chain_of_src, int64_t n, int n_chains, int64_t* seg_lo,
                                      int32_t* ntasks) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_chains) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (chain_of_src[mid] < c) lo = mid + 1; else hi = mid;
  }
  int64_t a = lo;
  hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (chain_of_src[mid] <= c) lo = mid + 1; else hi = mid;
  }
  seg_lo[2 * c] = a;
  seg_lo[2 * c + 1] = lo;
  ntasks[c] = (int32_t)((lo - a + 63) / 64);
}

__global__ void fill_tasks_kernel(const int64_t* seg, const int64_t* task_off, const int32_t* chain_model,
                                  int n_chains, int64_t base, WaveTask* tasks) {
  int c = blockIdx.x;
  if (c >= n_chains) return;
  int64_t lo = seg[2 * c], hi = seg[2 * c + 1];
  int64_t t0 = task_off[c];
  for (int64_t b = lo + 64 * (int64_t)threadIdx.x; b < hi; b += 64 * (int64_t)blockDim.x) {
    WaveTask t;
    t.model = chain_model[c];
    t.chain = c;
    t.begin = base + b;
    t.end = base + (b + 64 < hi ? b + 64 : hi);
    tasks[t0 + (b - lo) / 64] = t;
  }
}

struct TaskScratch {
  DevBuf seg, nt, off;
};
TaskScratch& tscratch() {
  static TaskScratch t;
  return t;
}

}  // namespace

int64_t build_wave_tasks(const int32_t* chain_of_src, int64_t begin, int64_t end, const int32_t* chain_model,
                         int n_chains, void* tasks_out, size_t cap_tasks, hipStream_t s, int32_t*) {
  const int64_t n = end - begin;
  if (n <= 0) return 0;
  TaskScratch& ts = tscratch();
  ts.seg.reserve(2 * (size_t)n_chains * sizeof(int64_t));
  ts.nt.reserve((size_t)n_chains * sizeof(int32_t));
  ts.off.reserve(((size_t)n_chains + 1) * sizeof(int64_t));
  hipLaunchKernelGGL(chain_segments_kernel, dim3(blocks_for(n_chains, 256)), dim3(256), 0, s, chain_of_src + begin,
                     n, n_chains, ts.seg.as<int64_t>(), ts.nt.as<int32_t>());
  HIP_CHECK(hipGetLastError());
  int64_t total = scan_counts(ts.nt.as<int32_t>(), n_chains, ts.off.as<int64_t>(), s);
  if ((size_t)total > cap_tasks) return -total;
  hipLaunchKernelGGL(fill_tasks_kernel, dim3(n_chains), dim3(64), 0, s, ts.seg.as<int64_t>(), ts.off.as<int64_t>(),
                     chain_model, n_chains, begin, reinterpret_cast<WaveTask*>(tasks_out));
  HIP_CHECK(hipGetLastError());
  return total;
}

}  // namespace imped
