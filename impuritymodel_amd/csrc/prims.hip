// Scans, sorts and wave-task construction for the basis builder.
//
// Sorting is what fixes the reference's basis order: every BFS layer is
// appended in ascending key order (``basis += sorted(basis_new)``,
// impurityModel/ed/finite.py:1749), and applyOp's insertion order decides the
// order of a transition operator's output (finite.py:1375-1388).  Records are
// sorted by stable LSD radix passes over the key words (least significant
// first) and then the chain id, using rocPRIM's device radix sort.
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "kernels.h"

namespace imped {

namespace {

struct Scratch {
  DevBuf temp, k0, k1, v1, k64a, k64b;
};
Scratch& scratch() {
  static Scratch s;
  return s;
}

__global__ void widen_kernel(const int32_t* in, int64_t* out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

__global__ void gather_word_kernel(const uint64_t* keys, int W, int w, const int32_t* perm, uint64_t* out,
                                   int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = keys[(size_t)perm[i] * W + w];
}

__global__ void gather_i32_kernel(const int32_t* src, const int32_t* perm, uint32_t* out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (uint32_t)src[perm[i]];
}

__global__ void gather_u64_kernel(const uint64_t* src, const int32_t* perm, uint64_t* out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = src[perm[i]];
}

template <class K>
void sort_pass(const K* keys_in, int32_t* perm, int64_t n, int begin_bit, int end_bit, hipStream_t s) {
  Scratch& sc = scratch();
  sc.k1.reserve(n * sizeof(K));
  sc.v1.reserve(n * sizeof(int32_t));
  size_t tb = 0;
  HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tb, keys_in, sc.k1.as<K>(), perm, sc.v1.as<int32_t>(), (size_t)n,
                                      begin_bit, end_bit, s));
  sc.temp.reserve(tb);
  HIP_CHECK(rocprim::radix_sort_pairs(sc.temp.p, tb, keys_in, sc.k1.as<K>(), perm, sc.v1.as<int32_t>(), (size_t)n,
                                      begin_bit, end_bit, s));
  HIP_CHECK(hipMemcpyAsync(perm, sc.v1.p, n * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
}

int bits_for(int64_t n) {
  int b = 1;
  while ((int64_t(1) << b) < n) ++b;
  return b;
}

void sort_words_then(int W, int n_orb, int n_prim_bits, const uint64_t* keys, const int32_t* prim, int32_t* perm,
                     int64_t n, hipStream_t s) {
  if (n <= 1) return;
  Scratch& sc = scratch();
  sc.k64a.reserve(n * sizeof(uint64_t));
  for (int w = W - 1; w >= 0; --w) {
    // Orbital bits of word w occupy its top bits; lower bits are zero.
    int used = n_orb - 64 * w;
    if (used <= 0) continue;
    if (used > 64) used = 64;
    hipLaunchKernelGGL(gather_word_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, keys, W, w, perm,
                       sc.k64a.as<uint64_t>(), n);
    sort_pass<uint64_t>(sc.k64a.as<uint64_t>(), perm, n, 64 - used, 64, s);
  }
  sc.k0.reserve(n * sizeof(uint32_t));
  hipLaunchKernelGGL(gather_i32_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, prim, perm, sc.k0.as<uint32_t>(),
                     n);
  sort_pass<uint32_t>(sc.k0.as<uint32_t>(), perm, n, 0, n_prim_bits, s);
}

}  // namespace

int64_t scan_counts(const int32_t* counts, int64_t n, int64_t* offsets, hipStream_t s) {
  if (n == 0) return 0;
  Scratch& sc = scratch();
  sc.k64b.reserve((n + 1) * sizeof(int64_t));
  int64_t* wide = sc.k64b.as<int64_t>();
  hipLaunchKernelGGL(widen_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, counts, wide, n);
  size_t tb = 0;
  HIP_CHECK(rocprim::exclusive_scan(nullptr, tb, wide, offsets, (int64_t)0, (size_t)n, rocprim::plus<int64_t>(), s));
  sc.temp.reserve(tb);
  HIP_CHECK(
      rocprim::exclusive_scan(sc.temp.p, tb, wide, offsets, (int64_t)0, (size_t)n, rocprim::plus<int64_t>(), s));
  int64_t last_off = 0;
  int32_t last_cnt = 0;
  HIP_CHECK(hipMemcpyAsync(&last_off, offsets + n - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(&last_cnt, counts + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  return last_off + last_cnt;
}

void sort_chain_keys(int W, int n_orb, int n_chain_bits, const uint64_t* keys, const int32_t* chain, int32_t* perm,
                     int64_t n, hipStream_t s) {
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
  sc.k64a.reserve(n * sizeof(uint64_t));
  hipLaunchKernelGGL(gather_u64_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, sec, perm, sc.k64a.as<uint64_t>(),
                     n);
  sort_pass<uint64_t>(sc.k64a.as<uint64_t>(), perm, n, 0, 64, s);
  sc.k0.reserve(n * sizeof(uint32_t));
  hipLaunchKernelGGL(gather_i32_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, prim, perm, sc.k0.as<uint32_t>(),
                     n);
  sort_pass<uint32_t>(sc.k0.as<uint32_t>(), perm, n, 0, n_prim_bits, s);
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
