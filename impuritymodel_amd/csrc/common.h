// Shared definitions of the MI355X exact-diagonalization engine (gfx950).
//
// Product states are W packed uint64 words; spin-orbital i lives in word i/64
// at bit 63 - i%64, so comparing words lexicographically orders states like
// the reference's big-endian bytes keys (product_state_representation.py:131-161).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#define IMPED_OK 0
#define IMPED_ERR_INVALID -1
#define IMPED_ERR_HIP -2
#define IMPED_ERR_CAPACITY -3
#define IMPED_ERR_BREAKDOWN -4
#define IMPED_ERR_NOCONV -5

#define IMPED_MAXW 2
#define IMPED_WAVE 64

namespace imped {

void set_error(const std::string& msg);

struct HipError {
  std::string msg;
};

#define HIP_CHECK(expr)                                                                  \
  do {                                                                                   \
    hipError_t e__ = (expr);                                                             \
    if (e__ != hipSuccess) {                                                             \
      throw ::imped::HipError{std::string(#expr) + ": " + hipGetErrorString(e__) +       \
                              " at " + __FILE__ + ":" + std::to_string(__LINE__)};       \
    }                                                                                    \
  } while (0)

// Device memory in size classes (powers of two up to 64 MiB, then 2 MiB
// steps) with a cache of idle blocks: a block given back is handed out again
// for the same class instead of returning to the runtime, so the handles a
// batched call creates and destroys cost no hipMalloc/hipFree after the
// first call.  All engine work runs in order on one stream, so a block's next
// owner is ordered after its previous owner's queued work.
size_t devmem_class(size_t bytes);
void* devmem_alloc(size_t class_bytes);
void devmem_free(void* p, size_t class_bytes);

// Device buffer owned by a handle; grows on demand, never shrinks.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) devmem_free(p, bytes);
  }
  void reserve(size_t nb) {
    if (nb <= bytes) return;
    if (p) devmem_free(p, bytes);
    p = nullptr;
    bytes = 0;
    const size_t alloc = devmem_class(nb);
    p = devmem_alloc(alloc);
    bytes = alloc;
  }
  // Grow while keeping the first `keep` bytes.
  void grow_keep(size_t nb, size_t keep, hipStream_t s) {
    if (nb <= bytes) return;
    const size_t alloc = devmem_class(bytes * 2 > nb ? bytes * 2 : nb);
    void* q = devmem_alloc(alloc);
    if (p && keep) HIP_CHECK(hipMemcpyAsync(q, p, keep, hipMemcpyDeviceToDevice, s));
    if (p) devmem_free(p, bytes);  // reused only by work queued after the copy
    p = q;
    bytes = alloc;
  }
  template <class T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
};

// Host-to-device copies.  Up to 1 MiB go through a page-locked ring: the
// bytes are copied into the ring at once and the device copy is queued
// without the host waiting for it (a pageable source makes the runtime
// stage and wait); a ring slot is reused only after stream_sync of its
// stream.  Larger copies, copies while the stream is being captured into a
// graph, and copies that find the ring full go straight from src.
void h2d_bytes(void* dst, const void* src, size_t bytes, hipStream_t s);

template <class T>
inline void h2d(DevBuf& b, const T* src, size_t n, hipStream_t s) {
  b.reserve(n * sizeof(T));
  if (n) h2d_bytes(b.p, src, n * sizeof(T), s);
}

// Several host arrays sent in one copy: add() places an array at a 256-byte
// aligned offset of the pack and returns that offset; upload() copies the
// whole pack into a device buffer, after which at<T>(d, off) is the array's
// device address.
struct HostPack {
  std::vector<char> bytes;
  template <class T>
  size_t add(const T* p, size_t n) {
    const size_t off = (bytes.size() + 255) & ~size_t(255);
    bytes.resize(off + std::max<size_t>(n, 1) * sizeof(T), 0);
    if (n) std::memcpy(bytes.data() + off, p, n * sizeof(T));
    return off;
  }
  void upload(DevBuf& d, hipStream_t s) const { h2d(d, bytes.data(), std::max<size_t>(bytes.size(), 1), s); }
  template <class T>
  static T* at(const DevBuf& d, size_t off) {
    return reinterpret_cast<T*>(static_cast<char*>(d.p) + off);
  }
};

// Device-to-host reads.  Small ones (<= 64 KiB) go through a pinned ring:
// the copy is queued like a kernel and dst is filled by the next
// stream_sync on the same stream (a pageable destination would make every
// small read a blocking staged copy of its own).  Larger reads go straight
// to dst; they too are complete only after stream_sync.
void d2h_bytes(void* dst, const void* src, size_t bytes, hipStream_t s);
// hipStreamSynchronize, then completes the staged reads of stream s.
void stream_sync(hipStream_t s);

template <class T>
inline void d2h(T* dst, const void* src, size_t n, hipStream_t s) {
  if (n) d2h_bytes(dst, src, n * sizeof(T), s);
}

// Launch shapes chosen per chain-dimension bucket (set by the autotuner).
struct Tuning {
  int64_t lanczos_wave_max_n = 4096;    // chains up to this size: one wave each
  int64_t lanczos_wg256_max_n = 65536;  // then 256 threads, above: 1024
  int64_t eigen_bt = 0;                 // eigensolver threads per chain (256, 512, 1024; 0: by the largest chain)
  int64_t eigen_wave_max_n = 0;         // eigensolver chains up to this size: one wave each (off)
  int64_t eigen_real = 1;               // real Krylov vectors when H is real
  int64_t eigen_dgks = 1;               // second Gram-Schmidt pass only when needed
  int64_t eigen_check_every = 12;       // Ritz convergence check interval
  int64_t eigen_check_start = 48;       // no Ritz check before min(this, n/2) steps
  int64_t eigen_dense_max_n = 0;        // eigensolver chains up to this size (<= 112): dense solver (0: off)
  int64_t flat_min_n = 32768;           // chains above this size use the grid-wide (flat) kernels
  int64_t spmv_lpr = 8;                 // lanes per CSR row in the flat SpMV (4..64)
  int64_t real_values = 1;              // grid-wide SpMV on a real CSR reads 8-byte real values
  int64_t lanczos_graph_steps = 16;     // flat Lanczos steps per replayed HIP graph (0: direct launches)
  int64_t eigen_lds_vectors = -1;       // cap on the eigensolver's LDS Krylov vectors (-1: as many as fit)
  int64_t small_sort = 1;               // one-workgroup sort for small key sets (0: digit passes always)
  int64_t mlp_graph = 0;                // surrogate predictions as one replayed HIP graph (0: direct launches)
  int64_t lanczos_lds = 1;              // reorthogonalized chains whose Krylov basis fits: basis in LDS
  int64_t lanczos_lds_waves = 4;        // waves per chain of the LDS-resident Lanczos (1, 2 or 4)
  int64_t lanczos_dense_max_n = 0;      // reorthogonalized chains up to this size (<= 112): Householder reduction (0: off)
};
extern Tuning g_tuning;

// True while launch timing is switched on (every launch then goes out directly).
bool timing_enabled();

// Optional HIP-event timing of kernel launches, accumulated per kernel name.
struct TimedLaunch {
  const char* name;
  hipStream_t s;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  TimedLaunch(const char* n, hipStream_t st);
  ~TimedLaunch();
};

inline unsigned blocks_for(int64_t n, int threads) {
  int64_t b = (n + threads - 1) / threads;
  return (unsigned)(b < 1 ? 1 : b);
}

// ------------------------------------------------------------------ device helpers

// Cross-lane move of a double by a DPP pattern (both 32-bit halves), lanes
// outside row_mask reading 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, ROW_MASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, ROW_MASK, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a wave, identical in every lane, with every lane
// active: DPP moves inside registers (quad swaps, half-row and row mirrors,
// then row broadcasts 15 and 31 collect the four rows in lane 63), read back
// from lane 63.  Fixed order, so runs are reproducible.
// Maximum over the wave, the same in every lane (wave-uniform).
__device__ __forceinline__ int wave_max_int(int x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x = max(x, __shfl_xor(x, off));
  return __builtin_amdgcn_readfirstlane(x);
}

__device__ __forceinline__ double wave_sum(double x) {
  x += dpp_move<0xB1, 0xF>(x);   // quad_perm [1,0,3,2]
  x += dpp_move<0x4E, 0xF>(x);   // quad_perm [2,3,0,1]
  x += dpp_move<0x141, 0xF>(x);  // row_half_mirror
  x += dpp_move<0x140, 0xF>(x);  // row_mirror
  x += dpp_move<0x142, 0xA>(x);  // row_bcast:15 into rows 1 and 3
  x += dpp_move<0x143, 0xC>(x);  // row_bcast:31 into rows 2 and 3
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), 63);
  return __hiloint2double(hi, lo);
}


struct Key {
  uint64_t w[IMPED_MAXW];
};

__device__ __forceinline__ uint32_t hash_key(uint32_t chain, const uint64_t* k, int W) {
  uint64_t h = 0x9E3779B97F4A7C15ull ^ (uint64_t)chain * 0xD6E8FEB86659FD93ull;
  for (int w = 0; w < W; ++w) {
    h ^= k[w] + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 31;
  }
  h *= 0x94D049BB133111EBull;
  h ^= h >> 29;
  return (uint32_t)h;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

// |z|^2 evaluated as abs(z)**2 in the reference (finite.py:1509).
__device__ __forceinline__ double abs2_ref(double2 z) {
  double a = hypot(z.x, z.y);
  return a * a;
}

// 1/(a + ib) by Smith's scaling (robust to over/underflow).
__device__ __forceinline__ double2 crecip(double2 d) {
  if (fabs(d.x) >= fabs(d.y)) {
    double r = d.y / d.x;
    double den = d.x + d.y * r;
    return make_double2(1.0 / den, -r / den);
  } else {
    double r = d.x / d.y;
    double den = d.x * r + d.y;
    return make_double2(r / den, -1.0 / den);
  }
}

}  // namespace imped
