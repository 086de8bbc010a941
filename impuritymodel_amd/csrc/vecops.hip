// Sparse multi-configurational states on the device.
//
// vector_apply: op|psi> for many (operator, state) pairs at once, with the
// semantics of applyOp (impurityModel/ed/finite.py:1357-1388, 1507-1510):
// contributions to one target are summed in source order, restrictions
// filter targets, |amp|^2 < slaterWeightMin is pruned after summation, and
// the result keeps the reference's dict insertion order -- targets ordered by
// their first contribution (source position, then the operator's term
// order).  Optional normalization as getSpectra (spectra.py:787-790).
//
// truncate_eigvecs: eigenvector dicts of eigensystem (finite.py:114-117),
// components with |v_i|^2 >= slaterWeightMin in basis order, not renormalized.
#include <cstring>

#include "krylov.h"

namespace imped {

namespace {

__global__ void gather_sources_kernel(int W, const int64_t* so_off, const int64_t* src_begin, int n_out,
                                      const uint64_t* keys, const double2* amps, uint64_t* out_keys,
                                      double2* out_amps, int32_t* out_pos) {
  const int o = blockIdx.x;
  if (o >= n_out) return;
  const int64_t a = so_off[o], b = so_off[o + 1], s0 = src_begin[o];
  for (int64_t i = a + threadIdx.x; i < b; i += blockDim.x) {
    const int64_t j = s0 + (i - a);
    for (int w = 0; w < W; ++w) out_keys[i * W + w] = keys[j * W + w];
    out_amps[i] = amps[j];
    out_pos[i] = (int32_t)(i - a);
  }
}

__global__ void iota_kernel(int32_t* p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (int32_t)i;
}

// Heads of equal-(vec, key) runs in sorted order; each head sums its run in
// order and flags survival.
template <int W>
__global__ void reduce_runs_kernel(const int32_t* perm, int64_t n, const uint64_t* keys, const int32_t* vec,
                                   const double2* val, const int32_t* src, const int32_t* first, double weight_min,
                                   int32_t* keep, double2* sum_out, uint64_t* order_key) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t p = perm[i];
  bool head = true;
  if (i > 0) {
    const int32_t q = perm[i - 1];
    bool same = vec[q] == vec[p];
    for (int w = 0; w < W && same; ++w) same = keys[(size_t)q * W + w] == keys[(size_t)p * W + w];
    head = !same;
  }
  if (!head) {
    keep[i] = 0;
    return;
  }
  double2 s = val[p];
  for (int64_t j = i + 1; j < n; ++j) {
    const int32_t q = perm[j];
    bool same = vec[q] == vec[p];
    for (int w = 0; w < W && same; ++w) same = keys[(size_t)q * W + w] == keys[(size_t)p * W + w];
    if (!same) break;
    s.x += val[q].x;
    s.y += val[q].y;
  }
  keep[i] = abs2_ref(s) >= weight_min ? 1 : 0;
  sum_out[i] = s;
  order_key[i] = ((uint64_t)(uint32_t)src[p] << 32) | (uint32_t)first[p];
}

__global__ void compact_runs_kernel(int W, const int32_t* perm, int64_t n, const int32_t* keep, const int64_t* pos,
                                    const uint64_t* keys, const int32_t* vec, const double2* sums,
                                    const uint64_t* order_key, uint64_t* ck, int32_t* cv, double2* ca,
                                    uint64_t* co) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !keep[i]) return;
  const int64_t d = pos[i];
  const int32_t p = perm[i];
  for (int w = 0; w < W; ++w) ck[d * W + w] = keys[(size_t)p * W + w];
  cv[d] = vec[p];
  ca[d] = sums[i];
  co[d] = order_key[i];
}

__global__ void final_gather_kernel(int W, const int32_t* perm, int64_t n, const uint64_t* ck, const double2* ca,
                                    const int32_t* cv, uint64_t* ok, double2* oa, int32_t* ov) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t p = perm[i];
  for (int w = 0; w < W; ++w) ok[i * W + w] = ck[(size_t)p * W + w];
  oa[i] = ca[p];
  ov[i] = cv[p];
}

__global__ void vec_counts_kernel(const int32_t* ov, int64_t n, int32_t* cnt) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(&cnt[ov[i]], 1);
}

// Sequential |amp|^2 sum per vector in stored order (norm2, finite.py:2023-2033),
// then optional division by the norm (spectra.py:787-790).
__global__ void norm_kernel(const int64_t* off, int V, double2* amps, double* norm2, int normalize) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  double s = 0.0;
  for (int64_t i = off[v]; i < off[v + 1]; ++i) s += abs2_ref(amps[i]);
  norm2[v] = s;
  if (normalize && s > 0.0) {
    const double nrm = sqrt(s);
    for (int64_t i = off[v]; i < off[v + 1]; ++i) amps[i] = make_double2(amps[i].x / nrm, amps[i].y / nrm);
  }
}

__global__ void trunc_flags_kernel(const double2* evecs, const int64_t* fl_off, const int64_t* vbase, int n_out,
                                   double weight_min, int32_t* flags) {
  const int o = blockIdx.y;
  const int64_t n = fl_off[o + 1] - fl_off[o];
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n_out || i >= n) return;
  flags[fl_off[o] + i] = abs2_ref(evecs[vbase[o] + i]) >= weight_min ? 1 : 0;
}

__global__ void trunc_scatter_kernel(int W, const double2* evecs, const uint64_t* keys, const int64_t* fl_off,
                                     const int64_t* vbase, const int64_t* kbase, int n_out, const int32_t* flags,
                                     const int64_t* pos, uint64_t* ok, double2* oa) {
  const int o = blockIdx.y;
  const int64_t n = fl_off[o + 1] - fl_off[o];
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n_out || i >= n) return;
  const int64_t f = fl_off[o] + i;
  if (!flags[f]) return;
  const int64_t d = pos[f];
  for (int w = 0; w < W; ++w) ok[d * W + w] = keys[(kbase[o] + i) * W + w];
  oa[d] = evecs[vbase[o] + i];
}

// Entries of a vector set with |amp|^2 >= weight_min (the pruning of
// spectra.py:1108-1111), kept in stored order.
__global__ void prune_flags_kernel(const double2* amps, int64_t n, double weight_min, int32_t* flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[i] = abs2_ref(amps[i]) >= weight_min ? 1 : 0;
}

__global__ void prune_scatter_kernel(int W, const uint64_t* keys, const double2* amps, int64_t n,
                                     const int32_t* flags, const int64_t* pos, uint64_t* ok, double2* oa) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flags[i]) return;
  const int64_t d = pos[i];
  for (int w = 0; w < W; ++w) ok[d * W + w] = keys[i * W + w];
  oa[d] = amps[i];
}

// out[o] = exclusive-scan position of entry at[o] (total when at[o] == n).
__global__ void gather_pos_kernel(const int64_t* pos, int64_t n, int64_t total, const int64_t* at, int m,
                                  int64_t* out) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o < m) out[o] = at[o] < n ? pos[at[o]] : total;
}

struct VScratch {
  DevBuf so_off, sbeg, skeys, samps, spos, tasks, counts, offs, ekeys, eval, evec, esrc, efirst, perm, keep, sums,
      okey, pos, ck, cv, ca, co, perm2, cnt, off, flags;
};
VScratch& vscratch() {
  static VScratch v;
  return v;
}

int bits_for(int64_t n) {
  int b = 1;
  while ((int64_t(1) << b) < n) ++b;
  return b;
}

// Exclusive scan of int32 flags into int64 positions (total returned).
int64_t scan_flags(const int32_t* flags, int64_t n, int64_t* pos, hipStream_t s) { return scan_counts(flags, n, pos, s); }

// pos_at[o] = scan position (in vscratch().pos over n flags, total kept) of
// segment start seg[o], o <= V, with one gather and one copy.
void segment_positions(const std::vector<int64_t>& seg, int64_t n, int64_t kept, std::vector<int64_t>& pos_at,
                       hipStream_t s) {
  VScratch& sc = vscratch();
  const int m = (int)seg.size();
  h2d(sc.off, seg.data(), m, s);
  sc.cnt.reserve((size_t)m * 8);
  hipLaunchKernelGGL(gather_pos_kernel, dim3(blocks_for(m, 256)), dim3(256), 0, s, sc.pos.as<int64_t>(), n, kept,
                     sc.off.as<int64_t>(), m, sc.cnt.as<int64_t>());
  HIP_CHECK(hipGetLastError());
  pos_at.resize(m);
  d2h(pos_at.data(), sc.cnt.p, m, s);
  stream_sync(s);
}

}  // namespace

void vector_apply(OpsView ops, int W, const std::vector<int64_t>& src_off, const uint64_t* src_keys,
                  const double2* src_amps, int n_out, const int32_t* out_src, const int32_t* out_model,
                  double weight_min, bool normalize, std::vector<int64_t>& out_off, DevBuf& out_keys,
                  DevBuf& out_amps, DevBuf& out_norm2, hipStream_t s, int n_orb) {
  VScratch& sc = vscratch();
  // sources of output vector o are the entries of its source vector
  std::vector<int64_t> so_off(n_out + 1, 0), sbeg(n_out);
  std::vector<WaveTask> tasks;
  for (int o = 0; o < n_out; ++o) {
    const int v = out_src[o];
    const int64_t len = src_off[v + 1] - src_off[v];
    sbeg[o] = src_off[v];
    so_off[o + 1] = so_off[o] + len;
    for (int64_t b = 0; b < len; b += 64)
      tasks.push_back(WaveTask{out_model[o], o, so_off[o] + b, so_off[o] + std::min<int64_t>(b + 64, len)});
  }
  const int64_t NS = so_off[n_out];
  out_off.assign(n_out + 1, 0);
  out_norm2.reserve(std::max(n_out, 1) * 8);
  if (NS == 0) {
    out_keys.reserve(8);
    out_amps.reserve(16);
    if (n_out) HIP_CHECK(hipMemsetAsync(out_norm2.p, 0, n_out * 8, s));
    stream_sync(s);
    return;
  }
  h2d(sc.so_off, so_off.data(), n_out + 1, s);
  h2d(sc.sbeg, sbeg.data(), n_out, s);
  sc.skeys.reserve(NS * W * 8);
  sc.samps.reserve(NS * 16);
  sc.spos.reserve(NS * 4);
  hipLaunchKernelGGL(gather_sources_kernel, dim3(n_out), dim3(256), 0, s, W, sc.so_off.as<int64_t>(),
                     sc.sbeg.as<int64_t>(), n_out, src_keys, src_amps, sc.skeys.as<uint64_t>(),
                     sc.samps.as<double2>(), sc.spos.as<int32_t>());
  HIP_CHECK(hipGetLastError());
  h2d(sc.tasks, tasks.data(), tasks.size(), s);
  ApplyArgs a{};
  a.ops = ops;
  a.tasks = sc.tasks.as<WaveTask>();
  a.n_tasks = (int64_t)tasks.size();
  a.src_keys = sc.skeys.as<uint64_t>();
  a.src_amps = sc.samps.as<double2>();
  a.src_pos = sc.spos.as<int32_t>();
  a.weight_min = weight_min;
  sc.counts.reserve(NS * 4);
  a.counts = sc.counts.as<int32_t>();
  launch_apply(W, MODE_VECTOR, false, a, s);
  sc.offs.reserve((NS + 1) * 8);
  const int64_t E = scan_counts(sc.counts.as<int32_t>(), NS, sc.offs.as<int64_t>(), s);
  if (E == 0) {
    out_keys.reserve(8);
    out_amps.reserve(16);
    HIP_CHECK(hipMemsetAsync(out_norm2.p, 0, std::max(n_out, 1) * 8, s));
    stream_sync(s);
    return;
  }
  if (E >= INT32_MAX) throw HipError{"vector_apply: too many entries"};
  sc.ekeys.reserve(E * W * 8);
  sc.eval.reserve(E * 16);
  sc.evec.reserve(E * 4);
  sc.esrc.reserve(E * 4);
  sc.efirst.reserve(E * 4);
  a.offsets = sc.offs.as<int64_t>();
  a.out_keys = sc.ekeys.as<uint64_t>();
  a.out_val = sc.eval.as<double2>();
  a.out_vec = sc.evec.as<int32_t>();
  a.out_src = sc.esrc.as<int32_t>();
  a.out_first = sc.efirst.as<int32_t>();
  launch_apply(W, MODE_VECTOR, true, a, s);
  // stable sort by (vec, key): equal targets stay in source order
  sc.perm.reserve(E * 4);
  hipLaunchKernelGGL(iota_kernel, dim3(blocks_for(E, 256)), dim3(256), 0, s, sc.perm.as<int32_t>(), E);
  sort_vec_keys_stable(W, n_orb, bits_for(n_out), sc.ekeys.as<uint64_t>(), sc.evec.as<int32_t>(),
                       sc.perm.as<int32_t>(), E, s);
  sc.keep.reserve(E * 4);
  sc.sums.reserve(E * 16);
  sc.okey.reserve(E * 8);
  if (W == 1)
    hipLaunchKernelGGL(reduce_runs_kernel<1>, dim3(blocks_for(E, 256)), dim3(256), 0, s, sc.perm.as<int32_t>(), E,
                       sc.ekeys.as<uint64_t>(), sc.evec.as<int32_t>(), sc.eval.as<double2>(), sc.esrc.as<int32_t>(),
                       sc.efirst.as<int32_t>(), weight_min, sc.keep.as<int32_t>(), sc.sums.as<double2>(),
                       sc.okey.as<uint64_t>());
  else
    hipLaunchKernelGGL(reduce_runs_kernel<2>, dim3(blocks_for(E, 256)), dim3(256), 0, s, sc.perm.as<int32_t>(), E,
                       sc.ekeys.as<uint64_t>(), sc.evec.as<int32_t>(), sc.eval.as<double2>(), sc.esrc.as<int32_t>(),
                       sc.efirst.as<int32_t>(), weight_min, sc.keep.as<int32_t>(), sc.sums.as<double2>(),
                       sc.okey.as<uint64_t>());
  HIP_CHECK(hipGetLastError());
  sc.pos.reserve((E + 1) * 8);
  const int64_t NU = scan_flags(sc.keep.as<int32_t>(), E, sc.pos.as<int64_t>(), s);
  if (NU == 0) {
    out_keys.reserve(8);
    out_amps.reserve(16);
    HIP_CHECK(hipMemsetAsync(out_norm2.p, 0, std::max(n_out, 1) * 8, s));
    stream_sync(s);
    return;
  }
  sc.ck.reserve(NU * W * 8);
  sc.cv.reserve(NU * 4);
  sc.ca.reserve(NU * 16);
  sc.co.reserve(NU * 8);
  hipLaunchKernelGGL(compact_runs_kernel, dim3(blocks_for(E, 256)), dim3(256), 0, s, W, sc.perm.as<int32_t>(), E,
                     sc.keep.as<int32_t>(), sc.pos.as<int64_t>(), sc.ekeys.as<uint64_t>(), sc.evec.as<int32_t>(),
                     sc.sums.as<double2>(), sc.okey.as<uint64_t>(), sc.ck.as<uint64_t>(), sc.cv.as<int32_t>(),
                     sc.ca.as<double2>(), sc.co.as<uint64_t>());
  // insertion order: (vec, first source position, first term)
  sc.perm2.reserve(NU * 4);
  hipLaunchKernelGGL(iota_kernel, dim3(blocks_for(NU, 256)), dim3(256), 0, s, sc.perm2.as<int32_t>(), NU);
  sort_pairs_u64_u32(sc.co.as<uint64_t>(), sc.cv.as<int32_t>(), bits_for(n_out), sc.perm2.as<int32_t>(), NU, s);
  out_keys.reserve(NU * W * 8);
  out_amps.reserve(NU * 16);
  DevBuf ov;
  ov.reserve(NU * 4);
  hipLaunchKernelGGL(final_gather_kernel, dim3(blocks_for(NU, 256)), dim3(256), 0, s, W, sc.perm2.as<int32_t>(), NU,
                     sc.ck.as<uint64_t>(), sc.ca.as<double2>(), sc.cv.as<int32_t>(), out_keys.as<uint64_t>(),
                     out_amps.as<double2>(), ov.as<int32_t>());
  sc.cnt.reserve(std::max(n_out, 1) * 4);
  HIP_CHECK(hipMemsetAsync(sc.cnt.p, 0, std::max(n_out, 1) * 4, s));
  hipLaunchKernelGGL(vec_counts_kernel, dim3(blocks_for(NU, 256)), dim3(256), 0, s, ov.as<int32_t>(), NU,
                     sc.cnt.as<int32_t>());
  std::vector<int32_t> cnt(n_out);
  d2h(cnt.data(), sc.cnt.p, n_out, s);
  stream_sync(s);
  for (int o = 0; o < n_out; ++o) out_off[o + 1] = out_off[o] + cnt[o];
  h2d(sc.off, out_off.data(), n_out + 1, s);
  hipLaunchKernelGGL(norm_kernel, dim3(blocks_for(n_out, 64)), dim3(64), 0, s, sc.off.as<int64_t>(), n_out,
                     out_amps.as<double2>(), out_norm2.as<double>(), normalize ? 1 : 0);
  HIP_CHECK(hipGetLastError());
  stream_sync(s);
}

void truncate_eigvecs(int W, const uint64_t* basis_keys, const double2* evecs, int nev,
                      const std::vector<int64_t>& row_base, const std::vector<int64_t>& chain_n, int n_out,
                      const int32_t* chain, const int32_t* which, double weight_min, std::vector<int64_t>& out_off,
                      DevBuf& out_keys, DevBuf& out_amps, hipStream_t s) {
  VScratch& sc = vscratch();
  std::vector<int64_t> fl_off(n_out + 1, 0), vbase(n_out), kbase(n_out);
  int64_t nmax = 1;
  for (int o = 0; o < n_out; ++o) {
    const int c = chain[o];
    fl_off[o + 1] = fl_off[o] + chain_n[c];
    vbase[o] = row_base[c] * nev + (int64_t)which[o] * chain_n[c];
    kbase[o] = row_base[c];
    nmax = std::max(nmax, chain_n[c]);
  }
  const int64_t NF = fl_off[n_out];
  out_off.assign(n_out + 1, 0);
  if (NF == 0 || n_out == 0) {
    out_keys.reserve(8);
    out_amps.reserve(16);
    return;
  }
  DevBuf d_fl, d_vb, d_kb;
  h2d(d_fl, fl_off.data(), n_out + 1, s);
  h2d(d_vb, vbase.data(), n_out, s);
  h2d(d_kb, kbase.data(), n_out, s);
  sc.flags.reserve(NF * 4);
  dim3 grid(blocks_for(nmax, 256), n_out);
  hipLaunchKernelGGL(trunc_flags_kernel, grid, dim3(256), 0, s, evecs, d_fl.as<int64_t>(), d_vb.as<int64_t>(), n_out,
                     weight_min, sc.flags.as<int32_t>());
  HIP_CHECK(hipGetLastError());
  sc.pos.reserve((NF + 1) * 8);
  const int64_t NU = scan_flags(sc.flags.as<int32_t>(), NF, sc.pos.as<int64_t>(), s);
  out_keys.reserve(std::max<int64_t>(NU, 1) * W * 8);
  out_amps.reserve(std::max<int64_t>(NU, 1) * 16);
  hipLaunchKernelGGL(trunc_scatter_kernel, grid, dim3(256), 0, s, W, evecs, basis_keys, d_fl.as<int64_t>(),
                     d_vb.as<int64_t>(), d_kb.as<int64_t>(), n_out, sc.flags.as<int32_t>(), sc.pos.as<int64_t>(),
                     out_keys.as<uint64_t>(), out_amps.as<double2>());
  HIP_CHECK(hipGetLastError());
  // per-output counts from the scan positions at segment starts
  std::vector<int64_t> pos_at(n_out + 1);
  segment_positions(fl_off, NF, NU, pos_at, s);
  for (int o = 0; o < n_out; ++o) out_off[o + 1] = out_off[o] + (pos_at[o + 1] - pos_at[o]);
}

void prune_vec(int W, const std::vector<int64_t>& off, const uint64_t* keys, const double2* amps, double weight_min,
               std::vector<int64_t>& out_off, DevBuf& out_keys, DevBuf& out_amps, hipStream_t s) {
  VScratch& sc = vscratch();
  const int V = (int)off.size() - 1;
  const int64_t N = off.empty() ? 0 : off.back();
  out_off.assign(V + 1, 0);
  if (N == 0) {
    out_keys.reserve(8);
    out_amps.reserve(16);
    return;
  }
  sc.flags.reserve(N * 4);
  hipLaunchKernelGGL(prune_flags_kernel, dim3(blocks_for(N, 256)), dim3(256), 0, s, amps, N, weight_min,
                     sc.flags.as<int32_t>());
  HIP_CHECK(hipGetLastError());
  sc.pos.reserve((N + 1) * 8);
  const int64_t NU = scan_flags(sc.flags.as<int32_t>(), N, sc.pos.as<int64_t>(), s);
  out_keys.reserve(std::max<int64_t>(NU, 1) * W * 8);
  out_amps.reserve(std::max<int64_t>(NU, 1) * 16);
  hipLaunchKernelGGL(prune_scatter_kernel, dim3(blocks_for(N, 256)), dim3(256), 0, s, W, keys, amps, N,
                     sc.flags.as<int32_t>(), sc.pos.as<int64_t>(), out_keys.as<uint64_t>(), out_amps.as<double2>());
  HIP_CHECK(hipGetLastError());
  std::vector<int64_t> pos_at(V + 1);
  segment_positions(off, N, NU, pos_at, s);
  for (int o = 0; o < V; ++o) out_off[o + 1] = out_off[o] + (pos_at[o + 1] - pos_at[o]);
}

}  // namespace imped
