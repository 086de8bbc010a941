// Krylov-space kernels of the engine.
//
// * lanczos_kernel: get_tridiagonal_krylov_vectors (impurityModel/ed/
//   finite.py:1974-1992) for every chain of a batch in one launch, one
//   workgroup per chain running all K steps (SpMV + alpha, w update + beta,
//   scale) with deterministic block reductions.  reorth = 1 adds two
//   classical Gram-Schmidt passes against all stored Krylov vectors.
// * eigen_kernel: replaces eigsh(which="SA") of eigensystem (finite.py:105):
//   Lanczos with full reorthogonalization from a seeded random vector, the
//   lowest Ritz pairs of the tridiagonal matrix by Sturm bisection and
//   inverse iteration, checked every few steps against the residual
//   |beta_m y_k[m-1]| <= tol * max(1, |lambda_k|).
// * solve_kernel: the resolvent solves of getRIXSmap (spectra.py:1066-1080)
//   for all incoming energies of a chain from one Krylov space (shifted
//   Lanczos/FOM with full reorthogonalization), residual-checked per shift.
// * cf_kernel: the continued fraction of getGreen (spectra.py:677-680) on the
//   omega grid for every chain, fused with the Boltzmann weights of
//   thermal_average (average.py:11-45) and the sum over transition operators
//   (spectra.py:233-238).
#include <algorithm>
#include <cmath>
#include <mutex>

#include "krylov.h"

namespace imped {

namespace {

// Sum over the workgroup, identical in every thread; fixed order, so runs
// are reproducible.  A one-wave workgroup needs no LDS round trip.
template <int BT>
__device__ __forceinline__ double block_sum(double x, double* lds) {
  x = wave_sum(x);
  if (BT == 64) return x;
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = x;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int k = 0; k < BT / 64; ++k) r += lds[k];
  __syncthreads();
  return r;
}

// block_sum with one barrier: lds holds two rows of BT/64 partials used in
// turn (parity flips on every call), so a call's writes never meet the
// previous call's reads; every thread of the workgroup makes the same
// sequence of calls.
template <int BT>
__device__ __forceinline__ double block_sum1(double x, double* lds, int& parity) {
  x = wave_sum(x);
  if (BT == 64) return x;
  double* row = lds + parity * (BT / 64);
  parity ^= 1;
  if ((threadIdx.x & 63) == 0) row[threadIdx.x >> 6] = x;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int k = 0; k < BT / 64; ++k) r += row[k];
  return r;
}

// Sum over the tpr lanes of a row (consecutive lanes of a quad), the same in
// each of them; every lane of the wave active.
__device__ __forceinline__ double row_lanes_sum(double x, int tpr) {
  if (tpr > 1) x += dpp_move<0xB1, 0xF>(x);  // quad_perm [1,0,3,2]
  if (tpr > 2) x += dpp_move<0x4E, 0xF>(x);  // quad_perm [2,3,0,1]
  return x;
}

// y = H x for the rows of one chain; returns the partial Re<x, y>.
template <int BT>
__device__ __forceinline__ double spmv_rows(const int64_t* rowptr, const int32_t* col, const double2* val,
                                            int64_t base, int64_t n, const double2* x, double2* y) {
  double pa = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += BT) {
    const int64_t e0 = rowptr[base + i], e1 = rowptr[base + i + 1];
    double ax = 0.0, ay = 0.0;
    for (int64_t e = e0; e < e1; ++e) {
      const double2 h = val[e];
      const double2 v = x[col[e]];
      ax += h.x * v.x - h.y * v.y;
      ay += h.x * v.y + h.y * v.x;
    }
    y[i] = make_double2(ax, ay);
    pa += ax * x[i].x + ay * x[i].y;
  }
  return pa;
}

// Classical Gram-Schmidt helpers (defined with the eigensolver below):
// cl[l] = <V_l, w> into LDS, then w -= sum_l cl[l] V_l returning |w|^2.
template <int BT, class T>
__device__ void dots_lds(const T* V, int m, int64_t n, const T* w, double2* cl, bool self);
template <int BT, class T>
__device__ double update_norm(const T* V, int m, int64_t n, T* w, const double2* cl, double* red);

struct LanczosArgs {
  const int64_t* row_base;
  const int64_t* chain_n;
  const int64_t* rowptr;
  const int32_t* col;
  const double2* val;
  const double2* v0;
  int K;
  int reorth;
  double2* ws;      // 3 vectors per chain at 3*row_base
  double2* vstore;  // reorth: K vectors per chain at K*row_base (coefficients in LDS)
  double* alpha;
  double* beta;
  int32_t* kout;
  int32_t* status;
  const int32_t* chain_list;  // chains handled by this launch, one per workgroup
};

template <int BT>
__global__ __launch_bounds__(BT) void lanczos_kernel(LanczosArgs a) {
  __shared__ double red[BT / 64];
  extern __shared__ double2 cl[];  // reorth: K + 2 Gram-Schmidt coefficients
  const int c = a.chain_list[blockIdx.x];
  const int64_t n = a.chain_n[c];
  const int64_t base = a.row_base[c];
  double* alpha = a.alpha + (size_t)c * a.K;
  double* beta = a.beta + (size_t)c * a.K;
  if (n == 0) {
    if (threadIdx.x == 0) {
      a.kout[c] = 0;
      a.status[c] = IMPED_OK;
    }
    return;
  }
  const int K = (int)(a.K < n ? a.K : n);
  double2* v0 = a.ws + 3 * base;
  double2* v1 = v0 + n;
  double2* w = v1 + n;
  double2* V = a.reorth ? a.vstore + (size_t)a.K * base : nullptr;
  for (int64_t i = threadIdx.x; i < n; i += BT) {
    v0[i] = a.v0[base + i];
    if (V) V[i] = v0[i];
  }
  __syncthreads();
  // wp = H v0; alpha0 = Re(conj(wp) . v0); w = wp - alpha0 v0
  double al = block_sum<BT>(spmv_rows<BT>(a.rowptr, a.col, a.val, base, n, v0, w), red);
  if (threadIdx.x == 0) alpha[0] = al;
  for (int64_t i = threadIdx.x; i < n; i += BT) w[i] = make_double2(w[i].x - al * v0[i].x, w[i].y - al * v0[i].y);
  __syncthreads();
  int status = IMPED_OK;
  int kdone = K;
  for (int j = 1; j < K; ++j) {
    double p = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += BT) p += w[i].x * w[i].x + w[i].y * w[i].y;
    p = block_sum<BT>(p, red);
    if (V) {
      // full reorthogonalization against V_0..V_{j-1}; a second classical
      // Gram-Schmidt pass when the first removed more than half of |w|^2
      dots_lds<BT, double2>(V, j, n, w, cl, false);
      double p1 = update_norm<BT, double2>(V, j, n, w, cl, red);
      if (p1 < 0.5 * p) {
        dots_lds<BT, double2>(V, j, n, w, cl, false);
        p1 = update_norm<BT, double2>(V, j, n, w, cl, red);
      }
      p = p1;
    }
    const double b = sqrt(p);
    if (threadIdx.x == 0) beta[j - 1] = b;
    if (b == 0.0) {
      status = IMPED_ERR_BREAKDOWN;
      kdone = j;
      break;
    }
    for (int64_t i = threadIdx.x; i < n; i += BT) {
      v1[i] = make_double2(w[i].x / b, w[i].y / b);
      if (V) V[(size_t)j * n + i] = v1[i];
    }
    __syncthreads();
    al = block_sum<BT>(spmv_rows<BT>(a.rowptr, a.col, a.val, base, n, v1, w), red);
    if (threadIdx.x == 0) alpha[j] = al;
    for (int64_t i = threadIdx.x; i < n; i += BT) {
      const double2 x = v1[i], y = v0[i];
      w[i] = make_double2(w[i].x - al * x.x - b * y.x, w[i].y - al * x.y - b * y.y);
      v0[i] = x;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a.kout[c] = kdone;
    a.status[c] = status;
  }
}

// ---- fully reorthogonalized Lanczos with the Krylov basis in LDS
//
// For a chain whose K' = min(K, n) Krylov vectors fit in LDS next to the
// residual, one wave runs the whole recurrence on LDS: V[l] at row l of
// stride s (the smallest s >= n with s = 1 mod 8, so both the row-parallel
// update, lane i reading V[l][i], and the vector-parallel dot pass, lane l
// reading V[l][i], touch 64 distinct LDS banks), then w, then the
// Gram-Schmidt coefficients.  The dot pass gives lane l the whole product
// <V_l, w> (a sequential sum over rows), so a step needs no per-vector
// wave reductions.  The SpMV, alpha, beta and three-term update are those
// of lanczos_kernel; v_{j-1} and v_j are rows j-1 and j of V.

__device__ __forceinline__ int64_t lds_row_stride(int64_t n) { return (n + 6) / 8 * 8 + 1; }

size_t lanczos_lds_bytes(int64_t n, int K) {
  const int64_t kk = K < n ? K : n;
  const int64_t s = (n + 6) / 8 * 8 + 1;
  return (size_t)(kk * s + s + kk + 2) * sizeof(double2);
}

__global__ __launch_bounds__(64) void lanczos_lds_kernel(LanczosArgs a) {
  extern __shared__ double2 lds[];
  const int c = a.chain_list[blockIdx.x];
  const int64_t n = a.chain_n[c];
  const int64_t base = a.row_base[c];
  double* alpha = a.alpha + (size_t)c * a.K;
  double* beta = a.beta + (size_t)c * a.K;
  const int lane = threadIdx.x;
  if (n == 0) {
    if (lane == 0) {
      a.kout[c] = 0;
      a.status[c] = IMPED_OK;
    }
    return;
  }
  const int K = (int)(a.K < n ? a.K : n);
  const int64_t s = lds_row_stride(n);
  double2* V = lds;
  double2* w = V + (size_t)K * s;
  double2* cl = w + s;
  const int64_t* rowptr = a.rowptr + base;
  // The first row of every lane (row = lane) keeps up to RC of its entries
  // in registers for all K steps, so most of the SpMV reads no memory; the
  // rest of the row, and rows from 64 on, are read from memory each step.
  constexpr int RC = 16;
  double2 hv[RC];
  int hc[RC];
  int64_t r0e1 = 0, r0rest = 0;
  int hn = 0;
  if (lane < n) {
    const int64_t e0 = rowptr[lane];
    r0e1 = rowptr[lane + 1];
    hn = (int)(r0e1 - e0 < RC ? r0e1 - e0 : RC);
    r0rest = e0 + hn;
#pragma unroll
    for (int k = 0; k < RC; ++k) {
      hv[k] = k < hn ? a.val[e0 + k] : make_double2(0.0, 0.0);
      hc[k] = k < hn ? a.col[e0 + k] : 0;
    }
  }
  // y = H x and Re<x, y> over the chain's rows (x, y in LDS)
  auto spmv = [&](const double2* x, double2* y) {
    double pa = 0.0;
    if (lane < n) {
      double ax = 0.0, ay = 0.0, bx = 0.0, by = 0.0;
#pragma unroll
      for (int k = 0; k < RC; k += 2) {
        if (k < hn) {
          const double2 v = x[hc[k]];
          ax += hv[k].x * v.x - hv[k].y * v.y;
          ay += hv[k].x * v.y + hv[k].y * v.x;
        }
        if (k + 1 < hn) {
          const double2 v = x[hc[k + 1]];
          bx += hv[k + 1].x * v.x - hv[k + 1].y * v.y;
          by += hv[k + 1].x * v.y + hv[k + 1].y * v.x;
        }
      }
      for (int64_t e = r0rest; e < r0e1; ++e) {
        const double2 h = a.val[e];
        const double2 v = x[a.col[e]];
        ax += h.x * v.x - h.y * v.y;
        ay += h.x * v.y + h.y * v.x;
      }
      ax += bx;
      ay += by;
      y[lane] = make_double2(ax, ay);
      pa += ax * x[lane].x + ay * x[lane].y;
    }
    for (int64_t i = lane + 64; i < n; i += 64) {
      const int64_t e0 = rowptr[i], e1 = rowptr[i + 1];
      double ax = 0.0, ay = 0.0;
      for (int64_t e = e0; e < e1; ++e) {
        const double2 h = a.val[e];
        const double2 v = x[a.col[e]];
        ax += h.x * v.x - h.y * v.y;
        ay += h.x * v.y + h.y * v.x;
      }
      y[i] = make_double2(ax, ay);
      pa += ax * x[i].x + ay * x[i].y;
    }
    return wave_sum(pa);
  };
  // one classical Gram-Schmidt pass against V_0..V_{m-1}; returns |w|^2.
  // Two partial sums per product (even and odd terms) halve the dependent
  // chain of additions.
  auto gram_schmidt = [&](int m) {
    for (int l = lane; l < m; l += 64) {
      const double2* vl = V + (size_t)l * s;
      double cx = 0.0, cy = 0.0, dx = 0.0, dy = 0.0;
      int64_t i = 0;
#pragma unroll 2
      for (; i + 1 < n; i += 2) {
        const double2 v = vl[i], b = w[i], v1 = vl[i + 1], b1 = w[i + 1];
        cx += v.x * b.x + v.y * b.y;
        cy += v.x * b.y - v.y * b.x;
        dx += v1.x * b1.x + v1.y * b1.y;
        dy += v1.x * b1.y - v1.y * b1.x;
      }
      if (i < n) {
        const double2 v = vl[i], b = w[i];
        cx += v.x * b.x + v.y * b.y;
        cy += v.x * b.y - v.y * b.x;
      }
      cl[l] = make_double2(cx + dx, cy + dy);
    }
    __syncthreads();
    double p = 0.0;
    for (int64_t i = lane; i < n; i += 64) {
      double ux = 0.0, uy = 0.0, tx = 0.0, ty = 0.0;
      int l = 0;
#pragma unroll 4
      for (; l + 1 < m; l += 2) {
        const double2 cc = cl[l], c1 = cl[l + 1];
        const double2 av = V[(size_t)l * s + i], a1 = V[(size_t)(l + 1) * s + i];
        ux += av.x * cc.x - av.y * cc.y;
        uy += av.x * cc.y + av.y * cc.x;
        tx += a1.x * c1.x - a1.y * c1.y;
        ty += a1.x * c1.y + a1.y * c1.x;
      }
      if (l < m) {
        const double2 cc = cl[l];
        const double2 av = V[(size_t)l * s + i];
        ux += av.x * cc.x - av.y * cc.y;
        uy += av.x * cc.y + av.y * cc.x;
      }
      const double wx = w[i].x - (ux + tx), wy = w[i].y - (uy + ty);
      w[i] = make_double2(wx, wy);
      p += wx * wx + wy * wy;
    }
    __syncthreads();
    return wave_sum(p);
  };
  for (int64_t i = lane; i < n; i += 64) V[i] = a.v0[base + i];
  __syncthreads();
  // wp = H v0; alpha0 = Re(conj(wp) . v0); w = wp - alpha0 v0
  double al = spmv(V, w);
  if (lane == 0) alpha[0] = al;
  for (int64_t i = lane; i < n; i += 64) w[i] = make_double2(w[i].x - al * V[i].x, w[i].y - al * V[i].y);
  __syncthreads();
  int status = IMPED_OK;
  int kdone = K;
  for (int j = 1; j < K; ++j) {
    double p = 0.0;
    for (int64_t i = lane; i < n; i += 64) p += w[i].x * w[i].x + w[i].y * w[i].y;
    p = wave_sum(p);
    // full reorthogonalization against V_0..V_{j-1}; a second pass when
    // the first removed more than half of |w|^2 (as lanczos_kernel)
    double p1 = gram_schmidt(j);
    if (p1 < 0.5 * p) p1 = gram_schmidt(j);
    const double b = sqrt(p1);
    if (lane == 0) beta[j - 1] = b;
    if (b == 0.0) {
      status = IMPED_ERR_BREAKDOWN;
      kdone = j;
      break;
    }
    double2* vj = V + (size_t)j * s;
    const double2* vp = vj - s;
    for (int64_t i = lane; i < n; i += 64) vj[i] = make_double2(w[i].x / b, w[i].y / b);
    __syncthreads();
    al = spmv(vj, w);
    if (lane == 0) alpha[j] = al;
    for (int64_t i = lane; i < n; i += 64) {
      const double2 x = vj[i], y = vp[i];
      w[i] = make_double2(w[i].x - al * x.x - b * y.x, w[i].y - al * x.y - b * y.y);
    }
    __syncthreads();
  }
  if (lane == 0) {
    a.kout[c] = kdone;
    a.status[c] = status;
  }
}

// ---- the same recurrence with NW waves per chain
//
// lanczos_lds_kernel leaves each step's Gram-Schmidt products as one long
// sequential sum per lane (n terms for a dot, m for an update) on a single
// wave.  Here NW waves share a chain: for the dots, wave g sums rows
// [g n/NW, (g+1) n/NW) of every <V_l, w> (lane l); for the update, wave g
// sums vectors [g m/NW, (g+1) m/NW) of every row (lane), each lane adding
// the NW dot partials of the coefficients it needs itself; the NW update
// partials of each row are then subtracted from w.  Each dependent chain is
// NW times shorter and the chain's CU runs NW times as many waves.  Row t
// of the SpMV (t < 64 NW) keeps up to RC entries of H in registers.
//
// The update is owned by rows: wave g updates rows [g n/NW, (g+1) n/NW) of
// w (the rows whose dot partials it formed), each row on a quad of lanes
// that splits the vectors (lane q takes l = q, q + 4, ...; DPP adds the
// quad), with the coefficients added from the NW dot partials once per wave
// into a row of its own.  A step takes four workgroup barriers: the SpMV
// reads w scaled by 1/beta (so no pass writes v_j before it) into a second
// residual buffer, |w|^2 is posted by the three-term update and summed
// after the next barrier, and a Gram-Schmidt pass needs one barrier after
// its dots and one for the sum of |w|^2.  LDS: V (K rows of stride s), w,
// w2, NW x K dot partials and NW x K coefficients.

size_t lanczos_ldsw_bytes(int64_t n, int K, int nw) {
  const int64_t kk = K < n ? K : n;
  const int64_t s = (n + 6) / 8 * 8 + 1;
  return (size_t)(kk * s + 2 * s + 2 * nw * kk) * sizeof(double2);
}

template <int NW>
__global__ __launch_bounds__(64 * NW) void lanczos_ldsw_kernel(LanczosArgs a) {
  constexpr int T = 64 * NW;
  constexpr int RC = 16;
  extern __shared__ double2 lds[];
  __shared__ double red[2][NW];
  const int c = a.chain_list[blockIdx.x];
  const int64_t n = a.chain_n[c];
  const int64_t base = a.row_base[c];
  double* alpha = a.alpha + (size_t)c * a.K;
  double* beta = a.beta + (size_t)c * a.K;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (n == 0) {
    if (tid == 0) {
      a.kout[c] = 0;
      a.status[c] = IMPED_OK;
    }
    return;
  }
  const int K = (int)(a.K < n ? a.K : n);
  const int64_t s = lds_row_stride(n);
  double2* V = lds;
  double2* w = V + (size_t)K * s;
  double2* w2 = w + s;
  double2* dotp = w2 + s;                 // [NW][K]
  double2* coefw = dotp + (size_t)NW * K;  // [NW][K], a row per wave
  const int64_t* rowptr = a.rowptr + base;
  // workgroup sums: post() leaves each wave's partial in a row of red
  // (rows used in turn), collect() adds them after a barrier; a row is
  // written again only two posts later, past a barrier that follows every
  // collect of it
  int rsel = 0;
  auto post = [&](double x) {
    x = wave_sum(x);
    if (lane == 0) red[rsel][wv] = x;
    const int r = rsel;
    rsel ^= 1;
    return r;
  };
  auto collect = [&](int r) {
    double t = 0.0;
#pragma unroll
    for (int g = 0; g < NW; ++g) t += red[r][g];
    return t;
  };
  // a row of H on tpr = 4, 2 or 1 threads (as many as n rows leave room
  // for): thread tid takes row tid / tpr and holds up to RC of its entries
  // tid % tpr, tid % tpr + tpr, ... in registers; DPP adds the row's lanes
  const int tpr = 4 * n <= T ? 4 : (2 * n <= T ? 2 : 1);
  const int64_t hrow = tid / tpr;
  double2 hv[RC];
  int hc[RC];
  int64_t r_end = 0, r_rest = 0;
  int hn = 0;
  if (hrow < n) {
    const int64_t e0 = rowptr[hrow] + tid % tpr;
    r_end = rowptr[hrow + 1];
    const int64_t mine = r_end > e0 ? (r_end - e0 + tpr - 1) / tpr : 0;
    hn = (int)(mine < RC ? mine : RC);
    r_rest = e0 + (int64_t)hn * tpr;
#pragma unroll
    for (int k = 0; k < RC; ++k) {
      hv[k] = k < hn ? a.val[e0 + (int64_t)k * tpr] : make_double2(0.0, 0.0);
      hc[k] = k < hn ? a.col[e0 + (int64_t)k * tpr] : 0;
    }
  }
  // the wave's longest register share: groups of four entries past it are
  // skipped (a wave-uniform branch); inside a group every load is issued
  // unconditionally (entries past a row's end hold 0 at column 0)
  const int hn_wave = wave_max_int(hn);
  // y = H (sc x) over the rows of this thread, and its part of Re<sc x, y>;
  // xj (when not null) receives sc x for those rows
  auto spmv = [&](const double2* x, double sc, double2* y, double2* xj) {
    double pa = 0.0;
    double ax = 0.0, ay = 0.0, bx = 0.0, by = 0.0;
    if (hrow < n) {
#pragma unroll
      for (int k0 = 0; k0 < RC; k0 += 4) {
        if (k0 < hn_wave) {
          double2 v[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) v[u] = x[hc[k0 + u]];
#pragma unroll
          for (int u = 0; u < 4; u += 2) {
            ax += hv[k0 + u].x * v[u].x - hv[k0 + u].y * v[u].y;
            ay += hv[k0 + u].x * v[u].y + hv[k0 + u].y * v[u].x;
            bx += hv[k0 + u + 1].x * v[u + 1].x - hv[k0 + u + 1].y * v[u + 1].y;
            by += hv[k0 + u + 1].x * v[u + 1].y + hv[k0 + u + 1].y * v[u + 1].x;
          }
        }
      }
      for (int64_t e = r_rest; e < r_end; e += tpr) {
        const double2 h = a.val[e];
        const double2 v = x[a.col[e]];
        ax += h.x * v.x - h.y * v.y;
        ay += h.x * v.y + h.y * v.x;
      }
    }
    ax = row_lanes_sum(ax + bx, tpr) * sc;
    ay = row_lanes_sum(ay + by, tpr) * sc;
    if (hrow < n && tid % tpr == 0) {
      y[hrow] = make_double2(ax, ay);
      const double2 xi = make_double2(x[hrow].x * sc, x[hrow].y * sc);
      if (xj) xj[hrow] = xi;
      pa += ax * xi.x + ay * xi.y;
    }
    for (int64_t i = tid + T; i < n; i += T) {
      const int64_t e0 = rowptr[i], e1 = rowptr[i + 1];
      double ax = 0.0, ay = 0.0;
      for (int64_t e = e0; e < e1; ++e) {
        const double2 h = a.val[e];
        const double2 v = x[a.col[e]];
        ax += h.x * v.x - h.y * v.y;
        ay += h.x * v.y + h.y * v.x;
      }
      ax *= sc;
      ay *= sc;
      y[i] = make_double2(ax, ay);
      const double2 xi = make_double2(x[i].x * sc, x[i].y * sc);
      if (xj) xj[i] = xi;
      pa += ax * xi.x + ay * xi.y;
    }
    return pa;
  };
  // one classical Gram-Schmidt pass of w against V_0..V_{m-1} (w complete
  // on entry); returns |w|^2 and leaves w complete
  auto gram_schmidt = [&](int m) {
    const int64_t r0 = n * wv / NW, r1 = n * (wv + 1) / NW;
    for (int l = lane; l < m; l += 64) {
      const double2* vl = V + (size_t)l * s;
      double cx = 0.0, cy = 0.0, dx = 0.0, dy = 0.0;
      int64_t i = r0;
#pragma unroll 4
      for (; i + 1 < r1; i += 2) {
        const double2 v = vl[i], b = w[i], v1 = vl[i + 1], b1 = w[i + 1];
        cx += v.x * b.x + v.y * b.y;
        cy += v.x * b.y - v.y * b.x;
        dx += v1.x * b1.x + v1.y * b1.y;
        dy += v1.x * b1.y - v1.y * b1.x;
      }
      if (i < r1) {
        const double2 v = vl[i], b = w[i];
        cx += v.x * b.x + v.y * b.y;
        cy += v.x * b.y - v.y * b.x;
      }
      dotp[(size_t)wv * K + l] = make_double2(cx + dx, cy + dy);
    }
    __syncthreads();
    // the coefficients, added from the NW dot partials by each wave into its
    // own row (read back by the same wave only, after its own writes)
    double2* cw = coefw + (size_t)wv * K;
    for (int l = lane; l < m; l += 64) {
      double2 t = dotp[l];
#pragma unroll
      for (int g = 1; g < NW; ++g) {
        const double2 u = dotp[(size_t)g * K + l];
        t.x += u.x;
        t.y += u.y;
      }
      cw[l] = t;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const int q = lane & 3;
    double p = 0.0;
    for (int64_t rb = r0; rb < r1; rb += 16) {
      const int64_t i = rb + (lane >> 2);
      const bool on = i < r1;
      const int64_t ii = on ? i : r0;  // idle lanes read a valid row, their sums are dropped
      double ux = 0.0, uy = 0.0, tx = 0.0, ty = 0.0;
      int l = q;
#pragma unroll 2
      for (; l + 4 < m; l += 8) {
        const double2 c0 = cw[l], c1 = cw[l + 4];
        const double2 v0 = V[(size_t)l * s + ii], v1 = V[(size_t)(l + 4) * s + ii];
        ux += v0.x * c0.x - v0.y * c0.y;
        uy += v0.x * c0.y + v0.y * c0.x;
        tx += v1.x * c1.x - v1.y * c1.y;
        ty += v1.x * c1.y + v1.y * c1.x;
      }
      if (l < m) {
        const double2 c0 = cw[l];
        const double2 v0 = V[(size_t)l * s + ii];
        ux += v0.x * c0.x - v0.y * c0.y;
        uy += v0.x * c0.y + v0.y * c0.x;
      }
      ux += tx;
      uy += ty;
      ux += dpp_move<0xB1, 0xF>(ux);  // quad_perm [1,0,3,2]
      uy += dpp_move<0xB1, 0xF>(uy);
      ux += dpp_move<0x4E, 0xF>(ux);  // quad_perm [2,3,0,1]
      uy += dpp_move<0x4E, 0xF>(uy);
      if (on && q == 0) {
        const double wx = w[i].x - ux, wy = w[i].y - uy;
        w[i] = make_double2(wx, wy);
        p += wx * wx + wy * wy;
      }
    }
    const int r = post(p);
    __syncthreads();
    return collect(r);
  };
  for (int64_t i = tid; i < n; i += T) V[i] = a.v0[base + i];
  __syncthreads();
  // w2 = H v0; alpha0 = Re(conj(w2) . v0); w = w2 - alpha0 v0
  int r = post(spmv(V, 1.0, w2, nullptr));
  __syncthreads();
  double al = collect(r);
  if (tid == 0) alpha[0] = al;
  double p = 0.0;
  for (int64_t i = tid; i < n; i += T) {
    const double wx = w2[i].x - al * V[i].x, wy = w2[i].y - al * V[i].y;
    w[i] = make_double2(wx, wy);
    p += wx * wx + wy * wy;
  }
  int rp = post(p);
  __syncthreads();
  int status = IMPED_OK;
  int kdone = K;
  for (int j = 1; j < K; ++j) {
    const double pw = collect(rp);
    // full reorthogonalization against V_0..V_{j-1}; a second pass when
    // the first removed more than half of |w|^2 (as lanczos_kernel)
    double p1 = gram_schmidt(j);
    if (p1 < 0.5 * pw) p1 = gram_schmidt(j);
    const double b = sqrt(p1);
    if (tid == 0) beta[j - 1] = b;
    if (b == 0.0) {
      status = IMPED_ERR_BREAKDOWN;
      kdone = j;
      break;
    }
    double2* vj = V + (size_t)j * s;
    const double2* vp = vj - s;
    // v_j = w / beta is formed inside the SpMV (rows by their owners)
    r = post(spmv(w, 1.0 / b, w2, vj));
    __syncthreads();
    al = collect(r);
    if (tid == 0) alpha[j] = al;
    p = 0.0;
    for (int64_t i = tid; i < n; i += T) {
      const double2 x = vj[i], y = vp[i];
      const double wx = w2[i].x - al * x.x - b * y.x, wy = w2[i].y - al * x.y - b * y.y;
      w[i] = make_double2(wx, wy);
      p += wx * wx + wy * wy;
    }
    rp = post(p);
    __syncthreads();
  }
  if (tid == 0) {
    a.kout[c] = kdone;
    a.status[c] = status;
  }
}

// ------------------------------------------------------------------ eigen

__device__ __forceinline__ uint64_t splitmix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__device__ __forceinline__ double u01(uint64_t x) { return (double)(splitmix(x) >> 11) * (1.0 / 9007199254740992.0); }

// 1/x to within an ulp: the hardware reciprocal estimate refined by two
// Newton steps (a short dependent chain, where a correctly rounded division
// is a long one).
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// Sturm counts at P shifts from the leading principal minors of T - x:
// d_0 = 1, d_1 = a_0 - x, d_{i+1} = (a_i - x) d_i - b_{i-1}^2 d_{i-1}; the
// number of eigenvalues of the m x m tridiagonal (a, b) below x is the
// number of sign changes in d_0, ..., d_m.  There is no division, so the
// dependent chain of a step is one FMA (the quotient form d_i / d_{i-1}
// needs a reciprocal per step); the P recurrences are interleaved.  Every 8
// steps the pair (d_{i-1}, d_i) is scaled by the power of two that brings
// the larger to [0.5, 1): exact, so no sign or ratio changes, and within 8
// steps the minors grow or shrink by at most (1 + |a - x| + b^2)^8 or b^16,
// far from the exponent range.  A minor that is exactly zero counts by its
// sign bit: its neighbours have opposite signs (b != 0 inside the chain), so
// the two changes around it count one either way.
template <int P>
__device__ __forceinline__ void sturm_counts(const double* a, const double* b, int m, const double* x, int* cnt) {
  double d0[P], d1[P];  // d_{i-1}, d_i
#pragma unroll
  for (int p = 0; p < P; ++p) {
    d0[p] = 1.0;
    d1[p] = a[0] - x[p];
    cnt[p] = (int)((unsigned)__double2hiint(d1[p]) >> 31);
  }
  auto step = [&](int i) {
    const double ai = a[i], bb = b[i - 1] * b[i - 1];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const double dn = fma(ai - x[p], d1[p], -bb * d0[p]);
      cnt[p] += (int)((unsigned)(__double2hiint(dn) ^ __double2hiint(d1[p])) >> 31);
      d0[p] = d1[p];
      d1[p] = dn;
    }
  };
  int i = 1;
  for (; i + 8 <= m; i += 8) {
#pragma unroll
    for (int u = 0; u < 8; ++u) step(i + u);
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int e = max(__builtin_amdgcn_frexp_exp(d0[p]), __builtin_amdgcn_frexp_exp(d1[p]));
      d0[p] = __builtin_amdgcn_ldexp(d0[p], -e);
      d1[p] = __builtin_amdgcn_ldexp(d1[p], -e);
    }
  }
  for (; i < m; ++i) step(i);
}

// Inverse iteration for the eigenvector of (a, b) at lambda; LU with partial
// pivoting of the shifted tridiagonal.  y and work (4m + m ints) are
// thread-private (LDS or global).  Every recurrence carries its running
// values in registers, so no iteration waits on a store it just made.
__device__ __forceinline__ void tri_inverse_iteration(const double* __restrict__ a, const double* __restrict__ b, int m, double lambda,
                                      double* __restrict__ y, double* __restrict__ dd, double* __restrict__ uu,
                                      double* __restrict__ ww, double* __restrict__ ll, int* __restrict__ piv,
                                      uint64_t seed) {
  double scale = 0.0;
#pragma unroll 8
  for (int i = 0; i < m; ++i) scale = fmax(scale, fabs(a[i]) + (i < m - 1 ? fabs(b[i]) : 0.0));
  const double tiny = 2.2e-16 * (scale > 0 ? scale : 1.0);
  // (d, u): diagonal and superdiagonal of the current row after elimination
  double d = a[0] - lambda, u = m > 1 ? b[0] : 0.0;
#pragma unroll 8
  for (int i = 0; i < m - 1; ++i) {
    const double sub = b[i];
    const double dn = a[i + 1] - lambda;
    const double un = i + 1 < m - 1 ? b[i + 1] : 0.0;
    // rows i and i+1 interchanged when |d| < |sub|; both cases by selects
    // (the lanes of different eigenvalues choose differently, and branches
    // would run both cases one after the other), one reciprocal
    const bool sw = !(fabs(d) >= fabs(sub));
    const double pv = sw ? sub : (d == 0.0 ? tiny : d);
    const double mult = (sw ? d : sub) * rcp_nr(pv);
    dd[i] = pv;
    uu[i] = sw ? dn : u;
    ww[i] = sw ? un : 0.0;
    ll[i] = mult;
    piv[i] = sw ? 1 : 0;
    const double d_keep = dn - mult * u, d_swap = u - mult * dn;
    const double u_swap = -mult * un;
    d = sw ? d_swap : d_keep;
    u = sw ? u_swap : un;
  }
  if (d == 0.0) d = tiny;
  dd[m - 1] = d;
  uu[m - 1] = 0.0;
  ww[m - 1] = 0.0;
  // back substitution multiplies by the guarded reciprocal pivots
#pragma unroll 8
  for (int i = 0; i < m; ++i) {
    double dv = dd[i];
    if (fabs(dv) < tiny) dv = dv < 0 ? -tiny : tiny;
    dd[i] = rcp_nr(dv);
  }
#pragma unroll 8
  for (int i = 0; i < m; ++i) y[i] = 0.5 + u01(seed + (uint64_t)i);
  for (int it = 0; it < 3; ++it) {
    // forward: L^-1 with the interchanges; yi is the running y[i]
    double yi = y[0];
#pragma unroll 8
    for (int i = 0; i < m - 1; ++i) {
      const double y1 = y[i + 1];
      const bool sw = piv[i] != 0;
      const double top = sw ? y1 : yi, bot = sw ? yi : y1;
      y[i] = top;
      yi = bot - ll[i] * top;
    }
    y[m - 1] = yi;
    // back substitution with U (bandwidth 3); y1, y2 = y[i+1], y[i+2]
    double y1 = 0.0, y2 = 0.0, nrm = 0.0;
#pragma unroll 8
    for (int i = m - 1; i >= 0; --i) {
      double r = y[i];
      if (i + 1 < m) r -= uu[i] * y1;
      if (i + 2 < m) r -= ww[i] * y2;
      const double v = r * dd[i];
      y[i] = v;
      nrm += v * v;
      y2 = y1;
      y1 = v;
    }
    const double inv = 1.0 / sqrt(nrm);
#pragma unroll 8
    for (int i = 0; i < m; ++i) y[i] *= inv;
  }
}

struct EigenArgs {
  const int64_t* row_base;
  const int64_t* chain_n;
  const int64_t* rowptr;
  const int32_t* col;
  const double2* val;
  int nev;
  int max_steps;
  double tol;
  uint64_t seed;
  void* vstore;        // per chain: msteps(c) * n vectors (double or double2), offset voff[c]
  const int64_t* voff;
  void* wbuf;          // per chain: n at row_base
  double2* coef;       // per chain: max_steps
  double* ab;          // per chain: 2*max_steps (alpha then beta)
  double* yk;          // per chain: nev*max_steps
  double* tw;          // per chain: nev * 5 * max_steps
  int* tpiv;           // per chain: nev * max_steps
  double* evals;
  double2* evecs;
  int32_t* steps;
  int32_t* conv;
  const int32_t* chain_list;
  int dgks;          // 1: second Gram-Schmidt pass only when needed
  int check_every;   // Ritz check interval (steps)
  int check_start;   // no check before this many steps (except at the last step)
  int lstride;       // LDS arrays: Gram-Schmidt coefficients (lstride + 2), then, when
  int lds_ab;        //   lds_ab = lds_work = 1, alpha/beta (2 lstride) and per eigenvalue y
  int lds_work;      //   (lstride), LU work (4 lstride) and pivots (lstride ints); lstride >= every
                     //   chain's steps
  int kl;            // with the LDS Krylov prefix: vectors kept in LDS (after w, 16-byte aligned)
};

// Element helpers: the eigensolver runs on complex (double2) vectors, or on
// real (double) vectors when H is real -- then a real start vector keeps
// every Krylov vector real and halves the memory traffic.
__device__ __forceinline__ double re_(double x) { return x; }
__device__ __forceinline__ double re_(double2 x) { return x.x; }
__device__ __forceinline__ double im_(double) { return 0.0; }
__device__ __forceinline__ double im_(double2 x) { return x.y; }
template <class T>
__device__ __forceinline__ T mk(double r, double i);
template <>
__device__ __forceinline__ double mk<double>(double r, double) {
  return r;
}
template <>
__device__ __forceinline__ double2 mk<double2>(double r, double i) {
  return make_double2(r, i);
}
template <class T>
constexpr bool is_real_v = sizeof(T) == sizeof(double);

// y = H x on one chain (T-typed vectors); returns the partial Re<x, y>.
template <int BT, class T>
__device__ __forceinline__ double spmv_t(const int64_t* rowptr, const int32_t* col, const double2* val, int64_t base,
                                         int64_t n, const T* x, T* y) {
  double pa = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += BT) {
    const int64_t e0 = rowptr[base + i], e1 = rowptr[base + i + 1];
    double ax = 0.0, ay = 0.0;
    // four nonzeros per round trip: the column and value loads issue
    // together (indices past the row clamped, their products skipped), the
    // products are added in row order
    for (int64_t e = e0; e < e1; e += 4) {
      int32_t cc[4];
      double2 hh[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int64_t k = e + u < e1 ? e + u : e1 - 1;
        cc[u] = col[k];
        hh[u] = val[k];
      }
      T vv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) vv[u] = x[cc[u]];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (e + u < e1) {
          const double2 h = hh[u];
          const T v = vv[u];
          if constexpr (is_real_v<T>) {
            ax += h.x * re_(v);
          } else {
            ax += h.x * re_(v) - h.y * im_(v);
            ay += h.x * im_(v) + h.y * re_(v);
          }
        }
      }
    }
    y[i] = mk<T>(ax, ay);
    pa += ax * re_(x[i]) + ay * im_(x[i]);
  }
  return pa;
}

// One classical Gram-Schmidt pass of w against V[0..m): coefficients by
// wave-level dot products, then one update sweep.
template <int BT, class T>
__device__ void cgs_pass(const T* V, int m, int64_t n, T* w, double2* coef) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  for (int l = wave; l < m; l += BT / 64) {
    const T* vl = V + (size_t)l * n;
    double cx = 0.0, cy = 0.0;
    for (int64_t i = lane; i < n; i += 64) {
      const T a = vl[i], b = w[i];
      cx += re_(a) * re_(b) + im_(a) * im_(b);  // conj(a) * b
      if constexpr (!is_real_v<T>) cy += re_(a) * im_(b) - im_(a) * re_(b);
    }
    cx = wave_sum(cx);
    if constexpr (!is_real_v<T>) cy = wave_sum(cy);
    if (lane == 0) coef[l] = make_double2(cx, cy);
  }
  __syncthreads();
  for (int64_t i = threadIdx.x; i < n; i += BT) {
    double wx = re_(w[i]), wy = im_(w[i]);
    for (int l = 0; l < m; ++l) {
      const double2 c = coef[l];
      const T a = V[(size_t)l * n + i];
      if constexpr (is_real_v<T>) {
        wx -= re_(a) * c.x;
      } else {
        wx -= re_(a) * c.x - im_(a) * c.y;
        wy -= re_(a) * c.y + im_(a) * c.x;
      }
    }
    w[i] = mk<T>(wx, wy);
  }
  __syncthreads();
}

template <int BT, class T>
__device__ __forceinline__ double norm2_t(const T* w, int64_t n, double* red) {
  double p = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += BT) p += re_(w[i]) * re_(w[i]) + im_(w[i]) * im_(w[i]);
  return block_sum<BT>(p, red);
}

constexpr int MAX_NEV = 16;

// cl[l] = <V_l, w> for l < m (and cl[m] = <w, w> when self), eight vectors
// per wave at a time so w is read once per eight dot products (and eight
// loads are in flight together); results in LDS.
template <int BT, class T>
__device__ void dots_lds(const T* V, int m, int64_t n, const T* w, double2* cl, bool self) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  constexpr int NW = BT / 64;
  constexpr int U = 8;
  const int total = m + (self ? 1 : 0);
  for (int l0 = wave * U; l0 < total; l0 += NW * U) {
    // loads of all eight vectors issue together: indices past m read V_{m-1}
    // (or w for the self product), sums past total are discarded
    double cx[U], cy[U];
    size_t off[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      cx[u] = cy[u] = 0.0;
      off[u] = (size_t)(l0 + u < m ? l0 + u : (m > 0 ? m - 1 : 0)) * n;
    }
    for (int64_t i = lane; i < n; i += 64) {
      const T b = w[i];
      T av[U];
#pragma unroll
      for (int u = 0; u < U; ++u) av[u] = m > 0 ? V[off[u] + i] : b;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (l0 + u >= m) av[u] = b;  // the self product (or a discarded sum)
        cx[u] += re_(av[u]) * re_(b) + im_(av[u]) * im_(b);
        if constexpr (!is_real_v<T>) cy[u] += re_(av[u]) * im_(b) - im_(av[u]) * re_(b);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const double x = wave_sum(cx[u]);
      const double y = is_real_v<T> ? 0.0 : wave_sum(cy[u]);
      if (lane == 0 && l0 + u < total) cl[l0 + u] = make_double2(x, y);
    }
  }
  __syncthreads();
}

// w -= sum_l cl[l] V_l over l < m; returns |w|^2 (same in every thread).
template <int BT, class T>
__device__ double update_norm(const T* V, int m, int64_t n, T* w, const double2* cl, double* red) {
  double p = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += BT) {
    double wx = re_(w[i]), wy = im_(w[i]);
    // unrolled so the loads of sixteen Krylov entries are in flight together
#pragma unroll 16
    for (int l = 0; l < m; ++l) {
      const double2 c = cl[l];
      const T av = V[(size_t)l * n + i];
      if constexpr (is_real_v<T>) {
        wx -= re_(av) * c.x;
      } else {
        wx -= re_(av) * c.x - im_(av) * c.y;
        wy -= re_(av) * c.y + im_(av) * c.x;
      }
    }
    w[i] = mk<T>(wx, wy);
    p += wx * wx + wy * wy;
  }
  return block_sum<BT>(p, red);
}

// Lowest nev eigenvalues of the symmetric tridiagonal (a, b) of size m by
// multisection: the workgroup splits into nev groups; each evaluates Sturm
// counts at G interior points of its interval per round and keeps the
// sub-interval where the count crosses k, shrinking it G+1 times per round.
// hint (or null): the same eigenvalues of a leading block of (a, b).  By
// Cauchy interlacing they bound the new ones from above, so a first round of
// points hint - scale 10^(-16 r / (G-2)) (r < G-1) and hint + 1e-14 scale
// brackets each new value within about its distance from the hint, and the
// ordinary rounds start from that bracket (the Gershgorin interval when the
// hint does not bracket it).
template <int BT>
__device__ __forceinline__ void ritz_values(const double* a, const double* b, int m, int nev, const double* hint, double* s_lam,
                            double* s_lo, double* s_hi, int* s_first, int* s_flag) {
  __shared__ double s_part[3][BT / 64];
  __shared__ double s_bounds[3];
  {
    double lo = INFINITY, hi = -INFINITY, sc = 0.0;
    for (int i = threadIdx.x; i < m; i += BT) {
      const double r = (i > 0 ? fabs(b[i - 1]) : 0.0) + (i < m - 1 ? fabs(b[i]) : 0.0);
      lo = fmin(lo, a[i] - r);
      hi = fmax(hi, a[i] + r);
      sc = fmax(sc, fabs(a[i]) + r);
    }
    for (int off = 32; off > 0; off >>= 1) {
      lo = fmin(lo, __shfl_xor(lo, off));
      hi = fmax(hi, __shfl_xor(hi, off));
      sc = fmax(sc, __shfl_xor(sc, off));
    }
    if ((threadIdx.x & 63) == 0) {
      s_part[0][threadIdx.x >> 6] = lo;
      s_part[1][threadIdx.x >> 6] = hi;
      s_part[2][threadIdx.x >> 6] = sc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int wv = 1; wv < BT / 64; ++wv) {
        lo = fmin(lo, s_part[0][wv]);
        hi = fmax(hi, s_part[1][wv]);
        sc = fmax(sc, s_part[2][wv]);
      }
      s_bounds[0] = lo - 1e-14 * sc - 1e-300;
      s_bounds[1] = hi + 1e-14 * sc + 1e-300;
      s_bounds[2] = sc;
    }
  }
  __syncthreads();
  const int G = BT / nev;
  const int g = threadIdx.x / G, r = threadIdx.x % G;
  const double scale = s_bounds[2];
  if ((int)threadIdx.x < nev) {
    s_lo[threadIdx.x] = s_bounds[0];
    s_hi[threadIdx.x] = s_bounds[1];
  }
  if (hint != nullptr) {
    auto hint_point = [&](int k, int rr) {
      return rr < G - 1 ? hint[k] - scale * exp10(-16.0 * (double)rr / (double)(G - 2)) : hint[k] + 1e-14 * scale;
    };
    if ((int)threadIdx.x < nev) s_first[threadIdx.x] = G;
    __syncthreads();
    if (g < nev) {
      double x = hint_point(g, r);
      int cnt;
      sturm_counts<1>(a, b, m, &x, &cnt);
      if (cnt > g) atomicMin(&s_first[g], r);
    }
    __syncthreads();
    if ((int)threadIdx.x < nev) {
      const int k = threadIdx.x, f = s_first[k];
      if (f < G) {  // otherwise the hint does not bound the value: Gershgorin interval
        if (f > 0) s_lo[k] = fmax(s_lo[k], hint_point(k, f - 1));
        s_hi[k] = fmin(s_hi[k], hint_point(k, f));
      }
    }
  }
  __syncthreads();
  for (int round = 0; round < 64; ++round) {
    if ((int)threadIdx.x < nev) s_first[threadIdx.x] = G;
    if (threadIdx.x == 0) *s_flag = 0;
    __syncthreads();
    if (g < nev) {
      const double lo = s_lo[g], hi = s_hi[g];
      double x = lo + (hi - lo) * (double)(r + 1) / (double)(G + 1);
      int cnt;
      sturm_counts<1>(a, b, m, &x, &cnt);
      if (cnt > g) atomicMin(&s_first[g], r);
    }
    __syncthreads();
    if ((int)threadIdx.x < nev) {
      const int k = threadIdx.x;
      const double lo = s_lo[k], hi = s_hi[k];
      const int f = s_first[k];
      const double nlo = f == 0 ? lo : lo + (hi - lo) * (double)f / (double)(G + 1);
      const double nhi = f == G ? hi : lo + (hi - lo) * (double)(f + 1) / (double)(G + 1);
      s_lo[k] = nlo;
      s_hi[k] = nhi;
      const double mid = 0.5 * (nlo + nhi);
      if (nhi - nlo > 2.2e-16 * fmax(fabs(nlo), fabs(nhi)) + 1e-300 && mid > nlo && mid < nhi) atomicOr(s_flag, 1);
    }
    __syncthreads();
    if (*s_flag == 0) break;
    __syncthreads();
  }
  if ((int)threadIdx.x < nev) s_lam[threadIdx.x] = 0.5 * (s_lo[threadIdx.x] + s_hi[threadIdx.x]);
  __syncthreads();
}

// Orthonormalize the inverse-iteration vectors of clustered eigenvalues
// (|lam_k - lam_q| <= 1e-3 max(1, |lam_k|)) in order, by wave 0 of the
// workgroup (lane-parallel dot products and updates).
__device__ __forceinline__ void ortho_clustered(double* Y, int ys, int m, int ne, const double* lam) {
  if (threadIdx.x >= 64) return;
  const int lane = threadIdx.x;
  for (int k = 1; k < ne; ++k) {
    double* yk = Y + (size_t)k * ys;
    for (int q = 0; q < k; ++q) {
      if (fabs(lam[k] - lam[q]) > 1e-3 * fmax(1.0, fabs(lam[k]))) continue;
      const double* yq = Y + (size_t)q * ys;
      double d = 0.0;
      for (int i = lane; i < m; i += 64) d += yq[i] * yk[i];
      d = wave_sum(d);
      double nn = 0.0;
      for (int i = lane; i < m; i += 64) {
        yk[i] -= d * yq[i];
        nn += yk[i] * yk[i];
      }
      const double inv = 1.0 / sqrt(wave_sum(nn));
      for (int i = lane; i < m; i += 64) yk[i] *= inv;
    }
  }
}

// ---- eigensolver step pieces over a split Krylov basis: vectors l < kl at
// VL + l n (LDS), the others at VG + l n (memory).  Each piece is inlined at
// call sites whose pointers have one known address space, and sums in the
// same order whatever the split, so the results do not depend on it.

// w = H vj - alpha vj - bprev vp (vp unused when !has_prev); returns |w|^2, alpha in *al.
template <int BT, class T>
__device__ __forceinline__ double three_term(const EigenArgs& a, int64_t base, int64_t n, const T* vj, const T* vp,
                                             bool has_prev, double bprev, T* w, double* red, double* al_out) {
  const double al = block_sum<BT>(spmv_t<BT, T>(a.rowptr, a.col, a.val, base, n, vj, w), red);
  double p0 = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += BT) {
    double xr = re_(w[i]) - al * re_(vj[i]);
    double xi = im_(w[i]) - al * im_(vj[i]);
    if (has_prev) {
      xr -= bprev * re_(vp[i]);
      xi -= bprev * im_(vp[i]);
    }
    w[i] = mk<T>(xr, xi);
    p0 += xr * xr + xi * xi;
  }
  *al_out = al;
  return block_sum<BT>(p0, red);
}

// The rows of H kept in registers.  A row gets tpr = 4, 2 or 1 threads (as
// many as n rows leave room for in the workgroup); thread t takes row
// t / tpr and holds up to HR of its entries t % tpr, t % tpr + tpr, ...
// (the rest of the row, and rows from BT on when tpr = 1, are read from
// memory).  The tpr partial products of a row are added by DPP.
template <int BT>
constexpr int hrow_len() {
  return BT <= 256 ? 16 : (BT <= 512 ? 8 : 4);
}
template <int HR>
struct HRow {
  static_assert(HR % 4 == 0, "register rows are read in groups of four");
  double2 v[HR];
  int c[HR];
  int cnt = 0, cnt_wave = 0;  // entries held; the wave's largest cnt
  int64_t rest = 0, end = 0;  // this thread's next entry past the registers; the row's end
  int tpr = 1;
};

template <int BT>
__device__ __forceinline__ int threads_per_row(int64_t n) {
  return 4 * n <= BT ? 4 : (2 * n <= BT ? 2 : 1);
}

template <int BT, int HR>
__device__ __forceinline__ void load_hrow(HRow<HR>& h, const int64_t* rowptr, const int32_t* col, const double2* val,
                                          int64_t base, int64_t n) {
  h.tpr = threads_per_row<BT>(n);
  const int64_t r = threadIdx.x / h.tpr;
  const int sub = threadIdx.x % h.tpr;
  if (r < n) {
    const int64_t e0 = rowptr[base + r] + sub;
    h.end = rowptr[base + r + 1];
    const int64_t mine = h.end > e0 ? (h.end - e0 + h.tpr - 1) / h.tpr : 0;
    h.cnt = (int)(mine < HR ? mine : HR);
    h.rest = e0 + (int64_t)h.cnt * h.tpr;
#pragma unroll
    for (int k = 0; k < HR; ++k) {
      h.v[k] = k < h.cnt ? val[e0 + (int64_t)k * h.tpr] : make_double2(0.0, 0.0);
      h.c[k] = k < h.cnt ? col[e0 + (int64_t)k * h.tpr] : 0;
    }
  }
  h.cnt_wave = wave_max_int(h.cnt);
}

// cl[l] = <V_l, w> for l in [l0, l0 + U) ∩ [l0, lend), one wave (l0 and
// lend wave-uniform).  Every lane loads all U vectors' entries (indices
// past lend clamped to lend - 1, their sums discarded), so the loads issue
// together instead of one bounds-checked branch at a time.  Each vector's
// sum is formed the same way whatever U.
template <int U, class T>
__device__ __forceinline__ void dots_wave(const T* V, int l0, int lend, int64_t n, const T* w, double2* cl) {
  const int lane = threadIdx.x & 63;
  double cx[U], cy[U];
  size_t off[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    cx[u] = cy[u] = 0.0;
    off[u] = (size_t)(l0 + u < lend ? l0 + u : lend - 1) * n;
  }
  for (int64_t i = lane; i < n; i += 64) {
    const T b = w[i];
    T av[U];
#pragma unroll
    for (int u = 0; u < U; ++u) av[u] = V[off[u] + i];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      cx[u] += re_(av[u]) * re_(b) + im_(av[u]) * im_(b);
      if constexpr (!is_real_v<T>) cy[u] += re_(av[u]) * im_(b) - im_(av[u]) * re_(b);
    }
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const double x = wave_sum(cx[u]);
    const double y = is_real_v<T> ? 0.0 : wave_sum(cy[u]);
    if (lane == 0 && l0 + u < lend) cl[l0 + u] = make_double2(x, y);
  }
}

template <int U, int NW, class T>
__device__ __forceinline__ void dots_chunks(const T* VL, int ma, const T* VG, int m, int64_t n, const T* w,
                                            double2* cl, int wave) {
  for (int l0 = wave * U; l0 < ma; l0 += NW * U) dots_wave<U, T>(VL, l0, ma, n, w, cl);
  for (int l0 = ma + wave * U; l0 < m; l0 += NW * U) dots_wave<U, T>(VG, l0, m, n, w, cl);
}

// cl[l] = <V_l, w> for l < m over the split basis; results in LDS.  Chunks
// of 2, 4 or 8 vectors per wave, the smallest that covers the m vectors in
// one round of the waves, so short bases keep every wave busy (each wave's
// cost is its loads plus two reductions per vector of its chunk).
template <int BT, class T>
__device__ __forceinline__ void dots_split(const T* VL, int kl, const T* VG, int m, int64_t n, const T* w,
                                           double2* cl) {
  constexpr int NW = BT / 64;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ma = m < kl ? m : kl;
  if (m <= 2 * NW)
    dots_chunks<2, NW, T>(VL, ma, VG, m, n, w, cl, wave);
  else if (m <= 4 * NW)
    dots_chunks<4, NW, T>(VL, ma, VG, m, n, w, cl, wave);
  else
    dots_chunks<8, NW, T>(VL, ma, VG, m, n, w, cl, wave);
  __syncthreads();
}

// w -= sum_l cl[l] V_l over l < m of the split basis; returns |w|^2.
template <int BT, class T>
__device__ __forceinline__ double update_split(const T* VL, int kl, const T* VG, int m, int64_t n, T* w,
                                               const double2* cl, double* red) {
  const int ma = m < kl ? m : kl;
  double p = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += BT) {
    double wx = re_(w[i]), wy = im_(w[i]);
#pragma unroll 16
    for (int l = 0; l < ma; ++l) {
      const double2 c = cl[l];
      const T av = VL[(size_t)l * n + i];
      if constexpr (is_real_v<T>) {
        wx -= re_(av) * c.x;
      } else {
        wx -= re_(av) * c.x - im_(av) * c.y;
        wy -= re_(av) * c.y + im_(av) * c.x;
      }
    }
#pragma unroll 16
    for (int l = ma; l < m; ++l) {
      const double2 c = cl[l];
      const T av = VG[(size_t)l * n + i];
      if constexpr (is_real_v<T>) {
        wx -= re_(av) * c.x;
      } else {
        wx -= re_(av) * c.x - im_(av) * c.y;
        wy -= re_(av) * c.y + im_(av) * c.x;
      }
    }
    w[i] = mk<T>(wx, wy);
    p += wx * wx + wy * wy;
  }
  return block_sum<BT>(p, red);
}

// x = sum_l y[l] V_l over l < m of the split basis, for one row i.
template <class T>
__device__ __forceinline__ void combine_split(const T* VL, int64_t sL, int kl, const T* VG, int m, int64_t n,
                                              int64_t i, const double* y, double* xr, double* xi) {
  const int ma = m < kl ? m : kl;
  double r = 0.0, q = 0.0;
#pragma unroll 8
  for (int l = 0; l < ma; ++l) {
    const T v = VL[(size_t)l * sL + i];
    r += y[l] * re_(v);
    q += y[l] * im_(v);
  }
#pragma unroll 8
  for (int l = ma; l < m; ++l) {
    const T v = VG[(size_t)l * n + i];
    r += y[l] * re_(v);
    q += y[l] * im_(v);
  }
  *xr = r;
  *xi = q;
}

// Dots of the Gram-Schmidt pass over the split basis with the work shared
// by the waves through partial sums in LDS (part: BT/64 rows of n), no
// per-vector wave reductions: wave g sums rows [g n/NW, (g+1) n/NW) of
// <V_l, w> for every l (lane l), the NW partials of each l are then added
// into cl.  The update (update_rows) is owned by rows, so it needs no
// partial sums.  The LDS prefix (l < kl) has row stride sL, the rest of the
// basis n.  The
// prefix and the rest are read in separate loops, each through a pointer
// whose address space the compiler knows, so the prefix is read with LDS
// instructions (a pointer that may point to either would be read with flat
// loads, which wait on both the memory and the LDS counters).

// <v, w> over rows [r0, r1); v and w each in one address space
template <class T>
__device__ __forceinline__ double2 dot_rows(const T* __restrict__ v, const T* __restrict__ w, int64_t r0, int64_t r1) {
  double cx = 0.0, cy = 0.0, dx = 0.0, dy = 0.0;
  int64_t i = r0;
#pragma unroll 4
  for (; i + 1 < r1; i += 2) {
    const T a = v[i], b = w[i], a1 = v[i + 1], b1 = w[i + 1];
    cx += re_(a) * re_(b) + im_(a) * im_(b);
    dx += re_(a1) * re_(b1) + im_(a1) * im_(b1);
    if constexpr (!is_real_v<T>) {
      cy += re_(a) * im_(b) - im_(a) * re_(b);
      dy += re_(a1) * im_(b1) - im_(a1) * re_(b1);
    }
  }
  if (i < r1) {
    const T a = v[i], b = w[i];
    cx += re_(a) * re_(b) + im_(a) * im_(b);
    if constexpr (!is_real_v<T>) cy += re_(a) * im_(b) - im_(a) * re_(b);
  }
  return make_double2(cx + dx, cy + dy);
}

template <int BT, class T>
__device__ __forceinline__ void dots_part(const T* VL, int64_t sL, int kl, const T* VG, int m, int64_t n, const T* w,
                                          double2* part, double2* cl) {
  constexpr int NW = BT / 64;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t r0 = n * wv / NW, r1 = n * (wv + 1) / NW;
  for (int l = lane; l < m; l += 64) {
    double2 d;
    if (l < kl)
      d = dot_rows(VL + (size_t)l * sL, w, r0, r1);
    else
      d = dot_rows(VG + (size_t)l * n, w, r0, r1);
    part[(size_t)wv * n + l] = d;
  }
  __syncthreads();
  for (int l = threadIdx.x; l < m; l += BT) {
    double2 t = part[l];
#pragma unroll
    for (int g = 1; g < NW; ++g) {
      const double2 u = part[(size_t)g * n + l];
      t.x += u.x;
      t.y += u.y;
    }
    cl[l] = t;
  }
  __syncthreads();
}

// Update of the Gram-Schmidt pass by rows: wave g owns rows [g n/NW,
// (g+1) n/NW) of w, in groups of 16 rows, each row on a quad of lanes that
// splits the vectors (lane q of the quad takes l = q, q + 4, ...); the quad
// adds its four sums by DPP and its lane 0 writes the row.  No partial sums
// cross waves, so the pass ends with the barrier of the norm's sum.
template <int BT, class T>
__device__ __forceinline__ double update_rows(const T* VL, int64_t sL, int kl, const T* VG, int m, int64_t n, T* w,
                                              const double2* cl, double* red2, int& parity) {
  constexpr int NW = BT / 64;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t r0 = n * wv / NW, r1 = n * (wv + 1) / NW;
  const int q = lane & 3;
  const int lm = m < kl ? m : kl;           // vectors [0, lm) in the LDS prefix
  const int lg = kl + ((q - kl) & 3);       // first l >= kl with l = q (mod 4)
  double p = 0.0;
  for (int64_t rb = r0; rb < r1; rb += 16) {
    const int64_t i = rb + (lane >> 2);
    const bool on = i < r1;
    const int64_t ii = on ? i : r0;  // idle lanes read a valid row, their sums are dropped
    double ux = 0.0, uy = 0.0, tx = 0.0, ty = 0.0;
    int l = q;
#pragma unroll 2
    for (; l + 4 < lm; l += 8) {
      const double2 c0 = cl[l], c1 = cl[l + 4];
      const T v0 = VL[(size_t)l * sL + ii], v1 = VL[(size_t)(l + 4) * sL + ii];
      ux += re_(v0) * c0.x - im_(v0) * c0.y;
      tx += re_(v1) * c1.x - im_(v1) * c1.y;
      if constexpr (!is_real_v<T>) {
        uy += re_(v0) * c0.y + im_(v0) * c0.x;
        ty += re_(v1) * c1.y + im_(v1) * c1.x;
      }
    }
    if (l < lm) {
      const double2 c0 = cl[l];
      const T v0 = VL[(size_t)l * sL + ii];
      ux += re_(v0) * c0.x - im_(v0) * c0.y;
      if constexpr (!is_real_v<T>) uy += re_(v0) * c0.y + im_(v0) * c0.x;
    }
    for (int lv = lg; lv < m; lv += 4) {
      const double2 c0 = cl[lv];
      const T v0 = VG[(size_t)lv * n + ii];
      ux += re_(v0) * c0.x - im_(v0) * c0.y;
      if constexpr (!is_real_v<T>) uy += re_(v0) * c0.y + im_(v0) * c0.x;
    }
    ux += tx;
    uy += ty;
    ux += dpp_move<0xB1, 0xF>(ux);  // quad_perm [1,0,3,2]
    ux += dpp_move<0x4E, 0xF>(ux);  // quad_perm [2,3,0,1]
    if constexpr (!is_real_v<T>) {
      uy += dpp_move<0xB1, 0xF>(uy);
      uy += dpp_move<0x4E, 0xF>(uy);
    }
    if (on && q == 0) {
      const double wx = re_(w[i]) - ux, wy = im_(w[i]) - uy;
      w[i] = mk<T>(wx, wy);
      p += wx * wx + wy * wy;
    }
  }
  return block_sum1<BT>(p, red2, parity);
}

// First half of a Lanczos step on the LDS layout: y = H v_j into y (the
// rows of H from registers, HRow; entries in groups of four up to the
// wave's longest register share, every load of a group issued together,
// entries past a row's end holding 0 at column 0), with v_j = x * s read
// from x = w (s = 1/beta_{j-1}; the start vector itself at j = 0, s = 1),
// the first thread of each row also storing the row of v_j at vdst (null
// at j = 0).  Returns the partial Re<v_j, y>.
template <int BT, int HR, class T, class VP>
__device__ __forceinline__ double spmv_scaled(const HRow<HR>& h, const int64_t* rowptr, const int32_t* col,
                                              const double2* val, int64_t base, int64_t n, const T* x, double sc,
                                              VP* vdst, T* y) {
  double pa = 0.0;
  auto row_out = [&](int64_t i, double ax, double ay) {
    const T xi = x[i];
    const double vr = re_(xi) * sc, vi = im_(xi) * sc;
    if (vdst) vdst[i] = mk<T>(vr, vi);
    ax *= sc;
    ay *= sc;
    y[i] = mk<T>(ax, ay);
    pa += ax * vr + ay * vi;
  };
  const int64_t r = threadIdx.x / h.tpr;
  double ax = 0.0, ay = 0.0, bx = 0.0, by = 0.0;
  if (r < n) {
#pragma unroll
    for (int k0 = 0; k0 < HR; k0 += 4) {
      if (k0 < h.cnt_wave) {
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = x[h.c[k0 + u]];
#pragma unroll
        for (int u = 0; u < 4; u += 2) {
          const double2 h0 = h.v[k0 + u], h1 = h.v[k0 + u + 1];
          ax += h0.x * re_(v[u]) - h0.y * im_(v[u]);
          bx += h1.x * re_(v[u + 1]) - h1.y * im_(v[u + 1]);
          if constexpr (!is_real_v<T>) {
            ay += h0.x * im_(v[u]) + h0.y * re_(v[u]);
            by += h1.x * im_(v[u + 1]) + h1.y * re_(v[u + 1]);
          }
        }
      }
    }
    for (int64_t e = h.rest; e < h.end; e += h.tpr) {
      const double2 hh = val[e];
      const T v = x[col[e]];
      ax += hh.x * re_(v) - hh.y * im_(v);
      if constexpr (!is_real_v<T>) ay += hh.x * im_(v) + hh.y * re_(v);
    }
  }
  ax = row_lanes_sum(ax + bx, h.tpr);
  if constexpr (!is_real_v<T>) ay = row_lanes_sum(ay + by, h.tpr);
  if (r < n && (int)threadIdx.x % h.tpr == 0) row_out(r, ax, ay);
  for (int64_t i = threadIdx.x + BT; i < n; i += BT) {
    const int64_t e0 = rowptr[base + i], e1 = rowptr[base + i + 1];
    double cx = 0.0, cy = 0.0;
    for (int64_t e = e0; e < e1; ++e) {
      const double2 hh = val[e];
      const T v = x[col[e]];
      cx += hh.x * re_(v) - hh.y * im_(v);
      if constexpr (!is_real_v<T>) cy += hh.x * im_(v) + hh.y * re_(v);
    }
    row_out(i, cx, cy);
  }
  return pa;
}

// LW: the tridiagonal and the Ritz work arrays live in LDS (a compile-time
// choice, so every access to them is an LDS instruction), else in memory.
// LV (with LW): the residual w and the first a.kl Krylov vectors live in LDS
// too, so the Gram-Schmidt passes over them wait on LDS, not on memory.
template <int BT, class T, bool LW, bool LV>
__global__ __launch_bounds__(BT) void eigen_kernel(EigenArgs a) {
  __shared__ double red[BT / 64];
  __shared__ double red2[2 * (BT / 64)];  // block_sum1's alternating rows
  int parity = 0;
  __shared__ int s_done;
  __shared__ double s_lam[MAX_NEV], s_lo[MAX_NEV], s_hi[MAX_NEV];
  __shared__ int s_first[MAX_NEV];
  __shared__ int s_flag;
  __shared__ double s_hint[MAX_NEV];  // Ritz values of the previous check
  __shared__ int s_full;
  int hint_ne = -1;
  extern __shared__ double2 cl[];  // lstride + 2 coefficients, then the optional arrays
  const int c = a.chain_list[blockIdx.x];
  const int64_t n = a.chain_n[c];
  const int64_t base = a.row_base[c];
  const int nev = a.nev;
  const int ms = a.max_steps;
  if (n == 0) {
    if (threadIdx.x == 0) {
      a.steps[c] = 0;
      a.conv[c] = 1;
    }
    return;
  }
  const int mmax = (int)(ms < n ? ms : n);
  T* VG = reinterpret_cast<T*>(a.vstore) + a.voff[c];
  double* alpha = a.ab + (size_t)c * 2 * ms;
  double* beta = alpha + ms;
  // tridiagonal and Ritz work arrays: LDS when they fit (lstride >= mmax)
  const int L = a.lstride;
  double* tail = reinterpret_cast<double*>(cl + L + 2);
  double* sa;
  double* sb;
  double* Y;
  double* twb;
  int* pivb;
  int ys;  // stride of eigenvalue k's arrays
  if constexpr (LW) {
    sa = tail;
    sb = tail + L;
    ys = L;
    Y = tail + 2 * L;
    twb = Y + (size_t)nev * L;
    pivb = reinterpret_cast<int*>(twb + (size_t)nev * 4 * L);
  } else {
    sa = alpha;
    sb = beta;
    ys = ms;
    Y = a.yk + (size_t)c * nev * ms;
    twb = a.tw + (size_t)c * nev * 5 * ms;
    pivb = a.tpiv + (size_t)c * nev * ms;
  }
  // split Krylov basis and residual: LDS prefix of kl vectors when LV
  T* VL;
  T* w;
  int kl;
  int64_t sL = n;           // row stride of the LDS prefix
  double2* part = nullptr;  // with LV: Gram-Schmidt partial sums, BT/64 rows of n
  if constexpr (LV) {
    const size_t off = ((size_t)((char*)(pivb + (size_t)nev * L) - (char*)cl) + 15) & ~(size_t)15;
    w = reinterpret_cast<T*>(reinterpret_cast<char*>(cl) + off);
    VL = w + 2 * n;  // after w and H v_j
    kl = a.kl;
    sL = lds_row_stride(n);
    // byte offset from the LDS base (an integer round trip of the address
    // would hide from the compiler that part is in LDS)
    const size_t poff = (off + ((size_t)2 * n + (size_t)kl * sL) * sizeof(T) + 15) & ~(size_t)15;
    part = reinterpret_cast<double2*>(reinterpret_cast<char*>(cl) + poff);
  } else {
    w = reinterpret_cast<T*>(a.wbuf) + base;
    VL = VG;
    kl = 0;
  }
  // with the LDS layout (small chains), each thread's row of H in registers
  HRow<hrow_len<BT>()> hrow;
  if constexpr (LV) load_hrow<BT>(hrow, a.rowptr, a.col, a.val, base, n);
  // seeded random start vector (real when T is real)
  {
    auto start = [&](T* V0) {
      double p = 0.0;
      for (int64_t i = threadIdx.x; i < n; i += BT) {
        const uint64_t h = a.seed * 0x100000001B3ull + (uint64_t)c * 0x9E3779B97F4A7C15ull + 2 * (uint64_t)i;
        const double xr = u01(h) - 0.5, xi = is_real_v<T> ? 0.0 : u01(h + 1) - 0.5;
        V0[i] = mk<T>(xr, xi);
        p += xr * xr + xi * xi;
      }
      const double nrm = sqrt(block_sum<BT>(p, red));
      for (int64_t i = threadIdx.x; i < n; i += BT) V0[i] = mk<T>(re_(V0[i]) / nrm, im_(V0[i]) / nrm);
    };
    if (kl > 0)
      start(VL);
    else
      start(VG);
  }
  __syncthreads();
  int m = 0;
  int converged = 0;
  // no Ritz check before min(check_start, n/2) steps (the last step is always checked)
  const int cstart = (int)(a.check_start < n / 2 ? a.check_start : n / 2);
  // Ritz pairs of T_m with m steps done and beta_m = b: eigenvalues by
  // workgroup multisection, then thread k refines eigenvector k by inverse
  // iteration; true when the chain is done (converged, or the last step)
  auto ritz_check = [&](bool last, double b) -> bool {
    // Ritz pairs of T_m: eigenvalues by workgroup multisection, then
    // thread k refines eigenvector k by inverse iteration
    const int ne = nev < m ? nev : m;
    const bool have_prev = ne == hint_ne;
    ritz_values<BT>(sa, sb, m, ne, have_prev ? s_hint : nullptr, s_lam, s_lo, s_hi, s_first, &s_flag);
    // The eigenvector work is skipped while some Ritz value still moved by
    // more than tol max(1, |lam|) since the previous check: its residual
    // cannot be below tol yet (the Ritz value error is of the order of the
    // squared residual over the gap), so the pair is checked next time.
    if (threadIdx.x == 0) s_full = (last || have_prev) ? 1 : 0;
    __syncthreads();
    if (!last && have_prev && (int)threadIdx.x < ne &&
        !(fabs(s_lam[threadIdx.x] - s_hint[threadIdx.x]) <= a.tol * fmax(1.0, fabs(s_lam[threadIdx.x]))))
      atomicAnd(&s_full, 0);
    __syncthreads();
    if ((int)threadIdx.x < ne) s_hint[threadIdx.x] = s_lam[threadIdx.x];
    hint_ne = ne;
    if (threadIdx.x == 0) s_done = 1;
    __syncthreads();
    if (s_full) {
      if ((int)threadIdx.x < nev && (int)threadIdx.x < m) {
        const int k = threadIdx.x;
        const double lam = s_lam[k];
        double* y = Y + (size_t)k * ys;
        double* tw = twb + (size_t)k * (LW ? 4 * ys : 5 * ys);
        int* piv = pivb + (size_t)k * ys;
        tri_inverse_iteration(sa, sb, m, lam, y, tw, tw + ys, tw + 2 * ys, tw + 3 * ys, piv,
                              a.seed + 7919 * (uint64_t)k);
        a.evals[(size_t)c * nev + k] = lam;
        const double res = fabs(b * y[m - 1]);
        if (!(res <= a.tol * fmax(1.0, fabs(lam)))) atomicAnd(&s_done, 0);
      }
      __syncthreads();
      // orthogonalize inverse-iteration vectors of clustered eigenvalues
      ortho_clustered(Y, ys, m, ne, s_lam);
      __syncthreads();
      if (s_done || last) {
        converged = s_done;
        return true;
      }
    }
    return false;
  };
  auto check_due = [&](bool last) { return last || (m >= nev && m >= cstart && m % a.check_every == 0); };
  if constexpr (LV) {
    // Lanczos step on the LDS layout with five workgroup barriers: the sum
    // of alpha after the SpMV, which also forms and stores v_j = w /
    // beta_{j-1} (the start vector at j = 0); the sum of |w|^2 after the
    // three-term update of each thread's own rows; the dots of the
    // Gram-Schmidt pass and their partial sums; the sum of |w|^2 after its
    // row-owned update.  Every thread has alpha and beta from the sums; the
    // tridiagonal goes to LDS for the Ritz check, which synchronizes first.
    constexpr int HR = hrow_len<BT>();
    T* w2 = w + n;  // H v_j
    double bprev = 0.0;
    for (int j = 0; j < mmax; ++j) {
      // three-term step: w = H v_j - alpha_j v_j - beta_{j-1} v_{j-1}, with
      // the big cancellation done before any Gram-Schmidt coefficient is
      // formed (coefficients taken from H v_j itself would carry errors of
      // order eps*|H v_j| and lose orthogonality geometrically)
      double pa;
      if (j == 0)
        pa = kl > 0 ? spmv_scaled<BT, HR, T, T>(hrow, a.rowptr, a.col, a.val, base, n, VL, 1.0, nullptr, w2)
                    : spmv_scaled<BT, HR, T, T>(hrow, a.rowptr, a.col, a.val, base, n, VG, 1.0, nullptr, w2);
      else if (j < kl)
        pa = spmv_scaled<BT, HR, T, T>(hrow, a.rowptr, a.col, a.val, base, n, w, 1.0 / bprev, VL + (size_t)j * sL, w2);
      else
        pa = spmv_scaled<BT, HR, T, T>(hrow, a.rowptr, a.col, a.val, base, n, w, 1.0 / bprev, VG + (size_t)j * n, w2);
      const double al = block_sum1<BT>(pa, red2, parity);
      double p0 = 0.0;
      auto three = [&](const T* vj, const T* vp, bool has_prev) {
        for (int64_t i = threadIdx.x; i < n; i += BT) {
          const T y = w2[i], v = vj[i];
          double xr = re_(y) - al * re_(v), xi = im_(y) - al * im_(v);
          if (has_prev) {
            const T u = vp[i];
            xr -= bprev * re_(u);
            xi -= bprev * im_(u);
          }
          w[i] = mk<T>(xr, xi);
          p0 += xr * xr + xi * xi;
        }
      };
      if (j == 0) {
        if (kl > 0)
          three(VL, VL, false);
        else
          three(VG, VG, false);
      } else if (j < kl) {
        three(VL + (size_t)j * sL, VL + (size_t)(j - 1) * sL, true);
      } else if (j == kl) {
        three(VG + (size_t)j * n, VL + (size_t)(j - 1) * sL, true);
      } else {
        three(VG + (size_t)j * n, VG + (size_t)(j - 1) * n, true);
      }
      p0 = block_sum1<BT>(p0, red2, parity);
      // full reorthogonalization against V_0..V_j; a second pass when the
      // first removed more than half of |w|^2 (Daniel-Gragg-Kaufman-Stewart),
      // always when the criterion is switched off
      dots_part<BT, T>(VL, sL, kl, VG, j + 1, n, w, part, cl);
      double p1 = update_rows<BT, T>(VL, sL, kl, VG, j + 1, n, w, cl, red2, parity);
      if (!a.dgks || p1 < 0.5 * p0) {
        dots_part<BT, T>(VL, sL, kl, VG, j + 1, n, w, part, cl);
        p1 = update_rows<BT, T>(VL, sL, kl, VG, j + 1, n, w, cl, red2, parity);
      }
      const double b = sqrt(p1);
      m = j + 1;
      const bool last = (m == mmax) || b <= 1e-13 * (fabs(al) + bprev + 1e-300);
      if (threadIdx.x == 0) {
        alpha[j] = al;
        sa[j] = al;
        beta[j] = b;
        sb[j] = b;
      }
      if (check_due(last)) {
        __syncthreads();
        if (ritz_check(last, b)) break;
      }
      if (last) break;
      bprev = b;
    }
  } else {
    for (int j = 0; j < mmax; ++j) {
      // three-term step (as above), then full reorthogonalization
      const double bprev = j > 0 ? sb[j - 1] : 0.0;
      double al, p0;
      if (j == 0)
        p0 = three_term<BT, T>(a, base, n, VG, VG, false, 0.0, w, red, &al);
      else
        p0 = three_term<BT, T>(a, base, n, VG + (size_t)j * n, VG + (size_t)(j - 1) * n, true, bprev, w, red, &al);
      dots_split<BT, T>(VL, kl, VG, j + 1, n, w, cl);
      double p1 = update_split<BT, T>(VL, kl, VG, j + 1, n, w, cl, red);
      if (!a.dgks || p1 < 0.5 * p0) {
        dots_split<BT, T>(VL, kl, VG, j + 1, n, w, cl);
        p1 = update_split<BT, T>(VL, kl, VG, j + 1, n, w, cl, red);
      }
      if (threadIdx.x == 0) {
        alpha[j] = al;
        sa[j] = al;
      }
      const double b = sqrt(p1);
      m = j + 1;
      const bool last = (m == mmax) || b <= 1e-13 * (fabs(al) + bprev + 1e-300);
      if (threadIdx.x == 0) {
        beta[j] = b;
        sb[j] = b;
      }
      __syncthreads();
      if (check_due(last) && ritz_check(last, b)) break;
      if (last) break;
      const double rb = 1.0 / b;
      T* vn = VG + (size_t)(j + 1) * n;
      for (int64_t i = threadIdx.x; i < n; i += BT) vn[i] = mk<T>(re_(w[i]) * rb, im_(w[i]) * rb);
      __syncthreads();
    }
  }
  // Ritz vectors x_k = sum_l y_k[l] V_l, normalized, stored complex
  double2* X = a.evecs + base * nev;
  const int kk = nev < m ? nev : m;
  for (int k = 0; k < kk; ++k) {
    const double* y = Y + (size_t)k * ys;
    double pn = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += BT) {
      double xr, xi;
      combine_split<T>(VL, sL, kl, VG, m, n, i, y, &xr, &xi);
      X[(size_t)k * n + i] = make_double2(xr, xi);
      pn += xr * xr + xi * xi;
    }
    const double nn = sqrt(block_sum<BT>(pn, red));
    for (int64_t i = threadIdx.x; i < n; i += BT)
      X[(size_t)k * n + i] = make_double2(X[(size_t)k * n + i].x / nn, X[(size_t)k * n + i].y / nn);
    __syncthreads();
  }
  for (int k = kk; k < nev; ++k) {
    for (int64_t i = threadIdx.x; i < n; i += BT) X[(size_t)k * n + i] = make_double2(0.0, 0.0);
    if (threadIdx.x == 0) a.evals[(size_t)c * nev + k] = NAN;
  }
  if (threadIdx.x == 0) {
    a.steps[c] = m;
    a.conv[c] = converged;
  }
}

// ------------------------------------------------------------------ dense eigensolver
//
// Lowest nev eigenpairs of a small chain without a Krylov iteration.  H goes
// into LDS as its packed lower triangle and is reduced to a real symmetric
// tridiagonal T = Q^H H Q by n - 1 Householder reflectors H_k = I - tau_k
// v_k v_k^H (v_k[0] = 1, acting on rows k+1..n-1), each mapping column k
// below the diagonal to beta_k e_1 with beta_k real (beta = -sign(Re x_0)
// |x|, tau = (beta - x_0) / beta, v = x / (x_0 - beta)).  Each reflector
// updates the trailing block as A <- A - v w^H - w v^H with y = A v and
// w = tau y - |tau|^2 (v^H y) v / 2 (v^H y is real for Hermitian A).  The
// eigenvalues of T come from the Sturm multisection and its eigenvectors z
// from the inverse iteration of the Krylov eigensolver's Ritz pairs; the
// eigenvectors of H are x = H_0 H_1 ... H_{n-2} z.  The cost is fixed
// (n - 1 reflectors, no convergence test), and degenerate eigenvalues come
// out exactly, with orthogonal vectors.
constexpr int kDenseMaxN = 112;

// Offset of (i, j), i >= j, in the column-major packed lower triangle.
__device__ __forceinline__ int dense_packed(int i, int j, int n) { return j * (2 * n - j + 1) / 2 + (i - j); }

template <int BT>
__global__ __launch_bounds__(BT) void dense_eigen_kernel(EigenArgs a) {
  __shared__ double red2[2 * (BT / 64)];
  int parity = 0;
  __shared__ double s_lam[MAX_NEV], s_lo[MAX_NEV], s_hi[MAX_NEV];
  __shared__ int s_first[MAX_NEV];
  __shared__ int s_flag;
  extern __shared__ double2 dsm[];
  const int c = a.chain_list[blockIdx.x];
  const int n = (int)a.chain_n[c];
  const int64_t base = a.row_base[c];
  const int nev = a.nev;
  if (n == 0) {
    if (threadIdx.x == 0) {
      a.steps[c] = 0;
      a.conv[c] = 1;
    }
    return;
  }
  const int np = n * (n + 1) / 2;
  double2* A = dsm;                                                  // packed lower triangle
  double2* yv = A + np;                                              // A v
  double2* wv = yv + n;                                              // w
  double2* tau = wv + n;                                             // tau_k
  double2* X = tau + n;                                              // nev vectors of H
  double* d = reinterpret_cast<double*>(X + (size_t)nev * n);        // diagonal of T
  double* e = d + n;                                                 // off-diagonal of T
  double* Y = e + n;                                                 // nev vectors of T
  double* tw = Y + (size_t)nev * n;                                  // inverse iteration work, 4n each
  int* piv = reinterpret_cast<int*>(tw + (size_t)nev * 4 * n);       // pivots, n each
  for (int q = threadIdx.x; q < np; q += BT) A[q] = make_double2(0.0, 0.0);
  __syncthreads();
  // the lower triangle from each row's own entries (one thread per row, so
  // no two threads write one element)
  for (int i = threadIdx.x; i < n; i += BT) {
    const int64_t e0 = a.rowptr[base + i], e1 = a.rowptr[base + i + 1];
    for (int64_t q = e0; q < e1; ++q) {
      const int j = a.col[q];
      if (j >= 0 && j <= i) {
        double2& t = A[dense_packed(i, j, n)];
        t.x += a.val[q].x;
        t.y += a.val[q].y;
      }
    }
  }
  __syncthreads();
  constexpr int TPR = BT >= 512 ? 4 : 2;  // threads per row of A v (BT / TPR >= kDenseMaxN rows)
  for (int k = 0; k + 1 < n; ++k) {
    const int m = n - k - 1;
    double2* col = A + dense_packed(k + 1, k, n);  // x = column k below the diagonal, then v
    const double2 x0 = col[0];
    const double dk = A[dense_packed(k, k, n)].x;
    double p = 0.0;
    for (int i = 1 + threadIdx.x; i < m; i += BT) p += col[i].x * col[i].x + col[i].y * col[i].y;
    const double xn2 = block_sum1<BT>(p, red2, parity);  // its barrier also orders the reads of x0 before v
    const bool refl = !(xn2 == 0.0 && x0.y == 0.0);
    double beta = x0.x;
    double2 tk = make_double2(0.0, 0.0), sc = make_double2(0.0, 0.0);
    if (refl) {
      const double nr = sqrt(x0.x * x0.x + x0.y * x0.y + xn2);
      beta = x0.x >= 0.0 ? -nr : nr;
      tk = make_double2((beta - x0.x) / beta, -x0.y / beta);
      const double ur = x0.x - beta, ui = x0.y, den = ur * ur + ui * ui;
      sc = make_double2(ur / den, -ui / den);
    }
    if (threadIdx.x == 0) {
      d[k] = dk;
      e[k] = beta;
      tau[k] = tk;
    }
    if (!refl) continue;  // the same in every thread
    for (int i = threadIdx.x; i < m; i += BT) {
      const double2 x = col[i];
      col[i] = i == 0 ? make_double2(1.0, 0.0) : make_double2(x.x * sc.x - x.y * sc.y, x.x * sc.y + x.y * sc.x);
    }
    __syncthreads();
    {
      // y = A v over the trailing block, TPR threads per row
      const int r = threadIdx.x / TPR, q = threadIdx.x % TPR;
      double yr = 0.0, yi = 0.0;
      if (r < m) {
        const int gr = k + 1 + r;
        // columns up to the diagonal: row gr of the lower triangle
#pragma unroll 4
        for (int s = q; s <= r; s += TPR) {
          const double2 h = A[dense_packed(gr, k + 1 + s, n)], v = col[s];
          yr += h.x * v.x - h.y * v.y;
          yi += h.x * v.y + h.y * v.x;
        }
        // columns past it: column gr below the diagonal, conjugated
        const int cg = dense_packed(gr, gr, n) - r;  // A[cg + s] = A[k+1+s][gr] for s > r
#pragma unroll 4
        for (int s = r + 1 + q; s < m; s += TPR) {
          const double2 h = A[cg + s], v = col[s];
          yr += h.x * v.x + h.y * v.y;
          yi += h.x * v.y - h.y * v.x;
        }
      }
      yr = row_lanes_sum(yr, TPR);
      yi = row_lanes_sum(yi, TPR);
      if (r < m && q == 0) yv[r] = make_double2(yr, yi);
    }
    __syncthreads();
    double pv = 0.0;
    for (int i = threadIdx.x; i < m; i += BT) {
      const double2 v = col[i], y = yv[i];
      pv += v.x * y.x + v.y * y.y;
    }
    const double hv = 0.5 * (tk.x * tk.x + tk.y * tk.y) * block_sum1<BT>(pv, red2, parity);
    for (int i = threadIdx.x; i < m; i += BT) {
      const double2 y = yv[i], v = col[i];
      wv[i] = make_double2(tk.x * y.x - tk.y * y.y - hv * v.x, tk.x * y.y + tk.y * y.x - hv * v.y);
    }
    __syncthreads();
    // A -= v w^H + w v^H on the trailing lower triangle: rows r and m-1-r
    // (r+1 and m-r elements) to one group of TU threads, v_r and w_r in
    // registers
    {
      constexpr int TU = BT / 64;  // 64 row pairs cover kDenseMaxN rows
      const int pr = threadIdx.x / TU, q = threadIdx.x % TU;
      if (pr < (m + 1) / 2) {
        for (int half = 0; half < 2; ++half) {
          const int r = half == 0 ? pr : m - 1 - pr;
          if (half == 1 && r == pr) break;  // the middle row of an odd m
          const double2 vr = col[r], wr = wv[r];
          const int gr = k + 1 + r;
#pragma unroll 4
          for (int s = q; s <= r; s += TU) {
            const double2 vs = col[s], ws = wv[s];
            double2& t = A[dense_packed(gr, k + 1 + s, n)];
            t.x -= vr.x * ws.x + vr.y * ws.y + wr.x * vs.x + wr.y * vs.y;
            t.y -= vr.y * ws.x - vr.x * ws.y + wr.y * vs.x - wr.x * vs.y;
          }
        }
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) d[n - 1] = A[dense_packed(n - 1, n - 1, n)].x;
  __syncthreads();
  const int ne = nev < n ? nev : n;
  ritz_values<BT>(d, e, n, ne, nullptr, s_lam, s_lo, s_hi, s_first, &s_flag);
  __syncthreads();
  if ((int)threadIdx.x < ne) {
    const int k = threadIdx.x;
    double* t = tw + (size_t)k * 4 * n;
    tri_inverse_iteration(d, e, n, s_lam[k], Y + (size_t)k * n, t, t + n, t + 2 * n, t + 3 * n, piv + (size_t)k * n,
                          a.seed + 7919 * (uint64_t)k);
    a.evals[(size_t)c * nev + k] = s_lam[k];
  }
  __syncthreads();
  ortho_clustered(Y, n, n, ne, s_lam);
  __syncthreads();
  // x = H_0 H_1 ... H_{n-2} z, one wave per vector (wave sums, no barriers)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double2* out = a.evecs + base * nev;
  for (int kv = wave; kv < ne; kv += BT / 64) {
    double2* x = X + (size_t)kv * n;
    const double* z = Y + (size_t)kv * n;
    for (int i = lane; i < n; i += 64) x[i] = make_double2(z[i], 0.0);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (int k = n - 2; k >= 0; --k) {
      const double2 t = tau[k];
      if (t.x == 0.0 && t.y == 0.0) continue;
      const int m = n - k - 1;
      const double2* v = A + dense_packed(k + 1, k, n);
      double2* xs = x + k + 1;
      double pr = 0.0, pi = 0.0;  // v^H x
      for (int i = lane; i < m; i += 64) {
        const double2 vi = v[i], xi = xs[i];
        pr += vi.x * xi.x + vi.y * xi.y;
        pi += vi.x * xi.y - vi.y * xi.x;
      }
      pr = wave_sum(pr);
      pi = wave_sum(pi);
      const double gr = t.x * pr - t.y * pi, gi = t.x * pi + t.y * pr;  // tau v^H x
      for (int i = lane; i < m; i += 64) {
        const double2 vi = v[i];
        double2 xi = xs[i];
        xi.x -= vi.x * gr - vi.y * gi;
        xi.y -= vi.x * gi + vi.y * gr;
        xs[i] = xi;
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    double nn = 0.0;
    for (int i = lane; i < n; i += 64) nn += x[i].x * x[i].x + x[i].y * x[i].y;
    const double inv = 1.0 / sqrt(wave_sum(nn));
    for (int i = lane; i < n; i += 64) out[(size_t)kv * n + i] = make_double2(x[i].x * inv, x[i].y * inv);
  }
  for (int k = ne; k < nev; ++k) {
    for (int i = threadIdx.x; i < n; i += BT) out[(size_t)k * n + i] = make_double2(0.0, 0.0);
    if (threadIdx.x == 0) a.evals[(size_t)c * nev + k] = NAN;
  }
  if (threadIdx.x == 0) {
    a.steps[c] = n;
    a.conv[c] = 1;
  }
}

// LDS bytes of dense_eigen_kernel for an n-row chain (layout as in the kernel).
size_t dense_eigen_bytes(int64_t n, int nev) {
  const size_t np = (size_t)n * (size_t)(n + 1) / 2, nn = (size_t)n, ne = (size_t)nev;
  return 16 * (np + 3 * nn + ne * nn) + 8 * (2 * nn + 5 * ne * nn) + 4 * ne * nn;
}

// ---- fully reorthogonalized Lanczos of a small chain as a Householder reduction
//
// In exact arithmetic, K steps of Lanczos with full reorthogonalization from
// a unit vector q give the leading K x K block of T = Q^H H Q for any unitary
// Q whose first column is q (the implicit Q theorem), up to the signs of the
// off-diagonal entries.  The kernel first applies the reflector that maps q
// to a multiple of e_1 to both sides of H (packed lower triangle in LDS, as
// in dense_eigen_kernel), then reduces the result to tridiagonal form by
// the reflectors of columns 0..K-2: alpha_j = T_jj, beta_j = |T_{j+1,j}|.
// No Krylov basis is stored (n(n+1)/2 + 2n complex values of LDS), so
// several chains share a CU (five of 60 rows).  A zero off-diagonal entry
// ends the chain as a zero beta ends the recurrence (kout,
// IMPED_ERR_BREAKDOWN).
template <int BT>
__global__ __launch_bounds__(BT) void house_lanczos_kernel(LanczosArgs a) {
  __shared__ double red2[2 * (BT / 64)];
  int parity = 0;
  extern __shared__ double2 dsm[];
  const int c = a.chain_list[blockIdx.x];
  const int n = (int)a.chain_n[c];
  const int64_t base = a.row_base[c];
  double* alpha = a.alpha + (size_t)c * a.K;
  double* beta = a.beta + (size_t)c * a.K;
  if (n == 0) {
    if (threadIdx.x == 0) {
      a.kout[c] = 0;
      a.status[c] = IMPED_OK;
    }
    return;
  }
  const int K = a.K < n ? a.K : n;
  const int np = n * (n + 1) / 2;
  double2* A = dsm;      // packed lower triangle
  double2* u = A + np;   // the start vector, then its reflector
  double2* yv = u + n;   // A v, then w in place (w_i needs only y_i and v_i)
  for (int q = threadIdx.x; q < np; q += BT) A[q] = make_double2(0.0, 0.0);
  for (int i = threadIdx.x; i < n; i += BT) u[i] = a.v0[base + i];
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += BT) {
    const int64_t e0 = a.rowptr[base + i], e1 = a.rowptr[base + i + 1];
    for (int64_t q = e0; q < e1; ++q) {
      const int j = a.col[q];
      if (j >= 0 && j <= i) {
        double2& t = A[dense_packed(i, j, n)];
        t.x += a.val[q].x;
        t.y += a.val[q].y;
      }
    }
  }
  __syncthreads();
  constexpr int TPR = BT >= 512 ? 4 : 2;  // threads per row of A v (BT / TPR >= kDenseMaxN rows)
  int kdone = K, status = IMPED_OK;
  bool stopped = false;  // a zero beta or a zero start vector ended the chain
  // reflector k acts on rows k+1..n-1; k = -1 is the start vector's
  for (int k = -1; k + 1 < K; ++k) {
    const int off = k + 1, m = n - off;
    double2* col = k < 0 ? u : A + dense_packed(k + 1, k, n);
    const double2 x0 = col[0];
    const double dk = k >= 0 ? A[dense_packed(k, k, n)].x : 0.0;
    double p = 0.0;
    for (int i = 1 + threadIdx.x; i < m; i += BT) p += col[i].x * col[i].x + col[i].y * col[i].y;
    const double xn2 = block_sum1<BT>(p, red2, parity);  // its barrier also orders the reads of x0 before v
    const bool refl = !(xn2 == 0.0 && x0.y == 0.0);
    double bk = x0.x;
    double2 tk = make_double2(0.0, 0.0), sc = make_double2(0.0, 0.0);
    if (refl) {
      const double nr = sqrt(x0.x * x0.x + x0.y * x0.y + xn2);
      bk = x0.x >= 0.0 ? -nr : nr;
      tk = make_double2((bk - x0.x) / bk, -x0.y / bk);
      const double ur = x0.x - bk, ui = x0.y, den = ur * ur + ui * ui;
      sc = make_double2(ur / den, -ui / den);
    }
    if (k >= 0) {
      if (threadIdx.x == 0) {
        alpha[k] = dk;
        beta[k] = fabs(bk);
      }
      if (bk == 0.0) {  // the same in every thread
        kdone = k + 1;
        status = IMPED_ERR_BREAKDOWN;
        stopped = true;
        break;
      }
    } else if (xn2 == 0.0 && x0.x == 0.0 && x0.y == 0.0) {
      // a zero start vector: the recurrence gives alpha_0 = 0, then beta_0 = 0
      if (threadIdx.x == 0) {
        alpha[0] = 0.0;
        if (K > 1) beta[0] = 0.0;
      }
      kdone = 1;
      status = K > 1 ? IMPED_ERR_BREAKDOWN : IMPED_OK;
      stopped = true;
      break;
    }
    if (!refl) continue;  // the same in every thread
    for (int i = threadIdx.x; i < m; i += BT) {
      const double2 x = col[i];
      col[i] = i == 0 ? make_double2(1.0, 0.0) : make_double2(x.x * sc.x - x.y * sc.y, x.x * sc.y + x.y * sc.x);
    }
    __syncthreads();
    {
      // y = A v over the trailing block, TPR threads per row
      const int r = threadIdx.x / TPR, q = threadIdx.x % TPR;
      double yr = 0.0, yi = 0.0;
      if (r < m) {
        const int gr = off + r;
#pragma unroll 4
        for (int s = q; s <= r; s += TPR) {
          const double2 h = A[dense_packed(gr, off + s, n)], v = col[s];
          yr += h.x * v.x - h.y * v.y;
          yi += h.x * v.y + h.y * v.x;
        }
        const int cg = dense_packed(gr, gr, n) - r;  // A[cg + s] = A[off+s][gr] for s > r
#pragma unroll 4
        for (int s = r + 1 + q; s < m; s += TPR) {
          const double2 h = A[cg + s], v = col[s];
          yr += h.x * v.x + h.y * v.y;
          yi += h.x * v.y - h.y * v.x;
        }
      }
      yr = row_lanes_sum(yr, TPR);
      yi = row_lanes_sum(yi, TPR);
      if (r < m && q == 0) yv[r] = make_double2(yr, yi);
    }
    __syncthreads();
    double pv = 0.0;
    for (int i = threadIdx.x; i < m; i += BT) {
      const double2 v = col[i], y = yv[i];
      pv += v.x * y.x + v.y * y.y;
    }
    const double hv = 0.5 * (tk.x * tk.x + tk.y * tk.y) * block_sum1<BT>(pv, red2, parity);
    for (int i = threadIdx.x; i < m; i += BT) {
      const double2 y = yv[i], v = col[i];
      yv[i] = make_double2(tk.x * y.x - tk.y * y.y - hv * v.x, tk.x * y.y + tk.y * y.x - hv * v.y);
    }
    __syncthreads();
    {
      // A -= v w^H + w v^H on the trailing lower triangle, rows r and m-1-r
      // to one group of TU threads
      constexpr int TU = BT / 64;
      const int pr = threadIdx.x / TU, q = threadIdx.x % TU;
      if (pr < (m + 1) / 2) {
        for (int half = 0; half < 2; ++half) {
          const int r = half == 0 ? pr : m - 1 - pr;
          if (half == 1 && r == pr) break;
          const double2 vr = col[r], wr = yv[r];
          const int gr = off + r;
#pragma unroll 4
          for (int s = q; s <= r; s += TU) {
            const double2 vs = col[s], ws = yv[s];
            double2& t = A[dense_packed(gr, off + s, n)];
            t.x -= vr.x * ws.x + vr.y * ws.y + wr.x * vs.x + wr.y * vs.y;
            t.y -= vr.y * ws.x - vr.x * ws.y + wr.y * vs.x - wr.x * vs.y;
          }
        }
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (!stopped) alpha[K - 1] = A[dense_packed(K - 1, K - 1, n)].x;
    a.kout[c] = kdone;
    a.status[c] = status;
  }
}

// LDS bytes of house_lanczos_kernel for an n-row chain.
size_t house_lanczos_bytes(int64_t n) { return 16 * ((size_t)n * (size_t)(n + 1) / 2 + 2 * (size_t)n); }

// ------------------------------------------------------------------ shifted solves
//
// (z_s - H) x_s = y for S complex shifts z_s of one chain (the incoming
// energies of getRIXSmap, spectra.py:1066-1080) share one Krylov space:
// Lanczos with full reorthogonalization from y/|y| gives H V = V T + b v e^T,
// and the Galerkin solution x_s = |y| V (z_s - T)^-1 e_1 has the residual
// |y| b |e_m^T (z_s - T)^-1 e_1|, so one SpMV per step serves every shift
// and the residual of each is known without forming x_s.

struct SolveArgs {
  const int64_t* row_base;
  const int64_t* chain_n;
  const int64_t* rowptr;
  const int32_t* col;
  const double2* val;
  const double2* y;    // right-hand side of every chain, one entry per row
  const double2* z;    // [C][S] shifts
  int S;
  int max_steps;
  double tol;          // relative residual |y - (z - H) x| / |y|
  double2* vstore;     // per chain: min(max_steps, n) * n Krylov vectors at voff[c]
  const int64_t* voff;
  double2* wbuf;       // per chain: n at row_base
  double* ab;          // [C][2 * max_steps]: alpha then beta
  double2* cz;         // [C][S][2 * max_steps]: pivots, then |y| (z - T)^-1 e_1
  int32_t* steps;
  double* resid;       // [C][S]
  int32_t* conv;
  const int32_t* chain_list;
  int dgks;
  int check_every;
};

// Last entry of (z - T_m)^-1 e_1 by downward elimination; the pivots
// d_i = z - a_i - b_{i-1}^2 / d_{i-1} keep |Im d_i| >= |Im z| > 0.
__device__ double2 tri_shift_last(const double* a, const double* b, int m, double2 z) {
  double2 d = make_double2(z.x - a[0], z.y);
  double2 f = make_double2(1.0, 0.0);
  for (int i = 1; i < m; ++i) {
    const double2 r = crecip(d);
    const double2 fr = cmul(f, r);
    f = make_double2(fr.x * b[i - 1], fr.y * b[i - 1]);
    const double bb = b[i - 1] * b[i - 1];
    d = make_double2(z.x - a[i] - bb * r.x, z.y - bb * r.y);
  }
  return cmul(f, crecip(d));
}

__device__ __forceinline__ unsigned long long res_bits(double r) {
  // non-negative doubles order like their bit patterns; NaN sorts above inf
  return (unsigned long long)__double_as_longlong(r != r ? __longlong_as_double(0x7ff8000000000000ll) : fabs(r));
}

template <int BT>
__global__ __launch_bounds__(BT) void solve_kernel(SolveArgs a) {
  __shared__ double red[BT / 64];
  __shared__ unsigned long long s_max;
  extern __shared__ double2 cl[];  // max_steps + 2 coefficients
  const int c = a.chain_list[blockIdx.x];
  const int64_t n = a.chain_n[c];
  const int64_t base = a.row_base[c];
  const int ms = a.max_steps;
  const int S = a.S;
  double* alpha = a.ab + (size_t)c * 2 * ms;
  double* beta = alpha + ms;
  double2* V = a.vstore + a.voff[c];
  double2* w = a.wbuf + base;
  double p = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += BT) {
    const double2 v = a.y[base + i];
    p += v.x * v.x + v.y * v.y;
  }
  const double ynorm = sqrt(block_sum<BT>(p, red));
  if (n == 0 || !(ynorm > 0.0)) {
    for (int s = threadIdx.x; s < S; s += BT) a.resid[(size_t)c * S + s] = 0.0;
    if (threadIdx.x == 0) {
      a.steps[c] = 0;
      a.conv[c] = 1;
    }
    return;
  }
  for (int64_t i = threadIdx.x; i < n; i += BT) {
    const double2 v = a.y[base + i];
    V[i] = make_double2(v.x / ynorm, v.y / ynorm);
  }
  __syncthreads();
  const int mmax = (int)(ms < n ? ms : n);
  int m = 0, converged = 0;
  for (int j = 0; j < mmax; ++j) {
    const double2* vj = V + (size_t)j * n;
    const double al = block_sum<BT>(spmv_t<BT, double2>(a.rowptr, a.col, a.val, base, n, vj, w), red);
    const double bprev = j > 0 ? beta[j - 1] : 0.0;
    double p0 = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += BT) {
      double xr = w[i].x - al * vj[i].x, xi = w[i].y - al * vj[i].y;
      if (j > 0) {
        xr -= bprev * vj[i - n].x;
        xi -= bprev * vj[i - n].y;
      }
      w[i] = make_double2(xr, xi);
      p0 += xr * xr + xi * xi;
    }
    p0 = block_sum<BT>(p0, red);
    dots_lds<BT, double2>(V, j + 1, n, w, cl, false);
    double p1 = update_norm<BT, double2>(V, j + 1, n, w, cl, red);
    if (!a.dgks || p1 < 0.5 * p0) {
      dots_lds<BT, double2>(V, j + 1, n, w, cl, false);
      p1 = update_norm<BT, double2>(V, j + 1, n, w, cl, red);
    }
    const double b = sqrt(p1);
    if (threadIdx.x == 0) {
      alpha[j] = al;
      beta[j] = b;
    }
    m = j + 1;
    // an invariant subspace (b ~ 0) holds the exact solutions
    const bool last = (m == mmax) || b <= 1e-13 * (fabs(al) + bprev + 1e-300);
    __syncthreads();
    if (last || m % a.check_every == 0) {
      if (threadIdx.x == 0) s_max = 0ull;
      __syncthreads();
      for (int s = threadIdx.x; s < S; s += BT) {
        const double2 l = tri_shift_last(alpha, beta, m, a.z[(size_t)c * S + s]);
        const double r = last && b <= 1e-13 * (fabs(al) + bprev + 1e-300) ? 0.0 : b * hypot(l.x, l.y);
        a.resid[(size_t)c * S + s] = r;
        atomicMax(&s_max, res_bits(r));
      }
      __syncthreads();
      const double worst = __longlong_as_double((long long)s_max);
      if (worst <= a.tol) {
        converged = 1;
        break;
      }
      if (last) break;
    }
    for (int64_t i = threadIdx.x; i < n; i += BT) V[(size_t)(j + 1) * n + i] = make_double2(w[i].x / b, w[i].y / b);
    __syncthreads();
  }
  // coefficients |y| (z_s - T_m)^-1 e_1: elimination pivots D, right-hand
  // side F, then back substitution into F
  for (int s = threadIdx.x; s < S; s += BT) {
    const double2 z = a.z[(size_t)c * S + s];
    double2* D = a.cz + ((size_t)c * S + s) * 2 * ms;
    double2* F = D + ms;
    D[0] = make_double2(z.x - alpha[0], z.y);
    F[0] = make_double2(ynorm, 0.0);
    for (int i = 1; i < m; ++i) {
      const double2 r = crecip(D[i - 1]);
      const double2 fr = cmul(F[i - 1], r);
      F[i] = make_double2(fr.x * beta[i - 1], fr.y * beta[i - 1]);
      const double bb = beta[i - 1] * beta[i - 1];
      D[i] = make_double2(z.x - alpha[i] - bb * r.x, z.y - bb * r.y);
    }
    F[m - 1] = cmul(F[m - 1], crecip(D[m - 1]));
    for (int i = m - 2; i >= 0; --i) {
      const double2 t = make_double2(F[i].x + beta[i] * F[i + 1].x, F[i].y + beta[i] * F[i + 1].y);
      F[i] = cmul(t, crecip(D[i]));
    }
  }
  if (threadIdx.x == 0) {
    a.steps[c] = m;
    a.conv[c] = converged;
  }
}

// x_s[i] = sum_l coef_s[l] V_l[i] for every shift; eight shifts per pass so
// each Krylov entry is read S/8 times (from cache after the first).
__global__ __launch_bounds__(256) void solve_combine_kernel(const int64_t* row_base, const int64_t* chain_n,
                                                            const int32_t* chain_list, const double2* vstore,
                                                            const int64_t* voff, const double2* cz,
                                                            const int32_t* steps, int S, int ms, double2* X) {
  const int c = chain_list[blockIdx.y];
  const int64_t n = chain_n[c];
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int m = steps[c];
  const double2* V = vstore + voff[c];
  double2* Xc = X + row_base[c] * S;
  for (int s0 = 0; s0 < S; s0 += 8) {
    double ax[8], ay[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) ax[u] = ay[u] = 0.0;
    for (int l = 0; l < m; ++l) {
      const double2 v = V[(size_t)l * n + i];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (s0 + u < S) {
          const double2 k = cz[((size_t)c * S + s0 + u) * 2 * ms + ms + l];
          ax[u] += k.x * v.x - k.y * v.y;
          ay[u] += k.x * v.y + k.y * v.x;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (s0 + u < S) Xc[(size_t)(s0 + u) * n + i] = make_double2(ax[u], ay[u]);
  }
}

// ------------------------------------------------------------------ continued fraction

struct CFArgs {
  const double* alpha;
  const double* beta;
  const int32_t* kout;
  int K;
  const double* energy;
  const double* norm2;
  const int32_t* sign;
  const int32_t* zflag;  // 1 on exactly one chain per thermal state of a spectrum
  const double* kT;               // per spectrum
  const int32_t* spec_chain_off;  // chains of spectrum q: spec_chains[off[q] .. off[q+1])
  const int32_t* spec_chains;
  const double* omega;
  int n_w;
  double delta;
  double2* g_out;  // optional [C][n_w]
  double* spec_out;
  int q0;          // first spectrum of this launch (grids hold at most 65535 rows)
  int tile;        // chains staged in LDS at a time
};

// 1/d for the continued fraction: conj(d) / |d|^2 with a Newton-refined
// hardware reciprocal.  |d|^2 stays in range here: Im d = +-delta != 0 and
// |d| is of the order of the spectrum's energies.
__device__ __forceinline__ double2 cf_recip(double2 d) {
  const double inv = rcp_nr(d.x * d.x + d.y * d.y);
  return make_double2(d.x * inv, -d.y * inv);
}

// One thread per (spectrum, omega); the spectrum's chains are summed in order.
// The workgroup's threads share the spectrum, so its chains' alpha and beta^2
// are staged in LDS, a.tile chains at a time, and every level of a thread's
// recurrence reads them there (a broadcast) instead of from memory.
__global__ __launch_bounds__(256) void cf_kernel(CFArgs a) {
  extern __shared__ double2 ab[];  // [a.tile][a.K]: (alpha_i, beta_i^2)
  const int q = a.q0 + blockIdx.y;
  const int iw = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = iw < a.n_w;
  const double om = act ? a.omega[iw] : 0.0;
  const int t0 = a.spec_chain_off[q], t1 = a.spec_chain_off[q + 1];
  // Boltzmann weights exp(-(e - e_min)/kT) / Z over the spectrum's states
  const double kT = a.kT[q];
  double emin = INFINITY;
  for (int t = t0; t < t1; ++t) emin = fmin(emin, a.energy[a.spec_chains[t]]);
  double Z = 0.0;
  for (int t = t0; t < t1; ++t)
    if (a.zflag[a.spec_chains[t]]) Z += exp(-(a.energy[a.spec_chains[t]] - emin) / kT);
  double acc = 0.0;
  for (int tb = t0; tb < t1; tb += a.tile) {
    const int nt = min(a.tile, t1 - tb);
    __syncthreads();  // the previous tile has been read
    for (int idx = threadIdx.x; idx < nt * a.K; idx += blockDim.x) {
      const int t = idx / a.K, i = idx - t * a.K;
      const size_t c = (size_t)a.spec_chains[tb + t];
      const double b = a.beta[c * a.K + i];
      ab[idx] = make_double2(a.alpha[c * a.K + i], b * b);
    }
    __syncthreads();
    if (!act) continue;
    for (int t = 0; t < nt; ++t) {
      const int c = a.spec_chains[tb + t];
      const int K = a.kout[c];
      const double sg = (double)a.sign[c];
      // z = sg*omega + i*sg*delta + e  (photoemission uses -omega, -delta)
      const double zr = sg * om + a.energy[c];
      const double zi = sg * a.delta;
      const double2* abt = ab + (size_t)t * a.K;
      double2 g = make_double2(0.0, 0.0);
      for (int i = K - 1; i >= 0; --i) {
        const double2 q2 = abt[i];
        // the last level has no tail (beta_{K-1} is not part of the chain)
        const double b2 = i < K - 1 ? q2.y : 0.0;
        g = cf_recip(make_double2(zr - q2.x - b2 * g.x, zi - b2 * g.y));
      }
      // gs = sign * norm2 * g ; spectrum contribution weight * (-Im gs)
      const double f = sg * a.norm2[c];
      const double2 gs = make_double2(f * g.x, f * g.y);
      if (a.g_out) a.g_out[(size_t)c * a.n_w + iw] = gs;
      acc += exp(-(a.energy[c] - emin) / kT) * (-gs.y);
    }
  }
  if (act) a.spec_out[(size_t)q * a.n_w + iw] = Z > 0.0 ? acc / Z : 0.0;
}

// ------------------------------------------------------------------ flat (multi-workgroup) Krylov
//
// For sectors too large for one workgroup (10^5..10^7 states): every
// Lanczos step is a short sequence of grid-wide launches on one chain, with
// alpha/beta kept on the device (no host round trip per step):
//   spmv (+ per-block Re<v, Hv>) -> reduce -> three-term update (+ |w|^2)
//   -> reduce (beta = sqrt) -> [Gram-Schmidt: dots -> reduce -> update]
//   -> scale into the next Krylov vector.
// The SpMV gives LPR lanes to a row (the autotuned launch shape).

constexpr int FT = 256;  // threads per flat block

// VR: the values are real and read from vre (8 bytes each) instead of val.
template <int LPR, class T, bool VR = false>
__device__ __forceinline__ void spmv_rows(const int64_t* rowptr, const int32_t* col, const double2* val, int64_t base,
                                          int64_t n, const T* x, T* y, double* partial, double* red,
                                          const double* vre = nullptr) {
  const int g = threadIdx.x / LPR, r = threadIdx.x % LPR;
  double pa = 0.0;
  // grid-stride over groups of FT/LPR rows, so the partial sums number at
  // most gridDim.x (the loop bound is uniform in the workgroup)
  for (int64_t rb = blockIdx.x; rb * (FT / LPR) < n; rb += gridDim.x) {
    const int64_t row = rb * (FT / LPR) + g;
    double ax = 0.0, ay = 0.0;
    if (row < n) {
      const int64_t e0 = rowptr[base + row], e1 = rowptr[base + row + 1];
      for (int64_t e = e0 + r; e < e1; e += LPR) {
        const T v = x[col[e]];
        if constexpr (VR) {
          const double h = vre[e];
          ax += h * re_(v);
          if constexpr (!is_real_v<T>) ay += h * im_(v);
        } else {
          const double2 h = val[e];
          if constexpr (is_real_v<T>) {
            ax += h.x * re_(v);
          } else {
            ax += h.x * re_(v) - h.y * im_(v);
            ay += h.x * im_(v) + h.y * re_(v);
          }
        }
      }
    }
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) {
      ax += __shfl_xor(ax, off);
      if constexpr (!is_real_v<T>) ay += __shfl_xor(ay, off);
    }
    if (row < n && r == 0) {
      y[row] = mk<T>(ax, ay);
      pa += ax * re_(x[row]) + ay * im_(x[row]);
    }
  }
  pa = block_sum<FT>(pa, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = pa;
}

template <int LPR, class T, bool VR>
__global__ __launch_bounds__(FT) void flat_spmv_kernel(const int64_t* rowptr, const int32_t* col, const double2* val,
                                                       const double* vre, int64_t base, int64_t n, const T* x, T* y,
                                                       double* partial) {
  __shared__ double red[FT / 64];
  spmv_rows<LPR, T, VR>(rowptr, col, val, base, n, x, y, partial, red, vre);
}

// out[k] = sum of partial[k*stride .. k*stride + count) in a fixed order;
// mode 1 takes the square root (beta), mode 2 stores into alpha and adds
// the value to an accumulator slot.
__global__ __launch_bounds__(FT) void flat_reduce_kernel(const double* partial, int64_t count, int64_t stride,
                                                         double* out, int sqrt_mode) {
  __shared__ double red[FT / 64];
  const int k = blockIdx.x;
  double p = 0.0;
  for (int64_t i = threadIdx.x; i < count; i += FT) p += partial[(int64_t)k * stride + i];
  p = block_sum<FT>(p, red);
  if (threadIdx.x == 0) out[k] = sqrt_mode ? sqrt(p) : p;
}

// w = w - alpha v - beta_prev v_prev (beta_prev may be null), partial |w|^2.
template <class T>
__device__ __forceinline__ void update_rows(T* w, const T* v, const T* vprev, double al, double bp, int64_t n,
                                            double* partial, double* red) {
  double p = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * FT + threadIdx.x; i < n; i += (int64_t)gridDim.x * FT) {
    double xr = re_(w[i]) - al * re_(v[i]);
    double xi = im_(w[i]) - al * im_(v[i]);
    if (vprev) {
      xr -= bp * re_(vprev[i]);
      xi -= bp * im_(vprev[i]);
    }
    w[i] = mk<T>(xr, xi);
    p += xr * xr + xi * xi;
  }
  p = block_sum<FT>(p, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = p;
}

template <class T>
__global__ __launch_bounds__(FT) void flat_update_kernel(T* w, const T* v, const T* vprev, const double* alpha,
                                                         const double* beta_prev, int64_t n, double* partial) {
  __shared__ double red[FT / 64];
  update_rows<T>(w, v, vprev, *alpha, beta_prev ? *beta_prev : 0.0, n, partial, red);
}

// dst = src / beta; a zero beta flags a breakdown at step `step`.
template <class T>
__global__ __launch_bounds__(FT) void flat_scale_kernel(const T* src, T* dst, const double* beta, int64_t n,
                                                        int32_t* status, int32_t* kout, int step) {
  const double b = *beta;
  if (b == 0.0) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && *status == IMPED_OK) {
      *status = IMPED_ERR_BREAKDOWN;
      *kout = step;
    }
    return;
  }
  for (int64_t i = (int64_t)blockIdx.x * FT + threadIdx.x; i < n; i += (int64_t)gridDim.x * FT)
    dst[i] = mk<T>(re_(src[i]) / b, im_(src[i]) / b);
}

// Gram-Schmidt coefficients: partial[l][blk] = sum over this block's rows of
// conj(V_l) w, for l in [0, m).  Rows are split into grid.x chunks; each
// block keeps 8 coefficients in flight.
template <class T>
__global__ __launch_bounds__(FT) void flat_dots_kernel(const T* V, int m, int64_t n, const T* w, double2* partial,
                                                       int64_t chunk) {
  __shared__ double2 part[FT / 64][8];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t i0 = (int64_t)blockIdx.x * chunk;
  const int64_t i1 = i0 + chunk < n ? i0 + chunk : n;
  for (int l0 = 0; l0 < m; l0 += 8) {
    // all eight loads issue together (indices past m clamped, sums discarded)
    double cx[8], cy[8];
    size_t off[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      cx[u] = cy[u] = 0.0;
      off[u] = (size_t)(l0 + u < m ? l0 + u : m - 1) * n;
    }
    for (int64_t i = i0 + threadIdx.x; i < i1; i += FT) {
      const T b = w[i];
      T av[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) av[u] = V[off[u] + i];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        cx[u] += re_(av[u]) * re_(b) + im_(av[u]) * im_(b);
        if constexpr (!is_real_v<T>) cy[u] += re_(av[u]) * im_(b) - im_(av[u]) * re_(b);
      }
    }
    // eight wave sums per wave into LDS, one barrier, eight threads finish
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const double x = wave_sum(cx[u]);
      const double y = is_real_v<T> ? 0.0 : wave_sum(cy[u]);
      if (lane == 0) part[wave][u] = make_double2(x, y);
    }
    __syncthreads();
    if ((int)threadIdx.x < 8 && l0 + (int)threadIdx.x < m) {
      double sx = 0.0, sy = 0.0;
      for (int wv = 0; wv < FT / 64; ++wv) {
        sx += part[wv][threadIdx.x].x;
        sy += part[wv][threadIdx.x].y;
      }
      partial[(size_t)(l0 + threadIdx.x) * gridDim.x + blockIdx.x] = make_double2(sx, sy);
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(FT) void flat_reduce_c_kernel(const double2* partial, int nb, double2* coef) {
  __shared__ double red[FT / 64];
  const int l = blockIdx.x;
  double x = 0.0, y = 0.0;
  for (int i = threadIdx.x; i < nb; i += FT) {
    x += partial[(size_t)l * nb + i].x;
    y += partial[(size_t)l * nb + i].y;
  }
  x = block_sum<FT>(x, red);
  y = block_sum<FT>(y, red);
  if (threadIdx.x == 0) coef[l] = make_double2(x, y);
}

// w -= sum_l coef[l] V_l, partial |w|^2 per block.
template <class T>
__global__ __launch_bounds__(FT) void flat_gs_update_kernel(const T* V, int m, int64_t n, T* w, const double2* coef,
                                                            double* partial) {
  __shared__ double red[FT / 64];
  double p = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * FT + threadIdx.x; i < n; i += (int64_t)gridDim.x * FT) {
    double wx = re_(w[i]), wy = im_(w[i]);
#pragma unroll 8
    for (int l = 0; l < m; ++l) {
      const double2 c = coef[l];
      const T a = V[(size_t)l * n + i];
      if constexpr (is_real_v<T>) {
        wx -= re_(a) * c.x;
      } else {
        wx -= re_(a) * c.x - im_(a) * c.y;
        wy -= re_(a) * c.y + im_(a) * c.x;
      }
    }
    w[i] = mk<T>(wx, wy);
    p += wx * wx + wy * wy;
  }
  p = block_sum<FT>(p, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = p;
}

// beta_j = sqrt(|w|^2) on the device: written to beta[j] and to the scale slot
__global__ void flat_set_beta_kernel(const double* p, double* beta_j, double* scale_slot) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const double b = sqrt(*p);
    *beta_j = b;
    *scale_slot = b;
  }
}

// Ritz check of T_m (one workgroup): eigenvalues, inverse-iteration vectors,
// residuals |beta_m y_k[m-1]| against tol; flags convergence.
__global__ __launch_bounds__(FT) void flat_ritz_kernel(const double* alpha, const double* beta, int m, int nev,
                                                       double tol, int ms, double* evals, double* Y, double* tw,
                                                       int* tpiv, int32_t* converged, uint64_t seed) {
  __shared__ double s_lam[MAX_NEV], s_lo[MAX_NEV], s_hi[MAX_NEV];
  __shared__ int s_first[MAX_NEV];
  __shared__ int s_flag, s_done;
  const int ne = nev < m ? nev : m;
  ritz_values<FT>(alpha, beta, m, ne, nullptr, s_lam, s_lo, s_hi, s_first, &s_flag);
  if (threadIdx.x == 0) s_done = 1;
  __syncthreads();
  if ((int)threadIdx.x < ne) {
    const int k = threadIdx.x;
    double* y = Y + (size_t)k * ms;
    double* t = tw + (size_t)k * 5 * ms;
    tri_inverse_iteration(alpha, beta, m, s_lam[k], y, t, t + ms, t + 2 * ms, t + 3 * ms, tpiv + (size_t)k * ms,
                          seed + 7919 * (uint64_t)k);
    evals[k] = s_lam[k];
    if (!(fabs(beta[m - 1] * y[m - 1]) <= tol * fmax(1.0, fabs(s_lam[k])))) atomicAnd(&s_done, 0);
  }
  __syncthreads();
  // orthogonalize inverse-iteration vectors of clustered eigenvalues
  ortho_clustered(Y, ms, m, ne, s_lam);
  if (threadIdx.x == 0) *converged = s_done;
}

// X_k = sum_l y_k[l] V_l for k < ne (complex output), partial |X_k|^2.
template <class T>
__global__ __launch_bounds__(FT) void flat_ritz_vectors_kernel(const T* V, int m, int64_t n, const double* Y, int ms,
                                                               int ne, double2* X, double* partial) {
  __shared__ double red[FT / 64];
  for (int k = 0; k < ne; ++k) {
    const double* y = Y + (size_t)k * ms;
    double p = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * FT + threadIdx.x; i < n; i += (int64_t)gridDim.x * FT) {
      double xr = 0.0, xi = 0.0;
      for (int l = 0; l < m; ++l) {
        const T v = V[(size_t)l * n + i];
        xr += y[l] * re_(v);
        xi += y[l] * im_(v);
      }
      X[(size_t)k * n + i] = make_double2(xr, xi);
      p += xr * xr + xi * xi;
    }
    p = block_sum<FT>(p, red);
    if (threadIdx.x == 0) partial[(size_t)k * gridDim.x + blockIdx.x] = p;
  }
}

__global__ __launch_bounds__(FT) void flat_normalize_kernel(double2* X, int64_t n, const double* norms2, int ne) {
  for (int k = 0; k < ne; ++k) {
    const double nn = sqrt(norms2[k]);
    for (int64_t i = (int64_t)blockIdx.x * FT + threadIdx.x; i < n; i += (int64_t)gridDim.x * FT)
      X[(size_t)k * n + i] = make_double2(X[(size_t)k * n + i].x / nn, X[(size_t)k * n + i].y / nn);
  }
}

template <class T>
__global__ __launch_bounds__(FT) void flat_random_kernel(T* v, int64_t n, uint64_t seed, int chain, double* partial) {
  __shared__ double red[FT / 64];
  double p = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * FT + threadIdx.x; i < n; i += (int64_t)gridDim.x * FT) {
    const uint64_t h = seed * 0x100000001B3ull + (uint64_t)chain * 0x9E3779B97F4A7C15ull + 2 * (uint64_t)i;
    const double xr = u01(h) - 0.5, xi = is_real_v<T> ? 0.0 : u01(h + 1) - 0.5;
    v[i] = mk<T>(xr, xi);
    p += xr * xr + xi * xi;
  }
  p = block_sum<FT>(p, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = p;
}

template <class T>
__global__ __launch_bounds__(FT) void flat_copy_kernel(const double2* src, T* dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * FT + threadIdx.x; i < n; i += (int64_t)gridDim.x * FT)
    dst[i] = mk<T>(src[i].x, src[i].y);
}

struct KScratch {
  DevBuf ws, vstore, coef, ab, yk, tw, tpiv, voff, conv, wbuf, clist;
  DevBuf cf_alpha_tmp, cf_in, gout, sout;
  DevBuf fpart, fcpart, fcoef, fscal;
  DevBuf sz, scz, sres, ssteps, eig_in;
};
KScratch& kscratch() {
  static KScratch k;
  return k;
}

int64_t chain_base(const KrylovBatch& kb, int c) {
  int64_t b = 0;
  for (int i = 0; i < c; ++i) b += kb.h_n[i];
  return b;
}

unsigned vec_blocks(int64_t n) { return (unsigned)std::min<int64_t>(std::max<int64_t>((n + FT - 1) / FT, 1), 2048); }

template <class T>
void launch_spmv(int lpr, const KrylovBatch& kb, int64_t base, int64_t n, const T* x, T* y, double* partial,
                 unsigned* nb_out, hipStream_t s) {
  const unsigned nb = (unsigned)std::min<int64_t>(std::max<int64_t>((n + FT / lpr - 1) / (FT / lpr), 1), 4096);
  *nb_out = nb;
  TimedLaunch tl("spmv", s);
#define IMPED_SPMV(L)                                                                                           \
  if (kb.val_re)                                                                                               \
    hipLaunchKernelGGL((flat_spmv_kernel<L, T, true>), dim3(nb), dim3(FT), 0, s, kb.rowptr, kb.col, kb.val,     \
                       kb.val_re, base, n, x, y, partial);                                                      \
  else                                                                                                         \
    hipLaunchKernelGGL((flat_spmv_kernel<L, T, false>), dim3(nb), dim3(FT), 0, s, kb.rowptr, kb.col, kb.val,    \
                       kb.val_re, base, n, x, y, partial)
  switch (lpr) {
    case 4: IMPED_SPMV(4); break;
    case 8: IMPED_SPMV(8); break;
    case 16: IMPED_SPMV(16); break;
    case 32: IMPED_SPMV(32); break;
    default: IMPED_SPMV(64); break;
  }
#undef IMPED_SPMV
  HIP_CHECK(hipGetLastError());
}

// ---- plain Lanczos of one large chain, replayed as a HIP graph ----
//
// Everything stays on the device (no per-step host synchronization), and
// every per-step quantity is read on the device: the step index j from a
// counter in the argument block, v_j from buffer (j+1)%2, v_{j-1} from
// buffer j%2, w from buffer 2, alpha[j] and beta[j-1] by index.  A step is
// five launches:
//   beta reduce (+ counter advance) -> scale w into v_j -> SpMV (+ per-block
//   Re<v, Hv>) -> alpha reduce -> three-term update (+ per-block |w|^2).
// The launches of G steps are captured once into a HIP graph whose only
// pointer argument is the argument block, and the graph is replayed K/G
// times; the last K%G steps are launched directly.  The graph depends on
// the grid shapes alone (powers of two), so one instantiation serves every
// chain of that shape.  With launch timing on, every step is launched
// directly so that each SpMV is timed.

struct StepArgs {
  double2* buf[3];
  double* alpha;
  double* beta;
  int32_t* status;
  int32_t* kout;
  double* part;
  const int64_t* rowptr;
  const int32_t* col;
  const double2* val;
  const double* val_re;  // real values (NULL: complex val)
  int64_t base, n;
  int32_t jc;  // index of the step opened last (-1 before the first)
  int32_t kc;  // steps of this chain
};

// Per-chain set-up: argument block, counter, kout = K, status = ok.
__global__ void step_init_kernel(StepArgs v, StepArgs* a) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *a = v;
    a->jc = -1;
    *v.kout = v.kc;
    *v.status = IMPED_OK;
  }
}

// Opens step j = jc + 1: beta[j-1] = sqrt(sum of the update's partials).
__global__ __launch_bounds__(FT) void step_beta_kernel(StepArgs* a, int64_t count) {
  __shared__ double red[FT / 64];
  const int j = a->jc + 1;
  double p = 0.0;
  if (j > 0)
    for (int64_t i = threadIdx.x; i < count; i += FT) p += a->part[i];
  p = block_sum<FT>(p, red);  // every thread has read jc before thread 0 advances it
  if (threadIdx.x == 0) {
    a->jc = j;
    if (j > 0) a->beta[j - 1] = sqrt(p);
  }
}

// v_j = w / beta[j-1] into the buffer of v_{j-2}; a zero beta flags a breakdown.
__global__ __launch_bounds__(FT) void step_scale_kernel(StepArgs* a) {
  const int j = a->jc;
  if (j == 0) return;
  const double b = a->beta[j - 1];
  if (b == 0.0) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && *a->status == IMPED_OK) {
      *a->status = IMPED_ERR_BREAKDOWN;
      *a->kout = j;
    }
    return;
  }
  const double2* src = a->buf[2];
  double2* dst = a->buf[(j + 1) % 2];
  const int64_t n = a->n;
  for (int64_t i = (int64_t)blockIdx.x * FT + threadIdx.x; i < n; i += (int64_t)gridDim.x * FT)
    dst[i] = make_double2(src[i].x / b, src[i].y / b);
}

template <int LPR, bool VR>
__global__ __launch_bounds__(FT) void step_spmv_kernel(StepArgs* a) {
  __shared__ double red[FT / 64];
  const int j = a->jc;
  spmv_rows<LPR, double2, VR>(a->rowptr, a->col, a->val, a->base, a->n, a->buf[(j + 1) % 2], a->buf[2], a->part, red,
                              a->val_re);
}

__global__ __launch_bounds__(FT) void step_alpha_kernel(StepArgs* a, int64_t count) {
  __shared__ double red[FT / 64];
  double p = 0.0;
  for (int64_t i = threadIdx.x; i < count; i += FT) p += a->part[i];
  p = block_sum<FT>(p, red);
  if (threadIdx.x == 0) a->alpha[a->jc] = p;
}

__global__ __launch_bounds__(FT) void step_update_kernel(StepArgs* a) {
  __shared__ double red[FT / 64];
  const int j = a->jc;
  update_rows<double2>(a->buf[2], a->buf[(j + 1) % 2], j > 0 ? a->buf[j % 2] : nullptr, a->alpha[j],
                       j > 0 ? a->beta[j - 1] : 0.0, a->n, a->part, red);
}

// Grid of a step kernel: the power of two covering `need` blocks, at most cap.
unsigned pow2_blocks(int64_t need, unsigned cap) {
  unsigned b = 1;
  while ((int64_t)b < need && b < cap) b <<= 1;
  return b;
}

void launch_lanczos_step(StepArgs* a, unsigned nbv, unsigned nbs, int lpr, bool real_vals, hipStream_t s) {
  hipLaunchKernelGGL(step_beta_kernel, dim3(1), dim3(FT), 0, s, a, (int64_t)nbv);
  hipLaunchKernelGGL(step_scale_kernel, dim3(nbv), dim3(FT), 0, s, a);
  {
    TimedLaunch tl("spmv", s);
#define IMPED_STEP_SPMV(L)                                                                  \
  if (real_vals)                                                                          \
    hipLaunchKernelGGL((step_spmv_kernel<L, true>), dim3(nbs), dim3(FT), 0, s, a);          \
  else                                                                                    \
    hipLaunchKernelGGL((step_spmv_kernel<L, false>), dim3(nbs), dim3(FT), 0, s, a)
    switch (lpr) {
      case 4: IMPED_STEP_SPMV(4); break;
      case 8: IMPED_STEP_SPMV(8); break;
      case 16: IMPED_STEP_SPMV(16); break;
      case 32: IMPED_STEP_SPMV(32); break;
      default: IMPED_STEP_SPMV(64); break;
    }
#undef IMPED_STEP_SPMV
  }
  hipLaunchKernelGGL(step_alpha_kernel, dim3(1), dim3(FT), 0, s, a, (int64_t)nbs);
  hipLaunchKernelGGL(step_update_kernel, dim3(nbv), dim3(FT), 0, s, a);
  HIP_CHECK(hipGetLastError());
}

// The argument block: allocated once and never moved, since instantiated
// graphs hold its address.
StepArgs* step_args_block() {
  static DevBuf blk;
  if (!blk.p) blk.reserve(sizeof(StepArgs));
  return blk.as<StepArgs>();
}

struct StepGraph {
  unsigned nbv, nbs;
  int lpr, G;
  bool real_vals;
  hipGraphExec_t exec;
};

// Instantiated G-step graph for a grid shape (captured on a private stream;
// at most 32 shapes are kept, the oldest destroyed after `s` has drained).
hipGraphExec_t lanczos_step_graph(StepArgs* a, unsigned nbv, unsigned nbs, int lpr, int G, bool real_vals,
                                  hipStream_t s) {
  static std::vector<StepGraph> cache;
  static hipStream_t cap = nullptr;
  for (const auto& g : cache)
    if (g.nbv == nbv && g.nbs == nbs && g.lpr == lpr && g.G == G && g.real_vals == real_vals) return g.exec;
  if (!cap) HIP_CHECK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
  HIP_CHECK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
  hipGraph_t graph = nullptr;
  try {
    for (int g = 0; g < G; ++g) launch_lanczos_step(a, nbv, nbs, lpr, real_vals, cap);
  } catch (...) {
    (void)hipStreamEndCapture(cap, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  HIP_CHECK(hipStreamEndCapture(cap, &graph));
  hipGraphExec_t exec = nullptr;
  const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  HIP_CHECK(ie);
  if (cache.size() >= 32) {
    stream_sync(s);
    HIP_CHECK(hipGraphExecDestroy(cache.front().exec));
    cache.erase(cache.begin());
  }
  cache.push_back(StepGraph{nbv, nbs, lpr, G, real_vals, exec});
  return exec;
}

// alpha/beta/kout/status point at this chain's slots.
void lanczos_flat_chain(const KrylovBatch& kb, int c, const double2* v0, int K, double* alpha, double* beta,
                        int32_t* kout, int32_t* status, hipStream_t s) {
  KScratch& sc = kscratch();
  const int64_t n = kb.h_n[c], base = chain_base(kb, c);
  const int Kc = (int)std::min<int64_t>(K, n);
  const int lpr = (int)g_tuning.spmv_lpr;
  sc.wbuf.reserve(3 * n * sizeof(double2));
  const unsigned nbv = pow2_blocks((n + FT - 1) / FT, 2048);
  const unsigned nbs = pow2_blocks((n + FT / lpr - 1) / (FT / lpr), 4096);
  sc.fpart.reserve(std::max<int64_t>((n + 3) / 4 + 1, 4096) * sizeof(double));
  StepArgs v{};
  for (int i = 0; i < 3; ++i) v.buf[i] = sc.wbuf.as<double2>() + i * n;
  v.alpha = alpha;
  v.beta = beta;
  v.status = status;
  v.kout = kout;
  v.part = sc.fpart.as<double>();
  v.rowptr = kb.rowptr;
  v.col = kb.col;
  v.val = kb.val;
  v.val_re = kb.val_re;
  v.base = base;
  v.n = n;
  v.kc = Kc;
  StepArgs* a = step_args_block();
  hipLaunchKernelGGL(step_init_kernel, dim3(1), dim3(64), 0, s, v, a);
  HIP_CHECK(hipMemcpyAsync(v.buf[1], v0 + base, n * sizeof(double2), hipMemcpyDeviceToDevice, s));
  const int G = (int)g_tuning.lanczos_graph_steps;
  int j = 0;
  if (G > 0 && !timing_enabled() && Kc >= G) {
    hipGraphExec_t exec = lanczos_step_graph(a, nbv, nbs, lpr, G, kb.val_re != nullptr, s);
    for (; j + G <= Kc; j += G) HIP_CHECK(hipGraphLaunch(exec, s));
  }
  for (; j < Kc; ++j) launch_lanczos_step(a, nbv, nbs, lpr, kb.val_re != nullptr, s);
}

// Lowest eigenpairs of one large chain: Lanczos with full reorthogonalization
// by grid-wide Gram-Schmidt, host-controlled (one small read-back per step).
template <class T>
int eigen_flat_chain(const KrylovBatch& kb, int c, int nev, double tol, int max_steps, uint64_t seed, double* evals,
                     double2* X, int32_t* steps_out, hipStream_t s) {
  KScratch& sc = kscratch();
  const int64_t n = kb.h_n[c], base = chain_base(kb, c);
  const int mmax = (int)std::min<int64_t>(max_steps, n);
  const int lpr = (int)g_tuning.spmv_lpr;
  const unsigned nbv = vec_blocks(n);
  sc.vstore.reserve((size_t)mmax * n * sizeof(T));
  sc.wbuf.reserve(n * sizeof(T));
  T* V = reinterpret_cast<T*>(sc.vstore.p);
  T* w = reinterpret_cast<T*>(sc.wbuf.p);
  // up to 2048 workgroups stream the Krylov basis in the dot products
  const int64_t chunk = std::max<int64_t>(FT * 4, (n + 2047) / 2048);
  const unsigned nbd = (unsigned)((n + chunk - 1) / chunk);
  sc.fpart.reserve(std::max<int64_t>((n + 3) / 4 + 1, 4096) * sizeof(double));
  sc.fcpart.reserve((size_t)(mmax + 1) * nbd * sizeof(double2));
  sc.fcoef.reserve((size_t)(mmax + 1) * sizeof(double2));
  sc.fscal.reserve(32 * sizeof(double));
  sc.ab.reserve((size_t)2 * mmax * sizeof(double));
  sc.yk.reserve((size_t)nev * mmax * sizeof(double));
  sc.tw.reserve((size_t)nev * 5 * mmax * sizeof(double));
  sc.tpiv.reserve((size_t)nev * mmax * sizeof(int));
  sc.conv.reserve(4);
  double* part = sc.fpart.as<double>();
  double* scal = sc.fscal.as<double>();
  double* alpha = sc.ab.as<double>();
  double* beta = alpha + mmax;
  hipLaunchKernelGGL((flat_random_kernel<T>), dim3(nbv), dim3(FT), 0, s, V, n, seed, c, part);
  hipLaunchKernelGGL(flat_reduce_kernel, dim3(1), dim3(FT), 0, s, part, (int64_t)nbv, (int64_t)0, scal, 1);
  hipLaunchKernelGGL((flat_scale_kernel<T>), dim3(nbv), dim3(FT), 0, s, V, V, scal, n, sc.conv.as<int32_t>(),
                     sc.conv.as<int32_t>(), 0);
  HIP_CHECK(hipGetLastError());
  int m = 0, converged = 0;
  double hb[4];
  auto gs_pass = [&](int mm) {
    hipLaunchKernelGGL((flat_dots_kernel<T>), dim3(nbd), dim3(FT), 0, s, V, mm, n, w, sc.fcpart.as<double2>(), chunk);
    hipLaunchKernelGGL(flat_reduce_c_kernel, dim3(mm), dim3(FT), 0, s, sc.fcpart.as<double2>(), (int)nbd,
                       sc.fcoef.as<double2>());
    hipLaunchKernelGGL((flat_gs_update_kernel<T>), dim3(nbv), dim3(FT), 0, s, V, mm, n, w, sc.fcoef.as<double2>(),
                       part);
    hipLaunchKernelGGL(flat_reduce_kernel, dim3(1), dim3(FT), 0, s, part, (int64_t)nbv, (int64_t)0, scal + 2, 0);
    HIP_CHECK(hipGetLastError());
  };
  for (int j = 0; j < mmax; ++j) {
    const T* vj = V + (size_t)j * n;
    unsigned nbs = 0;
    launch_spmv<T>(lpr, kb, base, n, vj, w, part, &nbs, s);
    hipLaunchKernelGGL(flat_reduce_kernel, dim3(1), dim3(FT), 0, s, part, (int64_t)nbs, (int64_t)0, alpha + j, 0);
    hipLaunchKernelGGL((flat_update_kernel<T>), dim3(nbv), dim3(FT), 0, s, w, vj, j > 0 ? vj - n : (const T*)nullptr,
                       alpha + j, j > 0 ? beta + j - 1 : nullptr, n, part);
    hipLaunchKernelGGL(flat_reduce_kernel, dim3(1), dim3(FT), 0, s, part, (int64_t)nbv, (int64_t)0, scal + 1, 0);
    gs_pass(j + 1);
    // one host read per step (|w|^2 before and after the first pass, alpha_j)
    // decides the DGKS second pass; beta_j is written on the device
    double al = 0.0;
    d2h_bytes(hb, scal, 3 * sizeof(double), s);
    d2h_bytes(&al, alpha + j, sizeof(double), s);
    stream_sync(s);
    double p1 = hb[2];
    if (!g_tuning.eigen_dgks || p1 < 0.5 * hb[1]) {
      gs_pass(j + 1);
      d2h_bytes(&p1, scal + 2, sizeof(double), s);
      stream_sync(s);
    }
    const double b = std::sqrt(p1);
    const double bprev = j > 0 ? hb[3] : 0.0;
    hb[3] = b;
    hipLaunchKernelGGL(flat_set_beta_kernel, dim3(1), dim3(64), 0, s, scal + 2, beta + j, scal + 3);
    HIP_CHECK(hipGetLastError());
    m = j + 1;
    const bool last = (m == mmax) || b <= 1e-13 * (std::fabs(al) + bprev + 1e-300);
    if (last || (m >= nev && m % g_tuning.eigen_check_every == 0)) {
      hipLaunchKernelGGL(flat_ritz_kernel, dim3(1), dim3(FT), 0, s, alpha, beta, m, nev, tol, mmax,
                         evals, sc.yk.as<double>(), sc.tw.as<double>(), sc.tpiv.as<int>(), sc.conv.as<int32_t>(),
                         seed);
      HIP_CHECK(hipGetLastError());
      int32_t conv = 0;
      d2h_bytes(&conv, sc.conv.p, 4, s);
      stream_sync(s);
      if (conv || last) {
        converged = conv;
        break;
      }
    }
    hipLaunchKernelGGL((flat_scale_kernel<T>), dim3(nbv), dim3(FT), 0, s, w, V + (size_t)(j + 1) * n, scal + 3, n,
                       sc.conv.as<int32_t>(), sc.conv.as<int32_t>(), 0);
    HIP_CHECK(hipGetLastError());
  }
  const int ne = std::min(nev, m);
  const unsigned nbr = vec_blocks(n);
  hipLaunchKernelGGL((flat_ritz_vectors_kernel<T>), dim3(nbr), dim3(FT), 0, s, V, m, n, sc.yk.as<double>(), mmax, ne,
                     X, part);
  hipLaunchKernelGGL(flat_reduce_kernel, dim3(ne), dim3(FT), 0, s, part, (int64_t)nbr, (int64_t)nbr, scal + 4, 0);
  hipLaunchKernelGGL(flat_normalize_kernel, dim3(nbr), dim3(FT), 0, s, X, n, scal + 4, ne);
  HIP_CHECK(hipGetLastError());
  if (ne < nev) {
    HIP_CHECK(hipMemsetAsync(X + (size_t)ne * n, 0, (size_t)(nev - ne) * n * sizeof(double2), s));
    std::vector<double> nanv(nev - ne, NAN);
    HIP_CHECK(hipMemcpyAsync(evals + ne, nanv.data(), nanv.size() * sizeof(double), hipMemcpyHostToDevice, s));
  }
  HIP_CHECK(hipMemcpyAsync(steps_out, &m, 4, hipMemcpyHostToDevice, s));
  stream_sync(s);
  return converged ? IMPED_OK : IMPED_ERR_NOCONV;
}

}  // namespace

namespace {

// LDS of one chain of lanczos_lds_kernel: two chains per CU (160 KiB).
constexpr size_t LDS_CHAIN_BYTES = 80 * 1024;

int lds_waves() {
  const int64_t w = g_tuning.lanczos_lds_waves;
  return w == 2 || w == 4 ? (int)w : 1;
}

// LDS of one chain of the LDS-resident Lanczos for the current wave count
size_t lds_chain_bytes(int64_t n, int K) {
  const int nw = lds_waves();
  return nw == 1 ? lanczos_lds_bytes(n, K) : lanczos_ldsw_bytes(n, K, nw);
}

template <class F>
void set_lds_limit(F* kernel) {
  HIP_CHECK(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)LDS_CHAIN_BYTES));
}

// nb chains (a.chain_list) on the LDS-resident Lanczos, `bytes` of LDS each
void launch_lanczos_lds(const LanczosArgs& a, unsigned nb, size_t bytes, hipStream_t s) {
  static std::once_flag once;
  std::call_once(once, [] {
    set_lds_limit(lanczos_lds_kernel);
    set_lds_limit(lanczos_ldsw_kernel<2>);
    set_lds_limit(lanczos_ldsw_kernel<4>);
  });
  switch (lds_waves()) {
    case 4:
      hipLaunchKernelGGL(lanczos_ldsw_kernel<4>, dim3(nb), dim3(256), bytes, s, a);
      break;
    case 2:
      hipLaunchKernelGGL(lanczos_ldsw_kernel<2>, dim3(nb), dim3(128), bytes, s, a);
      break;
    default:
      hipLaunchKernelGGL(lanczos_lds_kernel, dim3(nb), dim3(64), bytes, s, a);
  }
}

}  // namespace

void run_lanczos(const KrylovBatch& kb, const double2* v0, int K, int reorth, double* alpha, double* beta,
                 int32_t* kout, int32_t* status, hipStream_t s) {
  KScratch& sc = kscratch();
  int64_t R = 0;
  for (auto x : kb.h_n) R += x;
  sc.ws.reserve(std::max<int64_t>(3 * R, 1) * 16);
  LanczosArgs a{};
  a.row_base = kb.row_base;
  a.chain_n = kb.chain_n;
  a.rowptr = kb.rowptr;
  a.col = kb.col;
  a.val = kb.val;
  a.v0 = v0;
  a.K = K;
  a.reorth = reorth;
  a.ws = sc.ws.as<double2>();
  if (reorth) {
    // the Gram-Schmidt coefficients (K + 2 complex) live in LDS
    if (K > 3500) throw HipError{"lanczos: K must be <= 3500 with reorthogonalization"};
    sc.vstore.reserve(std::max<int64_t>(R * K, 1) * 16);
    a.vstore = sc.vstore.as<double2>();
  }
  const size_t shm = reorth ? (size_t)(K + 2) * sizeof(double2) : 0;
  a.alpha = alpha;
  a.beta = beta;
  a.kout = kout;
  a.status = status;
  if (kb.C == 0) return;
  // chains bucketed by dimension: with reorthogonalization, chains whose
  // Krylov basis fits LDS_CHAIN_BYTES get one wave with the basis in LDS;
  // otherwise one wave, 256 or 1024 threads per chain; the largest (plain
  // recurrence) run grid-wide, one chain at a time
  std::vector<int32_t> lists[5], flat;
  size_t lds_max = 0, house_max = 0;
  const int64_t house_lim = std::min<int64_t>(g_tuning.lanczos_dense_max_n, kDenseMaxN);
  for (int c = 0; c < kb.C; ++c) {
    const int64_t n = kb.h_n[c];
    if (!reorth && n > g_tuning.flat_min_n) {
      flat.push_back(c);
      continue;
    }
    if (reorth && n <= house_lim) {
      lists[4].push_back(c);
      house_max = std::max(house_max, house_lanczos_bytes(n));
      continue;
    }
    if (reorth && g_tuning.lanczos_lds && lds_chain_bytes(n, K) <= LDS_CHAIN_BYTES) {
      lists[3].push_back(c);
      lds_max = std::max(lds_max, lds_chain_bytes(n, K));
      continue;
    }
    const int bkt = n <= g_tuning.lanczos_wave_max_n ? 0 : (n <= g_tuning.lanczos_wg256_max_n ? 1 : 2);
    lists[bkt].push_back(c);
  }
  for (int c : flat)
    lanczos_flat_chain(kb, c, v0, K, alpha + (size_t)c * K, beta + (size_t)c * K, kout + c, status + c, s);
  std::vector<int32_t> all;
  for (auto& l : lists) all.insert(all.end(), l.begin(), l.end());
  h2d(sc.clist, all.data(), all.size(), s);
  size_t off = 0;
  for (int bkt = 0; bkt < 5; ++bkt) {
    const unsigned nb = (unsigned)lists[bkt].size();
    if (!nb) continue;
    a.chain_list = sc.clist.as<int32_t>() + off;
    off += nb;
    TimedLaunch tl("lanczos", s);
    if (bkt == 0)
      hipLaunchKernelGGL(lanczos_kernel<64>, dim3(nb), dim3(64), shm, s, a);
    else if (bkt == 1)
      hipLaunchKernelGGL(lanczos_kernel<256>, dim3(nb), dim3(256), shm, s, a);
    else if (bkt == 2)
      hipLaunchKernelGGL(lanczos_kernel<1024>, dim3(nb), dim3(1024), shm, s, a);
    else if (bkt == 3)
      launch_lanczos_lds(a, nb, lds_max, s);
    else {
      static std::once_flag once;
      std::call_once(once, [] {
        HIP_CHECK(hipFuncSetAttribute((const void*)house_lanczos_kernel<256>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
      });
      hipLaunchKernelGGL(house_lanczos_kernel<256>, dim3(nb), dim3(256), house_max, s, a);
    }
    HIP_CHECK(hipGetLastError());
  }
  // no wait: the chain list stays in the scratch buffer, and later copies
  // into it are queued behind these launches on the same stream
}

int run_eigen(const KrylovBatch& kb, int nev, double tol, int max_steps, uint64_t seed, double* evals,
              double2* evecs, int32_t* steps, int32_t* conv, bool real_values, hipStream_t s) {
  KScratch& sc = kscratch();
  const int C = kb.C;
  if (C == 0) return IMPED_OK;
  // max_steps bound: the Gram-Schmidt coefficients (max_steps + 2 complex)
  // live in LDS next to the kernel's static arrays
  if (nev > MAX_NEV || max_steps < 1 || max_steps > 3500) {
    set_error("eigensolver: nev must be <= 16 and max_steps in [1, 3500]");
    return IMPED_ERR_INVALID;
  }
  const bool real = real_values && g_tuning.eigen_real;
  // chains above flat_min_n are solved grid-wide, one at a time, after the
  // one-workgroup-per-chain launch of the others
  std::vector<int32_t> order, large;
  for (int c = 0; c < C; ++c) (kb.h_n[c] > g_tuning.flat_min_n ? large : order).push_back(c);
  // chains up to eigen_dense_max_n rows (at most kDenseMaxN, and the LDS
  // layout within one CU's share) go to the dense solver
  std::vector<int32_t> dense;
  if (g_tuning.eigen_dense_max_n > 0) {
    const int64_t lim = std::min<int64_t>(g_tuning.eigen_dense_max_n, kDenseMaxN);
    std::vector<int32_t> rest;
    for (int32_t c : order)
      (kb.h_n[c] <= lim && dense_eigen_bytes(kb.h_n[c], nev) <= (size_t)150 * 1024 ? dense : rest).push_back(c);
    order.swap(rest);
  }
  std::vector<int64_t> voff(C + 1, 0);
  int64_t R = 0;
  for (int c = 0; c < C; ++c) {
    const bool small = kb.h_n[c] <= g_tuning.flat_min_n;
    const int64_t mm = small ? std::min<int64_t>(max_steps, kb.h_n[c]) : 0;
    voff[c + 1] = voff[c] + mm * kb.h_n[c];
    R += small ? kb.h_n[c] : 0;
  }
  sc.vstore.reserve(std::max<int64_t>(voff[C], 1) * 16);
  sc.wbuf.reserve(std::max<int64_t>(R, 1) * 16);
  sc.coef.reserve((size_t)C * max_steps * 16);
  sc.ab.reserve((size_t)C * 2 * max_steps * 8);
  sc.yk.reserve((size_t)C * nev * max_steps * 8);
  sc.tw.reserve((size_t)C * nev * 5 * max_steps * 8);
  sc.tpiv.reserve((size_t)C * nev * max_steps * 4);
  EigenArgs a{};
  a.row_base = kb.row_base;
  a.chain_n = kb.chain_n;
  a.rowptr = kb.rowptr;
  a.col = kb.col;
  a.val = kb.val;
  a.nev = nev;
  a.max_steps = max_steps;
  a.tol = tol;
  a.seed = seed;
  a.vstore = sc.vstore.p;
  a.wbuf = sc.wbuf.p;
  a.coef = sc.coef.as<double2>();
  a.ab = sc.ab.as<double>();
  a.yk = sc.yk.as<double>();
  a.tw = sc.tw.as<double>();
  a.tpiv = sc.tpiv.as<int>();
  a.evals = evals;
  a.evecs = evecs;
  a.steps = steps;
  a.conv = conv;
  a.dgks = (int)g_tuning.eigen_dgks;
  a.check_every = (int)std::max<int64_t>(1, g_tuning.eigen_check_every);
  a.check_start = (int)std::max<int64_t>(0, g_tuning.eigen_check_start);
  // largest chains first so the longest workgroups start first; chains up to
  // eigen_wave_max_n get one wave each (reductions by shuffles, no barriers),
  // the others eigen_bt threads
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return kb.h_n[x] > kb.h_n[y]; });
  // vector offsets and the chain order in one copy
  HostPack pk;
  const size_t o_voff = pk.add(voff.data(), C + 1), o_list = pk.add(order.data(), order.size());
  const size_t o_dense = pk.add(dense.data(), dense.size());
  pk.upload(sc.eig_in, s);
  a.voff = HostPack::at<int64_t>(sc.eig_in, o_voff);
  int32_t* clist = HostPack::at<int32_t>(sc.eig_in, o_list);
  if (!dense.empty()) {
    int64_t nmax = 0;
    for (int32_t c : dense) nmax = std::max(nmax, kb.h_n[c]);
    static bool attr = false;
    if (!attr) {
      HIP_CHECK(hipFuncSetAttribute((const void*)dense_eigen_kernel<512>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    150 * 1024));
      attr = true;
    }
    TimedLaunch tl("eigen", s);
    a.chain_list = HostPack::at<int32_t>(sc.eig_in, o_dense);
    hipLaunchKernelGGL(dense_eigen_kernel<512>, dim3((unsigned)dense.size()), dim3(512),
                       dense_eigen_bytes(nmax, nev), s, a);
    HIP_CHECK(hipGetLastError());
  }
  size_t n_wide = 0;
  while (n_wide < order.size() && kb.h_n[order[n_wide]] > g_tuning.eigen_wave_max_n) ++n_wide;
  // LDS layout of a launch whose largest chain is order[first]: coefficients
  // always, alpha/beta and the Ritz work arrays when they fit in 60 KiB (the
  // 64 KiB per-workgroup limit less the kernel's static arrays)
  size_t shm = 0;
  bool lv = false;
  const size_t esz = real ? 8 : 16;
  // LDS per workgroup: the whole 160 KiB when every chain of the launch gets
  // its own CU, shared out when several must sit on one CU at once
  static int n_cu = 0;
  if (!n_cu) HIP_CHECK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, 0));
  auto layout = [&](size_t first, int bt, size_t n_chains) {
    const int64_t nmax = kb.h_n[order[first]];
    const int L = (int)std::min<int64_t>(max_steps, nmax);
    const size_t coef = (size_t)(L + 2) * sizeof(double2), ab = 2 * (size_t)L * 8;
    const size_t work = (size_t)nev * ((size_t)L * 8 + 4 * (size_t)L * 8 + (size_t)L * 4);
    a.lstride = L;
    a.lds_ab = a.lds_work = coef + ab + work <= 60 * 1024;
    shm = coef + (a.lds_work ? ab + work : 0);
    // Krylov prefix in LDS (after w): as many vectors as the share allows
    const int per_cu = (int)std::min<size_t>((n_chains + n_cu - 1) / n_cu, std::max(1, 2048 / bt));
    const size_t share = (size_t)(160 * 1024) / (size_t)per_cu;
    const size_t static_lds = 6 * 1024;  // the kernel's own arrays, with room to spare
    const size_t woff = (shm + 15) & ~(size_t)15;
    lv = false;
    a.kl = 0;
    if (a.lds_work && share > woff + static_lds) {
      // w and H v_j, then kl vectors of row stride sl, then the Gram-Schmidt partial
      // sums (bt/64 rows of nmax complex)
      const int64_t sl = (nmax + 6) / 8 * 8 + 1;
      const int64_t part_bytes = (int64_t)(bt / 64) * nmax * 16 + 16;
      const int64_t avail = (int64_t)share - (int64_t)static_lds - (int64_t)woff - 2 * nmax * (int64_t)esz - part_bytes;
      int64_t kl = avail > 0 ? std::min<int64_t>(avail / ((int64_t)esz * sl), std::min<int64_t>(max_steps, nmax)) : 0;
      const int64_t cap = g_tuning.eigen_lds_vectors;
      if (cap >= 0) kl = std::min(kl, cap);
      if (kl >= 8 || (cap >= 0 && kl >= 1)) {
        lv = true;
        a.kl = (int)kl;
        shm = woff + (size_t)(2 * nmax + kl * sl) * esz + (size_t)part_bytes;
      }
    }
  };
  // eigen_kernel<bt, T, LW, LV> for the layout just chosen
  auto launch = [&](int bt, dim3 g) {
#define IMPED_EIG_T(BT_, T_)                                                                               \
  if (lv) {                                                                                                \
    static bool attr = false;                                                                              \
    if (!attr) {                                                                                           \
      HIP_CHECK(hipFuncSetAttribute((const void*)eigen_kernel<BT_, T_, true, true>,                        \
                                    hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 6 * 1024));   \
      attr = true;                                                                                         \
    }                                                                                                      \
    hipLaunchKernelGGL((eigen_kernel<BT_, T_, true, true>), g, dim3(BT_), shm, s, a);                     \
  } else if (a.lds_work) {                                                                                 \
    hipLaunchKernelGGL((eigen_kernel<BT_, T_, true, false>), g, dim3(BT_), shm, s, a);                    \
  } else {                                                                                                 \
    hipLaunchKernelGGL((eigen_kernel<BT_, T_, false, false>), g, dim3(BT_), shm, s, a);                   \
  }
#define IMPED_EIG(BT_)      \
  if (real) {               \
    IMPED_EIG_T(BT_, double)  \
  } else {                  \
    IMPED_EIG_T(BT_, double2) \
  }
    if (bt == 64) {
      IMPED_EIG(64)
    } else if (bt == 256) {
      IMPED_EIG(256)
    } else if (bt == 512) {
      IMPED_EIG(512)
    } else {
      IMPED_EIG(1024)
    }
#undef IMPED_EIG
#undef IMPED_EIG_T
    HIP_CHECK(hipGetLastError());
  };
  if (n_wide < order.size()) {
    TimedLaunch tl("eigen", s);
    layout(n_wide, 64, order.size() - n_wide);
    a.chain_list = clist + n_wide;
    launch(64, dim3((unsigned)(order.size() - n_wide)));
  }
  a.chain_list = clist;
  if (n_wide > 0) {
    // eigen_bt 0 (automatic): a launch lasts as long as its largest chain,
    // which gets 1024 threads above 512 rows (fewer rows per thread), else 512
    // (fewer waves at each barrier)
    const int bt = g_tuning.eigen_bt ? (int)g_tuning.eigen_bt : (kb.h_n[order[0]] > 512 ? 1024 : 512);
    TimedLaunch tl("eigen", s);
    layout(0, bt, n_wide);
    launch(bt, dim3((unsigned)n_wide));
  }
  int rc = IMPED_OK;
  for (int c : large) {
    const int64_t base = chain_base(kb, c);
    double2* X = evecs + base * nev;
    TimedLaunch tl("eigen_flat", s);
    const int r = real ? eigen_flat_chain<double>(kb, c, nev, tol, max_steps, seed, evals + (size_t)c * nev, X,
                                                  steps + c, s)
                       : eigen_flat_chain<double2>(kb, c, nev, tol, max_steps, seed, evals + (size_t)c * nev, X,
                                                   steps + c, s);
    const int32_t ok = r == IMPED_OK;
    HIP_CHECK(hipMemcpyAsync(conv + c, &ok, sizeof(ok), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));  // ok lives on this stack frame
    if (r != IMPED_OK && r != IMPED_ERR_NOCONV) rc = r;
  }
  return rc;
}

int run_solve(const KrylovBatch& kb, const double2* y, int S, const double* z_host, double tol, int max_steps,
              double2* X, int32_t* steps_host, double* resid_host, hipStream_t s) {
  KScratch& sc = kscratch();
  const int C = kb.C;
  if (S < 1 || max_steps < 1 || max_steps > 3500) {
    set_error("shifted solve: need at least one shift and max_steps in [1, 3500]");
    return IMPED_ERR_INVALID;
  }
  for (size_t q = 0; q < (size_t)C * S; ++q)
    if (!(z_host[2 * q + 1] != 0.0) || !std::isfinite(z_host[2 * q]) || !std::isfinite(z_host[2 * q + 1])) {
      set_error("shifted solve: every shift needs a finite, nonzero imaginary part");
      return IMPED_ERR_INVALID;
    }
  if (C == 0) return IMPED_OK;
  std::vector<int64_t> voff(C + 1, 0);
  int64_t R = 0, nmax = 1;
  for (int c = 0; c < C; ++c) {
    voff[c + 1] = voff[c] + std::min<int64_t>(max_steps, kb.h_n[c]) * kb.h_n[c];
    R += kb.h_n[c];
    nmax = std::max(nmax, kb.h_n[c]);
  }
  h2d(sc.voff, voff.data(), C + 1, s);
  h2d(sc.sz, z_host, (size_t)C * S * 2, s);
  sc.vstore.reserve(std::max<int64_t>(voff[C], 1) * 16);
  sc.wbuf.reserve(std::max<int64_t>(R, 1) * 16);
  sc.ab.reserve((size_t)C * 2 * max_steps * 8);
  sc.scz.reserve((size_t)C * S * 2 * max_steps * 16);
  sc.sres.reserve((size_t)C * S * 8);
  sc.ssteps.reserve((size_t)C * 4);
  sc.conv.reserve((size_t)C * 4);
  std::vector<int32_t> order(C);
  for (int c = 0; c < C; ++c) order[c] = c;
  std::stable_sort(order.begin(), order.end(), [&](int p, int q) { return kb.h_n[p] > kb.h_n[q]; });
  h2d(sc.clist, order.data(), C, s);
  SolveArgs a{};
  a.row_base = kb.row_base;
  a.chain_n = kb.chain_n;
  a.rowptr = kb.rowptr;
  a.col = kb.col;
  a.val = kb.val;
  a.y = y;
  a.z = sc.sz.as<double2>();
  a.S = S;
  a.max_steps = max_steps;
  a.tol = tol;
  a.vstore = sc.vstore.as<double2>();
  a.voff = sc.voff.as<int64_t>();
  a.wbuf = sc.wbuf.as<double2>();
  a.ab = sc.ab.as<double>();
  a.cz = sc.scz.as<double2>();
  a.steps = sc.ssteps.as<int32_t>();
  a.resid = sc.sres.as<double>();
  a.conv = sc.conv.as<int32_t>();
  a.chain_list = sc.clist.as<int32_t>();
  a.dgks = (int)g_tuning.eigen_dgks;
  a.check_every = (int)std::max<int64_t>(1, g_tuning.eigen_check_every);
  const size_t shm = (size_t)(max_steps + 2) * sizeof(double2);
  {
    TimedLaunch tl("solve", s);
    if (g_tuning.eigen_bt == 256)
      hipLaunchKernelGGL(solve_kernel<256>, dim3(C), dim3(256), shm, s, a);
    else if (g_tuning.eigen_bt == 512 || g_tuning.eigen_bt == 0)  // automatic: 512
      hipLaunchKernelGGL(solve_kernel<512>, dim3(C), dim3(512), shm, s, a);
    else
      hipLaunchKernelGGL(solve_kernel<1024>, dim3(C), dim3(1024), shm, s, a);
    HIP_CHECK(hipGetLastError());
  }
  {
    TimedLaunch tl("solve_combine", s);
    hipLaunchKernelGGL(solve_combine_kernel, dim3(blocks_for(nmax, 256), C), dim3(256), 0, s, kb.row_base,
                       kb.chain_n, sc.clist.as<int32_t>(), sc.vstore.as<double2>(), sc.voff.as<int64_t>(),
                       sc.scz.as<double2>(), sc.ssteps.as<int32_t>(), S, max_steps, X);
    HIP_CHECK(hipGetLastError());
  }
  std::vector<int32_t> conv(C);
  d2h(conv.data(), sc.conv.p, C, s);
  if (steps_host) d2h(steps_host, sc.ssteps.p, C, s);
  if (resid_host) d2h(resid_host, sc.sres.p, (size_t)C * S, s);
  stream_sync(s);
  for (int c = 0; c < C; ++c)
    if (!conv[c]) return IMPED_ERR_NOCONV;
  return IMPED_OK;
}

void run_cf(int C, int K, const double* alpha, const double* beta, const int32_t* kout, const double* energy,
            const double* norm2, const int32_t* sign, const int32_t* zflag, const double* kT, const int32_t* spec_id,
            int n_spec, const double* omega, int n_w, double delta, double* g_out, double* spec_out, hipStream_t s,
            const double* norm2_dev) {
  KScratch& sc = kscratch();
  // chains grouped by spectrum, in chain order
  std::vector<int32_t> off(n_spec + 1, 0), chains;
  for (int c = 0; c < C; ++c)
    if (spec_id[c] >= 0 && spec_id[c] < n_spec) off[spec_id[c] + 1]++;
  for (int q = 0; q < n_spec; ++q) off[q + 1] += off[q];
  chains.assign(std::max(off[n_spec], 1), 0);
  std::vector<int32_t> fill(off.begin(), off.end() - 1);
  for (int c = 0; c < C; ++c)
    if (spec_id[c] >= 0 && spec_id[c] < n_spec) chains[fill[spec_id[c]]++] = c;
  // every per-call input in one copy
  HostPack pk;
  const size_t o_e = pk.add(energy, C), o_n = pk.add(norm2, norm2 ? C : 0), o_s = pk.add(sign, C),
               o_z = pk.add(zflag, C),
               o_off = pk.add(off.data(), n_spec + 1), o_ch = pk.add(chains.data(), chains.size()),
               o_om = pk.add(omega, n_w), o_kt = pk.add(kT, std::max(n_spec, 1));
  pk.upload(sc.cf_in, s);
  if (g_out) sc.gout.reserve((size_t)std::max(C, 1) * n_w * 16);
  sc.sout.reserve((size_t)std::max(n_spec, 1) * n_w * 8);
  CFArgs a{alpha, beta, kout, K, HostPack::at<double>(sc.cf_in, o_e),
           norm2 ? HostPack::at<double>(sc.cf_in, o_n) : norm2_dev,
           HostPack::at<int32_t>(sc.cf_in, o_s), HostPack::at<int32_t>(sc.cf_in, o_z),
           HostPack::at<double>(sc.cf_in, o_kt), HostPack::at<int32_t>(sc.cf_in, o_off),
           HostPack::at<int32_t>(sc.cf_in, o_ch), HostPack::at<double>(sc.cf_in, o_om), n_w, delta,
           g_out ? sc.gout.as<double2>() : nullptr, sc.sout.as<double>()};
  // LDS tile: as many of a spectrum's chains as 56 KiB holds, at least one
  // (a chain longer than 3584 levels takes up to 150 KiB)
  if ((size_t)K * sizeof(double2) > 150 * 1024)
    throw HipError{"continued fraction: at most 9600 Lanczos levels per chain"};
  int max_chains = 1;
  for (int q = 0; q < n_spec; ++q) max_chains = std::max(max_chains, off[q + 1] - off[q]);
  a.tile = (int)std::max<int64_t>(1, std::min<int64_t>(max_chains, (56 * 1024) / ((int64_t)K * 16)));
  const size_t shm = (size_t)a.tile * K * sizeof(double2);
  static std::once_flag big_lds;
  std::call_once(big_lds, [] {
    HIP_CHECK(hipFuncSetAttribute((const void*)cf_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
  });
  if (n_spec > 0 && n_w > 0) {
    TimedLaunch tl("cf", s);
    for (int q0 = 0; q0 < n_spec; q0 += 65535) {
      a.q0 = q0;
      const unsigned rows = (unsigned)std::min(65535, n_spec - q0);
      hipLaunchKernelGGL(cf_kernel, dim3(blocks_for(n_w, 256), rows), dim3(256), shm, s, a);
      HIP_CHECK(hipGetLastError());
    }
  }
  if (g_out) d2h(g_out, sc.gout.p, (size_t)C * n_w * 2, s);
  d2h(spec_out, sc.sout.p, (size_t)n_spec * n_w, s);
  stream_sync(s);
}

}  // namespace imped

// ------------------------------------------------------------------ allocation-free Lanczos
//
// imped_lanczos_noalloc runs the Lanczos kernel of imped_batch_lanczos on one
// chain whose CSR, start vector and outputs the caller holds in device
// memory, with every intermediate (a one-chain header and the Lanczos vectors)
// in the fixed device array `scratch` below instead of memory the library
// allocates.  The array's device address comes from hipGetSymbolAddress once
// per device; the array name itself is never used as an address in host code.

namespace imped {
namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // per device, for the life of the process
constexpr size_t kScratchHeader = 256;              // row_base[2], chain_n[1] (int64), chain_list[1] (int32)
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's scratch array (looked up on first use).
void* scratch_address(int dev) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

// The chain table of a one-chain launch: row_base = {0, n}, chain_n = {n},
// chain_list = {0} (written by a kernel: no host copies, nothing to clear).
__global__ void one_chain_header_kernel(int64_t* h, int64_t n) {
  h[0] = 0;
  h[1] = n;
  h[2] = n;
  reinterpret_cast<int32_t*>(h + 3)[0] = 0;
}

}  // namespace
}  // namespace imped

using imped::HipError;

extern "C" int imped_lanczos_noalloc(const int64_t* rowptr, const int32_t* col, const double* val, int64_t n,
                                     const double* v0, int K, int reorth, double* alpha, double* beta, int32_t* kout,
                                     int32_t* status, void* stream) {
  using namespace imped;
  try {
    if (!rowptr || !col || !val || !v0 || !alpha || !beta || !kout || !status || n < 1 || K < 1) {
      set_error("imped_lanczos_noalloc: null pointer, n < 1 or K < 1");
      return IMPED_ERR_INVALID;
    }
    if (reorth && K > 3500) {
      set_error("imped_lanczos_noalloc: K must be <= 3500 with reorthogonalization");
      return IMPED_ERR_INVALID;
    }
    const int64_t kk = std::min<int64_t>(K, n);
    const double need = (double)kScratchHeader + 48.0 * (double)n + (reorth ? 16.0 * (double)kk * (double)n : 0.0);
    if (need > (double)kScratchBytes) {
      set_error("imped_lanczos_noalloc: the chain needs " + std::to_string((long long)need) +
                " bytes of scratch (48 n, plus 16 n min(K, n) with reorthogonalization); the fixed scratch holds " +
                std::to_string((long long)kScratchBytes));
      return IMPED_ERR_CAPACITY;
    }
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
      set_error("imped_lanczos_noalloc: device index out of range");
      return IMPED_ERR_INVALID;
    }
    unsigned char* base = static_cast<unsigned char*>(scratch_address(dev));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    int64_t* head = reinterpret_cast<int64_t*>(base);
    hipLaunchKernelGGL(one_chain_header_kernel, dim3(1), dim3(1), 0, s, head, n);
    HIP_CHECK(hipGetLastError());
    LanczosArgs a{};
    a.row_base = head;
    a.chain_n = head + 2;
    a.chain_list = reinterpret_cast<const int32_t*>(head + 3);
    a.rowptr = rowptr;
    a.col = col;
    a.val = reinterpret_cast<const double2*>(val);
    a.v0 = reinterpret_cast<const double2*>(v0);
    a.K = K;
    a.reorth = reorth;
    a.ws = reinterpret_cast<double2*>(base + kScratchHeader);
    a.vstore = reorth ? a.ws + 3 * n : nullptr;
    a.alpha = alpha;
    a.beta = beta;
    a.kout = kout;
    a.status = status;
    const size_t shm = reorth ? (size_t)(K + 2) * sizeof(double2) : 0;
    // the size buckets of run_lanczos, so a chain gets the same kernel (and
    // the same arithmetic) as through imped_batch_lanczos
    if (reorth && g_tuning.lanczos_lds && lds_chain_bytes(n, K) <= LDS_CHAIN_BYTES)
      launch_lanczos_lds(a, 1, lds_chain_bytes(n, K), s);
    else if (n <= g_tuning.lanczos_wave_max_n)
      hipLaunchKernelGGL(lanczos_kernel<64>, dim3(1), dim3(64), shm, s, a);
    else if (n <= g_tuning.lanczos_wg256_max_n)
      hipLaunchKernelGGL(lanczos_kernel<256>, dim3(1), dim3(256), shm, s, a);
    else
      hipLaunchKernelGGL(lanczos_kernel<1024>, dim3(1), dim3(1024), shm, s, a);
    HIP_CHECK(hipGetLastError());
  } catch (const HipError& e) {
    set_error(e.msg);
    return IMPED_ERR_HIP;
  } catch (const std::exception& e) {
    set_error(e.what());
    return IMPED_ERR_INVALID;
  }
  return IMPED_OK;
}
