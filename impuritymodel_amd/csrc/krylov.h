// Krylov-space kernels: batched Lanczos, ground-state eigensolver,
// continued fractions with the thermal epilogue, and sparse-vector helpers.
#pragma once

#include "kernels.h"

namespace imped {

// CSR of all chains of a batch; rows of chain c are [row_base[c], row_base[c]+chain_n[c]),
// column indices are chain-local.
struct KrylovBatch {
  int C;
  const int64_t* row_base;
  const int64_t* chain_n;
  const int64_t* rowptr;
  const int32_t* col;
  const double2* val;
  std::vector<int64_t> h_n, h_nnz;
  // the real parts of val when every value is real (NULL otherwise): the
  // grid-wide SpMV then reads 8 bytes per value instead of 16
  const double* val_re = nullptr;
};

// Plain (reorth = 0) or fully reorthogonalized (reorth = 1) Lanczos from v0
// for every chain; alpha/beta are [C][K], kout[c] = min(K, n_c) (or the
// step of a beta == 0 breakdown, with status[c] = IMPED_ERR_BREAKDOWN).
void run_lanczos(const KrylovBatch& kb, const double2* v0, int K, int reorth, double* alpha, double* beta,
                 int32_t* kout, int32_t* status, hipStream_t s);

// Lowest nev eigenpairs of every chain.  evals [C][nev]; evecs: chain c's
// block starts at row_base[c]*nev, eigenvector k at offset k*n_c.
// real_values: every CSR value has zero imaginary part, so the solver may
// run on real vectors (when the eigen_real tuning flag allows it).  conv
// (device, [C]) gets 1 for a chain that met tol, else 0; nothing is read
// back: the caller reads conv with evals and steps.
int run_eigen(const KrylovBatch& kb, int nev, double tol, int max_steps, uint64_t seed, double* evals,
              double2* evecs, int32_t* steps, int32_t* conv, bool real_values, hipStream_t s);

// (z_{c,s} - H_c) x = y_c for S shifts per chain (z_host: [C][S] complex,
// every Im z != 0), one Krylov space per chain.  y: one entry per batch row.
// X: chain c's block starts at row_base[c]*S, shift s at offset s*n_c (the
// eigenvector layout with nev = S).  resid_host [C][S]: relative residuals.
// Returns IMPED_ERR_NOCONV when some shift stays above tol after max_steps.
int run_solve(const KrylovBatch& kb, const double2* y, int S, const double* z_host, double tol, int max_steps,
              double2* X, int32_t* steps_host, double* resid_host, hipStream_t s);

// Continued fractions of all chains on the omega grid plus the Boltzmann
// average over each spectrum's chains (host arrays are uploaded; with norm2
// NULL the chains' norms are read from norm2_dev on the device).
void run_cf(int C, int K, const double* alpha, const double* beta, const int32_t* kout, const double* energy,
            const double* norm2, const int32_t* sign, const int32_t* zflag, const double* kT, const int32_t* spec_id,
            int n_spec, const double* omega, int n_w, double delta, double* g_out, double* spec_out, hipStream_t s,
            const double* norm2_dev = nullptr);

// out vector o = op[out_model[o]] applied to source vector out_src[o].
void vector_apply(OpsView ops, int W, const std::vector<int64_t>& src_off, const uint64_t* src_keys,
                  const double2* src_amps, int n_out, const int32_t* out_src, const int32_t* out_model,
                  double weight_min, bool normalize, std::vector<int64_t>& out_off, DevBuf& out_keys,
                  DevBuf& out_amps, DevBuf& out_norm2, hipStream_t s, int n_orb);

// Entries of every vector of a set with |amp|^2 >= weight_min, in stored order.
void prune_vec(int W, const std::vector<int64_t>& off, const uint64_t* keys, const double2* amps, double weight_min,
               std::vector<int64_t>& out_off, DevBuf& out_keys, DevBuf& out_amps, hipStream_t s);

// Eigenvector components with |v_i|^2 >= weight_min, in basis order.
void truncate_eigvecs(int W, const uint64_t* basis_keys, const double2* evecs, int nev,
                      const std::vector<int64_t>& row_base, const std::vector<int64_t>& chain_n, int n_out,
                      const int32_t* chain, const int32_t* which, double weight_min, std::vector<int64_t>& out_off,
                      DevBuf& out_keys, DevBuf& out_amps, hipStream_t s);

}  // namespace imped
