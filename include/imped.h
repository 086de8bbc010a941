/* C ABI of libimped.so, the MI355X exact-diagonalization engine.
 *
 * impurityModel has no FFI: its hot path is a Python call surface.  Each
 * entry point below replaces one reference function (cited as
 * impurityModel/ed/<file>.py:<lines>); the Python package impuritymodel_amd
 * binds them with ctypes and keeps the reference's names, arguments and
 * exceptions (finite.applyOp, finite.expand_basis_and_hamiltonian, ...).
 *
 * Conventions
 *  - Product states are W packed uint64 words (W = 1 or 2), spin-orbital i in
 *    word i/64 at bit 63 - i%64; word-wise lexicographic order equals the
 *    order of the reference's bytes keys (product_state_representation.py:131-161).
 *  - Complex numbers are interleaved (re, im) doubles.
 *  - All pointers are host pointers unless stated otherwise; handles own
 *    their device memory.  Calls are synchronous on the library's stream.
 *  - Every function returns IMPED_OK (0) or a negative status; the message
 *    of the last error on the calling thread is imped_last_error().
 */
#ifndef IMPED_H
#define IMPED_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IMPED_ABI_VERSION 1

/* status codes */
#define IMPED_OK 0
#define IMPED_ERR_INVALID -1   /* bad argument */
#define IMPED_ERR_HIP -2       /* HIP runtime error */
#define IMPED_ERR_CAPACITY -3  /* size limit exceeded */
#define IMPED_ERR_BREAKDOWN -4 /* Lanczos beta == 0 (finite.py:1963, 1988) */
#define IMPED_ERR_NOCONV -5    /* eigensolver did not converge */

typedef struct imped_ops_s* imped_ops;
typedef struct imped_vec_s* imped_vec;
typedef struct imped_batch_s* imped_batch;
typedef struct imped_affine_s* imped_affine;

const char* imped_last_error(void);
int imped_abi_version(void);
/* Stream for all later calls (a hipStream_t; NULL = the null stream). */
int imped_set_stream(void* stream);
int imped_device_synchronize(void);

/* Launch timing: when enabled, every kernel launch of the engine is bracketed
 * by HIP events on the library's stream; imped_timing_query sums the elapsed
 * time of the launches recorded under a kernel name ("apply_count",
 * "apply_fill", "discover", "sort" (the span of one BFS layer's sort passes),
 * "csr_fill", "affine_fill", "affine_seed", "eigen", "eigen_flat", "lanczos",
 * "spmv", "solve", "solve_combine", "cf", "mlp_gemm", "mlp_adam").  No reference counterpart. */
int imped_timing_enable(int on);
int imped_timing_reset(void);
int imped_timing_query(const char* name, double* total_ms, int64_t* launches);

/* Launch-shape parameters picked per chain-dimension bucket; no reference
 * counterpart.  Names: "lanczos_wave_max_n", "lanczos_wg256_max_n" (Lanczos
 * threads per chain: 64 / 256 / 1024 by size), "eigen_wave_max_n" (eigensolver
 * chains up to this size get one wave), "eigen_bt" (threads per larger
 * eigensolver chain: 256, 512, 1024; 0, the default: 1024 when the launch's
 * largest chain has more than 512 rows, else 512), "eigen_real", "eigen_dgks",
 * "eigen_check_every" (Ritz convergence check every this many steps, 12 by
 * default), "eigen_check_start" (no Ritz convergence check before
 * min(this, n/2) steps of an n-row chain, 48 by default; the last step is
 * always checked), "eigen_dense_max_n" (eigensolver chains of up to this
 * many rows, at most 112, are solved densely: Householder tridiagonalization
 * in LDS, exact for degenerate eigenvalues, steps reported as n; 0, the
 * default: off), "lanczos_dense_max_n" (fully reorthogonalized Lanczos
 * chains of up to this many rows, at most 112, from unit start vectors, are
 * computed as a Householder reduction started by the reflector of the start
 * vector: the same alpha and beta, no Krylov basis in LDS, five 60-row
 * chains per CU; 0, the default: off), "flat_min_n" (chains above it use the grid-wide
 * kernels), "spmv_lpr" (lanes per row of the grid-wide SpMV: 4..64),
 * "real_values" (1, the default: when every value of a batch's CSR is real,
 * the grid-wide kernels read an 8-byte copy of the real parts, made once per
 * CSR; 0: the complex values),
 * "lanczos_graph_steps" (grid-wide Lanczos steps per replayed HIP graph,
 * 0..256; 0 launches every step directly, as does switched-on timing),
 * "eigen_lds_vectors" (cap on the Krylov vectors the eigensolver keeps in
 * LDS; -1 as many as fit, 0 none), "small_sort" (1: key sets of at most
 * 16384 records whose chain, key and position fit 63 bits are sorted by one
 * workgroup; 0: always the digit passes), "mlp_graph" (1: a surrogate
 * prediction's kernels run as one replayed HIP graph, which shortens the
 * GPU's span but costs the calling thread more than the launches it
 * replaces; 0, the default: direct launches, as with switched-on timing),
 * "lanczos_lds" (1, the default: a fully reorthogonalized chain whose Krylov
 * basis fits 80 KiB runs on one workgroup with the basis in LDS; 0: the
 * basis in memory), "lanczos_lds_waves" (waves of that workgroup: 1, 2 or
 * 4, the default; the Gram-Schmidt sums are split over the waves). */
int imped_set_tuning(const char* name, int64_t value);

/* The engine's device radix sort (the one that orders BFS layers and
 * applyOp outputs), exposed for testing: stable sort of the index array
 * perm[n] by keys[perm[i]] on bits [begin_bit, end_bit).  No reference
 * counterpart (the reference calls Python's sorted(), finite.py:1749). */
int imped_sort_u64(const uint64_t* keys, int64_t n, int begin_bit, int end_bit, int32_t* perm);

/* get_basis (finite.py:294-374) on the host: the restricted initial basis in
 * the reference's itertools order.  shell[10*l ..] per angular-momentum shell:
 * {first impurity index, impurity count, first valence-bath index, valence
 * count, first conduction-bath index, conduction count, n0imp, dnTol,
 * dnValBaths, dnConBaths}; shells combined like itertools.product (first
 * shell slowest).  keys == NULL only counts (*n_out). */
int imped_enumerate_basis(int n_orb, int W, int n_shells, const int32_t* shell, int64_t cap, int64_t* n_out,
                          uint64_t* keys);

/* Compiled operators of n_models models (the process dicts of applyOp,
 * finite.py:1301-1340, plus the occupation restrictions checked at
 * finite.py:1382-1385).  Model m owns groups [model_group_off[m],
 * model_group_off[m+1]); group g maps a state s to s ^ group_flip[g] and owns
 * terms [group_term_off[g], group_term_off[g+1]).  Term t is non-zero on s iff
 * (s & care) == val with care, val, sgn = term_masks[t*3W + {0, W, 2W}], has
 * sign (-1)^(popcount(s & sgn) + (term_meta[t] & 1)), value term_coef[2t..2t+1]
 * and dict position term_meta[t] >> 1.  Restriction r of model m:
 * restr_lohi[2r] <= popcount(s & restr_mask[rW..]) <= restr_lohi[2r+1]. */
int imped_ops_create(int W, int n_orb, int n_models, const int32_t* model_group_off, const int32_t* group_term_off,
                     const uint64_t* group_flip, const uint64_t* term_masks, const uint32_t* term_meta,
                     const double* term_coef, const int32_t* model_restr_off, const uint64_t* restr_mask,
                     const int32_t* restr_lohi, imped_ops* out);
/* Replaces the term values of an operator set with the same term structure
 * (n_terms = the term count given at creation, values in the same order): a
 * parameter sweep over one model family compiles its structure once.  Launches
 * already queued on the stream use the old values.  No reference
 * counterpart: the reference rebuilds its operator dicts per parameter set. */
int imped_ops_set_coef(imped_ops ops, const double* term_coef, int64_t n_terms);
/* Term values of an affine model family on the device: every model holds
 * terms_per_model terms of one structure, and model m < n_models gets
 * value[t] = base[t] + sum_p theta[m][p] dirs[p][t] (complex base/dirs as
 * interleaved doubles, real theta [n_models][n_params]); models from n_models
 * on keep their values.  Only theta changes from one parameter batch to the
 * next.  No reference counterpart (the reference assembles each hOp dict term
 * by term: get_spectra.py:123-176 with finite.py:271-374). */
int imped_ops_set_affine(imped_ops ops, int n_params, int64_t terms_per_model, const double* base, const double* dirs,
                         int n_models, const double* theta);
int imped_ops_destroy(imped_ops ops);

/* ---- sparse states (multi-configurational dicts {state: amplitude}) ---- */

/* n_vec vectors, vector v = entries [off[v], off[v+1]) of keys/amps. */
int imped_vec_create(int W, int n_vec, const int64_t* off, const uint64_t* keys, const double* amps, imped_vec* out);
int imped_vec_destroy(imped_vec vec);
int imped_vec_offsets(imped_vec vec, int64_t* off);
/* keys, amps (and norm2 per vector when computed) copied to host. */
int imped_vec_get(imped_vec vec, uint64_t* keys, double* amps, double* norm2);

/* applyOp(n, op, psi, slaterWeightMin, restrictions) (finite.py:1301-1511,
 * branches without opResult): output vector o is operator model out_model[o]
 * applied to source vector out_src[o].  Targets failing the model's
 * restrictions are dropped, |amp|^2 < weight_min is pruned after summation and
 * the dict insertion order is kept.  normalize != 0 divides each output by its
 * norm as getSpectra does (spectra.py:786-790); norm2 is stored either way. */
int imped_vec_apply(imped_ops ops, imped_vec src, int n_out, const int32_t* out_src, const int32_t* out_model,
                    double weight_min, int normalize, imped_vec* out);

/* The pruning of getRIXSmap after normalizing psi3 (spectra.py:1108-1111):
 * a new vector set with the entries of src whose |amp|^2 >= weight_min, in
 * stored order; src's per-vector norm2 is carried over. */
int imped_vec_prune(imped_vec src, double weight_min, imped_vec* out);

/* Eigenvector dicts of eigensystem (finite.py:114-117): output o holds the
 * components of eigenvector which[o] of chain chain[o] with
 * |v_i|^2 >= weight_min, in basis order, not renormalized.  After
 * imped_batch_solve, which[o] names a shift and the vector is that solution
 * (psi2 of getRIXSmap, spectra.py:1087-1090, with weight_min = 0). */
int imped_vec_from_eigen(imped_batch batch, int n_out, const int32_t* chain, const int32_t* which, double weight_min,
                         imped_vec* out);

/* ---- batches of Hilbert-space sectors ---- */

int imped_batch_create(imped_ops ops, int n_chains, const int32_t* chain_model, imped_batch* out);
int imped_batch_destroy(imped_batch batch);
/* Seed states of every chain (and optional start amplitudes). */
int imped_batch_set_seeds(imped_batch batch, const int64_t* seed_off, const uint64_t* seed_keys,
                          const double* seed_amps);
/* Seeds and start amplitudes from a vector set (vector c -> chain c). */
int imped_batch_seed_from_vec(imped_batch batch, imped_vec vec);

/* fixed_basis = 1: get_hamiltonian_matrix (finite.py:1514-1610): the seeds
 *   are the basis, the operator is applied without restrictions, targets
 *   outside the basis are dropped.
 * fixed_basis = 0: expand_basis_and_hamiltonian (finite.py:1696-1894): BFS
 *   closure of the seeds under the operator with the model's restrictions,
 *   seeds first, then each new layer in ascending key order.
 * Values with |h|^2 < weight_min are pruned (1e-12 in the reference).
 * max_layers > 0 stops the closure early (testing). */
int imped_batch_build(imped_batch batch, int fixed_basis, double weight_min, int max_layers);
/* Use a given CSR instead of building one (the matrix argument h of
 * get_tridiagonal_krylov_vectors, finite.py:1897): rows of H with chain-local
 * int32 columns, rowptr int64 over all chains (rows of chain c follow chain
 * c-1), complex values.  Columns are checked against the chain sizes. */
int imped_batch_set_csr(imped_batch batch, const int64_t* chain_n, const int64_t* rowptr, const int32_t* col,
                        const double* val);
/* chain_n[C], chain_nnz[C], totals[2] = {rows, nnz}; any may be NULL. */
int imped_batch_sizes(imped_batch batch, int64_t* chain_n, int64_t* chain_nnz, int64_t* totals);
/* Basis keys of all chains, rows*W words, chain after chain. */
int imped_batch_get_basis(imped_batch batch, uint64_t* keys);
/* CSR rows (H row r = conj of the operator applied to basis state r, i.e.
 * rowptr[rows+1] int64, chain-local int32 columns, complex values);
 * the reference's matrix h[i, j] = <i|H|j> is its conjugate transpose
 * mapped back: h[col, row] = conj(val). */
int imped_batch_get_csr(imped_batch batch, int64_t* rowptr, int32_t* col, double* val);

/* ---- fixed sparsity patterns with parameter-affine values ---- */

/* n_struct CSR structures, structure s with struct_n[s] rows: its rowptr
 * (struct_n[s]+1 entries, starting at 0) follows structure s-1's in rowptr,
 * its entries follow s-1's in col, base (complex) and, per parameter p,
 * dirs[p] (complex, [n_params][total entries]).  A structure's values for
 * parameters theta are base + sum_p theta_p dirs[p], entry by entry.  Used
 * for sectors whose basis and pattern do not change over a model family: the
 * Hamiltonian of a sector of an affine family, and (n_params = 0) a fixed
 * transition operator from a ground sector's rows (columns) into an excited
 * sector's rows.  No reference counterpart: the reference rebuilds every
 * sector per model (finite.py:1514-1610, 1696-1894). */
int imped_affine_create(int n_struct, const int64_t* struct_n, const int64_t* rowptr, const int32_t* col,
                        const double* base, int n_params, const double* dirs, imped_affine* out);
int imped_affine_destroy(imped_affine aff);
/* The batch's CSR from structures: chain c is square structure chain_struct[c]
 * with the values of parameter set chain_model[c] of theta [n_models][n_params],
 * entries with |h|^2 < weight_min set to zero (the pruning of
 * get_hamiltonian_matrix / expand_basis_and_hamiltonian).  Replaces
 * imped_batch_build for such sectors; formed on the device. */
int imped_batch_set_affine(imped_batch batch, imped_affine aff, const int32_t* chain_struct,
                           const int32_t* chain_model, int n_models, const double* theta, double weight_min);
/* Start vectors of the batch's chains on the device, for imped_batch_lanczos
 * with v0 = NULL: chain c gets T psi, T = structure chain_struct[c] of `trans`
 * (rows: chain c's rows; columns: rows of chain src_chain[c] of `ground`),
 * psi = eigenvector src_vec[c] of that chain after imped_batch_eigen with
 * components |psi_i|^2 < psi_weight_min dropped (the eigenvector dicts of
 * eigensystem, finite.py:114-117); entries |v|^2 < weight_min are pruned and
 * v is normalized, norm2[c] = |v|^2 before normalization (applyOp followed
 * by getSpectra's normalization, finite.py:1301-1511, spectra.py:786-790).
 * norm2 may be NULL: the norms then stay on the device (for imped_batch_cf
 * with norm2 NULL) and the call returns without waiting for the GPU. */
int imped_batch_seed_affine(imped_batch batch, imped_batch ground, imped_affine trans, const int32_t* src_chain,
                            const int32_t* src_vec, const int32_t* chain_struct, double psi_weight_min,
                            double weight_min, double* norm2);

/* The thermal, polarization-summed XAS of n_models parameter sets (theta
 * [n_models][n_params], kT [n_models] > 0) of an affine family from its
 * sector structures, in one call: the batched composition of eigensystem
 * (finite.py:58-118), the energy cut (get_spectra.py:302-303), applyOp of the
 * dipole operators, expand_basis_and_hamiltonian, Lanczos and getSpectra
 * with thermal_average (spectra.py:222-238, 743-806; average.py:11-45).
 *  - ground chain (m, k), k < n_comp: structure k of `ground`, theta row m,
 *    values pruned at h_weight_min; lowest nev eigenpairs (imped_batch_eigen
 *    with eig_tol, eig_max_steps, seed);
 *  - per model, the lowest n_psi_max of its components' eigenvalues in
 *    ascending order (ties in component order), kept while within
 *    energy_cut * kT[m] of the lowest;
 *  - per kept state (component k) and polarization t < n_pol, in (model,
 *    polarization, state) order, an excited chain: structure
 *    sector_index[k][t] of `sector` (theta row m), started from T psi with
 *    structure sector_index[k][t] of `trans` (imped_batch_seed_affine with
 *    psi_weight_min, weight_min);
 *  - Lanczos with K = min(krylov_size, largest excited chain) steps (reorth
 *    as imped_batch_lanczos), continued fractions on omega[n_w] with
 *    broadening delta, Boltzmann weights over each model's kept states.
 * Outputs: spectra [n_models][n_w], energies [n_models][n_psi_max] (the kept
 * ones, then NaN), n_kept [n_models]; may be NULL: eig_steps [n_models *
 * n_comp], chain_struct and kout [n_models * n_pol * n_psi_max],
 * info[4] = {excited chains, K, 0, 0}.  Two waits for the GPU per call (the
 * eigenvalues, then the spectra).  Returns IMPED_ERR_NOCONV or
 * IMPED_ERR_BREAKDOWN as imped_batch_eigen and imped_batch_lanczos report.
 * No reference counterpart as one call. */
int imped_family_xas(imped_affine ground, imped_affine sector, imped_affine trans, int n_comp, int n_pol,
                     const int32_t* sector_index, int n_models, const double* theta, const double* kT, int nev,
                     int n_psi_max, double energy_cut, int krylov_size, int reorth, double eig_tol,
                     int eig_max_steps, uint64_t seed, double h_weight_min, double psi_weight_min,
                     double weight_min, const double* omega, int n_w, double delta, double* spectra,
                     double* energies, int32_t* n_kept, int32_t* eig_steps, int32_t* chain_struct, int32_t* kout,
                     int32_t* info);

/* get_tridiagonal_krylov_vectors (finite.py:1897-1995) for every chain from
 * v0 (rows complex, or NULL = the start vectors of imped_batch_seed_affine
 * when the batch has them, else the seed amplitudes in the first rows of each
 * chain).  reorth = 0 is the reference's plain recurrence; reorth = 1 adds
 * full reorthogonalization (no reference counterpart).  alpha, beta are
 * [C][K]; kout[c] = min(K, n_c), or the step of a beta == 0 breakdown with
 * status[c] = IMPED_ERR_BREAKDOWN.  With alpha, beta, kout and status all
 * NULL the results stay on the device (for imped_batch_cf) and the call
 * returns without waiting for the GPU. */
int imped_batch_lanczos(imped_batch batch, const double* v0, int K, int reorth, double* alpha, double* beta,
                        int32_t* kout, int32_t* status);

/* Allocation-free variant of imped_batch_lanczos for one chain: the same
 * Lanczos kernel (plain recurrence, or reorth = 1) on a CSR, start vector and
 * outputs that the caller holds in DEVICE memory -- rowptr[n+1] int64, col
 * int32, val complex, v0[n] complex, alpha[K], beta[K], kout[1], status[1] --
 * queued on `stream` (a hipStream_t; NULL = the null stream) without
 * synchronizing.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops).  It calls no hipMalloc/hipFree: the Lanczos vectors live in
 * a fixed 64 MiB device array per GPU, which limits a chain to
 * 256 + 48 n (+ 16 n min(K, n) with reorthogonalization) bytes, i.e. about
 * 1.39e6 rows plain; larger chains return IMPED_ERR_CAPACITY.  Calls that
 * overlap on one GPU share that array and race; the array keeps the last
 * call's contents.  Returns after queuing the work. */
int imped_lanczos_noalloc(const int64_t* rowptr, const int32_t* col, const double* val, int64_t n, const double* v0,
                          int K, int reorth, double* alpha, double* beta, int32_t* kout, int32_t* status,
                          void* stream);

/* eigsh(h, k=nev, which="SA", tol) of eigensystem (finite.py:95-109) for every
 * chain: evals[C][nev] ascending, evecs (may be NULL) rows*nev complex with
 * chain c's block at row_base[c]*nev and eigenvector k at offset k*n_c.
 * Returns IMPED_ERR_NOCONV if some chain missed tol within max_steps. */
int imped_batch_eigen(imped_batch batch, int nev, double tol, int max_steps, uint64_t seed, double* evals,
                      double* evecs, int32_t* steps);

/* The resolvent solves of getRIXSmap (spectra.py:1066-1080, which calls
 * scipy.sparse.linalg.bicgstab on (wIn + i delta1 + e) I - H):
 *   (z_{c,s} I - H_c) x_{c,s} = y_c,  s < n_shift,
 * y_c = the chain's seed amplitudes (imped_batch_seed_from_vec), shifts
 * [C][n_shift] complex with nonzero imaginary parts.  All shifts of a chain
 * share one Krylov space (Lanczos with full reorthogonalization, Galerkin
 * solution); tol bounds the relative residual |y - (z - H) x| / |y| of every
 * shift.  Solutions replace the eigenvectors of the batch (layout of
 * imped_batch_eigen with nev = n_shift) for imped_vec_from_eigen.
 * steps[C], resid[C][n_shift] and x (a host copy of the solutions, complex,
 * rows * n_shift) may be NULL.  Returns IMPED_ERR_NOCONV (with the solutions
 * written) when some residual stays above tol. */
int imped_batch_solve(imped_batch batch, int n_shift, const double* shifts, double tol, int max_steps,
                      int32_t* steps, double* resid, double* x);

/* Continued fractions (spectra.py:677-680) of every chain after
 * imped_batch_lanczos, on omega[n_w] with broadening delta:
 *   g_c(w) = sign_c * norm2_c / (z - a0 - b0^2/(z - a1 - ...)),
 *   z = sign_c*(w + i delta) + energy_c
 * (sign -1 is the photoemission convention, spectra.py:136-137), written to
 * g_out[C][n_w] (complex, may be NULL); and spectra
 *   spec[q][w] = sum_{c in q} p_c * (-Im g_c(w)),
 *   p_c = exp(-(e_c - min e)/kT[q]) / Z, Z summed over chains with zflag = 1
 * (thermal_average, average.py:11-45, summed over operators).  norm2 NULL:
 * the norms imped_batch_seed_affine left on the device. */
int imped_batch_cf(imped_batch batch, const double* energy, const double* norm2, const int32_t* sign,
                   const int32_t* zflag, const double* kT, const int32_t* spec_id, int n_spec, const double* omega, int n_w,
                   double delta, double* g_out, double* spec_out);

/* ---- surrogate network (no reference counterpart) ----
 *
 * An ensemble of n_members fully connected networks of layer widths
 * dims[0..n_layers] with GELU between layers, for the parameters -> XAS
 * spectrum surrogate of the fitting loop.  GEMMs run on the matrix cores in
 * BF16 with FP32 accumulation; master weights, gradients and Adam moments are
 * FP32.  Member e's parameters are params[e*P .. e*P + P) (P from
 * imped_mlp_param_count): W_0 [dims[1]][dims[0]] row-major, b_0 [dims[1]],
 * W_1, b_1, ..., then zero padding. */
typedef struct imped_mlp_s* imped_mlp;
int imped_mlp_create(int n_members, int n_layers, const int32_t* dims, int max_batch, imped_mlp* out);
int imped_mlp_destroy(imped_mlp net);
int imped_mlp_param_count(imped_mlp net, int64_t* per_member);
/* Loads all members' parameters [n_members][P] and resets the Adam state. */
int imped_mlp_set_params(imped_mlp net, const float* params);
int imped_mlp_get_params(imped_mlp net, float* params);
/* y[e][b][dims[L]] = network e applied to x[b][dims[0]] (inputs rounded to BF16),
 * as predictions form it: hidden-layer GELU with erf by Abramowitz and Stegun
 * 7.1.26 (absolute error <= 1.5e-7, below the BF16 rounding of the
 * activations); training uses the full-precision erf. */
int imped_mlp_forward(imped_mlp net, const float* x, int batch, float* y);
/* Forward pass plus the ensemble statistics of each sample b, on the device:
 * member spectra m_e = max(y_e / scale, 0) / sum_d weights[d] max(y_e / scale, 0),
 * mean[b][d] = their average renormalized to sum_d weights[d] mean = 1, and
 * disagreement[b] = (1/n_members) sum_e sum_d weights[d] |m_e,d - mean[b][d]|
 * (trapezoid weights of the output grid give unit-area spectra and mean L1
 * distances).  No reference counterpart. */
int imped_mlp_predict(imped_mlp net, const float* x, int batch, const double* weights, double scale, double* mean,
                      double* disagreement);
/* The same statistics plus each sample's misfit to a target spectrum,
 * misfit[b] = sum_d weights[d] |mean[b][d] - target[d]| (the L1 distance of
 * unit-area spectra when target has unit area), computed on the device: a
 * fitting loop's Metropolis decisions read back 2 x batch doubles instead of
 * batch x dims[L].  mean may be NULL (the means stay on the device).  No
 * reference counterpart. */
int imped_mlp_misfit(imped_mlp net, const float* x, int batch, const double* weights, double scale,
                     const double* target, double* disagreement, double* misfit, double* mean);
/* Metropolis walkers of a surrogate fitting loop over n_params parameters
 * (the network's input width) in the box [lo, hi].  A step lets every walker
 * that is not waiting propose p = theta + step (hi - lo) z (z standard
 * normal; reflected at the box, then clipped), predicts all proposals in one
 * imped_mlp_misfit call (inputs 2 (p - lo) / (hi - lo) - 1, misfit to
 * target), and each trusted proposal (disagreement < threshold) is accepted
 * with probability min(1, exp(-(misfit - cur) / temperature)); untrusted
 * proposals make their walker wait until imped_walkers_decide answers them
 * (misfits from exact evaluation, walkers in ascending order).  Random
 * numbers: two SplitMix64 streams, seeded with seed (normals: two uniforms,
 * the top 53 bits of an output each, give two normals by Box-Muller, the
 * cosine one first) and with ~seed (one uniform per decision), each consumed in walker
 * order; normals are drawn ahead while the GPU predicts.  The host loop of a
 * fitting run with one library call per step; no reference counterpart (the
 * reference has no surrogate or fitting loop). */
typedef struct imped_walkers_s* imped_walkers;
int imped_walkers_create(imped_mlp net, int n_walkers, int n_params, const double* theta, const double* cur,
                         const double* lo, const double* hi, double step, double temperature, double threshold,
                         uint64_t seed, imped_walkers* out);
int imped_walkers_destroy(imped_walkers w);
/* One step; weights and target hold the network's output width, and
 * *n_waiting = walkers waiting for exact evaluation afterwards.  The walkers
 * keep a pointer to their network: destroy them first. */
int imped_walkers_step(imped_walkers w, const double* weights, double scale, const double* target,
                       int32_t* n_waiting);
/* The waiting walkers (ascending) and their proposals [count][n_params];
 * idx and proposals (either may be NULL) must hold n_walkers rows. */
int imped_walkers_waiting(imped_walkers w, int32_t* idx, double* proposals, int32_t* count);
int imped_walkers_decide(imped_walkers w, int32_t count, const int32_t* idx, const double* misfit);
/* theta [n_walkers][n_params], cur [n_walkers], best misfit so far and the
 * counters {predicted, served, accepted, waiting} (any may be NULL). */
int imped_walkers_state(imped_walkers w, double* theta, double* cur, double* best, int64_t* counters);
/* Rows [start, start + count) of the recorded served proposals' parameters
 * (the first record cap of them since the last counter reset). */
int imped_walkers_served(imped_walkers w, int64_t start, int64_t count, double* params);
/* Records at most max_rows served proposals (8 n_params bytes each) between
 * counter resets; later ones are counted but not recorded.  The default cap
 * is 2^20 rows (64 MiB at 8 parameters), so a long fitting run does not grow
 * host memory without bound.  No reference counterpart. */
int imped_walkers_set_record_cap(imped_walkers w, int64_t max_rows);
/* Zeroes the counters and forgets the served proposals. */
int imped_walkers_reset_counters(imped_walkers w);
/* Page-locked host memory (hipHostMalloc) for outputs read back every call:
 * device-to-host copies into it run at full link speed.  No reference
 * counterpart. */
int imped_pinned_alloc(int64_t bytes, void** out);
int imped_pinned_free(void* p);
/* Gradients [n_members][P] of the mean squared error mean_{b,j} (y_e - t)^2
 * over one batch (x, t), and the loss of each member (either may be NULL). */
int imped_mlp_gradients(imped_mlp net, const float* x, const float* t, int batch, float* grads, double* loss);
/* Training set kept on the device: x [n][dims[0]], t [n][dims[L]]. */
int imped_mlp_set_data(imped_mlp net, const float* x, const float* t, int64_t n);
/* steps Adam steps on minibatches batches[step][batch] (row indices of the
 * training set; every member sees the same rows); losses [steps][n_members]. */
int imped_mlp_train(imped_mlp net, const int32_t* batches, int steps, int batch, float lr, float beta1, float beta2,
                    float eps, double* losses);

#ifdef __cplusplus
}
#endif

#endif /* IMPED_H */
