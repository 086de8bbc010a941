"""Python handles over libimped.so: operator tables, sparse states, batches.

Thin objects that own a C handle each and convert between numpy arrays and
the C ABI of include/imped.h.  All numerical work happens in the library's
HIP kernels.
"""

import numpy as np

from . import opcompile
from ._lib import IMPED_ERR_BREAKDOWN, IMPED_ERR_NOCONV, C, check, lib, ptr


def _release(destroy, obj, _lib=lib):
    """Destroy a handle once.  The library is bound at definition time: at
    interpreter shutdown the module's globals may already be gone."""
    h = getattr(obj, "h", None)
    if h and _lib is not None:
        fn = getattr(_lib, destroy, None)
        if fn is not None:
            fn(h)
    obj.h = None


def _c64(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a.view(np.float64).reshape(-1) if a.size else np.zeros(2, np.float64)


class Ops:
    """Compiled operators of several models (one restriction set per model)."""

    def __init__(self, ops, restrictions, n_orb, W=None):
        self.n_orb = int(n_orb)
        self.W = W or opcompile.n_words(n_orb)
        self.M = len(ops)
        a = opcompile.compile_models(ops, restrictions, self.W)
        self._arrays = a
        h = C.c_void_p()
        check(
            lib.imped_ops_create(
                self.W, self.n_orb, self.M, ptr(a["model_group_off"]), ptr(a["group_term_off"]),
                ptr(_nz(a["group_flip"])), ptr(_nz(a["term_masks"])), ptr(_nz(a["term_meta"])),
                ptr(_nz(a["term_coef"])), ptr(a["model_restr_off"]), ptr(_nz(a["restr_mask"])),
                ptr(_nz(a["restr_lohi"])), C.byref(h),
            )
        )
        self.h = h

    @property
    def n_terms(self):
        return len(self._arrays["term_meta"])

    def set_coef(self, values):
        """Replace the term values (complex, compiled term order; same structure)."""
        v = _c64(values)
        check(lib.imped_ops_set_coef(self.h, ptr(v), len(v) // 2 if np.size(values) else 0))

    def set_affine(self, base, dirs, theta):
        """Values of the first len(theta) models: base [T1] + theta [m, P] @ dirs [P, T1],
        computed on the device (every model holds the same T1-term structure)."""
        base = np.asarray(base, np.complex128).reshape(-1)
        dirs = np.asarray(dirs, np.complex128).reshape(-1, len(base))
        theta = np.ascontiguousarray(theta, np.float64).reshape(-1, len(dirs))
        check(lib.imped_ops_set_affine(self.h, len(dirs), len(base), ptr(_c64(base)), ptr(_c64(dirs)), len(theta),
                                       ptr(_nz(theta.reshape(-1)))))

    def __del__(self, _release=_release):
        _release("imped_ops_destroy", self)


def _nz(a):
    """Non-empty contiguous array (the C side never reads past the counts)."""
    a = np.ascontiguousarray(a)
    if a.size == 0:
        return np.zeros(1, dtype=a.dtype)
    return a


class Vec:
    """A set of sparse states {key: amplitude} living on the device."""

    def __init__(self, W, handle=None, keys=None, amps=None, offsets=None):
        self.W = W
        if handle is None:
            keys = np.ascontiguousarray(keys, dtype=np.uint64).reshape(-1, W)
            offsets = np.ascontiguousarray(offsets, dtype=np.int64)
            h = C.c_void_p()
            check(lib.imped_vec_create(W, len(offsets) - 1, ptr(offsets), ptr(_nz(keys.reshape(-1))),
                                       ptr(_nz(_c64(amps))), C.byref(h)))
            handle = h
        self.h = handle
        self._off = None

    @property
    def offsets(self):
        if self._off is None:
            raise RuntimeError("offsets must be set by the producer")
        return self._off

    def set_count(self, V):
        off = np.zeros(V + 1, np.int64)
        check(lib.imped_vec_offsets(self.h, ptr(off)))
        self._off = off
        return self

    def norm2(self):
        """Per-vector |psi|^2 recorded by the producer (before normalization)."""
        off = self.offsets
        out = np.zeros(max(len(off) - 1, 1), np.float64)
        check(lib.imped_vec_get(self.h, None, None, ptr(out)))
        return out[: len(off) - 1]

    def prune(self, weight_min):
        """New vector set without the entries of |amp|^2 < weight_min."""
        h = C.c_void_p()
        check(lib.imped_vec_prune(self.h, float(weight_min), C.byref(h)))
        return Vec(self.W, handle=h).set_count(len(self.offsets) - 1)

    def get(self):
        """Return (keys [N, W] uint64, amps [N] complex, norm2 [V] or zeros)."""
        off = self.offsets
        N = int(off[-1])
        keys = np.zeros((max(N, 1), self.W), np.uint64)
        amps = np.zeros(max(N, 1), np.complex128)
        norm2 = np.zeros(max(len(off) - 1, 1), np.float64)
        check(lib.imped_vec_get(self.h, ptr(keys), ptr(amps.view(np.float64)), ptr(norm2)))
        return keys[:N], amps[:N], norm2[: len(off) - 1]

    def __del__(self, _release=_release):
        _release("imped_vec_destroy", self)


def vec_from_host(W, keys_list, amps_list):
    """Vec from lists of per-vector key arrays [n_v, W] and amplitude arrays."""
    off = np.zeros(len(keys_list) + 1, np.int64)
    for v, k in enumerate(keys_list):
        off[v + 1] = off[v] + len(k)
    keys = np.concatenate([np.asarray(k, np.uint64).reshape(-1, W) for k in keys_list]) if keys_list else np.zeros((0, W), np.uint64)
    amps = np.concatenate([np.asarray(a, np.complex128).reshape(-1) for a in amps_list]) if amps_list else np.zeros(0, complex)
    v = Vec(W, keys=keys, amps=amps, offsets=off)
    v._off = off
    return v


def vec_apply(ops, src, out_src, out_model, weight_min, normalize=False):
    """Output vector o = ops model out_model[o] applied to src vector out_src[o]."""
    out_src = np.ascontiguousarray(out_src, np.int32)
    out_model = np.ascontiguousarray(out_model, np.int32)
    h = C.c_void_p()
    check(lib.imped_vec_apply(ops.h, src.h, len(out_src), ptr(_nz(out_src)), ptr(_nz(out_model)),
                              float(weight_min), int(bool(normalize)), C.byref(h)))
    return Vec(ops.W, handle=h).set_count(len(out_src))


class Batch:
    """Chains of a batch: bases, CSR Hamiltonians, Lanczos and eigen results."""

    def __init__(self, ops, chain_model):
        self.ops = ops
        self.W = ops.W
        self.chain_model = np.ascontiguousarray(chain_model, np.int32)
        self.C = len(self.chain_model)
        h = C.c_void_p()
        check(lib.imped_batch_create(ops.h, self.C, ptr(_nz(self.chain_model)), C.byref(h)))
        self.h = h
        self.chain_n = None

    def __del__(self, _release=_release):
        _release("imped_batch_destroy", self)

    def set_seeds(self, keys_list, amps_list=None):
        off = np.zeros(self.C + 1, np.int64)
        for c, k in enumerate(keys_list):
            off[c + 1] = off[c] + len(k)
        keys = (np.concatenate([np.asarray(k, np.uint64).reshape(-1, self.W) for k in keys_list])
                if self.C else np.zeros((0, self.W), np.uint64))
        amps = None
        if amps_list is not None:
            amps = _c64(np.concatenate([np.asarray(a, np.complex128).reshape(-1) for a in amps_list]))
        check(lib.imped_batch_set_seeds(self.h, ptr(off), ptr(_nz(keys.reshape(-1))), ptr(amps)))

    def seed_from_vec(self, vec):
        check(lib.imped_batch_seed_from_vec(self.h, vec.h))

    def build(self, fixed_basis, weight_min=1e-12, max_layers=0):
        check(lib.imped_batch_build(self.h, int(bool(fixed_basis)), float(weight_min), int(max_layers)))
        self.chain_n = np.zeros(max(self.C, 1), np.int64)
        self.chain_nnz = np.zeros(max(self.C, 1), np.int64)
        tot = np.zeros(2, np.int64)
        check(lib.imped_batch_sizes(self.h, ptr(self.chain_n), ptr(self.chain_nnz), ptr(tot)))
        self.chain_n = self.chain_n[: self.C]
        self.chain_nnz = self.chain_nnz[: self.C]
        self.rows, self.nnz = int(tot[0]), int(tot[1])
        self.row_base = np.concatenate([[0], np.cumsum(self.chain_n)]).astype(np.int64)
        return self

    def set_csr(self, chain_n, rowptr, col, val):
        """Use a given CSR (rows of H, chain-local columns)."""
        chain_n = np.ascontiguousarray(chain_n, np.int64)
        rowptr = np.ascontiguousarray(rowptr, np.int64)
        col = np.ascontiguousarray(col, np.int32)
        check(lib.imped_batch_set_csr(self.h, ptr(_nz(chain_n)), ptr(rowptr), ptr(_nz(col)), ptr(_nz(_c64(val)))))
        self.chain_n = chain_n[: self.C]
        self.rows = int(chain_n[: self.C].sum())
        self.nnz = int(rowptr[-1])
        self.row_base = np.concatenate([[0], np.cumsum(self.chain_n)]).astype(np.int64)
        self.chain_nnz = np.array([rowptr[self.row_base[c + 1]] - rowptr[self.row_base[c]] for c in range(self.C)],
                                  np.int64)
        return self

    def set_affine(self, affine, chain_struct, chain_model, theta, weight_min=1e-12):
        """CSR of every chain from ``affine`` (an :class:`Affine`): chain c is
        structure chain_struct[c] with parameter row chain_model[c] of theta."""
        chain_struct = np.ascontiguousarray(chain_struct, np.int32)
        chain_model = np.ascontiguousarray(chain_model, np.int32)
        th = np.ascontiguousarray(np.asarray(theta, np.float64).reshape(-1, max(affine.P, 1))[:, : affine.P])
        check(lib.imped_batch_set_affine(self.h, affine.h, ptr(_nz(chain_struct)), ptr(_nz(chain_model)), len(th),
                                         ptr(_nz(th.reshape(-1))), float(weight_min)))
        self.chain_n = affine.n[chain_struct]
        self.chain_nnz = affine.nnz[chain_struct]
        self.rows, self.nnz = int(self.chain_n.sum()), int(self.chain_nnz.sum())
        self.row_base = np.concatenate([[0], np.cumsum(self.chain_n)]).astype(np.int64)
        return self

    def seed_affine(self, ground, trans, src_chain, src_vec, chain_struct, psi_weight_min=1e-7, weight_min=1e-7):
        """Start vectors T psi on the device (``imped_batch_seed_affine``):
        chain c from eigenvector src_vec[c] of chain src_chain[c] of the
        ``ground`` batch and structure chain_struct[c] of ``trans``.  Returns
        norm2 [C], |T psi|^2 before normalization."""
        norm2 = np.zeros(max(self.C, 1), np.float64)
        check(lib.imped_batch_seed_affine(self.h, ground.h, trans.h, ptr(_nz(np.ascontiguousarray(src_chain, np.int32))),
                                          ptr(_nz(np.ascontiguousarray(src_vec, np.int32))),
                                          ptr(_nz(np.ascontiguousarray(chain_struct, np.int32))),
                                          float(psi_weight_min), float(weight_min), ptr(norm2)))
        return norm2[: self.C]

    def basis_words(self):
        keys = np.zeros((max(self.rows, 1), self.W), np.uint64)
        check(lib.imped_batch_get_basis(self.h, ptr(keys)))
        return keys[: self.rows]

    def csr(self):
        rowptr = np.zeros(self.rows + 1, np.int64)
        col = np.zeros(max(self.nnz, 1), np.int32)
        val = np.zeros(max(self.nnz, 1), np.complex128)
        check(lib.imped_batch_get_csr(self.h, ptr(rowptr), ptr(col), ptr(val.view(np.float64))))
        return rowptr, col[: self.nnz], val[: self.nnz]

    def lanczos(self, K, reorth=0, v0=None, want_coeffs=True):
        """Lanczos chains of every chain (``imped_batch_lanczos``); alpha and
        beta stay on the device (for ``cf``) unless ``want_coeffs``."""
        alpha = np.zeros((max(self.C, 1), K), np.float64) if want_coeffs else None
        beta = np.zeros((max(self.C, 1), K), np.float64) if want_coeffs else None
        kout = np.zeros(max(self.C, 1), np.int32)
        status = np.zeros(max(self.C, 1), np.int32)
        v0a = None if v0 is None else _c64(v0)
        check(lib.imped_batch_lanczos(self.h, ptr(v0a), int(K), int(reorth), ptr(alpha), ptr(beta), ptr(kout),
                                      ptr(status)))
        self.K = K
        if not want_coeffs:
            return None, None, kout[: self.C], status[: self.C]
        return alpha[: self.C], beta[: self.C], kout[: self.C], status[: self.C]

    def eigen(self, nev, tol=1e-12, max_steps=600, seed=12345, want_vectors=True, allow_noconv=False):
        evals = np.zeros((max(self.C, 1), nev), np.float64)
        evecs = np.zeros(max(self.rows * nev, 1), np.complex128) if want_vectors else None
        steps = np.zeros(max(self.C, 1), np.int32)
        rc = check(
            lib.imped_batch_eigen(self.h, int(nev), float(tol), int(max_steps), int(seed), ptr(evals),
                                  None if evecs is None else ptr(evecs.view(np.float64)), ptr(steps)),
            allow=(IMPED_ERR_NOCONV,) if allow_noconv else (),
        )
        self.nev = nev
        return evals[: self.C], evecs, steps[: self.C], rc == 0

    def solve(self, shifts, tol=1e-12, max_steps=1000, allow_noconv=False, want_x=False):
        """(z I - H_c) x = seed vector of chain c for every shift z in
        ``shifts[c]`` ([C, S] complex); the solutions feed ``vec_from_eigen``.
        With ``want_x`` the solutions are also returned (chain c's block at
        row_base[c]*S, shift s at offset s*n_c)."""
        z = np.asarray(shifts, np.complex128).reshape(self.C, -1)
        S = z.shape[1]
        steps = np.zeros(max(self.C, 1), np.int32)
        resid = np.zeros((max(self.C, 1), S), np.float64)
        x = np.zeros(max(self.rows * S, 1), np.complex128) if want_x else None
        rc = check(lib.imped_batch_solve(self.h, int(S), ptr(_c64(z)), float(tol), int(max_steps), ptr(steps),
                                         ptr(resid), None if x is None else ptr(x.view(np.float64))),
                   allow=(IMPED_ERR_NOCONV,) if allow_noconv else ())
        self.nev = S
        out = (steps[: self.C], resid[: self.C], rc == 0)
        return out + (x,) if want_x else out

    def eigvec(self, evecs, c, k):
        """Eigenvector k of chain c from the flat evecs array."""
        n = int(self.chain_n[c])
        base = int(self.row_base[c]) * self.nev
        return evecs[base + k * n: base + (k + 1) * n]

    def vec_from_eigen(self, chains, which, weight_min):
        chains = np.ascontiguousarray(chains, np.int32)
        which = np.ascontiguousarray(which, np.int32)
        h = C.c_void_p()
        check(lib.imped_vec_from_eigen(self.h, len(chains), ptr(_nz(chains)), ptr(_nz(which)), float(weight_min),
                                       C.byref(h)))
        return Vec(self.W, handle=h).set_count(len(chains))

    def cf(self, energy, norm2, sign, zflag, kT, spec_id, n_spec, omega, delta, want_g=False):
        omega = np.ascontiguousarray(omega, np.float64)
        g = np.zeros((max(self.C, 1), len(omega)), np.complex128) if want_g else None
        spec = np.zeros((max(n_spec, 1), len(omega)), np.float64)
        check(lib.imped_batch_cf(
            self.h, ptr(_nz(np.asarray(energy, np.float64))), ptr(_nz(np.asarray(norm2, np.float64))),
            ptr(_nz(np.asarray(sign, np.int32))), ptr(_nz(np.asarray(zflag, np.int32))),
            ptr(np.ascontiguousarray(np.broadcast_to(np.asarray(kT, np.float64), (max(n_spec, 1),)))),
            ptr(_nz(np.asarray(spec_id, np.int32))), int(n_spec), ptr(omega), len(omega), float(delta),
            None if g is None else ptr(g.view(np.float64)), ptr(spec)))
        return (g[: self.C] if want_g else None), spec[:n_spec]


class Affine:
    """Fixed CSR structures with parameter-affine values (``imped_affine_*``).

    ``structs`` is a list of (rowptr, col, base, dirs) per structure: rowptr
    [n+1] from 0, col [nnz], base [nnz] complex, dirs [P, nnz] complex (P the
    same for all structures; P = 0 for fixed matrices)."""

    def __init__(self, structs, n_params=0):
        self.P = int(n_params)
        self.n = np.array([len(r) - 1 for r, _, _, _ in structs], np.int64)
        self.nnz = np.array([int(r[-1]) for r, _, _, _ in structs], np.int64)
        rowptr = np.concatenate([np.asarray(r, np.int64) for r, _, _, _ in structs])
        col = np.concatenate([np.asarray(c, np.int32) for _, c, _, _ in structs])
        base = _c64(np.concatenate([np.asarray(b, np.complex128) for _, _, b, _ in structs]))
        dirs = None
        if self.P:
            dirs = _c64(np.concatenate([np.asarray(d, np.complex128).reshape(self.P, -1) for _, _, _, d in structs],
                                       axis=1))
        h = C.c_void_p()
        check(lib.imped_affine_create(len(structs), ptr(self.n), ptr(rowptr), ptr(_nz(col)), ptr(_nz(base)), self.P,
                                      ptr(None if dirs is None else _nz(dirs)), C.byref(h)))
        self.h = h

    def __del__(self, _release=_release):
        _release("imped_affine_destroy", self)


def family_xas(ground, sector, trans, sector_index, theta, kT, omega, delta, nev=3, n_psi_max=3, energy_cut=10.0,
               krylov_size=200, reorth=1, eig_tol=1e-12, eig_max_steps=600, seed=12345, h_weight_min=1e-12,
               psi_weight_min=1e-7, weight_min=1e-7):
    """``imped_family_xas``: thermal, polarization-summed XAS of parameter
    sets from an affine family's sector structures (:class:`Affine` handles
    ``ground``, ``sector``, ``trans``; ``sector_index`` [n_comp, n_pol]) in
    one call.  Returns a dict with ``spectra`` [M, n_w], ``energies`` (list
    of kept eigenvalues per model), ``eig_steps`` [M * n_comp],
    ``chain_struct`` and ``kout`` of the excited chains, and ``K``."""
    sector_index = np.ascontiguousarray(sector_index, np.int32)
    n_comp, n_pol = sector_index.shape
    th = np.ascontiguousarray(np.asarray(theta, np.float64).reshape(len(theta), -1)[:, : ground.P])
    M = len(th)
    kT = np.ascontiguousarray(np.broadcast_to(np.asarray(kT, np.float64), (M,)))
    omega = np.ascontiguousarray(omega, np.float64)
    spectra = np.zeros((max(M, 1), len(omega)), np.float64)
    energies = np.zeros((max(M, 1), n_psi_max), np.float64)
    n_kept = np.zeros(max(M, 1), np.int32)
    eig_steps = np.zeros(max(M * n_comp, 1), np.int32)
    cap = max(M * n_pol * n_psi_max, 1)
    chain_struct = np.zeros(cap, np.int32)
    kout = np.zeros(cap, np.int32)
    info = np.zeros(4, np.int32)
    check(lib.imped_family_xas(ground.h, sector.h, trans.h, n_comp, n_pol, ptr(sector_index), M,
                               ptr(_nz(th.reshape(-1))), ptr(_nz(kT)), int(nev), int(n_psi_max), float(energy_cut),
                               int(krylov_size), int(reorth), float(eig_tol), int(eig_max_steps), int(seed),
                               float(h_weight_min), float(psi_weight_min), float(weight_min), ptr(_nz(omega)),
                               len(omega), float(delta), ptr(spectra), ptr(energies), ptr(n_kept), ptr(eig_steps),
                               ptr(chain_struct), ptr(kout), ptr(info)))
    X = int(info[0])
    return dict(spectra=spectra[:M], energies=[list(energies[m, : n_kept[m]]) for m in range(M)],
                eig_steps=eig_steps[: M * n_comp], chain_struct=chain_struct[:X], kout=kout[:X], K=int(info[1]))


BREAKDOWN = IMPED_ERR_BREAKDOWN


# ------------------------------------------------------------------ timing, tuning, enumeration


def timing_enable(on=True):
    check(lib.imped_timing_enable(int(bool(on))))


def timing_reset():
    check(lib.imped_timing_reset())


def timing_query(name):
    """(total milliseconds, launches) recorded for one kernel name."""
    ms = np.zeros(1, np.float64)
    k = np.zeros(1, np.int64)
    check(lib.imped_timing_query(name.encode(), ptr(ms), ptr(k)))
    return float(ms[0]), int(k[0])


def set_tuning(name, value):
    check(lib.imped_set_tuning(name.encode(), int(value)))


def enumerate_basis(n_orb, W, shells):
    """Restricted basis words [N, W] in get_basis order (host C++).

    ``shells``: list of 10-tuples per angular-momentum shell (see
    imped_enumerate_basis).
    """
    sh = np.ascontiguousarray(np.array(shells, np.int32).reshape(-1))
    n = np.zeros(1, np.int64)
    check(lib.imped_enumerate_basis(int(n_orb), int(W), len(shells), ptr(sh), 0, ptr(n), None))
    N = int(n[0])
    keys = np.zeros((max(N, 1), W), np.uint64)
    check(lib.imped_enumerate_basis(int(n_orb), int(W), len(shells), ptr(sh), N, ptr(n), ptr(keys)))
    return keys[:N]


def sort_u64(keys, begin_bit=0, end_bit=64, perm=None):
    """Stable device radix sort: permutation sorting ``keys`` on bits [begin, end)."""
    keys = np.ascontiguousarray(keys, np.uint64)
    n = len(keys)
    perm = np.arange(n, dtype=np.int32) if perm is None else np.ascontiguousarray(perm, np.int32).copy()
    check(lib.imped_sort_u64(ptr(_nz(keys)), n, int(begin_bit), int(end_bit), ptr(_nz(perm))))
    return perm
