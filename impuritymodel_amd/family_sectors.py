"""Sector cache of an affine model family: exact XAS without rebuilding sectors.

The batched XAS of a model family (``batched.xas_batch_ops``) rebuilds every
sector for every parameter set: the fixed-basis Hamiltonian of the ground
basis (get_hamiltonian_matrix, finite.py:1514-1610), the transition states
T_t|psi> (applyOp, finite.py:1301-1511) and the BFS closure with its CSR
Hamiltonian (expand_basis_and_hamiltonian, finite.py:1806-1894).  For a
family affine in its parameters, H(theta) = H_0 + sum_p theta_p H_p, none of
these sectors depends on theta:

* an entry of H(theta) on a fixed basis is an affine function of theta, so
  the union of the patterns of H_0, H_1, ... is the pattern of every member
  (an entry that vanishes for all theta has zero base and zero directions);
* the ground basis therefore splits into the same connected components for
  every member, and an eigenvector of H lies in one component (the lowest
  eigenpairs of the whole basis are the lowest of the components' merged);
* the closure of T_t|psi>, psi supported in component c, is the connected
  component of the excited-sector graph that T_t maps c into, whenever that
  set is connected (checked when the cache is built).

The cache holds, per ground component, the component's Hamiltonian as a CSR
structure with affine values, and per (component, polarization) the excited
sector's Hamiltonian and the transition matrix from the component into it.
A batch of parameter sets then costs: one fill launch for all ground chains
(``imped_batch_set_affine``), one eigensolver launch, one launch forming all
start vectors T psi (``imped_batch_seed_affine``, with the reference's
eigenvector truncation, pruning and normalization), one fill launch for all
excited chains, one Lanczos launch and one continued-fraction launch.

Sector bases keep the order of the BFS closure at a generic member; the
general path orders each member's closure from its own seeds.  Orders differ
only by a permutation of the basis, which changes the spectra by rounding
(with full reorthogonalization, about 1e-12 of the peak).
"""

import time

import numpy as np
import scipy.sparse
import scipy.sparse.csgraph

from .batched import POLARIZATIONS, XASResult, _dipole_tables, ground_bases
from .engine import Affine, Batch, Ops, family_xas, vec_apply, vec_from_host
from .models import K_B
from .operators import n_spin_orbitals

DENSE_MAX = 4096  # sector size up to which the cache is built with dense host matrices


class SectorCacheUnavailable(RuntimeError):
    """The family's sectors do not have the structure the cache needs."""


def _dense_pieces(batch, n_pieces, n):
    """Dense stored-layout matrices S_k (S_k[r, col] = val) of chains 0..n_pieces-1."""
    rowptr, col, val = batch.csr()
    out = []
    for k in range(n_pieces):
        r0, r1 = int(batch.row_base[k]), int(batch.row_base[k + 1])
        e0, e1 = int(rowptr[r0]), int(rowptr[r1])
        m = scipy.sparse.csr_matrix((val[e0:e1], col[e0:e1], rowptr[r0:r1 + 1] - e0), shape=(n, n))
        out.append(m.toarray())
    return out


def _affine_struct(pieces):
    """(rowptr, col, base, dirs) of the union pattern of pieces [H_0, H_1, ...]."""
    mask = np.zeros(pieces[0].shape, bool)
    for p in pieces:
        mask |= p != 0
    rows, cols = np.nonzero(mask)  # row-major, columns ascending
    n = pieces[0].shape[0]
    rowptr = np.zeros(n + 1, np.int64)
    np.add.at(rowptr, rows + 1, 1)
    rowptr = np.cumsum(rowptr)
    base = pieces[0][rows, cols]
    dirs = np.stack([p[rows, cols] for p in pieces[1:]]) if len(pieces) > 1 else np.zeros((0, len(rows)), complex)
    return rowptr, cols.astype(np.int32), base, dirs, mask


def _sector_pieces(ops_p, n_pieces, basis):
    """Dense pieces of the operator models 0..n_pieces-1 on a fixed basis (no pruning)."""
    n = len(basis)
    if n > DENSE_MAX:
        raise SectorCacheUnavailable("sector of %d states exceeds the dense cache build limit" % n)
    b = Batch(ops_p, np.arange(n_pieces, dtype=np.int32))
    b.set_seeds([basis] * n_pieces)
    b.build(fixed_basis=True, weight_min=0.0)
    if np.any(b.chain_n != n):
        raise SectorCacheUnavailable("fixed-basis build changed the sector size")
    return _dense_pieces(b, n_pieces, n)


class FamilySectors:
    """Sector cache of an :class:`impuritymodel_amd.surrogate.XASFamily`.

    Built once per family and word count with the general GPU build at a
    generic member ``theta0`` (drawn from the family's parameter ranges with
    ``seed``); raises :class:`SectorCacheUnavailable` when an excited sector
    is not connected or a sector is too large for the dense build."""

    def __init__(self, family, W=1, seed=0, nPsiMax=3):
        from .surrogate import PARAM_RANGES

        t0 = time.perf_counter()
        self.W = W
        n_orb = n_spin_orbitals(family.nBaths)
        tab = family.table(W)
        P = family.cmat.shape[0]
        self.P = P
        rng = np.random.default_rng(seed)
        lo, hi = PARAM_RANGES[:, 0], PARAM_RANGES[:, 1]
        theta0 = lo + (hi - lo) * rng.random(len(lo))
        coefs = [family.c0] + [family.cmat[p] for p in range(P)] + [family.coef(theta0[None])[0]]
        ops_p = Ops([tab.with_coef(c) for c in coefs], None, n_orb, W)
        generic = P + 1  # model index of theta0 in ops_p
        basis = ground_bases(family.models(theta0[None]), W)[0]
        self.ground_basis = basis
        n0 = len(basis)

        # ground components from the union pattern
        gp = _sector_pieces(ops_p, P + 1, basis)
        _, _, _, _, mask = _affine_struct(gp)
        ncomp, label = scipy.sparse.csgraph.connected_components(scipy.sparse.csr_matrix(mask | mask.T),
                                                                 directed=False)
        first = np.array([np.nonzero(label == c)[0][0] for c in range(ncomp)])
        comps = [np.nonzero(label == c)[0] for c in np.argsort(first)]
        self.comp_rows = comps
        self.comp_n = np.array([len(c) for c in comps], np.int64)
        if self.comp_n.min() < nPsiMax:
            raise SectorCacheUnavailable("a ground component has fewer than nPsiMax states")
        ground_structs = []
        for idx in comps:
            rp, col, base, dirs, _ = _affine_struct([p[np.ix_(idx, idx)] for p in gp])
            ground_structs.append((rp, col, base, dirs))

        # transition images T_t e_j of every ground state, then excited sectors
        m0 = family.models(theta0[None])[0]
        ops_t = Ops(_dipole_tables([m0], W), None, n_orb, W)
        nt = len(POLARIZATIONS)
        units = vec_from_host(W, [basis[j:j + 1] for j in range(n0)], [np.ones(1)] * n0)
        out_src = np.tile(np.arange(n0, dtype=np.int32), nt)
        out_model = np.repeat(np.arange(nt, dtype=np.int32), n0)
        img = vec_apply(ops_t, units, out_src, out_model, 0.0, normalize=False)
        keys, amps, _ = img.get()
        off = img.offsets

        sector_structs, trans_structs = [], []
        self.sector_index = np.zeros((ncomp, nt), np.int32)
        self.sector_basis = []
        for c, idx in enumerate(comps):
            for t in range(nt):
                cols_j, rows_k, vals = [], [], []
                seen = {}
                for jl, j in enumerate(idx):
                    o = t * n0 + j
                    for e in range(int(off[o]), int(off[o + 1])):
                        if amps[e] == 0:
                            continue
                        k = keys[e].tobytes()
                        if k not in seen:
                            seen[k] = keys[e]
                        cols_j.append(jl)
                        rows_k.append(k)
                        vals.append(amps[e])
                if not seen:
                    raise SectorCacheUnavailable("a polarization maps a ground component to nothing")
                seeds = np.array(list(seen.values()), np.uint64).reshape(-1, W)
                cb = Batch(ops_p, np.array([generic], np.int32))
                cb.set_seeds([seeds])
                cb.build(fixed_basis=False, weight_min=1e-12)
                sbasis = cb.basis_words()
                ns = len(sbasis)
                pos = {sbasis[i].tobytes(): i for i in range(ns)}
                sp = _sector_pieces(ops_p, P + 1, sbasis)
                rp, col, base, dirs, smask = _affine_struct(sp)
                nc, _ = scipy.sparse.csgraph.connected_components(scipy.sparse.csr_matrix(smask | smask.T),
                                                                  directed=False)
                if nc != 1:
                    raise SectorCacheUnavailable("excited sector (component %d, polarization %d) is not connected"
                                                 % (c, t))
                tm = np.zeros((ns, len(idx)), np.complex128)
                for r, jl, v in zip(rows_k, cols_j, vals):
                    tm[pos[r], jl] += v
                tr, tc = np.nonzero(tm)
                trp = np.cumsum(np.concatenate([[0], np.bincount(tr, minlength=ns)])).astype(np.int64)
                self.sector_index[c, t] = len(sector_structs)
                sector_structs.append((rp, col, base, dirs))
                trans_structs.append((trp, tc.astype(np.int32), tm[tr, tc], None))
                self.sector_basis.append(sbasis)
        self.ncomp, self.nt = ncomp, nt
        self.ground = Affine(ground_structs, P)
        self.sector = Affine(sector_structs, P)
        self.trans = Affine(trans_structs, 0)
        self.sector_n = self.sector.n.copy()
        self._dummy = Ops([{}], [None], 1)
        self.build_s = time.perf_counter() - t0

    def xas(self, thetas, omega, delta=0.2, krylovSize=200, nPsiMax=3, T=300.0, energy_cut=10.0, reorth=1,
            eig_tol=1e-12, eig_max_steps=600, seed=12345):
        """Thermal, polarization-summed XAS of parameter sets [M, >= P]
        (the spectra of ``batched.xas_batch_ops`` on the family's operators),
        all steps in one library call (``imped_family_xas``)."""
        t0 = time.perf_counter()
        th = np.ascontiguousarray(np.asarray(thetas, np.float64).reshape(len(thetas), -1)[:, : self.P])
        M = len(th)
        omega = np.asarray(omega, np.float64)
        T_m = np.broadcast_to(np.asarray(T, np.float64), (M,))
        r = family_xas(self.ground, self.sector, self.trans, self.sector_index, th, K_B * T_m, omega, delta,
                       nev=nPsiMax, n_psi_max=nPsiMax, energy_cut=energy_cut, krylov_size=krylovSize, reorth=reorth,
                       eig_tol=eig_tol, eig_max_steps=eig_max_steps, seed=seed)
        tim = {"total": time.perf_counter() - t0, "expand_build": 0.0}
        cs = r["chain_struct"]
        gdims = np.tile(self.comp_n, M)
        gnnz = np.tile(self.ground.nnz, M)
        return XASResult(omega, r["spectra"], r["energies"], tim, self.sector.n[cs], self.sector.nnz[cs],
                         r["eig_steps"], r["kout"], gdims, gnnz)

    def xas_steps(self, thetas, omega, delta=0.2, krylovSize=200, nPsiMax=3, T=300.0, energy_cut=10.0, reorth=1,
                  eig_tol=1e-12, eig_max_steps=600, seed=12345):
        """:meth:`xas` as a sequence of the batch entry points
        (``imped_batch_set_affine``, ``imped_batch_eigen``,
        ``imped_batch_seed_affine``, ``imped_batch_lanczos``,
        ``imped_batch_cf``) with the state selection in Python."""
        t0 = time.perf_counter()
        tim = {}
        th = np.ascontiguousarray(np.asarray(thetas, np.float64).reshape(len(thetas), -1)[:, : self.P])
        M = len(th)
        omega = np.asarray(omega, np.float64)
        T_m = np.broadcast_to(np.asarray(T, np.float64), (M,))
        nc = self.ncomp
        nev = nPsiMax
        gb = Batch(self._dummy, np.zeros(M * nc, np.int32))
        gb.set_affine(self.ground, np.tile(np.arange(nc, dtype=np.int32), M),
                      np.repeat(np.arange(M, dtype=np.int32), nc), th, weight_min=1e-12)
        tim["ground_build"] = time.perf_counter() - t0
        t1 = time.perf_counter()
        evals, _, steps, _ = gb.eigen(nev, tol=eig_tol, max_steps=eig_max_steps, seed=seed, want_vectors=False)
        tim["ground_eigen"] = time.perf_counter() - t1

        t1 = time.perf_counter()
        ev = np.asarray(evals, np.float64).reshape(M, nc * nev)
        order = np.argsort(ev, axis=1, kind="stable")[:, :nPsiMax]
        es = np.take_along_axis(ev, order, axis=1)
        keep = (es - es[:, :1]) < (energy_cut * K_B * T_m)[:, None]
        # chains ordered by (model, polarization, kept state)
        mm, kk = np.nonzero(keep)
        tt = np.repeat(np.arange(self.nt), len(mm))
        m3, k3 = np.tile(mm, self.nt), np.tile(kk, self.nt)
        o = np.lexsort((k3, tt, m3))
        model, t_o, k = m3[o], tt[o], k3[o]
        rank = order[model, k]
        comp, vec = rank // nev, rank % nev
        energy = es[model, k]
        src_chain = (model * nc + comp).astype(np.int32)
        struct = self.sector_index[comp, t_o].astype(np.int32)
        xb = Batch(self._dummy, np.zeros(len(model), np.int32))
        xb.set_affine(self.sector, struct, model.astype(np.int32), th, weight_min=1e-12)
        norm2 = xb.seed_affine(gb, self.trans, src_chain, vec.astype(np.int32), struct, 1e-7, 1e-7)
        tim["transition"] = time.perf_counter() - t1
        tim["expand_build"] = 0.0
        t1 = time.perf_counter()
        K = int(min(krylovSize, max(int(xb.chain_n.max()) if len(model) else 1, 1)))
        _, _, kout, status = xb.lanczos(K, reorth=reorth, want_coeffs=False)
        if np.any(status != 0) or np.any(norm2 == 0):
            raise ValueError("Warning: beta==0, implementation absent!")
        tim["lanczos"] = time.perf_counter() - t1
        t1 = time.perf_counter()
        _, spec = xb.cf(energy, norm2, np.ones(len(model), np.int32), (t_o == 0).astype(np.int32), K_B * T_m,
                        model.astype(np.int32), M, omega, delta)
        tim["cf"] = time.perf_counter() - t1
        tim["total"] = time.perf_counter() - t0
        energies = [list(es[m][keep[m]]) for m in range(M)]
        return XASResult(omega, spec, energies, tim, xb.chain_n.copy(), xb.chain_nnz.copy(), steps, kout,
                         gb.chain_n.copy(), gb.chain_nnz.copy())
