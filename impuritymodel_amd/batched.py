"""Batched XAS: many impurity models' spectra in the same device launches.

For M models this runs the reference's XAS chain (get_spectra.py:198-202
ground state, spectra.py:222-238 XAS) with every step batched over models:

1. fixed-basis H of all models' ground bases in one build
   (get_hamiltonian_matrix, finite.py:1514-1610),
2. lowest ``nPsiMax`` eigenpairs of all models in one eigensolver launch
   (eigensystem, finite.py:58-118), energy cut (get_spectra.py:302-303) and
   eigenvector truncation (finite.py:114-117),
3. all T_t|psi_i> (3 dipole polarizations x kept eigenstates x models) in
   one operator application, normalized (spectra.py:786-790),
4. the BFS closures and CSR Hamiltonians of all chains in one build
   (expand_basis_and_hamiltonian, finite.py:1806-1894),
5. all Lanczos chains in one launch (finite.py:1974-1992),
6. continued fractions, Boltzmann weights and the polarization sum of all
   models in one launch (spectra.py:677-680, average.py:11-45).

The Lanczos chains are fully reorthogonalized by default (``reorth=1``).  XAS
sectors of these models have at most a few hundred states, so a chain runs
K = min(krylovSize, n) = n steps; the reference's plain recurrence taken
that far has lost orthogonality, and its spectrum then depends on rounding
(two executions that only sum in different orders differ by about 1e-5 of
the peak).  With reorthogonalization the n-step continued fraction is the
exact resolvent, reproducible to 1e-10.  ``reorth=0`` keeps the reference's
plain recurrence.

``plain_xas`` computes the same spectrum one model and one chain at a time
with the same kernels (the unbatched GPU baseline).
"""

import time
from dataclasses import dataclass, field

import numpy as np

from . import opcompile
from .engine import Batch, Ops, vec_apply
from .finite import get_basis_words
from .models import K_B
from .operators import getDipoleOperators

POLARIZATIONS = ([1, 0, 0], [0, 1, 0], [0, 0, 1])


@dataclass
class XASResult:
    omega: np.ndarray
    spectra: np.ndarray          # [M, n_w] thermally averaged, polarization-summed -Im G
    energies: list               # per model: kept eigenvalues
    timings: dict = field(default_factory=dict)
    chain_dims: np.ndarray = None
    chain_nnz: np.ndarray = None
    eig_steps: np.ndarray = None
    kout: np.ndarray = None
    ground_dims: np.ndarray = None   # per eigensolver chain
    ground_nnz: np.ndarray = None


def _dipole_tables(models, W):
    """Dipole term tables (x, y, z) of every model with its restrictions."""
    out = []
    for m in models:
        cache = m.__dict__.setdefault("_dip_tables", {})
        if W not in cache:
            dip = getDipoleOperators(m.nBaths, POLARIZATIONS)
            cache[W] = [opcompile.TermTable(d, m.restrictions, W) for d in dip]
        out.extend(cache[W])
    return out


def ground_bases(models, W):
    """get_basis of every model as [N, W] words (native enumeration)."""
    return [get_basis_words(m.nBaths, m.nValBaths, m.dnValBaths, m.dnConBaths, m.dnTols, m.n0imps, W)
            for m in models]


def xas_batch(models, omega=None, delta=0.2, krylovSize=200, nPsiMax=3, T=300.0, energy_cut=10.0, reorth=1,
              eig_tol=1e-12, eig_max_steps=600, seed=12345, bases=None):
    """Thermal, polarization-summed XAS of every model (all steps batched)."""
    t0 = time.perf_counter()
    tim = {}
    if omega is None:
        omega = np.linspace(-25.0, 25.0, 2000)
    omega = np.asarray(omega, dtype=np.float64)
    n_orb = max(m.n_spin_orbitals for m in models)
    W = opcompile.n_words(n_orb)
    ops_h = Ops([m.term_table(W) for m in models], None, n_orb, W)
    ops_t = Ops(_dipole_tables(models, W), None, n_orb, W)
    if bases is None:
        bases = ground_bases(models, W)
    tim["host_setup"] = time.perf_counter() - t0
    return xas_batch_ops(ops_h, ops_t, bases, omega, delta, krylovSize, nPsiMax, T, energy_cut, reorth, eig_tol,
                         eig_max_steps, seed, tim, t0)


def xas_batch_ops(ops_h, ops_t, bases, omega, delta=0.2, krylovSize=200, nPsiMax=3, T=300.0, energy_cut=10.0,
                  reorth=1, eig_tol=1e-12, eig_max_steps=600, seed=12345, tim=None, t0=None):
    """xas_batch on compiled operators: model m < len(bases) of ``ops_h`` is the
    Hamiltonian, models 3m..3m+2 of ``ops_t`` its x, y, z dipole operators,
    bases[m] its ground basis (the operator sets may hold more models)."""
    t0 = time.perf_counter() if t0 is None else t0
    tim = {} if tim is None else tim
    omega = np.asarray(omega, dtype=np.float64)
    M = len(bases)
    t1 = time.perf_counter()
    gb = Batch(ops_h, np.arange(M, dtype=np.int32))
    gb.set_seeds(bases)
    gb.build(fixed_basis=True, weight_min=1e-12)
    tim["ground_build"] = time.perf_counter() - t1
    t1 = time.perf_counter()
    nev = nPsiMax
    evals, _, steps, ok = gb.eigen(nev, tol=eig_tol, max_steps=eig_max_steps, seed=seed, want_vectors=False)
    tim["ground_eigen"] = time.perf_counter() - t1

    t1 = time.perf_counter()
    T_m = np.broadcast_to(np.asarray(T, np.float64), (M,))
    ev = np.asarray(evals, np.float64).reshape(M, nev)
    keep = (ev - ev[:, :1]) < (energy_cut * K_B * T_m)[:, None]
    psi_chain, psi_which = (x.astype(np.int32) for x in np.nonzero(keep))  # model-major, state ascending
    psis = gb.vec_from_eigen(psi_chain, psi_which, 1e-7)
    nt = len(POLARIZATIONS)
    # output vectors ordered by (model, polarization, kept state)
    n_psi = len(psi_chain)
    kk = np.tile(np.arange(n_psi), nt)
    tt = np.repeat(np.arange(nt), n_psi)
    order = np.lexsort((kk, tt, psi_chain[kk]))
    out_src = kk[order].astype(np.int32)
    t_o = tt[order]
    chain_model = psi_chain[out_src]
    out_model = (nt * chain_model + t_o).astype(np.int32)
    energy = ev[chain_model, psi_which[out_src]]
    spec_id = chain_model
    zflag = (t_o == 0).astype(np.int32)
    kept = [np.nonzero(keep[m])[0] for m in range(M)]
    seeds = vec_apply(ops_t, psis, out_src, out_model, 1e-7, normalize=True)
    norm2 = seeds.norm2()
    tim["transition"] = time.perf_counter() - t1

    t1 = time.perf_counter()
    xb = Batch(ops_h, np.array(chain_model, np.int32))
    xb.seed_from_vec(seeds)
    xb.build(fixed_basis=False, weight_min=1e-12)
    tim["expand_build"] = time.perf_counter() - t1
    t1 = time.perf_counter()
    K = int(min(krylovSize, max(int(xb.chain_n.max()), 1)))
    _, _, kout, status = xb.lanczos(K, reorth=reorth, want_coeffs=False)
    if np.any(status != 0):
        raise ValueError("Warning: beta==0, implementation absent!")
    tim["lanczos"] = time.perf_counter() - t1
    t1 = time.perf_counter()
    _, spec = xb.cf(np.array(energy), norm2, np.ones(len(energy), np.int32), np.array(zflag, np.int32), K_B * T_m,
                    np.array(spec_id, np.int32), M, omega, delta)
    tim["cf"] = time.perf_counter() - t1
    tim["total"] = time.perf_counter() - t0
    return XASResult(omega, spec, [[evals[m][i] for i in kept[m]] for m in range(M)], tim, xb.chain_n.copy(),
                     xb.chain_nnz.copy(), steps, kout, gb.chain_n.copy(), gb.chain_nnz.copy())


def state_spectra(ops_h, ops_t, psis, energies, omega, delta, krylovSize=150, T=300.0, sign=1, reorth=0,
                  weight_min=1e-7):
    """Thermal, operator-summed spectrum of one model from device eigenstates.

    The ``getSpectra`` + ``thermal_average`` + operator sum of
    simulate_spectra (spectra.py:124-238) for transition operators ``ops_t``
    (one Ops model per operator) on eigenstates ``psis`` (a Vec, one vector per
    state, energies given): all (operator, state) chains in one batch.
    ``sign = -1`` is the photoemission convention (-w, -delta, negated).
    Returns ``(spectrum [n_w], info)`` with stage times and chain sizes.
    """
    info = {}
    nt, ne = ops_t.M, len(energies)
    omega = np.asarray(omega, np.float64)
    t0 = time.perf_counter()
    out_src = np.array([i for t in range(nt) for i in range(ne)], np.int32)
    out_model = np.array([t for t in range(nt) for i in range(ne)], np.int32)
    seeds = vec_apply(ops_t, psis, out_src, out_model, weight_min, normalize=True)
    norm2 = seeds.norm2()
    info["transition_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    xb = Batch(ops_h, np.zeros(len(out_src), np.int32))
    xb.seed_from_vec(seeds)
    xb.build(fixed_basis=False, weight_min=1e-12)
    info["build_s"] = time.perf_counter() - t0
    info["chain_n"] = xb.chain_n.copy()
    info["chain_nnz"] = xb.chain_nnz.copy()
    t0 = time.perf_counter()
    K = int(min(krylovSize, max(int(xb.chain_n.max()), 1)))
    _, _, kout, status = xb.lanczos(K, reorth=reorth, want_coeffs=False)
    if np.any(status != 0):
        raise ValueError("Warning: beta==0, implementation absent!")
    info["lanczos_s"] = time.perf_counter() - t0
    info["kout"] = kout.copy()
    t0 = time.perf_counter()
    energy = np.asarray(energies, np.float64)[out_src]
    zflag = (out_model == 0).astype(np.int32)
    _, spec = xb.cf(energy, norm2, np.full(len(out_src), sign, np.int32), zflag, K_B * T,
                    np.zeros(len(out_src), np.int32), 1, omega, delta)
    info["cf_s"] = time.perf_counter() - t0
    return spec[0], info


def plain_xas(model, omega=None, delta=0.2, krylovSize=200, nPsiMax=3, T=300.0, energy_cut=10.0, reorth=1,
              eig_tol=1e-12, eig_max_steps=600, seed=12345, basis=None):
    """The same spectrum with one launch sequence per chain (unbatched GPU path)."""
    if omega is None:
        omega = np.linspace(-25.0, 25.0, 2000)
    omega = np.asarray(omega, dtype=np.float64)
    n = model.n_spin_orbitals
    W = opcompile.n_words(n)
    ops_h = Ops([model.hOp], [model.restrictions], n, W)
    dip = getDipoleOperators(model.nBaths, POLARIZATIONS)
    if basis is None:
        basis = ground_bases([model], W)[0]
    gb = Batch(ops_h, [0])
    gb.set_seeds([basis])
    gb.build(fixed_basis=True)
    evals, _, _, _ = gb.eigen(nPsiMax, tol=eig_tol, max_steps=eig_max_steps, seed=seed, want_vectors=False)
    e = evals[0]
    kept = [i for i in range(nPsiMax) if e[i] - e[0] < energy_cut * K_B * T]
    acc = np.zeros(len(omega))
    Z = 0.0
    e0 = min(e[i] for i in kept)
    for i in kept:
        Z += np.exp(-(e[i] - e0) / (K_B * T))
    for t, d in enumerate(dip):
        ops_t = Ops([d], [model.restrictions], n, W)
        for i in kept:
            psi = gb.vec_from_eigen([0], [i], 1e-7)
            seed_vec = vec_apply(ops_t, psi, [0], [0], 1e-7, normalize=True)
            _, _, norm2 = seed_vec.get()
            xb = Batch(ops_h, [0])
            xb.seed_from_vec(seed_vec)
            xb.build(fixed_basis=False)
            K = int(min(krylovSize, max(int(xb.chain_n[0]), 1)))
            xb.lanczos(K, reorth=reorth, want_coeffs=False)
            _, spec = xb.cf([e[i]], norm2, [1], [1], K_B * T, [0], 1, omega, delta)
            acc += np.exp(-(e[i] - e0) / (K_B * T)) * spec[0]
    return acc / Z
