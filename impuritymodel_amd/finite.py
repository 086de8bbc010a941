"""Drop-in replacement of impurityModel.ed.finite's hot path on the GPU.

Same names, argument order, defaults, return types and exceptions as the
reference (impurityModel/ed/finite.py); states are the reference's ``bytes``
keys and operators its process dicts.  Every operator application, basis
expansion, Hamiltonian build, Lanczos recurrence and eigensolve runs in
libimped.so's HIP kernels; this module converts dicts to packed words and
back.

The ``h_dict`` memo of the BFS functions is filled as the reference fills
it: after a closure on the device, every basis state's H|PS> (restricted,
pruned at 1e-12) is written into the caller's dict from the device CSR, so
the reference's own composition ``expand_basis`` ->
``get_hamiltonian_matrix_from_h_dict`` works unchanged.  The device path does
not read the memo back (H|PS> is recomputed on the GPU).
``parallelization_mode``/MPI arguments are accepted and ignored (one GPU).
"""

from collections import OrderedDict  # noqa: F401  (re-exported for callers)

import numpy as np
import scipy.sparse

from . import opcompile
from .engine import Batch, Ops, vec_apply, vec_from_host
from .operators import (  # noqa: F401  (reference names)
    addOps,
    assert_hermitian,
    c2i,
    daggerOp,
    dc_MLFT,
    gauntC,
    get2p3dSlaterCondonUop,
    get_spherical_2_cubic_matrix,
    getSOCop,
    getU,
    getUop,
    gethHfieldop,
    i2c,
)

# ------------------------------------------------------------------ helpers

_ops_cache = {}


def _ops_for(n, op, restrictions=None):
    """Compiled single-model Ops for an operator dict (cached per call site)."""
    key = (n, id(op), len(op), None if restrictions is None else tuple(sorted((tuple(sorted(k)), v) for k, v in restrictions.items())))
    ent = _ops_cache.get(key)
    if ent is not None and ent[0] is op:
        return ent[1]
    o = Ops([op], [restrictions], n)
    if len(_ops_cache) > 64:
        _ops_cache.clear()
    _ops_cache[key] = (op, o)
    return o


def _words(keys, n):
    return opcompile.bytes_to_words(list(keys), n)


def _psi_arrays(psi, n):
    keys = list(psi.keys())
    amps = np.array([complex(a) for a in psi.values()], np.complex128)
    return keys, _words(keys, n), amps


# ------------------------------------------------------------------ operator application


def applyOp(n_spin_orbitals, op, psi, slaterWeightMin=1e-12, restrictions=None, opResult=None):
    """Return op|psi> (finite.py:1301-1511) computed on the GPU.

    Without ``opResult`` the whole application (bit tests, signs, merge of
    equal targets in the reference's order, restrictions, pruning) runs in
    ``imped_vec_apply``.  With ``opResult`` the memo semantics of the reference
    are kept: states already in the memo use its stored (possibly pruned)
    result, new states are applied on the GPU one by one and stored pruned at
    ``|value|^2 < slaterWeightMin``; the sum is assembled in psi's order.
    """
    n = n_spin_orbitals
    if len(psi) == 0:
        return {}
    ops = _ops_for(n, op, restrictions)
    W = ops.W
    if opResult is None:
        _, words, amps = _psi_arrays(psi, n)
        src = vec_from_host(W, [words], [amps])
        out = vec_apply(ops, src, [0], [0], slaterWeightMin)
        k, a, _ = out.get()
        return dict(zip(opcompile.words_to_bytes(k, n), a.tolist()))
    new_states = [s for s in psi if s not in opResult]
    per_state = {}
    if new_states:
        words = _words(new_states, n)
        src = vec_from_host(W, [words[i:i + 1] for i in range(len(new_states))],
                            [np.ones(1, complex)] * len(new_states))
        out = vec_apply(ops, src, np.arange(len(new_states)), np.zeros(len(new_states), np.int32), 0.0)
        k, a, _ = out.get()
        off = out.offsets
        kb = opcompile.words_to_bytes(k, n)
        for i, s in enumerate(new_states):
            per_state[s] = list(zip(kb[off[i]:off[i + 1]], a[off[i]:off[i + 1]].tolist()))
    psi_new = {}
    for state, amp in psi.items():
        if state in per_state:
            memo = {}
            for t, v in per_state[state]:
                psi_new[t] = psi_new[t] + amp * v if t in psi_new else amp * v
                memo[t] = v
            opResult[state] = {t: v for t, v in memo.items() if abs(v) ** 2 >= slaterWeightMin}
        else:
            for t, v in opResult[state].items():
                psi_new[t] = psi_new[t] + v * amp if t in psi_new else v * amp
    return {s: a for s, a in psi_new.items() if abs(a) ** 2 >= slaterWeightMin}


def c(n_spin_orbitals, i, psi):
    """c_i^dagger |psi> (finite.py:449-473) via applyOp on the GPU."""
    return applyOp(n_spin_orbitals, {((i, "c"),): 1}, psi, slaterWeightMin=0.0)


def a(n_spin_orbitals, i, psi):
    """c_i |psi> (finite.py:422-446) via applyOp on the GPU."""
    return applyOp(n_spin_orbitals, {((i, "a"),): 1}, psi, slaterWeightMin=0.0)


def inner(a, b):
    """<a|b> of two dict states (finite.py:377-394); host dict algebra."""
    acc = 0
    for state, amp in b.items():
        if state in a:
            acc += np.conj(a[state]) * amp
    return acc


def addToFirst(psi1, psi2, mul=1):
    """psi1 += mul*psi2 in place (finite.py:397-419); host dict algebra."""
    for s, amp in psi2.items():
        psi1[s] = psi1[s] + amp * mul if s in psi1 else amp * mul


def add(psi1, psi2, mul=1):
    """psi1 + mul*psi2 (finite.py:1998-2020); host dict algebra."""
    psi = psi1.copy()
    for s, amp in psi2.items():
        psi[s] = psi[s] + mul * amp if s in psi else mul * amp
    return psi


def norm2(psi):
    """<psi|psi> (finite.py:2023-2033); host dict algebra."""
    return sum(abs(amp) ** 2 for amp in psi.values())


# ------------------------------------------------------------------ bases and Hamiltonians


def basis_shells(nBaths, valBaths, dnValBaths, dnConBaths, dnTol, n0imp):
    """Per-shell enumeration parameters of get_basis for imped_enumerate_basis.

    Impurity spin-orbitals (l, s, m) of a shell are consecutive in c2i order,
    and so are its valence and conduction bath states.
    """
    shells = []
    for l in nBaths.keys():
        assert valBaths[l] <= nBaths[l]
        imp_first = c2i(nBaths, (l, 0, -l))
        nv = valBaths[l]
        nc = nBaths[l] - valBaths[l]
        v_first = c2i(nBaths, (l, 0)) if nv > 0 else 0
        c_first = c2i(nBaths, (l, nv)) if nc > 0 else 0
        shells.append((imp_first, 2 * (2 * l + 1), v_first, nv, c_first, nc, n0imp[l], dnTol[l], dnValBaths[l],
                       dnConBaths[l]))
    return shells


def get_basis_words(nBaths, valBaths, dnValBaths, dnConBaths, dnTol, n0imp, W=None):
    """get_basis as packed words [N, W] (native enumeration)."""
    from .engine import enumerate_basis

    n = sum(2 * (2 * ang + 1) + nb for ang, nb in nBaths.items())
    W = W or opcompile.n_words(n)
    return enumerate_basis(n, W, basis_shells(nBaths, valBaths, dnValBaths, dnConBaths, dnTol, n0imp))


def get_basis(nBaths, valBaths, dnValBaths, dnConBaths, dnTol, n0imp):
    """Restricted initial basis as bytes keys in itertools order (finite.py:294-374).

    Enumerated by libimped's host enumerator: for each shell the nested
    combinations of impurity, valence and conduction indices per occupation
    partition, shells combined as itertools.product, not sorted.
    """
    n = sum(2 * (2 * ang + 1) + nb for ang, nb in nBaths.items())
    words = get_basis_words(nBaths, valBaths, dnValBaths, dnConBaths, dnTol, n0imp)
    return tuple(opcompile.words_to_bytes(words, n))


def _reference_csr(b, chain=0):
    """scipy CSR h[i, j] = <i|H|j> of one chain from the device rows."""
    rowptr, col, val = b.csr()
    r0, r1 = int(b.row_base[chain]), int(b.row_base[chain + 1])
    e0, e1 = int(rowptr[r0]), int(rowptr[r1])
    n = r1 - r0
    rows = np.repeat(np.arange(n), np.diff(rowptr[r0:r1 + 1]))
    # device row r holds conj(<t|H|r>) at column t -> reference entry (t, r)
    return scipy.sparse.csr_matrix((np.conj(val[e0:e1]), (col[e0:e1], rows)), shape=(n, n))


def get_hamiltonian_matrix(n_spin_orbitals, hOp, basis, mode="sparse_MPI"):
    """H in a fixed basis (finite.py:1514-1610) built on the GPU.

    Each column is H applied to one basis state without restrictions, targets
    outside the basis dropped, |h|^2 < 1e-12 pruned.  ``mode`` "dense_*"
    returns a dense ndarray, the sparse modes a scipy CSR matrix.
    """
    ops = _ops_for(n_spin_orbitals, hOp, None)
    b = Batch(ops, [0])
    b.set_seeds([_words(basis, n_spin_orbitals)])
    b.build(fixed_basis=True, weight_min=1e-12)
    h = _reference_csr(b)
    if mode in ("dense_serial", "dense_MPI"):
        return h.toarray()
    return h


def _fill_h_dict(b, basis, h_dict):
    """h_dict[basis[j]] = {basis[k]: <k|H|j>} from the device rows of chain 0
    (finite.py:1744-1745: the memo holds H|PS> for every expanded state)."""
    rowptr, col, val = b.csr()
    vals = np.conj(val).tolist()
    for j, s in enumerate(basis):
        if s not in h_dict:
            lo, hi = int(rowptr[j]), int(rowptr[j + 1])
            h_dict[s] = {basis[k]: v for k, v in zip(col[lo:hi].tolist(), vals[lo:hi])}


def expand_basis(n_spin_orbitals, h_dict, hOp, basis0, restrictions, parallelization_mode="serial"):
    """BFS closure of basis0 under hOp with restrictions (finite.py:1696-1803), on the GPU.

    Fills ``h_dict`` (when not None) with H|PS> of every basis state not yet
    in it, as the reference does."""
    if parallelization_mode not in ("serial", "H_build"):
        raise Exception("Wrong parallelization parameter.")
    b = _expanded_batch(n_spin_orbitals, hOp, list(basis0), restrictions)
    basis = tuple(opcompile.words_to_bytes(b.basis_words(), n_spin_orbitals))
    if h_dict is not None:
        _fill_h_dict(b, basis, h_dict)
    return basis


def _expanded_batch(n, hOp, seeds, restrictions, seed_amps=None):
    ops = _ops_for(n, hOp, restrictions)
    b = Batch(ops, [0])
    b.set_seeds([_words(seeds, n)], None if seed_amps is None else [seed_amps])
    b.build(fixed_basis=False, weight_min=1e-12)
    return b


def get_hamiltonian_matrix_from_h_dict(h_dict, basis, parallelization_mode="serial", return_h_local=False,
                                       mode="sparse"):
    """CSR (or dense) matrix from an H|PS> memo (finite.py:1613-1693).

    A host format conversion of a dict the caller already holds.  As in the
    reference, the serial mode raises ``KeyError`` for a basis state missing
    from ``h_dict``; "H_build" uses the states present (finite.py:1671-1682).
    """
    if parallelization_mode == "serial":
        assert return_h_local is False
    if mode not in ("sparse", "dense") or parallelization_mode not in ("serial", "H_build") or (
            mode == "dense" and parallelization_mode == "H_build"):
        raise Exception("Wrong input parameters")
    n = len(basis)
    basis_index = {basis[i]: i for i in range(n)}
    data, row, col = [], [], []
    if parallelization_mode == "serial":
        columns = range(n)
    else:
        present = set(basis).intersection(h_dict.keys())
        columns = [j for j in range(n) if basis[j] in present]
    for j in columns:
        for k, v in h_dict[basis[j]].items():
            data.append(v)
            col.append(j)
            row.append(basis_index[k])
    h = scipy.sparse.csr_matrix((np.array(data, complex), (np.array(row, np.int64), np.array(col, np.int64))),
                                shape=(n, n))
    return (h.toarray() if mode == "dense" else h), basis_index


def expand_basis_and_hamiltonian(n_spin_orbitals, h_dict, hOp, basis0, restrictions, parallelization_mode="serial",
                                 return_h_local=False):
    """(CSR h, basis_index) of the BFS closure (finite.py:1806-1894), built on the GPU."""
    if parallelization_mode not in ("serial", "H_build"):
        raise Exception("Wrong parallelization parameter.")
    b = _expanded_batch(n_spin_orbitals, hOp, list(basis0), restrictions)
    basis = opcompile.words_to_bytes(b.basis_words(), n_spin_orbitals)
    if h_dict is not None:
        _fill_h_dict(b, basis, h_dict)
    return _reference_csr(b), {s: i for i, s in enumerate(basis)}


# ------------------------------------------------------------------ Krylov


def _batch_from_matrix(h):
    """A one-chain batch holding the rows of a given scipy matrix h."""
    h = scipy.sparse.csr_matrix(h)
    n = h.shape[0]
    dummy = Ops([{}], [None], 1)
    b = Batch(dummy, [0])
    b._dummy_ops = dummy
    b.set_csr([n], h.indptr.astype(np.int64), h.indices.astype(np.int32), h.data.astype(np.complex128))
    return b


def get_tridiagonal_krylov_vectors(h, psi0, krylovSize, h_local=False, mode="sparse", reorth="none"):
    """alpha, beta of the Krylov tridiagonal matrix (finite.py:1897-1995) on the GPU.

    Plain three-term Lanczos as the reference (``reorth="none"``); ``"full"``
    adds full reorthogonalization.  beta == 0 raises the reference's
    ``ValueError("Warning: beta==0, implementation absent!")``.
    """
    b = _batch_from_matrix(h)
    n = len(psi0)
    K = min(krylovSize, n)
    alpha, beta, kout, status = b.lanczos(K, reorth=1 if reorth == "full" else 0,
                                          v0=np.asarray(psi0, dtype=np.complex128))
    if status[0] != 0:
        raise ValueError("Warning: beta==0, implementation absent!")
    return alpha[0, :K].copy(), beta[0, : K - 1].copy()


def eigensystem(n_spin_orbitals, hOp, basis, nPsiMax, groundDiagMode="Lanczos", eigenValueTol=1e-9,
                slaterWeightMin=1e-7):
    """Lowest eigenpairs as (es, psis) (finite.py:58-118), solved on the GPU.

    The fixed-basis H is built on the device and diagonalized by Lanczos with
    full reorthogonalization; "full" runs it to the whole space.  psis keep
    the components with |v_i|^2 >= slaterWeightMin, not renormalized.
    """
    if groundDiagMode not in ("full", "Lanczos"):
        print("Wrong diagonalization mode")
    n = n_spin_orbitals
    ops = _ops_for(n, hOp, None)
    b = Batch(ops, [0])
    b.set_seeds([_words(basis, n)])
    b.build(fixed_basis=True, weight_min=1e-12)
    nev = min(nPsiMax, len(basis))
    if groundDiagMode == "full":
        es, _, _, _ = b.eigen(nev, tol=0.0, max_steps=len(basis), want_vectors=True, allow_noconv=True)
    else:
        es, _, _, _ = b.eigen(nev, tol=eigenValueTol, max_steps=max(600, 4 * nev), want_vectors=True)
    vec = b.vec_from_eigen([0] * nev, list(range(nev)), slaterWeightMin)
    k, amps, _ = vec.get()
    off = vec.offsets
    kb = opcompile.words_to_bytes(k, n)
    psis = [dict(zip(kb[off[i]:off[i + 1]], amps[off[i]:off[i + 1]].tolist())) for i in range(nev)]
    return np.array(es[0]), psis
