"""Hamiltonian matrices and BFS basis expansion (oracle).

Restates the serial branches of impurityModel/ed/finite.py:
``get_hamiltonian_matrix`` (1514-1610), ``get_hamiltonian_matrix_from_h_dict``
(1613-1693), ``expand_basis`` (1696-1803) and
``expand_basis_and_hamiltonian`` (1806-1894).  Matrices are scipy CSR with
``h[i, j] = <i|H|j>``: column j holds the targets of H|basis[j]>.
"""

import numpy as np
import scipy.sparse

from .applyop import apply_op


def get_hamiltonian_matrix(n, hOp, basis):
    """H on a fixed basis (finite.py:1514-1610, sparse branch).

    Each column is ``applyOp(hOp, {basis[j]: 1})`` without restrictions at
    the default ``slaterWeightMin`` 1e-12; targets outside the basis are
    dropped.
    """
    nb = len(basis)
    index = {b: i for i, b in enumerate(basis)}
    data, row, col = [], [], []
    for j in range(nb):
        res = apply_op(n, hOp, {basis[j]: 1})
        for k, v in res.items():
            if k in index:
                data.append(v)
                col.append(j)
                row.append(index[k])
    return scipy.sparse.csr_matrix(
        (np.array(data, dtype=complex), (np.array(row, dtype=np.int64), np.array(col, dtype=np.int64))),
        shape=(nb, nb),
    )


def expand_basis(n, h_dict, hOp, basis0, restrictions):
    """BFS closure of ``basis0`` under H (finite.py:1696-1750, serial).

    Seeds keep their order; each new layer is appended in ascending order.
    ``h_dict`` memoizes ``applyOp(hOp, {b: 1}, restrictions=...)`` per state.
    """
    basis = list(basis0)
    i = 0
    nb = len(basis)
    while i < nb:
        basis_set = frozenset(basis)
        new = set()
        for b in basis[i:nb]:
            if b in h_dict:
                res = h_dict[b]
            else:
                res = apply_op(n, hOp, {b: 1}, restrictions=restrictions)
                h_dict[b] = res
            new.update(set(res.keys()).difference(basis_set))
        i = nb
        basis += sorted(new)
        nb = len(basis)
    return tuple(basis)


def get_hamiltonian_matrix_from_h_dict(h_dict, basis):
    """CSR from the memo (finite.py:1613-1670, serial sparse branch)."""
    nb = len(basis)
    index = {b: i for i, b in enumerate(basis)}
    data, row, col = [], [], []
    for j in range(nb):
        for k, v in h_dict[basis[j]].items():
            data.append(v)
            col.append(j)
            row.append(index[k])
    h = scipy.sparse.csr_matrix(
        (np.array(data, dtype=complex), (np.array(row, dtype=np.int64), np.array(col, dtype=np.int64))),
        shape=(nb, nb),
    )
    return h, index


def expand_basis_and_hamiltonian(n, h_dict, hOp, basis0, restrictions):
    """finite.py:1806-1894 (serial): returns ``(csr h, basis_index)``."""
    basis = expand_basis(n, h_dict, hOp, basis0, restrictions)
    return get_hamiltonian_matrix_from_h_dict(h_dict, basis)
