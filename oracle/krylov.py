"""Lanczos tridiagonalization, continued fraction, ground states (oracle).

Restates ``get_tridiagonal_krylov_vectors`` (impurityModel/ed/finite.py:
1897-1995, serial branch 1974-1992), the continued fraction of ``getGreen``
(impurityModel/ed/spectra.py:677-680) and ``eigensystem`` (finite.py:58-118).
"""

from math import sqrt

import numpy as np
import scipy.sparse.linalg


def lanczos(h, psi0, krylovSize, reorth="none"):
    """Return ``(alpha, beta)`` of the Krylov tridiagonal matrix.

    ``reorth="none"`` is the reference's plain three-term recurrence:
    alpha_j = Re(conj(H v_j) . v_j), beta_j = ||w||, beta == 0 raises
    ``ValueError("Warning: beta==0, implementation absent!")``, K = min(K, n).
    ``reorth="full"`` (no reference counterpart) additionally removes the
    components of w along all previous Krylov vectors twice (classical
    Gram-Schmidt, two passes) before beta is taken.
    """
    n = len(psi0)
    K = min(krylovSize, n)
    alpha = np.zeros(K, dtype=float)
    beta = np.zeros(K - 1, dtype=float)
    v0 = np.array(psi0, dtype=complex)
    V = [v0] if reorth == "full" else None
    wp = h.dot(v0)
    alpha[0] = np.dot(np.conj(wp), v0).real
    w = wp - alpha[0] * v0
    for j in range(1, K):
        if reorth == "full":
            Vm = np.array(V)
            for _ in range(2):
                w = w - Vm.T @ (np.conj(Vm) @ w)
        beta[j - 1] = sqrt(np.sum(np.abs(w) ** 2))
        if beta[j - 1] != 0:
            v1 = w / beta[j - 1]
        else:
            raise ValueError("Warning: beta==0, implementation absent!")
        wp = h.dot(v1)
        alpha[j] = np.dot(np.conj(wp), v1).real
        w = wp - alpha[j] * v1 - beta[j - 1] * v0
        v0 = v1
        if reorth == "full":
            V.append(v1)
    return alpha, beta


def continued_fraction(alpha, beta, omega, delta, e):
    """G(w) = 1/(z - a0 - b0^2/(z - a1 - ...)), z = w + i delta + e.

    Bottom-up evaluation as spectra.py:677-680.
    """
    K = len(alpha)
    z = np.asarray(omega) + 1j * delta + e
    g = None
    for i in range(K - 1, -1, -1):
        g = 1.0 / (z - alpha[i]) if i == K - 1 else 1.0 / (z - alpha[i] - beta[i] ** 2 * g)
    return g


def eigensystem(n_spin_orbitals, hOp, basis, nPsiMax, groundDiagMode="Lanczos", eigenValueTol=1e-9,
                slaterWeightMin=1e-7, h=None):
    """Lowest eigenpairs and truncated eigenvector dicts (finite.py:58-118).

    ``h`` may be passed when the caller already has the fixed-basis matrix.
    Eigenvectors keep only components with |v_i|^2 >= slaterWeightMin and are
    not renormalized (finite.py:114-117).
    """
    from .hamiltonian import get_hamiltonian_matrix

    if h is None:
        h = get_hamiltonian_matrix(n_spin_orbitals, hOp, basis)
    if groundDiagMode == "full":
        es, vecs = np.linalg.eigh(h.toarray())
        es = es[:nPsiMax]
        vecs = vecs[:, :nPsiMax]
    else:
        es, vecs = scipy.sparse.linalg.eigsh(h, k=nPsiMax, which="SA", tol=eigenValueTol)
        idx = np.argsort(es)
        es = es[idx]
        vecs = vecs[:, idx]
    psis = [
        {basis[i]: vecs[i, vi] for i in range(len(basis)) if slaterWeightMin <= abs(vecs[i, vi]) ** 2}
        for vi in range(len(es))
    ]
    return es, psis
