"""Static expectation values of the 3d shell (oracle).

TEST INFRASTRUCTURE ONLY (see the package header).  Restates
impurityModel/ed/finite.py:843-1298 on int-keyed dicts: Lz, Sz (diagonal
sums, 843-893), S^2 and L^2 as |Sz psi|^2 + (|S+ psi|^2 + |S- psi|^2)/2
(895-935 with the apply functions 1115-1298), the trace of the 3d density
matrix (937-964), the density matrix <psi|c_j^dag c_i|psi> (967-1012), its
cubic-harmonics form (1015-1060) and the eg/t2g occupations (1063-1112).
Spin index 1 is spin up (finite.py:891).
"""

from collections import OrderedDict
from math import sqrt

import numpy as np

from .applyop import add, apply_op, inner, norm2
from .basis import c2i

L3D = 2


def get_spherical_2_cubic_matrix():
    """d-shell spherical -> cubic harmonics matrix (finite.py:220-268, l = 2):
    column j is cubic harmonic j in the spherical basis m = -2..2."""
    r2 = 1 / np.sqrt(2)
    u = np.zeros((5, 5), dtype=complex)
    u[2, 0] = 1
    u[0, 1] = u[4, 1] = r2
    u[1, 2] = u[3, 2] = -1j * r2
    u[1, 3], u[3, 3] = r2, -r2
    u[0, 4], u[4, 4] = 1j * r2, -1j * r2
    return u


def _n(nBaths):
    return sum(2 * (2 * ang + 1) + nb for ang, nb in nBaths.items())


def _occupied(state, n):
    return [i for i in range(n) if (state >> (n - 1 - i)) & 1]


def _spin_orb(nBaths):
    """index -> (s, m) of the 3d spin-orbitals."""
    return {c2i(nBaths, (L3D, s, m)): (s, m) for s in range(2) for m in range(-L3D, L3D + 1)}


def getLz3d(nBaths, psi):
    n, so = _n(nBaths), _spin_orb(nBaths)
    return sum(sum(so[i][1] for i in _occupied(st, n) if i in so) * abs(a) ** 2 for st, a in psi.items())


def getSz3d(nBaths, psi):
    n, so = _n(nBaths), _spin_orb(nBaths)
    return sum(sum((0.5 if so[i][0] == 1 else -0.5) for i in _occupied(st, n) if i in so) * abs(a) ** 2
               for st, a in psi.items())


def getTraceDensityMatrix(nBaths, psi, l=2):
    n = _n(nBaths)
    idx = [c2i(nBaths, (l, s, m)) for s in range(2) for m in range(-l, l + 1)]
    return sum(sum(1 for i in idx if (st >> (n - 1 - i)) & 1) * abs(a) ** 2 for st, a in psi.items())


def _apply_pairs(nBaths, psi, pairs):
    """sum_k coef_k c_i^dag c_j psi over pairs (i_so, j_so, coef), applied as a then c."""
    n = _n(nBaths)
    out = {}
    for (i, j, coef) in pairs:
        term = apply_op(n, {((c2i(nBaths, i), "c"), (c2i(nBaths, j), "a")): 1.0}, psi, slaterWeightMin=0.0)
        out = add(out, term, coef)
    return out


def applySz3d(nBaths, psi):
    return _apply_pairs(nBaths, psi, [((2, s, m), (2, s, m), 0.5 if s == 1 else -0.5) for s in range(2)
                                      for m in range(-2, 3)])


def applyLz3d(nBaths, psi):
    return _apply_pairs(nBaths, psi, [((2, s, m), (2, s, m), m) for s in range(2) for m in range(-2, 3)])


def applySplus3d(nBaths, psi):
    return _apply_pairs(nBaths, psi, [((2, 1, m), (2, 0, m), 1.0) for m in range(-2, 3)])


def applySminus3d(nBaths, psi):
    return _apply_pairs(nBaths, psi, [((2, 0, m), (2, 1, m), 1.0) for m in range(-2, 3)])


def applyLplus3d(nBaths, psi):
    return _apply_pairs(nBaths, psi, [((2, s, m + 1), (2, s, m), sqrt((2 - m) * (2 + m + 1))) for s in range(2)
                                      for m in range(-2, 2)])


def applyLminus3d(nBaths, psi):
    return _apply_pairs(nBaths, psi, [((2, s, m - 1), (2, s, m), sqrt((2 + m) * (2 - m + 1))) for s in range(2)
                                      for m in range(-1, 3)])


def getSsqr3d(nBaths, psi):
    return (norm2(applySz3d(nBaths, psi)) + 0.5 * (norm2(applySplus3d(nBaths, psi))
                                                    + norm2(applySminus3d(nBaths, psi))))


def getLsqr3d(nBaths, psi):
    return (norm2(applyLz3d(nBaths, psi)) + 0.5 * (norm2(applyLplus3d(nBaths, psi))
                                                    + norm2(applyLminus3d(nBaths, psi))))


def getDensityMatrix(nBaths, psi, l=2):
    """{((l,si,mi),(l,sj,mj)): <psi|c_j^dag c_i|psi>} in the reference's loop order."""
    n = _n(nBaths)
    rho = OrderedDict()
    for si in range(2):
        for sj in range(2):
            for mi in range(-l, l + 1):
                for mj in range(-l, l + 1):
                    i, j = c2i(nBaths, (l, si, mi)), c2i(nBaths, (l, sj, mj))
                    new = apply_op(n, {((j, "c"), (i, "a")): 1.0}, psi, slaterWeightMin=0.0)
                    v = inner(psi, new)
                    if v != 0:
                        rho[((l, si, mi), (l, sj, mj))] = v
    return rho


def getDensityMatrixCubic(nBaths, psi):
    rho = getDensityMatrix(nBaths, psi)
    u = get_spherical_2_cubic_matrix()
    out = OrderedDict()
    for i in range(5):
        for j in range(5):
            for si in range(2):
                for sj in range(2):
                    for k, mk in enumerate(range(-2, 3)):
                        for m, mm in enumerate(range(-2, 3)):
                            e = ((2, si, mm), (2, sj, mk))
                            if e in rho:
                                t = np.conj(u[m, i]) * rho[e] * u[k, j]
                                if t != 0:
                                    key = ((si, i), (sj, j))
                                    out[key] = out[key] + t if key in out else t
    return out


def getEgT2gOccupation(nBaths, psi):
    """[eg_dn, eg_up, t2g_dn, t2g_up] from the spin-diagonal density matrix in cubic harmonics."""
    n = _n(nBaths)
    u = get_spherical_2_cubic_matrix()
    occ = [0, 0, 0, 0]
    for i in range(5):
        for j, mj in enumerate(range(-2, 3)):
            for k, mk in enumerate(range(-2, 3)):
                for s in range(2):
                    jj, kk = c2i(nBaths, (2, s, mj)), c2i(nBaths, (2, s, mk))
                    new = apply_op(n, {((jj, "c"), (kk, "a")): 1.0}, psi, slaterWeightMin=0.0)
                    v = u[j, i] * np.conj(u[k, i]) * inner(psi, new)
                    occ[(0 if i < 2 else 2) + s] += v
    return [o.real if abs(o.imag) < 1e-12 else o for o in occ]
