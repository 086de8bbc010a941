"""Static expectation values of the 3d shell on the GPU (drop-in).

Same names and return values as impurityModel/ed/finite.py:121-175 and
843-1112: Lz, Sz, S^2, L^2, the trace of the 3d density matrix, the density
matrix in spherical and cubic harmonics, the eg/t2g occupations and the two
printers.  Every operator a call needs (Sz, S+, S-, Lz, L+, L-, the number
operator, or the 100 pairs c_j^dag c_i of the density matrix) is applied to
every state in one ``imped_vec_apply``; norms come from the device, inner
products <psi|O psi> from a key join on the host.
"""

from collections import OrderedDict
from math import sqrt

import numpy as np

from . import opcompile
from .average import k_B, thermal_average
from .engine import Ops, vec_apply, vec_from_host
from .finite import _words
from .operators import c2i, get_spherical_2_cubic_matrix, n_spin_orbitals


def _pairs_op(nBaths, pairs):
    """sum coef c_i^dag c_j over (i_spin_orbital, j_spin_orbital, coef)."""
    op = {}
    for i, j, coef in pairs:
        key = ((c2i(nBaths, i), "c"), (c2i(nBaths, j), "a"))
        op[key] = op.get(key, 0.0) + coef
    return op


def _sz(nb):
    return _pairs_op(nb, [((2, s, m), (2, s, m), 0.5 if s == 1 else -0.5) for s in range(2) for m in range(-2, 3)])


def _lz(nb):
    return _pairs_op(nb, [((2, s, m), (2, s, m), float(m)) for s in range(2) for m in range(-2, 3)])


def _splus(nb):
    return _pairs_op(nb, [((2, 1, m), (2, 0, m), 1.0) for m in range(-2, 3)])


def _sminus(nb):
    return _pairs_op(nb, [((2, 0, m), (2, 1, m), 1.0) for m in range(-2, 3)])


def _lplus(nb):
    return _pairs_op(nb, [((2, s, m + 1), (2, s, m), sqrt((2 - m) * (3 + m))) for s in range(2) for m in range(-2, 2)])


def _lminus(nb):
    return _pairs_op(nb, [((2, s, m - 1), (2, s, m), sqrt((2 + m) * (3 - m))) for s in range(2) for m in range(-1, 3)])


def _n3d(nb, l=2):
    return _pairs_op(nb, [((l, s, m), (l, s, m), 1.0) for s in range(2) for m in range(-l, l + 1)])


def _apply_all(nBaths, psis, ops):
    """Every operator on every state: (src keys/amps per state, out Vec, out offsets);
    output k * len(psis) + p is operator k applied to state p."""
    n = n_spin_orbitals(nBaths)
    W = opcompile.n_words(n)
    keys = [_words(list(p.keys()), n) for p in psis]
    amps = [np.array([complex(a) for a in p.values()], np.complex128) for p in psis]
    src = vec_from_host(W, keys, amps)
    tab = Ops(ops, None, n, W)
    P = len(psis)
    out = vec_apply(tab, src, np.tile(np.arange(P, dtype=np.int32), len(ops)),
                    np.repeat(np.arange(len(ops), dtype=np.int32), P), 0.0)
    return keys, amps, out


def _inner(ka, aa, kb, ab):
    """<a|b> over matching keys."""
    if len(ka) == 0 or len(kb) == 0:
        return 0j
    W = ka.shape[1]
    dt = np.dtype((np.void, 8 * W))
    va = np.ascontiguousarray(ka).view(dt).ravel()
    vb = np.ascontiguousarray(kb).view(dt).ravel()
    _, ia, ib = np.intersect1d(va, vb, assume_unique=True, return_indices=True)
    return complex(np.sum(np.conj(aa[ia]) * ab[ib]))


def _expect(nBaths, psis, ops):
    """<psi_p|O_k|psi_p> [k][p]."""
    keys, amps, out = _apply_all(nBaths, psis, ops)
    ok, oa, _ = out.get()
    off = out.offsets
    P = len(psis)
    r = np.zeros((len(ops), P), complex)
    for k in range(len(ops)):
        for p in range(P):
            o = k * P + p
            r[k, p] = _inner(keys[p], amps[p], ok[off[o]:off[o + 1]], oa[off[o]:off[o + 1]])
    return r


def _norms(nBaths, psis, ops):
    """|O_k psi_p|^2 [k][p] from the device."""
    _, _, out = _apply_all(nBaths, psis, ops)
    return out.norm2().reshape(len(ops), len(psis))


def getLz3d(nBaths, psi):
    """<psi|Lz_3d|psi> (finite.py:843-866)."""
    return float(_expect(nBaths, [psi], [_lz(nBaths)])[0, 0].real)


def getSz3d(nBaths, psi):
    """<psi|Sz_3d|psi> (finite.py:869-892)."""
    return float(_expect(nBaths, [psi], [_sz(nBaths)])[0, 0].real)


def getSsqr3d(nBaths, psi, tol=1e-8):
    """|Sz psi|^2 + (|S+ psi|^2 + |S- psi|^2)/2 (finite.py:895-913)."""
    nz, npl, nmi = _norms(nBaths, [psi], [_sz(nBaths), _splus(nBaths), _sminus(nBaths)])[:, 0]
    return float(nz + 0.5 * (npl + nmi))


def getLsqr3d(nBaths, psi, tol=1e-8):
    """|Lz psi|^2 + (|L+ psi|^2 + |L- psi|^2)/2 (finite.py:916-934)."""
    nz, npl, nmi = _norms(nBaths, [psi], [_lz(nBaths), _lplus(nBaths), _lminus(nBaths)])[:, 0]
    return float(nz + 0.5 * (npl + nmi))


def getTraceDensityMatrix(nBaths, psi, l=2):
    """<psi|sum_i n_i|psi> over the 3d shell (finite.py:937-964)."""
    return float(_expect(nBaths, [psi], [_n3d(nBaths, l)])[0, 0].real)


def _rho_ops(nBaths, l=2):
    keys, ops = [], []
    for si in range(2):
        for sj in range(2):
            for mi in range(-l, l + 1):
                for mj in range(-l, l + 1):
                    keys.append(((l, si, mi), (l, sj, mj)))
                    ops.append({((c2i(nBaths, (l, sj, mj)), "c"), (c2i(nBaths, (l, si, mi)), "a")): 1.0})
    return keys, ops


def getDensityMatrix(nBaths, psi, l=2):
    """{((l,si,mi),(l,sj,mj)): <psi|c_j^dag c_i|psi>}, nonzero entries in the
    reference's loop order (finite.py:967-1012)."""
    keys, ops = _rho_ops(nBaths, l)
    vals = _expect(nBaths, [psi], ops)[:, 0]
    return OrderedDict((k, v) for k, v in zip(keys, vals) if v != 0)


def getDensityMatrixCubic(nBaths, psi):
    """Density matrix in cubic harmonics (finite.py:1015-1060)."""
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
    """[eg_dn, eg_up, t2g_dn, t2g_up] (finite.py:1063-1112)."""
    rho = getDensityMatrix(nBaths, psi)
    u = get_spherical_2_cubic_matrix()
    occ = [0j, 0j, 0j, 0j]
    for i in range(5):
        for j, mj in enumerate(range(-2, 3)):
            for k, mk in enumerate(range(-2, 3)):
                for s in range(2):
                    v = rho.get(((2, s, mk), (2, s, mj)), 0)
                    occ[(0 if i < 2 else 2) + s] += u[j, i] * np.conj(u[k, i]) * v
    out = []
    for o in occ:
        if abs(o.imag) < 1e-12:
            out.append(o.real)
        else:
            print("Warning: Complex occupation")
            out.append(o)
    return out


def printExpValues(nBaths, es, psis, n=None):
    """E0 and a table of N(3d), eg/t2g occupations, Lz, Sz, L^2, S^2 (finite.py:121-145)."""
    if n is None:
        n = len(es)
    es = np.asarray(es)
    print("E0 = {:7.4f}".format(es[0]))
    print("  i  E-E0  N(3d) N(egDn) N(egUp) N(t2gDn) N(t2gUp) Lz(3d) Sz(3d) L^2(3d) S^2(3d)")
    for i, (e, psi) in enumerate(zip(es[:n] - es[0], psis[:n])):
        oc = getEgT2gOccupation(nBaths, psi)
        print(("{:3d} {:6.3f} {:5.2f} {:6.3f} {:7.3f} {:8.3f} {:7.3f} {:7.2f} {:6.2f} {:7.2f} {:7.2f}").format(
            i, e, getTraceDensityMatrix(nBaths, psi), oc[0], oc[1], oc[2], oc[3], getLz3d(nBaths, psi),
            getSz3d(nBaths, psi), getLsqr3d(nBaths, psi), getSsqr3d(nBaths, psi)))
    print("\n")


def printThermalExpValues(nBaths, es, psis, T=300, cutOff=10):
    """Boltzmann averages of the same quantities (finite.py:148-175)."""
    e = np.asarray(es) - es[0]
    mask = e < cutOff * k_B * T
    e = e[mask]
    psis = [p for p, keep in zip(psis, mask) if keep]
    occs = thermal_average(e, np.array([getEgT2gOccupation(nBaths, p) for p in psis]), T=T)
    print("<E-E0> = {:4.3f}".format(thermal_average(e, e, T=T)))
    print("<N(3d)> = {:4.3f}".format(thermal_average(e, [getTraceDensityMatrix(nBaths, p) for p in psis], T=T)))
    print("<N(egDn)> = {:4.3f}".format(occs[0]))
    print("<N(egUp)> = {:4.3f}".format(occs[1]))
    print("<N(t2gDn)> = {:4.3f}".format(occs[2]))
    print("<N(t2gUp)> = {:4.3f}".format(occs[3]))
    print("<Lz(3d)> = {:4.3f}".format(thermal_average(e, [getLz3d(nBaths, p) for p in psis], T=T)))
    print("<Sz(3d)> = {:4.3f}".format(thermal_average(e, [getSz3d(nBaths, p) for p in psis], T=T)))
    print("<L^2(3d)> = {:4.3f}".format(thermal_average(e, [getLsqr3d(nBaths, p) for p in psis], T=T)))
    print("<S^2(3d)> = {:4.3f}".format(thermal_average(e, [getSsqr3d(nBaths, p) for p in psis], T=T)))
