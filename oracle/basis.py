"""Restricted initial basis and spin-orbital indexing (oracle).

Restates ``c2i`` (impurityModel/ed/finite.py:731-780) for integer angular
momenta and ``get_basis`` (finite.py:294-374).
"""

import itertools

from .states import tuple2int


def c2i(nBaths, spinOrb):
    """Index of (l, s, m) or bath (l, b) (finite.py:731-780)."""
    i = 0
    for lp in nBaths.keys():
        for sp in range(2):
            for mp in range(-lp, lp + 1):
                if (lp, sp, mp) == spinOrb:
                    return i
                i += 1
    for lp, nb in nBaths.items():
        for b in range(nb):
            if (lp, b) == spinOrb:
                return i
            i += 1
    raise Exception("Can not find index corresponding to spin-orbital state")


def n_spin_orbitals(nBaths):
    """finite.py:363 / get_spectra.py:175"""
    return sum(2 * (2 * l + 1) + nb for l, nb in nBaths.items())


def get_basis(nBaths, valBaths, dnValBaths, dnConBaths, dnTol, n0imp):
    """Restricted basis as int states in itertools order (finite.py:294-374).

    For each shell l: loops dnVal, dnCon with |dnVal-dnCon| <= dnTol, takes
    combinations of impurity, valence and conduction indices, nested in that
    order; the shells are combined with ``itertools.product`` and each
    configuration's sorted tuple becomes one state.  The list is not sorted.
    """
    for l in nBaths.keys():
        assert valBaths[l] <= nBaths[l]
    basisL = {}
    for l in nBaths.keys():
        basisL[l] = []
        for dnVal in range(dnValBaths[l] + 1):
            for dnCon in range(dnConBaths[l] + 1):
                dn = dnVal - dnCon
                if abs(dn) <= dnTol[l]:
                    nImp = n0imp[l] + dn
                    nVal = valBaths[l] - dnVal
                    nCon = dnCon
                    assert nVal <= valBaths[l]
                    assert nCon <= nBaths[l] - valBaths[l]
                    assert nImp <= 2 * (2 * l + 1)
                    imp = [c2i(nBaths, (l, s, m)) for s in range(2) for m in range(-l, l + 1)]
                    bImp = tuple(itertools.combinations(imp, nImp))
                    if valBaths[l] == 0:
                        bVal = ((),)
                    else:
                        val = [c2i(nBaths, (l, b)) for b in range(valBaths[l])]
                        bVal = tuple(itertools.combinations(val, nVal))
                    if nBaths[l] - valBaths[l] == 0:
                        bCon = ((),)
                    else:
                        con = [c2i(nBaths, (l, b)) for b in range(valBaths[l], nBaths[l])]
                        bCon = tuple(itertools.combinations(con, nCon))
                    for a in bImp:
                        for b in bVal:
                            for c in bCon:
                                basisL[l].append(a + b + c)
    n = n_spin_orbitals(nBaths)
    basis = []
    for conf in itertools.product(*basisL.values()):
        basis.append(tuple2int(tuple(sorted(itertools.chain.from_iterable(conf))), n))
    return tuple(basis)


def get_restrictions(nBaths, nValBaths, n0imps, dnTols, dnValBaths, dnConBaths, l=2):
    """Excited-state occupation restrictions (get_spectra.py:154-169)."""
    r = {}
    idx = frozenset(c2i(nBaths, (l, s, m)) for s in range(2) for m in range(-l, l + 1))
    r[idx] = (n0imps[l] - 1, n0imps[l] + dnTols[l] + 1)
    idx = frozenset(c2i(nBaths, (l, b)) for b in range(nValBaths[l]))
    r[idx] = (nValBaths[l] - dnValBaths[l], nValBaths[l])
    idx = frozenset(c2i(nBaths, (l, b)) for b in range(nValBaths[l], nBaths[l]))
    r[idx] = (0, dnConBaths[l])
    return r
