"""Operator application on multi-configurational states (oracle).

Restates ``applyOp`` (impurityModel/ed/finite.py:1301-1511) and the dict
vector algebra (finite.py:377-473, 1998-2033) on int-keyed dicts.
"""

from .states import create, remove


def restriction_masks(n, restrictions):
    """Turn ``{frozenset(indices): (lo, hi)}`` into ``[(mask, lo, hi)]``.

    ``len(restriction & occupied)`` (finite.py:1382-1385) equals the popcount
    of ``state & mask`` in the integer form.
    """
    if restrictions is None:
        return None
    out = []
    for idx, (lo, hi) in restrictions.items():
        m = 0
        for i in idx:
            m |= 1 << (n - 1 - i)
        out.append((m, lo, hi))
    return out


def _passes(state, rmasks):
    for m, lo, hi in rmasks:
        k = bin(state & m).count("1")
        if k < lo or hi < k:
            return False
    return True


def _apply_process(n, process, state):
    """Apply one process tuple right to left (finite.py:1365-1373).

    Returns ``(target, sign)`` or ``(None, 0)`` when Pauli-blocked.
    """
    s = state
    sign_tot = 1
    for i, action in process[::-1]:
        if action == "a":
            s, sign = remove(n, i, s)
        elif action == "c":
            s, sign = create(n, i, s)
        else:
            raise Exception("Operator type unknown: {}".format(action))
        if sign == 0:
            return None, 0
        sign_tot *= sign
    return s, sign_tot


def apply_op(n, op, psi, slaterWeightMin=1e-12, restrictions=None, opResult=None):
    """Return op|psi> as an int-keyed dict (finite.py:1301-1511).

    All four branches of the reference are kept: with/without restrictions
    and with/without the ``opResult`` memo.  Restrictions are checked when a
    target is first inserted; the memo entries are pruned per source state at
    ``|h*sign|^2 < slaterWeightMin`` after they are built (finite.py:1459-1464,
    1499-1504); the result is pruned at the end (finite.py:1507-1510).
    """
    rm = restriction_masks(n, restrictions)
    psi_new = {}
    if opResult is None:
        for state, amp in psi.items():
            for process, h in op.items():
                t, sgn = _apply_process(n, process, state)
                if t is None:
                    continue
                if t in psi_new:
                    psi_new[t] += amp * h * sgn
                elif rm is None or _passes(t, rm):
                    psi_new[t] = amp * h * sgn
    else:
        for state, amp in psi.items():
            if state in opResult:
                for s, a in opResult[state].items():
                    if s in psi_new:
                        psi_new[s] += a * amp
                    else:
                        psi_new[s] = a * amp
                continue
            memo = {}
            opResult[state] = memo
            for process, h in op.items():
                t, sgn = _apply_process(n, process, state)
                if t is None:
                    continue
                if rm is None:
                    if t in memo:
                        memo[t] += h * sgn
                    else:
                        memo[t] = h * sgn
                    if t in psi_new:
                        psi_new[t] += amp * h * sgn
                    else:
                        psi_new[t] = amp * h * sgn
                else:
                    if t in psi_new:
                        psi_new[t] += amp * h * sgn
                        if t in memo:
                            memo[t] += h * sgn
                        else:
                            memo[t] = h * sgn
                    elif _passes(t, rm):
                        psi_new[t] = amp * h * sgn
                        memo[t] = h * sgn
            for ps, a in list(memo.items()):
                if abs(a) ** 2 < slaterWeightMin:
                    memo.pop(ps)
    for state, a in list(psi_new.items()):
        if abs(a) ** 2 < slaterWeightMin:
            psi_new.pop(state)
    return psi_new


def inner(a, b):
    """<a|b> (finite.py:377-394)"""
    acc = 0
    for s, amp in b.items():
        if s in a:
            acc += a[s].conjugate() * amp
    return acc


def add(psi1, psi2, mul=1):
    """psi1 + mul*psi2 (finite.py:1998-2020)"""
    psi = dict(psi1)
    for s, a in psi2.items():
        psi[s] = psi[s] + mul * a if s in psi else mul * a
    return psi


def norm2(psi):
    """<psi|psi> (finite.py:2023-2033)"""
    return sum(abs(a) ** 2 for a in psi.values())
