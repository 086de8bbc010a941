"""Second-quantized operator builders (host side).

Operators are dicts ``{process: value}`` where a process is a tuple of
``(index, 'c')`` / ``(index, 'a')`` steps applied right to left, exactly the
format of impurityModel (impurityModel/ed/finite.py:1301-1340).  These
builders produce the small host tables that the HIP engine compiles; they are
setup code, not the hot path.

Angular algebra is computed here from closed forms: Wigner 3j symbols by the
Racah formula in exact rational arithmetic, which gives the Gaunt
coefficients that the reference takes from ``sympy.physics.wigner.gaunt``
(finite.py:476-482).
"""

from fractions import Fraction
from functools import lru_cache
from math import factorial, pi, sqrt

import numpy as np

# ---------------------------------------------------------------- angular algebra


def _wigner3j_sq_signed(j1, j2, j3, m1, m2, m3):
    """Return ``(sign, square)`` of the Wigner 3j symbol for integer j, m.

    Racah formula: (-1)^(j1-j2-m3) sqrt(D * prod (j+-m)!) * sum_t (-1)^t / (...)
    with D the triangle coefficient.  The square is an exact ``Fraction``.
    """
    if m1 + m2 + m3 != 0:
        return 0, Fraction(0)
    if j3 < abs(j1 - j2) or j3 > j1 + j2:
        return 0, Fraction(0)
    if abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0, Fraction(0)
    f = factorial
    tri = Fraction(f(j1 + j2 - j3) * f(j1 - j2 + j3) * f(-j1 + j2 + j3), f(j1 + j2 + j3 + 1))
    pref = tri * f(j1 + m1) * f(j1 - m1) * f(j2 + m2) * f(j2 - m2) * f(j3 + m3) * f(j3 - m3)
    tmin = max(0, j2 - j3 - m1, j1 - j3 + m2)
    tmax = min(j1 + j2 - j3, j1 - m1, j2 + m2)
    s = Fraction(0)
    for t in range(tmin, tmax + 1):
        den = f(t) * f(j3 - j2 + t + m1) * f(j3 - j1 + t - m2) * f(j1 + j2 - j3 - t) * f(j1 - t - m1) * f(j2 - t + m2)
        s += Fraction((-1) ** t, den)
    if s == 0:
        return 0, Fraction(0)
    sign = (-1) ** ((j1 - j2 - m3) % 2) * (1 if s > 0 else -1)
    return sign, pref * s * s


def wigner3j(j1, j2, j3, m1, m2, m3):
    sign, sq = _wigner3j_sq_signed(j1, j2, j3, m1, m2, m3)
    return 0.0 if sign == 0 else sign * sqrt(sq)


@lru_cache(maxsize=None)
def gauntC(k, l, m, lp, mp, prec=16):
    """Coulomb Gaunt coefficient c^k(l m, l' m') (finite.py:476-482).

    The reference evaluates sqrt(4 pi/(2k+1)) (-1)^m gaunt(l, k, l', -m, m-m', m').
    With gaunt = sqrt((2l+1)(2k+1)(2l'+1)/(4 pi)) 3j(l k l'; 0 0 0) 3j(l k l'; -m, m-m', m')
    the 4 pi cancels, so c^2 is the exact rational
    (2l+1)(2l'+1) 3j0^2 3jm^2.  ``prec`` is accepted for signature parity.
    """
    s0, q0 = _wigner3j_sq_signed(l, k, lp, 0, 0, 0)
    s1, q1 = _wigner3j_sq_signed(l, k, lp, -m, m - mp, mp)
    if s0 == 0 or s1 == 0:
        return 0.0
    sign = s0 * s1 * (-1) ** (m % 2)
    return float(sign * sqrt(Fraction((2 * l + 1) * (2 * lp + 1)) * q0 * q1))


# ---------------------------------------------------------------- index maps


def c2i(nBaths, spinOrb):
    """Index of spin-orbital (l, s, m) or bath state (l, b) (finite.py:731-780).

    Impurity spin-orbitals come first in ``nBaths`` key order, spin-major,
    m ascending; bath states follow in the same key order.
    """
    i = 0
    for lp in nBaths.keys():
        lps = lp if isinstance(lp, tuple) else (lp,)
        for lpi in lps:
            for sp in range(2):
                for mp in range(-lpi, lpi + 1):
                    if (lpi, sp, mp) == spinOrb:
                        return i
                    i += 1
    for lp, nb in nBaths.items():
        for b in range(nb):
            if (lp, b) == spinOrb:
                return i
            i += 1
    raise Exception("Can not find index corresponding to spin-orbital state")


def i2c(nBaths, i):
    """Inverse of :func:`c2i` (finite.py:783-840)."""
    k = 0
    for lp in nBaths.keys():
        lps = lp if isinstance(lp, tuple) else (lp,)
        for lpi in lps:
            if i - k < 2 * (2 * lpi + 1):
                for sp in range(2):
                    for mp in range(-lpi, lpi + 1):
                        if k == i:
                            return (lpi, sp, mp)
                        k += 1
            k += 2 * (2 * lpi + 1)
    for lp, nb in nBaths.items():
        b = i - k
        if b < nb:
            return (lp, b)
        k += nb
    raise Exception("Can not find spin-orbital state corresponding to index.")


@lru_cache(maxsize=64)
def _index_map(nbaths_items):
    """{spin-orbital tuple: c2i index} for every spin-orbital of a layout."""
    nBaths = dict(nbaths_items)
    n = n_spin_orbitals(nBaths)
    return {i2c(nBaths, i): i for i in range(n)}


def n_spin_orbitals(nBaths):
    """Total spin-orbital count (get_spectra.py:175)."""
    return sum(2 * (2 * ang + 1) + nb for ang, nb in nBaths.items())


# ---------------------------------------------------------------- operator algebra


def addOps(ops):
    """Sum operators, dropping zero values (finite.py:629-651)."""
    out = {}
    for op in ops:
        for proc, value in op.items():
            if value != 0:
                out[proc] = out[proc] + value if proc in out else value
    return out


def daggerOp(op):
    """Hermitian conjugate (finite.py:271-287)."""
    out = {}
    for proc, value in op.items():
        new = []
        for e in proc[::-1]:
            if e[1] == "a":
                new.append((e[0], "c"))
            elif e[1] == "c":
                new.append((e[0], "a"))
            else:
                raise Exception("Operator type unknown: {}".format(e[1]))
        out[tuple(new)] = value.conjugate()
    return out


def assert_hermitian(op):
    """finite.py:290-291"""
    assert daggerOp(op) == op


# ---------------------------------------------------------------- Hamiltonian pieces


def getU(l1, m1, l2, m2, l3, m3, l4, m4, R):
    """u = sum_k R_k c^k(l1 m1, l4 m4) c^k(l3 m3, l2 m2) if m1+m2 = m3+m4 (finite.py:485-534)."""
    if m1 + m2 != m3 + m4:
        return 0
    u = 0
    for k, Rk in enumerate(R):
        u += Rk * gauntC(k, l1, m1, l4, m4) * gauntC(k, l3, m3, l2, m2)
    return u


def getUop(l1, l2, l3, l4, R):
    """Spin-polarized Slater-Condon U operator with value u/2 (finite.py:586-626)."""
    uDict = {}
    for m1 in range(-l1, l1 + 1):
        for m2 in range(-l2, l2 + 1):
            for m3 in range(-l3, l3 + 1):
                for m4 in range(-l4, l4 + 1):
                    u = getU(l1, m1, l2, m2, l3, m3, l4, m4, R)
                    if u != 0:
                        for s in range(2):
                            for sp in range(2):
                                proc = (((l1, s, m1), "c"), ((l2, sp, m2), "c"), ((l3, sp, m3), "a"), ((l4, s, m4), "a"))
                                if not (s == sp and ((l1, m1) == (l2, m2) or (l3, m3) == (l4, m4))):
                                    uDict[proc] = u / 2.0
    return uDict


def get2p3dSlaterCondonUop(Fdd=(9, 0, 8, 0, 6), Fpp=(20, 0, 8), Fpd=(10, 0, 8), Gpd=(0, 3, 0, 2)):
    """Fdd + Fpp + Fpd + Gpd operators summed (finite.py:654-681)."""
    FddOp = getUop(2, 2, 2, 2, Fdd)
    FppOp = getUop(1, 1, 1, 1, Fpp)
    FpdOp = addOps([getUop(1, 2, 2, 1, Fpd), getUop(2, 1, 1, 2, Fpd)])
    GpdOp = addOps([getUop(1, 2, 1, 2, Gpd), getUop(2, 1, 2, 1, Gpd)])
    return addOps([FddOp, FppOp, FpdOp, GpdOp])


def getSOCop(xi, l=2):
    """Spin-orbit coupling xi l.s of one shell (finite.py:684-705)."""
    op = {}
    for m in range(-l, l + 1):
        for s in range(2):
            op[(((l, s, m), "c"), ((l, s, m), "a"))] = xi * m * (1 / 2.0 if s == 1 else -1 / 2.0)
    for m in range(-l, l):
        value = xi / 2.0 * sqrt((l - m) * (l + m + 1))
        op[(((l, 1, m), "c"), ((l, 0, m + 1), "a"))] = value
        op[(((l, 0, m + 1), "c"), ((l, 1, m), "a"))] = value
    return op


def gethHfieldop(hx, hy, hz, l=2):
    """Zeeman term of one shell (finite.py:708-728)."""
    op = {}
    for m in range(-l, l + 1):
        op[(((l, 1, m), "c"), ((l, 0, m), "a"))] = hx / 2
        op[(((l, 0, m), "c"), ((l, 1, m), "a"))] = hx / 2
        op[(((l, 1, m), "c"), ((l, 0, m), "a"))] += -hy * 1j / 2
        op[(((l, 0, m), "c"), ((l, 1, m), "a"))] += hy * 1j / 2
        for s in range(2):
            op[(((l, s, m), "c"), ((l, s, m), "a"))] = hz / 2 if s == 1 else -hz / 2
    return op


def dc_MLFT(n3d_i, c, Fdd, n2p_i=None, Fpd=None, Gpd=None):
    """MLFT double counting (finite.py:177-217)."""
    if not int(n3d_i) == n3d_i:
        raise ValueError("3d occupation should be an integer")
    if n2p_i is not None and int(n2p_i) != n2p_i:
        raise ValueError("2p occupation should be an integer")
    Udd = Fdd[0] - 14.0 / 441 * (Fdd[2] + Fdd[4])
    if n2p_i is None and Fpd is None and Gpd is None:
        return Udd * n3d_i - c
    if n2p_i == 6 and Fpd is not None and Gpd is not None:
        Upd = Fpd[0] - (1 / 15.0) * Gpd[1] - (3 / 70.0) * Gpd[3]
        return [Udd * n3d_i + Upd * n2p_i - c, Upd * (n3d_i + 1) - c]
    raise ValueError("double counting input wrong.")


def get_spherical_2_cubic_matrix(spinpol=False, l=2):
    """Unitary spherical -> cubic harmonics matrix (finite.py:220-268)."""
    r2 = 1 / np.sqrt(2)
    if l == 1:
        u = np.zeros((3, 3), dtype=complex)
        u[0, 0] = 1j * r2
        u[2, 0] = 1j * r2
        u[0, 1] = r2
        u[2, 1] = -r2
        u[1, 2] = 1
    elif l == 2:
        u = np.zeros((5, 5), dtype=complex)
        u[2, 0] = 1
        u[[0, -1], 1] = r2
        u[1, 2] = -1j * r2
        u[-2, 2] = -1j * r2
        u[1, 3] = r2
        u[-2, 3] = -r2
        u[0, 4] = 1j * r2
        u[-1, 4] = -1j * r2
    else:
        raise ValueError("l must be 1 or 2")
    if spinpol:
        n, m = np.shape(u)
        U = np.zeros((2 * n, 2 * m), dtype=complex)
        U[0:n, 0:m] = u
        U[n:, m:] = u
        u = U
    return u


def get_hamiltonian_operator_from_h0(nBaths, slaterCondon, SOCs, DCinfo, hField, h0_operator):
    """hOp = U + hField + SOC(2p) + SOC(3d) - DC + h0, integer-indexed.

    Same sum and order as get_spectra.py:335-418, with the non-interacting
    part passed in as an operator dict keyed by (l, s, m)/(l, b) tuples.
    """
    Fdd, Fpp, Fpd, Gpd = slaterCondon
    xi_2p, xi_3d = SOCs
    n0imps, ctc = DCinfo
    hx, hy, hz = hField
    uOp = get2p3dSlaterCondonUop(Fdd=Fdd, Fpp=Fpp, Fpd=Fpd, Gpd=Gpd)
    soc2p = getSOCop(xi_2p, l=1)
    soc3d = getSOCop(xi_3d, l=2)
    dc = dc_MLFT(n3d_i=n0imps[2], c=ctc, Fdd=Fdd, n2p_i=n0imps[1], Fpd=Fpd, Gpd=Gpd)
    eDC = {}
    for il, l in enumerate([2, 1]):
        for s in range(2):
            for m in range(-l, l + 1):
                eDC[(((l, s, m), "c"), ((l, s, m), "a"))] = -dc[il]
    hH = gethHfieldop(hx, hy, hz, l=2)
    hsum = addOps([uOp, hH, soc2p, soc3d, eDC, h0_operator])
    index = _index_map(tuple(nBaths.items()))
    hOp = {}
    for proc, value in hsum.items():
        hOp[tuple((index[so], action) for so, action in proc)] = value
    assert_hermitian(hOp)
    return hOp


def get_bath_h0(e_imp, tenDq, eps_bath, v_bath, l=2):
    """Non-interacting 3d + bath operator of a crystal-field model.

    Impurity: e_eg = e_imp + 3/5 10Dq, e_t2g = e_imp - 2/5 10Dq in the cubic
    basis, rotated to spherical harmonics (get_spectra_using_CF.py:476-494).
    Bath state b couples to the spherical 3d spin-orbital number b mod 10
    (spin-major, m ascending) with on-site energy eps_bath[b] and hopping
    v_bath[b] -- the one-bath-state-per-spin-orbital layout of
    get_spectra_using_CF.py:517-546 generalized to several bath states per
    spin-orbital.
    """
    e_eg = e_imp + 3 / 5 * tenDq
    e_t2g = e_imp - 2 / 5 * tenDq
    h = np.zeros((2 * l + 1, 2 * l + 1))
    np.fill_diagonal(h, (e_eg, e_eg, e_t2g, e_t2g, e_t2g))
    u = get_spherical_2_cubic_matrix(spinpol=False, l=l)
    h = np.dot(u, np.dot(h, np.conj(u.T)))
    # the rotation can leave h[i, j] and conj(h[j, i]) one ulp apart
    h = 0.5 * (h + np.conj(h.T))
    op = {}
    for i, mi in enumerate(range(-l, l + 1)):
        for j, mj in enumerate(range(-l, l + 1)):
            if h[i, j] != 0:
                for s in range(2):
                    op[(((l, s, mi), "c"), ((l, s, mj), "a"))] = h[i, j]
    nso = 2 * (2 * l + 1)
    for b in range(len(eps_bath)):
        k = b % nso
        s, m = divmod(k, 2 * l + 1)
        m -= l
        v = complex(v_bath[b])
        if eps_bath[b] != 0:
            op[(((l, b), "c"), ((l, b), "a"))] = complex(eps_bath[b])
        if v != 0:
            op[(((l, b), "c"), ((l, s, m), "a"))] = v
            op[(((l, s, m), "c"), ((l, b), "a"))] = v.conjugate()
    return op


def get_CF_hamiltonian(nBaths, nValBaths, cf, bath_state_basis="spherical"):
    """Non-interacting 3d + bath operator from crystal-field parameters
    (get_spectra_using_CF.py:433-563).

    ``cf`` is the ten-tuple of ``readers.read_h0_CF_file`` (or a JSON file
    name).  In cubic harmonics every 3d orbital couples to one valence bath
    state (and, with 10 conduction bath states, one conduction state): on-site
    e_imp + 3/5 10Dq (eg) or e_imp - 2/5 10Dq (t2g), bath energies and
    hoppings per eg/t2g.  The impurity block is rotated to spherical
    harmonics; the bath states are spherical (``u_bath = u``) or kept cubic.
    Bath state of spin s, orbital i: valence s*5 + i, conduction 10 + s*5 + i.
    Returns the sum impurity + hopping + bath-energy operators, keyed by
    (l, s, m) / (l, b) spin-orbitals.
    """
    if isinstance(cf, str):
        from .readers import read_h0_CF_file

        cf = read_h0_CF_file(cf)
    (e_imp, tenDq, e_val_eg, e_val_t2g, e_con_eg, e_con_t2g, v_val_eg, v_val_t2g, v_con_eg, v_con_t2g) = cf
    l = 2
    nl = 2 * l + 1

    def cubic(eg, t2g):
        return np.diag([eg, eg, t2g, t2g, t2g]).astype(float)

    u = get_spherical_2_cubic_matrix(spinpol=False, l=l)
    if bath_state_basis == "spherical":
        u_bath = u
    elif bath_state_basis == "cubic":
        u_bath = np.eye(nl)
    else:
        raise Exception("Design of this basis is not (yet) implemented.")
    def herm(h):
        # u D u^dagger is Hermitian up to the rounding of the products; the
        # mean with its adjoint (a change of at most one ulp) makes the
        # operator pass the exact-equality hermiticity check of
        # get_hamiltonian_operator whatever order the products sum in
        return 0.5 * (h + np.conj(h.T))

    h_imp = herm(u @ cubic(e_imp + 3 / 5 * tenDq, e_imp - 2 / 5 * tenDq) @ np.conj(u.T))
    v_val = u_bath @ cubic(v_val_eg, v_val_t2g) @ np.conj(u.T)
    v_con = u_bath @ cubic(v_con_eg, v_con_t2g) @ np.conj(u.T)
    e_val = herm(u_bath @ cubic(e_val_eg, e_val_t2g) @ np.conj(u_bath.T))
    e_con = herm(u_bath @ cubic(e_con_eg, e_con_t2g) @ np.conj(u_bath.T))
    with_con = nBaths[l] - nValBaths[l] == 10
    ms = list(range(-l, l + 1))
    # impurity block, spin by spin (i, j over m ascending)
    h_imp_op = {}
    for i, mi in enumerate(ms):
        for j, mj in enumerate(ms):
            if h_imp[i, j] != 0:
                for s in range(2):
                    h_imp_op[(((l, s, mi), "c"), ((l, s, mj), "a"))] = h_imp[i, j]
    # hoppings and bath energies in the reference's insertion order:
    # spin, then bath row i, then impurity column j; conduction terms follow
    # the valence terms of the same (s, i, j)
    hopp, ebath = {}, {}
    for s in range(2):
        for i in range(nl):
            bi_val = s * nl + i
            bi_con = 2 * nl + bi_val
            for j, mj in enumerate(ms):
                bj_val = s * nl + j
                bj_con = 2 * nl + bj_val
                blocks = [(v_val, e_val, bi_val, bj_val)]
                if with_con:
                    blocks.append((v_con, e_con, bi_con, bj_con))
                for vmat, emat, bi, bj in blocks:
                    v = vmat[i, j]
                    if v != 0:
                        hopp[(((l, bi), "c"), ((l, s, mj), "a"))] = v
                        hopp[(((l, s, mj), "c"), ((l, bi), "a"))] = v.conjugate()
                    if emat[i, j] != 0:
                        ebath[(((l, bi), "c"), ((l, bj), "a"))] = emat[i, j]
    return addOps([h_imp_op, hopp, ebath])


# ---------------------------------------------------------------- transition operators


def getDipoleOperator(nBaths, n):
    """Dipole 2p -> 3d operator for polarization n (spectra.py:341-382)."""
    t = {}
    nd = {-1: (n[0] + 1j * n[1]) / sqrt(2), 0: n[2], 1: (-n[0] + 1j * n[1]) / sqrt(2)}
    l1, l2 = nBaths.keys()
    for m in range(-l2, l2 + 1):
        for mp in range(-l1, l1 + 1):
            for s in range(2):
                if abs(m - mp) <= 1:
                    tij = gauntC(1, l2, m, l1, mp, prec=16)
                    tij *= nd[m - mp]
                    if tij != 0:
                        i = c2i(nBaths, (l2, s, m))
                        j = c2i(nBaths, (l1, s, mp))
                        t[((i, "c"), (j, "a"))] = tij
    return t


def getDipoleOperators(nBaths, ns):
    """spectra.py:302-320"""
    return [getDipoleOperator(nBaths, n) for n in ns]


def getDaggeredDipoleOperators(nBaths, ns):
    """spectra.py:323-338"""
    return [daggerOp(getDipoleOperator(nBaths, n)) for n in ns]


def getInversePhotoEmissionOperators(nBaths, l=2):
    """{c_i^dagger} of one shell (spectra.py:511-528)."""
    return [{((c2i(nBaths, (l, s, m)), "c"),): 1} for s in range(2) for m in range(-l, l + 1)]


def getPhotoEmissionOperators(nBaths, l=2):
    """{c_i} of one shell (spectra.py:531-548)."""
    return [{((c2i(nBaths, (l, s, m)), "a"),): 1} for s in range(2) for m in range(-l, l + 1)]


def radial_integral(k, qNorm, Ri, Rj, r):
    """int conj(Ri) j_k(|q| r) Rj r^2 dr by the trapezoid rule (spectra.py:488)."""
    from scipy.special import spherical_jn

    f = np.conj(Ri) * spherical_jn(k, qNorm * r) * Rj * r**2
    d = np.diff(r)
    return (d * (f[1:] + f[:-1]) / 2.0).sum()


def _sph_harm(m, k, phi, theta):
    """Y_k^m(theta, phi) with the reference's argument order (spectra.py:497)."""
    import scipy.special as sp

    if hasattr(sp, "sph_harm_y"):
        return sp.sph_harm_y(k, m, theta, phi)
    return sp.sph_harm(m, k, phi, theta)


def getNIXSOperator(nBaths, q, li, lj, Ri, Rj, r, kmin=1):
    """NIXS transition operator (spectra.py:437-508).

    T = sum_k R_k i^k sqrt(2k+1) conj(Y_k^m(q)) c^k(li mi, lj mj), summed over
    k in [kmin, li+lj] with li+lj+k even.  The azimuth uses the reference's
    arccos(qx/(|q| sin theta)), so phi lies in [0, pi].
    """
    q = np.array(q, dtype=float)
    qNorm = np.linalg.norm(q)
    theta = np.arccos(q[2] / qNorm)
    phi = np.arccos(q[0] / (qNorm * np.sin(theta)))
    t = {}
    for k in range(kmin, abs(li + lj) + 1):
        if (li + lj + k) % 2 == 0:
            Rk = radial_integral(k, qNorm, Ri, Rj, r)
            for mi in range(-li, li + 1):
                for mj in range(-lj, lj + 1):
                    m = mi - mj
                    if abs(m) <= k:
                        tij = Rk
                        tij *= 1j**k * sqrt(2 * k + 1)
                        tij *= np.conj(_sph_harm(m, k, phi, theta))
                        tij *= gauntC(k, li, mi, lj, mj, prec=16)
                        if tij != 0:
                            for s in range(2):
                                i = c2i(nBaths, (li, s, mi))
                                j = c2i(nBaths, (lj, s, mj))
                                proc = ((i, "c"), (j, "a"))
                                t[proc] = t[proc] + tij if proc in t else tij
    return t


def getNIXSOperators(nBaths, qs, li, lj, Ri, Rj, r, kmin=1):
    """spectra.py:385-434"""
    return [getNIXSOperator(nBaths, q, li, lj, Ri, Rj, r, kmin) for q in qs]
