"""Green's functions, spectra and thermal averages (oracle).

Restates ``getGreen`` (impurityModel/ed/spectra.py:551-681, sparse branch),
``getSpectra`` (spectra.py:684-806, one-rank ``H_build`` branch 778-803),
``getRIXSmap`` (spectra.py:809-1130, ``wIn`` branch 1012-1129), the
XAS/PS/XPS/RIXS composition of ``simulate_spectra`` (spectra.py:124-293) and
``thermal_average`` (impurityModel/ed/average.py:11-45).
"""

from math import sqrt

import numpy as np

from .applyop import apply_op, norm2
from .hamiltonian import expand_basis_and_hamiltonian
from .krylov import continued_fraction, lanczos

# scipy.constants "Boltzmann constant in eV/K" (average.py:8), exact SI value.
k_B = 1.380649e-23 / 1.602176634e-19


def thermal_average(energies, observable, T=300):
    """Boltzmann-weighted average over axis 0 (average.py:11-45)."""
    if len(energies) != np.shape(observable)[0]:
        raise ValueError("Passed array is not of the right shape")
    z = 0
    o_average = 0
    shift = np.min(energies)
    for e, o in zip(energies, observable):
        weight = np.exp(-(e - shift) / (k_B * T))
        z += weight
        o_average += weight * o
    return o_average / z


def get_green(n, e, psi, hOp, omega, delta, krylovSize, slaterWeightMin, restrictions=None, h_dict=None,
              reorth="none", return_ab=False, sizes=None):
    """<psi|(omega + i delta + e - H)^-1|psi> by Lanczos (spectra.py:551-681).

    ``sizes`` (a list) receives the BFS basis size, as the reference prints
    it (finite.py:1879-1886)."""
    g = np.zeros(len(omega), dtype=complex)
    if len(psi) == 0:
        return (g, None, None) if return_ab else g
    if h_dict is None:
        h_dict = {}
    h, index = expand_basis_and_hamiltonian(n, h_dict, hOp, psi.keys(), restrictions)
    nb = len(index)
    if sizes is not None:
        sizes.append(nb)
    psi0 = np.zeros(nb, dtype=complex)
    for ps, amp in psi.items():
        psi0[index[ps]] = amp
    K = min(krylovSize, nb)
    alpha, beta = lanczos(h, psi0, K, reorth=reorth)
    g = continued_fraction(alpha, beta, omega, delta, e)
    return (g, alpha, beta) if return_ab else g


def get_spectra(n, hOp, tOps, psis, es, w, delta, restrictions=None, krylovSize=150, slaterWeightMin=1e-7,
                reorth="none", sizes=None):
    """gs[i, t, :] = |T_t psi_i|^2 G_i(normalized T_t psi_i) (spectra.py:684-806).

    One ``h_dict`` memo is shared by all chains; each transition operator has
    its own ``t_big`` memo shared over eigenstates (spectra.py:745, 780-786).
    """
    ne = len(es)
    gs = np.zeros((ne, len(tOps), len(w)), dtype=complex)
    h = {}
    for t, tOp in enumerate(tOps):
        t_big = {}
        for i in range(ne):
            psiR = apply_op(n, tOp, psis[i], slaterWeightMin, restrictions, t_big)
            normalization = sqrt(norm2(psiR))
            for s in psiR.keys():
                psiR[s] /= normalization
            gs[i, t, :] = normalization**2 * get_green(
                n, es[i], psiR, hOp, w, delta, krylovSize, slaterWeightMin, restrictions, h, reorth=reorth,
                sizes=sizes)
    return gs


def rixs_map(n, hOp, tOpsIn, tOpsOut, psis, es, wIns, wLoss, delta1, delta2, restrictions=None, krylovSize=150,
             slaterWeightMin=1e-7, alias_memo=False, reorth="none"):
    """RIXS Green's functions gs[iE, tIn, tOut, iwIn, :] (spectra.py:809-1130).

    Per (tIn, eigenstate): psi1 = T_in psi (memo per tIn, 1012-1021), the
    core-hole sector H by BFS expansion from psi1's keys (one h_dict shared
    by all, 1024-1045), y = psi1 as a vector (1051-1054); per incoming
    energy x = (wIn + i delta1 + e - H)^-1 y (1066-1080) and psi2 = x as a
    dict over the nonzero entries in basis order (1087-1090); per outgoing
    operator psi3 = T_out psi2, normalized, pruned at slaterWeightMin
    (1094-1111), and norm^2 * getGreen(e, psi3, wLoss, delta2) (1113-1125).

    The reference solves with scipy's ``bicgstab`` at its default tolerance
    (relative residual 1e-5), an approximation of the exact solution; this
    restatement solves exactly (sparse LU), which is what a tight-tolerance
    solver converges to.  ``alias_memo`` reproduces the reference's single
    memo shared by all out-going operators (``[{}] * len(tOpsOut)``,
    spectra.py:910): the first operator fills it for every state it meets,
    so later operators reuse the first one's T_out|state>.
    """
    import scipy.sparse
    import scipy.sparse.linalg

    ne = len(es)
    gs = np.zeros((ne, len(tOpsIn), len(tOpsOut), len(wIns), len(wLoss)), dtype=complex)
    h_exc = {}
    h_ground = {}
    shared = {}
    t_out_memo = [shared] * len(tOpsOut) if alias_memo else [{} for _ in tOpsOut]
    for tIn, tOpIn in enumerate(tOpsIn):
        t_in_memo = {}
        for iE in range(ne):
            e = es[iE]
            psi1 = apply_op(n, tOpIn, psis[iE], slaterWeightMin, restrictions, t_in_memo)
            h, index = expand_basis_and_hamiltonian(n, h_exc, hOp, psi1.keys(), restrictions)
            nb = len(index)
            y = np.zeros(nb, dtype=complex)
            for ps, amp in psi1.items():
                y[index[ps]] = amp
            states = [None] * nb
            for ps, i in index.items():
                states[i] = ps
            for iw, wIn in enumerate(wIns):
                a = scipy.sparse.identity(nb, dtype=complex, format="csc") * (wIn + 1j * delta1 + e) - h.tocsc()
                x = scipy.sparse.linalg.spsolve(a, y) if nb else np.zeros(0, complex)
                x = np.atleast_1d(x)
                psi2 = {states[i]: amp for i, amp in enumerate(x) if amp != 0}
                for tOut, tOpOut in enumerate(tOpsOut):
                    psi3 = apply_op(n, tOpOut, psi2, slaterWeightMin, restrictions, t_out_memo[tOut])
                    normalization = sqrt(norm2(psi3))
                    for s in psi3.keys():
                        psi3[s] /= normalization
                    for s, amp in list(psi3.items()):
                        if abs(amp) ** 2 < slaterWeightMin:
                            psi3.pop(s)
                    gs[iE, tIn, tOut, iw, :] = normalization**2 * get_green(
                        n, e, psi3, hOp, wLoss, delta2, krylovSize, slaterWeightMin, restrictions, h_ground,
                        reorth=reorth)
    return gs


def rixs(n, hOp, tOpsIn, tOpsOut, psis, es, wIns, wLoss, delta1, delta2, restrictions, T=300, krylovSize=150,
         alias_memo=False, reorth="none"):
    """RIXS branch of simulate_spectra (spectra.py:251-293).

    Returns ``(gs, a, aSum)``: the Green's functions, the thermal average of
    -Im gs over eigenstates and its sum over in/out operators, [nwIn, nwLoss].
    """
    gs = rixs_map(n, hOp, tOpsIn, tOpsOut, psis, es, wIns, wLoss, delta1, delta2, restrictions, krylovSize,
                  alias_memo=alias_memo, reorth=reorth)
    a = thermal_average(es[: np.shape(gs)[0]], -gs.imag, T=T)
    return gs, a, np.sum(a, axis=(0, 1))


def rixs_bin(wIns, wLoss, aSum):
    """The float32 RIXS.bin table (spectra.py:287-293): [0,0] = len(wIn),
    first row wIn, first column wLoss, body aSum transposed to [wLoss, wIn]."""
    tmp = np.zeros((len(wLoss) + 1, len(wIns) + 1), dtype=np.float32)
    tmp[0, 0] = len(wIns)
    tmp[0, 1:] = wIns
    tmp[1:, 0] = wLoss
    tmp[1:, 1:] = np.asarray(aSum).T
    return tmp


def energy_cut(es, T=300, cut=10.0):
    """Keep states with e - e0 < cut*k_B*T (get_spectra.py:134, 302-303)."""
    lim = cut * k_B * T
    return tuple(e for e in es if e - es[0] < lim)


def xas(n, hOp, tOps, psis, es, w, delta, restrictions, T=300, krylovSize=150, reorth="none", sizes=None):
    """XAS branch of simulate_spectra (spectra.py:222-238).

    Returns ``(gs, a, aSum)``: raw Green's functions, thermally averaged
    -Im per polarization, and their sum.
    """
    gs = get_spectra(n, hOp, tOps, psis, es, w, delta, restrictions, krylovSize, reorth=reorth, sizes=sizes)
    a = thermal_average(es[: np.shape(gs)[0]], -gs.imag, T=T)
    return gs, a, np.sum(a, axis=0)


def photoemission(n, hOp, tOpsPS, tOpsIPS, psis, es, w, delta, restrictions, T=300, krylovSize=150,
                  reorth="none", sizes=None):
    """PS + IPS branch of simulate_spectra (spectra.py:124-149).

    ``tOpsIPS`` may be None for the core XPS branch (spectra.py:163-179).
    """
    if tOpsIPS is None:
        gs = -get_spectra(n, hOp, tOpsPS, psis, es, -w, -delta, restrictions, krylovSize, reorth=reorth, sizes=sizes)
    else:
        # the reference's order: inverse photoemission first (spectra.py:129-138)
        g_ips = get_spectra(n, hOp, tOpsIPS, psis, es, w, delta, restrictions, krylovSize, reorth=reorth, sizes=sizes)
        gs = -get_spectra(n, hOp, tOpsPS, psis, es, -w, -delta, restrictions, krylovSize, reorth=reorth,
                          sizes=sizes) + g_ips
    a = thermal_average(es[: np.shape(gs)[0]], -gs.imag, T=T)
    return gs, a, np.sum(a, axis=0)
