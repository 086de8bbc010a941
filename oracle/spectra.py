"""Green's functions, spectra and thermal averages (oracle).

Restates ``getGreen`` (impurityModel/ed/spectra.py:551-681, sparse branch),
``getSpectra`` (spectra.py:684-806, one-rank ``H_build`` branch 778-803),
the XAS/PS/XPS composition of ``simulate_spectra`` (spectra.py:124-238) and
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
              reorth="none", return_ab=False):
    """<psi|(omega + i delta + e - H)^-1|psi> by Lanczos (spectra.py:551-681)."""
    g = np.zeros(len(omega), dtype=complex)
    if len(psi) == 0:
        return (g, None, None) if return_ab else g
    if h_dict is None:
        h_dict = {}
    h, index = expand_basis_and_hamiltonian(n, h_dict, hOp, psi.keys(), restrictions)
    nb = len(index)
    psi0 = np.zeros(nb, dtype=complex)
    for ps, amp in psi.items():
        psi0[index[ps]] = amp
    K = min(krylovSize, nb)
    alpha, beta = lanczos(h, psi0, K, reorth=reorth)
    g = continued_fraction(alpha, beta, omega, delta, e)
    return (g, alpha, beta) if return_ab else g


def get_spectra(n, hOp, tOps, psis, es, w, delta, restrictions=None, krylovSize=150, slaterWeightMin=1e-7,
                reorth="none"):
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
                n, es[i], psiR, hOp, w, delta, krylovSize, slaterWeightMin, restrictions, h, reorth=reorth
            )
    return gs


def energy_cut(es, T=300, cut=10.0):
    """Keep states with e - e0 < cut*k_B*T (get_spectra.py:134, 302-303)."""
    lim = cut * k_B * T
    return tuple(e for e in es if e - es[0] < lim)


def xas(n, hOp, tOps, psis, es, w, delta, restrictions, T=300, krylovSize=150, reorth="none"):
    """XAS branch of simulate_spectra (spectra.py:222-238).

    Returns ``(gs, a, aSum)``: raw Green's functions, thermally averaged
    -Im per polarization, and their sum.
    """
    gs = get_spectra(n, hOp, tOps, psis, es, w, delta, restrictions, krylovSize, reorth=reorth)
    a = thermal_average(es[: np.shape(gs)[0]], -gs.imag, T=T)
    return gs, a, np.sum(a, axis=0)


def photoemission(n, hOp, tOpsPS, tOpsIPS, psis, es, w, delta, restrictions, T=300, krylovSize=150,
                  reorth="none"):
    """PS + IPS branch of simulate_spectra (spectra.py:124-149).

    ``tOpsIPS`` may be None for the core XPS branch (spectra.py:163-179).
    """
    gs = -get_spectra(n, hOp, tOpsPS, psis, es, -w, -delta, restrictions, krylovSize, reorth=reorth)
    if tOpsIPS is not None:
        gs = gs + get_spectra(n, hOp, tOpsIPS, psis, es, w, delta, restrictions, krylovSize, reorth=reorth)
    a = thermal_average(es[: np.shape(gs)[0]], -gs.imag, T=T)
    return gs, a, np.sum(a, axis=0)
