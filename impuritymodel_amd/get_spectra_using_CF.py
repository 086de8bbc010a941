"""Hamiltonian operators from crystal-field parameters
(impurityModel/ed/get_spectra_using_CF.py:335-627), without its command line.

``get_hamiltonian_operator_using_CF`` sums the same pieces as
``get_spectra.get_hamiltonian_operator`` but builds the non-interacting part
from the ten CF parameters of a JSON file (``read_h0_CF_file``, with the
reference's Ni-in-NiO defaults) instead of a pickled h0.
"""

from .operators import get_CF_hamiltonian, get_hamiltonian_operator_from_h0
from .readers import read_h0_CF_file  # noqa: F401  (reference name)


def get_hamiltonian_operator_using_CF(nBaths, nValBaths, slaterCondon, SOCs, DCinfo, hField, h0_CF_filename,
                                      bath_state_basis="spherical"):
    """hOp = U + hField + SOC(2p) + SOC(3d) - DC + h0_CF, integer-indexed
    (get_spectra_using_CF.py:335-430).  ``h0_CF_filename`` may also be the
    ten-tuple of CF parameters."""
    h0 = get_CF_hamiltonian(nBaths, nValBaths, h0_CF_filename, bath_state_basis)
    return get_hamiltonian_operator_from_h0(nBaths, slaterCondon, SOCs, DCinfo, hField, h0)
