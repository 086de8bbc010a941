"""Thermal averaging (impurityModel/ed/average.py).

The spectra path applies the Boltzmann weights on the device, fused into the
continued-fraction kernel (``imped_batch_cf``).  This host function keeps the
reference's standalone helper for callers that average their own arrays.
"""

import numpy as np

k_B = 1.380649e-23 / 1.602176634e-19  # eV/K, scipy.constants value (average.py:8)


def thermal_average(energies, observable, T=300):
    """Boltzmann average over axis 0 (average.py:11-45)."""
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
