"""The spectra run of impurityModel/ed/get_spectra.py, without its command line.

``main`` takes the reference's parameters (get_spectra.py:21-332) and runs the
same composition on the GPU engine: Hamiltonian operator from an h0 pickle,
restricted ground basis, lowest eigenstates, static expectation values, the
``data.npz`` run record, the energy cut and ``simulate_spectra`` (PS/IPS, XPS,
NIXS, XAS, RIXS) with its ``.dat``/``RIXS.bin`` files and spectra datasets.
The argument parser of the reference (get_spectra.py:427-618) is not
reproduced; its defaults are ``DEFAULTS`` here.
"""

import os
import time
from collections import OrderedDict

import numpy as np

from . import finite, spectra
from .average import k_B
from .models import get_restrictions
from .observables import printExpValues, printThermalExpValues
from .operators import get_hamiltonian_operator_from_h0
from .output import write_spectra
from .readers import read_pickled_file, read_radial_file, save_data_npz

# Command-line defaults of the reference (get_spectra.py:440-590).
DEFAULTS = dict(
    ls=(1, 2),
    nBaths=(0, 10),
    nValBaths=(0, 10),
    n0imps=(6, 8),
    dnTols=(0, 2),
    dnValBaths=(0, 2),
    dnConBaths=(0, 0),
    Fdd=(7.5, 0, 9.9, 0, 6.6),
    Fpp=(0.0, 0.0, 0.0),
    Fpd=(8.9, 0, 6.8),
    Gpd=(0.0, 5.0, 0, 2.8),
    xi_2p=11.629,
    xi_3d=0.096,
    chargeTransferCorrection=1.5,
    hField=(0, 0, 0.0001),
    nPsiMax=5,
    nPrintSlaterWeights=3,
    tolPrintOccupation=0.5,
    T=300,
    energy_cut=10,
    delta=0.2,
    deltaRIXS=0.050,
    deltaNIXS=0.100,
)


def get_hamiltonian_operator(nBaths, nValBaths, slaterCondon, SOCs, DCinfo, hField, h0_filename):
    """hOp = U + hField + SOC(2p) + SOC(3d) - DC + h0, integer-indexed
    (get_spectra.py:335-418); h0 read from a pickle file by the allow-list
    reader.  Asserts hermiticity as the reference does."""
    h0_operator = read_pickled_file(h0_filename)
    return get_hamiltonian_operator_from_h0(nBaths, slaterCondon, SOCs, DCinfo, hField, h0_operator)


def spectra_grids():
    """Energy meshes, polarizations and NIXS momenta of the run
    (get_spectra.py:136-152)."""
    return dict(
        w=np.linspace(-25, 25, 3000),
        epsilons=[[1, 0, 0], [0, 1, 0], [0, 0, 1]],
        epsilonsRIXSin=[[1, 0, 0], [0, 1, 0], [0, 0, 1]],
        epsilonsRIXSout=[[1, 0, 0], [0, 1, 0], [0, 0, 1]],
        wIn=np.linspace(-10, 20, 50),
        wLoss=np.linspace(-2, 12, 4000),
        qsNIXS=[2 * np.array([1, 1, 1]) / np.sqrt(3), 7 * np.array([1, 1, 1]) / np.sqrt(3)],
        liNIXS=2,
        ljNIXS=2,
    )


class _Datasets:
    """Collects ``create_dataset`` calls; written once as spectra.h5 (or .npz)."""

    def __init__(self):
        self.data = {}

    def create_dataset(self, name, data):
        self.data[name] = np.asarray(data)


def main(h0_filename, radial_filename, ls, nBaths, nValBaths, n0imps, dnTols, dnValBaths, dnConBaths, Fdd, Fpp, Fpd,
         Gpd, xi_2p, xi_3d, chargeTransferCorrection, hField, nPsiMax, nPrintSlaterWeights, tolPrintOccupation, T,
         energy_cut, delta, deltaRIXS, deltaNIXS, output_dir="."):
    """Ground state, then all spectra (get_spectra.py:21-332).

    Writes ``data.npz``, ``spectra.h5`` (``spectra.npz`` without h5py) and the
    ``.dat``/``RIXS.bin`` spectra into ``output_dir``.  Returns a dict with
    ``es`` (all eigenvalues), ``psis``, ``hOp``, ``basis``, ``restrictions``
    and ``spectra`` (simulate_spectra's {section: (gs, a, aSum)}).
    ``nPrintSlaterWeights`` and ``tolPrintOccupation`` are accepted for the
    signature; the product-state and density-matrix listings are not printed.
    """
    t0 = time.time()
    nBaths = OrderedDict(zip(ls, nBaths))
    nValBaths = OrderedDict(zip(ls, nValBaths))
    n0imps = OrderedDict(zip(ls, n0imps))
    dnTols = OrderedDict(zip(ls, dnTols))
    dnValBaths = OrderedDict(zip(ls, dnValBaths))
    dnConBaths = OrderedDict(zip(ls, dnConBaths))
    energy_cut *= k_B * T
    g = spectra_grids()
    restrictions = get_restrictions(nBaths, nValBaths, n0imps, dnTols, dnValBaths, dnConBaths)
    radialMesh, RiNIXS = read_radial_file(radial_filename)
    RjNIXS = np.copy(RiNIXS)
    n_spin_orbitals = sum(2 * (2 * ang + 1) + nBath for ang, nBath in nBaths.items())
    print("#spin-orbitals:", n_spin_orbitals)
    hOp = get_hamiltonian_operator(nBaths, nValBaths, [Fdd, Fpp, Fpd, Gpd], [xi_2p, xi_3d],
                                   [n0imps, chargeTransferCorrection], hField, h0_filename)
    print("{:d} processes in the Hamiltonian.".format(len(hOp)))
    basis = finite.get_basis(nBaths, nValBaths, dnValBaths, dnConBaths, dnTols, n0imps)
    print("#basis states = {:d}".format(len(basis)))
    es, psis = finite.eigensystem(n_spin_orbitals, hOp, basis, nPsiMax)
    print("time(ground_state) = {:.2f} seconds \n".format(time.time() - t0))
    t0 = time.time()
    printThermalExpValues(nBaths, es, psis)
    printExpValues(nBaths, es, psis)
    save_data_npz(os.path.join(output_dir, "data"), ls=ls, nBaths=nBaths, nValBaths=nValBaths, n0imps=n0imps,
                  dnTols=dnTols, dnValBaths=dnValBaths, dnConBaths=dnConBaths, Fdd=Fdd, Fpp=Fpp, Fpd=Fpd, Gpd=Gpd,
                  xi_2p=xi_2p, xi_3d=xi_3d, chargeTransferCorrection=chargeTransferCorrection, hField=hField,
                  h0_filename=h0_filename, nPsiMax=nPsiMax, T=T, energy_cut=energy_cut, delta=delta,
                  restrictions=restrictions, epsilons=g["epsilons"], epsilonsRIXSin=g["epsilonsRIXSin"],
                  epsilonsRIXSout=g["epsilonsRIXSout"], deltaRIXS=deltaRIXS, deltaNIXS=deltaNIXS,
                  n_spin_orbitals=n_spin_orbitals, hOp=hOp)
    h5f = _Datasets()
    for name, value in (("E", es), ("w", g["w"]), ("wIn", g["wIn"]), ("wLoss", g["wLoss"]), ("qsNIXS", g["qsNIXS"]),
                        ("r", radialMesh), ("RiNIXS", RiNIXS), ("RjNIXS", RjNIXS)):
        h5f.create_dataset(name, data=value)
    print("time(expectation values) = {:.2f} seconds \n".format(time.time() - t0))
    all_es = np.array(es)
    es = tuple(e for e in es if e - es[0] < energy_cut)
    psis = tuple(psis[i] for i in range(len(es)))
    print("Consider {:d} eigenstates for the spectra \n".format(len(es)))
    out = spectra.simulate_spectra(es, psis, hOp, T, g["w"], delta, g["epsilons"], g["wLoss"], deltaNIXS,
                                   g["qsNIXS"], g["liNIXS"], g["ljNIXS"], RiNIXS, RjNIXS, radialMesh, g["wIn"],
                                   deltaRIXS, g["epsilonsRIXSin"], g["epsilonsRIXSout"], restrictions, h5f, nBaths,
                                   output_dir=output_dir)
    write_spectra(os.path.join(output_dir, "spectra.h5"), **h5f.data)
    return dict(es=all_es, psis=psis, hOp=hOp, basis=basis, restrictions=restrictions, spectra=out)
