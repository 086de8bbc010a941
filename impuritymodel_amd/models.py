"""Seeded NiO-like Anderson impurity models (host side).

A model is the reference's 2p + 3d + ligand-bath layout (``c2i`` order:
2p spin-orbitals 0-5, 3d 6-15, bath 16...), with the Hamiltonian summed as in
get_spectra.py:335-418 and the occupation restrictions of
get_spectra.py:154-169.  The non-interacting part is the crystal-field form
of get_spectra_using_CF.py:433-563 with random bath energies and hoppings
drawn from ``numpy.random.default_rng(seed)``.
"""

from collections import OrderedDict
from dataclasses import dataclass, field

import numpy as np

from . import operators as ops

K_B = 1.380649e-23 / 1.602176634e-19  # eV/K, scipy.constants value (average.py:8)


@dataclass
class ImpurityModel:
    """One seeded NiO-like impurity model and everything needed for XAS."""

    n_bath: int
    hOp: dict
    nBaths: OrderedDict
    nValBaths: OrderedDict
    n0imps: OrderedDict
    dnTols: OrderedDict
    dnValBaths: OrderedDict
    dnConBaths: OrderedDict
    restrictions: dict
    params: dict = field(default_factory=dict)

    @property
    def n_spin_orbitals(self):
        return ops.n_spin_orbitals(self.nBaths)

    def term_table(self, W):
        """Compiled H term table (with restrictions) for W-word states.

        Built once per model: the model is an immutable input, its operator
        dict is converted to the engine's table format the first time a batch
        needs it.
        """
        from .opcompile import TermTable

        cache = self.__dict__.setdefault("_tables", {})
        if W not in cache:
            cache[W] = TermTable(self.hOp, self.restrictions, W)
        return cache[W]


def make_model(
    n_bath=10,
    seed=0,
    tenDq=0.6,
    e_imp=-1.31796,
    Fdd=(8.0, 0.0, 10.0, 0.0, 6.2),
    Fpp=(0.0, 0.0, 0.0),
    Fpd=(8.5, 0.0, 6.0),
    Gpd=(0.0, 4.5, 0.0, 2.5),
    xi_2p=11.5,
    xi_3d=0.08,
    chargeTransferCorrection=1.5,
    hField=(0.0, 0.0, 1e-4),
    eps_range=(-3.0, -1.0),
    v_range=(0.5, 2.0),
    n_con=0,
    dnValBath=2,
    dnConBath=0,
    dnTol=2,
    n0_3d=8,
):
    """Build a NiO-like model with ``n_bath`` valence bath spin-orbitals.

    Defaults are the small CPU-runnable case: 6 2p + 10 3d + 10 bath
    spin-orbitals, Slater integrals Fdd (8, 10, 6.2), Fpd (8.5, 6.0), Gpd
    (4.5, 2.5), 2p SOC 11.5 eV, 3d SOC 0.08 eV, 10Dq 0.6 eV, bath energies
    U(-3, -1) and hoppings U(0.5, 2) eV.  ``n_con`` adds conduction bath
    states (energies U(1, 3) eV) coupled like the valence ones.
    """
    rng = np.random.default_rng(seed)
    nb_tot = n_bath + n_con
    nBaths = OrderedDict([(1, 0), (2, nb_tot)])
    nValBaths = OrderedDict([(1, 0), (2, n_bath)])
    n0imps = OrderedDict([(1, 6), (2, n0_3d)])
    dnTols = OrderedDict([(1, 0), (2, dnTol)])
    dnValBaths = OrderedDict([(1, 0), (2, dnValBath)])
    dnConBaths = OrderedDict([(1, 0), (2, dnConBath)])
    eps = rng.uniform(eps_range[0], eps_range[1], n_bath)
    vv = rng.uniform(v_range[0], v_range[1], n_bath)
    if n_con:
        eps = np.concatenate([eps, rng.uniform(1.0, 3.0, n_con)])
        vv = np.concatenate([vv, rng.uniform(v_range[0], v_range[1], n_con)])
    h0 = ops.get_bath_h0(e_imp, tenDq, eps, vv)
    hOp = ops.get_hamiltonian_operator_from_h0(
        nBaths, [Fdd, Fpp, Fpd, Gpd], [xi_2p, xi_3d], [n0imps, chargeTransferCorrection], hField, h0
    )
    restrictions = get_restrictions(nBaths, nValBaths, n0imps, dnTols, dnValBaths, dnConBaths)
    params = dict(seed=seed, tenDq=tenDq, e_imp=e_imp, Fdd=Fdd, Fpd=Fpd, Gpd=Gpd, xi_2p=xi_2p, xi_3d=xi_3d,
                  chargeTransferCorrection=chargeTransferCorrection, hField=hField)
    return ImpurityModel(n_bath, hOp, nBaths, nValBaths, n0imps, dnTols, dnValBaths, dnConBaths, restrictions, params)


def get_restrictions(nBaths, nValBaths, n0imps, dnTols, dnValBaths, dnConBaths, l=2):
    """Excited-state restrictions {frozenset(indices): (lo, hi)} (get_spectra.py:154-169)."""
    r = {}
    idx = frozenset(ops.c2i(nBaths, (l, s, m)) for s in range(2) for m in range(-l, l + 1))
    r[idx] = (n0imps[l] - 1, n0imps[l] + dnTols[l] + 1)
    idx = frozenset(ops.c2i(nBaths, (l, b)) for b in range(nValBaths[l]))
    r[idx] = (nValBaths[l] - dnValBaths[l], nValBaths[l])
    idx = frozenset(ops.c2i(nBaths, (l, b)) for b in range(nValBaths[l], nBaths[l]))
    r[idx] = (0, dnConBaths[l])
    return r


def large_model(n_val=50, n_con=10, seed=7):
    """The large single instance: 6 2p + 10 3d + 50 valence + 10 conduction bath
    spin-orbitals (76, two words per state), bath energies U(-6, 4) eV (valence)
    and hoppings U(0.2, 1.5) eV, at most 2 valence holes and 2 conduction
    electrons.  Restrictions for the excited sectors: 3d occupancy 7..10, 2p
    core 5..6, valence >= n_val - 2, conduction <= 2."""
    m = make_model(n_bath=n_val, n_con=n_con, seed=seed, dnValBath=2, dnConBath=2, eps_range=(-6.0, 4.0),
                   v_range=(0.2, 1.5))
    nb = m.nBaths
    r = {
        frozenset(ops.c2i(nb, (2, s, mm)) for s in range(2) for mm in range(-2, 3)): (7, 10),
        frozenset(ops.c2i(nb, (1, s, mm)) for s in range(2) for mm in range(-1, 2)): (5, 6),
        frozenset(ops.c2i(nb, (2, b)) for b in range(n_val)): (n_val - 2, n_val),
        frozenset(ops.c2i(nb, (2, b)) for b in range(n_val, n_val + n_con)): (0, 2),
    }
    m.restrictions = r
    m.__dict__.pop("_tables", None)
    return m


def batched_params(n_models, seed=0):
    """Parameter draws of the batched workload: 10Dq U(0.3, 1.5), F0 U(5, 9),
    charge transfer U(1, 8) eV (used as the MLFT correction c), V U(0.5, 3)
    and bath size from {10, 20, 30, 40, 50}."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_models):
        out.append(
            dict(
                n_bath=int(rng.choice([10, 20, 30, 40, 50])),
                seed=int(rng.integers(0, 2**31 - 1)),
                tenDq=float(rng.uniform(0.3, 1.5)),
                F0=float(rng.uniform(5.0, 9.0)),
                delta=float(rng.uniform(1.0, 8.0)),
                vmax=float(rng.uniform(0.5, 3.0)),
            )
        )
    return out


def model_from_params(p):
    """NiO-like model from one :func:`batched_params` draw."""
    return make_model(
        n_bath=p["n_bath"],
        seed=p["seed"],
        tenDq=p["tenDq"],
        Fdd=(p["F0"], 0.0, 10.0, 0.0, 6.2),
        chargeTransferCorrection=p["delta"],
        v_range=(0.5, max(0.5, p["vmax"])),
    )
