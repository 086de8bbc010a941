"""Fixtures of the reference's golden 50-bath NiO run (tests/golden/nio50_refout.*).

Decodes the operators and parameters written by
``tests/golden/make_refout_fixtures.py`` and gives the run's grids, so the
CPU oracle tests and the GPU parity tests set up the same problem.
"""

import json
import os
import pickle
from collections import OrderedDict

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
_cache = {}


def fixture():
    if "json" not in _cache:
        with open(os.path.join(GOLDEN, "nio50_refout.json")) as fh:
            _cache["json"] = json.load(fh)
    return _cache["json"]


def tables():
    """The reference's PS/XPS/NIXS/XAS .dat tables (float) and RIXS.bin rows."""
    if "npz" not in _cache:
        z = np.load(os.path.join(GOLDEN, "nio50_refout.npz"))
        out = {k: z[k].astype(np.float64) / 1e4 for k in ("PS", "XPS", "NIXS", "XAS")}
        out["RIXS_head"] = z["RIXS_head"]
        out["RIXS_rows"] = z["RIXS_rows"]
        out["RIXS_row_index"] = z["RIXS_row_index"]
        _cache["npz"] = out
    return _cache["npz"]


def _orb(x):
    return tuple(x) if isinstance(x, list) else x


def decode_op(terms, numpy_scalars=False):
    """Operator dict from the fixture encoding (insertion order kept)."""
    op = {}
    for proc, re_, im in terms:
        key = tuple((_orb(o), a) for o, a in proc)
        op[key] = np.complex128(complex(re_, im)) if numpy_scalars else complex(re_, im)
    return op


def params():
    p = fixture()["params"]
    out = dict(p)
    for k in ("nBaths", "nValBaths", "n0imps", "dnTols", "dnValBaths", "dnConBaths"):
        out[k] = OrderedDict((a, b) for a, b in p[k])
    return out


def h0():
    return decode_op(fixture()["h0"], numpy_scalars=True)


def hop():
    """The reference run's hOp (data.npz), 1132 terms."""
    return decode_op(fixture()["hOp"])


def restrictions():
    return {frozenset(k): tuple(v) for k, v in fixture()["restrictions"]}


def write_h0_pickle(path):
    """The h0 operator as a pickle of numpy complex values, like the
    reference's h0/h0_NiO_50bath.pickle (for the pickle reader)."""
    with open(path, "wb") as fh:
        pickle.dump(h0(), fh, protocol=3)
    return path


def radial_path():
    return os.path.join(GOLDEN, "Ni3d.dat")


def grids():
    """Meshes, polarizations and NIXS momenta of the run (get_spectra.py:136-152)."""
    return dict(
        w=np.linspace(-25, 25, 3000),
        wIn=np.linspace(-10, 20, 50),
        wLoss=np.linspace(-2, 12, 4000),
        epsilons=[[1, 0, 0], [0, 1, 0], [0, 0, 1]],
        qsNIXS=[2 * np.array([1, 1, 1]) / np.sqrt(3), 7 * np.array([1, 1, 1]) / np.sqrt(3)],
    )


def log():
    return fixture()["log"]
