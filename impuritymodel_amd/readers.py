"""Readers and writers of the reference's input and output data files.

* ``read_pickled_file`` -- the non-interacting Hamiltonians ``h0/*.pickle``
  (get_spectra.py:421-424), through an allow-list unpickler.
* ``load_data_npz`` / ``save_data_npz`` -- the run record ``data.npz``
  (get_spectra.py:254-284), whose dict-valued entries are object arrays.
* ``read_radial_file`` -- the radial orbitals ``radialOrbitals/*.dat``
  (get_spectra.py:172-173).
* ``read_h0_CF_file`` -- the crystal-field parameter JSON files ``h0/*_CF.json``
  with the reference's defaults (get_spectra_using_CF.py:566-627).

The reference reads its pickles with ``pickle.load``, which executes whatever
the file names.  These files hold only numpy scalars and arrays, ordered
dicts and frozensets (and Python complex numbers in files this package
writes), so the unpickler here resolves exactly those globals and refuses
every other one before anything is constructed: a file naming any other
callable raises ``pickle.UnpicklingError`` and runs nothing.
"""

import importlib
import io
import json
import pickle
import zipfile

import numpy as np

# (module, name) as written in the files -> module to resolve it in.  Files
# written by numpy 1.x name numpy.core.multiarray, which numpy 2 moved.
_ALLOWED = {
    ("numpy.core.multiarray", "scalar"),
    ("numpy.core.multiarray", "_reconstruct"),
    ("numpy._core.multiarray", "scalar"),
    ("numpy._core.multiarray", "_reconstruct"),
    ("numpy", "dtype"),
    ("numpy", "ndarray"),
    ("collections", "OrderedDict"),
    ("builtins", "frozenset"),
    ("builtins", "complex"),
}


def _resolve_module(module):
    if module == "numpy.core.multiarray":
        try:
            return importlib.import_module("numpy._core.multiarray")
        except ImportError:  # numpy 1.x
            return importlib.import_module("numpy.core.multiarray")
    return importlib.import_module(module)


class AllowListUnpickler(pickle.Unpickler):
    """Unpickler that constructs only numpy arrays/scalars/dtypes, OrderedDict,
    frozenset and complex; any other global in the stream raises
    UnpicklingError."""

    def find_class(self, module, name):
        if (module, name) not in _ALLOWED:
            raise pickle.UnpicklingError("refused global %s.%s (not an allowed data type)" % (module, name))
        return getattr(_resolve_module(module), name)


def loads(data):
    """Unpickle bytes with the allow-list unpickler."""
    return AllowListUnpickler(io.BytesIO(data)).load()


def read_pickled_file(filename: str):
    """Content of a pickled operator file such as ``h0/h0_NiO_50bath.pickle``
    (get_spectra.py:421-424), loaded with the allow-list unpickler."""
    with open(filename, "rb") as handle:
        return AllowListUnpickler(handle).load()


def _read_npy(raw):
    fp = io.BytesIO(raw)
    version = np.lib.format.read_magic(fp)
    if version == (1, 0):
        shape, fortran, dtype = np.lib.format.read_array_header_1_0(fp)
    else:
        shape, fortran, dtype = np.lib.format.read_array_header_2_0(fp)
    if dtype.hasobject:
        # object arrays are pickled lists; the allow-list decides what is built
        return AllowListUnpickler(fp).load()
    count = int(np.prod(shape)) if shape else 1
    arr = np.frombuffer(fp.read(count * dtype.itemsize), dtype=dtype, count=count)
    return arr.reshape(shape, order="F" if fortran else "C").copy()


def load_data_npz(path):
    """All entries of a ``data.npz`` run record as a dict of arrays.

    Object entries (``nBaths``, ``restrictions``, ``hOp``, ...) are 0-d
    object arrays around the dict, as numpy wrote them
    (get_spectra.py:254-284); ``.item()`` gives the dict.
    """
    out = {}
    with zipfile.ZipFile(path) as z:
        for name in z.namelist():
            key = name[:-4] if name.endswith(".npy") else name
            out[key] = _read_npy(z.read(name))
    return out


def save_data_npz(path, **entries):
    """Write a run record like get_spectra.py:254-284 (``np.savez_compressed``).

    Dict values (nBaths, restrictions, hOp, ...) are stored as 0-d object
    arrays, as numpy does for the reference, so the reference's own tooling
    reads the file with ``np.load(..., allow_pickle=True)``.
    """
    arrays = {}
    for k, v in entries.items():
        if isinstance(v, dict):
            a = np.empty((), dtype=object)
            a[()] = v
            arrays[k] = a
        else:
            arrays[k] = np.asarray(v)
    np.savez_compressed(path, **arrays)


def read_radial_file(filename):
    """(radialMesh, R) of a two-column radial orbital file (get_spectra.py:172-173)."""
    radial_mesh, r = np.loadtxt(filename).T
    return radial_mesh, r


# Defaults of read_h0_CF_file (get_spectra_using_CF.py:604-615): Ni in NiO.
CF_DEFAULTS = (
    ("e_imp", -1.31796),
    ("e_deltaO_imp", 0.60422),
    ("e_val_eg", -4.4),
    ("e_val_t2g", -6.5),
    ("e_con_eg", 3),
    ("e_con_t2g", 2),
    ("v_val_eg", 1.883),
    ("v_val_t2g", 1.395),
    ("v_con_eg", 0.6),
    ("v_con_t2g", 0.4),
)


def cf_parameters(parameters):
    """The ten crystal-field parameters in the reference's order, missing ones
    taken from the defaults (get_spectra_using_CF.py:566-627)."""
    return tuple(parameters.get(name, default) for name, default in CF_DEFAULTS)


def read_h0_CF_file(h0_CF_filename):
    """(e_imp, e_deltaO_imp, e_val_eg, e_val_t2g, e_con_eg, e_con_t2g,
    v_val_eg, v_val_t2g, v_con_eg, v_con_t2g) from a CF JSON file
    (get_spectra_using_CF.py:566-627)."""
    with open(h0_CF_filename, "r") as fh:
        return cf_parameters(json.loads(fh.read()))
