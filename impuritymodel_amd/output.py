"""Spectra files in the reference's formats.

* ``write_dat``: the ``PS.dat`` / ``XPS.dat`` / ``NIXS.dat`` / ``XAS.dat``
  text tables of simulate_spectra (spectra.py:150-157, 180-187, 209-216,
  239-246): columns energy, operator sum, one column per transition
  operator, ``%8.4f``, header ``E  sum  T1  T2  T3 ...``.
* ``write_rixs_bin``: ``RIXS.bin`` (spectra.py:284-293), float32 table
  [len(wLoss)+1, len(wIn)+1] with [0, 0] = len(wIn), first row wIn, first
  column wLoss and the body the thermally averaged map summed over
  polarizations, transposed to [wLoss, wIn].
* ``write_spectra``: the datasets of ``spectra.h5`` (get_spectra.py:287-295,
  spectra.py:146-147, ...).  HDF5 needs h5py; without it the same dataset
  names go to an ``.npz`` archive.
"""

import numpy as np

DAT_HEADER = "E  sum  T1  T2  T3 ..."


def write_dat(path, w, a):
    """w [n_w], a [n_ops, n_w] thermally averaged -Im per operator."""
    a = np.atleast_2d(np.asarray(a, np.float64))
    cols = [np.asarray(w, np.float64), a.sum(axis=0)] + [a[i] for i in range(a.shape[0])]
    np.savetxt(path, np.array(cols).T, fmt="%8.4f", header=DAT_HEADER)


def read_dat(path):
    """(w, sum, per-operator [n_ops, n_w]) from a .dat table."""
    t = np.loadtxt(path)
    return t[:, 0], t[:, 1], t[:, 2:].T


def rixs_table(wIn, wLoss, aSum):
    """The float32 RIXS.bin table for a map aSum [len(wIn), len(wLoss)]."""
    t = np.zeros((len(wLoss) + 1, len(wIn) + 1), dtype=np.float32)
    t[0, 0] = len(wIn)
    t[0, 1:] = wIn
    t[1:, 0] = wLoss
    t[1:, 1:] = np.asarray(aSum).T
    return t


def write_rixs_bin(path, wIn, wLoss, aSum):
    rixs_table(wIn, wLoss, aSum).tofile(path)


def read_rixs_bin(path, n_w_in):
    """(wIn, wLoss, aSum [len(wIn), len(wLoss)]) from RIXS.bin with n_w_in incoming energies."""
    t = np.fromfile(path, dtype=np.float32).reshape(-1, n_w_in + 1)
    return t[0, 1:], t[1:, 0], t[1:, 1:].T


def write_spectra(path, **datasets):
    """spectra.h5 datasets (E, w, wIn, wLoss, XAS, XASthermal, ...).

    Writes HDF5 when h5py is importable, else an .npz archive with the same
    dataset names; returns the path written."""
    try:
        import h5py
    except ImportError:
        out = path[:-3] + ".npz" if path.endswith(".h5") else path + ".npz"
        np.savez(out, **{k: np.asarray(v) for k, v in datasets.items()})
        return out
    with h5py.File(path, "w") as f:
        for k, v in datasets.items():
            f.create_dataset(k, data=np.asarray(v))
    return path
