"""The reference's output formats (CPU).

Pinned by the first lines of the reference's own Ni_NiO_50bath output
(tests/golden/refout_XAS_head.dat, refout_RIXS_head.npy, written by
tests/golden/make_format_fixtures.py).
"""

import os

import numpy as np

from impuritymodel_amd import output
from oracle import spectra as ospec

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_dat_table_reproduces_reference_text(tmp_path):
    with open(os.path.join(GOLD, "refout_XAS_head.dat")) as f:
        ref = f.read()
    t = np.loadtxt(os.path.join(GOLD, "refout_XAS_head.dat"))
    p = tmp_path / "XAS.dat"
    output.write_dat(str(p), t[:, 0], t[:, 2:].T)
    assert p.read_text() == ref
    w, s, a = output.read_dat(str(p))
    assert np.array_equal(w, t[:, 0]) and np.array_equal(s, t[:, 1]) and a.shape == (3, 5)


def test_rixs_bin_layout_matches_reference(tmp_path):
    ref = np.load(os.path.join(GOLD, "refout_RIXS_head.npy"))
    assert ref[0, 0] == 50 and np.all(np.diff(ref[0, 1:]) > 0) and np.all(np.diff(ref[1:, 0]) > 0)
    w_in, w_loss, body = ref[0, 1:], ref[1:, 0], ref[1:, 1:].T
    t = output.rixs_table(w_in, w_loss, body)
    assert t.dtype == np.float32 and np.array_equal(t, ref)
    p = tmp_path / "RIXS.bin"
    output.write_rixs_bin(str(p), w_in, w_loss, body)
    assert p.stat().st_size == ref.size * 4
    a, b, c = output.read_rixs_bin(str(p), 50)
    assert np.array_equal(a, w_in) and np.array_equal(b, w_loss) and np.array_equal(c, body)
    # the oracle's restatement of the same layout
    assert np.array_equal(ospec.rixs_bin(w_in, w_loss, body), ref)


def test_spectra_datasets_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    ds = dict(E=np.array([-1.0, -0.5]), w=np.linspace(-1, 1, 5), XAS=rng.normal(size=(2, 3, 5)),
              XASthermal=rng.normal(size=(3, 5)))
    out = output.write_spectra(str(tmp_path / "spectra.h5"), **ds)
    if out.endswith(".npz"):
        z = np.load(out)
        for k, v in ds.items():
            assert np.array_equal(z[k], v)
