"""The GPU engine against the reference's golden 50-bath NiO run, end to end.

``impuritymodel_amd.get_spectra.main`` runs the reference's composition
(get_spectra.py:21-332) on the fixture inputs: the run's parameters, its h0
(written back to a pickle and read by the allow-list reader) and the Ni 3d
radial orbital.  Every number the reference recorded is compared
(tests/golden/nio50_refout.*, from referenceOutput/Ni_NiO_50bath):

* hOp (1132 terms) in data.npz, the ground basis size, E0 and E - E0 of the
  5 lowest states, the eigenvectors' product-state counts and leading
  weights, the static expectation values (output.txt:38-54);
* the BFS basis size of every chain of every spectrum (1452 closures);
* PS.dat, XPS.dat, NIXS.dat and XAS.dat (%8.4f: 5e-5 plus a margin; PS 5e-4,
  see test_refout_oracle.py for why its unconverged K = 150 photoemission
  chains are looser) and RIXS.bin (float32).

RIXS: the reference solves (wIn + i delta + e - H) x = T_in psi with scipy's
bicgstab at its default relative residual 1e-5, and its out-going chains run
150 plain Lanczos steps on 870-900 states with 0.05 eV broadening; the engine
solves to 1e-12.  The map is compared at every fourth loss energy (the
fixture's rows) against its maximum.
"""

import numpy as np
import pytest

import refout
from impuritymodel_amd import get_spectra, readers
from impuritymodel_amd.average import k_B, thermal_average
from impuritymodel_amd.observables import (getEgT2gOccupation, getLsqr3d, getLz3d, getSsqr3d, getSz3d,
                                           getTraceDensityMatrix)
from impuritymodel_amd.output import read_dat, read_rixs_bin

pytestmark = pytest.mark.gpu

DAT_TOL = 5e-5 + 2e-6
PS_TOL = 5e-4
RIXS_REL_TOL = 2e-3


@pytest.fixture(scope="module")
def run(gpu, tmp_path_factory):
    d = tmp_path_factory.mktemp("refout")
    p = refout.params()
    h0 = refout.write_h0_pickle(str(d / "h0_NiO_50bath.pickle"))

    def pair(key):
        return tuple(p[key].values())

    res = get_spectra.main(
        h0, refout.radial_path(), ls=(1, 2), nBaths=pair("nBaths"), nValBaths=pair("nValBaths"),
        n0imps=pair("n0imps"), dnTols=pair("dnTols"), dnValBaths=pair("dnValBaths"), dnConBaths=pair("dnConBaths"),
        Fdd=tuple(p["Fdd"]), Fpp=tuple(p["Fpp"]), Fpd=tuple(p["Fpd"]), Gpd=tuple(p["Gpd"]), xi_2p=p["xi_2p"],
        xi_3d=p["xi_3d"], chargeTransferCorrection=p["chargeTransferCorrection"], hField=tuple(p["hField"]),
        nPsiMax=p["nPsiMax"], nPrintSlaterWeights=3, tolPrintOccupation=0.5, T=p["T"], energy_cut=10,
        delta=p["delta"], deltaRIXS=p["deltaRIXS"], deltaNIXS=p["deltaNIXS"], output_dir=str(d))
    return d, res


def test_refout_ground_state(run):
    d, res = run
    log = refout.log()
    ref = refout.hop()
    assert list(res["hOp"].keys()) == list(ref.keys())
    assert max(abs(res["hOp"][k] - ref[k]) for k in ref) < 2e-15
    assert len(res["basis"]) == log["n_basis"] == 1770
    es = res["es"]
    assert abs(es[0] - log["E0"]) <= 5e-5
    assert np.all(np.abs((es - es[0]) - np.array(log["table"])[:, 1]) <= 5e-4)
    assert len(res["psis"]) == log["n_spectra_states"] == 3
    # the run record reads back with the allow-list loader
    rec = readers.load_data_npz(str(d / "data.npz"))
    back = rec["hOp"].item()
    assert list(back.keys()) == list(ref.keys())
    assert rec["restrictions"].item() == refout.restrictions()


def test_refout_eigenvectors_and_expectation_values(run):
    """Eigenvector dicts truncated at |v|^2 >= 1e-7 (finite.py:114-117) have
    the log's product-state counts, and the static observables computed on
    the device match the printed table and thermal averages."""
    _, res = run
    log = refout.log()
    p = refout.params()
    nb = p["nBaths"]
    from impuritymodel_amd import finite

    basis = res["basis"]
    es, psis = finite.eigensystem(p["n_spin_orbitals"], res["hOp"], basis, p["nPsiMax"])
    assert [len(x) for x in psis] == log["product_state_counts"]
    for psi, lead in zip(psis, log["leading_weights"]):
        w = np.sort(np.abs(np.array(list(psi.values()))) ** 2)[::-1][:3]
        assert np.allclose(w, lead, atol=2e-6)
    table = np.array(log["table"])
    rows = []
    for psi in psis:
        occ = getEgT2gOccupation(nb, psi)
        rows.append([np.real(getTraceDensityMatrix(nb, psi)), *np.real(occ), np.real(getLz3d(nb, psi)),
                     np.real(getSz3d(nb, psi)), np.real(getLsqr3d(nb, psi)), np.real(getSsqr3d(nb, psi))])
    rows = np.array(rows)
    half = np.array([5e-3, 5e-4, 5e-4, 5e-4, 5e-4, 5e-3, 5e-3, 5e-3, 5e-3]) + 1e-6
    assert np.all(np.abs(rows - table[:, 2:]) <= half), np.abs(rows - table[:, 2:])
    keep = [i for i in range(len(es)) if es[i] - es[0] < 10 * k_B * p["T"]]
    avg = thermal_average(np.asarray(es)[keep], rows[keep], T=p["T"])
    names = ["N(3d)", "N(egDn)", "N(egUp)", "N(t2gDn)", "N(t2gUp)", "Lz(3d)", "Sz(3d)", "L^2(3d)", "S^2(3d)"]
    for name, v in zip(names, avg):
        assert abs(v - log["thermal"][name]) <= 5e-4 + 1e-6, (name, v)


def test_refout_bfs_sizes(run):
    _, res = run
    sizes = refout.log()["bfs_sizes"]
    info = res["spectra"]["info"]
    for name in ("PS", "XPS", "NIXS", "XAS"):
        assert info[name]["chain_n"] == sizes[name], name
    # RIXS log order: per (T_in, eigenstate) the core-hole sector, then its
    # out-going chains per (wIn, T_out)
    core = [int(x) for x in info["RIXS"]["core_n"]]
    out = [int(x) for x in info["RIXS"]["out_n"]]
    per = len(out) // len(core)
    seq = []
    for c in range(len(core)):
        seq.append(core[c])
        seq.extend(out[c * per:(c + 1) * per])
    assert seq == sizes["RIXS"]


@pytest.mark.parametrize("name", ["XAS", "XPS", "NIXS", "PS"])
def test_refout_dat_files(run, name):
    """The thermally averaged spectra per operator and their sum against the
    reference's %8.4f file (5e-5 is its rounding); the file written here holds
    the same values rounded the same way."""
    d, res = run
    ref = refout.tables()[name]
    _, a, s = res["spectra"][name]
    tol = PS_TOL if name == "PS" else DAT_TOL
    err = max(np.max(np.abs(a - ref[:, 2:].T)), np.max(np.abs(s - ref[:, 1])))
    print("%s.dat: max |diff| = %.3g (tolerance %.3g)" % (name, err, tol))
    assert err <= tol
    w, fs, fa = read_dat(str(d / (name + ".dat")))
    assert np.array_equal(w, ref[:, 0])
    assert np.max(np.abs(fa - a)) <= 5e-5 + 1e-9 and np.max(np.abs(fs - s)) <= 5e-5 + 1e-9


def test_refout_rixs_bin(run):
    d, _ = run
    t = refout.tables()
    head = t["RIXS_head"]
    n_w_in = int(head[0])
    wIn, wLoss, a = read_rixs_bin(str(d / "RIXS.bin"), n_w_in)
    assert np.array_equal(wIn, head[1:])
    rows = t["RIXS_rows"]
    idx = t["RIXS_row_index"]
    assert np.array_equal(wLoss[idx], rows[:, 0])
    ref = rows[:, 1:].T  # [wIn, loss row]
    got = a[:, idx]
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print("RIXS.bin: max |diff| / max = %.3g" % err)
    assert err <= RIXS_REL_TOL
