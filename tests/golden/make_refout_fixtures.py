"""Write the fixtures of the reference's golden 50-bath NiO run.

The reference's end-to-end record (``impurityModel/test/referenceOutput/
Ni_NiO_50bath/``, made by ``scripts/run_Ni_NiO_Xbath.sh 50``) and its input
``h0/h0_NiO_50bath.pickle`` are data.  This script reads them in this
container -- the pickles through the package's allow-list unpickler, which
constructs nothing but numpy values, ordered dicts and frozensets -- and
writes them as fixtures that travel with the tests:

* ``nio50_refout.json``: the run's parameters, the h0 operator (224 terms),
  the full ``hOp`` of ``data.npz`` (1132 terms, insertion order kept), its
  restrictions, and the numbers of ``output.txt``: E0, the per-eigenstate
  table, product-state counts and leading weights, the NIXS radial
  integrals, and every BFS basis size in call order per spectrum section.
* ``nio50_refout.npz``: ``PS.dat``, ``XPS.dat``, ``NIXS.dat`` and ``XAS.dat``
  as exact integers (value x 1e4; the files are ``%8.4f``) and ``RIXS.bin``
  (float32) at every 4th energy-loss row.

Run: ``python tests/golden/make_refout_fixtures.py`` (needs the reference
tree; the GPU tests only read the written files).
"""

import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from impuritymodel_amd.readers import load_data_npz, read_pickled_file  # noqa: E402

REF = "/root/reference"
RUN = os.path.join(REF, "impurityModel/test/referenceOutput/Ni_NiO_50bath")
RIXS_ROW_STEP = 4


def _orb(x):
    return list(x) if isinstance(x, tuple) else int(x)


def encode_op(op):
    """[[ [[orbital, action], ...], re, im ], ...] in insertion order."""
    return [[[[_orb(o), a] for o, a in proc], float(np.real(v)), float(np.imag(v))] for proc, v in op.items()]


def parse_log(path):
    lines = open(path).read().splitlines()
    out = {}
    for ln in lines:
        m = re.match(r"E0 = (-?\d+\.\d+)", ln)
        if m:
            out["E0"] = float(m.group(1))
        m = re.match(r"#basis states = (\d+)", ln)
        if m:
            out["n_basis"] = int(m.group(1))
        m = re.match(r"(\d+) processes in the Hamiltonian", ln)
        if m:
            out["n_processes"] = int(m.group(1))
        m = re.match(r"<#Hamiltonian elements/column> = (\d+)", ln)
        if m:
            out["nnz_per_column"] = int(m.group(1))
        m = re.match(r"Consider (\d+) eigenstates", ln)
        if m:
            out["n_spectra_states"] = int(m.group(1))
    # per-eigenstate table after the header "  i  E-E0 ..."
    k = next(i for i, ln in enumerate(lines) if ln.strip().startswith("i  E-E0"))
    out["table_columns"] = lines[k].split()
    rows = []
    for ln in lines[k + 1:]:
        if not ln.strip():
            break
        rows.append([float(x) for x in ln.split()])
    out["table"] = rows
    # thermal expectation values "<name> = value" before E0
    thermal = {}
    for ln in lines[:k]:
        m = re.match(r"<(.+)> = (-?\d+\.\d+)$", ln.strip())
        if m:
            thermal[m.group(1)] = float(m.group(2))
    out["thermal"] = thermal
    # product-state counts and leading weights of the eigenstates
    counts, weights = [], []
    for i, ln in enumerate(lines):
        m = re.match(r"Consists of (\d+) product states", ln)
        if m:
            counts.append(int(m.group(1)))
            wl = lines[i + 2].strip().strip("[]")
            weights.append([float(x) for x in wl.split()])
    out["product_state_counts"] = counts
    out["leading_weights"] = weights
    # NIXS radial integrals
    out["nixs_rintegrals"] = [float(m.group(2)) for m in
                              (re.match(r"Rintegral\(k= (\d) \) = (\S+)", ln) for ln in lines) if m]
    # BFS basis sizes per section
    sections = {"PS": "Create 3d inverse photoemission", "XPS": "Create core 2p x-ray",
                "NIXS": "Create NIXS spectra", "XAS": "Create XAS spectra", "RIXS": "Create RIXS spectra"}
    starts = sorted((next(i for i, ln in enumerate(lines) if ln.startswith(tag)), name)
                    for name, tag in sections.items())
    sizes = {}
    for j, (s, name) in enumerate(starts):
        e = starts[j + 1][0] if j + 1 < len(starts) else len(lines)
        sizes[name] = [int(m.group(1)) for m in
                       (re.match(r"Hamiltonian basis sizes: len\(basis_index\) = (\d+)", ln) for ln in lines[s:e]) if m]
    out["bfs_sizes"] = sizes
    return out


def main():
    data = load_data_npz(os.path.join(RUN, "data.npz"))
    h0_name = os.path.basename(str(data["h0_filename"]))
    h0 = read_pickled_file(os.path.join(REF, "h0", h0_name))
    params = {}
    for k in ("ls", "Fdd", "Fpp", "Fpd", "Gpd", "hField", "epsilons", "epsilonsRIXSin", "epsilonsRIXSout"):
        params[k] = np.asarray(data[k]).tolist()
    for k in ("xi_2p", "xi_3d", "chargeTransferCorrection", "energy_cut", "delta", "deltaRIXS", "deltaNIXS"):
        params[k] = float(data[k])
    for k in ("nPsiMax", "T", "n_spin_orbitals"):
        params[k] = int(data[k])
    for k in ("nBaths", "nValBaths", "n0imps", "dnTols", "dnValBaths", "dnConBaths"):
        params[k] = [[int(a), int(b)] for a, b in data[k].item().items()]
    params["h0_filename"] = h0_name
    restrictions = [[sorted(int(i) for i in key), list(v)] for key, v in data["restrictions"].item().items()]
    fix = dict(params=params, h0=encode_op(h0), hOp=encode_op(data["hOp"].item()), restrictions=restrictions,
               log=parse_log(os.path.join(RUN, "output.txt")))
    with open(os.path.join(HERE, "nio50_refout.json"), "w") as fh:
        json.dump(fix, fh, separators=(",", ":"))
    arrays = {}
    for name in ("PS", "XPS", "NIXS", "XAS"):
        t = np.loadtxt(os.path.join(RUN, name + ".dat"))
        arrays[name] = np.rint(t * 1e4).astype(np.int32)
    rixs = np.fromfile(os.path.join(RUN, "RIXS.bin"), dtype=np.float32)
    n_w_in = int(rixs[0])
    rixs = rixs.reshape(-1, n_w_in + 1)
    arrays["RIXS_head"] = rixs[0]
    arrays["RIXS_rows"] = rixs[1::RIXS_ROW_STEP]
    arrays["RIXS_row_index"] = np.arange(0, rixs.shape[0] - 1, RIXS_ROW_STEP)
    np.savez_compressed(os.path.join(HERE, "nio50_refout.npz"), **arrays)
    print("wrote", os.path.join(HERE, "nio50_refout.json"), os.path.join(HERE, "nio50_refout.npz"))


if __name__ == "__main__":
    main()
