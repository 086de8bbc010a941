"""GPU parity: the HIP engine against the CPU oracle and the golden vectors.

Bars: product states, insertion order, bases and sparsity patterns
bit-exact; matrix elements 1e-12; eigenvalues 1e-10; spectra 1e-8 relative
(to the largest value); Lanczos alpha/beta 1e-10 while the recurrence is
orthogonal (K = 40 on a 66-spin-orbital chain).
"""

import os

import numpy as np
import pytest
import scipy.sparse

from oracle import applyop, basis as obasis, hamiltonian as oham, krylov as okry, spectra as ospec
from oracle import states as psr

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def words_to_ints(w, n):
    from impuritymodel_amd import opcompile

    return opcompile.words_to_ints(np.asarray(w, np.uint64), n)


def ints_to_bytes(states, n):
    return [psr.int2bytes(s, n) for s in states]


def bytes_to_ints(keys, n):
    return [psr.bytes2int(k, n) for k in keys]


# ------------------------------------------------------------------ applyOp


def test_applyop_reference_kats(gpu):
    from impuritymodel_amd import finite

    vac = "000000"
    psi = {psr.str2bytes(vac): 7}
    for i in range(6):
        assert finite.applyOp(6, {((i, "c"),): 9}, psi) == {psr.str2bytes(vac[:i] + "1" + vac[i + 1:]): 63}
        assert finite.applyOp(6, {((i, "a"),): 9}, psi) == {}
    for i in range(6):
        for j in range(6):
            r = finite.applyOp(6, {((i, "c"), (j, "c")): 9}, psi)
            if i == j:
                assert r == {}
            else:
                a, b = min(i, j), max(i, j)
                s = vac[:a] + "1" + vac[a + 1:b] + "1" + vac[b + 1:]
                assert r == {psr.str2bytes(s): 7 * 9 * (2 * (i < j) - 1)}
            rr = finite.applyOp(6, {((i, "c"), (j, "c")): 9}, psi, restrictions={frozenset([1, 4, 5]): (0, 1)})
            if i in (1, 4, 5) and j in (1, 4, 5):
                assert rr == {}
    ps = "011000"
    r = finite.applyOp(6, {((1, "c"), (1, "a")): 9, ((2, "c"), (2, "a")): 11}, {psr.str2bytes(ps): 7})
    assert r == {psr.str2bytes(ps): 140}


def test_applyop_opresult_memo(gpu):
    from impuritymodel_amd import finite

    vac = psr.str2bytes("000000")
    op = {((0, "c"),): 9, ((3, "c"),): 11}
    memo = {}
    r = finite.applyOp(6, op, {vac: 7}, opResult=memo)
    assert r == {psr.str2bytes("100000"): 63, psr.str2bytes("000100"): 77}
    assert memo == {vac: {psr.str2bytes("100000"): 9, psr.str2bytes("000100"): 11}}
    wrong = {vac: {psr.str2bytes("111111"): 2}}
    assert finite.applyOp(6, op, {vac: 7}, opResult=wrong) == {psr.str2bytes("111111"): 14}


@pytest.mark.parametrize("n", [10, 26, 66, 100])
def test_applyop_random_vs_oracle(gpu, n):
    from impuritymodel_amd import finite

    rng = np.random.default_rng(100 + n)
    for trial in range(12):
        op = {}
        for _ in range(int(rng.integers(1, 30))):
            k = int(rng.choice([1, 2, 4]))
            proc = tuple((int(rng.integers(n)), str(rng.choice(["c", "a"]))) for _ in range(k))
            op[proc] = complex(rng.normal(), rng.normal())
        states = {int.from_bytes(rng.bytes(16), "big") >> (128 - n) for _ in range(int(rng.integers(1, 40)))}
        psi = {s: complex(rng.normal(), rng.normal()) for s in states}
        restr = None
        if trial % 2:
            idx = frozenset(int(x) for x in rng.choice(n, size=min(n, 8), replace=False))
            restr = {idx: (2, 5)}
        ref = applyop.apply_op(n, op, psi, 1e-12, restr)
        got = finite.applyOp(n, op, dict(zip(ints_to_bytes(list(psi), n), psi.values())), 1e-12, restr)
        assert bytes_to_ints(list(got), n) == list(ref), "insertion order / membership differs"
        np.testing.assert_allclose(list(got.values()), list(ref.values()), rtol=0, atol=1e-12)


# ------------------------------------------------------------------ Hamiltonians


def test_fixed_basis_hamiltonian_golden(gpu):
    from impuritymodel_amd import finite
    from impuritymodel_amd.models import make_model

    g = np.load(os.path.join(GOLD, "ground_10bath_s7.npz"))
    m = make_model(n_bath=10, seed=int(g["seed"]))
    n = m.n_spin_orbitals
    b = finite.get_basis(m.nBaths, m.nValBaths, m.dnValBaths, m.dnConBaths, m.dnTols, m.n0imps)
    assert bytes_to_ints(b, n) == words_to_ints(g["basis"], n)
    h = finite.get_hamiltonian_matrix(n, m.hOp, b).tocoo()
    got = {(r, c): v for r, c, v in zip(h.row, h.col, h.data)}
    ref = {(r, c): v for r, c, v in zip(g["h_row"], g["h_col"], g["h_val"])}
    assert set(got) == set(ref)
    assert max(abs(got[k] - ref[k]) for k in ref) <= 1e-12


def test_expand_chain_golden_w2(gpu):
    from impuritymodel_amd import finite
    from impuritymodel_amd.models import make_model

    g = np.load(os.path.join(GOLD, "chain_50bath_s2.npz"))
    m = make_model(n_bath=50, seed=int(g["seed"]))
    n = m.n_spin_orbitals
    seeds = ints_to_bytes(words_to_ints(g["seed_keys"], n), n)
    h, index = finite.expand_basis_and_hamiltonian(n, {}, m.hOp, seeds, m.restrictions)
    assert bytes_to_ints(list(index), n) == words_to_ints(g["basis"], n)
    h = h.tocoo()
    got = {(r, c): v for r, c, v in zip(h.row, h.col, h.data)}
    ref = {(r, c): v for r, c, v in zip(g["h_row"], g["h_col"], g["h_val"])}
    assert set(got) == set(ref)
    assert max(abs(got[k] - ref[k]) for k in ref) <= 1e-12
    psi0 = np.zeros(len(index), complex)
    vals = g["seed_vals"] / np.sqrt(np.sum(np.abs(g["seed_vals"]) ** 2))
    for s, v in zip(seeds, vals):
        psi0[index[s]] = v
    a, b = finite.get_tridiagonal_krylov_vectors(scipy.sparse.csr_matrix(h), psi0, 40)
    # Plain Lanczos amplifies summation-order differences step by step:
    # 1e-12 relative over the first 10 steps, 1e-9 relative over all 40.
    scale = np.abs(g["alpha"]).max()
    np.testing.assert_allclose(a[:10], g["alpha"][:10], rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(b[:10], g["beta"][:10], rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(a, g["alpha"], rtol=0, atol=1e-9 * scale)
    np.testing.assert_allclose(b, g["beta"], rtol=0, atol=1e-9 * scale)


def test_expand_many_seeds_vs_oracle(gpu):
    from impuritymodel_amd import finite
    from impuritymodel_amd.models import make_model

    m = make_model(n_bath=20, seed=11, n_con=4, dnConBath=1)
    n = m.n_spin_orbitals
    b = obasis.get_basis(m.nBaths, m.nValBaths, m.dnValBaths, m.dnConBaths, m.dnTols, m.n0imps)
    rs = applyop.restriction_masks(n, m.restrictions)
    seeds = [s for s in b[::37] if applyop._passes(s, rs)][:12]
    ho, io = oham.expand_basis_and_hamiltonian(n, {}, m.hOp, seeds, m.restrictions)
    hg, ig = finite.expand_basis_and_hamiltonian(n, {}, m.hOp, ints_to_bytes(seeds, n), m.restrictions)
    assert bytes_to_ints(list(ig), n) == list(io)
    d = (hg - ho)
    assert (abs(d).max() if d.nnz else 0.0) <= 1e-12
    assert hg.nnz == ho.nnz


# ------------------------------------------------------------------ Krylov


def test_eigensystem_vs_dense(gpu):
    from impuritymodel_amd import finite
    from impuritymodel_amd.models import make_model

    g = np.load(os.path.join(GOLD, "ground_10bath_s7.npz"))
    m = make_model(n_bath=10, seed=int(g["seed"]))
    n = m.n_spin_orbitals
    b = finite.get_basis(m.nBaths, m.nValBaths, m.dnValBaths, m.dnConBaths, m.dnTols, m.n0imps)
    es, psis = finite.eigensystem(n, m.hOp, b, 3, eigenValueTol=1e-12)
    np.testing.assert_allclose(es, g["evals"], rtol=0, atol=1e-10)
    bo = obasis.get_basis(m.nBaths, m.nValBaths, m.dnValBaths, m.dnConBaths, m.dnTols, m.n0imps)
    _, po = okry.eigensystem(n, m.hOp, bo, 3, groundDiagMode="full")
    for pg, p in zip(psis, po):
        keys = bytes_to_ints(list(pg), n)
        assert keys == [k for k in bo if k in p]  # basis order, same truncation
        # eigenvectors agree up to a global phase (components of the
        # truncated dicts compared elementwise)
        vg = np.array([pg[psr.int2bytes(k, n)] for k in keys])
        vo = np.array([p[k] for k in keys])
        ov = np.vdot(vg, vo)
        np.testing.assert_allclose(vg * (ov / abs(ov)), vo, rtol=0, atol=1e-8)


@pytest.mark.parametrize("hy", [0.0, 2e-3])
def test_eigensolver_variants_vs_dense(gpu, hy):
    """Real (hy = 0) and complex (hy != 0) Hamiltonians; every launch shape,
    one or two Gram-Schmidt passes, real or complex Krylov vectors."""
    from impuritymodel_amd import engine
    from impuritymodel_amd.batched import ground_bases
    from impuritymodel_amd.models import make_model

    models = [make_model(n_bath=nb, seed=20 + nb, hField=(0.0, hy, 1e-4)) for nb in (10, 20, 30)]
    ops = engine.Ops([m.hOp for m in models], [m.restrictions for m in models], max(m.n_spin_orbitals for m in models))
    b = engine.Batch(ops, np.arange(3, dtype=np.int32))
    b.set_seeds(ground_bases(models, ops.W))
    b.build(fixed_basis=True)
    rowptr, col, val = b.csr()
    ref = []
    for c in range(3):
        r0, r1 = b.row_base[c], b.row_base[c + 1]
        n = r1 - r0
        dense = np.zeros((n, n), complex)
        for r in range(n):
            for e in range(rowptr[r0 + r], rowptr[r0 + r + 1]):
                dense[r, col[e]] = val[e]
        ref.append(np.linalg.eigvalsh(dense)[:3])
    try:
        for bt in (0, 256, 1024):
            for real in (0, 1):
                for dgks in (0, 1):
                    engine.set_tuning("eigen_bt", bt)
                    engine.set_tuning("eigen_real", real)
                    engine.set_tuning("eigen_dgks", dgks)
                    ev, _, steps, ok = b.eigen(3, tol=1e-12)
                    assert ok
                    for c in range(3):
                        np.testing.assert_allclose(ev[c], ref[c], rtol=0, atol=1e-10)
    finally:
        engine.set_tuning("eigen_bt", 0)
        engine.set_tuning("eigen_real", 1)
        engine.set_tuning("eigen_dgks", 1)


@pytest.mark.parametrize("hy", [0.0, 2e-3])
def test_eigensolver_lds_krylov_prefix_split(gpu, hy):
    """The Krylov basis split between LDS (first k vectors) and memory gives
    the same eigenpairs whatever k: none, 1, 5, 20, as many as fit; against
    dense diagonalization too."""
    from impuritymodel_amd import engine
    from impuritymodel_amd.batched import ground_bases
    from impuritymodel_amd.models import make_model

    models = [make_model(n_bath=nb, seed=30 + nb, hField=(0.0, hy, 1e-4)) for nb in (10, 20)]
    ops = engine.Ops([m.hOp for m in models], [m.restrictions for m in models], max(m.n_spin_orbitals for m in models))
    b = engine.Batch(ops, np.arange(2, dtype=np.int32))
    b.set_seeds(ground_bases(models, ops.W))
    b.build(fixed_basis=True)
    rowptr, col, val = b.csr()
    ref = []
    for c in range(2):
        r0, r1 = b.row_base[c], b.row_base[c + 1]
        dense = np.zeros((r1 - r0, r1 - r0), complex)
        for r in range(r1 - r0):
            for e in range(rowptr[r0 + r], rowptr[r0 + r + 1]):
                dense[r, col[e]] = val[e]
        ref.append(np.linalg.eigvalsh(dense)[:3])
    try:
        engine.set_tuning("eigen_lds_vectors", 0)
        ev0, vec0, st0, ok = b.eigen(3, tol=1e-12)
        assert ok
        for c in range(2):
            np.testing.assert_allclose(ev0[c], ref[c], rtol=0, atol=1e-10)
        for cap in (1, 5, 20, -1):
            engine.set_tuning("eigen_lds_vectors", cap)
            ev, vec, st, ok = b.eigen(3, tol=1e-12)
            assert ok and np.array_equal(st, st0)
            np.testing.assert_allclose(ev, ev0, rtol=0, atol=1e-13)
            np.testing.assert_allclose(vec, vec0, rtol=0, atol=1e-11)
    finally:
        engine.set_tuning("eigen_lds_vectors", -1)


def test_flat_kernels_match_workgroup_kernels(gpu):
    """The grid-wide (flat) Lanczos and eigensolver used for large sectors give
    the workgroup kernels' results; every SpMV lanes-per-row shape."""
    from impuritymodel_amd import engine
    from impuritymodel_amd.batched import ground_bases
    from impuritymodel_amd.models import make_model

    models = [make_model(n_bath=nb, seed=40 + nb, hField=(0.0, hy, 1e-4)) for nb, hy in ((20, 0.0), (30, 1e-3))]
    ops = engine.Ops([m.hOp for m in models], [m.restrictions for m in models], max(m.n_spin_orbitals for m in models))
    b = engine.Batch(ops, np.arange(2, dtype=np.int32))
    b.set_seeds(ground_bases(models, ops.W))
    b.build(fixed_basis=True)
    ref_ev, _, _, ok = b.eigen(3, tol=1e-12)
    assert ok
    rng = np.random.default_rng(1)
    v0 = rng.normal(size=b.rows) + 1j * rng.normal(size=b.rows)
    for c in range(2):
        r0, r1 = b.row_base[c], b.row_base[c + 1]
        v0[r0:r1] /= np.linalg.norm(v0[r0:r1])
    ra, rb, rk, rs = b.lanczos(12, v0=v0)
    try:
        engine.set_tuning("flat_min_n", 0)
        for lpr in (4, 8, 16, 32, 64):
            engine.set_tuning("spmv_lpr", lpr)
            fa, fb, fk, fs = b.lanczos(12, v0=v0)
            assert np.array_equal(fk, rk) and np.array_equal(fs, rs)
            np.testing.assert_allclose(fa, ra, rtol=1e-11, atol=0)
            np.testing.assert_allclose(fb[:, :11], rb[:, :11], rtol=1e-10, atol=0)
        for real in (0, 1):
            engine.set_tuning("eigen_real", real)
            ev, evecs, steps, ok = b.eigen(3, tol=1e-12)
            assert ok
            np.testing.assert_allclose(ev, ref_ev, rtol=0, atol=1e-10)
    finally:
        engine.set_tuning("flat_min_n", 32768)
        engine.set_tuning("spmv_lpr", 8)
        engine.set_tuning("eigen_real", 1)


def test_flat_lanczos_graph_replay_matches_direct(gpu):
    """The grid-wide Lanczos replayed as HIP graphs of G steps (with a direct
    remainder) gives bit-identical alpha/beta to direct launches, for several
    G and two chains that reuse one instantiated graph; a breakdown inside a
    replayed graph is flagged at the same step as with direct launches."""
    from impuritymodel_amd import engine, finite
    from impuritymodel_amd.batched import ground_bases
    from impuritymodel_amd.models import make_model

    models = [make_model(n_bath=nb, seed=40 + nb, hField=(0.0, hy, 1e-4)) for nb, hy in ((20, 0.0), (30, 1e-3))]
    ops = engine.Ops([m.hOp for m in models], [m.restrictions for m in models], max(m.n_spin_orbitals for m in models))
    b = engine.Batch(ops, np.arange(2, dtype=np.int32))
    b.set_seeds(ground_bases(models, ops.W))
    b.build(fixed_basis=True)
    rng = np.random.default_rng(2)
    v0 = rng.normal(size=b.rows) + 1j * rng.normal(size=b.rows)
    for c in range(2):
        r0, r1 = b.row_base[c], b.row_base[c + 1]
        v0[r0:r1] /= np.linalg.norm(v0[r0:r1])
    K = 45
    wa, wb, wk, ws = b.lanczos(K, v0=v0)  # workgroup kernels
    try:
        engine.set_tuning("flat_min_n", 0)
        engine.set_tuning("lanczos_graph_steps", 0)
        da, db, dk, ds = b.lanczos(K, v0=v0)
        np.testing.assert_allclose(da, wa, rtol=1e-11, atol=0)
        np.testing.assert_allclose(db[:, :K - 1], wb[:, :K - 1], rtol=1e-10, atol=0)
        for G in (16, 7, 1, 45):
            engine.set_tuning("lanczos_graph_steps", G)
            for _ in range(2):  # the second call replays the cached graph
                ga, gb_, gk, gs = b.lanczos(K, v0=v0)
                assert np.array_equal(ga, da) and np.array_equal(gb_, db)
                assert np.array_equal(gk, dk) and np.array_equal(gs, ds)
        engine.set_tuning("lanczos_graph_steps", 1)
        d = scipy.sparse.csr_matrix(np.diag([1.0, 2.0, 3.0]).astype(complex))
        with pytest.raises(ValueError, match="beta==0"):
            finite.get_tridiagonal_krylov_vectors(d, np.array([1, 0, 0], complex), 3)
        a3, b3 = finite.get_tridiagonal_krylov_vectors(d, np.ones(3, complex) / np.sqrt(3), 3)
        ao, bo = okry.lanczos(d, np.ones(3, complex) / np.sqrt(3), 3)
        np.testing.assert_allclose(a3, ao, atol=1e-12)
        np.testing.assert_allclose(b3, bo, atol=1e-12)
    finally:
        engine.set_tuning("flat_min_n", 32768)
        engine.set_tuning("lanczos_graph_steps", 16)


def test_lanczos_random_hermitian_and_breakdown(gpu):
    from impuritymodel_amd import finite

    rng = np.random.default_rng(5)
    n = 60
    a = scipy.sparse.random(n, n, density=0.1, random_state=1) + 1j * scipy.sparse.random(n, n, density=0.1,
                                                                                         random_state=2)
    h = scipy.sparse.csr_matrix((a + a.conj().T) * 0.5 + scipy.sparse.diags(rng.uniform(-2, 2, n)))
    psi = rng.normal(size=n) + 1j * rng.normal(size=n)
    psi /= np.linalg.norm(psi)
    for K in (1, 2, 15):
        ag, bg = finite.get_tridiagonal_krylov_vectors(h, psi, K)
        ao, bo = okry.lanczos(h, psi, K)
        np.testing.assert_allclose(ag, ao, atol=1e-12)
        np.testing.assert_allclose(bg, bo, atol=1e-12)
    ag, bg = finite.get_tridiagonal_krylov_vectors(h, psi, n, reorth="full")
    ao, bo = okry.lanczos(h, psi, n, reorth="full")
    np.testing.assert_allclose(ag, ao, atol=1e-9)
    np.testing.assert_allclose(bg, bo, atol=1e-9)
    d = scipy.sparse.csr_matrix(np.diag([1.0, 2.0, 3.0]).astype(complex))
    with pytest.raises(ValueError, match="beta==0"):
        finite.get_tridiagonal_krylov_vectors(d, np.array([1, 0, 0], complex), 3)


def test_green_function_vs_oracle(gpu):
    from impuritymodel_amd import spectra
    from impuritymodel_amd.models import make_model

    m = make_model(n_bath=10, seed=2)
    n = m.n_spin_orbitals
    b = obasis.get_basis(m.nBaths, m.nValBaths, m.dnValBaths, m.dnConBaths, m.dnTols, m.n0imps)
    rs = applyop.restriction_masks(n, m.restrictions)
    psi = {s: 1.0 / np.sqrt(3) for s in [x for x in b if applyop._passes(x, rs)][:3]}
    w = np.linspace(-10, 10, 301)
    go = ospec.get_green(n, -700.0, psi, m.hOp, w, 0.3, 500, 1e-7, m.restrictions, reorth="full")
    gg = spectra.getGreen(n, -700.0, dict(zip(ints_to_bytes(list(psi), n), psi.values())), m.hOp, w, 0.3, 500, 1e-7,
                          m.restrictions, reorth="full")
    np.testing.assert_allclose(gg, go, rtol=0, atol=1e-8 * np.abs(go).max())


# ------------------------------------------------------------------ spectra


def test_xas_batched_vs_golden_and_plain(gpu):
    from impuritymodel_amd.batched import plain_xas, xas_batch
    from impuritymodel_amd.models import make_model

    g = np.load(os.path.join(GOLD, "ground_10bath_s7.npz"))
    m = make_model(n_bath=10, seed=int(g["seed"]))
    w = g["omega"]
    res = xas_batch([m], omega=w, krylovSize=200, nPsiMax=3, reorth=1)
    scale = np.abs(g["xas_sum"]).max()
    np.testing.assert_allclose(res.spectra[0], g["xas_sum"], rtol=0, atol=1e-8 * scale)
    plain = plain_xas(m, omega=w, krylovSize=200, nPsiMax=3, reorth=1)
    np.testing.assert_allclose(plain, res.spectra[0], rtol=0, atol=1e-10 * scale)


def test_photoemission_spectra_vs_oracle(gpu):
    from impuritymodel_amd import finite, spectra
    from impuritymodel_amd.models import make_model
    from impuritymodel_amd.operators import getInversePhotoEmissionOperators, getPhotoEmissionOperators

    m = make_model(n_bath=10, seed=4)
    n = m.n_spin_orbitals
    bo = obasis.get_basis(m.nBaths, m.nValBaths, m.dnValBaths, m.dnConBaths, m.dnTols, m.n0imps)
    es, po = okry.eigensystem(n, m.hOp, bo, 2, groundDiagMode="full")
    w = np.linspace(-15, 15, 301)
    tps = getPhotoEmissionOperators(m.nBaths, l=2)[:3]
    tips = getInversePhotoEmissionOperators(m.nBaths, l=2)[:3]
    gs_o, _, _ = ospec.photoemission(n, m.hOp, tps, tips, po, es, w, 0.2, m.restrictions, reorth="full")
    pg = [dict(zip(ints_to_bytes(list(p), n), p.values())) for p in po]
    gs_g = -spectra.getSpectra(n, m.hOp, tps, pg, es, -w, -0.2, m.restrictions, reorth="full")
    gs_g = gs_g + spectra.getSpectra(n, m.hOp, tips, pg, es, w, 0.2, m.restrictions, reorth="full")
    scale = np.abs(gs_o).max()
    np.testing.assert_allclose(gs_g, gs_o, rtol=0, atol=1e-8 * scale)


def test_shifted_solve_vs_dense(gpu):
    """imped_batch_solve: (z - H) x = y for many shifts per chain against a
    dense solve; chains of different sizes in one launch."""
    from impuritymodel_amd import engine

    rng = np.random.default_rng(5)
    sizes = [1, 7, 60, 300]
    mats = []
    for n in sizes:
        a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        mats.append(0.5 * (a + a.conj().T) + np.diag(rng.uniform(-5, 5, n)))
    rowptr, col, val = [0], [], []
    for h in mats:
        for r in range(h.shape[0]):
            nz = np.nonzero(h[r])[0]
            col.extend(nz.tolist())
            # the library stores row r as the conjugate of H's column r
            val.extend(np.conj(h[nz, r]).tolist())
            rowptr.append(len(col))
    ops = engine.Ops([{((0, "c"), (0, "a")): 1.0}], [None], 8)
    b = engine.Batch(ops, np.zeros(len(sizes), np.int32))
    # right-hand sides are the seed amplitudes, placed on each chain's first rows
    ys = [rng.normal(size=n) + 1j * rng.normal(size=n) for n in sizes]
    keys = [np.arange(n, dtype=np.uint64).reshape(-1, 1) for n in sizes]
    b.set_seeds(keys, ys)
    b.set_csr(np.array(sizes), np.array(rowptr), np.array(col), np.array(val))
    S = 5
    shifts = rng.uniform(-4, 4, (len(sizes), S)) + 1j * rng.choice([0.3, -0.5, 1.0], (len(sizes), S))
    steps, resid, ok, X = b.solve(shifts, tol=1e-12, max_steps=400, want_x=True)
    assert ok and np.all(resid <= 1e-12)
    assert np.all(steps >= 1) and np.all(steps <= np.array(sizes))
    for c, (h, y) in enumerate(zip(mats, ys)):
        n = len(y)
        for s in range(S):
            x = np.linalg.solve(shifts[c, s] * np.eye(n) - h, y)
            base = int(b.row_base[c]) * S + s * n
            np.testing.assert_allclose(X[base:base + n], x, rtol=0, atol=1e-10 * np.abs(x).max())
    # a batch given only a CSR has no basis keys to attach to the solutions
    with pytest.raises(Exception, match="no basis keys"):
        b.vec_from_eigen(np.zeros(1, np.int32), np.zeros(1, np.int32), 0.0)


@pytest.mark.parametrize("alias", [False, True])
def test_rixs_map_vs_oracle(gpu, alias):
    from impuritymodel_amd import spectra
    from impuritymodel_amd.models import make_model
    from impuritymodel_amd.operators import getDaggeredDipoleOperators, getDipoleOperators

    m = make_model(n_bath=10, seed=6)
    n = m.n_spin_orbitals
    bo = obasis.get_basis(m.nBaths, m.nValBaths, m.dnValBaths, m.dnConBaths, m.dnTols, m.n0imps)
    es, po = okry.eigensystem(n, m.hOp, bo, 2, groundDiagMode="full")
    eps = [[1, 0, 0], [0, 1, 0], [0, 0, 1]]
    t_in = getDipoleOperators(m.nBaths, eps)
    t_out = getDaggeredDipoleOperators(m.nBaths, eps)
    w_in = np.array([-2.0, 0.5, 3.0])
    w_loss = np.linspace(-2, 10, 121)
    gs_o = ospec.rixs_map(n, m.hOp, t_in, t_out, po, es, w_in, w_loss, 0.5, 0.1, m.restrictions, 150,
                          alias_memo=alias, reorth="full")
    pg = [dict(zip(ints_to_bytes(list(p), n), p.values())) for p in po]
    gs_g = spectra.getRIXSmap(n, m.hOp, t_in, t_out, pg, es, w_in, w_loss, 0.5, 0.1, m.restrictions, 150,
                              alias_memo=alias, reorth="full")
    scale = np.abs(gs_o).max()
    assert scale > 0
    np.testing.assert_allclose(gs_g, gs_o, rtol=0, atol=1e-7 * scale)


def test_static_observables_vs_oracle(gpu, capsys):
    from impuritymodel_amd import observables as gobs
    from impuritymodel_amd.models import make_model
    from oracle import observables as oobs

    m = make_model(n_bath=10, seed=8)
    n = m.n_spin_orbitals
    bo = obasis.get_basis(m.nBaths, m.nValBaths, m.dnValBaths, m.dnConBaths, m.dnTols, m.n0imps)
    es, po = okry.eigensystem(n, m.hOp, bo, 3, groundDiagMode="full")
    for p in po:
        pg = dict(zip(ints_to_bytes(list(p), n), p.values()))
        for name in ("getLz3d", "getSz3d", "getSsqr3d", "getLsqr3d", "getTraceDensityMatrix"):
            assert abs(getattr(gobs, name)(m.nBaths, pg) - getattr(oobs, name)(m.nBaths, p)) < 1e-11, name
        rg, ro = gobs.getDensityMatrix(m.nBaths, pg), oobs.getDensityMatrix(m.nBaths, p)
        assert set(rg) == set(ro)
        assert max(abs(rg[k] - ro[k]) for k in ro) < 1e-12
        np.testing.assert_allclose(gobs.getEgT2gOccupation(m.nBaths, pg), oobs.getEgT2gOccupation(m.nBaths, p),
                                   rtol=0, atol=1e-12)
        cg, co = gobs.getDensityMatrixCubic(m.nBaths, pg), oobs.getDensityMatrixCubic(m.nBaths, p)
        assert set(cg) == set(co) and max(abs(cg[k] - co[k]) for k in co) < 1e-12
    pgs = [dict(zip(ints_to_bytes(list(p), n), p.values())) for p in po]
    gobs.printExpValues(m.nBaths, np.array(es), pgs)
    gobs.printThermalExpValues(m.nBaths, np.array(es), pgs)
    out = capsys.readouterr().out
    assert "E0 =" in out and "<L^2(3d)>" in out


def test_batched_equals_single_chain_runs(gpu):
    """Many mixed-size models (W = 1 and 2) in one batch reproduce the per-model
    runs: same chain bases, CSR and spectra."""
    from impuritymodel_amd.batched import xas_batch
    from impuritymodel_amd.models import batched_params, model_from_params

    P = batched_params(24, seed=3)
    models = [model_from_params(p) for p in P]
    w = np.linspace(-25, 25, 300)
    big = xas_batch(models, omega=w, krylovSize=200, reorth=1)
    for k in (0, 5, 11, 17, 23):
        one = xas_batch([models[k]], omega=w, krylovSize=200, reorth=1)
        scale = np.abs(one.spectra[0]).max()
        np.testing.assert_allclose(big.spectra[k], one.spectra[0], rtol=0, atol=1e-10 * scale)


def test_small_sort_matches_digit_passes(gpu):
    """The one-workgroup sort of small key sets orders BFS layers and vector
    entries exactly as the digit passes do: identical chain sizes, CSR and
    spectra with it on and off (W = 1 and W = 2 models in one batch)."""
    from impuritymodel_amd import engine
    from impuritymodel_amd.batched import xas_batch
    from impuritymodel_amd.models import batched_params, model_from_params

    models = [model_from_params(p) for p in batched_params(12, seed=8)][:6]
    out = {}
    try:
        for flag in (0, 1):
            engine.set_tuning("small_sort", flag)
            r = xas_batch(models, omega=np.linspace(-10, 20, 400), delta=0.3, krylovSize=60, reorth=1)
            out[flag] = r
    finally:
        engine.set_tuning("small_sort", 1)
    assert np.array_equal(out[0].chain_dims, out[1].chain_dims)
    assert np.array_equal(out[0].chain_nnz, out[1].chain_nnz)
    assert np.array_equal(out[0].spectra, out[1].spectra)
