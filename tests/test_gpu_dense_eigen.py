"""Dense eigensolver (Householder tridiagonalization in LDS) on the GPU.

Chains of up to ``eigen_dense_max_n`` rows (at most 112) are solved by
``dense_eigen_kernel`` instead of the Lanczos eigensolver.  Checked against
numpy's ``eigvalsh`` (lowest eigenvalues to 1e-11 of the spectral scale),
by the residual |H x - lambda x| and the orthonormality of the returned
vectors (1e-10), on a degenerate block-diagonal chain (whose repeated lowest
eigenvalue a single Krylov space cannot hold twice), and on the exact
spectra of the fitting family (dense against Lanczos ground states, 1e-10 of
the peak).  The same reduction started from a reflector of the start
vector replaces the fully reorthogonalized Lanczos recurrence of small
chains (``lanczos_dense_max_n``): checked against the LDS-resident
recurrence and on the family spectra.
"""

import numpy as np
import pytest
import scipy.sparse

pytestmark = pytest.mark.gpu


def _hermitian(rng, n, density=0.35):
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    a[rng.random((n, n)) > density] = 0
    return a + a.conj().T + np.diag(rng.standard_normal(n))


def _batch(mats):
    from impuritymodel_amd.engine import Batch, Ops

    rowptr, cols, vals = [0], [], []
    for h in mats:
        h = scipy.sparse.csr_matrix(h)
        rowptr.extend((h.indptr[1:] + rowptr[-1]).tolist())
        cols.append(h.indices)
        vals.append(h.data)
    b = Batch(Ops([{}], [None], 1), np.zeros(len(mats), np.int32))
    b.set_csr([len(h) for h in mats], np.array(rowptr, np.int64), np.concatenate(cols).astype(np.int32),
              np.concatenate(vals).astype(np.complex128))
    return b


@pytest.fixture
def dense_on(gpu):
    from impuritymodel_amd import engine

    engine.set_tuning("eigen_dense_max_n", 112)
    try:
        yield
    finally:
        engine.set_tuning("eigen_dense_max_n", 0)


def test_dense_eigenpairs_vs_numpy(dense_on):
    rng = np.random.default_rng(21)
    sizes = [1, 2, 3, 7, 33, 64, 90, 100, 112, 113]  # 113 rows: the Lanczos eigensolver
    mats = [_hermitian(rng, n) for n in sizes]
    # two equal blocks: the lowest eigenvalue twice
    blk = _hermitian(rng, 20)
    mats.append(np.block([[blk, np.zeros((20, 20))], [np.zeros((20, 20)), blk]]))
    b = _batch(mats)
    nev = 3
    ev, vec, steps, ok = b.eigen(nev)
    assert ok
    for c, h in enumerate(mats):
        n = len(h)
        ref = np.linalg.eigvalsh(h)[:nev]
        scale = max(1.0, np.abs(np.linalg.eigvalsh(h)).max())
        ne = min(nev, n)
        np.testing.assert_allclose(ev[c, :ne], ref[:ne], rtol=0, atol=1e-11 * scale)
        if n <= 112:
            assert steps[c] == n  # the dense solver reports n
        X = np.stack([b.eigvec(vec, c, k) for k in range(ne)], axis=1)
        assert np.abs(h @ X - X * ev[c, :ne]).max() <= 1e-10 * scale
        assert np.abs(X.conj().T @ X - np.eye(ne)).max() <= 1e-10
    deg = ev[-1]
    assert abs(deg[0] - deg[1]) <= 1e-11 * max(1.0, abs(deg[0]))


def test_dense_ground_states_give_the_same_family_spectra(gpu):
    """exact_spectra with the family's ground components (90 and 100 rows)
    solved densely against the default Lanczos eigensolver."""
    from impuritymodel_amd import engine
    from impuritymodel_amd import surrogate as S

    fam = S.XASFamily()
    rng = np.random.default_rng(31)
    lo, hi = S.PARAM_RANGES[:, 0], S.PARAM_RANGES[:, 1]
    th = lo + (hi - lo) * rng.random((24, 8))
    th[:8, 7] = 10.0
    ref, _ = S.exact_spectra(fam, th)
    engine.set_tuning("eigen_dense_max_n", 112)
    try:
        got, _ = S.exact_spectra(fam, th)
    finally:
        engine.set_tuning("eigen_dense_max_n", 0)
    err = (np.abs(got - ref) / np.abs(ref).max(axis=1, keepdims=True)).max()
    print("dense vs Lanczos ground states: %.3g of the peak" % err)
    assert err <= 1e-10


@pytest.mark.parametrize("K", [112, 40])
def test_householder_lanczos_matches_lds_lanczos(gpu, K):
    """Fully reorthogonalized Lanczos of chains of up to 112 rows as a
    Householder reduction (``lanczos_dense_max_n``) against the LDS-resident
    recurrence, from unit start vectors: alpha and beta to 1e-10 of the
    spectral scale, the same kout and status; a zero start vector ends both
    at kout = 1 with a zero beta."""
    from impuritymodel_amd import engine

    rng = np.random.default_rng(41)
    sizes = [1, 2, 7, 33, 60, 64, 65, 97, 112, 10]
    mats = [_hermitian(rng, n) for n in sizes]
    b = _batch(mats)
    v0 = []
    for c, n in enumerate(sizes):
        v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        v0.append(np.zeros(n, complex) if c == len(sizes) - 1 else v / np.linalg.norm(v))
    v0 = np.concatenate(v0)
    a1, b1, k1, s1 = b.lanczos(K, reorth=1, v0=v0)
    engine.set_tuning("lanczos_dense_max_n", 112)
    try:
        a2, b2, k2, s2 = b.lanczos(K, reorth=1, v0=v0)
    finally:
        engine.set_tuning("lanczos_dense_max_n", 0)
    assert np.array_equal(k1, k2) and np.array_equal(s1, s2)
    assert k2[-1] == 1
    for c, h in enumerate(mats):
        kk = int(k1[c])
        scale = max(1.0, np.abs(np.linalg.eigvalsh(h)).max())
        np.testing.assert_allclose(a2[c, :kk], a1[c, :kk], rtol=0, atol=1e-10 * scale)
        nb = kk - 1 if kk == min(K, len(h)) else kk
        np.testing.assert_allclose(b2[c, :nb], b1[c, :nb], rtol=0, atol=1e-10 * scale)


def test_householder_lanczos_gives_the_same_family_spectra(gpu):
    """exact_spectra with the excited chains (60 rows, K = 60) reduced by
    Householder reflectors against the LDS-resident Lanczos."""
    from impuritymodel_amd import engine
    from impuritymodel_amd import surrogate as S

    fam = S.XASFamily()
    rng = np.random.default_rng(43)
    lo, hi = S.PARAM_RANGES[:, 0], S.PARAM_RANGES[:, 1]
    th = lo + (hi - lo) * rng.random((24, 8))
    th[:8, 7] = 10.0
    ref, _ = S.exact_spectra(fam, th)
    engine.set_tuning("lanczos_dense_max_n", 112)
    try:
        got, _ = S.exact_spectra(fam, th)
    finally:
        engine.set_tuning("lanczos_dense_max_n", 0)
    err = (np.abs(got - ref) / np.abs(ref).max(axis=1, keepdims=True)).max()
    print("Householder vs LDS Lanczos excited chains: %.3g of the peak" % err)
    assert err <= 1e-10
