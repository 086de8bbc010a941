"""Sector cache of the XAS model family against the general path and the oracle.

The general path (``batched.xas_batch_ops``) rebuilds every sector per
parameter set; the cache (``family_sectors.FamilySectors``) forms only the
values (``imped_batch_set_affine``) and the start vectors
(``imped_batch_seed_affine``) on fixed patterns.  Same spectra: the two
differ by a permutation of each excited basis, a rounding-level change with
fully reorthogonalized chains (bar 1e-10 of the peak), and both equal the
oracle's fully reorthogonalized XAS at 1e-8 of the peak.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _thetas(n, seed, T=None):
    from impuritymodel_amd import surrogate as S

    rng = np.random.default_rng(seed)
    lo, hi = S.PARAM_RANGES[:, 0], S.PARAM_RANGES[:, 1]
    th = lo + (hi - lo) * rng.random((n, 8))
    if T is not None:
        th[:, 7] = T
    return th


@pytest.fixture(scope="module")
def family(gpu):
    from impuritymodel_amd import surrogate as S

    fam = S.XASFamily()
    assert fam.sectors(1) is not None
    return fam


def test_cache_sectors_are_the_general_path_sectors(family):
    """Ground components partition the ground basis, and the excited chains of
    the general path have the sizes of the cached excited sectors, chain by
    chain."""
    from impuritymodel_amd import surrogate as S
    from impuritymodel_amd.batched import xas_batch_ops

    cache = family.sectors(1)
    n0 = len(cache.ground_basis)
    assert cache.comp_n.sum() == n0
    assert sorted(np.concatenate(cache.comp_rows).tolist()) == list(range(n0))
    for b, n in zip(cache.sector_basis, cache.sector_n):
        assert len(set(map(bytes, b))) == n
    th = _thetas(6, 21)
    ops_h, ops_t = family.batch_ops(th, 1)
    gen = xas_batch_ops(ops_h, ops_t, [cache.ground_basis] * len(th), S.OMEGA, delta=S.DELTA, T=th[:, 7])
    fast = cache.xas(th, S.OMEGA, delta=S.DELTA, T=th[:, 7])
    assert np.array_equal(fast.chain_dims, gen.chain_dims)
    assert np.array_equal(fast.chain_nnz, gen.chain_nnz)
    print("ground components", cache.comp_n.tolist(), "excited sectors", cache.sector_n.tolist(),
          "cache build %.3f s" % cache.build_s)


@pytest.mark.parametrize("T", [300.0, 10.0, None])
def test_cache_path_equals_general_path(family, T):
    from impuritymodel_amd import surrogate as S
    from impuritymodel_amd.batched import xas_batch_ops

    cache = family.sectors(1)
    th = _thetas(16, 5, T)
    ops_h, ops_t = family.batch_ops(th, 1)
    gen = xas_batch_ops(ops_h, ops_t, [cache.ground_basis] * len(th), S.OMEGA, delta=S.DELTA, T=th[:, 7])
    fast = cache.xas(th, S.OMEGA, delta=S.DELTA, T=th[:, 7])
    for m in range(len(th)):
        assert len(fast.energies[m]) == len(gen.energies[m])
        np.testing.assert_allclose(fast.energies[m], gen.energies[m], rtol=0, atol=1e-10)
    scale = np.abs(gen.spectra).max(axis=1, keepdims=True)
    err = (np.abs(fast.spectra - gen.spectra) / scale).max()
    print("cache vs general path: %.3g of the peak" % err)
    assert err <= 1e-10


def test_one_call_equals_entry_point_sequence(family):
    """imped_family_xas (one call, state selection in C++) against the same
    composition from the batch entry points with the selection in Python:
    same kernels on the same inputs, so the same spectra."""
    from impuritymodel_amd import surrogate as S

    cache = family.sectors(1)
    for T in (300.0, 10.0):
        th = _thetas(12, 9, T)
        one = cache.xas(th, S.OMEGA, delta=S.DELTA, T=th[:, 7])
        seq = cache.xas_steps(th, S.OMEGA, delta=S.DELTA, T=th[:, 7])
        for m in range(len(th)):
            assert one.energies[m] == seq.energies[m]
        assert np.array_equal(one.chain_dims, seq.chain_dims)
        assert np.array_equal(one.kout, seq.kout)
        assert np.array_equal(one.eig_steps, seq.eig_steps)
        err = np.abs(one.spectra - seq.spectra).max() / np.abs(seq.spectra).max()
        print("one call vs entry-point sequence (T=%g): %.3g of the peak" % (T, err))
        assert err <= 1e-14


def test_cache_path_vs_oracle(family):
    """exact_spectra (cache path) against the oracle at 1e-8 of the peak."""
    from oracle import basis as obasis
    from oracle import krylov as okry
    from oracle import spectra as ospec

    from impuritymodel_amd import surrogate as S
    from impuritymodel_amd.operators import getDipoleOperators

    th = _thetas(2, 17)
    got, res = S.exact_spectra(family, th)
    assert res.timings["expand_build"] == 0.0  # the cache path ran
    tops = getDipoleOperators(family.nBaths, ([1, 0, 0], [0, 1, 0], [0, 0, 1]))
    b = obasis.get_basis(family.nBaths, family.nValBaths, family.dnValBaths, family.dnConBaths, family.dnTols,
                         family.n0imps)
    for i in range(len(th)):
        hOp = family.hop(th[i])
        es, po = okry.eigensystem(26, hOp, b, 3, groundDiagMode="full")
        es = ospec.energy_cut(tuple(es), T=th[i, 7])
        _, _, ref = ospec.xas(26, hOp, tops, po[:len(es)], es, S.OMEGA, S.DELTA, family.restrictions, T=th[i, 7],
                              krylovSize=200, reorth="full")
        ref = S.area_normalize(ref[None])[0]
        np.testing.assert_allclose(got[i], ref, rtol=0, atol=1e-8 * np.abs(ref).max())


def test_affine_entry_points_reject_bad_maps(family):
    from impuritymodel_amd._lib import IMPED_ERR_INVALID, ImpedError
    from impuritymodel_amd.engine import Batch

    cache = family.sectors(1)
    b = Batch(cache._dummy, np.zeros(2, np.int32))
    with pytest.raises(ImpedError) as e:
        b.set_affine(cache.ground, [0, cache.ground.n.size], [0, 0], np.zeros((1, cache.P)))
    assert e.value.code == IMPED_ERR_INVALID
    with pytest.raises(ImpedError):
        b.set_affine(cache.trans, [0, 0], [0, 0], np.zeros((1, 0)))  # not square
    b.set_affine(cache.ground, [0, 1], [0, 0], np.zeros((1, cache.P)))
    x = Batch(cache._dummy, np.zeros(1, np.int32))
    x.set_affine(cache.sector, [cache.sector_index[0, 0]], [0], np.zeros((1, cache.P)))
    with pytest.raises(ImpedError):  # the ground batch has no eigenvectors yet
        x.seed_affine(b, cache.trans, [0], [0], [cache.sector_index[0, 0]])
    b.eigen(3, want_vectors=False)
    with pytest.raises(ImpedError):  # eigenvector index beyond nev
        x.seed_affine(b, cache.trans, [0], [3], [cache.sector_index[0, 0]])
    n2 = x.seed_affine(b, cache.trans, [0], [0], [cache.sector_index[0, 0]])
    assert n2[0] > 0
