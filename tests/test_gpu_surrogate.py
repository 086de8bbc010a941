"""GPU checks of the surrogate network kernels and the family model builder.

The BF16 matrix-core kernels are compared with the oracle's BF16 emulation
(the same rounding points, float64 sums): 99.9% of outputs within 1e-3 of
the largest value and all within 1e-2, gradients to 2e-2 relative
(Frobenius) per layer -- a transposed operand, a wrong row or a missing GELU
derivative gives O(1) errors.
"""

import numpy as np
import pytest

from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


def _layer_slices(dims):
    out, off = [], 0
    for l in range(len(dims) - 1):
        n = dims[l] * dims[l + 1] + dims[l + 1]
        out.append(slice(off, off + n))
        off += n
    return out


@pytest.mark.parametrize("dims,batch", [((8, 72, 40, 100), 37), ((8, 512, 512, 512, 1000), 256)])
def test_mlp_forward_and_gradients_vs_oracle(gpu, dims, batch):
    from impuritymodel_amd.surrogate import Ensemble

    E = 3
    net = Ensemble(E, dims, max_batch=64, seed=4)
    rng = np.random.default_rng(5)
    params = net.init_params(7)
    params[:, :] += rng.normal(0, 0.01, params.shape).astype(np.float32)  # non-zero biases
    net.set_params(params)
    x = rng.uniform(-1, 1, (batch, dims[0])).astype(np.float32)
    t = rng.normal(size=(batch, dims[-1])).astype(np.float32)
    y = net.forward(x)
    loss, g = net.gradients(x, t)
    for e in range(E):
        ref, _ = omlp.forward(params[e], dims, x, bf16=True)
        # FP32 (GPU) and float64 (oracle) pre-activations can round a hidden
        # activation to neighbouring BF16 values (2^-8 apart); such rare flips
        # move single outputs by a few 1e-3 of the largest
        diff = np.abs(y[e] - ref)
        scale = np.abs(ref).max()
        assert np.quantile(diff, 0.999) <= 1e-3 * scale
        assert diff.max() <= 1e-2 * scale
        lref, gref = omlp.loss_and_grad(params[e], dims, x, t, bf16=True)
        assert abs(loss[e] - lref) < 1e-3 * lref
        for sl in _layer_slices(dims):
            err = np.linalg.norm(g[e, sl] - gref[sl]) / np.linalg.norm(gref[sl])
            assert err < 2e-2, (e, sl, err)
        # padding past the last layer stays zero
        assert not np.any(g[e, _layer_slices(dims)[-1].stop:])


def test_mlp_training_reduces_loss(gpu):
    from impuritymodel_amd.surrogate import Ensemble

    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, (4096, 8)).astype(np.float32)
    A = 0.5 * rng.normal(size=(8, 64))
    t = np.sin(x @ A).astype(np.float32)
    net = Ensemble(2, (8, 128, 128, 64), max_batch=512, seed=1)
    losses = net.train(x, t, epochs=60, batch=512, lr=3e-3, lr_final=1e-4)
    assert losses.shape == (480, 2)
    assert np.all(losses[-20:].mean(axis=0) < 0.1 * losses[:5].mean(axis=0))
    # the network trained on the GPU predicts held-out points
    xv = rng.uniform(-1, 1, (512, 8)).astype(np.float32)
    yv = net.forward(xv)
    mse = ((yv - np.sin(xv @ A)[None]) ** 2).mean(axis=(1, 2))
    assert np.all(mse < 0.2 * np.var(np.sin(xv @ A)))


def test_family_spectra_match_assembled_models(gpu):
    """Family members (one matrix product per batch) give the spectra of the
    same models assembled term by term through the operator builders."""
    from impuritymodel_amd.batched import xas_batch
    from impuritymodel_amd.models import ImpurityModel
    from impuritymodel_amd.surrogate import DELTA, OMEGA, XASFamily, area_normalize, exact_spectra, sample_params

    fam = XASFamily()
    P = sample_params(6, seed=3)
    fast, _ = exact_spectra(fam, P, reorth=1)
    slow = []
    for p in P:
        m = ImpurityModel(10, fam.hop(p), fam.nBaths, fam.nValBaths, fam.n0imps, fam.dnTols, fam.dnValBaths,
                          fam.dnConBaths, fam.restrictions)
        r = xas_batch([m], omega=OMEGA, delta=DELTA, krylovSize=200, T=p[7], reorth=1)
        slow.append(area_normalize(r.spectra[0]))
    np.testing.assert_allclose(fast, np.array(slow), rtol=0, atol=1e-8 * np.abs(slow).max())


@pytest.mark.parametrize("E,D,H", [(5, 1000, 64), (3, 200, 64), (9, 1000, 64), (2, 1100, 64), (5, 1000, 512)])
def test_ensemble_predict_statistics_vs_numpy(gpu, E, D, H):
    """The device ensemble statistics and misfits equal the same quantities
    computed in numpy (trapezoid rule, np.trapezoid) from the members' raw
    outputs (two-launch forward pass).  The cases cover the register kernel
    (one and four values per thread), the workgroup kernel (more than 8
    members; more than 1024 outputs), and, with 512 hidden units, the
    prediction that forms the first hidden layer inside the second layer's
    launch."""
    from impuritymodel_amd.surrogate import Ensemble, trapezoid_weights

    omega = np.linspace(-5.0, 15.0, D)
    scale = float(omega[-1] - omega[0])
    dims = (8, H, H, D)
    net = Ensemble(E, dims, max_batch=64, seed=2)
    rng = np.random.default_rng(3)
    p = net.init_params(3)
    p += rng.normal(0, 0.02, p.shape).astype(np.float32)
    net.set_params(p)
    w = trapezoid_weights(omega)
    for B in (37, 5):
        x = rng.uniform(-1, 1, (B, 8)).astype(np.float32)
        y = np.maximum(net.forward(x).astype(np.float64) / scale, 0.0)
        members = y / np.trapezoid(y, omega, axis=-1)[..., None]
        mean = members.mean(axis=0)
        mean = mean / np.trapezoid(mean, omega, axis=-1)[:, None]
        dis = np.trapezoid(np.abs(members - mean[None]), omega, axis=-1).mean(axis=0)
        gm, gd = net.predict(x, w, scale)
        np.testing.assert_allclose(gm, mean, rtol=1e-12, atol=1e-14 * np.abs(mean).max())
        np.testing.assert_allclose(gd, dis, rtol=1e-11, atol=1e-15)
        target = np.abs(np.sin(omega)) + 0.1
        target /= np.trapezoid(target, omega)
        l1 = np.trapezoid(np.abs(mean - target[None]), omega, axis=-1)
        md, ml = net.misfit(x, w, scale, target)
        np.testing.assert_allclose(md, dis, rtol=1e-11, atol=1e-15)
        np.testing.assert_allclose(ml, l1, rtol=1e-11, atol=1e-15)
        md2, ml2, mm = net.misfit(x, w, scale, target, want_mean=True)
        assert np.array_equal(md2, md) and np.array_equal(ml2, ml) and np.array_equal(mm, gm)


def test_surrogate_rows_independent_of_batch(gpu):
    """A sample's predicted spectrum does not depend on the other samples of
    its batch (bit for bit), so re-predicting served parameters reproduces
    the spectra a fitting loop used."""
    from impuritymodel_amd.surrogate import OMEGA, SCALE, Ensemble, trapezoid_weights

    net = Ensemble(5, (8, 512, 512, 512, len(OMEGA)), max_batch=1024, seed=6)
    rng = np.random.default_rng(8)
    x = rng.uniform(-1, 1, (300, 8)).astype(np.float32)
    w = trapezoid_weights(OMEGA)
    mean, dis = net.predict(x, w, SCALE)
    rows = np.sort(rng.choice(300, 41, replace=False))
    m2, d2 = net.predict(x[rows], w, SCALE)
    assert np.array_equal(m2, mean[rows]) and np.array_equal(d2, dis[rows])


def test_predict_graph_replays_match_direct_launches(gpu):
    """The prediction runs as a replayed HIP graph when kernel timing is off
    and as direct launches when it is on: every replay, with other work and
    batch-size changes in between, equals the direct launches' results."""
    from impuritymodel_amd import engine
    from impuritymodel_amd.surrogate import OMEGA, SCALE, Ensemble, trapezoid_weights

    net = Ensemble(5, (8, 512, 512, 512, len(OMEGA)), max_batch=1024, seed=9)
    rng = np.random.default_rng(10)
    w = trapezoid_weights(OMEGA)
    target = np.abs(np.cos(OMEGA)) + 0.2
    target /= target @ w
    try:
        for B in (256, 37, 256, 256, 300):
            x = rng.uniform(-1, 1, (B, 8)).astype(np.float32)
            engine.timing_enable(True)
            d0, l0, m0 = net.misfit(x, w, SCALE, target, want_mean=True)
            engine.timing_enable(False)
            net.forward(x[:5])  # unrelated work between the replays
            d1, l1, m1 = net.misfit(x, w, SCALE, target, want_mean=True)
            d2, l2 = net.misfit(x, w, SCALE, target)
            assert np.array_equal(d0, d1) and np.array_equal(l0, l1) and np.array_equal(m0, m1)
            assert np.array_equal(d0, d2) and np.array_equal(l0, l2)
    finally:
        engine.timing_enable(False)


def test_fitting_loop_device_misfit_matches_host(gpu):
    """The fitting loop's decisions from device misfits follow the same walk
    as decisions from spectra read back and compared on the host."""
    from impuritymodel_amd import surrogate as S

    fam = S.XASFamily()
    sur = S.Surrogate(fam, S.Ensemble(5, (8, 512, 512, 512, len(S.OMEGA)), max_batch=1024, seed=0))
    sur.threshold = np.inf  # every proposal served by the (untrained) network
    target = np.abs(np.sin(S.OMEGA)) + 0.1
    runs = []
    for keep in (True, False):
        loop = S.FittingLoop(fam, target, sur, n_walkers=64, seed=3, keep_spectra=keep)
        for _ in range(6):
            loop.step()
        loop.finish()
        runs.append((loop.theta.copy(), loop.cur.copy(), loop.served()))
    (t1, c1, (p1, s1)), (t2, c2, (p2, s2)) = runs
    assert np.array_equal(t1, t2) and np.array_equal(p1, p2)
    np.testing.assert_allclose(c2, c1, rtol=1e-12, atol=0)
    assert np.array_equal(s1, s2)


def test_native_walkers_follow_python_loop(gpu):
    """The library's walkers (one call per step) take the walk of the Python
    fitting loop driven by the same random stream: the same proposals served,
    waiting and accepted, the same exact flushes, the same final states."""
    from walker_rng import WalkerStreams

    from impuritymodel_amd import surrogate as S

    fam = S.XASFamily()
    sur = S.Surrogate(fam, S.Ensemble(5, (8, 512, 512, 512, len(S.OMEGA)), max_batch=1024, seed=0))
    _, dis = sur.predict(S.sample_params(256, seed=7))
    sur.threshold = float(np.median(dis))  # about half the proposals wait for exact ED
    target = np.abs(np.sin(S.OMEGA)) + 0.1
    py = S.FittingLoop(fam, target, sur, n_walkers=48, seed=5, flush_at=12, rng=WalkerStreams(5))
    nat = S.NativeFittingLoop(fam, target, sur, n_walkers=48, seed=5, flush_at=12)
    for _ in range(6):
        py.step()
        nat.step()
    py.finish()
    nat.finish()
    assert (nat.n_surrogate, nat.n_exact, nat.n_flushes, nat.n_predicted) == \
        (py.n_surrogate, py.n_exact, py.n_flushes, py.n_predicted)
    assert nat.n_surrogate > 0 and nat.n_flushes > 1
    # Box-Muller through numpy's and the C library's log and cos may differ in the last bit
    np.testing.assert_allclose(nat.theta, py.theta, rtol=1e-12, atol=0)
    np.testing.assert_allclose(nat.cur, py.cur, rtol=1e-9, atol=1e-15)
    assert abs(nat.best - py.best) <= 1e-9 * py.best
    (pp, ps), (np_, ns) = py.served(), nat.served()
    np.testing.assert_allclose(np_, pp, rtol=1e-12, atol=0)
    np.testing.assert_allclose(ns, ps, rtol=1e-5, atol=1e-7 * np.abs(ps).max())


def test_affine_family_operators_on_device(gpu):
    """Operators of a parameter batch formed on the device (c0 + theta @ cmat)
    give the spectra of per-model compiled tables; a second, smaller batch
    reuses the cached operators; set_coef replaces values in place."""
    from impuritymodel_amd.batched import xas_batch
    from impuritymodel_amd.engine import Batch
    from impuritymodel_amd.surrogate import DELTA, OMEGA, XASFamily, area_normalize, exact_spectra, sample_params

    fam = XASFamily()
    for n, seed in ((20, 4), (9, 5)):
        P = sample_params(n, seed=seed)
        fast, _ = exact_spectra(fam, P, reorth=1)
        r = xas_batch(fam.models(P), omega=OMEGA, delta=DELTA, krylovSize=200, T=P[:, 7], reorth=1)
        ref = area_normalize(r.spectra)
        np.testing.assert_allclose(fast, ref, rtol=0, atol=1e-9 * np.abs(ref).max())
    # set_coef: the operators of the last batch, values rewritten from the host
    P = sample_params(3, seed=6)
    ops_h, _ = fam.batch_ops(P)
    tab = fam.table(1)
    cf = fam.coef(P)
    vals = np.concatenate([cf[i % 3][tab.src] for i in range(ops_h.M)])
    ops_h.set_coef(vals)
    b = Batch(ops_h, np.arange(3, dtype=np.int32))
    b.set_seeds([fam._bases[1]] * 3)
    b.build(fixed_basis=True)
    ev, _, _, ok = b.eigen(2, want_vectors=False)
    b2 = Batch(fam.batch_ops(P)[0], np.arange(3, dtype=np.int32))
    b2.set_seeds([fam._bases[1]] * 3)
    b2.build(fixed_basis=True)
    ev2, _, _, ok2 = b2.eigen(2, want_vectors=False)
    assert ok and ok2
    np.testing.assert_allclose(ev, ev2, rtol=0, atol=1e-10)
    with pytest.raises(Exception):
        ops_h.set_coef(vals[:-1])
