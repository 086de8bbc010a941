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
