"""GPU: the autotuner picks among its candidates, and the tuned shapes give
the same results as the defaults (eigenvalues 1e-10, Lanczos alpha 1e-10)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_autotune_choices_and_invariance(gpu, tmp_path):
    from impuritymodel_amd import autotune, engine
    from impuritymodel_amd.batched import ground_bases
    from impuritymodel_amd.engine import Batch, Ops
    from impuritymodel_amd.models import batched_params, model_from_params

    models = [model_from_params(p) for p in batched_params(10, seed=5) if p["n_bath"] <= 30][:5]
    n_orb = max(m.n_spin_orbitals for m in models)
    ops = Ops([m.hOp for m in models], [m.restrictions for m in models], n_orb)
    gb = Batch(ops, np.arange(len(models), dtype=np.int32))
    gb.set_seeds(ground_bases(models, ops.W))
    gb.build(fixed_basis=True)
    autotune.reset()
    ref_ev, _, _, ok = gb.eigen(3, want_vectors=False)
    assert ok
    rng = np.random.default_rng(0)
    v0 = rng.normal(size=gb.rows) + 1j * rng.normal(size=gb.rows)
    for c in range(gb.C):
        v0[gb.row_base[c]:gb.row_base[c + 1]] /= np.linalg.norm(v0[gb.row_base[c]:gb.row_base[c + 1]])
    ref_a, _, _, _ = gb.lanczos(30, v0=v0)
    try:
        engine.set_tuning("flat_min_n", 200)  # the larger ground sectors take the grid-wide path
        table = autotune.autotune(lanczos_batch=gb, lanczos_K=30, eigen_batch=gb, large_batch=gb, large_K=20,
                                  repeats=1, v0=v0, large_v0=v0)
        for knob, row in table.items():
            assert row["best"] in autotune.CANDIDATES[knob]
            assert len(row["ms"]) == len(autotune.CANDIDATES[knob]) and all(t > 0 for t in row["ms"].values())
        ev, _, _, ok = gb.eigen(3, want_vectors=False)
        assert ok
        np.testing.assert_allclose(ev, ref_ev, rtol=0, atol=1e-10)
        a, _, _, _ = gb.lanczos(30, v0=v0)
        np.testing.assert_allclose(a[:, :10], ref_a[:, :10], rtol=1e-10, atol=0)
        p = autotune.save(table, "test", path=str(tmp_path / "t.json"))
        autotune.reset()
        assert autotune.load("test", path=p) == table
    finally:
        engine.set_tuning("flat_min_n", 32768)
        autotune.reset()
