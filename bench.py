"""Benchmark: XAS spectra/s of batched exact diagonalization on one MI355X.

Workload (one "step"): the batched-ED configuration -- 128 NiO-like Anderson
impurity models with parameters drawn from a seed (10Dq U(0.3, 1.5), F0 U(5, 9),
charge transfer U(1, 8) eV, V U(0.5, 3), bath size from {10, 20, 30, 40, 50}),
each giving its thermal, polarization-summed XAS: ground basis, H build,
lowest 3 eigenstates (tol 1e-12), energy cut, 3 dipole polarizations x kept
eigenstates Lanczos chains of 200 steps, continued fractions on 2000 omega
points, Boltzmann average.  complex128 throughout.  The operator dicts are
built once (synthetic input); everything from operator compilation to the
spectra on the host is timed.

Prints one JSON line (see the contract in the repository docs) with a
roofline figure for the dominant kernel (HIP-event timed inside the timed
steps) and a CPU baseline: the oracle (pure Python/numpy restatement of the
reference) on a bounded sample of models, one model per worker process.
"""

import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "XAS spectra/s on 1 MI355X: batched+surrogate vs plain GPU ED @ L1 err≤1e-2"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_lanczos(n, nnz, k):
    """20 nnz + 4(n+1) + 48 n bytes per Lanczos step (complex128 values,
    int32 columns/rowptr, read v_j, v_{j-1}, write w)."""
    n = np.asarray(n, np.float64)
    nnz = np.asarray(nnz, np.float64)
    k = np.asarray(k, np.float64)
    return float(np.sum(k * (20 * nnz + 4 * (n + 1) + 48 * n)))


def algorithmic_bytes_eigen(n, nnz, steps, nev):
    """Per step j (j Krylov vectors stored): the Lanczos step bytes plus two
    classical Gram-Schmidt passes that each read the j stored vectors twice
    (dots, update): 64 j n; plus the Ritz vectors (nev x steps x n x 16)."""
    tot = 0.0
    for ni, zi, si in zip(n, nnz, steps):
        s = float(si)
        tot += s * (20 * zi + 4 * (ni + 1) + 48 * ni) + 64.0 * ni * s * (s + 1) / 2 + 16.0 * nev * s * ni
    return tot


def _oracle_xas(args):
    """One model's XAS with the CPU oracle (reference algorithm, ARPACK ground state)."""
    p, omega, K = args
    sys.path.insert(0, ROOT)
    from impuritymodel_amd.models import model_from_params
    from impuritymodel_amd.operators import getDipoleOperators
    from oracle import basis, krylov, spectra

    m = model_from_params(p)
    n = m.n_spin_orbitals
    b = basis.get_basis(m.nBaths, m.nValBaths, m.dnValBaths, m.dnConBaths, m.dnTols, m.n0imps)
    es, psis = krylov.eigensystem(n, m.hOp, b, 3, eigenValueTol=1e-12)
    es = spectra.energy_cut(tuple(es))
    tops = getDipoleOperators(m.nBaths, ([1, 0, 0], [0, 1, 0], [0, 0, 1]))
    _, _, s = spectra.xas(n, m.hOp, tops, psis[: len(es)], es, omega, 0.2, m.restrictions, krylovSize=K)
    return s


def cpu_baseline(params, omega, K, workers, budget_s=25.0):
    """Oracle spectra/s over a bounded sample: one model per worker."""
    sample = params[:workers]
    t0 = time.perf_counter()
    ctx = mp.get_context("spawn")
    with ctx.Pool(workers) as pool:
        pool.map(_oracle_xas, [(p, omega, K) for p in sample], chunksize=1)
    dt = time.perf_counter() - t0
    return dict(value=len(sample) / dt, unit="spectra/s", cores=workers, kind="port",
                sample="%d models of the same parameter distribution (bath sizes %s), one per worker process, "
                       "%.1f s wall" % (len(sample), sorted(p["n_bath"] for p in sample), dt))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--models", type=int, default=128)
    ap.add_argument("--omega", type=int, default=2000)
    ap.add_argument("--krylov", type=int, default=200)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-workers", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the engine runs on a single MI355X; "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    from impuritymodel_amd import engine
    from impuritymodel_amd.batched import xas_batch
    from impuritymodel_amd.models import batched_params, model_from_params

    params = batched_params(args.models, seed=0)
    models = [model_from_params(p) for p in params]
    omega = np.linspace(-25.0, 25.0, args.omega)

    for _ in range(args.warmup):
        res = xas_batch(models, omega=omega, krylovSize=args.krylov)
    torch.cuda.synchronize()
    engine.timing_enable(True)
    engine.timing_reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = xas_batch(models, omega=omega, krylovSize=args.krylov)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    kernels = {k: engine.timing_query(k) for k in ("eigen", "lanczos", "apply_count", "apply_fill", "cf")}
    engine.timing_enable(False)
    value = args.models * args.steps / dt

    # dominant kernel and its algorithmic bytes per launch
    dom = max(kernels, key=lambda k: kernels[k][0])
    ms_total, launches = kernels[dom]
    if dom == "eigen":
        from impuritymodel_amd.batched import ground_bases
        from impuritymodel_amd.engine import Batch, Ops

        # sizes of the ground-state sectors (identical every step)
        n_orb = max(m.n_spin_orbitals for m in models)
        gb_ops = Ops([m.hOp for m in models], [m.restrictions for m in models], n_orb)
        gb = Batch(gb_ops, np.arange(len(models), dtype=np.int32))
        gb.set_seeds(ground_bases(models, gb_ops.W))
        gb.build(fixed_basis=True)
        per_step = algorithmic_bytes_eigen(gb.chain_n, gb.chain_nnz, res.eig_steps, 3)
    elif dom == "lanczos":
        per_step = algorithmic_bytes_lanczos(res.chain_dims, res.chain_nnz, res.kout)
    else:
        per_step = None
    roofline = None
    if per_step is not None:
        launches_per_step = launches / args.steps
        per_launch = per_step / launches_per_step
        avg_ms = ms_total / launches
        achieved = per_launch / (avg_ms * 1e-3) / 1e9
        roofline = dict(bound="hbm", kernel=dom + "_kernel", achieved=round(achieved, 2), peak=HBM_PEAK_GBS,
                        unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                        avg_launch_ms=round(avg_ms, 4), algorithmic_bytes_per_launch=per_launch)

    cpu = None
    if not args.no_cpu_baseline:
        workers = args.cpu_workers or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(params, omega, args.krylov, workers)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "spectra/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded impurity-model parameters)",
        "config": {
            "workload": "batched exact-diagonalization XAS, %d NiO-like impurity models per step (bath sizes "
                        "10-50 spin-orbitals, 26-66 spin-orbitals), 3 polarizations x <=3 thermal states, "
                        "Lanczos %d steps, %d omega points, complex128; exact path (no surrogate)"
                        % (args.models, args.krylov, args.omega),
            "models_per_step": args.models,
            "chains_per_step": int(len(res.chain_dims)),
            "chain_dim_range": [int(res.chain_dims.min()), int(res.chain_dims.max())],
            "kernel_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in kernels.items()},
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
