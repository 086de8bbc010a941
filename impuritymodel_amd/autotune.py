"""Autotuner of the engine's launch shapes (no reference counterpart).

The Krylov kernels take their launch shape from ``imped_set_tuning`` knobs
chosen per chain-dimension bucket:

* ``lanczos_wave_max_n``: chains up to this size run the Lanczos recurrence
  in one wave (shuffle reductions, no barriers); larger ones get 256 or 1024
  threads;
* ``eigen_bt``, ``eigen_check_every``, ``eigen_wave_max_n``: threads per
  eigensolver chain, Ritz-check interval, one-wave bucket;
* ``spmv_lpr``: lanes per CSR row of the grid-wide SpMV (sectors above
  ``flat_min_n``).

``autotune`` times each candidate on the caller's own built batches (best of
``repeats`` runs, HIP-event kernel time), keeps the fastest, sets it, and
returns the table.  Every candidate computes the same result (the parity
tests run all of them); only the time differs.  ``save``/``load`` keep the
choices per device in a JSON file.
"""

import json
import os
import time

import numpy as np

from . import engine

DEFAULTS = {
    "lanczos_wave_max_n": 4096,
    "lanczos_wg256_max_n": 65536,
    "eigen_bt": 0,
    "eigen_check_every": 12,
    "eigen_wave_max_n": 0,
    "spmv_lpr": 8,
    "lanczos_graph_steps": 16,
    "eigen_lds_vectors": -1,
    "small_sort": 1,
    "mlp_graph": 0,
}

CANDIDATES = {
    "lanczos_wave_max_n": [512, 1024, 4096, 16384],
    "eigen_bt": [256, 512, 1024],
    "eigen_check_every": [4, 8, 12, 16],
    "eigen_wave_max_n": [0, 256, 1024],
    "spmv_lpr": [4, 8, 16, 32],
}


def _timed(fn, kernels, repeats):
    """Best-of-repeats kernel time (ms) of fn(), from the engine's launch timing."""
    best = np.inf
    for _ in range(repeats):
        engine.timing_reset()
        t0 = time.perf_counter()
        fn()
        wall = time.perf_counter() - t0
        ms = sum(engine.timing_query(k)[0] for k in kernels)
        best = min(best, ms if ms > 0 else wall * 1e3)
    return best


def _sweep(knob, values, fn, kernels, repeats, table):
    times = {}
    for v in values:
        engine.set_tuning(knob, v)
        times[v] = _timed(fn, kernels, repeats)
    best = min(times, key=times.get)
    engine.set_tuning(knob, best)
    table[knob] = dict(best=best, ms={str(k): round(t, 4) for k, t in times.items()})
    return best


def autotune(lanczos_batch=None, lanczos_K=200, eigen_batch=None, nev=3, large_batch=None, large_K=50,
             repeats=2, v0=None, large_v0=None):
    """Tune the knobs on built batches; returns {knob: {best, ms per candidate}}.

    lanczos_batch: a built Batch (chains of the workload's sizes; start
    vectors from its seeds or ``v0``); eigen_batch: a built fixed-basis Batch
    (ground-state sectors); large_batch: a built Batch whose chains exceed
    ``flat_min_n`` (start vectors from its seeds or ``large_v0``).  Launch
    timing is switched on while tuning and off afterwards.
    """
    table = {}
    engine.timing_enable(True)
    try:
        if lanczos_batch is not None:
            _sweep("lanczos_wave_max_n", CANDIDATES["lanczos_wave_max_n"],
                   lambda: lanczos_batch.lanczos(lanczos_K, v0=v0), ["lanczos"], repeats, table)
        if eigen_batch is not None:
            run = lambda: eigen_batch.eigen(nev, want_vectors=False, allow_noconv=True)  # noqa: E731
            for knob in ("eigen_bt", "eigen_check_every", "eigen_wave_max_n"):
                _sweep(knob, CANDIDATES[knob], run, ["eigen"], repeats, table)
        if large_batch is not None:
            _sweep("spmv_lpr", CANDIDATES["spmv_lpr"], lambda: large_batch.lanczos(large_K, v0=large_v0), ["spmv"],
                   repeats, table)
    finally:
        engine.timing_enable(False)
    return table


def apply(choices):
    """Set knobs from {knob: value} or an autotune table."""
    for k, v in choices.items():
        engine.set_tuning(k, int(v["best"] if isinstance(v, dict) else v))


def reset():
    apply(DEFAULTS)


def _path(path):
    return path or os.environ.get("IMPED_TUNING_FILE",
                                  os.path.join(os.path.expanduser("~"), ".cache", "impuritymodel_amd", "tuning.json"))


def save(table, key, path=None):
    """Store a tuning table under key (e.g. device and workload name)."""
    p = _path(path)
    os.makedirs(os.path.dirname(p), exist_ok=True)
    data = {}
    if os.path.exists(p):
        with open(p) as f:
            data = json.load(f)
    data[key] = table
    with open(p, "w") as f:
        json.dump(data, f, indent=1)
    return p


def load(key, path=None):
    """Apply a stored table; returns it (None when absent)."""
    p = _path(path)
    if not os.path.exists(p):
        return None
    with open(p) as f:
        table = json.load(f).get(key)
    if table:
        apply(table)
    return table
