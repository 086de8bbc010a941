"""The engine's device radix sort against numpy's stable argsort."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 255, 2047, 2048, 2049, 2095, 10000, 300001])
def test_radix_sort_stable(gpu, n):
    from impuritymodel_amd.engine import sort_u64

    rng = np.random.default_rng(n)
    for begin, end, dup in ((0, 64, False), (24, 64, False), (24, 64, True), (0, 8, True), (62, 64, True)):
        keys = rng.integers(0, 2**63, size=n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, size=n,
                                                                                        dtype=np.uint64)
        if dup:
            keys = keys[rng.integers(0, max(1, n // 7), size=n)]
        mask = np.uint64(((1 << (end - begin)) - 1) << begin) if end - begin < 64 else np.uint64(2**64 - 1)
        ref = np.argsort(keys & mask, kind="stable")
        got = sort_u64(keys, begin, end)
        assert np.array_equal(got, ref), (n, begin, end, dup)
