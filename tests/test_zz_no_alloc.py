"""imped_lanczos_noalloc (the allocation-free Lanczos entry point) against
imped_batch_lanczos on the same CSR and start vector.

The two entry points run the same kernel for chains up to the grid-wide
threshold, so alpha, beta, kout and status must agree bit for bit.  Two calls
in a row, the second on a smaller chain, check that the shared scratch array
carries nothing over; a chain beyond the scratch capacity is refused.
"""

import ctypes as C

import numpy as np
import pytest
import scipy.sparse

pytestmark = pytest.mark.gpu


def _hermitian(n, seed, density):
    rng = np.random.default_rng(seed)
    a = scipy.sparse.random(n, n, density=density, random_state=rng, data_rvs=lambda k: rng.standard_normal(k))
    b = scipy.sparse.random(n, n, density=density, random_state=rng, data_rvs=lambda k: rng.standard_normal(k))
    m = (a + 1j * b).tocsr()
    h = (m + m.conj().T + scipy.sparse.diags(rng.standard_normal(n) * 3)).tocsr()
    h.sort_indices()
    v0 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    return h, v0 / np.linalg.norm(v0)


def _noalloc(h, v0, K, reorth):
    import torch

    from impuritymodel_amd._lib import check, lib

    dev = torch.device("cuda:0")
    rowptr = torch.from_numpy(h.indptr.astype(np.int64)).to(dev)
    col = torch.from_numpy(h.indices.astype(np.int32)).to(dev)
    val = torch.from_numpy(np.ascontiguousarray(h.data.astype(np.complex128)).view(np.float64)).to(dev)
    v = torch.from_numpy(np.ascontiguousarray(v0.astype(np.complex128)).view(np.float64)).to(dev)
    alpha = torch.full((K,), np.nan, dtype=torch.float64, device=dev)
    beta = torch.full((K,), np.nan, dtype=torch.float64, device=dev)
    ks = torch.full((2,), -7, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    n = h.shape[0]
    check(lib.imped_lanczos_noalloc(C.c_void_p(rowptr.data_ptr()), C.c_void_p(col.data_ptr()),
                                    C.c_void_p(val.data_ptr()), n, C.c_void_p(v.data_ptr()), K, reorth,
                                    C.c_void_p(alpha.data_ptr()), C.c_void_p(beta.data_ptr()),
                                    C.c_void_p(ks.data_ptr()), C.c_void_p(ks.data_ptr() + 4), None))
    torch.cuda.synchronize()
    return alpha.cpu().numpy(), beta.cpu().numpy(), ks.cpu().numpy()


def _batch(h, v0, K, reorth):
    from impuritymodel_amd.finite import _batch_from_matrix

    b = _batch_from_matrix(h)
    alpha, beta, kout, status = b.lanczos(K, reorth=reorth, v0=v0)
    return alpha[0], beta[0], int(kout[0]), int(status[0])


@pytest.mark.parametrize("reorth", [0, 1])
def test_noalloc_lanczos_equals_batch_entry_point(gpu, reorth):
    # a 256-thread chain, then smaller one-wave chains on the same scratch
    # (the last one, with reorthogonalization, keeps its basis in LDS)
    for n, K, seed in ((9000, 120, 1), (700, 200, 2), (60, 60, 3)):
        h, v0 = _hermitian(n, seed, 6.0 / n)
        a_ref, b_ref, k_ref, s_ref = _batch(h, v0, K, reorth)
        a, b, ks = _noalloc(h, v0, K, reorth)
        kk = min(K, n)
        print("n=%d K=%d reorth=%d kout=%d status=%d" % (n, K, reorth, ks[0], ks[1]), flush=True)
        assert (int(ks[0]), int(ks[1])) == (k_ref, s_ref) == (kk, 0)
        assert np.array_equal(a[:kk], a_ref[:kk])
        assert np.array_equal(b[:kk - 1], b_ref[:kk - 1])


def test_noalloc_lanczos_refuses_chain_beyond_scratch(gpu):
    import torch

    from impuritymodel_amd._lib import IMPED_ERR_CAPACITY, lib

    dev = torch.device("cuda:0")
    small = torch.zeros(16, dtype=torch.float64, device=dev)
    p = C.c_void_p(small.data_ptr())
    # 48 n bytes exceed the 64 MiB scratch: refused before any launch
    rc = lib.imped_lanczos_noalloc(p, p, p, 2_000_000, p, 10, 0, p, p, p, p, None)
    assert rc == IMPED_ERR_CAPACITY
    assert b"scratch" in lib.imped_last_error()
    # with reorthogonalization the Krylov vectors count too
    rc = lib.imped_lanczos_noalloc(p, p, p, 30_000, p, 200, 1, p, p, p, p, None)
    assert rc == IMPED_ERR_CAPACITY
