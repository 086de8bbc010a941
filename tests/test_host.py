"""Host-side product code on the CPU: the term compiler, state conversions,
the native basis enumerator, and the C ABI exports of libimped.so."""

import ctypes
import os
import re

import numpy as np
import pytest

from impuritymodel_amd import opcompile
from oracle import applyop, basis

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _eval_compiled(n, op, psi, W):
    """Evaluate the compiled masks in Python (the kernel's arithmetic)."""
    go, gf, masks, meta, _ = opcompile._structure(tuple(op.keys()), W)
    vals = list(op.values())
    shift = 64 * W - n

    def big(words):
        v = 0
        for w in range(W):
            v = (v << 64) | int(words[w])
        return v

    out = {}
    for s, amp in psi.items():
        S = s << shift
        for g in range(len(go) - 1):
            acc, hit = 0, False
            for t in range(go[g], go[g + 1]):
                care, val, sgn = big(masks[t][:W]), big(masks[t][W:2 * W]), big(masks[t][2 * W:])
                if (S & care) == val:
                    par = (bin(S & sgn).count("1") + int(meta[t])) & 1
                    acc += (-1 if par else 1) * vals[int(meta[t]) >> 1]
                    hit = True
            if hit:
                tgt = (S ^ big(gf[g])) >> shift
                out[tgt] = out.get(tgt, 0) + amp * acc
    return {k: v for k, v in out.items() if abs(v) ** 2 >= 1e-12}


@pytest.mark.parametrize("n", [6, 10, 64, 66, 100])
def test_compiled_terms_match_oracle(n):
    rng = np.random.default_rng(n)
    W = opcompile.n_words(n)
    for _ in range(40):
        op = {}
        for _ in range(int(rng.integers(1, 8))):
            k = int(rng.choice([1, 2, 4]))
            proc = tuple((int(rng.integers(n)), str(rng.choice(["c", "a"]))) for _ in range(k))
            op[proc] = complex(rng.normal(), rng.normal())
        psi = {int.from_bytes(rng.bytes(16), "big") >> (128 - n): complex(rng.normal(), rng.normal())
               for _ in range(3)}
        ref = applyop.apply_op(n, op, psi)
        got = _eval_compiled(n, op, psi, W)
        assert set(ref) == set(got)
        for k in ref:
            assert abs(ref[k] - got[k]) < 1e-12


def test_group_order_and_merge():
    # two processes with the same flip merge into one group in dict order
    op = {((1, "c"), (2, "a")): 1.0, ((3, "c"),): 2.0, ((2, "c"), (1, "a")): 3.0, ((1, "c"), (2, "a"), (3, "c"), (3, "a")): 5.0}
    go, gf, masks, meta, src = opcompile._structure(tuple(op.keys()), 1)
    assert list(src) == [0, 2, 3, 1]
    assert list(go) == [0, 3, 4]


def test_words_bytes_roundtrip():
    rng = np.random.default_rng(1)
    for n in (7, 26, 64, 66, 76, 128):
        nb = (n + 7) // 8
        keys = []
        for _ in range(20):
            v = int.from_bytes(rng.bytes(16), "big") >> (128 - n)
            keys.append((v << (8 * nb - n)).to_bytes(nb, "big"))
        w = opcompile.bytes_to_words(keys, n)
        assert opcompile.words_to_bytes(w, n) == keys
        ints = [int.from_bytes(k, "big") >> (8 * nb - n) for k in keys]
        assert opcompile.words_to_ints(opcompile.ints_to_words(ints, n), n) == ints
        # word order == bytes order
        order_w = sorted(range(20), key=lambda i: tuple(int(x) for x in w[i]))
        order_b = sorted(range(20), key=lambda i: keys[i])
        assert order_w == order_b


def test_native_basis_enumeration_matches_oracle():
    from impuritymodel_amd import finite
    from impuritymodel_amd.models import make_model

    for kw in (dict(n_bath=10), dict(n_bath=20, n_con=10, dnConBath=1), dict(n_bath=50)):
        m = make_model(seed=0, **kw)
        args = (m.nBaths, m.nValBaths, m.dnValBaths, m.dnConBaths, m.dnTols, m.n0imps)
        n = m.n_spin_orbitals
        got = [int.from_bytes(x, "big") >> (8 * len(x) - n) for x in finite.get_basis(*args)]
        assert got == list(basis.get_basis(*args))


def test_c_abi_exports_every_declared_symbol():
    from impuritymodel_amd import _lib

    hdr = open(os.path.join(ROOT, "include", "imped.h")).read()
    names = set(re.findall(r"^\s*(?:int|const char\*)\s+(imped_\w+)\s*\(", hdr, re.M))
    assert len(names) >= 20
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in names:
        assert hasattr(lib, name), name
    # the Python binding declares argtypes for every function of the header
    assert names - {"imped_last_error"} <= set(_lib.SIGNATURES)


def test_c_abi_rejects_null_arguments():
    """Entry points check their handles and arrays before touching memory or
    the GPU: NULL arguments return IMPED_ERR_INVALID with a message."""
    from impuritymodel_amd._lib import IMPED_ERR_INVALID, lib

    w = np.ones(4)
    assert lib.imped_mlp_predict(None, None, 4, None, 1.0, None, None) == IMPED_ERR_INVALID
    assert b"imped_mlp_predict" in lib.imped_last_error()
    assert lib.imped_walkers_decide(None, 1, None, None) == IMPED_ERR_INVALID
    assert lib.imped_walkers_state(None, None, None, None, None) == IMPED_ERR_INVALID
    assert lib.imped_walkers_served(None, 0, 1, None) == IMPED_ERR_INVALID
    assert lib.imped_walkers_reset_counters(None) == IMPED_ERR_INVALID
    assert lib.imped_walkers_set_record_cap(None, 10) == IMPED_ERR_INVALID
    assert lib.imped_walkers_step(None, w.ctypes.data, 1.0, w.ctypes.data, None) == IMPED_ERR_INVALID
    assert lib.imped_walkers_waiting(None, None, None, None) == IMPED_ERR_INVALID
    assert lib.imped_lanczos_noalloc(None, None, None, 10, None, 5, 0, None, None, None, None, None) == \
        IMPED_ERR_INVALID
    assert lib.imped_affine_create(0, None, None, None, None, 0, None, None) == IMPED_ERR_INVALID
    assert lib.imped_batch_set_affine(None, None, None, None, 0, None, 0.0) == IMPED_ERR_INVALID
    assert lib.imped_batch_seed_affine(None, None, None, None, None, None, 0.0, 0.0, None) == IMPED_ERR_INVALID
    assert lib.imped_family_xas(None, None, None, 1, 1, None, 0, None, None, 1, 1, 1.0, 1, 1, 1e-12, 10, 0, 0.0, 0.0,
                                0.0, None, 0, 0.1, None, None, None, None, None, None, None) == IMPED_ERR_INVALID
