"""Known-answer tests of the oracle, taken from the reference's own tests.

impurityModel/test/test_product_state_representation.py:11-152,
test_create.py:5-92, test_remove.py:5-73 and test_applyOp.py:5-155, restated
on the oracle's int/bytes forms (no bitarray).
"""

from oracle import applyop
from oracle import states as psr


def test_psr_conversions():
    n = 7
    assert psr.tuple2str((2, 5), n) == "0010010"
    assert psr.str2tuple("0010010") == (2, 5)
    assert psr.tuple2int((2, 5), n) == 18
    assert psr.int2tuple(18, n) == (2, 5)
    assert psr.tuple2bytes((2, 5), n) == b"$"
    assert psr.bytes2tuple(b"$", n) == (2, 5)
    assert psr.str2bytes("0010010") == b"$"
    assert psr.int2bytes(18, n) == b"$"
    assert psr.bytes2int(b"$", n) == 18
    assert psr.int2str(18, n) == "0010010"


def test_create_sequence():
    n = 7
    s, amp = psr.create(n, 5, 0)
    assert amp == 1
    s, amp = psr.create(n, 2, s)
    assert amp == 1
    assert psr.int2tuple(s, n) == (2, 5)
    s, amp = psr.create(n, 3, s)
    assert amp == -1
    assert psr.int2tuple(s, n) == (2, 3, 5)
    s2, amp = psr.create(n, 3, s)
    assert amp == 0 and s2 == -1


def test_remove_sequence():
    n = 7
    s = psr.tuple2int((2, 3, 5), n)
    s, amp = psr.remove(n, 3, s)
    assert amp == -1 and psr.int2tuple(s, n) == (2, 5)
    s, amp = psr.remove(n, 2, s)
    assert amp == 1 and psr.int2tuple(s, n) == (5,)
    s2, amp = psr.remove(n, 4, s)
    assert amp == 0 and s2 == -1


def _b(s):
    return psr.bytes2int(psr.str2bytes(s), len(s))


def test_applyop_create_one():
    vac = "000000"
    psi = {_b(vac): 7}
    for i in range(6):
        r = applyop.apply_op(6, {((i, "c"),): 9}, psi)
        assert r == {_b(vac[:i] + "1" + vac[i + 1:]): 63}


def test_applyop_create_on_one_electron():
    ps = "001000"
    psi = {_b(ps): 7}
    for i in range(6):
        r = applyop.apply_op(6, {((i, "c"),): 9}, psi)
        if i == 2:
            assert r == {}
        else:
            assert r == {_b(ps[:i] + "1" + ps[i + 1:]): 7 * 9 * (2 * (i < 2) - 1)}


def test_applyop_create_two():
    vac = "000000"
    psi = {_b(vac): 7}
    for i in range(6):
        for j in range(6):
            r = applyop.apply_op(6, {((i, "c"), (j, "c")): 9}, psi)
            if i == j:
                assert r == {}
            else:
                a, b = min(i, j), max(i, j)
                s = vac[:a] + "1" + vac[a + 1:b] + "1" + vac[b + 1:]
                assert r == {_b(s): 7 * 9 * (2 * (i < j) - 1)}


def test_applyop_two_processes_and_same_target():
    vac = "000000"
    psi = {_b(vac): 7}
    for i in range(6):
        for j in range(6):
            if i != j:
                r = applyop.apply_op(6, {((i, "c"),): 9, ((j, "c"),): 11}, psi)
                assert r == {_b(vac[:i] + "1" + vac[i + 1:]): 63, _b(vac[:j] + "1" + vac[j + 1:]): 77}
    ps = "011000"
    r = applyop.apply_op(6, {((1, "c"), (1, "a")): 9, ((2, "c"), (2, "a")): 11}, {_b(ps): 7})
    assert r == {_b(ps): 7 * 20}
    for i in range(6):
        assert applyop.apply_op(6, {((i, "a"),): 9}, {_b(vac): 7}) == {}


def test_applyop_opresult_memo_trusted():
    vac = "000000"
    psi = {_b(vac): 7}
    op = {((0, "c"),): 9, ((3, "c"),): 11}
    memo = {}
    r = applyop.apply_op(6, op, psi, opResult=memo)
    assert r == {_b("100000"): 63, _b("000100"): 77}
    assert memo == {_b(vac): {_b("100000"): 9, _b("000100"): 11}}
    assert applyop.apply_op(6, op, psi, opResult=memo) == r
    wrong = {_b(vac): {_b("111111"): 2}}
    assert applyop.apply_op(6, op, psi, opResult=wrong) == {_b("111111"): 14}


def test_applyop_restrictions():
    restrictions = {frozenset([1, 4, 5]): (0, 1)}
    vac = "000000"
    psi = {_b(vac): 7}
    for i in range(6):
        for j in range(6):
            r = applyop.apply_op(6, {((i, "c"), (j, "c")): 9}, psi, restrictions=restrictions)
            if i == j or (i in (1, 4, 5) and j in (1, 4, 5)):
                assert r == {}
            else:
                a, b = min(i, j), max(i, j)
                s = vac[:a] + "1" + vac[a + 1:b] + "1" + vac[b + 1:]
                assert r == {_b(s): 7 * 9 * (2 * (i < j) - 1)}
