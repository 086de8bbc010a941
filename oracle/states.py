"""Product-state encodings and single fermion operators (oracle).

Restates ``impurityModel/ed/product_state_representation.py`` (int, str,
tuple and bytes forms) and the integer variants of ``create.py`` and
``remove.py``.  The bitarray form is not needed: its bytes are produced from
the integer form directly.
"""


def tuple2str(t, n):
    """product_state_representation.py:78-96"""
    occ = set(t)
    return "".join("1" if i in occ else "0" for i in range(n))


def str2tuple(s):
    """product_state_representation.py:64-75"""
    return tuple(pos for pos, ch in enumerate(s) if ch == "1")


def str2int(s):
    """product_state_representation.py:35-44"""
    return int(s, 2)


def int2str(i, n):
    """product_state_representation.py:47-61"""
    s = bin(i)[2:]
    return "0" * (n - len(s)) + s


def tuple2int(t, n):
    """product_state_representation.py:99-112 (orbital i at bit n-1-i)"""
    v = 0
    for i in t:
        v |= 1 << (n - 1 - i)
    return v


def int2tuple(v, n):
    """product_state_representation.py:115-128"""
    return tuple(i for i in range(n) if (v >> (n - 1 - i)) & 1)


def int2bytes(v, n):
    """Bytes key of the reference (product_state_representation.py:131-141,
    303-315): big-endian bitarray of length n, zero-padded to whole bytes."""
    nbytes = (n + 7) // 8
    return (v << (8 * nbytes - n)).to_bytes(nbytes, "big")


def bytes2int(b, n):
    """product_state_representation.py:318-331"""
    nbytes = len(b)
    return int.from_bytes(b, "big") >> (8 * nbytes - n)


def tuple2bytes(t, n):
    """product_state_representation.py:246-258"""
    return int2bytes(tuple2int(t, n), n)


def bytes2tuple(b, n):
    """product_state_representation.py:261-273"""
    return int2tuple(bytes2int(b, n), n)


def str2bytes(s):
    """product_state_representation.py:276-286"""
    return int2bytes(str2int(s), len(s))


def create(n, i, state):
    """c_i^dagger on an int state: create.py:58-92.

    Returns ``(new_state, amp)``; ``(-1, 0)`` when orbital ``i`` is occupied.
    The sign is (-1)^(number of occupied orbitals j < i).
    """
    bit = 1 << (n - 1 - i)
    if state & bit:
        return -1, 0
    amp = 1 if bin(state >> (n - i)).count("1") % 2 == 0 else -1
    return state | bit, amp


def remove(n, i, state):
    """c_i on an int state: remove.py:57-92.

    Returns ``(new_state, amp)``; ``(-1, 0)`` when orbital ``i`` is empty.
    """
    bit = 1 << (n - 1 - i)
    if not state & bit:
        return -1, 0
    amp = 1 if bin(state >> (n - i)).count("1") % 2 == 0 else -1
    return state ^ bit, amp
