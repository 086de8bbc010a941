"""Product-state formats of impurityModel (product_state_representation.py).

Keys are the reference's ``bytes`` (big-endian bitarray, zero-padded); the
engine's packed form is ``opcompile.bytes_to_words``.  Only the conversions
callers need without bitarray are provided: tuple, str, int and bytes.
"""

from .opcompile import bytes_to_words, words_to_bytes  # noqa: F401


def tuple2str(t, n):
    occ = set(t)
    return "".join("1" if i in occ else "0" for i in range(n))


def str2tuple(s):
    return tuple(pos for pos, ch in enumerate(s) if ch == "1")


def str2int(s):
    return int(s, 2)


def int2str(i, n):
    s = bin(i)[2:]
    return "0" * (n - len(s)) + s


def tuple2int(t, n):
    return str2int(tuple2str(t, n))


def int2tuple(i, n):
    return str2tuple(int2str(i, n))


def int2bytes(i, n):
    nb = (n + 7) // 8
    return (i << (8 * nb - n)).to_bytes(nb, "big")


def bytes2int(b, n):
    return int.from_bytes(b, "big") >> (8 * len(b) - n)


def tuple2bytes(t, n):
    return int2bytes(tuple2int(t, n), n)


def bytes2tuple(b, n):
    return int2tuple(bytes2int(b, n), n)


def str2bytes(s):
    return int2bytes(str2int(s), len(s))


def bytes2str(b, n):
    return int2str(bytes2int(b, n), n)
