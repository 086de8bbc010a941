"""Compile second-quantized operator dicts into the engine's term tables.

A process ``((i1, a1), ..., (ik, ak))`` acts right to left
(impurityModel/ed/finite.py:1365): each step removes ('a') or creates ('c') an
electron at spin-orbital ``i`` with sign (-1)^(occupied orbitals j < i) of the
current state (create.py:142-148, remove.py:141-147), and the product is zero
as soon as a step is Pauli-blocked.  Walking the steps symbolically turns a
process into

* ``care``/``val``: the original occupations it requires
  (zero on s unless (s & care) == val),
* ``flip``: the orbitals whose occupation it changes (target = s ^ flip),
* ``sgn``/``parity``: sign = (-1)^(popcount(s & sgn) + parity), because the
  sign of every step is the parity of s XOR (flips so far) on the orbitals
  below i, and parity is linear over XOR.

Processes are grouped by flip mask in the dict's first-occurrence order; each
term keeps its dict position, which the engine uses to reproduce applyOp's
summation and insertion order.  Structures are cached per tuple of processes,
so models that differ only in values compile once.

State words: W uint64 per state, orbital i at bit 63 - i % 64 of word i // 64,
i.e. the reference's bytes key zero-padded to 8W bytes read as big-endian
words.
"""

from functools import lru_cache

import numpy as np


def n_words(n_orb):
    return max(1, (n_orb + 63) // 64)


def _bit(i, W):
    return 1 << (64 * W - 1 - i)


def _prefix(i, W):
    """Mask of orbitals j < i."""
    top = 64 * W
    return ((1 << top) - 1) ^ ((1 << (top - i)) - 1)


def _words(x, W):
    return [(x >> (64 * (W - 1 - w))) & 0xFFFFFFFFFFFFFFFF for w in range(W)]


def compile_process(process, W):
    """Return (care, val, flip, sgn, parity) ints, or None if identically zero."""
    known = {}
    care = val = sgn = flips = 0
    parity = 0
    for i, action in process[::-1]:
        if action == "a":
            need = 1
        elif action == "c":
            need = 0
        else:
            raise Exception("Operator type unknown: {}".format(action))
        b = _bit(i, W)
        if i in known:
            if known[i] != need:
                return None
        else:
            care |= b
            if need:
                val |= b
        p = _prefix(i, W)
        sgn ^= p
        parity ^= bin(flips & p).count("1") & 1
        flips ^= b
        known[i] = 1 - need
    # sign bits on required orbitals are constants
    parity ^= bin(sgn & care & val).count("1") & 1
    sgn &= ~care
    return care, val, flips, sgn, parity


@lru_cache(maxsize=256)
def _structure(processes, W):
    """Grouped term structure of a tuple of processes (cached)."""
    groups = {}
    order = []
    for t, proc in enumerate(processes):
        c = compile_process(proc, W)
        if c is None:
            continue
        care, val, flip, sgn, par = c
        if flip not in groups:
            groups[flip] = []
            order.append(flip)
        groups[flip].append((t, care, val, sgn, par))
    T = sum(len(groups[f]) for f in order)
    G = len(order)
    group_term_off = np.zeros(G + 1, np.int32)
    group_flip = np.zeros((G, W), np.uint64)
    masks = np.zeros((T, 3, W), np.uint64)
    meta = np.zeros(T, np.uint32)
    src_index = np.zeros(T, np.int64)
    k = 0
    for g, f in enumerate(order):
        group_flip[g] = _words(f, W)
        for t, care, val, sgn, par in groups[f]:
            masks[k, 0] = _words(care, W)
            masks[k, 1] = _words(val, W)
            masks[k, 2] = _words(sgn, W)
            meta[k] = (t << 1) | par
            src_index[k] = t
            k += 1
        group_term_off[g + 1] = k
    return group_term_off, group_flip, masks.reshape(T, 3 * W), meta, src_index


def restriction_arrays(restrictions, W):
    """{frozenset(indices): (lo, hi)} -> (masks [R, W], lohi [R, 2])."""
    if not restrictions:
        return np.zeros((0, W), np.uint64), np.zeros((0, 2), np.int32)
    masks, lohi = [], []
    for idx, (lo, hi) in restrictions.items():
        m = 0
        for i in idx:
            m |= _bit(i, W)
        masks.append(_words(m, W))
        lohi.append((int(lo), int(hi)))
    return np.array(masks, np.uint64).reshape(-1, W), np.array(lohi, np.int32).reshape(-1, 2)


class TermTable:
    """One operator's compiled term table (with its restrictions) for W words."""

    __slots__ = ("W", "group_term_off", "group_flip", "masks", "meta", "coef", "rmask", "rlohi", "src")

    def __init__(self, op, restrictions, W):
        g_off, g_flip, masks, met, src = _structure(tuple(op.keys()), W)
        vals = np.fromiter(op.values(), dtype=np.complex128, count=len(op)) if len(op) else np.zeros(0, complex)
        self.W = W
        self.group_term_off = g_off
        self.group_flip = g_flip
        self.masks = masks
        self.meta = met
        self.src = src
        self.coef = vals[src]
        self.rmask, self.rlohi = restriction_arrays(restrictions, W)

    def with_coef(self, values):
        """The same term structure with new values (given in dict-key order)."""
        t = TermTable.__new__(TermTable)
        for name in TermTable.__slots__:
            setattr(t, name, getattr(self, name))
        t.coef = np.asarray(values, np.complex128)[self.src]
        return t


def _compile_periodic(tts, W):
    """compile_models for a list whose term structures repeat with a period p
    (tables sharing their arrays, as a model family or the per-model dipole
    operators give): the structure arrays are tiled, only the values are
    gathered per model.  None when the list has no such period."""
    M = len(tts)
    if M < 2:
        return None
    p = 1
    while p <= M // 2 and tts[p].masks is not tts[0].masks:
        p += 1
    if p > M // 2 or M % p:
        return None
    for i in range(p, M):
        a, b = tts[i], tts[i % p]
        if a.masks is not b.masks or a.rmask is not b.rmask or a.rlohi is not b.rlohi:
            return None
    reps = M // p
    # structure of one period
    g_cnt = np.array([len(t.group_term_off) - 1 for t in tts[:p]], np.int64)
    t_cnt = np.array([len(t.coef) for t in tts[:p]], np.int64)
    r_cnt = np.array([len(t.rmask) for t in tts[:p]], np.int64)
    T = int(t_cnt.sum())
    g_off = np.concatenate([t.group_term_off[:-1] + o for t, o in zip(tts[:p], np.concatenate([[0], np.cumsum(t_cnt)[:-1]]))])
    group_term_off = (g_off[None, :].astype(np.int64) + T * np.arange(reps)[:, None]).reshape(-1)
    group_term_off = np.concatenate([group_term_off, [T * reps]]).astype(np.int32)
    coef = np.concatenate([t.coef for t in tts]) if T else np.zeros(0, complex)
    return dict(
        model_group_off=np.concatenate([[0], np.cumsum(np.tile(g_cnt, reps))]).astype(np.int32),
        group_term_off=group_term_off,
        group_flip=np.ascontiguousarray(np.tile(np.concatenate([t.group_flip for t in tts[:p]]).reshape(-1, W),
                                                (reps, 1)), np.uint64),
        term_masks=np.ascontiguousarray(np.tile(np.concatenate([t.masks for t in tts[:p]]).reshape(-1, 3 * W),
                                                (reps, 1)), np.uint64),
        term_meta=np.ascontiguousarray(np.tile(np.concatenate([t.meta for t in tts[:p]]), reps), np.uint32),
        term_coef=np.ascontiguousarray(np.stack([coef.real, coef.imag], axis=-1).reshape(-1), np.float64),
        model_restr_off=np.concatenate([[0], np.cumsum(np.tile(r_cnt, reps))]).astype(np.int32),
        restr_mask=np.ascontiguousarray(np.tile(np.concatenate([t.rmask for t in tts[:p]]).reshape(-1, W),
                                                (reps, 1)), np.uint64),
        restr_lohi=np.ascontiguousarray(np.tile(np.concatenate([t.rlohi for t in tts[:p]]).reshape(-1), reps),
                                        np.int32),
    )


def compile_models(ops, restrictions, W):
    """Concatenate the term tables of several operators.

    ``ops`` is a list of operator dicts (or ready ``TermTable`` objects),
    ``restrictions`` a list (same length) of restriction dicts or None.
    Returns a dict of arrays matching ``imped_ops_create``.
    """
    tts = []
    for m, op in enumerate(ops):
        if isinstance(op, TermTable):
            assert op.W == W, "term table compiled for another word count"
            tts.append(op)
        else:
            tts.append(TermTable(op, restrictions[m] if restrictions is not None else None, W))
    tiled = _compile_periodic(tts, W)
    if tiled is not None:
        return tiled
    mg, gt, gf, tm, meta, coef, mr, rm, rl = [0], [], [], [], [], [], [0], [], []
    t_base = 0
    for tt in tts:
        g_off = tt.group_term_off
        gt.append(g_off[:-1] + t_base if len(g_off) > 1 else np.zeros(0, np.int32))
        gf.append(tt.group_flip)
        tm.append(tt.masks)
        meta.append(tt.meta)
        coef.append(tt.coef)
        t_base += len(tt.coef)
        mg.append(mg[-1] + len(g_off) - 1)
        rm.append(tt.rmask)
        rl.append(tt.rlohi)
        mr.append(mr[-1] + len(tt.rmask))
    group_term_off = np.concatenate(gt + [np.array([t_base], np.int32)]).astype(np.int32)
    coef = np.concatenate(coef) if coef else np.zeros(0, complex)
    return dict(
        model_group_off=np.array(mg, np.int32),
        group_term_off=group_term_off,
        group_flip=np.ascontiguousarray(np.concatenate(gf).reshape(-1, W), np.uint64),
        term_masks=np.ascontiguousarray(np.concatenate(tm).reshape(-1, 3 * W), np.uint64),
        term_meta=np.ascontiguousarray(np.concatenate(meta), np.uint32),
        term_coef=np.ascontiguousarray(np.stack([coef.real, coef.imag], axis=-1).reshape(-1), np.float64),
        model_restr_off=np.array(mr, np.int32),
        restr_mask=np.ascontiguousarray(np.concatenate(rm).reshape(-1, W), np.uint64),
        restr_lohi=np.ascontiguousarray(np.concatenate(rl).reshape(-1), np.int32),
    )


# ------------------------------------------------------------------ state conversions


def bytes_to_words(keys, n_orb, W=None):
    """Reference bytes keys -> uint64 array [N, W]."""
    W = W or n_words(n_orb)
    nb = (n_orb + 7) // 8
    N = len(keys)
    if N == 0:
        return np.zeros((0, W), np.uint64)
    raw = np.frombuffer(b"".join(keys), dtype=np.uint8).reshape(N, nb)
    pad = np.zeros((N, 8 * W), np.uint8)
    pad[:, :nb] = raw
    return pad.view(">u8").astype(np.uint64).reshape(N, W)


def words_to_bytes(words, n_orb):
    """uint64 array [N, W] -> list of reference bytes keys."""
    nb = (n_orb + 7) // 8
    words = np.ascontiguousarray(words, dtype=np.uint64)
    N = words.shape[0]
    if N == 0:
        return []
    raw = words.astype(">u8").view(np.uint8).reshape(N, -1)[:, :nb]
    b = raw.tobytes()
    return [b[i * nb:(i + 1) * nb] for i in range(N)]


def ints_to_words(states, n_orb, W=None):
    """Reference integer-form states (orbital i at bit n-1-i) -> [N, W] words."""
    W = W or n_words(n_orb)
    shift = 64 * W - n_orb
    out = np.zeros((len(states), W), np.uint64)
    for k, s in enumerate(states):
        v = int(s) << shift
        for w in range(W):
            out[k, w] = (v >> (64 * (W - 1 - w))) & 0xFFFFFFFFFFFFFFFF
    return out


def words_to_ints(words, n_orb):
    W = words.shape[1] if words.ndim == 2 else 1
    shift = 64 * W - n_orb
    out = []
    for row in np.asarray(words, dtype=np.uint64).reshape(-1, W):
        v = 0
        for w in range(W):
            v = (v << 64) | int(row[w])
        out.append(v >> shift)
    return out
