"""The random streams of libimped's fitting-loop walkers, restated in numpy
with the interface of numpy's generator (normal(size=), uniform(size=)), so
the Python fitting loop can be driven by the same numbers: normals from the
stream seeded with seed, decision uniforms from the one seeded with ~seed.

SplitMix64 (Steele, Lea and Flood, OOPSLA 2014): state += 0x9E3779B97F4A7C15,
output = the state through the 64-bit finalizer; uniforms are the top 53
bits times 2^-53; two uniforms (u1 = 1 - first, u2 = second) give two
normals by Box-Muller, r cos(2 pi u2) first and r sin(2 pi u2) second, with
r = sqrt(-2 ln u1).
"""

import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)


class SplitMixStream:
    def __init__(self, seed):
        self.state = np.uint64(int(seed) % (1 << 64))
        self._spare = None  # the sine normal of the last pair, not yet used

    def next_u64(self, n):
        s = self.state + GAMMA * np.arange(1, n + 1, dtype=np.uint64)
        if n:
            self.state = s[-1]
        z = (s ^ (s >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))

    def uniform(self, size):
        n = int(np.prod(size))
        return ((self.next_u64(n) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53).reshape(size)

    def normal(self, size):
        n = int(np.prod(size))
        out = []
        if n and self._spare is not None:
            out.append(np.array([self._spare]))
            self._spare = None
        m = n - sum(len(o) for o in out)
        if m:
            pairs = (m + 1) // 2
            u = self.uniform(2 * pairs)
            r = np.sqrt(-2.0 * np.log(1.0 - u[0::2]))
            a = 2.0 * np.pi * u[1::2]
            z = np.empty(2 * pairs)
            z[0::2] = r * np.cos(a)
            z[1::2] = r * np.sin(a)
            out.append(z[:m])
            if m % 2:
                self._spare = z[-1]
        return (np.concatenate(out) if out else np.zeros(0)).reshape(size)


class WalkerStreams:
    """normal() from SplitMix64(seed), uniform() from SplitMix64(~seed)."""

    def __init__(self, seed):
        self._n = SplitMixStream(seed)
        self._u = SplitMixStream(~int(seed) % (1 << 64))

    def normal(self, size):
        return self._n.normal(size)

    def uniform(self, size):
        return self._u.uniform(size)
