"""The walkers' random stream (SplitMix64) against the generator's published
test vector, and the shape of its normals."""

import numpy as np

from walker_rng import SplitMixStream


def test_splitmix64_published_vector():
    # first outputs of SplitMix64 seeded with 1234567 (the reference vector
    # distributed with the generator's C implementation)
    ref = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
           16408922859458223821]
    s = SplitMixStream(1234567)
    assert [int(v) for v in s.next_u64(3)] + [int(v) for v in s.next_u64(2)] == ref


def test_stream_uniforms_and_normals():
    s = SplitMixStream(42)
    u = s.uniform(200000)
    assert u.min() >= 0.0 and u.max() < 1.0 and abs(u.mean() - 0.5) < 5e-3
    z = SplitMixStream(43).normal((100000, 2))
    assert z.shape == (100000, 2)
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1.0) < 0.02
    # two normals per two uniforms (cosine, then sine), the same sequence however it is drawn
    a, b = SplitMixStream(7), SplitMixStream(7)
    u = b.uniform(6)
    r, t = np.sqrt(-2.0 * np.log(1.0 - u[0::2])), 2.0 * np.pi * u[1::2]
    ref = np.stack([r * np.cos(t), r * np.sin(t)], axis=1).ravel()
    got = np.concatenate([a.normal(1), a.normal((2, 1)).ravel(), a.normal(3)])
    np.testing.assert_array_equal(got, ref)
