"""Pins the oracle's noise stream: Philox4x32-10 against the published
known-answer vectors (Salmon et al. SC'11 / Random123 kat_vectors), and the
uniform / Box-Muller transforms by their defining properties."""

from __future__ import annotations

import numpy as np

from oracle import philox

KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


def test_philox_known_answers():
    for ctr, key, want in KAT:
        got = philox.philox4x32_10(np.array([ctr], np.uint32), np.array(key, np.uint32))[0]
        assert tuple(int(v) for v in got) == want


def test_philox_batch_matches_scalar():
    rng = np.random.default_rng(1)
    ctr = rng.integers(0, 2**32, (64, 4), dtype=np.uint64).astype(np.uint32)
    key = np.array([123, 456], np.uint32)
    batch = philox.philox4x32_10(ctr, key)
    for i in range(0, 64, 7):
        one = philox.philox4x32_10(ctr[i : i + 1], key)[0]
        assert np.array_equal(one, batch[i])


def test_uniforms_exact_in_float32_and_open_interval():
    u = philox.uniforms(99, np.arange(100000), 3, 1)
    assert np.all(u > 0) and np.all(u < 1)
    assert np.array_equal(u.astype(np.float32).astype(np.float64), u)
    assert abs(u.mean() - 0.5) < 0.005


def test_normals_moments_and_independence_of_streams():
    n = philox.normals(7, np.arange(200000), 0, 0)
    assert abs(n.mean()) < 0.01
    assert abs(n.std() - 1.0) < 0.01
    assert abs(np.corrcoef(n[:, 0], n[:, 2])[0, 1]) < 0.01
    m = philox.normals(7, np.arange(200000), 0, 1)
    assert abs(np.corrcoef(n[:, 0], m[:, 0])[0, 1]) < 0.01
    # a different step index is a different stream, the same counter repeats
    assert not np.array_equal(philox.normals(7, np.arange(10), 1, 0), philox.normals(7, np.arange(10), 0, 0))
    assert np.array_equal(philox.normals(7, np.arange(10), 1, 0), philox.normals(7, np.arange(10), 1, 0))
