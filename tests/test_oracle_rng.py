"""Pins the oracle's integer RNG (assets/shaders/random.wgsl:1-68).

Checks: an independent numpy uint32 evaluation of each hash on 1M inputs,
committed golden vectors, and the uniform-float quirk (1.0 reachable,
random.wgsl:16-19)."""
import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).parent / "golden" / "rng.json"
M32 = np.uint64(0xFFFFFFFF)


def np_pcg(x):
    x = x.astype(np.uint64)
    st = (x * np.uint64(747796405) + np.uint64(2891336453)) & M32
    w = (((st >> ((st >> np.uint64(28)) + np.uint64(4))) ^ st) * np.uint64(277803737)) & M32
    return ((w >> np.uint64(22)) ^ w).astype(np.uint32)


def np_xorshift(x):
    s = x.astype(np.uint64)
    s = (s ^ (s << np.uint64(13))) & M32
    s = s ^ (s >> np.uint64(17))
    s = (s ^ (s << np.uint64(5))) & M32
    return s.astype(np.uint32)


def np_hash_combine(x, y):
    x = x.astype(np.uint64)
    y = y.astype(np.uint64)
    v = (((x * np.uint64(1664525) + y + np.uint64(1013904223)) & M32) * np.uint64(1664525)) & M32
    v ^= v >> np.uint64(11)
    v ^= (v << np.uint64(7)) & np.uint64(0x9D2C5680)
    v ^= (v << np.uint64(15)) & np.uint64(0xEFC60000)
    v ^= v >> np.uint64(18)
    return v.astype(np.uint32)


def test_hashes_match_numpy(oracle):
    rng = np.random.default_rng(7)
    xs = rng.integers(0, 2**32, size=4096, dtype=np.uint64).astype(np.uint32)
    xs[:6] = [0, 1, 2, 0xFFFFFFFF, 0x80000000, 12345]
    ys = rng.integers(0, 2**32, size=4096, dtype=np.uint64).astype(np.uint32)
    pcg = np.array([oracle.lib.or_pcg_hash(int(x)) for x in xs], np.uint32)
    xsh = np.array([oracle.lib.or_xor_shift(int(x)) for x in xs], np.uint32)
    hc = np.array([oracle.lib.or_hash_combine(int(x), int(y)) for x, y in zip(xs, ys)], np.uint32)
    np.testing.assert_array_equal(pcg, np_pcg(xs))
    np.testing.assert_array_equal(xsh, np_xorshift(xs))
    np.testing.assert_array_equal(hc, np_hash_combine(xs, ys))


def test_uniform_chain_matches_numpy(oracle):
    for seed in (0, 1, 0xDEADBEEF):
        got = oracle.uniform_floats(seed, 1000)
        st = np.uint32(seed)
        ref = []
        for _ in range(1000):
            st = np_pcg(np.array([st], np.uint32))[0]
            ref.append(np.float32(st) / np.float32(4294967296.0))
        np.testing.assert_array_equal(got, np.array(ref, np.float32))
        assert got.min() >= 0.0 and got.max() <= 1.0


def test_uniform_reaches_one(oracle):
    # f32(state) rounds to 2^32 for states >= 2^32 - 128, so the draw is 1.0.
    # pcg_hash is a bijection, so 128 of 2^32 inputs land there (~8 per 2^28).
    found = None
    for chunk in range(64):
        xs = np.arange(chunk << 22, (chunk + 1) << 22, dtype=np.uint64).astype(np.uint32)
        idx = np.nonzero(np_pcg(xs) >= np.uint32(0xFFFFFF80))[0]
        if idx.size:
            found = int(xs[idx[0]])
            break
    assert found is not None
    assert oracle.uniform_floats(found, 1)[0] == np.float32(1.0)


def test_golden_vectors(oracle):
    g = json.loads(GOLDEN.read_text())
    for x, v in zip(g["inputs"], g["pcg_hash"]):
        assert oracle.lib.or_pcg_hash(x) == v
    for x, v in zip(g["inputs"], g["xor_shift_u32"]):
        assert oracle.lib.or_xor_shift(x) == v
    for x, v in zip(g["inputs"], g["wangh_hash"]):
        assert oracle.lib.or_wangh_hash(x) == v
    for (x, y), v in zip(g["pairs"], g["hash_combine"]):
        assert oracle.lib.or_hash_combine(x, y) == v
    np.testing.assert_array_equal(oracle.uniform_floats(g["chain_seed"], len(g["chain"])),
                                  np.array(g["chain"], np.float32))
