"""Writes tests/golden/rng.json: RNG golden vectors.

The integers are computed here with plain Python int arithmetic modulo 2^32
straight from the text of assets/shaders/random.wgsl (the reference itself
cannot run in this environment: it is WGSL inside a Rust crate)."""
import json
import struct
from pathlib import Path

M = 0xFFFFFFFF


def pcg(x):
    st = (x * 747796405 + 2891336453) & M
    w = (((st >> ((st >> 28) + 4)) ^ st) * 277803737) & M
    return (w >> 22) ^ w


def xs(x):
    s = (x ^ (x << 13)) & M
    s ^= s >> 17
    s = (s ^ (s << 5)) & M
    return s


def wangh(x):
    s = (x ^ 61) ^ (x >> 16)
    s = (s * 9) & M
    s ^= s >> 4
    s = (s * 0x27D4EB2D) & M
    s ^= s >> 15
    return s


def hc(x, y):
    v = (((x * 1664525 + y + 1013904223) & M) * 1664525) & M
    v ^= v >> 11
    v ^= (v << 7) & 0x9D2C5680
    v ^= (v << 15) & 0xEFC60000
    v ^= v >> 18
    return v & M


def f32(x):
    return struct.unpack("f", struct.pack("f", x))[0]


def main():
    inputs = [0, 1, 2, 3, 42, 1337, 65535, 65536, 0x7FFFFFFF, 0x80000000, 0xDEADBEEF, 0xFFFFFFFF]
    pairs = [[0, 0], [1, 2], [0xFFFFFFFF, 7], [123456789, 987654321]]
    chain = []
    st = 2024
    for _ in range(16):
        st = pcg(st)
        chain.append(f32(f32(float(st)) / 4294967296.0))
    out = {
        "inputs": inputs,
        "pcg_hash": [pcg(x) for x in inputs],
        "xor_shift_u32": [xs(x) for x in inputs],
        "wangh_hash": [wangh(x) for x in inputs],
        "pairs": pairs,
        "hash_combine": [hc(x, y) for x, y in pairs],
        "chain_seed": 2024,
        "chain": chain,
    }
    Path(__file__).with_name("rng.json").write_text(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
