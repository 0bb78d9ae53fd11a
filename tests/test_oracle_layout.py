"""Pins the oracle's output layout: 64x64 block-major order
(crates/appearance-render-loop/assets/shaders/block.wgsl:3-15) and the ACES
fit of resolve.wgsl:29-38."""
import numpy as np


def test_block_index_is_permutation(oracle):
    w, h = 192, 128
    idx = np.array([[oracle.lib.or_block_index(x, y, w) for x in range(w)] for y in range(h)])
    assert sorted(idx.ravel().tolist()) == list(range(w * h))
    # known entries: blocks of 4096 px, row-major inside a block
    assert oracle.lib.or_block_index(0, 0, w) == 0
    assert oracle.lib.or_block_index(1, 0, w) == 1
    assert oracle.lib.or_block_index(0, 1, w) == 64
    assert oracle.lib.or_block_index(64, 0, w) == 4096
    assert oracle.lib.or_block_index(0, 64, w) == 3 * 4096
    assert oracle.lib.or_block_index(191, 127, w) == w * h - 1


def test_aces_closed_form(oracle):
    assert oracle.lib.or_aces(0.0) == 0.0
    assert oracle.lib.or_aces(1e6) == 1.0
    x = np.float32(0.5)
    expect = (x * (np.float32(2.51) * x + np.float32(0.03))) / (
        x * (np.float32(2.43) * x + np.float32(0.59)) + np.float32(0.14))
    assert abs(oracle.lib.or_aces(0.5) - float(expect)) < 1e-7
    xs = np.linspace(0, 20, 200)
    ys = [oracle.lib.or_aces(float(v)) for v in xs]
    assert all(b >= a for a, b in zip(ys, ys[1:]))
