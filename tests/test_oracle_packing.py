"""Pins the oracle's packing codecs (packing.wgsl:11-76).

Rgb9e5: exact for values on the shared-exponent grid, error below one grid
step (max channel * 2^-8), clamping at 65408 and of negatives, MiniEngine's
shared exponent.  Oct30: angular error bound, exact axes, antipodes,
and the 15-bit field layout."""
import numpy as np


def test_rgb9e5_roundtrip_error(oracle):
    rng = np.random.default_rng(1)
    rgb = (rng.random((20000, 3)) * np.exp(rng.uniform(-12, 10, (20000, 1)))).astype(np.float32)
    dec = oracle.rgb9e5_unpack(oracle.rgb9e5_pack(rgb))
    mx = rgb.max(axis=1, keepdims=True)
    assert np.all(np.abs(dec - rgb) <= mx * 2.0 ** -8 + 2.0 ** -24)


def test_rgb9e5_exact_grid_values(oracle):
    vals = np.array([[1, 1, 1], [0.5, 0.25, 0.125], [256, 128, 1], [0, 0, 0], [3, 5, 7],
                     [65408, 0, 0]], np.float32)
    dec = oracle.rgb9e5_unpack(oracle.rgb9e5_pack(vals))
    np.testing.assert_array_equal(dec, vals)


def test_rgb9e5_clamps(oracle):
    # 2.0 is below half a grid step (128) once the max channel is 65408
    dec = oracle.rgb9e5_unpack(oracle.rgb9e5_pack(np.array([[1e9, -5, 2.0]], np.float32)))
    np.testing.assert_array_equal(dec, [[65408.0, 0.0, 0.0]])
    dec = oracle.rgb9e5_unpack(oracle.rgb9e5_pack(np.array([[1e9, -5, 384.0]], np.float32)))
    np.testing.assert_array_equal(dec, [[65408.0, 0.0, 384.0]])
    # smallest shared exponent: 2^-16 * 2^-8 steps, tiny values flush to 0
    dec = oracle.rgb9e5_unpack(oracle.rgb9e5_pack(np.array([[1e-12, 0, 0]], np.float32)))
    np.testing.assert_array_equal(dec, [[0, 0, 0]])


def test_rgb9e5_field_layout(oracle):
    # 1.0 in every channel: exponent field 15 + 1 = 16, mantissas 256.
    p = int(oracle.rgb9e5_pack(np.array([[1, 1, 1]], np.float32))[0])
    assert p >> 27 == 16
    assert p & 0x1FF == 256 and (p >> 9) & 0x1FF == 256 and (p >> 18) & 0x1FF == 256
    p = int(oracle.rgb9e5_pack(np.array([[1, 0, 0]], np.float32))[0])
    assert (p >> 9) & 0x1FF == 0 and (p >> 18) & 0x1FF == 0


def test_oct30_angular_error(oracle):
    rng = np.random.default_rng(2)
    v = rng.normal(size=(50000, 3)).astype(np.float32)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    dec = oracle.oct30_unpack(oracle.oct30_pack(v))
    a, b = dec.astype(np.float64), v.astype(np.float64)
    ang = np.arctan2(np.linalg.norm(np.cross(a, b), axis=1), np.sum(a * b, axis=1))
    assert np.max(ang) < 2e-4
    np.testing.assert_allclose(np.linalg.norm(dec, axis=1), 1.0, atol=1e-6)


def test_oct30_axes_and_layout(oracle):
    axes = np.array([[0, 0, 1], [0, 0, -1], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0]],
                    np.float32)
    p = oracle.oct30_pack(axes)
    dec = oracle.oct30_unpack(p)
    # 32767 is odd, so the square's centre 0.5 sits on a tie (16383.5 -> 16384)
    # and axes decode 2^-15 off: within 4e-5, not exact.
    np.testing.assert_allclose(dec, axes, atol=4e-5)
    # +z maps to the centre of the square: both 15-bit fields round(0.5*32767) = 16384 (tie -> even)
    assert int(p[0]) & 0x7FFF == 16384 and (int(p[0]) >> 15) & 0x7FFF == 16384
    # +x maps to (1, 0.5)
    assert int(p[2]) & 0x7FFF == 32767
    assert int(p[0]) >> 30 == 0


def test_oct30_ties_round_to_even(oracle):
    """WGSL round() is ties-to-even (packing.wgsl:33-34); Rust's f32::round
    in the reference's host twin is ties-away (crates/appearance-packing/src/lib.rs:55-56).
    These directions put the x field exactly on k + 0.5 with k even (the
    float32 chain |x|+|y|+|z|, x/s, *0.5+0.5, *32767 is exact for them), where
    the two rules differ: the codec follows WGSL, k, not k + 1."""
    f = np.float32
    for k, a in [(16386, 0.00018310546875), (16388, 0.00030517578125), (16384, 6.103515625e-05)]:
        a = f(a)
        b = f(f(1.0) - a)
        s = f(a + b)
        ex = f(f(f(a / s) * f(0.5)) + f(0.5))
        assert f(ex * f(32767.0)) == f(k + 0.5)      # an exact tie
        p = int(oracle.oct30_pack(np.array([[a, b, 0.0]], np.float32))[0])
        assert p & 0x7FFF == k, (k, p & 0x7FFF)
        assert k % 2 == 0
