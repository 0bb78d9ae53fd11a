"""The sky panorama (the reference's HDR sky texture, sky.rs:93-105 and
Sky::sky, shared/sky_bindings.wgsl:68-70) on the CPU:

* libapt's Radiance HDR decoder against an independent RGBE encoder
  (tests/hdr_util.py) with flat, run-length and repeat-marker scanlines,
  header variants and damaged files;
* the oracle's panorama lookup against the closed form of
  unit_vector_to_panorama_coords (assets/shaders/math.wgsl:101-106): a
  panorama holding its own texel coordinates returns, through bilinear
  filtering, the sample point (u W - 1/2, v H - 1/2) with u = (atan2(z, x) +
  pi) / 2 pi and v = acos(y) / pi;
* light transport under a panorama: a Lambertian plane under a sky of
  radiance cos(theta) above the horizon and none below has L = 2 rho / 3,
  under a constant sky L = rho L_sky; a panorama turned upside down goes dark.
"""
import numpy as np
import pytest

from appearance_amd._ffi import Sky
from appearance_amd.scene import decode_hdr
from tests import hdr_util
from tests._oracle import oracle_render, sky_probe
from tests.scenes_util import material, sun_plane


@pytest.mark.parametrize("style,w,h", [("flat", 7, 5), ("flat", 33, 4), ("rle", 8, 3), ("rle", 300, 6),
                                       ("old", 40, 5)])
def test_hdr_decoding(style, w, h):
    rng = np.random.default_rng(w * 7 + h)
    rgb = np.exp(rng.normal(0, 3, (h, w, 3)))
    rgb[:, w // 3: w // 3 + 5] = rgb[:, w // 3: w // 3 + 1]   # runs
    rgb[0, :3] = 0.0
    q = hdr_util.to_rgbe(rgb)
    data = hdr_util.encode(q, style)
    got = decode_hdr(data)
    assert got.shape == (h, w, 3)
    np.testing.assert_array_equal(got, hdr_util.from_rgbe(q))
    # the encoding is Ward's shared exponent: relative error below 2^-8 of the largest channel
    np.testing.assert_allclose(got, rgb, rtol=0, atol=float(rgb.max()) * 2 ** -7)


def test_hdr_header_variants_and_errors():
    q = hdr_util.to_rgbe(np.full((2, 9, 3), 0.75))
    for header in (b"#?RGBE\n", b"#?RADIANCE\n# a comment\nEXPOSURE=1.0\nFORMAT=32-bit_rle_rgbe\n"):
        np.testing.assert_array_equal(decode_hdr(hdr_util.encode(q, "rle", header)), hdr_util.from_rgbe(q))
    good = hdr_util.encode(q, "rle")
    bad = [b"P6\n" + good[11:],                                              # not Radiance
           good.replace(b"32-bit_rle_rgbe", b"32-bit_rle_xyze"),             # XYZ texels
           good.replace(b"-Y 2 +X 9", b"+Y 2 +X 9"),                         # flipped orientation
           good[:-5]]                                                        # truncated
    for data in bad:
        with pytest.raises(RuntimeError):
            decode_hdr(data)


def coords_panorama(w, h):
    """Texel (i, j) holds (i, j, 1): bilinear filtering returns the sample point."""
    j, i = np.mgrid[0:h, 0:w].astype(np.float32)
    return np.stack([i, j, np.ones_like(i)], -1)


def test_panorama_coordinates(oracle):
    w, h = 64, 32
    pano = coords_panorama(w, h)
    rng = np.random.default_rng(0)
    sky = Sky()
    dirs = [np.array(d, np.float64) for d in ((1, 0.2, 0), (0, 0.3, -1), (0, -0.4, 1), (0.3, 0.9, 0.2))]
    dirs += [rng.normal(size=3) for _ in range(50)]
    for d in dirs:
        d /= np.linalg.norm(d)
        u = (np.arctan2(d[2], d[0]) + np.pi) / (2 * np.pi)
        v = np.arccos(d[1]) / np.pi
        x, y = u * w - 0.5, v * h - 0.5
        if not (0.01 < x < w - 1.01 and 0.01 < y < h - 1.01):
            continue   # the repeat seams mix the first and last texels
        got = sky_probe(oracle, sky, d.astype(np.float32), pano)
        np.testing.assert_allclose(got, [x, y, 1.0], atol=2e-4, err_msg=str(d))
    # named directions: +x at u = 1/2, -z at 1/4, +z at 3/4; the horizon at v = 1/2
    for d, u in (((1, 0, 0), 0.5), ((0, 0, -1), 0.25), ((0, 0, 1), 0.75)):
        got = sky_probe(oracle, sky, np.array(d, np.float32), pano)
        np.testing.assert_allclose(got[:2], [u * w - 0.5, 0.5 * h - 0.5], atol=2e-4)


def test_panorama_replaces_the_gradient_not_the_sun(oracle):
    sky = Sky()
    sky.zenith[:] = [0.2, 0.4, 0.9]
    sky.horizon[:] = [0.9, 0.9, 0.9]
    sky.sun_enabled = 1.0
    sky.sun_direction[:] = [0, -1, 0]
    sky.sun_intensity = 50.0
    sky.sun_size = 0.5
    up = np.array([0, 1, 0], np.float32)
    np.testing.assert_allclose(sky_probe(oracle, sky, up), [0.2, 0.4, 0.9], rtol=1e-6)
    pano = np.full((8, 16, 3), [3.0, 2.0, 1.0], np.float32)
    np.testing.assert_allclose(sky_probe(oracle, sky, up, pano), [3.0, 2.0, 1.0], rtol=1e-6)


def cos_sky(w=64, h=256, upside_down=False):
    theta = (np.arange(h) + 0.5) / h * np.pi
    L = np.maximum(np.cos(theta), 0.0)
    if upside_down:
        L = L[::-1]
    return np.repeat(np.repeat(L[:, None, None], w, 1), 3, 2).astype(np.float32)


def plane_under_sky(rho=0.5):
    s = sun_plane(rho, (0.0, -1.0, 0.0))
    s.sky.sun_enabled = 0.0
    s.materials = [material(color=(rho,) * 3, bsdf_model=1)]
    return s


def test_lambertian_plane_under_a_cosine_sky(oracle):
    s = plane_under_sky(0.5)
    img, _ = oracle_render(oracle, s, 48, 48, spp=32, max_bounces=2, rr_start=100, sky_pano=cos_sky())
    want = 2 * 0.5 / 3
    assert abs(img.mean() - want) / want < 5e-3, (img.mean(), want)
    dark, _ = oracle_render(oracle, s, 48, 48, spp=8, max_bounces=2, rr_start=100,
                            sky_pano=cos_sky(upside_down=True))
    assert dark.mean() < 1e-3 * want


def test_lambertian_plane_under_a_constant_sky(oracle):
    s = plane_under_sky(0.7)
    img, _ = oracle_render(oracle, s, 32, 32, spp=8, max_bounces=2, rr_start=100,
                           sky_pano=np.full((16, 32, 3), [1.5, 1.0, 0.25], np.float32))
    np.testing.assert_allclose(img.reshape(-1, 3).mean(0), 0.7 * np.array([1.5, 1.0, 0.25]), rtol=2e-3)
