"""Pins of the oracle's ReSTIR integrator and SVGF passes (CPU only).

- With reuse off the ReSTIR estimator is unbiased: its mean over many frames
  equals the plain oracle path tracer at pixel centres, pixel by pixel within
  a z-score bound (the plain tracer is pinned by the furnace and FP64 checks).
- With reuse on, on a static camera, it stays unbiased: the mean over many
  independent frame sequences of the last frame's image equals the plain
  tracer for DI temporal reuse (M cap engaged), DI spatial reuse (unbiased
  combination with its Z count and the solid-angle Jacobian of reused light
  samples), GI temporal reuse, and all of them together; planted mistakes in
  Z, 1/M and the Jacobian make the check fail.
- GI spatial reuse (restir_gi_spatial.wgsl, off in the reference's
  configuration) is unbiased where the reconnection Jacobian it leaves out is
  1: a floor lit by a panorama sky; Z counting one domain fails that check.
- The a-trous pass keeps a constant image constant, reproduces a linear ramp
  inside a flat region, and spreads an impulse with the closed-form weights
  h(dx) h(dy) / (1 + 55/64), h = (3/8, 1/4, 1/16), at the pass's step size.
- The G-buffer's FP16 rounding equals numpy's float16 conversion.
"""
import numpy as np
import pytest

from appearance_amd.renderer import default_config
from appearance_amd.scene import generate_scene
from tests._oracle import OracleRestir, atrous_probe, f16_round, oracle_render


def test_f16_round_matches_numpy(oracle):
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * np.float32(10.0) ** rng.integers(-9, 6, 20000),
        np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, 1e9, -1e9, 2.0 ** -24, 2.0 ** -25, 2.0 ** -25 * 1.5,
                  6.1e-5, 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, np.inf, -np.inf], np.float32)])
    got = f16_round(oracle, x)
    with np.errstate(over="ignore"):
        want = x.astype(np.float16).astype(np.float32)
    np.testing.assert_array_equal(got, want)
    nan = f16_round(oracle, np.array([np.nan], np.float32))
    assert np.isnan(nan[0])


def _zscores(oracle, name, seed, w, h, bounces, frames, ref_batches, ref_spp):
    s = generate_scene(name, seed=seed)
    cfg = default_config(restir_di=0, restir_gi=0, firefly_filter=0)
    r = OracleRestir(oracle, s, w, h)
    acc = np.zeros((h, w, 3))
    acc2 = np.zeros((h, w, 3))
    for f in range(frames):
        img = r.frame(cfg, frame_idx=f, max_bounces=bounces).astype(np.float64)
        acc += img
        acc2 += img * img
    m = acc / frames
    var = np.maximum(acc2 / frames - m * m, 0.0) / frames
    refs = np.array([oracle_render(oracle, s, w, h, spp=ref_spp, max_bounces=bounces, rr_start=3,
                                   jitter=False, frame_idx=1000 + b)[0] for b in range(ref_batches)],
                    np.float64)
    rm = refs.mean(0)
    rv = refs.var(0, ddof=1) / ref_batches
    sd = np.sqrt(var + rv + (1e-6 * np.abs(rm)) ** 2 + 1e-20)
    return m, rm, sd


@pytest.mark.parametrize("name,seed,w,h,bounces", [("cornell", 0, 32, 31, 2), ("cornell", 0, 24, 23, 4),
                                                   ("disney", 3, 24, 23, 3)])
def test_restir_without_reuse_is_unbiased(oracle, name, seed, w, h, bounces):
    # Odd/even frame sizes keep pixel centres off the box's corner edges,
    # where the two tracers' different ray offsets could pick different walls.
    m, rm, sd = _zscores(oracle, name, seed, w, h, bounces, frames=256, ref_batches=8, ref_spp=256)
    z = (m - rm) / sd
    rel = abs(m.mean() - rm.mean()) / rm.mean()
    print(f"[restir] {name} {w}x{h} depth {bounces}: frame mean {m.mean():.6f} plain {rm.mean():.6f} "
          f"rel {rel:.2e} z rms {np.sqrt((z ** 2).mean()):.2f} max |z| {np.abs(z).max():.1f}")
    assert rel < 2e-3
    assert np.sqrt((z ** 2).mean()) < 1.6
    assert (np.abs(z) > 5).mean() < 2e-3
    # the check has the power to see a 2% bias
    zb = (m * 1.02 - rm) / sd
    assert np.sqrt((zb ** 2).mean()) > 1.6


def _plain_reference(oracle, s, w, h, bounces, bvh, batches=8, spp=256, sky_pano=None):
    refs = np.array([oracle_render(oracle, s, w, h, spp=spp, max_bounces=bounces, rr_start=3,
                                   jitter=False, frame_idx=1000 + b, bvh=bvh, sky_pano=sky_pano)[0]
                     for b in range(batches)],
                    np.float64)
    return refs.mean(0), refs.var(0, ddof=1) / batches


def _reuse_sequences(oracle, s, w, h, bounces, cfg, sequences, frames, bvh, sky_pano=None):
    """Mean and variance of the mean of the last frame of `sequences`
    independent frame sequences (a fresh reservoir history each, seeds far
    apart) on the scene's static camera."""
    acc = np.zeros((h, w, 3))
    acc2 = np.zeros((h, w, 3))
    for k in range(sequences):
        r = OracleRestir(oracle, s, w, h, bvh=bvh, sky_pano=sky_pano)
        for f in range(frames):
            img = r.frame(cfg, frame_idx=f, seed=7919 * k + 13, max_bounces=bounces, threads=1)
        img = img.astype(np.float64)
        acc += img
        acc2 += img * img
    m = acc / sequences
    return m, np.maximum(acc2 / sequences - m * m, 0.0) / sequences


def _reuse_z(m, var, rm, rv):
    sd = np.sqrt(var + rv + (1e-6 * np.abs(rm)) ** 2 + 1e-20)
    z = (m - rm) / sd
    return float(np.sqrt((z ** 2).mean())), float((np.abs(z) > 5).mean()), \
        float(np.sqrt((((m * 1.02 - rm) / sd) ** 2).mean()))


# Reuse modes: config overrides and frames per sequence.  The caps are small
# so they engage within the sequence (M grows by the 5 candidates of a frame).
REUSE_MODES = {
    "di_temporal": (dict(restir_di=1, restir_gi=0, di_spatial_neighbours=0, di_mcap=2), 4),
    "di_spatial": (dict(restir_di=1, restir_gi=0, di_spatial_neighbours=5, di_spatial_radius=6.0), 1),
    "gi_temporal": (dict(restir_di=0, restir_gi=1, gi_mcap=2), 4),
    "all": (dict(restir_di=1, restir_gi=1, di_spatial_neighbours=5, di_spatial_radius=6.0, di_mcap=2,
                 gi_mcap=3), 4),
}


@pytest.fixture(scope="module")
def cornell_reuse(oracle):
    s = generate_scene("cornell")
    bvh = oracle.bvh(np.ascontiguousarray(s.positions, np.float32))
    return s, bvh, _plain_reference(oracle, s, 24, 23, 2, bvh)


@pytest.mark.parametrize("mode", list(REUSE_MODES))
def test_restir_with_reuse_is_unbiased(oracle, cornell_reuse, mode):
    s, bvh, (rm, rv) = cornell_reuse
    kw, frames = REUSE_MODES[mode]
    cfg = default_config(firefly_filter=0, **kw)
    m, var = _reuse_sequences(oracle, s, 24, 23, 2, cfg, 512, frames, bvh)
    zrms, tail, power = _reuse_z(m, var, rm, rv)
    rel = abs(m.mean() - rm.mean()) / rm.mean()
    print(f"[restir] cornell {mode}: mean {m.mean():.6f} plain {rm.mean():.6f} rel {rel:.2e} z rms {zrms:.2f}")
    assert rel < 3e-3
    assert zrms < 1.5
    assert tail < 5e-3
    assert power > 1.6


@pytest.mark.parametrize("mode,plant_kind", [("di_spatial", 1), ("di_spatial", 5), ("di_spatial", 6),
                                             ("di_temporal", 3)])
def test_restir_reuse_planted_errors_are_caught(oracle, cornell_reuse, mode, plant_kind):
    """Z = M1 + M2, the two domains' M swapped in Z, no Jacobian on reused
    light samples, no 1/M: each biases the reuse check beyond its bound."""
    from tests.test_oracle_reservoir import plant
    s, bvh, (rm, rv) = cornell_reuse
    kw, frames = REUSE_MODES[mode]
    cfg = default_config(firefly_filter=0, **kw)
    with plant(oracle, plant_kind):
        m, var = _reuse_sequences(oracle, s, 24, 23, 2, cfg, 512, frames, bvh)
    zrms, _, _ = _reuse_z(m, var, rm, rv)
    print(f"[restir] cornell {mode} with plant {plant_kind}: z rms {zrms:.2f}")
    assert zrms > 2.5


def test_restir_with_all_reuse_is_unbiased_on_the_many_light_scene(oracle):
    """The Disney spheres with 4096 emissive triangles and the sun: every
    reuse pass on, three frames per sequence."""
    s = generate_scene("restir", seed=4)
    bvh = oracle.bvh(np.ascontiguousarray(s.positions, np.float32))
    rm, rv = _plain_reference(oracle, s, 24, 23, 2, bvh)
    kw, _ = REUSE_MODES["all"]
    m, var = _reuse_sequences(oracle, s, 24, 23, 2, default_config(firefly_filter=0, **kw), 512, 3, bvh)
    zrms, tail, power = _reuse_z(m, var, rm, rv)
    rel = abs(m.mean() - rm.mean()) / rm.mean()
    print(f"[restir] many-light all reuse: mean {m.mean():.6f} plain {rm.mean():.6f} rel {rel:.2e} "
          f"z rms {zrms:.2f}")
    assert rel < 6e-3
    assert zrms < 1.5
    assert tail < 5e-3
    assert power > 1.6


GI_SPATIAL = dict(restir_di=0, restir_gi=1, gi_mcap=2, gi_spatial_passes=2)


@pytest.fixture(scope="module")
def sky_plane_reuse(oracle):
    """A grey Lambertian floor under a cosine-weighted panorama sky (no sun), seen from 2 m:
    every indirect sample is a sky point 1000 units out, so a neighbour's
    sample seen from this pixel is the same direction to within 1e-3 and the
    reconnection Jacobian that GI spatial reuse leaves out
    (inline_path_tracer.wgsl:321) is 1 to that accuracy."""
    from tests.test_sky_panorama import cos_sky, plane_under_sky
    s = plane_under_sky(0.5)
    bvh = oracle.bvh(np.ascontiguousarray(s.positions, np.float32))
    pano = cos_sky()
    return s, bvh, pano, _plain_reference(oracle, s, 24, 23, 2, bvh, batches=8, spp=128, sky_pano=pano)


def test_restir_gi_spatial_reuse_is_unbiased_under_a_sky(oracle, sky_plane_reuse):
    """Two GI spatial passes after GI temporal reuse (M cap engaged), two
    frames per sequence: the unbiased combination's Z counts both domains."""
    s, bvh, pano, (rm, rv) = sky_plane_reuse
    m, var = _reuse_sequences(oracle, s, 24, 23, 2, default_config(firefly_filter=0, **GI_SPATIAL), 256, 2,
                              bvh, pano)
    zrms, tail, power = _reuse_z(m, var, rm, rv)
    rel = abs(m.mean() - rm.mean()) / rm.mean()
    print(f"[restir] sky plane gi spatial: mean {m.mean():.6f} plain {rm.mean():.6f} rel {rel:.2e} "
          f"z rms {zrms:.2f}")
    assert rel < 3e-3
    assert zrms < 1.5
    assert tail < 5e-3
    assert power > 1.6


def test_restir_gi_spatial_planted_z_error_is_caught(oracle, sky_plane_reuse):
    """Z counting the pixel's own domain only (W too large by (M1 + M2) / M1)."""
    from tests.test_oracle_reservoir import plant
    s, bvh, pano, (rm, rv) = sky_plane_reuse
    with plant(oracle, 9):
        m, var = _reuse_sequences(oracle, s, 24, 23, 2, default_config(firefly_filter=0, **GI_SPATIAL), 256, 2,
                                  bvh, pano)
    zrms, _, _ = _reuse_z(m, var, rm, rv)
    print(f"[restir] sky plane gi spatial with plant 9: z rms {zrms:.2f}")
    assert zrms > 2.5


def test_restir_gi_spatial_passes_change_the_frame(oracle):
    """On the Cornell box the passes run (the frame differs from GI temporal
    alone, same seeds) and stay finite; zero passes is GI temporal exactly."""
    s = generate_scene("cornell")
    imgs = {}
    for n in (0, 1, 3):
        r = OracleRestir(oracle, s, 24, 23)
        cfg = default_config(firefly_filter=0, **dict(GI_SPATIAL, gi_spatial_passes=n))
        for f in range(2):
            imgs[n] = r.frame(cfg, frame_idx=f, seed=5, max_bounces=2, threads=1)
    r = OracleRestir(oracle, s, 24, 23)
    cfg = default_config(firefly_filter=0, restir_di=0, restir_gi=1, gi_mcap=2)
    for f in range(2):
        base = r.frame(cfg, frame_idx=f, seed=5, max_bounces=2, threads=1)
    np.testing.assert_array_equal(imgs[0], base)
    for n in (1, 3):
        assert np.isfinite(imgs[n]).all()
        assert not np.array_equal(imgs[n], base)


def _flat(h, w):
    depth = np.ones((h, w), np.float32)
    nrm = np.zeros((h, w, 3), np.float32)
    nrm[..., 2] = 1.0
    return depth, nrm


@pytest.mark.parametrize("pass_idx", [0, 1, 2, 3, 4])
def test_atrous_constant_image_is_fixed(oracle, pass_idx):
    h, w = 40, 44
    depth, nrm = _flat(h, w)
    col = np.broadcast_to(np.array([0.3, 1.7, 0.05], np.float32), (h, w, 3)).copy()
    var = np.random.default_rng(pass_idx).random((h, w)).astype(np.float32)
    out = atrous_probe(oracle, col, var, depth, nrm, pass_idx)
    np.testing.assert_allclose(out, col, rtol=2e-6)


@pytest.mark.parametrize("pass_idx", [0, 1, 2])
def test_atrous_linear_ramp_inside(oracle, pass_idx):
    h, w = 48, 48
    depth, nrm = _flat(h, w)
    x = np.arange(w, dtype=np.float32)[None, :, None]
    y = np.arange(h, dtype=np.float32)[:, None, None]
    col = (np.float32(0.5) + np.float32(0.01) * x + np.float32(0.004) * y) * np.ones((1, 1, 3), np.float32)
    col = np.ascontiguousarray(col, np.float32)
    var = np.full((h, w), 0.02, np.float32)
    out = atrous_probe(oracle, col, var, depth, nrm, pass_idx)
    b = 2 << pass_idx
    np.testing.assert_allclose(out[b:h - b, b:w - b], col[b:h - b, b:w - b], rtol=2e-6)


@pytest.mark.parametrize("pass_idx", [0, 1, 3])
def test_atrous_impulse_weights(oracle, pass_idx):
    h, w = 72, 72
    depth, nrm = _flat(h, w)
    step = 1 << pass_idx
    hk = {0: 3 / 8, 1: 1 / 4, 2: 1 / 16}
    cy = cx = 36
    for dy, dx in [(0, 1), (1, 1), (2, 0), (-2, -1), (2, 2), (-1, 0)]:
        col = np.zeros((h, w, 3), np.float32)
        col[cy + dy * step, cx + dx * step] = 1.0
        var = np.full((h, w), 1e12, np.float32)  # luma stopping off
        out = atrous_probe(oracle, col, var, depth, nrm, pass_idx)
        want = hk[abs(dx)] * hk[abs(dy)] / (1.0 + 55.0 / 64.0)
        np.testing.assert_allclose(out[cy, cx], want, rtol=1e-5)
        # a tap between the kernel's taps (step > 1) contributes nothing
        if step > 1:
            col2 = np.zeros((h, w, 3), np.float32)
            col2[cy + 1, cx] = 1.0
            assert np.all(atrous_probe(oracle, col2, var, depth, nrm, pass_idx)[cy, cx] == 0.0)


def test_atrous_edge_stopping(oracle):
    """Depth and normal stopping: a neighbour at another depth d is weighted by
    exp(-d^2 / 0.02), one with a normal at angle a by cos(a)^128."""
    h, w = 16, 16
    depth, nrm = _flat(h, w)
    var = np.full((h, w), 1e12, np.float32)
    col = np.zeros((h, w, 3), np.float32)
    col[8, 9] = 1.0
    depth2 = depth.copy()
    depth2[8, 9] = 1.25
    out = atrous_probe(oracle, col, var, depth2, nrm, 0)
    wd = np.exp(-(0.25 ** 2) / 0.02)
    tot = 1.0 + 55.0 / 64.0 - (1 / 4 * 3 / 8) * (1 - wd)
    np.testing.assert_allclose(out[8, 8, 0], (1 / 4 * 3 / 8) * wd / tot, rtol=1e-5)
    nrm2 = nrm.copy()
    a = 0.1
    nrm2[8, 9] = [np.sin(a), 0.0, np.cos(a)]
    out = atrous_probe(oracle, col, var, depth, nrm2, 0)
    c = float(np.float32(np.cos(a)))
    wn = c ** 128
    tot = 1.0 + 55.0 / 64.0 - (1 / 4 * 3 / 8) * (1 - wn)
    np.testing.assert_allclose(out[8, 8, 0], (1 / 4 * 3 / 8) * wn / tot, rtol=1e-4)


# ---- firefly filter, SVGF temporal pass and TAA: closed-form pins -------------
# The oracle's restatements of firefly_filter.wgsl, svgf_temporal.wgsl and
# taa.wgsl run through or_post_probe on a static camera facing a flat wall.

def _luma(c):
    return 0.2126 * c[..., 0] + 0.7152 * c[..., 1] + 0.0722 * c[..., 2]


def test_firefly_constant_image_is_fixed(oracle):
    from tests._oracle import post_probe
    img = np.broadcast_to(np.array([0.3, 0.5, 0.2], np.float32), (1, 9, 11, 3)).copy()
    out, _ = post_probe(oracle, 0, img)
    np.testing.assert_array_equal(out, img[0])


def test_firefly_clamps_an_impulse_to_its_neighbourhood(oracle):
    from tests._oracle import post_probe
    base = np.array([0.3, 0.5, 0.2], np.float32)
    img = np.broadcast_to(base, (1, 9, 11, 3)).copy()
    img[0, 4, 5] = [40.0, 35.0, 50.0]      # a firefly: replaced by the neighbourhood's brightest
    img[0, 2, 2] = [0.0, 0.001, 0.0]       # a dark outlier: replaced by its darkest neighbour
    out, _ = post_probe(oracle, 0, img)
    np.testing.assert_array_equal(out[4, 5], base)
    np.testing.assert_array_equal(out[2, 2], base)
    mask = np.ones((9, 11), bool)
    mask[4, 5] = mask[2, 2] = False
    np.testing.assert_array_equal(out[mask], img[0][mask])  # in range of their neighbours: untouched


def test_firefly_ignores_sky(oracle):
    from tests._oracle import post_probe
    rng = np.random.default_rng(3)
    img = rng.uniform(0, 5, (1, 7, 8, 3)).astype(np.float32)
    sky = np.zeros((7, 8), np.uint8)
    sky[:, :4] = 1
    img[0, 3, 6] = 1e4                      # surrounded by sky on one side only
    out, _ = post_probe(oracle, 0, img, sky=sky)
    np.testing.assert_array_equal(out[:, :4], img[0][:, :4])   # sky passes through
    sky2 = np.ones((7, 8), np.uint8)
    sky2[3, 3] = 0                          # a surface pixel whose neighbours are all sky
    out2, _ = post_probe(oracle, 0, img, sky=sky2)
    np.testing.assert_array_equal(out2[3, 3], img[0, 3, 3])


def test_svgf_temporal_first_frame_passes_through(oracle):
    from tests._oracle import post_probe
    rng = np.random.default_rng(4)
    img = rng.uniform(0, 2, (1, 6, 9, 3)).astype(np.float32)
    out, var = post_probe(oracle, 1, img)
    np.testing.assert_array_equal(out, img[0])
    assert np.all(var == 0.0)


@pytest.mark.parametrize("cap", [4, 24])
def test_svgf_temporal_is_a_running_mean_then_an_ema(oracle, cap):
    """Frame-constant colours c_0..c_{F-1} on a static view: the history
    length k grows 0, 1, .. up to the cap and alpha = 1 / (1 + k), so the
    output is the mean of the frames seen while k <= cap and an exponential
    average with weight 1 / (1 + cap) after that; the variance is that of the
    luma moments."""
    from tests._oracle import post_probe
    rng = np.random.default_rng(5)
    frames = 12
    cols = rng.uniform(0.1, 3.0, (frames, 3))
    imgs = np.broadcast_to(cols[:, None, None, :], (frames, 8, 10, 3)).astype(np.float32)
    m = np.zeros(3)
    m1 = m2 = 0.0
    for f in range(frames):
        a = 1.0 / (1.0 + min(f, cap))
        c = imgs[f, 0, 0].astype(np.float64)
        m = m + (c - m) * a
        lum = _luma(c)
        m1 = m1 + (lum - m1) * a
        m2 = m2 + (lum * lum - m2) * a
        out, var = post_probe(oracle, 1, imgs[:f + 1], svgf_history=cap)
        inner = out[1:-1, 1:-1]
        np.testing.assert_allclose(inner, np.broadcast_to(m, inner.shape), rtol=2e-6)
        np.testing.assert_allclose(var[1:-1, 1:-1], max(m2 - m1 * m1, 0.0), rtol=1e-3, atol=1e-6)
        if f == frames - 1 and cap >= frames:
            np.testing.assert_allclose(m, imgs[:, 0, 0].astype(np.float64).mean(axis=0), rtol=1e-12)


def test_svgf_temporal_sky_has_no_history(oracle):
    from tests._oracle import post_probe
    sky = np.zeros((6, 7), np.uint8)
    sky[2:4, 2:5] = 1
    imgs = np.stack([np.full((6, 7, 3), v, np.float32) for v in (1.0, 3.0, 5.0)])
    out, var = post_probe(oracle, 1, imgs, sky=sky)
    np.testing.assert_array_equal(out[2:4, 2:5], imgs[-1][2:4, 2:5])
    assert np.all(var[2:4, 2:5] == 0.0)


def test_taa_first_frame_passes_through_and_constant_frames_are_fixed(oracle):
    from tests._oracle import post_probe
    rng = np.random.default_rng(6)
    img = rng.uniform(0, 2, (1, 6, 9, 3)).astype(np.float32)
    out, _ = post_probe(oracle, 2, img)
    np.testing.assert_array_equal(out, img[0])
    # a constant frame after any history: the neighbourhood box has zero
    # extent, so the history clips to its mean and the output is the frame
    # (up to the box's FP32 width: the variance sum(x^2) - sum(x)^2 / n of
    # equal values cancels to a few ulps, whose square root is ~3e-4 |c|)
    c = np.array([0.7, 0.2, 1.3], np.float32)
    imgs = np.concatenate([img, np.broadcast_to(c, (1, 6, 9, 3))])
    out, _ = post_probe(oracle, 2, imgs)
    np.testing.assert_allclose(out[1:-1, 1:-1], np.broadcast_to(c, (4, 7, 3)), rtol=2e-3)
    # ... and with the same constant as history the box test leaves it alone
    imgs = np.broadcast_to(c, (3, 6, 9, 3)).copy()
    out, _ = post_probe(oracle, 2, imgs)
    np.testing.assert_allclose(out, imgs[-1], rtol=1e-6)


def test_taa_clips_history_to_the_neighbourhood_box(oracle):
    """One interior pixel's 3x3 neighbourhood with known mean and deviation:
    the history (a constant far outside the box) is clipped onto the box
    surface along the line to the box centre, and blended with the current
    pixel by luma-weighted factors 0.2 and 0.8 (Mitchell-Netravali weights
    vanish at the neighbours: the filter is evaluated at 2 |dx| = 2, its
    support's edge, so the reconstruction is the centre pixel itself)."""
    from tests._oracle import post_probe
    rng = np.random.default_rng(7)
    cur = rng.uniform(0.2, 1.0, (7, 7, 3)).astype(np.float32)
    hist_c = np.array([5.0, -3.0, 4.0], np.float32)
    imgs = np.stack([np.broadcast_to(hist_c, (7, 7, 3)), cur])
    out, _ = post_probe(oracle, 2, imgs)
    y, x = 3, 3
    nb = cur[y - 1:y + 2, x - 1:x + 2].reshape(9, 3).astype(np.float64)
    mean = nb.mean(axis=0)
    sd = nb.std(axis=0, ddof=1)
    r = hist_c - mean
    k = np.max(np.abs(r / sd))
    clipped = mean + r / k if k > 1 else hist_c
    assert np.isclose(np.max(np.abs((clipped - mean) / sd)), 1.0)   # on the box surface
    rec = cur[y, x].astype(np.float64)
    cw = 0.2 / (1.0 + _luma(rec))
    hw = 0.8 / (1.0 + _luma(clipped))
    expect = (rec * cw + clipped * hw) / (cw + hw)
    np.testing.assert_allclose(out[y, x], expect, rtol=2e-5)
