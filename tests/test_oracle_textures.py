"""Pins of the oracle's textured materials (CPU only): the sampler, the
material texture slots, uv interpolation and normal mapping of
shared/material/material_pool_bindings.wgsl:15-154 and
helpers/inline_path_tracer.wgsl:82-165, each checked against what the
mathematics fixes rather than a retyped formula:

* bilinear repeat sampling: texel centres return the texel, midpoints the
  mean of the two texels, a bilinear function of the texel indices is
  reproduced exactly between centres, the sampler has period 1 in u and v,
  and the seam at u = 0 mixes the last and the first column;
* srgb_to_linear against the IEC 61966-2-1 anchors (128 -> 0.2158605,
  188 -> 0.5028864) and its linear toe; the colour's alpha is the luminance;
  emission is linear;
* every factor reads its own channel (metallic B, roughness G, clearcoat R,
  clearcoat roughness G, transmission R, sheen R) of a texture whose four
  channels all differ;
* normal maps: axis texels map onto the tangent, the bitangent and the
  normal; the map is equivariant under rotations of the frame; scale 1/2 is
  the bisector of the normal and the map; the clearcoat normal is the bare
  normal, the mapped normal, or its own map as the material's textures say;
* the renderer's light transport with textures: a sun-lit Lambertian plane
  under a base-colour texture, under a normal map (the cosine to the bent
  normal), an emitter seen directly through an emission texture, and a floor
  lit through one;
* the ReSTIR oracle with reuse off is unbiased against the plain oracle on a
  room that uses every texture slot.
"""
import ctypes as C

import numpy as np
import pytest

from appearance_amd._ffi import Sky
from appearance_amd.renderer import default_config
from appearance_amd.scene import SceneData
from tests._oracle import OracleRestir, oracle_render, shading_probe, tex_sample
from tests.scenes_util import camera, material, sun_plane, sun_plane_radiance

NO_TEX = 0xFFFFFFFF


def rgba(r, g, b, a=255):
    return np.uint32(r | (g << 8) | (b << 16) | (a << 24))


def tex_of(rng, h, w):
    return rng.integers(0, 2 ** 32, (h, w), dtype=np.uint64).astype(np.uint32)


def channels(t):
    t = np.asarray(t, np.uint32)
    return np.stack([(t >> (8 * k)) & 0xFF for k in range(4)], -1).astype(np.float64) / 255.0


# ---- sampler -------------------------------------------------------------------

def test_texel_centres_and_channel_order(oracle):
    rng = np.random.default_rng(0)
    t = tex_of(rng, 5, 7)
    for j in range(5):
        for i in range(7):
            got = tex_sample(oracle, t, (i + 0.5) / 7, (j + 0.5) / 5)
            np.testing.assert_allclose(got, channels(t[j, i]), atol=2e-6)
    one = np.array([[0x04030201]], np.uint32)
    np.testing.assert_allclose(tex_sample(oracle, one, 0.3, 0.8), np.array([1, 2, 3, 4]) / 255.0, atol=1e-7)


def test_midpoints_average_neighbours(oracle):
    rng = np.random.default_rng(1)
    t = tex_of(rng, 4, 8)
    c = channels(t)
    for i in range(7):
        got = tex_sample(oracle, t, (i + 1) / 8, 2.5 / 4)
        np.testing.assert_allclose(got, 0.5 * (c[2, i] + c[2, i + 1]), atol=2e-6)
    for j in range(3):
        got = tex_sample(oracle, t, 5.5 / 8, (j + 1) / 4)
        np.testing.assert_allclose(got, 0.5 * (c[j, 5] + c[j + 1, 5]), atol=2e-6)


def test_bilinear_function_is_reproduced(oracle):
    """f(i, j) = a + b i + c j + d i j on the texel grid: bilinear filtering
    between the centres returns f at the sample point exactly."""
    h, w = 6, 8
    j, i = np.mgrid[0:h, 0:w].astype(np.float64)
    f = 10 + 15 * i + 12 * j + 2 * i * j          # 10 .. 245, exact bytes
    t = (f.astype(np.uint32) * 0x010101 + 0xFF000000).astype(np.uint32)
    rng = np.random.default_rng(2)
    for _ in range(200):
        x, y = rng.uniform(0, w - 1), rng.uniform(0, h - 1)   # texel-space point inside the centres
        u, v = np.float32((x + 0.5) / w), np.float32((y + 0.5) / h)
        xs, ys = float(u) * w - 0.5, float(v) * h - 0.5
        got = tex_sample(oracle, t, u, v)
        want = (10 + 15 * xs + 12 * ys + 2 * xs * ys) / 255.0
        np.testing.assert_allclose(got[:3], want, rtol=0, atol=3e-6)


def test_repeat_period_and_seam(oracle):
    rng = np.random.default_rng(3)
    t = tex_of(rng, 8, 16)
    for _ in range(100):
        u, v = rng.integers(0, 4096) / 4096.0, rng.integers(0, 4096) / 4096.0
        base = tex_sample(oracle, t, u, v)
        for du, dv in ((1, 0), (0, 1), (-2, 3), (5, -1)):
            np.testing.assert_allclose(tex_sample(oracle, t, u + du, v + dv), base, atol=2e-6)
    c = channels(t)
    # u = 0 sits half way between the last and the first column
    np.testing.assert_allclose(tex_sample(oracle, t, 0.0, 3.5 / 8), 0.5 * (c[3, 15] + c[3, 0]), atol=2e-6)
    np.testing.assert_allclose(tex_sample(oracle, t, 6.5 / 16, 0.0), 0.5 * (c[7, 6] + c[0, 6]), atol=2e-6)


# ---- material slots ------------------------------------------------------------

def tri_scene(textures, uvs=((0, 0), (1, 0), (0, 1)), tan=None, bit=None):
    """One triangle in z = 0 with the given uvs and textures (a SceneData)."""
    pos = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32)
    s = SceneData(pos, np.tile([0, 0, 1], (1, 3)).astype(np.float32), np.zeros(1, np.uint32), [material()],
                  camera((0, 0, 1), (0, 0, 0)), Sky(), uvs=np.asarray(uvs, np.float32).reshape(1, 6),
                  tangents=np.zeros((1, 12), np.float32), textures=[np.asarray(t, np.uint32) for t in textures])
    if tan is None:
        tan = np.tile([1, 0, 0], 3)
    if bit is None:
        bit = np.tile([0, 1, 0], 3)
    return s, (np.asarray(tan, np.float32).reshape(1, 9), np.asarray(bit, np.float32).reshape(1, 9))


def probe(oracle, s, frames, mat, n=(0, 0, 1), bu=0.25, bv=0.25):
    return shading_probe(oracle, s, mat, 0, bu, bv, np.asarray(n, np.float32), *frames)


def srgb(c):
    return np.where(c < 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


def test_srgb_anchors_alpha_and_linear_emission(oracle):
    for c, want in ((0, 0.0), (255, 1.0), (128, 0.2158605), (188, 0.5028864), (10, 10 / 255 / 12.92)):
        s, fr = tri_scene([[[rgba(c, c, c, 128)]]])
        m = material(color=(1.0, 0.5, 0.25), color_texture=0, emission=(2.0, 2.0, 2.0), emission_texture=0,
                     sheen=1.0, sheen_tint=(1.0, 1.0, 1.0), sheen_tint_texture=0)
        p = probe(oracle, s, fr, m)
        np.testing.assert_allclose(p["color"], np.array([1.0, 0.5, 0.25]) * want, rtol=2e-6, atol=1e-9)
        np.testing.assert_allclose(p["sheen_tint"], [want] * 3, rtol=2e-6, atol=1e-9)
        np.testing.assert_allclose(p["emission"], [2.0 * c / 255] * 3, rtol=2e-6)   # no sRGB on emission
        np.testing.assert_allclose(p["luminance"], 128 / 255, rtol=2e-6)
    # untextured: the factors, luminance 1
    s, fr = tri_scene([[[rgba(9, 9, 9)]]])
    p = probe(oracle, s, fr, material(color=(0.3, 0.2, 0.1), emission=(1.0, 2.0, 3.0)))
    np.testing.assert_allclose(p["color"], [0.3, 0.2, 0.1], rtol=1e-7)
    np.testing.assert_allclose(p["emission"], [1, 2, 3])
    assert p["luminance"] == 1.0


def test_each_factor_reads_its_channel(oracle):
    r, g, b, a = 51, 102, 153, 204
    s, fr = tri_scene([[[rgba(r, g, b, a)]]])
    m = material(metallic=0.8, roughness=0.6, metallic_roughness_texture=0, clearcoat=0.9, clearcoat_texture=0,
                 clearcoat_roughness=0.7, clearcoat_roughness_texture=0, transmission=0.4, transmission_texture=0,
                 sheen=0.5, sheen_texture=0)
    p = probe(oracle, s, fr, m)
    for name, factor, ch in (("metallic", 0.8, b), ("roughness", 0.6, g), ("clearcoat", 0.9, r),
                             ("clearcoat_roughness", 0.7, g), ("transmission", 0.4, r), ("sheen", 0.5, r)):
        np.testing.assert_allclose(p[name], factor * ch / 255, rtol=2e-6, err_msg=name)


def test_uv_interpolation(oracle):
    uvs = ((0.25, 2.0), (-1.5, 0.5), (3.0, -0.75))
    s, fr = tri_scene([[[rgba(1, 1, 1)]]], uvs=uvs)
    U = np.array(uvs)
    for bu, bv in ((0, 0), (1, 0), (0, 1), (1 / 3, 1 / 3), (0.1, 0.7)):
        p = probe(oracle, s, fr, material(), bu=bu, bv=bv)
        np.testing.assert_allclose(p["uv"], U[0] * (1 - bu - bv) + U[1] * bu + U[2] * bv, atol=2e-6)


# ---- normal mapping --------------------------------------------------------------

def nmap_of(oracle, texel, n=(0, 0, 1), tan=(1, 0, 0), bit=(0, 1, 0), scale=1.0, cc=None):
    tex = [[[texel]]]
    if cc is not None:
        tex.append([[cc]])
    s, fr = tri_scene(tex, tan=np.tile(tan, 3), bit=np.tile(bit, 3))
    m = material(normal_texture=0, normal_scale=scale)
    if cc is not None:
        m.clearcoat, m.clearcoat_normal_texture = 1.0, 1
    return probe(oracle, s, fr, m, n=n)


def test_axis_texels_map_onto_the_frame(oracle):
    for texel, axis in ((rgba(255, 128, 128), (1, 0, 0)), (rgba(0, 128, 128), (-1, 0, 0)),
                        (rgba(128, 255, 128), (0, 1, 0)), (rgba(128, 0, 128), (0, -1, 0)),
                        (rgba(128, 128, 255), (0, 0, 1))):
        p = nmap_of(oracle, texel)
        assert abs(np.linalg.norm(p["nmap"]) - 1) < 1e-6
        assert p["nmap"] @ np.array(axis) > 0.9999, (texel, p["nmap"])


def test_normal_map_is_rotation_equivariant(oracle):
    rng = np.random.default_rng(5)
    texel = rgba(190, 70, 220)
    base = nmap_of(oracle, texel)["nmap"].astype(np.float64)
    for _ in range(5):
        q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        if np.linalg.det(q) < 0:
            q[:, 0] = -q[:, 0]
        p = nmap_of(oracle, texel, n=q[:, 2], tan=q[:, 0], bit=q[:, 1])
        np.testing.assert_allclose(p["nmap"], q @ base, atol=3e-6)


def test_normal_scale(oracle):
    texel = rgba(230, 60, 200)
    full = nmap_of(oracle, texel)["nmap"].astype(np.float64)
    half = nmap_of(oracle, texel, scale=0.5)["nmap"].astype(np.float64)
    n = np.array([0.0, 0.0, 1.0])
    # the bisector of n and the full map
    assert abs(np.dot(np.cross(n, full), half)) < 1e-6
    assert abs(np.arccos(np.clip(n @ half, -1, 1)) - np.arccos(np.clip(half @ full, -1, 1))) < 1e-5
    np.testing.assert_array_equal(nmap_of(oracle, texel, scale=0.0)["nmap"], n)   # scale 0: no mapping


def test_clearcoat_normal_choice(oracle):
    texel, other = rgba(230, 60, 200), rgba(40, 200, 180)
    n = np.array([0, 0, 1], np.float32)
    # no clearcoat normal texture: the bare normal
    np.testing.assert_array_equal(nmap_of(oracle, texel, scale=0.5)["ncc"], n)
    # its own texture: its own map, without the normal scale
    p = nmap_of(oracle, texel, scale=0.5, cc=other)
    np.testing.assert_array_equal(p["ncc"], nmap_of(oracle, other)["nmap"])
    # the same texture as the normal map: the mapped (and scaled) normal
    s, fr = tri_scene([[[texel]]])
    m = material(normal_texture=0, normal_scale=0.5, clearcoat=1.0, clearcoat_normal_texture=0)
    p = probe(oracle, s, fr, m)
    np.testing.assert_array_equal(p["ncc"], p["nmap"])
    np.testing.assert_array_equal(p["nmap"], nmap_of(oracle, texel, scale=0.5)["nmap"])


def test_no_tangents_no_mapping(oracle):
    s, _ = tri_scene([[[rgba(255, 128, 128)]]])
    s.tangents = None
    m = material(normal_texture=0)
    p = shading_probe(oracle, s, m, 0, 0.2, 0.2, np.array([0, 0, 1], np.float32))
    np.testing.assert_array_equal(p["nmap"], [0, 0, 1])


# ---- light transport --------------------------------------------------------------

def textured_plane(s, textures, mat, tan=(1, 0, 0), bit=(0, 0, -1)):
    """sun_plane's floor with uvs, a tangent frame and textures."""
    s.materials = [mat]
    s.uvs = np.tile([0.1, 0.2, 0.3, 0.2, 0.2, 0.4], (s.n_tris, 1)).astype(np.float32)
    s.tangents = np.zeros((s.n_tris, 12), np.float32)
    s.textures = [np.asarray(t, np.uint32) for t in textures]
    return (np.tile(tan, (s.n_tris, 3)).astype(np.float32), np.tile(bit, (s.n_tris, 3)).astype(np.float32))


def test_sun_plane_under_a_colour_texture(oracle):
    s = sun_plane(0.6, (-0.3, -1.0, 0.2))
    fr = textured_plane(s, [[[rgba(128, 128, 128)]]], material(color=(0.6,) * 3, bsdf_model=1, color_texture=0))
    img, _ = oracle_render(oracle, s, 64, 64, spp=8, max_bounces=1, rr_start=100, frames=fr)
    want = sun_plane_radiance(0.6 * 0.2158605, s.sky)
    assert abs(img.mean() - want) / want < 2e-3, (img.mean(), want)


def test_sun_plane_under_a_normal_map(oracle):
    """Lambertian floor (normal +y, tangent +x, bitangent -z) whose normal map
    bends the shading normal: L = rho L_sun sin^2(alpha) (l . n'), n' the
    normalised T x + B y + N z of the texel's (x, y, z) in [-1, 1]."""
    s = sun_plane(0.6, (-0.3, -1.0, 0.2))
    texel = (160, 90, 240)
    fr = textured_plane(s, [[[rgba(*texel)]]], material(color=(0.6,) * 3, bsdf_model=1, normal_texture=0))
    img, _ = oracle_render(oracle, s, 64, 64, spp=8, max_bounces=1, rr_start=100, frames=fr)
    x, y, z = (np.array(texel) / 255.0) * 2 - 1
    nb = np.array([1.0, 0, 0]) * x + np.array([0, 0, -1.0]) * y + np.array([0, 1.0, 0]) * z
    nb /= np.linalg.norm(nb)
    l = -np.array(s.sky.sun_direction[:], np.float64)
    l /= np.linalg.norm(l)
    flat = sun_plane_radiance(0.6, s.sky)
    want = flat * (l @ nb) / l[1]
    assert abs(want / flat - 1) > 0.05   # the map matters
    assert abs(img.mean() - want) / want < 2e-3, (img.mean(), want, flat)


def emitter_and_floor(emission, tex=None):
    """An emissive quad facing down over a grey floor, camera looking at both."""
    quad = lambda a, b, c, d: [np.concatenate([a, b, c]), np.concatenate([a, c, d])]
    v = lambda *x: np.array(x, np.float32)
    pos = np.array(quad(v(-5, 0, 5), v(5, 0, 5), v(5, 0, -5), v(-5, 0, -5)) +
                   quad(v(-1, 2, -1), v(1, 2, -1), v(1, 2, 1), v(-1, 2, 1)), np.float32)
    mats = [material(color=(0.5,) * 3, bsdf_model=1), material(emission=emission)]
    s = SceneData(pos, None, np.array([0, 0, 1, 1], np.uint32), mats, camera((0, 1.0, 4.0), (0, 1.0, 0)), Sky())
    if tex is not None:
        mats[1].emission_texture = 0
        s.materials = mats
        s.uvs = np.tile([0.3, 0.3, 0.6, 0.3, 0.6, 0.6], (4, 1)).astype(np.float32)
        s.textures = [np.asarray(tex, np.uint32)]
    return s


def test_emission_texture_seen_and_sampled(oracle):
    c = 77
    ref = emitter_and_floor((4.0, 4.0, 4.0 * c / 255))
    tex = emitter_and_floor((4.0, 4.0, 4.0), [[rgba(255, 255, c)]])
    a, _ = oracle_render(oracle, ref, 48, 40, spp=16, max_bounces=3, rr_start=100)
    b, _ = oracle_render(oracle, tex, 48, 40, spp=16, max_bounces=3, rr_start=100)
    assert a.mean() > 0.05
    np.testing.assert_allclose(b, a, rtol=2e-5, atol=1e-7)


def test_restir_without_reuse_is_unbiased_on_the_textured_room(oracle, tmp_path):
    """The room's colour (sRGB, alpha as luminance), metallic-roughness,
    clearcoat, clearcoat-roughness, sheen, sheen-tint and emission textures
    (the emission texture also at NEE's light points).  The normal maps and
    the transmission texture are left out here, because with them the two
    integrators legitimately differ. (1) A fractional transmission (bilinear
    between 0 and 1 texels) makes the reference's Disney evaluate() (one
    branch past 0.5, disney_bsdf.wgsl:567-609) and sample() (the transmission
    lobe with probability t, :380-382) describe different BSDFs, which the
    two estimators weight differently. (2) Bent normals send continuation
    rays between the shading and the geometric horizon, where the plain
    integrator offsets the origin along the shading normal and the ReSTIR one
    along the ray (trace.wgsl:92-99)."""
    from appearance_amd.scene import load_gltf
    from tests.gltf_util import textured_room
    textured_room(tmp_path / "room.glb")
    s = load_gltf(tmp_path / "room.glb")
    for m in s.materials:
        m.normal_texture = m.clearcoat_normal_texture = m.transmission_texture = NO_TEX
    w, h, bounces = 24, 19, 3
    cam = camera((0.3, 1.4, 3.6), (0.0, 1.0, -2.0), 60.0)
    cfg = default_config(restir_di=0, restir_gi=0, firefly_filter=0)
    r = OracleRestir(oracle, s, w, h)
    frames = 256
    acc = np.zeros((h, w, 3))
    acc2 = np.zeros((h, w, 3))
    for f in range(frames):
        img = r.frame(cfg, camera=cam, frame_idx=f, max_bounces=bounces).astype(np.float64)
        acc += img
        acc2 += img * img
    m = acc / frames
    var = np.maximum(acc2 / frames - m * m, 0.0) / frames
    refs = np.array([oracle_render(oracle, s, w, h, spp=256, max_bounces=bounces, rr_start=3, jitter=False,
                                   frame_idx=1000 + b, camera=cam)[0] for b in range(8)], np.float64)
    rm, rv = refs.mean(0), refs.var(0, ddof=1) / 8
    sd = np.sqrt(var + rv + (1e-6 * np.abs(rm)) ** 2 + 1e-20)
    z = (m - rm) / sd
    rel = abs(m.mean() - rm.mean()) / rm.mean()
    zrms = float(np.sqrt((z ** 2).mean()))
    print(f"[textures] room ReSTIR reuse off: mean {m.mean():.5f} plain {rm.mean():.5f} rel {rel:.2e} "
          f"z rms {zrms:.2f}")
    assert rel < 3e-3
    assert zrms < 1.6
    assert np.sqrt((((m * 1.02 - rm) / sd) ** 2).mean()) > 1.6
