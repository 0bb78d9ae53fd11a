"""ctypes loader of the CPU oracle (oracle/liboracle.so).  Test-side only."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "oracle" / "liboracle.so"

_P = C.c_void_p
_SIGS = {
    "or_pcg_hash": (C.c_uint32, [C.c_uint32]),
    "or_xor_shift": (C.c_uint32, [C.c_uint32]),
    "or_hash_combine": (C.c_uint32, [C.c_uint32, C.c_uint32]),
    "or_wangh_hash": (C.c_uint32, [C.c_uint32]),
    "or_uniform_floats": (None, [C.c_uint32, C.c_uint32, _P]),
    "or_rgb9e5_pack": (None, [_P, C.c_uint32, _P]),
    "or_rgb9e5_unpack": (None, [_P, C.c_uint32, _P]),
    "or_oct30_pack": (None, [_P, C.c_uint32, _P]),
    "or_oct30_unpack": (None, [_P, C.c_uint32, _P]),
    "or_tri_test": (C.c_int, [_P, _P, _P]),
    "or_bvh_build": (_P, [_P, C.c_uint32]),
    "or_bvh_destroy": (None, [_P]),
    "or_bvh_node_count": (C.c_uint32, [_P]),
    "or_trace": (None, [_P, _P, C.c_uint32, C.c_int, _P, _P, C.c_int]),
    "or_brute": (None, [_P, C.c_uint32, _P, C.c_uint32, C.c_int, _P]),
    "or_inst_create": (_P, [_P, _P, C.c_uint32]),
    "or_inst_destroy": (None, [_P]),
    "or_inst_trace": (None, [_P, _P, C.c_uint32, _P, C.c_uint32, C.c_int, C.c_int, _P, C.c_int]),
    "or_affine_inverse": (C.c_int, [_P, _P]),
    "or_block_index": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32]),
    "or_aces": (C.c_float, [C.c_float]),
}


def _vp(a: np.ndarray) -> C.c_void_p:
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


class Oracle:
    def __init__(self, lib: C.CDLL):
        self.lib = lib
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name, None)
            if fn is not None:
                fn.restype = res
                fn.argtypes = args

    def uniform_floats(self, state: int, n: int) -> np.ndarray:
        out = np.empty(n, np.float32)
        self.lib.or_uniform_floats(state, n, _vp(out))
        return out

    def rgb9e5_pack(self, rgb: np.ndarray) -> np.ndarray:
        rgb = np.ascontiguousarray(rgb, np.float32).reshape(-1, 3)
        out = np.empty(rgb.shape[0], np.uint32)
        self.lib.or_rgb9e5_pack(_vp(rgb), rgb.shape[0], _vp(out))
        return out

    def rgb9e5_unpack(self, d: np.ndarray) -> np.ndarray:
        d = np.ascontiguousarray(d, np.uint32).reshape(-1)
        out = np.empty((d.shape[0], 3), np.float32)
        self.lib.or_rgb9e5_unpack(_vp(d), d.shape[0], _vp(out))
        return out

    def oct30_pack(self, v: np.ndarray) -> np.ndarray:
        v = np.ascontiguousarray(v, np.float32).reshape(-1, 3)
        out = np.empty(v.shape[0], np.uint32)
        self.lib.or_oct30_pack(_vp(v), v.shape[0], _vp(out))
        return out

    def oct30_unpack(self, p: np.ndarray) -> np.ndarray:
        p = np.ascontiguousarray(p, np.uint32).reshape(-1)
        out = np.empty((p.shape[0], 3), np.float32)
        self.lib.or_oct30_unpack(_vp(p), p.shape[0], _vp(out))
        return out

    def tri_test(self, ray: np.ndarray, tri: np.ndarray):
        ray = np.ascontiguousarray(ray, np.float32).reshape(8)
        tri = np.ascontiguousarray(tri, np.float32).reshape(9)
        out = np.zeros(3, np.float32)
        ok = self.lib.or_tri_test(_vp(ray), _vp(tri), _vp(out))
        return (bool(ok), out)

    def bvh(self, tris: np.ndarray) -> "OracleBvh":
        return OracleBvh(self, tris)

    def brute(self, tris: np.ndarray, rays: np.ndarray, any_hit: bool = False) -> np.ndarray:
        tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        rays = np.ascontiguousarray(rays).view(np.float32).reshape(-1, 8)
        n = rays.shape[0]
        out = np.empty(n if any_hit else (n, 4), np.float32)
        self.lib.or_brute(_vp(tris), tris.shape[0], _vp(rays), n, int(any_hit), _vp(out))
        return out


    def affine_inverse(self, o2w: np.ndarray):
        """World-to-object 3x4 of an object-to-world 3x4 (float32), or None if singular."""
        o2w = np.ascontiguousarray(o2w, np.float32).reshape(12)
        out = np.empty(12, np.float32)
        return out.reshape(3, 4) if self.lib.or_affine_inverse(_vp(o2w), _vp(out)) else None

    def instances(self, blas_tris) -> "OracleInstances":
        return OracleInstances(self, blas_tris)


class OracleInstances:
    """BLASes (a list of (n, 9) triangle arrays) traced through instances."""

    def __init__(self, o: Oracle, blas_tris):
        self.o = o
        mats = [np.ascontiguousarray(t, np.float32).reshape(-1, 9) for t in blas_tris]
        self.first = np.zeros(len(mats) + 1, np.uint32)
        self.first[1:] = np.cumsum([m.shape[0] for m in mats])
        self.tris = np.ascontiguousarray(np.concatenate(mats, 0))
        self.h = C.c_void_p(o.lib.or_inst_create(_vp(self.tris), _vp(self.first), len(mats)))

    def trace(self, instances: np.ndarray, rays: np.ndarray, any_hit: bool = False, brute: bool = False,
              threads: int = 0) -> np.ndarray:
        inst = np.ascontiguousarray(instances)
        assert inst.dtype.itemsize == 64
        rays = np.ascontiguousarray(rays).view(np.float32).reshape(-1, 8)
        n = rays.shape[0]
        out = np.empty(n if any_hit else (n, 4), np.float32)
        self.o.lib.or_inst_trace(self.h, _vp(inst), inst.shape[0], _vp(rays), n, int(any_hit), int(brute),
                                 _vp(out), threads)
        return out

    def __del__(self):
        try:
            self.o.lib.or_inst_destroy(self.h)
        except Exception:
            pass


class OracleBvh:
    def __init__(self, o: Oracle, tris: np.ndarray):
        self.o = o
        self.tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        self.h = C.c_void_p(o.lib.or_bvh_build(_vp(self.tris), self.tris.shape[0]))
        self.last_counts = (0, 0)

    def trace(self, rays: np.ndarray, any_hit: bool = False, threads: int = 0) -> np.ndarray:
        rays = np.ascontiguousarray(rays).view(np.float32).reshape(-1, 8)
        n = rays.shape[0]
        out = np.empty(n if any_hit else (n, 4), np.float32)
        counts = np.zeros(2, np.uint64)
        self.o.lib.or_trace(self.h, _vp(rays), n, int(any_hit), _vp(out), _vp(counts), threads)
        self.last_counts = (int(counts[0]), int(counts[1]))
        return out

    def __del__(self):
        try:
            self.o.lib.or_bvh_destroy(self.h)
        except Exception:
            pass


def hits_as_struct(h: np.ndarray):
    """(n,4) float32 oracle hits -> t, u, v (float32) and prim (uint32)."""
    h = np.ascontiguousarray(h, np.float32)
    return h[:, 0], h[:, 1], h[:, 2], h[:, 3].view(np.uint32)


def load_oracle() -> Oracle:
    if not LIB.exists():
        import subprocess
        import sys
        subprocess.run([sys.executable, str(ROOT / "oracle" / "build.py")], check=True)
    return Oracle(C.CDLL(str(LIB)))


class _Attribs(C.Structure):
    """ocr::Attribs (oracle/oracle_pt.h)."""
    _fields_ = [("uv", C.c_void_p), ("tan", C.c_void_p), ("bit", C.c_void_p), ("tex", C.c_void_p),
                ("n_tex", C.c_uint32), ("pad", C.c_uint32), ("sky_rgb", C.c_void_p), ("sky_w", C.c_uint32),
                ("sky_h", C.c_uint32)]


def _unit_rows(v: np.ndarray) -> np.ndarray:
    v = np.asarray(v, np.float32)
    ln = np.sqrt((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2])
    out = np.zeros_like(v)
    nz = ln > 0
    out[nz] = v[nz] / ln[nz, None]
    return out


def vertex_frames(scene):
    """The per-vertex tangent frame a textured scene's triangles carry on the
    device: unit tangent (glTF TANGENT xyz) and unit bitangent (given, else
    cross(n, t) * w with the vertex normal, or the winding normal without
    normals; vertex_pool_bindings.wgsl:43-46); (n, 9) float32 each."""
    if scene.tangents is None:
        return None, None
    q = np.asarray(scene.tangents, np.float32).reshape(-1, 4)
    if scene.normals is not None:
        n3 = np.asarray(scene.normals, np.float32).reshape(-1, 3)
    else:
        p = np.asarray(scene.positions, np.float32).reshape(-1, 3, 3)
        e1, e2 = p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]
        fn = np.stack([e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1], e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2],
                       e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]], 1)
        n3 = np.repeat(fn, 3, axis=0)
    if scene.bitangents is not None:
        b = np.asarray(scene.bitangents, np.float32).reshape(-1, 3)
    else:
        b = np.stack([(n3[:, 1] * q[:, 2] - n3[:, 2] * q[:, 1]) * q[:, 3],
                      (n3[:, 2] * q[:, 0] - n3[:, 0] * q[:, 2]) * q[:, 3],
                      (n3[:, 0] * q[:, 1] - n3[:, 1] * q[:, 0]) * q[:, 3]], 1).astype(np.float32)
    return (np.ascontiguousarray(_unit_rows(q[:, :3]).reshape(-1, 9)),
            np.ascontiguousarray(_unit_rows(b).reshape(-1, 9)))


def oracle_attribs(scene, tan=None, bit=None, sky_pano=None):
    """(ocr::Attribs pointer or None, keep-alive) for a SceneData: uvs, the
    per-vertex frames (vertex_frames, or tan/bit given in world space), the
    textures, and a sky panorama ((h, w, 3) float32) if given."""
    if not getattr(scene, "textured", False) and sky_pano is None:
        return None, []
    from appearance_amd._ffi import Texture
    keep = []
    a = _Attribs()
    if sky_pano is not None:
        pano = np.ascontiguousarray(sky_pano, np.float32)
        keep.append(pano)
        a.sky_rgb, a.sky_w, a.sky_h = pano.ctypes.data, pano.shape[1], pano.shape[0]
    if not getattr(scene, "textured", False):
        keep.append(a)
        return C.cast(C.pointer(a), C.c_void_p), keep
    if scene.uvs is not None:
        uv = np.ascontiguousarray(scene.uvs, np.float32)
        keep.append(uv)
        a.uv = uv.ctypes.data
    if tan is None and bit is None:
        tan, bit = vertex_frames(scene)
    if tan is not None:
        tan = np.ascontiguousarray(tan, np.float32)
        bit = np.ascontiguousarray(bit, np.float32)
        keep += [tan, bit]
        a.tan, a.bit = tan.ctypes.data, bit.ctypes.data
    if scene.textures:
        texs = [np.ascontiguousarray(t, np.uint32) for t in scene.textures]
        tv = (Texture * len(texs))(*[Texture(t.shape[1], t.shape[0], t.ctypes.data_as(C.POINTER(C.c_uint32)))
                                     for t in texs])
        keep += texs + [tv]
        a.tex = C.cast(tv, C.c_void_p)
        a.n_tex = len(texs)
    keep.append(a)
    return C.cast(C.pointer(a), C.c_void_p), keep


def sky_probe(o: Oracle, sky, d, pano=None) -> np.ndarray:
    """The oracle's sky radiance (no sun) along direction d, with a panorama if given."""
    at, keep = oracle_attribs(None, sky_pano=pano) if pano is not None else (None, [])
    out = np.empty(3, np.float32)
    dd = np.ascontiguousarray(d, np.float32)
    fn = o.lib.or_sky_probe
    fn.argtypes = [C.c_void_p, C.c_void_p, _P, _P]
    fn(C.cast(C.pointer(sky), C.c_void_p), at, _vp(dd), _vp(out))
    del keep
    return out


def tex_sample(o: Oracle, tex: np.ndarray, u: float, v: float) -> np.ndarray:
    """The oracle's bilinear repeat sample of an (h, w) uint32 RGBA8 texture."""
    from appearance_amd._ffi import Texture
    t = np.ascontiguousarray(tex, np.uint32)
    d = Texture(t.shape[1], t.shape[0], t.ctypes.data_as(C.POINTER(C.c_uint32)))
    out = np.empty(4, np.float32)
    fn = o.lib.or_tex_sample
    fn.argtypes = [C.c_void_p, C.c_float, C.c_float, _P]
    fn(C.cast(C.pointer(d), C.c_void_p), u, v, _vp(out))
    return out


SHADING_FIELDS = ["uv", "nmap", "ncc", "color", "luminance", "metallic", "roughness", "emission", "transmission",
                  "sheen", "sheen_tint", "clearcoat", "clearcoat_roughness"]
_SHADING_SIZES = [2, 3, 3, 3, 1, 1, 1, 3, 1, 1, 3, 1, 1]


def shading_probe(o: Oracle, scene, mat, prim: int, bu: float, bv: float, n, tan=None, bit=None) -> dict:
    """The oracle's shading inputs at a hit: uv, mapped and clearcoat normals, evaluated material."""
    at, keep = oracle_attribs(scene, tan, bit)
    out = np.empty(26, np.float32)
    nn = np.ascontiguousarray(n, np.float32)
    fn = o.lib.or_shading_probe
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_float, C.c_float, _P, _P]
    fn(C.cast(C.pointer(mat), C.c_void_p), at, prim, bu, bv, _vp(nn), _vp(out))
    res, k = {}, 0
    for name, sz in zip(SHADING_FIELDS, _SHADING_SIZES):
        res[name] = out[k] if sz == 1 else out[k:k + sz].copy()
        k += sz
    return res


def oracle_render(o: Oracle, scene, width: int, height: int, spp: int, max_bounces: int,
                  rr_start: int, seed: int = 0, frame_idx: int = 0, row_start: int = 0,
                  rows: int | None = None, pixels=None, precision: int = 32, threads: int = 0,
                  bvh=None, camera=None, sky=None, jitter: bool = True, tally: bool = False, frames=None,
                  sky_pano=None):
    """Run the oracle path tracer on a SceneData; returns (mean radiance, counts):
    counts = (extension rays, shadow rays), or with tally=True a dict that
    adds the box and triangle tests of both kinds of query (FP32).  A textured
    scene's uvs, tangent frames and textures go along; frames = (tan, bit)
    gives the world-space per-vertex frames instead of vertex_frames(scene);
    sky_pano an (h, w, 3) sky panorama."""
    import ctypes as C
    from appearance_amd._ffi import Material
    rows = height - row_start if rows is None else rows
    pos = np.ascontiguousarray(scene.positions, np.float32)
    nrm = None if scene.normals is None else np.ascontiguousarray(scene.normals, np.float32)
    mids = np.ascontiguousarray(scene.material_ids, np.uint32)
    mats = (Material * len(scene.materials))(*scene.materials)
    b = bvh if bvh is not None else o.bvh(pos)
    cam = camera if camera is not None else scene.camera
    sk = sky if sky is not None else scene.sky
    cfg = np.array([width, height, row_start, rows, spp, max_bounces, rr_start, seed, frame_idx,
                    1 if jitter else 0], np.uint32)
    if pixels is not None:
        pixels = np.ascontiguousarray(pixels, np.uint32)
        out = np.empty((pixels.shape[0], 3), np.float32)
    else:
        out = np.empty((rows, width, 3), np.float32)
    counts = np.zeros(6, np.uint64)
    at, keep = oracle_attribs(scene, *(frames or (None, None)), sky_pano=sky_pano)
    fn = o.lib.or_render_counted
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_int,
                   C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]
    rc = fn(_vp(pos), _vp(nrm) if nrm is not None else None, _vp(mids), C.cast(mats, C.c_void_p),
            pos.shape[0], b.h, C.cast(C.pointer(sk), C.c_void_p), C.cast(C.pointer(cam), C.c_void_p),
            _vp(cfg), _vp(pixels) if pixels is not None else None,
            0 if pixels is None else pixels.shape[0], precision, threads, _vp(out), _vp(counts), 6, at)
    del keep
    assert rc == 0
    if tally:
        keys = ["extension_rays", "shadow_rays", "closest_box_tests", "closest_tri_tests", "any_box_tests",
                "any_tri_tests"]
        return out, {k: int(v) for k, v in zip(keys, counts)}
    return out, (int(counts[0]), int(counts[1]))


class OracleRestir:
    """The oracle's ReSTIR integrator with its own per-pixel state (reservoirs,
    G-buffer, denoiser history) carried from frame to frame."""

    FIELDS = ["di_candidates", "di_mcap", "di_spatial_neighbours", "gi_mcap", "restir_di", "restir_gi",
              "firefly_filter", "svgf", "taa", "svgf_history", "svgf_iterations"]

    def __init__(self, o: Oracle, scene, width: int, height: int, row_start: int = 0,
                 rows: int | None = None, bvh=None, frames=None, sky_pano=None):
        self.o = o
        self.scene = scene
        self.attribs, self._keep = oracle_attribs(scene, *(frames or (None, None)), sky_pano=sky_pano)
        self.width, self.height, self.row_start = width, height, row_start
        self.rows = height - row_start if rows is None else rows
        self.pos = np.ascontiguousarray(scene.positions, np.float32)
        self.nrm = None if scene.normals is None else np.ascontiguousarray(scene.normals, np.float32)
        self.mids = np.ascontiguousarray(scene.material_ids, np.uint32)
        from appearance_amd._ffi import Material
        self.mats = (Material * len(scene.materials))(*scene.materials)
        self.bvh = bvh if bvh is not None else o.bvh(self.pos)
        fn = o.lib.or_restir_create
        fn.restype = C.c_void_p
        fn.argtypes = [C.c_uint32, C.c_uint32]
        self.h = C.c_void_p(fn(width, self.rows))
        self.rays = (0, 0)

    def frame(self, cfg, camera=None, sky=None, spp: int = 1, max_bounces: int = 2, seed: int = 0,
              frame_idx: int = 0, threads: int = 0) -> np.ndarray:
        """cfg: an appearance_amd Config (or any object with the ReSTIR fields)."""
        cam = camera if camera is not None else self.scene.camera
        sk = sky if sky is not None else self.scene.sky
        radius = np.array([cfg.di_spatial_radius], np.float32).view(np.uint32)[0]
        c = np.array([self.width, self.height, self.row_start, self.rows, spp, max_bounces, seed, frame_idx,
                      cfg.di_candidates, cfg.di_mcap, cfg.di_spatial_neighbours, cfg.gi_mcap,
                      cfg.restir_di, cfg.restir_gi, cfg.firefly_filter, cfg.svgf, cfg.taa,
                      cfg.svgf_history, cfg.svgf_iterations, radius, getattr(cfg, "gi_spatial_passes", 0)],
                     np.uint32)
        out = np.empty((self.rows, self.width, 3), np.float32)
        rays = np.zeros(2, np.uint64)
        fn = self.o.lib.or_restir_frame
        fn.restype = C.c_int
        fn.argtypes = [C.c_void_p, _P, _P, _P, _P, C.c_uint32, _P, _P, _P, _P, C.c_int, _P, _P, _P]
        r = fn(self.h, _vp(self.pos), _vp(self.nrm) if self.nrm is not None else None, _vp(self.mids),
               C.cast(self.mats, C.c_void_p), self.pos.shape[0], self.bvh.h, C.cast(C.pointer(sk), C.c_void_p),
               C.cast(C.pointer(cam), C.c_void_p), _vp(c), threads, _vp(out), _vp(rays), self.attribs)
        assert r == 0
        self.rays = (int(rays[0]), int(rays[1]))
        return out

    def __del__(self):
        try:
            fn = self.o.lib.or_restir_destroy
            fn.argtypes = [C.c_void_p]
            fn(self.h)
        except Exception:
            pass


def f16_round(o: Oracle, x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32).reshape(-1)
    out = np.empty_like(x)
    fn = o.lib.or_f16_round
    fn.argtypes = [_P, C.c_uint32, _P]
    fn(_vp(x), x.shape[0], _vp(out))
    return out


def atrous_probe(o: Oracle, col, var, depth, nrm, pass_idx: int) -> np.ndarray:
    h, w = depth.shape
    col = np.ascontiguousarray(col, np.float32)
    var = np.ascontiguousarray(var, np.float32)
    depth = np.ascontiguousarray(depth, np.float32)
    nrm = np.ascontiguousarray(nrm, np.float32)
    out = np.empty((h, w, 3), np.float32)
    fn = o.lib.or_atrous_probe
    fn.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, _P, _P, _P, _P, _P]
    fn(w, h, pass_idx, _vp(col), _vp(var), _vp(depth), _vp(nrm), _vp(out))
    return out


def post_probe(o: Oracle, kind: int, imgs, sky=None, svgf_history: int = 4):
    """kind 0 firefly (frame 0), 1 SVGF temporal, 2 TAA over the frames of imgs
    (frames, h, w, 3) on a static camera facing a flat wall; returns (colour,
    variance) of the last frame."""
    imgs = np.ascontiguousarray(imgs, np.float32)
    f, h, w, _ = imgs.shape
    skyb = None if sky is None else np.ascontiguousarray(sky, np.uint8)
    out = np.empty((h, w, 3), np.float32)
    var = np.empty((h, w), np.float32)
    fn = o.lib.or_post_probe
    fn.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _P, _P, C.c_uint32, _P, _P]
    fn.restype = C.c_int
    rc = fn(kind, w, h, f, _vp(imgs), _vp(skyb) if skyb is not None else None, svgf_history, _vp(out), _vp(var))
    assert rc == 0
    return out, var


def jpeg_block(o: Oracle, rgbx, quality: int) -> bytes:
    """The oracle's JFIF file of one (64, 64, 4) RGBX block."""
    a = np.ascontiguousarray(rgbx, np.uint8)
    assert a.shape == (64, 64, 4)
    out = np.empty(1 << 17, np.uint8)
    fn = o.lib.or_jpeg_block
    fn.argtypes = [_P, C.c_int, _P, C.c_size_t]
    fn.restype = C.c_size_t
    n = fn(_vp(a), int(quality), _vp(out), out.size)
    assert n > 0
    return out[:n].tobytes()
