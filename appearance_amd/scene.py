"""Procedural scenes (libapt's apt_scene_generate) as numpy arrays."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._ffi import Camera, Material, MeshAttribs, MeshDesc, Sky, Texture, check, lib


@dataclass
class SceneData:
    positions: np.ndarray      # (n, 9) float32
    normals: np.ndarray | None # (n, 9) float32
    material_ids: np.ndarray   # (n,) uint32
    materials: list            # list of Material (ctypes, 160 B)
    camera: Camera
    sky: Sky
    uvs: np.ndarray | None = None         # (n, 6) float32
    tangents: np.ndarray | None = None    # (n, 12) float32: xyz, handedness per vertex
    bitangents: np.ndarray | None = None  # (n, 9) float32
    textures: list = field(default_factory=list)  # (height, width) uint32 RGBA8 arrays

    @property
    def textured(self) -> bool:
        return self.uvs is not None or self.tangents is not None or bool(self.textures)

    def attribs(self):
        """(MeshAttribs, keep-alive list) for apt_scene_set_attribs."""
        keep = []
        a = MeshAttribs()
        for name, arr in (("uvs", self.uvs), ("tangents", self.tangents), ("bitangents", self.bitangents)):
            if arr is not None:
                arr = np.ascontiguousarray(arr, np.float32)
                keep.append(arr)
                setattr(a, name, arr.ctypes.data_as(C.POINTER(C.c_float)))
        if self.textures:
            texs = [np.ascontiguousarray(t, np.uint32) for t in self.textures]
            keep += texs
            tv = (Texture * len(texs))(*[Texture(t.shape[1], t.shape[0], t.ctypes.data_as(C.POINTER(C.c_uint32)))
                                         for t in texs])
            keep.append(tv)
            a.textures = tv
            a.n_textures = len(texs)
        return a, keep

    @property
    def n_tris(self) -> int:
        return self.positions.shape[0]

    def materials_array(self) -> np.ndarray:
        """Materials as an (m, 40) float32 view of the 160 B descriptors."""
        buf = (Material * len(self.materials))(*self.materials)
        return np.frombuffer(bytes(buf), dtype=np.float32).reshape(len(self.materials), 40).copy()


def load_gltf(path) -> SceneData:
    """A glTF 2.0 model (.glb / .gltf) read by libapt (apt_scene_load_gltf)."""
    L = lib()
    L.apt_scene_load_gltf.restype = C.c_int
    L.apt_scene_load_gltf.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
    h = C.c_void_p()
    check(L.apt_scene_load_gltf(str(path).encode(), C.byref(h)), "apt_scene_load_gltf")
    return _view(h)


def decode_hdr(data: bytes) -> np.ndarray:
    """A Radiance HDR (RGBE) image decoded by libapt: (height, width, 3) float32."""
    w, h = C.c_uint32(), C.c_uint32()
    out = C.POINTER(C.c_float)()
    check(lib().apt_image_decode_hdr(data, len(data), C.byref(w), C.byref(h), C.byref(out)), "apt_image_decode_hdr")
    try:
        return np.ctypeslib.as_array(out, shape=(h.value * w.value * 3,)).reshape(h.value, w.value, 3).copy()
    finally:
        lib().apt_image_free(out)


def decode_image(data: bytes) -> np.ndarray:
    """A PNG or baseline JPEG texture image decoded by libapt: (height, width, 4) uint8 RGBA."""
    w, h = C.c_uint32(), C.c_uint32()
    check(lib().apt_image_decode_rgba8(data, len(data), C.byref(w), C.byref(h), None, 0), "apt_image_decode_rgba8")
    out = np.empty((h.value, w.value, 4), np.uint8)
    check(lib().apt_image_decode_rgba8(data, len(data), C.byref(w), C.byref(h), C.c_void_p(out.ctypes.data),
                                       out.nbytes), "apt_image_decode_rgba8")
    return out


def load_hdr(path) -> np.ndarray:
    """A Radiance .hdr file as a (height, width, 3) float32 sky panorama."""
    with open(path, "rb") as f:
        return decode_hdr(f.read())


def generate_scene(name: str, seed: int = 0) -> SceneData:
    h = C.c_void_p()
    check(lib().apt_scene_generate(name.encode(), seed, C.byref(h)), "apt_scene_generate")
    return _view(h)


def _view(h) -> SceneData:
    try:
        m, cam, sky = MeshDesc(), Camera(), Sky()
        check(lib().apt_scene_data_view(h, C.byref(m), C.byref(cam), C.byref(sky)),
              "apt_scene_data_view")
        n = m.n_tris
        if n == 0:
            pos, nrm, mids = np.zeros((0, 9), np.float32), None, np.zeros(0, np.uint32)
        else:
            pos = np.ctypeslib.as_array(m.positions, shape=(n * 9,)).reshape(n, 9).copy()
            nrm = None
            if m.normals:
                nrm = np.ctypeslib.as_array(m.normals, shape=(n * 9,)).reshape(n, 9).copy()
            mids = np.ctypeslib.as_array(m.material_ids, shape=(n,)).copy()
        mats = [Material.from_buffer_copy(m.materials[i]) for i in range(m.n_materials)]
        a = MeshAttribs()
        check(lib().apt_scene_data_attribs(h, C.byref(a)), "apt_scene_data_attribs")

        def arr(ptr, k):
            return np.ctypeslib.as_array(ptr, shape=(n * k,)).reshape(n, k).copy() if ptr and n else None

        texs = []
        for i in range(a.n_textures):
            t = a.textures[i]
            texs.append(np.ctypeslib.as_array(t.rgba8, shape=(t.height * t.width,)).reshape(t.height, t.width).copy())
        return SceneData(pos, nrm, mids, mats, cam, sky, arr(a.uvs, 6), arr(a.tangents, 12), arr(a.bitangents, 9),
                         texs)
    finally:
        lib().apt_scene_data_destroy(h)
