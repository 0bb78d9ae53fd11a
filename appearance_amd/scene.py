"""Procedural scenes (libapt's apt_scene_generate) as numpy arrays."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._ffi import Camera, Material, MeshDesc, Sky, check, lib


@dataclass
class SceneData:
    positions: np.ndarray      # (n, 9) float32
    normals: np.ndarray | None # (n, 9) float32
    material_ids: np.ndarray   # (n,) uint32
    materials: list            # list of Material (ctypes, 160 B)
    camera: Camera
    sky: Sky

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
        return SceneData(pos, nrm, mids, mats, cam, sky)
    finally:
        lib().apt_scene_data_destroy(h)
