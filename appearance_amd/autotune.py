"""Per-scene autotuner (libapt's apt_autotune): BVH width x LDS stack depth x
wavefront path budget, timed on the device, written to a per-scene JSON."""
from __future__ import annotations

import ctypes as C
import json
from pathlib import Path

import numpy as np

from ._ffi import Camera, Material, MeshDesc, Sky, check, lib


class TuneOpts(C.Structure):
    _fields_ = [("probe_width", C.c_uint32), ("probe_height", C.c_uint32),
                ("probe_spp", C.c_uint32), ("probe_bounces", C.c_uint32), ("repeats", C.c_uint32)]


class TuneResult(C.Structure):
    _fields_ = [("bvh_width", C.c_uint32), ("lds_stack_depth", C.c_uint32),
                ("queue_chunk", C.c_uint32), ("trials", C.c_uint32), ("best_ms", C.c_double)]


def tune(scene, probe=(512, 256), spp: int = 2, bounces: int = 4, repeats: int = 2,
         json_path: str | Path | None = None) -> dict:
    L = lib()
    L.apt_autotune.restype = C.c_int
    L.apt_autotune.argtypes = [C.POINTER(MeshDesc), C.POINTER(Sky), C.POINTER(Camera),
                               C.POINTER(TuneOpts), C.POINTER(TuneResult), C.c_char_p, C.c_size_t]
    pos = np.ascontiguousarray(scene.positions, np.float32)
    nrm = None if scene.normals is None else np.ascontiguousarray(scene.normals, np.float32)
    mids = np.ascontiguousarray(scene.material_ids, np.uint32)
    mats = (Material * len(scene.materials))(*scene.materials)
    md = MeshDesc()
    md.positions = pos.ctypes.data_as(C.POINTER(C.c_float))
    md.normals = nrm.ctypes.data_as(C.POINTER(C.c_float)) if nrm is not None else None
    md.material_ids = mids.ctypes.data_as(C.POINTER(C.c_uint32))
    md.n_tris = pos.shape[0]
    md.n_materials = len(scene.materials)
    md.materials = mats
    opts = TuneOpts(probe[0], probe[1], spp, bounces, repeats)
    res = TuneResult()
    buf = C.create_string_buffer(1 << 16)
    check(L.apt_autotune(C.byref(md), C.byref(scene.sky), C.byref(scene.camera), C.byref(opts),
                         C.byref(res), buf, len(buf)), "apt_autotune")
    report = json.loads(buf.value.decode())
    if json_path is not None:
        Path(json_path).write_text(json.dumps(report, indent=1))
    return report
