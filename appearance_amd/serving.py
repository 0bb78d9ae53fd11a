"""Python view of libapt's serving loop (apt_server_* in include/apt.h).

A Server keeps scenes resident on the GPU and renders submitted requests with
continuous batching; results are FP32 linear radiance means (h, w, 3).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._ffi import Camera, Material, MeshDesc, Sky, check, lib


class ServerConfig(C.Structure):
    _fields_ = [("path_slots", C.c_uint32), ("max_active", C.c_uint32),
                ("accum_budget", C.c_uint64), ("max_bounces", C.c_uint32),
                ("tile_paths", C.c_uint32), ("bvh_width", C.c_uint32),
                ("lds_stack_depth", C.c_uint32), ("policy", C.c_uint32), ("persistent", C.c_uint32)]


class Request(C.Structure):
    _fields_ = [("scene_id", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("spp", C.c_uint32), ("max_bounces", C.c_uint32), ("rr_start_bounce", C.c_uint32),
                ("seed", C.c_uint32), ("_pad", C.c_uint32), ("camera", Camera)]


class Result(C.Structure):
    _fields_ = [("state", C.c_uint32), ("status", C.c_uint32), ("submit_ms", C.c_double),
                ("admit_ms", C.c_double), ("done_ms", C.c_double), ("paths", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [("batches", C.c_uint64), ("paths", C.c_uint64), ("requests_done", C.c_uint64),
                ("active", C.c_uint32), ("queued", C.c_uint32), ("gpu_busy_ms", C.c_double)]


QUEUED, ACTIVE, DONE, CANCELLED, FAILED = range(5)

_P = C.c_void_p
_SIGS = {
    "apt_server_create": (C.c_int, [C.POINTER(ServerConfig), C.POINTER(_P)]),
    "apt_server_destroy": (None, [_P]),
    "apt_server_add_scene": (C.c_int, [_P, C.POINTER(MeshDesc), C.POINTER(Sky),
                                       C.POINTER(C.c_uint32)]),
    "apt_submit": (C.c_int, [_P, C.POINTER(Request), _P, C.POINTER(C.c_uint64)]),
    "apt_poll": (C.c_int, [_P, C.c_uint64, C.POINTER(Result)]),
    "apt_wait": (C.c_int, [_P, C.c_uint64, C.c_uint32, C.POINTER(Result)]),
    "apt_cancel": (C.c_int, [_P, C.c_uint64]),
    "apt_release": (C.c_int, [_P, C.c_uint64]),
    "apt_server_get_stats": (C.c_int, [_P, C.POINTER(ServerStats)]),
}


def _lib():
    L = lib()
    if not getattr(L, "_srv_sigs", False):
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        L._srv_sigs = True
    return L


class Server:
    def __init__(self, path_slots: int = 0, max_active: int = 0, accum_budget: int = 0,
                 max_bounces: int = 0, tile_paths: int = 0, bvh_width: int = 0,
                 lds_stack_depth: int = 0, policy: int = 0, persistent: int = 0):
        cfg = ServerConfig(path_slots, max_active, accum_budget, max_bounces, tile_paths, bvh_width,
                           lds_stack_depth, policy, persistent)
        h = C.c_void_p()
        check(_lib().apt_server_create(C.byref(cfg), C.byref(h)), "apt_server_create")
        self._h = h
        self._out = {}

    def add_scene(self, scene) -> int:
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
        sid = C.c_uint32()
        check(_lib().apt_server_add_scene(self._h, C.byref(md), C.byref(scene.sky), C.byref(sid)),
              "apt_server_add_scene")
        return sid.value

    def submit(self, scene_id: int, width: int, height: int, spp: int, max_bounces: int,
               camera: Camera, rr_start_bounce: int = 3, seed: int = 0) -> int:
        out = np.zeros((height, width, 3), np.float32)
        r = Request(scene_id, width, height, spp, max_bounces, rr_start_bounce, seed, 0, camera)
        rid = C.c_uint64()
        check(_lib().apt_submit(self._h, C.byref(r), C.c_void_p(out.ctypes.data), C.byref(rid)),
              "apt_submit")
        self._out[rid.value] = out
        return rid.value

    def poll(self, rid: int) -> Result:
        res = Result()
        check(_lib().apt_poll(self._h, rid, C.byref(res)), "apt_poll")
        return res

    def wait(self, rid: int, timeout_ms: int = 600000) -> Result:
        res = Result()
        check(_lib().apt_wait(self._h, rid, timeout_ms, C.byref(res)), "apt_wait")
        return res

    def result(self, rid: int) -> np.ndarray:
        return self._out[rid]

    def cancel(self, rid: int) -> None:
        check(_lib().apt_cancel(self._h, rid), "apt_cancel")

    def release(self, rid: int) -> np.ndarray:
        """Forget a finished request (apt_release); returns its result."""
        check(_lib().apt_release(self._h, rid), "apt_release")
        return self._out.pop(rid)

    def stats(self) -> dict:
        s = ServerStats()
        check(_lib().apt_server_get_stats(self._h, C.byref(s)), "apt_server_get_stats")
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib().apt_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- the serving workload ---------------------------------------------------

RESOLUTIONS = [(256, 256), (512, 512), (1280, 720), (1920, 1080)]
SPPS = [1, 4, 16, 64]


def mixed_requests(n: int, rate: float, seed: int = 5) -> list[dict]:
    """The serving configuration's request stream: resolution drawn from
    {256^2, 512^2, 1280x720, 1920x1080}, spp from {1, 4, 16, 64}, depth 4..8,
    one of four resident scenes, a seeded orbit angle for the camera and
    Poisson arrivals at `rate` requests/s (t in seconds)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        w, h = RESOLUTIONS[rng.integers(4)]
        out.append({"t": t, "scene": int(rng.integers(4)), "w": w, "h": h,
                    "spp": SPPS[rng.integers(4)], "depth": int(rng.integers(4, 9)),
                    "angle": float(rng.uniform(0, 2 * np.pi)), "seed": i})
    return out


def orbit_camera(scene, angle: float) -> Camera:
    """The scene's camera turned by angle/4 - 1/8 radians about a point 5
    units ahead of it (a seeded pose per request)."""
    from .world import look_at

    m = np.array(scene.camera.transform[:], np.float32).reshape(4, 4)  # rows are columns
    eye = m[3, :3].astype(np.float64)
    back = m[2, :3].astype(np.float64)
    target = eye - back * 5.0
    off = eye - target
    c, s_ = np.cos(angle * 0.25 - 0.125), np.sin(angle * 0.25 - 0.125)
    rot = np.array([off[0] * c + off[2] * s_, off[1], -off[0] * s_ + off[2] * c])
    cam = Camera()
    cam.transform[:] = look_at(tuple(target + rot), tuple(target)).T.ravel().tolist()
    cam.fov_deg = scene.camera.fov_deg
    cam.near_plane = 0.01
    cam.far_plane = 1000.0
    return cam
