"""Python view of libapt's renderer, mirroring the reference's interface.

  PathTracerGpu        crates/appearance-path-tracer-gpu/src/lib.rs:250-615
                       (new, handle_visible_world_action, render)
  NodeRenderer         crates/appearance-render-loop/src/node.rs:14-25
  DistributedRenderer  crates/appearance-distributed-renderer/src/lib.rs:11-82

render() hands the RGBA8 64x64-block-layout pixels to result_callback, as the
reference does, and returns nothing; render_accum() additionally returns the
FP32 linear radiance mean used for parity checks.  Every call goes through the
C ABI; there is no Python or CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import world
from ._ffi import BLOCK_CB, RESULT_CB, Config, MeshDesc, RenderStats, Sky, check, lib, vptr


def default_config(**overrides) -> Config:
    c = Config()
    lib().apt_config_default(C.byref(c))
    for k, v in overrides.items():
        if not hasattr(c, k):
            raise AttributeError(f"unknown config field {k}")
        setattr(c, k, v)
    return c


INTEGRATOR_PLAIN = 0
INTEGRATOR_RESTIR = 1


def plain_config(**overrides) -> Config:
    """default_config with the wavefront path tracer (APT_INTEGRATOR_PLAIN)
    and the ReSTIR switches and post chain off."""
    kw = dict(integrator=INTEGRATOR_PLAIN, restir_di=0, restir_gi=0, firefly_filter=0, svgf=0, taa=0)
    kw.update(overrides)
    return default_config(**kw)


class PathTracerGpu:
    def __init__(self, config: Config | None = None, **overrides):
        cfg = config if config is not None else default_config(**overrides)
        h = C.c_void_p()
        check(lib().apt_create(C.byref(cfg), C.byref(h)), "apt_create")
        self._h = h
        self.config = cfg
        self._keep = []

    def set_config(self, **overrides) -> None:
        for k, v in overrides.items():
            setattr(self.config, k, v)
        check(lib().apt_set_config(self._h, C.byref(self.config)), "apt_set_config")

    def handle_visible_world_action(self, action) -> None:
        data = action.to_bytes()
        buf = C.create_string_buffer(data, len(data))
        check(lib().apt_world_action(self._h, action.ty, buf, len(data)), "apt_world_action")

    def set_scene(self, scene, sky=None) -> None:
        """Replace the scene by a SceneData (appearance_amd.scene) or arrays."""
        pos = np.ascontiguousarray(scene.positions, np.float32)
        nrm = None if scene.normals is None else np.ascontiguousarray(scene.normals, np.float32)
        mids = np.ascontiguousarray(scene.material_ids, np.uint32)
        from ._ffi import Material
        mats = (Material * len(scene.materials))(*scene.materials)
        md = MeshDesc()
        md.positions = pos.ctypes.data_as(C.POINTER(C.c_float))
        md.normals = nrm.ctypes.data_as(C.POINTER(C.c_float)) if nrm is not None else None
        md.material_ids = mids.ctypes.data_as(C.POINTER(C.c_uint32))
        md.n_tris = pos.shape[0]
        md.n_materials = len(scene.materials)
        md.materials = mats
        sk = sky if sky is not None else getattr(scene, "sky", None)
        if getattr(scene, "textured", False):
            at, _keep = scene.attribs()
            check(lib().apt_scene_set_attribs(self._h, C.byref(md), C.byref(at),
                                              C.byref(sk) if sk is not None else None), "apt_scene_set_attribs")
            return
        check(lib().apt_scene_set(self._h, C.byref(md), C.byref(sk) if sk is not None else None),
              "apt_scene_set")

    def set_sky_panorama(self, rgb) -> None:
        """The equirectangular sky panorama ((height, width, 3) linear RGB,
        row 0 at the zenith; None: the analytic gradient)."""
        if rgb is None:
            check(lib().apt_set_sky_panorama(self._h, 0, 0, None), "apt_set_sky_panorama")
            return
        a = np.ascontiguousarray(rgb, np.float32)
        assert a.ndim == 3 and a.shape[2] == 3
        check(lib().apt_set_sky_panorama(self._h, a.shape[1], a.shape[0], a.ctypes.data_as(C.POINTER(C.c_float))),
              "apt_set_sky_panorama")

    def set_camera(self, camera) -> None:
        """Apply a CameraUpdate (world.CameraUpdate or the ABI Camera struct)."""
        if isinstance(camera, world.CameraUpdate):
            self.handle_visible_world_action(camera)
            return
        data = bytes(camera)
        buf = C.create_string_buffer(data, len(data))
        check(lib().apt_world_action(self._h, world.CAMERA_UPDATE, buf, len(data)), "apt_world_action")

    def render(self, resolution, start_row: int, end_row: int, result_callback) -> None:
        w, h = int(resolution[0]), int(resolution[1])

        def _cb(ptr, n, _user):
            result_callback(C.string_at(ptr, n))

        cb = RESULT_CB(_cb)
        check(lib().apt_render_cb(self._h, w, h, start_row, end_row, cb, None), "apt_render_cb")

    def render_accum(self, resolution, start_row: int = 0, end_row: int | None = None,
                     with_rgba: bool = True):
        """Render; return (rgba8 block layout bytes or None, FP32 mean radiance (rows, w, 3))."""
        w, h = int(resolution[0]), int(resolution[1])
        end_row = h if end_row is None else end_row
        rows = end_row - start_row
        rgba = np.empty(w * rows * 4, np.uint8) if with_rgba else None
        acc = np.empty((rows, w, 3), np.float32)
        check(lib().apt_render(self._h, w, h, start_row, end_row,
                               vptr(rgba) if with_rgba else None, rgba.size if with_rgba else 0,
                               vptr(acc), acc.size), "apt_render")
        return rgba, acc

    def render_rgba(self, resolution, start_row: int = 0, end_row: int | None = None) -> np.ndarray:
        """Render; return only the RGBA8 block-layout bytes (the render-node output)."""
        w, h = int(resolution[0]), int(resolution[1])
        end_row = h if end_row is None else end_row
        rgba = np.empty(w * (end_row - start_row) * 4, np.uint8)
        check(lib().apt_render(self._h, w, h, start_row, end_row, vptr(rgba), rgba.size, None, 0),
              "apt_render")
        return rgba

    def render_device(self, resolution, start_row: int = 0, end_row: int | None = None) -> None:
        """Render into the renderer's device buffers only (both outputs NULL):
        the frame is resolved in HBM and nothing crosses PCIe."""
        w, h = int(resolution[0]), int(resolution[1])
        end_row = h if end_row is None else end_row
        check(lib().apt_render(self._h, w, h, start_row, end_row, None, 0, None, 0), "apt_render")

    def start_render(self, data, block_callback, jpeg_quality: int | None = None) -> None:
        """Node::start_render (node.rs:44-105) without the transport: data is a
        world.StartRenderData or its 20 bytes; block_callback(row, column_block,
        frame_idx, pixels) receives each 64x64 RGBA8 block of the slab, or with
        jpeg_quality (the node uses 95) the block's JPEG file, compressed on
        the device (RenderPartialFinishedData::compressed_pixel_bytes)."""
        raw = data.to_bytes() if hasattr(data, "to_bytes") else bytes(data)

        def _cb(row, col, frame, ptr, n, _user):
            block_callback(row, col, frame, C.string_at(ptr, n))

        cb = BLOCK_CB(_cb)
        buf = C.create_string_buffer(raw, len(raw))
        if jpeg_quality is None:
            check(lib().apt_start_render(self._h, buf, len(raw), cb, None), "apt_start_render")
        else:
            check(lib().apt_start_render_jpeg(self._h, buf, len(raw), int(jpeg_quality), cb, None),
                  "apt_start_render_jpeg")

    def set_frame_index(self, frame_idx: int) -> None:
        """Restart the frame sequence at frame_idx (seeds frame * spp + sample)."""
        check(lib().apt_set_frame_index(self._h, frame_idx), "apt_set_frame_index")

    def stats(self) -> dict:
        s = RenderStats()
        check(lib().apt_get_render_stats(self._h, C.byref(s)), "apt_get_render_stats")
        return {f: getattr(s, f) for f, _ in RenderStats._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().apt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class NodeRenderer:
    """The trait a render node drives (node.rs:14-25)."""

    def visible_world_action(self, action) -> None:
        raise NotImplementedError

    def render(self, resolution, start_row: int, end_row: int, result_callback) -> None:
        raise NotImplementedError


class DistributedRenderer(NodeRenderer):
    """NodeRenderer backed by libapt (distributed-renderer/src/lib.rs:60-82)."""

    def __init__(self, config: Config | None = None, **overrides):
        self.path_tracer = PathTracerGpu(config, **overrides)

    def visible_world_action(self, action) -> None:
        self.path_tracer.handle_visible_world_action(action)

    def render(self, resolution, start_row: int, end_row: int, result_callback) -> None:
        self.path_tracer.render(resolution, start_row, end_row, result_callback)


def block_to_linear(pixels: bytes | np.ndarray, width: int, rows: int) -> np.ndarray:
    """Undo the 64x64 block layout (crates/appearance-render-loop/assets/shaders/block.wgsl)."""
    px = np.frombuffer(bytes(pixels), np.uint8).reshape(-1, 4)
    B = 64
    y, x = np.mgrid[0:rows, 0:width]
    bid = (y // B) * ((width + B - 1) // B) + x // B
    src = bid * B * B + (y % B) * B + (x % B)
    return px[src.ravel()].reshape(rows, width, 4)
