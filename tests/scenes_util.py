"""Small hand-built scenes for the integrator tests (closed forms)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from appearance_amd._ffi import Camera, Material, Sky, lib
from appearance_amd.scene import SceneData
from appearance_amd.world import look_at


def material(**kw) -> Material:
    m = Material()
    lib().apt_material_default(C.byref(m))
    for k, v in kw.items():
        if isinstance(v, (tuple, list)):
            getattr(m, k)[:] = v
        else:
            setattr(m, k, v)
    return m


def camera(eye, target, fov=60.0) -> Camera:
    c = Camera()
    t = look_at(eye, target)
    c.transform[:] = t.T.ravel().tolist()
    c.fov_deg = fov
    c.near_plane = 0.01
    c.far_plane = 1000.0
    return c


def quad(a, b, c, d):
    return [np.concatenate([a, b, c]), np.concatenate([a, c, d])]


def closed_box(albedo: float, emission: float, half: float = 1.0) -> SceneData:
    """Closed cube, every face Lambertian with the same albedo and emitting
    inwards: radiance is Le / (1 - albedo) everywhere inside."""
    h = half
    P = {(sx, sy, sz): np.array([sx * h, sy * h, sz * h], np.float32)
         for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)}
    faces = []
    # each quad wound so its normal points into the box
    faces += quad(P[(-1, -1, -1)], P[(1, -1, -1)], P[(1, -1, 1)], P[(-1, -1, 1)])  # floor, +y
    faces += quad(P[(-1, 1, -1)], P[(-1, 1, 1)], P[(1, 1, 1)], P[(1, 1, -1)])      # ceiling, -y
    faces += quad(P[(-1, -1, -1)], P[(-1, -1, 1)], P[(-1, 1, 1)], P[(-1, 1, -1)])  # x = -1, +x
    faces += quad(P[(1, -1, -1)], P[(1, 1, -1)], P[(1, 1, 1)], P[(1, -1, 1)])      # x = +1, -x
    faces += quad(P[(-1, -1, -1)], P[(-1, 1, -1)], P[(1, 1, -1)], P[(1, -1, -1)])  # z = -1, +z
    faces += quad(P[(-1, -1, 1)], P[(1, -1, 1)], P[(1, 1, 1)], P[(-1, 1, 1)])      # z = +1, -z
    pos = np.array(faces, np.float32)
    # orient every triangle's winding normal towards the centre (inwards)
    for t in pos:
        a, b, c = t[0:3], t[3:6], t[6:9]
        if np.cross(b - a, c - a) @ (a + b + c) > 0:
            t[3:6], t[6:9] = c.copy(), b.copy()
    m = material(color=(albedo,) * 3, emission=(emission,) * 3, bsdf_model=1)
    sky = Sky()
    return SceneData(pos, None, np.zeros(len(pos), np.uint32), [m],
                     camera((0.0, 0.0, 0.3), (0.3, -0.2, -1.0)), sky)


def sun_plane(albedo: float, sun_dir_travel, sun_size: float = 0.5, intensity: float = 100.0):
    """A large Lambertian floor lit only by the sun (black sky), seen from above."""
    G = 50.0
    pos = np.array(quad(np.array([-G, 0, G], np.float32), np.array([G, 0, G], np.float32),
                        np.array([G, 0, -G], np.float32), np.array([-G, 0, -G], np.float32)),
                   np.float32)
    m = material(color=(albedo,) * 3, bsdf_model=1)
    sky = Sky()
    d = np.asarray(sun_dir_travel, np.float64)
    d /= np.linalg.norm(d)
    sky.sun_direction[:] = d.tolist()
    sky.sun_size = sun_size
    sky.sun_color[:] = [1.0, 1.0, 1.0]
    sky.sun_intensity = intensity
    sky.sun_enabled = 1.0
    return SceneData(pos, None, np.zeros(2, np.uint32), [m], camera((0.0, 2.0, 0.0), (0.1, 0.0, -0.3)),
                     sky)


def sun_plane_radiance(albedo, sky) -> float:
    """Closed form: rho * L_sun * sin^2(alpha) * cos(theta_l) for a cap of half
    angle alpha fully above the surface (integral of cos over the cap)."""
    l = -np.array(sky.sun_direction[:], np.float64)
    l /= np.linalg.norm(l)
    cut = np.pi / 1.95
    Ls = sky.sun_intensity * max(0.0, 1 - np.exp(-((cut - np.arccos(l[1])) / 1.4)))
    alpha = sky.sun_size * 0.1
    return albedo * Ls * np.sin(alpha) ** 2 * l[1]


def deep_tree_triangles(n: int = 600, base: float = 1.5) -> np.ndarray:
    """Tiny triangles at x = base^-k, k = 0..n-1: each binned-SAH split peels a
    few triangles off the far end, so the tree is a long caterpillar whose
    4- and 8-wide collapses need more traversal stack than the kernels hold
    (the 2-wide one fits)."""
    x = base ** (-np.arange(n, dtype=np.float64))
    s = x * 0.01
    t = np.zeros((n, 9), np.float64)
    t[:, 0] = x
    t[:, 3] = x + s
    t[:, 6] = x
    t[:, 7] = s
    return np.ascontiguousarray(t, np.float32)


def deep_tree_scene() -> SceneData:
    pos = deep_tree_triangles()
    m = material(color=(0.5,) * 3, bsdf_model=1)
    return SceneData(pos, None, np.zeros(len(pos), np.uint32), [m],
                     camera((0.5, 0.0, 2.0), (0.5, 0.0, 0.0)), Sky())
