"""The C-ABI library loads on a host without a GPU, exports every function
declared in include/apt.h, and its host-only entry points (procedural scene
and ray generation, defaults) behave deterministically."""
import ctypes as C
import re
from pathlib import Path

import numpy as np

from appearance_amd import _ffi
from appearance_amd.raytrace import rays_bvh_stress
from appearance_amd.scene import generate_scene

HEADER = Path(__file__).resolve().parents[1] / "include" / "apt.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(apt_[a-z0-9_]+)\s*\(", text))
    return sorted(n for n in names if not n.endswith("_cb") or n == "apt_render_cb")


def test_exports_every_declared_symbol():
    lib = _ffi.lib()
    names = declared_functions()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_abi_version_and_errors():
    lib = _ffi.lib()
    assert lib.apt_abi_version() == 1
    h = C.c_void_p()
    st = lib.apt_scene_generate(b"no_such_scene", 0, C.byref(h))
    assert st == 6
    assert b"no_such_scene" in lib.apt_last_error()


def test_cornell_scene():
    s = generate_scene("cornell")
    assert s.n_tris == 36
    assert len(s.materials) == 4
    assert tuple(s.materials[3].emission) == (17.0, 12.0, 4.0)
    assert s.materials[0].bsdf_model == 1
    assert s.camera.fov_deg == 40.0


def test_bvh_stress_scene_deterministic():
    a = generate_scene("bvh_stress:4096", seed=1)
    b = generate_scene("bvh_stress:4096", seed=1)
    c = generate_scene("bvh_stress:4096", seed=2)
    assert a.positions.shape == (4096, 9)
    np.testing.assert_array_equal(a.positions, b.positions)
    assert not np.array_equal(a.positions, c.positions)
    cen = a.positions.reshape(-1, 3, 3).mean(axis=1)
    assert np.all(np.abs(cen) < 10.5)
    e1 = a.positions[:, 3:6] - a.positions[:, 0:3]
    ln = np.log(np.linalg.norm(e1, axis=1))
    assert abs(ln.mean() + 2.0) < 0.05 and abs(ln.std() - 0.5) < 0.05


def test_disney_scene_shape():
    s = generate_scene("disney", seed=3)
    assert s.n_tris == 256 * 2048 + 2
    assert len(s.materials) == 257
    assert s.normals is not None


def test_rays_bvh_stress():
    r = rays_bvh_stress(seed=2, side=64, n_incoherent=1000)
    assert r.shape == (64 * 64 + 1000,)
    d = r["dir"]
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    assert np.all(r["org"][:4096] == np.array([0, 0, -25], np.float32))
    assert np.all(np.abs(r["org"][4096:]) <= 10.0)
    assert np.all(r["tmax"] == np.float32(1e30))


def test_start_render_data_bytes():
    from appearance_amd import world
    d = world.StartRenderData(1920, 1088, 64, 512, 9)
    b = d.to_bytes()
    assert len(b) == 20 and b[:4] == (1920).to_bytes(4, "little")
    assert world.StartRenderData.from_bytes(b) == d
