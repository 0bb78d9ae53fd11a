"""GPU integrator parity: libapt's wavefront path tracer (through the C ABI)
against the oracle path tracer on the same scene, camera and per-pixel seeds.

Tolerances: relMSE = sum((gpu - ref)^2) / sum(ref^2) over the frame.  GPU and
oracle draw the same random numbers, so paths only diverge where the device
and host transcendental functions round differently; the bar is relMSE
<= 1e-3 (the radiance parity tolerance, stated at 4096 spp) and tighter where stated."""
import numpy as np
import pytest

from appearance_amd import world
from appearance_amd.renderer import PathTracerGpu, block_to_linear, default_config, plain_config
from appearance_amd.scene import generate_scene
from tests._oracle import oracle_render
from tests.scenes_util import closed_box, sun_plane, sun_plane_radiance

pytestmark = pytest.mark.gpu


def relmse(a, b):
    return float(np.sum((a.astype(np.float64) - b) ** 2) / np.sum(b.astype(np.float64) ** 2))


def gpu_render(scene, w, h, spp, bounces, rr, seed=0, **cfg):
    pt = PathTracerGpu(plain_config(sample_count=spp, max_bounces=bounces, rr_start_bounce=rr,
                                      seed=seed, **cfg))
    pt.set_scene(scene)
    pt.set_camera(scene.camera)
    rgba, acc = pt.render_accum((w, h))
    return pt, rgba, acc


def test_cornell_vs_oracle_fp32(gpu_available, oracle):
    s = generate_scene("cornell")
    print("[render] cornell 512x512 16 spp", flush=True)
    pt, rgba, acc = gpu_render(s, 512, 512, 16, 6, 3)
    ref, counts = oracle_render(oracle, s, 512, 512, spp=16, max_bounces=6, rr_start=3)
    r = relmse(acc, ref)
    print(f"[render] cornell relMSE gpu vs oracle fp32 = {r:.3e}", flush=True)
    assert r < 1e-4
    st = pt.stats()
    assert st["paths"] == 512 * 512 * 16
    # the two tracers trace the same rays but for diverged paths
    assert abs(st["extension_rays"] - counts[0]) / counts[0] < 1e-3


def test_cornell_high_spp_vs_fp64(gpu_available, oracle):
    """The parity run at 4096 spp.  (1) Same seeds: the GPU's full frame
    against the FP64 oracle on a 256-pixel subset (rounding divergence only).
    (2) Against the converged answer: the GPU's 4096-spp frame and the FP32
    oracle's (same seeds) both against the FP64 2^16-spp mean of the
    committed 4096-pixel fixture (tests/golden/make_cornell_fp64.py, other
    seeds).  Both errors are stated; each is the noise of a 4096-spp estimate
    (per-pixel z rms 1), and the two agree with each other to relMSE 1e-3."""
    from tests.test_oracle_render import cornell_fp64_fixture, fp64_zscores
    s = generate_scene("cornell")
    print("[render] cornell 512x512 4096 spp", flush=True)
    _, _, acc = gpu_render(s, 512, 512, 4096, 6, 3)
    pix = np.random.default_rng(0).choice(512 * 512, 256, replace=False).astype(np.uint32)
    ref64, _ = oracle_render(oracle, s, 512, 512, spp=4096, max_bounces=6, rr_start=3,
                             pixels=pix, precision=64)
    g = acc.reshape(-1, 3)[pix]
    r = relmse(g, ref64)
    print(f"[render] cornell 4096 spp relMSE gpu vs oracle fp64, same seeds = {r:.3e}", flush=True)
    assert r <= 1e-3
    fpix, mean, bvar, nb = cornell_fp64_fixture()
    print("[render] cornell fp32 oracle 4096 spp on the fixture's 4096 pixels", flush=True)
    f32, _ = oracle_render(oracle, s, 512, 512, spp=4096, max_bounces=6, rr_start=3, pixels=fpix)
    gpu = acc.reshape(-1, 3)[fpix].astype(np.float64)
    f32 = f32.astype(np.float64)
    r_gpu, r_f32, r_pair = relmse(gpu, mean), relmse(f32, mean), relmse(gpu, f32)
    z_gpu = fp64_zscores(gpu, mean, bvar, nb, 4096)
    z_f32 = fp64_zscores(f32, mean, bvar, nb, 4096)
    zg, zf = float(np.sqrt((z_gpu ** 2).mean())), float(np.sqrt((z_f32 ** 2).mean()))
    noise = float(np.sum(bvar + bvar / nb) / np.sum(mean ** 2))
    print(f"[render] against the FP64 2^16-spp mean: relMSE gpu {r_gpu:.3e} (z rms {zg:.3f}), "
          f"fp32 oracle {r_f32:.3e} (z rms {zf:.3f}), expected noise {noise:.3e}; gpu vs fp32 {r_pair:.3e}",
          flush=True)
    assert 0.9 < zg < 1.1 and 0.9 < zf < 1.1
    assert r_pair <= 1e-3
    assert abs(gpu.mean() - mean.mean()) / mean.mean() < 2e-3


def test_furnace(gpu_available):
    s = closed_box(0.5, 1.0)
    _, _, acc = gpu_render(s, 64, 64, 64, 48, 3)
    assert abs(acc.mean() - 2.0) / 2.0 < 0.01, acc.mean()


def test_sun_plane(gpu_available):
    s = sun_plane(0.6, (-0.3, -1.0, 0.2))
    _, _, acc = gpu_render(s, 64, 64, 16, 1, 100)
    expect = sun_plane_radiance(0.6, s.sky)
    assert abs(acc.mean() - expect) / expect < 2e-3


def test_disney_vs_oracle(gpu_available, oracle):
    s = generate_scene("disney", seed=3)
    print("[render] disney 256x128 4 spp", flush=True)
    _, _, acc = gpu_render(s, 256, 128, 4, 8, 3)
    ref, _ = oracle_render(oracle, s, 256, 128, spp=4, max_bounces=8, rr_start=3)
    r = relmse(acc, ref)
    print(f"[render] disney relMSE gpu vs oracle fp32 = {r:.3e}", flush=True)
    assert r < 1e-3
    assert acc.mean() > 0.05


def test_disney_full_frame(gpu_available, oracle):
    """The disney workload at its full size (1920x1088, 4 spp, depth 8, RR
    from bounce 3).  Size-independent properties: one 2^24-path pass, four
    one-sample passes (2^20-path budget) and two slabs give the same frame
    bit for bit (per-pixel seeds, samples summed in order); and 16384
    sampled pixels match the oracle's same pixels, same seeds, to relMSE
    <= 1e-3."""
    s = generate_scene("disney", seed=3)
    w, h = 1920, 1088
    print("[render] disney 1920x1088 4 spp, one pass", flush=True)
    pt_full, _, full = gpu_render(s, w, h, 4, 8, 3)
    assert np.isfinite(full).all() and (full >= 0).all() and full.mean() > 0.05
    # the frame's ray counts equal the oracle's on the same seeds (the
    # committed per-sample figures the bench's algorithmic bytes come from)
    import json
    from pathlib import Path
    fix = json.loads((Path(__file__).parent / "golden" / "disney_sample_counts.json").read_text())
    st = pt_full.stats()
    for k in ("extension_rays", "shadow_rays"):
        got, want = st[k], fix["counts"][k]
        print(f"[render] disney {k}: device {got}, oracle {want}", flush=True)
        assert abs(got - want) <= 2e-4 * want
    assert st["ms_extend"] > 0 and st["ms_shade"] > 0 and st["ms_shadow"] > 0
    assert st["ms_extend"] + st["ms_shade"] + st["ms_shadow"] <= st["ms_total"] * 1.001
    print("[render] disney 1920x1088 4 spp, one sample per pass", flush=True)
    _, _, passes = gpu_render(s, w, h, 4, 8, 3, queue_chunk=1 << 20)
    np.testing.assert_array_equal(passes, full)
    for r0, r1 in ((0, 512), (512, 1088)):
        print(f"[render] disney slab rows {r0}..{r1}", flush=True)
        pt = PathTracerGpu(plain_config(sample_count=4, max_bounces=8, rr_start_bounce=3))
        pt.set_scene(s)
        pt.set_camera(s.camera)
        _, slab = pt.render_accum((w, h), r0, r1)
        np.testing.assert_array_equal(slab, full[r0:r1])
    print("[render] disney oracle on 16384 sampled pixels", flush=True)
    pix = np.random.default_rng(7).choice(w * h, 16384, replace=False).astype(np.uint32)
    ref, _ = oracle_render(oracle, s, w, h, spp=4, max_bounces=8, rr_start=3, pixels=pix)
    r = relmse(full.reshape(-1, 3)[pix], ref)
    print(f"[render] disney full frame relMSE on sampled pixels = {r:.3e}", flush=True)
    assert r < 1e-3


def test_slab_and_determinism(gpu_available):
    s = generate_scene("cornell")
    pt = PathTracerGpu(plain_config(sample_count=4, max_bounces=4))
    pt.set_scene(s)
    pt.set_camera(s.camera)
    _, full = pt.render_accum((128, 128))
    _, again = pt.render_accum((128, 128))
    pt2 = PathTracerGpu(plain_config(sample_count=4, max_bounces=4))
    pt2.set_scene(s)
    pt2.set_camera(s.camera)
    _, slab = pt2.render_accum((128, 128), 64, 128)
    # frame_idx advanced on pt; reseed a fresh renderer for the comparison
    np.testing.assert_array_equal(slab, full[64:128])
    assert not np.array_equal(again, full)  # next frame, new seeds


def test_block_layout_and_tonemap(gpu_available):
    s = generate_scene("cornell")
    pt = PathTracerGpu(plain_config(sample_count=2, max_bounces=3))
    pt.set_scene(s)
    pt.set_camera(s.camera)
    rgba, acc = pt.render_accum((192, 128))
    lin = block_to_linear(rgba, 192, 128)
    x = acc.astype(np.float32)
    tm = (x * (np.float32(2.51) * x + np.float32(0.03))) / (x * (np.float32(2.43) * x + np.float32(0.59)) + np.float32(0.14))
    expect = np.rint(np.clip(tm, 0, 1) * 255).astype(np.int32)
    assert np.abs(lin[..., :3].astype(np.int32) - expect).max() <= 1
    assert np.all(lin[..., 3] == 255)
    got = []
    pt.render((192, 128), 0, 128, lambda px: got.append(px))
    assert len(got) == 1 and len(got[0]) == 192 * 128 * 4


def test_world_actions_match_scene_set(gpu_available):
    s = generate_scene("cornell")
    a = PathTracerGpu(plain_config(sample_count=2, max_bounces=4))
    a.set_scene(s)
    a.set_camera(s.camera)
    _, ref = a.render_accum((64, 64))
    b = PathTracerGpu(plain_config(sample_count=2, max_bounces=4))
    b.handle_visible_world_action(world.SpawnModel("procedural://cornell?seed=0"))
    cam = world.CameraUpdate(np.array(s.camera.transform[:], np.float32).reshape(4, 4).T, 40.0)
    b.handle_visible_world_action(cam)
    _, got = b.render_accum((64, 64))
    np.testing.assert_array_equal(got, ref)
    b.handle_visible_world_action(world.Clear())
    with pytest.raises(Exception):
        b.render_accum((64, 64))


def _model_xform(angle_deg, tx):
    """Column-major-friendly 4x4 (row i = basis row): rotation about y, then a move along x."""
    a = np.deg2rad(angle_deg)
    m = np.eye(4, dtype=np.float32)
    m[0, 0], m[0, 2], m[2, 0], m[2, 2] = np.cos(a), np.sin(a), -np.sin(a), np.cos(a)
    m[0, 3] = tx
    return m


def test_model_actions_equal_direct_spawn(gpu_available):
    """TransformModel and DestroyModel (scene_resources/mod.rs:159-276 rebuilds
    the TLAS after them) leave the renderer as if the surviving models had
    been spawned at their final transforms: the image is bit-identical, and
    each action's scene rebuild (TLAS and placement) is timed in the stats."""
    import uuid
    cam = world.CameraUpdate(world.look_at((0.0, 1.0, 6.5), (0.0, 1.0, 0.0)), 50.0)
    ua, ub = uuid.UUID(int=1), uuid.UUID(int=2)
    t_b0, t_b1 = _model_xform(0.0, 2.5), _model_xform(30.0, -2.2)

    def renderer(*spawns):
        pt = PathTracerGpu(plain_config(sample_count=2, max_bounces=4))
        for u, t in spawns:
            pt.handle_visible_world_action(world.SpawnModel("procedural://cornell?seed=0", t, u))
        pt.handle_visible_world_action(cam)
        return pt

    a = renderer((ua, np.eye(4, dtype=np.float32)), (ub, t_b0))
    _, before = a.render_accum((128, 64))
    assert a.stats()["ms_scene"] > 0.0
    a.handle_visible_world_action(world.TransformModel(ub, t_b1))
    a.set_frame_index(0)
    _, moved = a.render_accum((128, 64))
    _, ref = renderer((ua, np.eye(4, dtype=np.float32)), (ub, t_b1)).render_accum((128, 64))
    np.testing.assert_array_equal(moved, ref)
    assert not np.array_equal(moved, before)
    a.handle_visible_world_action(world.DestroyModel(ua))
    a.set_frame_index(0)
    _, left = a.render_accum((128, 64))
    _, ref = renderer((ub, t_b1)).render_accum((128, 64))
    np.testing.assert_array_equal(left, ref)
    with pytest.raises(Exception):
        a.handle_visible_world_action(world.DestroyModel(ua))
    with pytest.raises(Exception):
        a.handle_visible_world_action(world.TransformModel(ua, t_b0))


def test_progressive_accumulation(gpu_available):
    s = generate_scene("cornell")
    one = PathTracerGpu(plain_config(sample_count=2, max_bounces=4, accum_frames=1))
    one.set_scene(s)
    one.set_camera(s.camera)
    _, f0 = one.render_accum((64, 64))
    _, f01 = one.render_accum((64, 64))
    two = PathTracerGpu(plain_config(sample_count=2, max_bounces=4, frame_idx=1))
    two.set_scene(s)
    two.set_camera(s.camera)
    _, f1 = two.render_accum((64, 64))
    np.testing.assert_allclose(f01, (f0 + f1) / 2, rtol=2e-6, atol=1e-6)


def test_invalid_requests_rejected(gpu_available):
    s = generate_scene("cornell")
    pt = PathTracerGpu(plain_config())
    pt.set_scene(s)
    for args in [((100, 64),), ((64, 64), 0, 32), ((64, 64), 64, 128)]:
        with pytest.raises(Exception):
            pt.render_accum(*args)


def test_autotune_cornell(gpu_available, tmp_path):
    from appearance_amd.autotune import tune
    s = generate_scene("cornell")
    rep = tune(s, probe=(128, 64), spp=2, bounces=4, repeats=1, json_path=tmp_path / "t.json")
    assert len(rep["candidates"]) == 18
    assert all(c["identical"] for c in rep["candidates"])
    best = rep["best"]
    assert best["bvh_width"] in (2, 4, 8) and best["lds_stack_depth"] in (8, 12, 16)
    assert (tmp_path / "t.json").exists()


def test_spawn_transform_matches_numpy_placement(gpu_available):
    """SpawnModel with a rotation, a translation and a non-uniform scale
    (the device placement kernel and the inverse-transpose normal matrix)
    against apt_scene_set fed the same model placed in numpy: positions
    M p in float32 with the kernel's order of operations, vertex normals
    through the inverse transpose of M's 3x3 part and renormalised.  The
    scenes are the Disney spheres (vertex normals) and the Cornell box
    (geometric normals).  The spawned model is traced in its object space
    (a two-level scene) and the placed one in world space, so hit distances
    differ in their last bits: the frames agree to relMSE 1e-6 (measured
    about 3e-7), not bit for bit."""
    from appearance_amd.scene import SceneData
    a = np.radians(35.0)
    rot = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])
    b = np.radians(-12.0)
    rot = rot @ np.array([[1, 0, 0], [0, np.cos(b), -np.sin(b)], [0, np.sin(b), np.cos(b)]])
    m = np.eye(4)
    m[:3, :3] = rot @ np.diag([1.3, 0.7, 1.15])
    m[:3, 3] = [0.4, -0.25, 0.6]
    m = m.astype(np.float32)
    f = np.float32
    for name, seed in (("disney", 3), ("cornell", 0)):
        s = generate_scene(name, seed=seed)
        p = s.positions.reshape(-1, 3).astype(np.float32)
        world_p = np.empty_like(p)
        for r in range(3):
            world_p[:, r] = (m[r, 0] * p[:, 0] + m[r, 1] * p[:, 1]) + m[r, 2] * p[:, 2] + m[r, 3]
        nrm = None
        if s.normals is not None:
            nm = np.linalg.inv(m[:3, :3].astype(np.float64)).T.astype(np.float32)
            q = s.normals.reshape(-1, 3).astype(np.float32)
            v = np.empty_like(q)
            for r in range(3):
                v[:, r] = (nm[r, 0] * q[:, 0] + nm[r, 1] * q[:, 1]) + nm[r, 2] * q[:, 2]
            ln = np.sqrt((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2])
            nrm = np.where(ln[:, None] > f(0), v / ln[:, None], v).reshape(-1, 9).astype(np.float32)
        placed = SceneData(world_p.reshape(-1, 9), nrm, s.material_ids, s.materials, s.camera, s.sky)
        spawn = PathTracerGpu(plain_config(sample_count=2, max_bounces=4))
        spawn.handle_visible_world_action(world.SpawnModel(f"procedural://{name}?seed={seed}", m))
        spawn.set_camera(s.camera)
        _, got = spawn.render_accum((256, 128), with_rgba=False)
        direct = PathTracerGpu(plain_config(sample_count=2, max_bounces=4))
        direct.set_scene(placed, sky=s.sky)
        direct.set_camera(s.camera)
        _, want = direct.render_accum((256, 128), with_rgba=False)
        r = relmse(got, want)
        print(f"[render] {name} spawned with a transform vs numpy placement: relMSE {r:.3e}, "
              f"bit-identical {np.array_equal(got, want)}", flush=True)
        assert want.mean() > 0.01
        assert r < 1e-6


def test_start_render_job(gpu_available):
    """apt_start_render: the 20-byte StartRenderData job renders the slab and
    hands over its 64x64 blocks in Node::start_render's order (node.rs:55-93),
    each tagged with its row, column block and the job's frame_idx; put
    together they are the slab that apt_render returns."""
    s = generate_scene("cornell")
    job = world.StartRenderData(192, 256, 64, 192, frame_idx=7)
    a = PathTracerGpu(plain_config(sample_count=2, max_bounces=3))
    a.set_scene(s)
    a.set_camera(s.camera)
    blocks = []
    a.start_render(job.to_bytes(), lambda row, col, frame, px: blocks.append((row, col, frame, px)))
    assert [(r, c) for r, c, _, _ in blocks] == [(64 + 64 * by, bx) for by in range(2) for bx in range(3)]
    assert all(f == 7 and len(px) == 64 * 64 * 4 for _, _, f, px in blocks)
    b = PathTracerGpu(plain_config(sample_count=2, max_bounces=3))
    b.set_scene(s)
    b.set_camera(s.camera)
    rgba = b.render_rgba((192, 256), 64, 192)
    assert b"".join(px for _, _, _, px in blocks) == bytes(rgba)
    with pytest.raises(Exception):
        a.start_render(world.StartRenderData(100, 256, 0, 64).to_bytes(), lambda *x: None)
    with pytest.raises(Exception):
        a.start_render(job.to_bytes()[:16], lambda *x: None)


def test_gltf_spawn_matches_oracle(gpu_available, oracle, tmp_path):
    """SpawnModel with a glTF file (tests/gltf_util.demo_scene: node
    hierarchy, generated normals, an emissive quad with emissive strength, a
    transmissive panel with ior and volume, clearcoat / sheen / specular)
    renders the same frame as apt_scene_set fed apt_scene_load_gltf's
    triangles, and matches the FP32 oracle on them (relMSE <= 1e-3)."""
    from appearance_amd._ffi import Sky
    from appearance_amd.scene import load_gltf
    from tests.gltf_util import demo_scene
    from tests.scenes_util import camera
    p = tmp_path / "demo.glb"
    demo_scene(p)
    s = load_gltf(p)
    cam = camera((0.5, 2.2, 6.0), (0.0, 0.8, 0.0), 45.0)
    pt = PathTracerGpu(plain_config(sample_count=32, max_bounces=6))
    pt.handle_visible_world_action(world.SpawnModel(str(p)))
    pt.set_camera(cam)
    _, spawned = pt.render_accum((192, 128), with_rgba=False)
    direct = PathTracerGpu(plain_config(sample_count=32, max_bounces=6))
    direct.set_scene(s, sky=Sky())
    direct.set_camera(cam)
    _, want = direct.render_accum((192, 128), with_rgba=False)
    np.testing.assert_array_equal(spawned, want)
    ref, _ = oracle_render(oracle, s, 192, 128, spp=32, max_bounces=6, rr_start=3, camera=cam, sky=Sky())
    r = relmse(spawned, ref)
    print(f"[render] glTF demo scene spawned: mean {spawned.mean():.4f}, relMSE vs oracle {r:.3e}", flush=True)
    assert spawned.mean() > 0.01
    assert r < 1e-3


def test_cornell_conor_spawn_matches_oracle(gpu_available, oracle):
    """The reference render host's Cornell box (assets/CornellBoxConor.glb,
    apps/render-host/src/lib.rs:138), re-encoded without its texture as
    tests/golden/cornell_conor.glb, spawned through SpawnModel and rendered
    against the FP32 oracle on the same triangles (relMSE <= 1e-3)."""
    from pathlib import Path
    from appearance_amd._ffi import Sky
    from appearance_amd.scene import load_gltf
    from tests.scenes_util import camera
    p = Path(__file__).resolve().parent / "golden" / "cornell_conor.glb"
    s = load_gltf(p)
    cam = camera((0.0, 1.0, 3.6), (0.0, 1.0, 0.0), 45.0)
    pt = PathTracerGpu(plain_config(sample_count=32, max_bounces=8))
    pt.handle_visible_world_action(world.SpawnModel(str(p)))
    pt.set_camera(cam)
    _, got = pt.render_accum((128, 128), with_rgba=False)
    ref, _ = oracle_render(oracle, s, 128, 128, spp=32, max_bounces=8, rr_start=3, camera=cam, sky=Sky())
    r = relmse(got, ref)
    print(f"[render] CornellBoxConor spawned: mean {got.mean():.4f}, relMSE vs oracle {r:.3e}", flush=True)
    assert got.mean() > 0.05
    assert r < 1e-3
