"""Benchmark of the path-tracing hot path on one MI355X.

Default workload (disney): the Disney-BSDF scene (256 UV spheres of 2048
triangles + a 2-triangle ground, 524,290 triangles, seed 3) with the analytic
sky and sun, 1920x1088 (1080 rounded up to the 64-row block layout), 4 spp,
depth 8, Russian roulette from bounce 3, FP32, the wavefront path tracer.
One step is one frame (apt_render) with the scene resident in HBM and the
frame resolved into device buffers; `value` is Msamples/s over the timed
frames, Mrays/s (extension + shadow rays) beside it.  The same frame through
the render-node entry (RGBA8 block layout read back to host memory) is
reported as msamples_s_render_node.

roofline: the dominant kernel, the extend stage (k_extend, closest-hit
traversal of the extension rays): its algorithmic bytes per frame (the
extension rays per sample times the bytes per closest-hit ray of the canonical
BVH2, tests/golden/disney_sample_counts.json) over its device time per frame
(HIP events around each launch, on the launch stream).

Beside it: the BVH-stress traversal (1,048,576 random triangles, 2,097,152
rays, closest and any hit) as `bvh_stress`, and the CPU baseline: the FP32
oracle path tracer on a bounded pixel sample of the same frame.

  python bench.py [--gpus 1] [--steps K] [--warmup W]
                  [--workload disney|bvh_stress|restir|serving|bvh_build|jpeg]
                  [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Msamples/s & Mrays/s at 1080p; served p50/p99 latency vs one-at-a-time"
HBM_PEAK_GBS = 8000.0
COUNTS = ROOT / "tests" / "golden" / "bvh_stress_counts.json"
DISNEY_COUNTS = ROOT / "tests" / "golden" / "disney_sample_counts.json"


COPY_BYTES = 256 << 20
TRAFFIC = ROOT / "profiles" / "trace_closest_bvh_stress_traffic.md"
EXTEND_TRAFFIC = ROOT / "profiles" / "extend_disney_traffic.md"
JPEG_TRAFFIC = ROOT / "profiles" / "jpeg_node_blocks_traffic.md"


def measured_traffic(path=TRAFFIC, key="traffic_bytes_per_launch"):
    """HBM bytes per launch from a committed FETCH_SIZE/WRITE_SIZE counter
    summary (already scaled by the copy-kernel calibration), or None."""
    import re

    if not path.exists():
        return None
    m = re.search(key + r"\s*=\s*([0-9.eE+]+)", path.read_text())
    return float(m.group(1)) if m else None


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def calibrate_copy() -> None:
    """One 256 MiB fill and one 256 MiB device-to-device copy: the known-size
    kernels whose HBM counter readings scale a --pmc run's."""
    import torch

    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def host_cores() -> int:
    """Threads for the CPU baseline: the cores this process may run on, at
    most 16 (a one-GPU share of the box)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model() -> tuple[str, set]:
    name, flags = "unknown", set()
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name") and name == "unknown":
                name = line.split(":", 1)[1].strip()
            elif line.startswith("flags") and not flags:
                flags = set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return name, flags


def baseline_oracle():
    """The oracle build the CPU baseline times: -O3 -march=x86-64-v3 where
    the host has AVX2, FMA and BMI2 (the two builds give the same bits), else
    the generic -O3 build.  Returns (oracle, build flags, cpu model)."""
    import ctypes as C

    sys.path.insert(0, str(ROOT / "oracle"))
    import build as oracle_build  # noqa: E402  (oracle/build.py: the flag lists only)
    sys.path.pop(0)
    from tests._oracle import LIB, Oracle, load_oracle

    name, flags = cpu_model()
    v3 = oracle_build.OUT_V3
    if {"avx2", "fma", "bmi2"} <= flags and v3.exists():
        return Oracle(C.CDLL(str(v3))), oracle_build.baseline_flags(), name
    return load_oracle(), " ".join(oracle_build.FLAGS[:1] + ["-ffp-contract=off"]) + " (generic x86-64)", name


def bench_bvh_stress(args, steps: int | None = None):
    """The BVH-stress traversal: closest and any hit over all 2,097,152 rays
    with rays and hits resident in HBM.  Returns its result dict."""
    import numpy as np
    import torch

    from appearance_amd.raytrace import BUILDER_GPU, Bvh, rays_bvh_stress
    from appearance_amd.scene import generate_scene

    steps = args.steps if steps is None else steps
    log("bvh_stress: generating the scene and rays")
    scene = generate_scene("bvh_stress", seed=1)
    rays = rays_bvh_stress(seed=2, side=1024, n_incoherent=1 << 20)
    n = rays.shape[0]
    # Built twice: the first build in a process also loads the builder's code
    # objects; the second one's time is the one reported.
    bvh = Bvh(scene.positions, width=args.width, lds_stack=args.stress_lds_stack, builder=BUILDER_GPU)
    del bvh
    bvh = Bvh(scene.positions, width=args.width, lds_stack=args.stress_lds_stack, builder=BUILDER_GPU)
    st = bvh.stats()
    d_rays = torch.from_numpy(rays.view(np.uint8).copy()).cuda()
    d_hits = torch.empty(n * 16, dtype=torch.uint8, device="cuda")
    d_occ = torch.empty(n, dtype=torch.int32, device="cuda")
    # An explicit stream (default priority): the kernels and the timing events
    # must share it, and torch's default stream has the null handle, which
    # libapt would replace by its own stream.
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    if args.calibrate_copy:
        calibrate_copy()
    for _ in range(args.warmup):
        bvh.trace_closest_device(d_rays, n, d_hits, stream)
    torch.cuda.synchronize()
    log(f"bvh_stress: timing {steps} closest-hit launches")
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        starts[k].record(stream)
        bvh.trace_closest_device(d_rays, n, d_hits, stream)
        ends[k].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = [a.elapsed_time(b) for a, b in zip(starts, ends)]
    avg_ms = sum(launch_ms) / len(launch_ms)
    hits = d_hits.view(torch.int32).view(n, 4)[:, 3].cpu().numpy().view(np.uint32)
    n_hit = int((hits != 0xFFFFFFFF).sum())
    for _ in range(2):
        bvh.trace_any_device(d_rays, n, d_occ, stream)
    ea, eb = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ea.record(stream)
    for _ in range(steps):
        bvh.trace_any_device(d_rays, n, d_occ, stream)
    eb.record(stream)
    torch.cuda.synchronize()
    any_ms = ea.elapsed_time(eb) / steps
    counts = json.loads(COUNTS.read_text())
    bpr = counts["closest_all"]["alg_bytes_per_ray"]
    achieved = bpr * n / (avg_ms * 1e-3) / 1e9
    return {
        "scene": scene, "rays": rays, "wall": wall, "steps": steps, "n": n, "bpr": bpr,
        "mrays_s_closest": n * steps / wall / 1e6,
        "mrays_s_any": n / (any_ms * 1e-3) / 1e6,
        "closest_hits": n_hit,
        "config": {"n_tris": scene.n_tris, "n_rays": n, "bvh_width": args.width, "lds_stack": args.stress_lds_stack,
                   "bvh_nodes": st["n_nodes"], "bvh_build_ms": round(st["build_ms"], 1)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "alg_bytes_per_ray": bpr, "kernel_avg_ms": avg_ms},
    }


def cpu_baseline_trace(tris, rays):
    """The oracle's BVH2 traversal on the host cores, on a bounded sample."""
    import numpy as np

    o, flags, model = baseline_oracle()
    threads = host_cores()
    t0 = time.time()
    b = o.bvh(tris)
    build_s = time.time() - t0
    n_half = 1 << 20
    rng = np.random.default_rng(0)
    # 1/4 of each ray set, so the sample has the workload's coherent/incoherent mix
    k = n_half // 4
    idx = np.concatenate([rng.choice(n_half, k, replace=False),
                          n_half + rng.choice(n_half, k, replace=False)])
    sample = np.ascontiguousarray(rays.view(np.float32).reshape(-1, 8)[np.sort(idx)])
    b.trace(sample[:4096], threads=threads)  # warm-up
    times = []
    for _ in range(5):
        t0 = time.time()
        b.trace(sample, threads=threads)
        times.append(time.time() - t0)
    t = sorted(times)[2]
    return {"value": sample.shape[0] / t / 1e6, "unit": "Mrays/s", "cores": threads,
            "kind": "port", "cpu_model": model, "build_flags": flags,
            "sample": f"{sample.shape[0]} rays (1/4 of the coherent and incoherent sets), "
                      f"closest hit, oracle BVH2 built in {build_s:.1f} s, warm-up then median of 5"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=4)
    ap.add_argument("--lds-stack", type=int, default=8)
    ap.add_argument("--stress-lds-stack", type=int, default=12,
                    help="LDS stack entries of the BVH-stress traversal (its incoherent rays into random "
                         "triangles overflow 8 entries; 12 is faster there and all but ends the overflow writes)")
    ap.add_argument("--workload", default="disney",
                    choices=["disney", "bvh_stress", "restir", "serving", "bvh_build", "jpeg"])
    ap.add_argument("--requests", type=int, default=96)
    ap.add_argument("--rates", default="20,200")
    ap.add_argument("--persistent", type=int, default=0,
                    help="serving: 1 = one persistent kernel per batch, 0 = wavefront stages")
    ap.add_argument("--policy", type=int, default=0,
                    help="serving: 0 = one-batch requests shortest first, then the rest in "
                         "arrival order; 1 = FIFO; 2 = shortest remaining work first")
    ap.add_argument("--gi-spatial-passes", type=int, default=0,
                    help="restir: ReSTIR GI spatial passes (0, as the reference ships its configuration)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-bvh-stress", action="store_true",
                    help="disney: skip the traversal leg reported beside the headline")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 256 MiB device-to-device copy before timing (the known-size "
                         "kernel that scales the HBM counters of a --pmc run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              f"(slabs, rays and requests would shard per GPU with no exchange)", file=sys.stderr)
        return 2

    if args.workload == "disney":
        return bench_disney(args)
    if args.workload == "serving":
        return bench_serving(args)
    if args.workload == "restir":
        return bench_restir(args)
    if args.workload == "bvh_build":
        return bench_bvh_build(args)
    if args.workload == "jpeg":
        return bench_node_jpeg(args)

    import torch

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    r = bench_bvh_stress(args)
    out = {
        "metric": METRIC,
        "value": r["mrays_s_closest"],
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": r["steps"],
        "warmup": args.warmup,
        "ms_per_step": r["wall"] * 1e3 / r["steps"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "bvh_stress: 1,048,576 triangles, 2,097,152 rays "
                               "(1M coherent 1024x1024 primary + 1M incoherent), closest hit",
                   **r["config"]},
        "roofline": r["roofline"],
        "closest_hits": r["closest_hits"],
        "any_hit_mrays_s": r["mrays_s_any"],
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle BVH2 on host cores)")
        out["cpu_baseline"] = cpu_baseline_trace(r["scene"].positions, r["rays"])
    print(json.dumps(out), flush=True)
    return 0


def bench_bvh_build(args) -> int:
    """One step = one device binned-SAH build of the 1,048,576-triangle
    bvh_stress soup (triangles resident in HBM) into a 4-wide BVH; the host
    builder's result is checked byte for byte."""
    import numpy as np
    import torch

    from appearance_amd.raytrace import BUILDER_GPU, BUILDER_HOST, Bvh
    from appearance_amd.scene import generate_scene

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    log("generating bvh_stress scene")
    scene = generate_scene("bvh_stress", seed=1)
    n = scene.n_tris
    d_tri = torch.from_numpy(np.ascontiguousarray(scene.positions, np.float32).reshape(-1)).cuda()
    for _ in range(args.warmup):
        Bvh.from_device(d_tri, n, width=args.width, builder=BUILDER_GPU)
    torch.cuda.synchronize()
    log(f"timing {args.steps} device builds")
    ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b = Bvh.from_device(d_tri, n, width=args.width, builder=BUILDER_GPU)
        ms.append(b.stats()["build_ms"])
    wall = time.perf_counter() - t0
    log("host build and byte comparison")
    h = Bvh(scene.positions, width=args.width, builder=BUILDER_HOST)
    nh, th = h.export()
    ng, tg = b.export()
    same = bool(np.array_equal(nh, ng) and np.array_equal(th.view(np.uint32), tg.view(np.uint32)))
    out = {
        "metric": METRIC, "value": n * args.steps / wall / 1e6, "unit": "Mtris/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"bvh_build: {n:,} triangles, device binned SAH (32 bins, leaf <= 4), "
                               f"{args.width}-wide collapse", "bvh_width": args.width},
        "device_build_ms": sorted(ms)[len(ms) // 2], "host_build_ms": h.stats()["build_ms"],
        "bvh_nodes": h.stats()["n_nodes"], "identical_to_host_build": same,
    }
    if not args.no_cpu_baseline:
        from tests._oracle import load_oracle
        o = load_oracle()
        t0 = time.time()
        o.bvh(scene.positions)
        dt = time.time() - t0
        out["cpu_baseline"] = {"value": n / dt / 1e6, "unit": "Mtris/s", "cores": 1, "kind": "port",
                               "sample": f"the oracle's binned-SAH BVH2 build of all {n:,} triangles, "
                                         f"one thread, {dt:.1f} s"}
    print(json.dumps(out), flush=True)
    return 0


def cpu_baseline_render(scene, width, height, spp, bounces, rr, target_s=3.0):
    """The FP32 oracle path tracer on the host cores, on a bounded random pixel
    sample of the same frame (same scene, camera and per-pixel seeds): one
    warm-up run sizes the sample to about target_s per run, then the median of
    five runs."""
    import numpy as np

    from tests._oracle import oracle_render

    o, flags, model = baseline_oracle()
    threads = host_cores()
    t0 = time.time()
    b = o.bvh(scene.positions)
    build_s = time.time() - t0
    order = np.random.default_rng(0).permutation(width * height).astype(np.uint32)
    probe = order[:8192]
    t0 = time.time()
    oracle_render(o, scene, width, height, spp=spp, max_bounces=bounces, rr_start=rr, pixels=probe,
                  bvh=b, threads=threads)
    rate = probe.shape[0] / max(time.time() - t0, 1e-3)
    n = int(min(width * height, max(8192, rate * target_s)))
    pix = np.sort(order[:n])
    times = []
    for _ in range(5):
        t0 = time.time()
        oracle_render(o, scene, width, height, spp=spp, max_bounces=bounces, rr_start=rr, pixels=pix,
                      bvh=b, threads=threads)
        times.append(time.time() - t0)
    t = sorted(times)[2]
    return {"value": n * spp / t / 1e6, "unit": "Msamples/s", "cores": threads, "kind": "port",
            "cpu_model": model, "build_flags": flags,
            "sample": f"{n} random pixels of the {width}x{height} frame x {spp} spp (same scene, camera and "
                      f"seeds), FP32 oracle path tracer, BVH2 built in {build_s:.1f} s; warm-up, then median "
                      f"of 5 runs ({', '.join(f'{x:.2f}' for x in times)} s)"}


def bench_jpeg(pt, W, H, steps, cpu=True):
    """The node's per-block JPEG compression (turbojpeg::compress(image, 95,
    Sub2x2) of every 64x64 block, node.rs:65-79) of the rendered frame: the
    device kernels on the frame's blocks resident in HBM (HIP events around
    the three launches), the whole node step apt_start_render_jpeg (render,
    compress, files to host), and on the host the same blocks through
    libjpeg-turbo (Pillow's build, the library the reference's turbojpeg
    crate wraps, one thread as the node's callback runs it) and through the
    oracle port."""
    import io

    import numpy as np

    from appearance_amd import jpeg, world

    frame = pt.render_rgba((W, H))
    blocks = np.frombuffer(frame, np.uint8).reshape(-1, 64, 64, 4)
    n = blocks.shape[0]
    jpeg.compress_blocks(blocks, 95)  # warm-up: buffers, tables
    dev = []
    for _ in range(steps):
        files, ms = jpeg.compress_blocks(blocks, 95, timing=True)
        dev.append(ms)
    dev_ms = sorted(dev)[len(dev) // 2]
    file_bytes = sum(len(f) for f in files)
    # a full-chip batch: the frame's blocks repeated to 16384 (a 2^26-pixel slab)
    big = blocks[np.arange(jpeg.MAX_BLOCKS) % n]
    jpeg.compress_blocks(big, 95)
    big_ms = sorted(jpeg.compress_blocks(big, 95, timing=True)[1] for _ in range(5))[2]
    job = world.StartRenderData(width=W, height=H, row_start=0, row_end=H, frame_idx=0)
    got = []
    pt.start_render(job, lambda r, c, f, b: got.append(len(b)), jpeg_quality=95)
    k = max(3, steps // 4)
    t0 = time.perf_counter()
    for _ in range(k):
        pt.start_render(job, lambda r, c, f, b: None, jpeg_quality=95)
    node_ms = (time.perf_counter() - t0) * 1e3 / k
    t0 = time.perf_counter()
    for _ in range(k):
        pt.start_render(job, lambda r, c, f, b: None)
    node_raw_ms = (time.perf_counter() - t0) * 1e3 / k
    in_bytes = n * 64 * 64 * 4
    alg = in_bytes + file_bytes
    out = {
        "blocks": n, "quality": 95, "bytes_out": file_bytes,
        "compression_ratio": in_bytes / file_bytes,
        "device_ms": dev_ms, "mpix_s": n * 4096 / (dev_ms * 1e-3) / 1e6,
        "device_ms_16384_blocks": big_ms, "mpix_s_16384_blocks": jpeg.MAX_BLOCKS * 4096 / (big_ms * 1e-3) / 1e6,
        "node_step_ms": {"start_render_jpeg": node_ms, "start_render_rgba": node_raw_ms},
        "roofline": {"bound": "hbm", "kernel": "k_jpeg_blocks + k_jpeg_offsets + k_jpeg_pack",
                     "achieved": alg / (dev_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": alg / (dev_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "alg_bytes_per_launch": alg,
                     "traffic": measured_traffic(JPEG_TRAFFIC, "traffic_bytes_per_frame"),
                     "alg_bytes": "16384 B of RGBX read + the file written, per block"},
    }
    if cpu:
        from PIL import Image, features

        from tests._oracle import jpeg_block

        o, flags, model = baseline_oracle()
        imgs = [Image.fromarray(np.ascontiguousarray(b[..., :3])) for b in blocks]
        t0 = time.perf_counter()
        reps = 0
        while time.perf_counter() - t0 < 3.0:
            for im in imgs:
                im.save(io.BytesIO(), "JPEG", quality=95, subsampling=2)
            reps += 1
        lt_s = (time.perf_counter() - t0) / reps
        t0 = time.perf_counter()
        reps = 0
        while time.perf_counter() - t0 < 3.0:
            for b in blocks:
                jpeg_block(o, b, 95)
            reps += 1
        port_s = (time.perf_counter() - t0) / reps
        out["cpu_libjpeg_turbo"] = {
            "value": n * 4096 / lt_s / 1e6, "unit": "Mpix/s", "cores": 1, "kind": "reference library",
            "ms_per_frame": lt_s * 1e3,
            "sample": f"the {n} blocks of the frame, each compressed by libjpeg-turbo "
                      f"{features.version_feature('libjpeg_turbo')} (Pillow, RGB, q95, 4:2:0, SIMD) as the "
                      f"node's callback does, one thread, frames repeated for >= 3 s"}
        out["cpu_baseline"] = {
            "value": n * 4096 / port_s / 1e6, "unit": "Mpix/s", "cores": 1, "kind": "port",
            "ms_per_frame": port_s * 1e3, "cpu_model": model, "build_flags": flags,
            "sample": f"the {n} blocks of the frame through the oracle (scalar C++), one thread, "
                      f"frames repeated for >= 3 s"}
    return out


def bench_node_jpeg(args) -> int:
    """The render node's JPEG step alone (--workload jpeg): the disney frame
    at 1920x1088 rendered once, then its 510 blocks compressed on the device
    `--steps` times (the bench_jpeg leg)."""
    import torch

    from appearance_amd.renderer import PathTracerGpu, plain_config
    from appearance_amd.scene import generate_scene

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    W, H = 1920, 1088
    log("generating disney scene")
    scene = generate_scene("disney", seed=3)
    pt = PathTracerGpu(plain_config(sample_count=4, max_bounces=8, rr_start_bounce=3))
    pt.set_scene(scene)
    pt.set_camera(scene.camera)
    if args.calibrate_copy:
        calibrate_copy()
    t0 = time.perf_counter()
    r = bench_jpeg(pt, W, H, steps=args.steps, cpu=not args.no_cpu_baseline)
    wall = time.perf_counter() - t0
    out = {
        "metric": "render-node JPEG blocks: Mpix/s compressed on the device", "value": r["mpix_s"],
        "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": r["device_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": f"the disney frame {W}x{H} (4 spp, depth 8) as 510 64x64 RGBX blocks, "
                               f"JPEG q95 4:2:0 per block", "blocks": r["blocks"], "quality": 95},
        "wall_s": wall, **{k: v for k, v in r.items() if k not in ("mpix_s",)},
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_disney(args) -> int:
    import torch

    from appearance_amd.renderer import PathTracerGpu, plain_config
    from appearance_amd.scene import generate_scene

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    W, H, SPP, DEPTH, RR = 1920, 1088, 4, 8, 3  # 1088: slab heights are multiples of 64
    counts = json.loads(DISNEY_COUNTS.read_text())
    log("generating disney scene")
    scene = generate_scene("disney", seed=3)
    pt = PathTracerGpu(plain_config(sample_count=SPP, max_bounces=DEPTH, rr_start_bounce=RR,
                                      bvh_width=args.width, lds_stack_depth=args.lds_stack))
    pt.set_scene(scene)
    pt.set_camera(scene.camera)
    if args.calibrate_copy:
        calibrate_copy()
    for _ in range(args.warmup):
        pt.render_device((W, H))
    torch.cuda.synchronize()
    log(f"timing {args.steps} frames")
    frames = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pt.render_device((W, H))
        frames.append(pt.stats())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    samples = W * H * SPP
    rays = sum(f["extension_rays"] + f["shadow_rays"] for f in frames)
    ext = sum(f["extension_rays"] for f in frames) / len(frames)
    dev_ms = sum(f["ms_total"] for f in frames) / len(frames)
    ext_ms = sum(f["ms_extend"] for f in frames) / len(frames)
    shade_ms = sum(f["ms_shade"] for f in frames) / len(frames)
    shadow_ms = sum(f["ms_shadow"] for f in frames) / len(frames)
    # the render-node entry: the same frames resolved to RGBA8 in the block
    # layout and read back to host memory
    k_rn = max(3, args.steps // 4)
    t1 = time.perf_counter()
    for _ in range(k_rn):
        pt.render_rgba((W, H))
    wall_rn = time.perf_counter() - t1
    # extend stage: bytes per frame over its device time per frame (8 launches)
    ext_bytes = counts["alg_bytes"]["extend_stage_per_sample"] * samples
    achieved = ext_bytes / (ext_ms * 1e-3) / 1e9
    frame_bytes = counts["alg_bytes"]["sample"] * samples
    out = {
        "metric": METRIC,
        "value": samples * args.steps / wall / 1e6,
        "unit": "Msamples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"disney: 256 UV spheres x 2048 triangles + ground (524,290 triangles, seed 3), "
                               f"analytic sky + sun, {W}x{H} (1080 rounded up to 64-row blocks), {SPP} spp, "
                               f"depth {DEPTH}, RR from bounce {RR}, wavefront path tracer",
                   "width": W, "height": H, "spp": SPP, "max_bounces": DEPTH,
                   "bvh_width": args.width, "lds_stack": args.lds_stack,
                   "gi_spatial_passes": args.gi_spatial_passes},
        "mrays_s": rays / wall / 1e6,
        "device_ms_per_frame": dev_ms,
        "stage_ms_per_frame": {"extend": ext_ms, "shade": shade_ms, "shadow": shadow_ms,
                               "other": max(0.0, dev_ms - ext_ms - shade_ms - shadow_ms)},
        "extension_rays_per_sample": {"device": ext / samples,
                                      "oracle": counts["per_sample"]["extension_rays"]},
        "msamples_s_render_node": samples * k_rn / wall_rn / 1e6,
        "roofline": {"bound": "hbm", "kernel": "k_extend (closest-hit traversal of the extension rays)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(EXTEND_TRAFFIC, "traffic_bytes_per_launch"),
                     "alg_bytes_per_launch": ext_bytes / DEPTH,
                     "alg_bytes_per_ray": counts["alg_bytes"]["closest_ray"],
                     "kernel_avg_ms": ext_ms / DEPTH, "launches_per_frame": DEPTH},
        "frame_roofline": {"alg_bytes_per_sample": counts["alg_bytes"]["sample"],
                           "achieved": frame_bytes / (dev_ms * 1e-3) / 1e9,
                           "frac": frame_bytes / (dev_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
    }
    log("render-node JPEG blocks")
    out["node_jpeg"] = bench_jpeg(pt, W, H, steps=10, cpu=not args.no_cpu_baseline)
    if not args.no_bvh_stress:
        r = bench_bvh_stress(args, steps=10)
        out["bvh_stress"] = {"mrays_s_closest": r["mrays_s_closest"], "mrays_s_any": r["mrays_s_any"],
                             "closest_hits": r["closest_hits"], "roofline": r["roofline"], **r["config"]}
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle path tracer on host cores)")
        out["cpu_baseline"] = cpu_baseline_render(scene, W, H, SPP, DEPTH, RR)
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)
    return 0


def restir_orbit(k, frames=60):
    """Camera k of the 60-frame orbit around the many-light scene."""
    import math

    from tests.scenes_util import camera

    a = 2.0 * math.pi * (k % frames) / 240.0
    return camera((13.5 * math.sin(a), 7.5, 0.5 + 13.5 * math.cos(a)), (0.0, 0.0, 0.5), 40.0)


def restir_config(**kw):
    from appearance_amd.renderer import default_config

    return default_config(di_candidates=32, di_mcap=20, di_spatial_neighbours=5, di_spatial_radius=30.0,
                          restir_di=1, restir_gi=1, svgf=1, svgf_iterations=5, svgf_history=4, **kw)


def cpu_baseline_restir(scene, frames=3, w=480, h=270):
    """The oracle's ReSTIR integrator on the host cores on a bounded sample:
    a few frames of the orbit at 1/16 of the 1080p pixel count."""
    from tests._oracle import OracleRestir

    o, flags, model = baseline_oracle()
    threads = host_cores()
    r = OracleRestir(o, scene, w, h)
    cfg = restir_config()
    r.frame(cfg, camera=restir_orbit(0), frame_idx=0, threads=threads)  # warm-up, builds history
    t0 = time.time()
    for f in range(1, frames + 1):
        r.frame(cfg, camera=restir_orbit(f), frame_idx=f, threads=threads)
    t = time.time() - t0
    return {"value": w * h * frames / t / 1e6, "unit": "Msamples/s", "cores": threads, "kind": "port",
            "cpu_model": model, "build_flags": flags,
            "sample": f"{frames} frames of the orbit at {w}x{h} (1/16 of the 1080p pixels), 1 spp, "
                      f"same settings, FP32 oracle ReSTIR integrator"}


def bench_restir(args) -> int:
    import torch

    from appearance_amd.renderer import PathTracerGpu
    from appearance_amd.scene import generate_scene

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    W, H = 1920, 1088  # RGBA8 block layout: slab heights are multiples of 64
    log("generating the many-light scene")
    scene = generate_scene("restir", seed=3)
    pt = PathTracerGpu(restir_config(bvh_width=args.width, lds_stack_depth=args.lds_stack,
                                     gi_spatial_passes=args.gi_spatial_passes))
    pt.set_scene(scene)
    k = 0
    for _ in range(max(1, args.warmup)):
        pt.set_camera(restir_orbit(k))
        pt.render_rgba((W, H))
        k += 1
    torch.cuda.synchronize()
    log(f"timing {args.steps} frames of the 60-frame orbit")
    dev_ms, rays = [], 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pt.set_camera(restir_orbit(k))
        pt.render_rgba((W, H))
        k += 1
        st = pt.stats()
        dev_ms.append(st["ms_total"])
        rays += st["extension_rays"] + st["shadow_rays"]
    wall = time.perf_counter() - t0
    out = {
        "metric": METRIC,
        "value": W * H * args.steps / wall / 1e6,
        "unit": "Msamples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"restir: disney scene + 4096 emissive triangles (528,386 triangles), {W}x{H}, "
                               "1 spp per frame, camera orbit, ReSTIR DI (32 candidates, M cap 20, 5 taps in "
                               "30 px) + GI reuse, depth 2, firefly filter, SVGF 5 passes / 4 frames",
                   "bvh_width": args.width, "lds_stack": args.lds_stack,
                   "gi_spatial_passes": args.gi_spatial_passes},
        "mrays_s": rays / wall / 1e6,
        "device_ms_per_frame": sum(dev_ms) / len(dev_ms),
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle ReSTIR integrator on host cores)")
        out["cpu_baseline"] = cpu_baseline_restir(scene)
    print(json.dumps(out), flush=True)
    return 0


def size_classes(reqs, lat_ms, span_s):
    """Requests per size class (samples <= 1 M, <= 16 M, more): count,
    requests/s over the run, p50 and p99 latency."""
    import numpy as np

    out = {}
    for name, lo, hi in (("small <=1M samples", 0, 1 << 20), ("medium <=16M", 1 << 20, 1 << 24),
                         ("large >16M", 1 << 24, 1 << 62)):
        ls = [l for r, l in zip(reqs, lat_ms) if lo < r["w"] * r["h"] * r["spp"] <= hi]
        if ls:
            out[name] = {"requests": len(ls), "req_s": len(ls) / span_s,
                         "p50_ms": float(np.percentile(ls, 50)), "p99_ms": float(np.percentile(ls, 99))}
    return out


def bench_serving(args) -> int:
    import threading

    import numpy as np
    import torch

    from appearance_amd.renderer import PathTracerGpu, plain_config
    from appearance_amd.scene import generate_scene
    from appearance_amd.serving import DONE, Server, mixed_requests, orbit_camera

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    log("generating the four resident scenes")
    names = ["cornell", "bvh_stress", "disney", "restir"]
    scenes = [generate_scene(n, seed=[0, 1, 3, 3][i]) for i, n in enumerate(names)]
    sky = scenes[2].sky
    scenes[1].sky = sky  # the triangle soup has no emitters: light it with the sky
    srv = Server(max_active=64, policy=args.policy, persistent=args.persistent)
    ids = [srv.add_scene(sc) for sc in scenes]
    plain = []
    for sc in scenes:
        pt = PathTracerGpu(plain_config())
        pt.set_scene(sc)
        plain.append(pt)
    report = {"rates": {}}
    total_samples = None
    rates = [float(r) for r in args.rates.split(",")]
    for rate in rates:
        reqs = mixed_requests(args.requests, rate)
        for r in reqs:
            r["cam"] = orbit_camera(scenes[r["scene"]], r["angle"])
        samples = sum(r["w"] * r["h"] * r["spp"] for r in reqs)
        total_samples = samples
        # served: a client thread submits at the arrival times
        log(f"served stream: {len(reqs)} requests at {rate} req/s, {samples/1e6:.0f} Msamples")
        rids = [None] * len(reqs)
        t0 = time.perf_counter()

        def client():
            for i, r in enumerate(reqs):
                if i % 128 == 127:
                    log(f"served: {i + 1} of {len(reqs)} requests submitted")
                d = r["t"] - (time.perf_counter() - t0)
                if d > 0:
                    time.sleep(d)
                rids[i] = srv.submit(ids[r["scene"]], r["w"], r["h"], r["spp"], r["depth"], r["cam"],
                                     seed=r["seed"])

        busy0 = srv.stats()["gpu_busy_ms"]
        th = threading.Thread(target=client)
        th.start()
        th.join()
        lat = []
        last, first = 0.0, float("inf")
        for k, rid in enumerate(rids):
            if k % 128 == 127:
                log(f"served: {k + 1} of {len(rids)} requests done")
            res = srv.wait(rid)
            assert res.state == DONE
            lat.append(res.done_ms - res.submit_ms)
            last = max(last, res.done_ms)
            first = min(first, res.submit_ms)
            srv.release(rid)
        span_s = (last - first) / 1e3
        busy_s = (srv.stats()["gpu_busy_ms"] - busy0) / 1e3
        served = {"msamples_s": samples / span_s / 1e6, "p50_ms": float(np.percentile(lat, 50)),
                  "p99_ms": float(np.percentile(lat, 99)), "makespan_s": span_s,
                  "gpu_busy_s": busy_s, "gpu_busy_frac": busy_s / span_s,
                  "by_size": size_classes(reqs, lat, span_s)}
        # one at a time: FIFO through the plain renderer
        log("one-at-a-time baseline")
        lat_b = []
        dev_b = 0.0
        t0 = time.perf_counter()
        for k, r in enumerate(reqs):
            if k % 128 == 127:
                log(f"one at a time: {k + 1} of {len(reqs)} requests done")
            d = r["t"] - (time.perf_counter() - t0)
            if d > 0:
                time.sleep(d)
            pt = plain[r["scene"]]
            pt.set_config(sample_count=r["spp"], max_bounces=r["depth"], seed=r["seed"], frame_idx=0)
            pt.set_camera(r["cam"])
            pt.render_accum((r["w"], r["h"]), with_rgba=False)
            lat_b.append((time.perf_counter() - t0 - r["t"]) * 1e3)
            dev_b += pt.stats()["ms_total"] / 1e3
        span_b = time.perf_counter() - t0 - reqs[0]["t"]
        base = {"msamples_s": samples / span_b / 1e6, "p50_ms": float(np.percentile(lat_b, 50)),
                "p99_ms": float(np.percentile(lat_b, 99)), "makespan_s": span_b,
                "gpu_busy_s": dev_b, "gpu_busy_frac": dev_b / span_b,
                "by_size": size_classes(reqs, lat_b, span_b)}
        report["rates"][str(rate)] = {"served": served, "one_at_a_time": base,
                                      "throughput_ratio": served["msamples_s"] / base["msamples_s"]}
        log(f"rate {rate}: served {served}, baseline {base}")
    top = report["rates"][str(rates[-1])]
    out = {
        "metric": METRIC,
        "value": top["served"]["msamples_s"],
        "unit": "Msamples/s",
        "n_gpus": 1, "steps": args.requests, "warmup": 0,
        "ms_per_step": top["served"]["makespan_s"] * 1e3 / args.requests,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"serving: {args.requests} requests, Poisson arrivals, 4 resident "
                               f"scenes, res in {{256^2,512^2,1280x720,1920x1080}}, spp in "
                               f"{{1,4,16,64}}, depth 4..8", "total_samples": total_samples, "policy": args.policy,
                               "persistent": args.persistent},
        "serving": report,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
