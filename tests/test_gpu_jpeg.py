"""Device JPEG compression of render-node blocks against the oracle.

The node compresses every 64x64 RGBX block it sends with
turbojpeg::compress(image, 95, Sub2x2) (node.rs:65-79).  libapt's kernels
must write the oracle's bytes exactly (the oracle is pinned to
libjpeg-turbo's output in test_oracle_jpeg.py), for single blocks, large
batches, and the blocks apt_start_render_jpeg sends for a rendered slab."""
from __future__ import annotations

import ctypes as C
import io

import numpy as np
import pytest

from tests._oracle import jpeg_block
from tests.jpeg_util import random_blocks, sample_blocks

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def jpeg(gpu_available):
    from appearance_amd import jpeg as j
    return j


@pytest.mark.parametrize("quality", [95, 50, 100, 1])
def test_sample_blocks_bit_exact(oracle, jpeg, quality):
    blocks = sample_blocks()
    names = sorted(blocks)
    print(f"jpeg sample blocks q={quality}", flush=True)
    got = jpeg.compress_blocks(np.stack([blocks[n] for n in names]), quality)
    for n, f in zip(names, got):
        assert f == jpeg_block(oracle, blocks[n], quality), (n, quality)


def test_every_quality_bit_exact(oracle, jpeg):
    """Qualities 1..100 (each scales the Annex K tables differently, with
    the 8-bit clamp at the low end): three blocks per quality in one call."""
    blocks = sample_blocks()
    picks = np.stack([blocks["smooth_noise"], blocks["edge"], blocks["noise"]])
    print("jpeg qualities 1..100", flush=True)
    for q in range(1, 101):
        got = jpeg.compress_blocks(picks, q)
        for i in range(len(picks)):
            assert got[i] == jpeg_block(oracle, picks[i], q), (q, i)


def test_single_block_api(oracle, jpeg):
    b = sample_blocks()["edge"]
    assert jpeg.compress(b) == jpeg_block(oracle, b, 95)


def test_batch_of_4096_bit_exact(oracle, jpeg):
    blocks = random_blocks(4096, seed=3)
    print("jpeg batch 4096", flush=True)
    got, ms = jpeg.compress_blocks(blocks, 95, timing=True)
    assert ms > 0.0
    for i in range(len(blocks)):
        assert got[i] == jpeg_block(oracle, blocks[i], 95), i


def test_max_batch_decodes_and_matches_samples(oracle, jpeg):
    """APT_JPEG_MAX_BLOCKS blocks in one call (a 2^26-pixel slab): every file
    is a well-formed JPEG, and a spread of them equals the oracle's."""
    n = jpeg.MAX_BLOCKS
    base = random_blocks(256, seed=11)
    blocks = base[np.arange(n) % 256].copy()
    blocks[:, 0, 0, 0] = (np.arange(n) // 256).astype(np.uint8)  # no two blocks alike
    print(f"jpeg batch {n}", flush=True)
    got = jpeg.compress_blocks(blocks, 95)
    assert len(got) == n
    for f in got:
        assert f[:2] == b"\xff\xd8" and f[-2:] == b"\xff\xd9"
    for i in list(range(0, n, 509)) + [n - 1]:
        assert got[i] == jpeg_block(oracle, blocks[i], 95), i


def test_errors(jpeg):
    from appearance_amd._ffi import lib
    L = lib()
    b = np.zeros((1, 64, 64, 4), np.uint8)
    out = np.empty(16, np.uint8)
    offs = (C.c_uint64 * 2)()
    # quality outside 1..100
    for q in (0, 101):
        assert L.apt_jpeg_compress_blocks(C.c_void_p(b.ctypes.data), 1, q, C.c_void_p(out.ctypes.data), out.size,
                                          offs, None) == 1
    # no blocks, too many blocks
    assert L.apt_jpeg_compress_blocks(C.c_void_p(b.ctypes.data), 0, 95, C.c_void_p(out.ctypes.data), out.size,
                                      offs, None) == 1
    assert L.apt_jpeg_compress_blocks(C.c_void_p(b.ctypes.data), jpeg.MAX_BLOCKS + 1, 95,
                                      C.c_void_p(out.ctypes.data), out.size, offs, None) == 1
    # an output that is too small reports the size it needs
    assert L.apt_jpeg_compress_blocks(C.c_void_p(b.ctypes.data), 1, 95, C.c_void_p(out.ctypes.data), out.size,
                                      offs, None) == 1
    assert offs[1] == len(jpeg.compress(b[0])) > out.size


@pytest.mark.parametrize("integrator", ["plain", "restir"])
def test_start_render_jpeg_sends_the_blocks_compressed(oracle, gpu_available, integrator):
    """apt_start_render_jpeg's callbacks carry the oracle's JPEG of each
    RGBA block apt_start_render hands over for the same frame, in the same
    order with the same row, column and frame fields (the plain integrator,
    and the default ReSTIR one with its post chain)."""
    from appearance_amd import world
    from appearance_amd.renderer import PathTracerGpu, default_config, plain_config
    from appearance_amd.scene import generate_scene
    s = generate_scene("cornell")
    cfg = plain_config(sample_count=2, max_bounces=3) if integrator == "plain" else default_config()
    job = world.StartRenderData(width=256, height=256, row_start=64, row_end=192, frame_idx=7)
    raw, comp = [], []
    # two fresh renderers: the same first frame (ReSTIR keeps history across frames)
    renderers = []
    for out, q in ((raw, None), (comp, 95)):
        pt = PathTracerGpu(cfg)
        pt.set_scene(s)
        pt.set_camera(s.camera)
        pt.set_frame_index(3)
        print(f"start_render{'_jpeg' if q else ''} 256x128 slab, {integrator}", flush=True)
        pt.start_render(job, lambda r, c, f, px, out=out: out.append((r, c, f, px)), jpeg_quality=q)
        renderers.append(pt)
    assert [x[:3] for x in raw] == [x[:3] for x in comp]
    assert len(comp) == 8
    for (r, c, f, px), (_, _, _, jp) in zip(raw, comp):
        blk = np.frombuffer(px, np.uint8).reshape(64, 64, 4)
        assert jp == jpeg_block(oracle, blk, 95), (r, c)
    # the host side decodes them (host.rs:341-350)
    from PIL import Image
    img = np.asarray(Image.open(io.BytesIO(comp[0][3])).convert("RGB"))
    assert img.shape == (64, 64, 3)
    # a slab that is not whole blocks is refused like apt_start_render's
    bad = world.StartRenderData(width=250, height=256, row_start=0, row_end=64, frame_idx=0)
    with pytest.raises(Exception):
        renderers[1].start_render(bad, lambda *a: None, jpeg_quality=95)
    for pt in renderers:
        pt.close()
