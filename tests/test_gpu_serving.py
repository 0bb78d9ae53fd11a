"""Serving loop (continuous batching) on the GPU through the C ABI:
results equal the plain renderer's, are bit-identical alone and in a mix,
cancellation and bad requests behave."""
import numpy as np
import pytest

from appearance_amd.renderer import PathTracerGpu, plain_config
from appearance_amd.scene import generate_scene
from appearance_amd.serving import CANCELLED, DONE, Server
from tests.scenes_util import camera

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scenes():
    return {"cornell": generate_scene("cornell"), "disney": generate_scene("disney", seed=3)}


def plain(scene, w, h, spp, bounces, rr=3, seed=0):
    pt = PathTracerGpu(plain_config(sample_count=spp, max_bounces=bounces, rr_start_bounce=rr,
                                      seed=seed))
    pt.set_scene(scene)
    pt.set_camera(scene.camera)
    return pt.render_accum((w, h), with_rgba=False)[1]


def test_matches_plain_renderer(gpu_available, scenes):
    srv = Server(path_slots=1 << 18)
    sid = srv.add_scene(scenes["cornell"])
    rid = srv.submit(sid, 160, 96, 8, 6, scenes["cornell"].camera)
    res = srv.wait(rid)
    assert res.state == DONE and res.paths == 160 * 96 * 8
    ref = plain(scenes["cornell"], 160, 96, 8, 6)
    np.testing.assert_allclose(srv.result(rid), ref, rtol=1e-5, atol=1e-6)


def test_alone_equals_mixed(gpu_available, scenes):
    cams = [scenes["disney"].camera, camera((3.0, 4.0, 9.0), (0.0, 0.5, 0.0), 50.0)]
    alone_srv = Server(path_slots=1 << 18)
    d_alone = alone_srv.add_scene(scenes["disney"])
    rid = alone_srv.submit(d_alone, 192, 128, 4, 6, cams[1], seed=7)
    alone_srv.wait(rid)
    alone = alone_srv.result(rid).copy()
    alone_srv.close()

    mix = Server(path_slots=1 << 18, tile_paths=8192)
    c = mix.add_scene(scenes["cornell"])
    d = mix.add_scene(scenes["disney"])
    ids = [mix.submit(c, 256, 256, 2, 4, scenes["cornell"].camera),
           mix.submit(d, 320, 200, 1, 8, cams[0]),
           mix.submit(d, 192, 128, 4, 6, cams[1], seed=7),
           mix.submit(c, 64, 64, 16, 6, scenes["cornell"].camera)]
    for i in ids:
        assert mix.wait(i).state == DONE
    np.testing.assert_array_equal(mix.result(ids[2]), alone)
    st = mix.stats()
    assert st["requests_done"] == 4 and st["batches"] >= 1


def test_cancel_and_bad_requests(gpu_available, scenes):
    srv = Server(path_slots=1 << 16, max_active=2)
    sid = srv.add_scene(scenes["cornell"])
    big = srv.submit(sid, 1024, 1024, 64, 8, scenes["cornell"].camera)
    small = srv.submit(sid, 64, 64, 1, 2, scenes["cornell"].camera)
    srv.cancel(big)
    assert srv.wait(big).state == CANCELLED
    assert srv.wait(small).state == DONE
    with pytest.raises(Exception):
        srv.submit(99, 64, 64, 1, 2, scenes["cornell"].camera)
    with pytest.raises(Exception):
        srv.submit(sid, 64, 64, 1, 99, scenes["cornell"].camera)
    with pytest.raises(Exception):
        srv.submit(sid, 0, 64, 1, 2, scenes["cornell"].camera)
