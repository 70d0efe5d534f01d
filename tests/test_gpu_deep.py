"""Paths deeper than 16 bounces (the reference takes any maxDepth prop,
render/device_cuda.cpp:68, render/diffuse_integrator_single.h:33).

Scene: a 2 x 2 x 16 tunnel closed at z = -8 and open at z = +8, with a sphere
inside, seen from inside looking at the closed end, so most paths bounce many
times before they escape through the far opening.  Bar: the accumulator is
bit-identical to the oracle's (kernel numerics) at max depth 32 on both frame
pipelines and in the batched form, the ray count equals the oracle's, and the
per-depth statistics put the depths >= 15 into the last bin.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W, H = 48, 32


def _tunnel(scenes):
    idx, pos = scenes.box_tris((-1.0, -1.0, -8.0), (1.0, 1.0, 8.0))
    keep = np.ones(len(idx), bool)
    keep[2:4] = False  # the z = +8 face: the tunnel's opening
    f, v = scenes.icosphere(2)
    return scenes.merge([(idx[keep], pos), (f, (v * 0.45 + np.array([0.3, -0.4, -5.0])).astype(np.float32))])


def _camera(orc, w, h):
    return orc.camera_from_view([0.0, 0.2, 6.0], 0.05, 0.0, math.radians(50.0), w / h)


@pytest.fixture(scope="module")
def tunnel(scenes, orc):
    mesh = _tunnel(scenes)
    return mesh, orc.Scene(mesh.idx, mesh.pos)


@pytest.mark.parametrize("pipeline", ["fused", "wavefront"])
def test_depth_32_bit_exact(prt, orc, tunnel, pipeline):
    mesh, sc = tunnel
    cam = _camera(orc, W, H)
    spp, depth, albedo = 2, 32, 0.9
    with prt.HipDevice(0) as d:
        d.upload_scene(mesh.idx, mesh.pos)
        d.build()
        d.set_pipeline(pipeline)
        d.set_camera(cam)
        d.init_frame(W, H)
        stats = [d.render("diffuseHip", max_depth=depth, albedo=albedo) for _ in range(spp)]
        got = d.accum()
        if pipeline == "fused":
            d.set_profiling(counters=True)
            d.clear_frame()
            prof = [d.render("diffuseHip", max_depth=depth, albedo=albedo) for _ in range(spp)]
            assert np.array_equal(d.accum(), got)
        else:
            prof = stats
    ref = sc.render(cam, W, H, spp=spp, max_depth=depth, albedo=albedo, renderer=1, numerics=0)
    assert np.array_equal(got, ref["accum"]), f"{int((got != ref['accum']).any(-1).sum())} pixels differ"
    assert sum(s["ray"] for s in stats) == ref["rays"]
    queue = np.sum([s["queue"] for s in prof], axis=0)
    assert queue.sum() == ref["rays"]
    assert queue[0] == spp * W * H
    assert queue[15] > 0, "no ray reached depth 15"


def test_wavefront_stops_when_no_path_is_left(prt, orc, scenes):
    """An open scene at max depth 100: every path escapes within a few
    bounces, and the wavefront frame ends at the first 16-depth check."""
    mesh = scenes.scene_convex_sphere(2)
    cam = orc.camera_from_view([0.0, 0.0, 2.5], 0.0, 0.0, math.radians(50.0), W / H)
    with prt.HipDevice(0) as d:
        d.upload_scene(mesh.idx, mesh.pos)
        d.build()
        d.set_camera(cam)
        d.init_frame(W, H)
        out = {}
        for pipeline in ("wavefront", "fused"):
            d.set_pipeline(pipeline)
            d.clear_frame()
            d.set_profiling(timing=True)
            st = d.render("diffuseHip", max_depth=100, albedo=0.8)
            out[pipeline] = (d.accum(), st)
    ref = orc.Scene(mesh.idx, mesh.pos).render(cam, W, H, spp=1, max_depth=100, albedo=0.8, renderer=1, numerics=0)
    for pipeline, (img, st) in out.items():
        assert np.array_equal(img, ref["accum"]), pipeline
        assert st["ray"] == ref["rays"], pipeline
    assert out["wavefront"][1]["traceMs"] > 0 and out["wavefront"][1]["shadeMs"] > 0


def test_batch_depth_32_equals_plain(prt, orc, tunnel):
    mesh, _ = tunnel
    cam = _camera(orc, W, H)
    job = {"scene": 0, "width": W, "height": H, "spp": 3, "max_depth": 32, "pass0": 5, "albedo": 0.9,
           "camera": cam}
    with prt.HipDevice(0) as d:
        d.batch_add_scene(mesh.idx, mesh.pos)
        d.batch_set_jobs([job])
        st = d.batch_render()
        got = d.batch_accum(0)
    with prt.HipDevice(0) as d:
        d.upload_scene(mesh.idx, mesh.pos)
        d.build()
        d.set_camera(cam)
        d.init_frame(W, H)
        d.set_pass(5)
        rays = sum(d.render("diffuseHip", max_depth=32, albedo=0.9)["ray"] for _ in range(3))
        ref = d.accum()
    assert np.array_equal(got, ref)
    assert st["ray"] == rays


def test_depth_limits(prt, orc, scenes):
    mesh = scenes.scene_convex_sphere(1)
    cam = orc.camera_from_view([0.0, 0.0, 2.5], 0.0, 0.0, math.radians(50.0), 1.0)
    with prt.HipDevice(0) as d:
        d.upload_scene(mesh.idx, mesh.pos)
        d.build()
        d.set_camera(cam)
        d.init_frame(8, 8)
        d.render("diffuseHip", max_depth=1024, albedo=0.8)
        for bad in (-1, 1025):
            with pytest.raises(RuntimeError, match="max_depth"):
                d.render("diffuseHip", max_depth=bad)
