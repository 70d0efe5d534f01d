"""GPU parity of the thin-lens camera (ThinLensCamera::getRay,
camera/thin_lens_camera.h:38-44): primary rays and sampler states
bit-identical to the oracle, and diffuse / diffuse2 renders bit-identical on
both frame pipelines and in the batched form."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cam(orc, scenes, name, w, h, radius, focus):
    v = scenes.CONFIGS[name]["view"]
    return orc.camera20(orc.camera_from_view(v["pos"], v["angle_x"], v["angle_y"], v["fov_y"], w / h), radius, focus)


def test_thin_lens_raygen_bit_exact(dev, orc, scenes):
    for (w, h, p, r, f) in [(96, 64, 0, 0.05, 2.0), (100, 60, 7, 0.2, 0.7)]:
        cam = _cam(orc, scenes, "heightfield_box_512", w, h, r, f)
        dev.set_camera(cam)
        dev.init_frame(w, h)
        rays, states = dev.debug_raygen(p)
        ref_rays, ref_states = orc.raygen(cam, w, h, p)
        assert np.array_equal(rays.view(np.uint32), ref_rays.view(np.uint32))
        assert np.array_equal(states, ref_states)


@pytest.mark.parametrize("pipeline,renderer", [("fused", "diffuseHip"), ("wavefront", "diffuseHip"),
                                               ("fused", "diffuse2Hip")])
def test_thin_lens_render_bit_exact(dev, orc, scenes, pipeline, renderer):
    mesh = scenes.build_scene("spheres_1080p")
    dev.upload_scene(mesh.idx, mesh.pos)
    dev.build()
    w, h = 240, 136
    cam = _cam(orc, scenes, "spheres_1080p", w, h, 0.03, 2.2)
    dev.set_camera(cam)
    dev.init_frame(w, h)
    dev.set_pipeline(pipeline)
    try:
        rays = sum(dev.render(renderer, max_depth=8, albedo=0.7)["ray"] for _ in range(2))
    finally:
        dev.set_pipeline("fused")
    kind = {"diffuseHip": 1, "diffuse2Hip": 2}[renderer]
    ref = orc.Scene(mesh.idx, mesh.pos).render(cam, w, h, spp=2, max_depth=8, albedo=0.7, renderer=kind, numerics=0)
    assert np.array_equal(dev.accum(), ref["accum"])
    assert rays == ref["rays"]


def test_thin_lens_batch_equals_plain(prt, orc, scenes):
    mesh = scenes.scene_heightfield_box(1001, cells=16)
    cam = _cam(orc, scenes, "heightfield_box_512", 64, 48, 0.1, 3.0)
    job = {"scene": 0, "width": 64, "height": 48, "spp": 3, "max_depth": 4, "pass0": 2, "albedo": 0.8, "camera": cam}
    with prt.HipDevice(0) as dev:
        dev.batch_add_scene(mesh.idx, mesh.pos)
        dev.batch_set_jobs([job])
        dev.batch_render()
        got = dev.batch_accum(0)
    ref = orc.Scene(mesh.idx, mesh.pos).render(cam, 64, 48, spp=3, max_depth=4, albedo=0.8, renderer=1, numerics=0,
                                               pass0=2)
    assert np.array_equal(got, ref["accum"])
