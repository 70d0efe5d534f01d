"""GPU parity of the full shading context (per-vertex normals and texture
coordinates, geometry/triangle_mesh_cuda.cuh:41-130) and of the "Fast"
renderers (render/device_cuda.cpp:79-90): every image bit-identical to the
oracle's renderer in kernel numerics with the same attributes, on both frame
pipelines, in the batched form, through the RGB-albedo forward and adjoint,
and through `protoray render` of a .mesh that carries the attributes.
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "embree-benchmark-protoray_amd", "protoray")


@pytest.fixture(scope="module")
def smooth(scenes, orc):
    m = scenes.scene_smooth()
    s = orc.Scene(m.idx, m.pos)
    s.set_attributes(m.normals, m.texcoords)
    return m, s


def _cam(orc, scenes, w, h):
    v = scenes.SMOOTH_VIEW
    return orc.camera_from_view(v["pos"], v["angle_x"], v["angle_y"], v["fov_y"], w / h)


def _setup(dev, orc, scenes, m, w, h, attributes=True):
    dev.upload_scene(m.idx, m.pos)
    if attributes:
        dev.set_attributes(m.normals, m.texcoords)
    dev.build()
    cam = _cam(orc, scenes, w, h)
    dev.set_camera(cam)
    dev.init_frame(w, h)
    return cam


@pytest.mark.parametrize("pipeline,renderer", [("fused", "primaryHip"), ("fused", "diffuseHip"),
                                               ("fused", "diffuse2Hip"), ("wavefront", "primaryHip"),
                                               ("wavefront", "diffuseHip")])
def test_attribute_render_bit_exact(dev, orc, scenes, smooth, pipeline, renderer):
    m, sc = smooth
    w, h = 160, 120
    cam = _setup(dev, orc, scenes, m, w, h)
    dev.set_pipeline(pipeline)
    try:
        rays = sum(dev.render(renderer, max_depth=6, albedo=0.75)["ray"] for _ in range(2))
    finally:
        dev.set_pipeline("fused")
    kind = {"primaryHip": 0, "diffuseHip": 1, "diffuse2Hip": 2}[renderer]
    ref = sc.render(cam, w, h, spp=2, max_depth=6, albedo=0.75, renderer=kind, numerics=0)
    got = dev.accum()
    assert rays == ref["rays"]
    assert np.array_equal(got, ref["accum"]), f"{int((got != ref['accum']).any(-1).sum())} pixels differ"


@pytest.mark.parametrize("pipeline,renderer", [("fused", "primaryFastHip"), ("fused", "diffuseFastHip"),
                                               ("fused", "diffuse2FastHip"), ("wavefront", "diffuseFastHip")])
def test_fast_renderers(dev, orc, scenes, smooth, pipeline, renderer):
    """Attributes ignored, each pass overwrites: after passes 0..2 the
    accumulator holds pass 2 alone, as rendered on the bare mesh."""
    m, _ = smooth
    w, h = 96, 72
    cam = _setup(dev, orc, scenes, m, w, h)
    dev.set_pipeline(pipeline)
    try:
        for _ in range(3):
            dev.render(renderer, max_depth=5, albedo=0.8)
    finally:
        dev.set_pipeline("fused")
    kind = {"primaryFastHip": 0, "diffuseFastHip": 1, "diffuse2FastHip": 2}[renderer]
    ref = orc.Scene(m.idx, m.pos).render(cam, w, h, spp=1, max_depth=5, albedo=0.8, renderer=kind, numerics=0,
                                         pass0=2)
    got = dev.accum()
    assert np.array_equal(got, ref["accum"])
    assert (got[:, :, 3] == 1.0).all()


def test_upload_clears_attributes(dev, orc, scenes, smooth):
    m, _ = smooth
    w, h = 64, 48
    cam = _setup(dev, orc, scenes, m, w, h)
    cam = _setup(dev, orc, scenes, m, w, h, attributes=False)   # re-upload: no attributes
    dev.render("primaryHip")
    ref = orc.Scene(m.idx, m.pos).render(cam, w, h, spp=1, max_depth=0, albedo=0.8, renderer=0, numerics=0)
    assert np.array_equal(dev.accum(), ref["accum"])
    with pytest.raises(ValueError):
        dev.set_attributes(m.normals[:-1], None)


def test_batch_with_attributes_equals_plain(prt, orc, scenes, smooth):
    m, sc = smooth
    w, h = 80, 60
    cam = _cam(orc, scenes, w, h)
    bare = scenes.scene_heightfield_box(1001, cells=16)
    vb = scenes.CONFIGS["heightfield_box_512"]["view"]
    cam_b = orc.camera_from_view(vb["pos"], vb["angle_x"], vb["angle_y"], vb["fov_y"], 48 / 40)
    jobs = [{"scene": 0, "width": w, "height": h, "spp": 3, "max_depth": 5, "pass0": 1, "albedo": 0.7, "camera": cam},
            {"scene": 1, "width": 48, "height": 40, "spp": 2, "max_depth": 4, "pass0": 0, "albedo": 0.8,
             "camera": cam_b}]
    with prt.HipDevice(0) as bdev:
        s0, _ = bdev.batch_add_scene(m.idx, m.pos)
        bdev.batch_set_attributes(s0, m.normals, m.texcoords)
        bdev.batch_add_scene(bare.idx, bare.pos)
        bdev.batch_set_jobs(jobs)
        bdev.batch_render()
        got0, got1 = bdev.batch_accum(0), bdev.batch_accum(1)
    ref0 = sc.render(cam, w, h, spp=3, max_depth=5, albedo=0.7, renderer=1, numerics=0, pass0=1)
    ref1 = orc.Scene(bare.idx, bare.pos).render(cam_b, 48, 40, spp=2, max_depth=4, albedo=0.8, renderer=1,
                                                numerics=0)
    assert np.array_equal(got0, ref0["accum"])
    assert np.array_equal(got1, ref1["accum"])


def test_albedo_forward_and_adjoint_with_attributes(dev, orc, scenes, smooth):
    """Tolerances as the attribute-free adjoint tests: forward 2e-6 relative
    per pixel, adjoint 1e-4 relative against the oracle's exact double
    adjoint over the same recorded paths."""
    m, sc = smooth
    w, h, spp, depth = 64, 48, 3, 5
    cam = _setup(dev, orc, scenes, m, w, h)
    a = scenes.albedo_set(m.tri_count, 7)
    dev.set_albedo(a)
    img, st = dev.render_albedo_image(pass0=2, spp=spp, max_depth=depth)
    tape = orc.PathTape(sc, cam, w, h, pass0=2, spp=spp, max_depth=depth)
    ref = tape.image(a.astype(np.float64))
    assert st["ray"] == tape.rays
    assert np.allclose(img, ref, rtol=2e-6, atol=1e-7)
    dldi = np.random.default_rng(3).normal(size=(h, w, 3)).astype(np.float32)
    grad, _ = dev.adjoint_albedo(pass0=2, spp=spp, max_depth=depth, dldi=dldi)
    gref = tape.adjoint(a.astype(np.float64), dldi.astype(np.float64))
    assert np.linalg.norm(grad - gref) <= 1e-4 * np.linalg.norm(gref)


def test_cli_renders_mesh_attributes(tmp_path, orc, scenes, smooth):
    m, sc = smooth
    scenes.save_mesh(str(tmp_path / "smooth.mesh"), m)
    v = scenes.SMOOTH_VIEW
    (tmp_path / "smooth.view").write_bytes(scenes.views_to_bytes([dict(v, radius=0.0, focus=1.0)]))
    r = subprocess.run([EXE, "render", "smooth.mesh", "-dev", "hip", "-r", "primary", "-size", "64,48", "-spp", "1",
                        "-o", "out"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    ppm = (tmp_path / "out.ppm").read_bytes()
    head = b"P6\n64 48\n255\n"
    got = np.frombuffer(ppm[len(head):], np.uint8).reshape(48, 64, 3)
    cam = orc.camera_from_view(v["pos"], v["angle_x"], v["angle_y"], v["fov_y"], 64 / 48)
    ref = sc.render(cam, 64, 48, spp=1, max_depth=0, albedo=0.8, renderer=0, numerics=0)["accum"]
    c = ref[:, :, :3] / ref[:, :, 3:]
    expect = np.clip((np.power(c.astype(np.float32), np.float32(1 / 2.2)) * 255).astype(np.int32), 0, 255)
    # the pixel bytes are B, G, R (image/image_io.cpp:55-76 writes the BGR8 surface as is)
    assert np.abs(got[:, :, ::-1].astype(int) - expect).max() <= 1
