"""GPU parity of the any-hit query and of diffuse2Hip (next-event shadow rays).

Bars:
  - occluded: the occlusion answer (distance >= 0 or -1) identical to the
    kernel-numerics oracle for every ray; a reported distance lies in
    [tnear, tfar];
  - diffuse2Hip renders: accumulation buffers bit-identical to the oracle's
    Diffuse2IntegratorSingle restatement in kernel numerics, same ray count
    (extension + shadow rays); known answers on a convex mesh;
  - the wavefront pipeline refuses diffuse2 with an error.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
FLT_MAX = np.float32(3.402823466e38)


def _cam(orc, scenes, name, w, h):
    v = scenes.CONFIGS[name]["view"]
    return orc.camera_from_view(v["pos"], v["angle_x"], v["angle_y"], v["fov_y"], w / h)


def _load(dev, mesh):
    dev.upload_scene(mesh.idx, mesh.pos)
    return dev.build()


def test_occluded_parity(dev, orc, scenes):
    mesh = scenes.scene_heightfield_box(1001)
    _load(dev, mesh)
    sc = orc.Scene(mesh.idx, mesh.pos)
    rng = np.random.default_rng(31)
    n = 200_000
    lo, hi = mesh.pos.min(0), mesh.pos.max(0)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = rng.uniform(lo, hi, size=(n, 3))
    rays[:, 4:7] = d
    rays[:, 3] = rng.uniform(0.0, 0.05, n)
    rays[:, 7] = np.where(rng.uniform(size=n) < 0.5, FLT_MAX, rng.uniform(0.01, 1.0, n)).astype(np.float32)
    rays[::101, 7] = -1.0
    got = dev.occluded(rays)
    ref = sc.occluded_fp32(rays)
    assert np.array_equal(got >= 0, ref >= 0), f"{int(((got >= 0) != (ref >= 0)).sum())} occlusion answers differ"
    occ = got >= 0
    assert (got[occ] >= rays[occ, 3]).all() and (got[occ] <= rays[occ, 7]).all()
    assert (got[::101] == -1).all()
    assert 0.2 < occ.mean() < 0.95
    # a blocker exists exactly where the closest-hit query finds a hit
    hits = dev.intersect(rays)
    assert np.array_equal(occ, hits[:, 1].view(np.int32) >= 0)


def test_occluded_watertight_icosphere(dev, orc, scenes):
    mesh = scenes.scene_convex_sphere(3)
    _load(dev, mesh)
    pos = mesh.pos.astype(np.float64)
    tgt = np.concatenate([pos] + [0.5 * (pos[mesh.idx[:, a]] + pos[mesh.idx[:, b]]) for a, b in [(0, 1), (1, 2), (2, 0)]])
    g = np.random.default_rng(6).normal(size=tgt.shape)
    nrm = tgt / np.linalg.norm(tgt, axis=1, keepdims=True)
    d = -nrm + 0.6 * g / np.linalg.norm(g, axis=1, keepdims=True)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((len(tgt), 8), np.float32)
    rays[:, 0:3] = tgt - 3 * d
    rays[:, 4:7] = d
    rays[:, 7] = FLT_MAX
    assert (dev.occluded(rays) >= 0).all(), "a shadow ray through a shared edge or vertex passed"


def test_render_diffuse2_bit_exact(dev, orc, scenes):
    mesh = scenes.scene_heightfield_box(1001)
    _load(dev, mesh)
    w, h = 160, 100   # partial 8x8 tiles in y
    cam = _cam(orc, scenes, "heightfield_box_512", w, h)
    dev.set_camera(cam)
    dev.init_frame(w, h)
    rays = 0
    for _ in range(3):
        rays += dev.render("diffuse2Hip", max_depth=4, albedo=0.8)["ray"]
    ref = orc.Scene(mesh.idx, mesh.pos).render(cam, w, h, spp=3, max_depth=4, albedo=0.8, renderer=2, numerics=0)
    got = dev.accum()
    assert np.array_equal(got, ref["accum"]), f"{int((got != ref['accum']).any(-1).sum())} pixels differ"
    assert rays == ref["rays"], "ray counts differ from the oracle"


def test_render_diffuse2_spheres_bit_exact(dev, orc, scenes):
    mesh = scenes.build_scene("spheres_1080p")
    _load(dev, mesh)
    w, h = 240, 136
    cam = _cam(orc, scenes, "spheres_1080p", w, h)
    dev.set_camera(cam)
    dev.init_frame(w, h)
    st = dev.render("diffuse2Hip", max_depth=8, albedo=0.7)
    ref = orc.Scene(mesh.idx, mesh.pos).render(cam, w, h, spp=1, max_depth=8, albedo=0.7, renderer=2, numerics=0)
    assert np.array_equal(dev.accum(), ref["accum"])
    assert st["ray"] == ref["rays"]


def test_diffuse2_known_answers(dev, orc, scenes):
    sph = scenes.scene_convex_sphere(2)
    _load(dev, sph)
    cam = orc.camera_from_view([0, 0, 3], 0, 0, math.radians(45), 1.0)
    dev.set_camera(cam)
    dev.init_frame(48, 48)
    st = dev.render("diffuse2Hip", max_depth=3, albedo=0.8)
    vals = dev.accum()[:, :, 0]
    assert set(np.unique(vals).tolist()) == {np.float32(0.8), 1.0}
    hit = vals != 1.0
    assert st["ray"] == int(hit.sum()) * 3 + int((~hit).sum())


def test_diffuse2_needs_fused_pipeline(dev, orc, scenes):
    sph = scenes.scene_convex_sphere(1)
    _load(dev, sph)
    dev.set_camera(orc.camera_from_view([0, 0, 3], 0, 0, math.radians(45), 1.0))
    dev.init_frame(16, 16)
    dev.set_pipeline("wavefront")
    try:
        with pytest.raises(RuntimeError, match="fused"):
            dev.render("diffuse2Hip", max_depth=2)
    finally:
        dev.set_pipeline("fused")
