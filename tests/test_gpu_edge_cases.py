"""Edge cases of the C ABI on the GPU: empty and degenerate inputs, tiny
scenes whose BVH8 is one node, rays that cannot hit, the smallest frames, and
the error paths (each call fails with a message and the context stays usable).
Hit records are compared bit for bit with the oracle's brute force (every
triangle tested), renders with the oracle's single-ray renderer.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
FLT_MAX = np.float32(3.402823466e38)


def _rays(org, d, tnear=0.0, tfar=FLT_MAX):
    org = np.atleast_2d(np.asarray(org, np.float32))
    d = np.atleast_2d(np.asarray(d, np.float64))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    r = np.zeros((max(len(org), len(d)), 8), np.float32)
    r[:, 0:3], r[:, 4:7] = org, d
    r[:, 3], r[:, 7] = tnear, tfar
    return r


def _fan_rays(n, seed, centre=(0.3, 0.3, 0.0)):
    rng = np.random.default_rng(seed)
    org = np.tile([0.25, 0.25, 2.0], (n, 1)) + rng.normal(scale=0.5, size=(n, 3))
    tgt = np.asarray(centre) + rng.normal(scale=0.6, size=(n, 3))
    return _rays(org, tgt - org)


def test_zero_rays_and_unhittable_rays(dev, orc, scenes):
    mesh = scenes.scene_heightfield_box(1001, cells=8)
    dev.upload_scene(mesh.idx, mesh.pos)
    dev.build()
    assert dev.intersect(np.zeros((0, 8), np.float32)).shape == (0, 4)
    assert dev.occluded(np.zeros((0, 8), np.float32)).shape == (0,)
    down = _rays([[0.1, 1.5, 0.2]] * 4, [[0.0, -1.0, 0.0]] * 4)
    down[0, 7] = -1.0                    # tfar < 0: no query
    down[1, 3], down[1, 7] = 5.0, 4.0    # tnear > tfar
    down[2, 7] = 1e-3                    # ends before the floor
    hits = dev.intersect(down)
    sc = orc.Scene(mesh.idx, mesh.pos)
    ref = sc.intersect_fp32_brute(down)
    assert np.array_equal(hits.view(np.uint32), ref.view(np.uint32))
    assert (hits[:3, 1].view(np.int32) == -1).all() and (hits[:3, 0] == -1.0).all()
    assert hits[3, 1].view(np.int32) >= 0
    occ = dev.occluded(down)
    assert (occ[:3] == -1.0).all() and occ[3] >= 0.0


@pytest.mark.parametrize("ntris", [1, 2, 3, 4])
def test_tiny_scenes(dev, orc, ntris):
    """One to four triangles: a single BVH8 node, leaf slots only."""
    rng = np.random.default_rng(ntris)
    pos = rng.uniform(-1, 1, size=(3 * ntris, 3)).astype(np.float32)
    pos[:, 2] *= 0.2
    idx = np.arange(3 * ntris, dtype=np.int32).reshape(-1, 3)
    dev.upload_scene(idx, pos)
    st = dev.build()
    assert st["tris"] == ntris and st["nodes"] == 1
    rays = _fan_rays(4000, 10 + ntris, centre=pos.mean(0))
    sc = orc.Scene(idx, pos)
    hits = dev.intersect(rays)
    assert np.array_equal(hits.view(np.uint32), sc.intersect_fp32_brute(rays).view(np.uint32))
    assert (hits[:, 1].view(np.int32) >= 0).any()
    cam = orc.camera_from_view([0.0, 0.0, 3.0], 0.0, 0.0, 0.9, 1.0)
    dev.set_camera(cam)
    dev.init_frame(24, 24)
    dev.render("diffuseHip", max_depth=3, albedo=0.6)
    ref = sc.render(cam, 24, 24, spp=1, max_depth=3, albedo=0.6, renderer=1, numerics=0)
    assert np.array_equal(dev.accum(), ref["accum"])


def test_degenerate_triangles(dev, orc, scenes):
    """Zero-area triangles (a repeated vertex, collinear vertices, a point)
    among ordinary ones are never hit and do not disturb the others."""
    mesh = scenes.scene_heightfield_box(1001, cells=8)
    pos = mesh.pos.copy()
    extra_p = np.array([[0.1, 0.5, 0.1], [0.4, 0.5, 0.1], [0.1, 0.5, 0.1],      # repeated vertex
                        [0.0, 0.6, 0.0], [0.2, 0.6, 0.2], [0.4, 0.6, 0.4],      # collinear
                        [0.3, 0.7, 0.3], [0.3, 0.7, 0.3], [0.3, 0.7, 0.3]],     # a point
                       np.float32)
    base = len(pos)
    pos = np.concatenate([pos, extra_p])
    idx = np.concatenate([mesh.idx, base + np.arange(9, dtype=np.int32).reshape(3, 3)])
    dev.upload_scene(idx, pos)
    dev.build()
    rng = np.random.default_rng(8)
    org = rng.uniform([-1, 0.3, -1], [1, 1.9, 1], size=(20000, 3))
    rays = _rays(org, rng.normal(size=(20000, 3)))
    hits = dev.intersect(rays)
    sc = orc.Scene(idx, pos)
    assert np.array_equal(hits.view(np.uint32), sc.intersect_fp32_brute(rays).view(np.uint32))
    assert not np.isin(hits[:, 1].view(np.int32), len(mesh.idx) + np.arange(3)).any()


def test_one_pixel_and_thin_frames(dev, orc, scenes):
    mesh = scenes.scene_heightfield_box(1001, cells=16)
    dev.upload_scene(mesh.idx, mesh.pos)
    dev.build()
    sc = orc.Scene(mesh.idx, mesh.pos)
    v = scenes.CONFIGS["heightfield_box_512"]["view"]
    for (w, h) in [(1, 1), (1, 37), (53, 1), (9, 9)]:
        cam = orc.camera_from_view(v["pos"], v["angle_x"], v["angle_y"], v["fov_y"], w / h)
        dev.set_camera(cam)
        dev.init_frame(w, h)
        for _ in range(3):
            dev.render("diffuseHip", max_depth=4, albedo=0.8)
        ref = sc.render(cam, w, h, spp=3, max_depth=4, albedo=0.8, renderer=1, numerics=0)
        assert np.array_equal(dev.accum(), ref["accum"]), (w, h)


def test_error_paths_leave_the_context_usable(prt, orc, scenes):
    mesh = scenes.scene_heightfield_box(1001, cells=4)
    with prt.HipDevice(0) as d:
        with pytest.raises(RuntimeError, match="no BVH"):
            d.intersect(np.zeros((4, 8), np.float32))
        with pytest.raises(RuntimeError, match="out of range"):
            d.upload_scene(np.array([[0, 1, 99]], np.int32), mesh.pos)
        d.upload_scene(np.zeros((0, 3), np.int32), mesh.pos)
        with pytest.raises(RuntimeError, match="no triangles"):
            d.build()
        d.upload_scene(mesh.idx, mesh.pos)
        d.build()
        with pytest.raises(RuntimeError, match="no camera"):
            d.render("diffuseHip")
        cam = orc.camera_from_view([0.0, 1.0, 3.0], 0.3, 0.0, 0.8, 1.0)
        d.set_camera(cam)
        with pytest.raises(RuntimeError, match="no frame"):
            d.render("diffuseHip")
        d.init_frame(16, 16)
        with pytest.raises(RuntimeError, match="max_depth"):
            d.render("diffuseHip", max_depth=1025)
        with pytest.raises(RuntimeError, match="invalid renderer"):
            d.lib.prt_hip_render(d.ctx, 7, 4, 0.8, None) == 0 or d._check(1, "render")
        with pytest.raises(ValueError):
            d.render("noSuchRenderer")
        with pytest.raises(RuntimeError, match="set_bvh_collapse"):
            d.lib.prt_hip_set_bvh_collapse(d.ctx, 7, 1.0) == 0 or d._check(1, "set_bvh_collapse")
        with pytest.raises(RuntimeError, match="set_bvh_collapse"):
            d.lib.prt_hip_set_bvh_collapse(d.ctx, 1, -1.0) == 0 or d._check(1, "set_bvh_collapse")
        with pytest.raises(RuntimeError, match="set_tail"):
            d.lib.prt_hip_set_tail(d.ctx, 9) == 0 or d._check(1, "set_tail")
        with pytest.raises(ValueError):
            d.set_tail("sideways")
        # still usable after every failure
        d.render("diffuseHip", max_depth=4, albedo=0.8)
        ref = orc.Scene(mesh.idx, mesh.pos).render(cam, 16, 16, spp=1, max_depth=4, albedo=0.8, renderer=1,
                                                   numerics=0)
        assert np.array_equal(d.accum(), ref["accum"])


@pytest.mark.parametrize("scale", [1e-3, 1.0, 1e4])
def test_axis_parallel_and_grazing_rays(dev, orc, scenes, scale):
    """Rays along the axes, rays with one zero direction component (the
    traversal clamps that reciprocal to 1e20) and rays that run inside box
    faces or along box edges, over 27 boxes on an integer grid at three
    scales (the node box exponents change with the scale).  The child boxes
    are 8-bit quantized and tested with a rounding slack; every hit record
    must equal the oracle's brute force (every triangle tested) bit for bit,
    so a box test that rejects a child it should keep shows up as a miss or
    a farther hit."""
    parts = []
    for i in range(3):
        for j in range(3):
            for k in range(3):
                lo = np.array([2 * i, 2 * j, 2 * k], np.float64) * scale
                parts.append(scenes.box_tris(tuple(lo), tuple(lo + scale)))
    mesh = scenes.merge(parts)
    dev.upload_scene(mesh.idx, mesh.pos)
    dev.build()
    # origins on the grid of face planes (0, 0.5, 1, ... 5 units), outside
    # and inside the boxes, and directions along the axes or with one zero
    vals = np.arange(-1.0, 6.0, 0.5)
    g = np.stack(np.meshgrid(vals, vals, [-1.0, 0.0, 2.5], indexing="ij"), -1).reshape(-1, 3)
    dirs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
            (1, 1, 0), (1, -1, 0), (0, 1, 1), (0, 1, -1), (1, 0, 1), (-1, 0, 1), (1, 2, 3)]
    org, d = [], []
    for dv in dirs:
        for o in g:
            # permute the origin's axes so that every axis sees the grid values
            for perm in ((0, 1, 2), (2, 0, 1), (1, 2, 0)):
                org.append(o[list(perm)])
                d.append(dv)
    org = np.asarray(org) * scale
    rays = _rays(org, np.asarray(d, np.float64))
    hits = dev.intersect(rays)
    ref = orc.Scene(mesh.idx, mesh.pos).intersect_fp32_brute(rays)
    bad = np.flatnonzero((hits.view(np.uint32) != ref.view(np.uint32)).any(1))
    assert bad.size == 0, f"{bad.size} of {len(rays)} rays differ, first {rays[bad[:3]]}"
    assert (hits[:, 1].view(np.int32) >= 0).sum() > len(rays) // 10
