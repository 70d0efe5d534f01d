"""Parity at the benchmark workloads' full sizes.

- spheres_1080p (1,310,722 triangles, 1920x1080, depth 8): two full frames
  bit-identical to the oracle's single-ray renderer in kernel numerics, same
  ray count; the same two passes in one batched launch equal them too.
- city (15,960,000 triangles, the 1.33 M-box workload): the BVH8 holds every
  triangle once with its own vertices; 200,000 random rays give bit-identical
  closest-hit records to the oracle (overlapping boxes produce coincident
  faces, so this exercises the (t, primId) tie rule at scale); a 960x540
  frame at depth 8 is bit-identical to the oracle's; a 1080p chain with the
  frame lookahead equals the chain without it.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
FLT_MAX = np.float32(3.402823466e38)


def _cam(orc, scenes, name, w, h):
    v = scenes.CONFIGS[name]["view"]
    return orc.camera_from_view(v["pos"], v["angle_x"], v["angle_y"], v["fov_y"], w / h)


def test_spheres_full_frames_bit_exact(prt, dev, orc, scenes):
    cfg = scenes.CONFIGS["spheres_1080p"]
    mesh = scenes.build_scene("spheres_1080p")
    w, h, depth, albedo = cfg["width"], cfg["height"], cfg["max_depth"], cfg["albedo"]
    dev.upload_scene(mesh.idx, mesh.pos)
    dev.build()
    cam = _cam(orc, scenes, "spheres_1080p", w, h)
    dev.set_camera(cam)
    dev.init_frame(w, h)
    rays = sum(dev.render("diffuseHip", max_depth=depth, albedo=albedo)["ray"] for _ in range(2))
    got = dev.accum()
    ref = orc.Scene(mesh.idx, mesh.pos).render(cam, w, h, spp=2, max_depth=depth, albedo=albedo, renderer=1,
                                               numerics=0)
    assert rays == ref["rays"]
    assert np.array_equal(got, ref["accum"]), f"{int((got != ref['accum']).any(-1).sum())} pixels differ"
    with prt.HipDevice(0) as bdev:
        bdev.batch_add_scene(mesh.idx, mesh.pos)
        bdev.batch_set_jobs([{"scene": 0, "width": w, "height": h, "spp": 2, "max_depth": depth, "pass0": 0,
                              "albedo": albedo, "camera": cam}])
        st = bdev.batch_render()
        assert st["ray"] == rays
        assert np.array_equal(bdev.batch_accum(0), got)


@pytest.fixture(scope="module")
def city(scenes, orc):
    mesh = scenes.build_scene("city_4k")
    return mesh, orc.Scene(mesh.idx, mesh.pos)


def test_city_bvh_records(dev, city):
    mesh, _ = city
    dev.upload_scene(mesh.idx, mesh.pos)
    st = dev.build()
    assert st["tris"] == mesh.tri_count == 15_960_000
    _, tris, prims, depth = dev.bvh()
    assert depth >= 1
    assert np.array_equal(np.sort(prims), np.arange(mesh.tri_count))
    rng = np.random.default_rng(3)
    pick = rng.choice(mesh.tri_count, size=200_000, replace=False)
    assert np.array_equal(tris[pick], mesh.pos[mesh.idx[prims[pick]]].reshape(-1, 9))


def test_city_intersect_and_frame_bit_exact(dev, orc, scenes, city):
    mesh, sc = city
    dev.upload_scene(mesh.idx, mesh.pos)
    dev.build()
    rng = np.random.default_rng(17)
    n = 200_000
    lo, hi = mesh.pos.min(0), mesh.pos.max(0)
    d = rng.normal(size=(n, 3))
    d[:, 1] = -np.abs(d[:, 1])   # mostly downwards, into the towers
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = rng.uniform(lo, [hi[0], 3.0, hi[2]], size=(n, 3))
    rays[:, 4:7] = d
    rays[:, 7] = FLT_MAX
    hits = dev.intersect(rays)
    ref = sc.intersect_fp32(rays)
    assert np.array_equal(hits.view(np.uint32), ref.view(np.uint32)), \
        f"{int((hits.view(np.uint32) != ref.view(np.uint32)).any(1).sum())} hit records differ"
    assert (hits[:, 1].view(np.int32) >= 0).mean() > 0.5
    cfg = scenes.CONFIGS["city_4k"]
    w, h = 960, 540
    cam = _cam(orc, scenes, "city_4k", w, h)
    dev.set_camera(cam)
    dev.init_frame(w, h)
    st = dev.render("diffuseHip", max_depth=cfg["max_depth"], albedo=cfg["albedo"])
    r = sc.render(cam, w, h, spp=1, max_depth=cfg["max_depth"], albedo=cfg["albedo"], renderer=1, numerics=0)
    assert st["ray"] == r["rays"]
    assert np.array_equal(dev.accum(), r["accum"])


def test_city_lookahead_chain_bit_exact(dev, orc, scenes, city):
    """The city's nodes and triangles exceed 256 MiB, so its diffuse launches
    run one workgroup per CU fewer than the spheres': a chain of 4 calls at
    1920x1080 with the frame lookahead equals the same calls without it, call
    by call (accumulator and ray count)."""
    mesh, _ = city
    dev.upload_scene(mesh.idx, mesh.pos)
    dev.build()
    cfg = scenes.CONFIGS["city_4k"]
    w, h = 1920, 1080
    cam = _cam(orc, scenes, "city_4k", w, h)
    out = {}
    try:
        for look in (True, False):
            dev.set_lookahead(look)
            dev.set_camera(cam)
            dev.init_frame(w, h)
            out[look] = []
            for _ in range(4):
                st = dev.render("diffuseHip", max_depth=cfg["max_depth"], albedo=cfg["albedo"])
                out[look].append((dev.accum(), st["ray"]))
    finally:
        dev.set_lookahead(True)
    for k, ((a, ra), (b, rb)) in enumerate(zip(out[True], out[False])):
        assert ra == rb, f"call {k}: ray count {ra} with lookahead, {rb} without"
        assert np.array_equal(a, b), f"call {k}: {int((a != b).any(-1).sum())} pixels differ"
