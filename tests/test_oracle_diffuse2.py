"""Oracle: the any-hit query and the diffuse2 renderer (next-event shadow rays).

Pins, each independent of the oracle's own formulas:
  - any hit over the BVH == any hit by testing every triangle, and a blocker
    exists exactly when the closest-hit query over the same [tnear, tfar] finds
    a hit (both use the same watertight triangle test, different searches);
  - the occlusion answer agrees with the FP64 intersector except for rays
    within 1e-5 of a triangle edge;
  - known answers of Diffuse2IntegratorSingle (render/diffuse2_integrator_single.h:42-87):
    a convex mesh under a white sky gives exactly R on hit pixels (shadow ray
    0.5 R + escaping extension 0.5 R) and 1 on misses; maxDepth 0 gives 0 / 1;
  - diffuse and diffuse2 are unbiased estimators of the same image (the MIS
    weights 1/2 + 1/2 split the sky term between the shadow ray and the
    extension ray, both cosine-sampled), so their means agree within Monte
    Carlo noise.
"""
import math

import numpy as np
import pytest


def _soup(rng, T=300):
    c = rng.uniform(-1, 1, size=(T, 1, 3))
    tris = (c + 0.15 * rng.normal(size=(T, 3, 3))).astype(np.float32)
    return np.arange(3 * T, dtype=np.int32).reshape(T, 3), tris.reshape(-1, 3)


def _rays(rng, n, tnear=None, tfar=None):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = rng.uniform(-2, 2, size=(n, 3))
    r[:, 4:7] = d
    r[:, 3] = 0.0 if tnear is None else tnear
    r[:, 7] = np.float32(3.4e38) if tfar is None else tfar
    return r


def test_occluded_bvh_equals_brute_and_closest(orc):
    rng = np.random.default_rng(21)
    idx, pos = _soup(rng)
    sc = orc.Scene(idx, pos)
    n = 4000
    rays = _rays(rng, n, tnear=rng.uniform(0, 0.5, n).astype(np.float32),
                 tfar=rng.uniform(0.2, 4.0, n).astype(np.float32))
    rays[::97, 7] = -1.0  # inactive rays
    a = sc.occluded_fp32(rays)
    b = sc.occluded_fp32(rays, brute=True)
    assert np.array_equal(a >= 0, b >= 0)
    closest = sc.intersect_fp32_brute(rays)
    has_hit = closest[:, 1].view(np.int32) >= 0
    assert np.array_equal(a >= 0, has_hit)
    assert (a[::97] == -1).all()
    assert 0.05 < has_hit.mean() < 0.9
    # a reported blocker lies in [tnear, tfar]
    occ = a >= 0
    assert (a[occ] >= rays[occ, 3]).all() and (a[occ] <= rays[occ, 7]).all()


def test_occluded_agrees_with_fp64_away_from_edges(orc):
    rng = np.random.default_rng(22)
    idx, pos = _soup(rng)
    sc = orc.Scene(idx, pos)
    rays = _rays(rng, 3000)
    occ = sc.occluded_fp32(rays) >= 0
    r64 = sc.intersect_fp64(rays)
    diff = occ != (r64["prim"] >= 0)
    assert (r64["edge"][diff] < 1e-5).all()


def test_occluded_watertight_on_shared_edges(orc, scenes):
    mesh = scenes.scene_convex_sphere(3)
    sc = orc.Scene(mesh.idx, mesh.pos)
    pos = mesh.pos.astype(np.float64)
    tgt = np.concatenate([pos] + [0.5 * (pos[mesh.idx[:, a]] + pos[mesh.idx[:, b]]) for a, b in [(0, 1), (1, 2), (2, 0)]])
    g = np.random.default_rng(5).normal(size=tgt.shape)
    nrm = tgt / np.linalg.norm(tgt, axis=1, keepdims=True)
    d = -nrm + 0.6 * g / np.linalg.norm(g, axis=1, keepdims=True)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((len(tgt), 8), np.float32)
    rays[:, 0:3] = tgt - 3 * d
    rays[:, 4:7] = d
    rays[:, 7] = np.float32(3.4e38)
    assert (sc.occluded_fp32(rays) >= 0).all()
    # stopping short of the sphere: nothing blocks
    rays[:, 7] = 1.5
    assert (sc.occluded_fp32(rays) < 0).all()


@pytest.mark.parametrize("numerics", [0, 1])
def test_diffuse2_known_answers(orc, scenes, numerics):
    w = h = 48
    sph = scenes.scene_convex_sphere(2)
    sc = orc.Scene(sph.idx, sph.pos)
    cam = orc.camera_from_view([0, 0, 3], 0, 0, math.radians(45), 1.0)
    r = sc.render(cam, w, h, spp=1, max_depth=3, albedo=0.8, renderer=2, numerics=numerics)
    vals = r["accum"][:, :, 0]
    assert set(np.unique(vals).tolist()) == {np.float32(0.8), 1.0}
    hit = vals != 1.0
    # hit pixel: camera ray + shadow ray + escaping extension ray; miss: 1 ray
    assert r["rays"] == int(hit.sum()) * 3 + int((~hit).sum())
    r0 = sc.render(cam, w, h, spp=1, max_depth=0, albedo=0.8, renderer=2, numerics=numerics)
    assert set(np.unique(r0["accum"][:, :, 0]).tolist()) == {0.0, 1.0}
    box = scenes.scene_closed_box()
    sb = orc.Scene(box.idx, box.pos)
    cam2 = orc.camera_from_view([0, 0, 0.5], 0.3, 0.2, math.radians(60), 1.0)
    rb = sb.render(cam2, w, h, spp=2, max_depth=5, albedo=0.8, renderer=2, numerics=numerics)
    # nothing reaches the sky from inside a closed box, up to the rare
    # self-intersection leak of the reference's ray offset
    assert (rb["accum"][:, :, 0] == 0).mean() > 0.995
    assert rb["rays"] >= 0.99 * 2 * w * h * 11  # 6 extension + 5 shadow rays per path


def test_diffuse2_and_diffuse_estimate_the_same_image(orc, scenes):
    cfg = scenes.CONFIGS["heightfield_box_512"]
    mesh = scenes.build_scene("heightfield_box_512")
    sc = orc.Scene(mesh.idx, mesh.pos)
    w = h = 32
    v = cfg["view"]
    cam = orc.camera_from_view(v["pos"], v["angle_x"], v["angle_y"], v["fov_y"], 1.0)
    spp = 256
    stats = []
    for renderer in (1, 2):
        r = sc.render(cam, w, h, spp=spp, max_depth=4, albedo=0.8, renderer=renderer, numerics=1)
        s = r["accum"][:, :, 0].astype(np.float64)
        q = r["sumsq"].astype(np.float64)
        mean = s / spp
        var = np.maximum(q / spp - mean ** 2, 0.0) / spp   # variance of the pixel mean
        stats.append((mean, var))
    (m1, v1), (m2, v2) = stats
    z_img = (m1.mean() - m2.mean()) / math.sqrt((v1.sum() + v2.sum()) / (w * h) ** 2)
    assert abs(z_img) < 5.0
    z_pix = (m1 - m2) / np.sqrt(v1 + v2 + 1e-30)
    assert (np.abs(z_pix) > 3.0).mean() < 0.02
    # next-event estimation lowers the noise of this open scene
    assert v2.sum() < v1.sum()
