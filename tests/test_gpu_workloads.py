"""Parity on the benchmark workloads at their stated sizes.

Bars (tolerances written here):
  - config 1 (heightfield in an open box, 8,202 triangles) at its full
    512x512 x 4 spp x depth 4: bit-identical to the kernel-numerics oracle on
    both frame pipelines, same ray count;
  - config 2 (spheres_1080p, 1.31 M triangles): all 2,073,600 primary rays of
    a 1080p pass against the FP64 intersector -- primId identical except rays
    within 1e-6 (times the scene extent) of an edge or with a second surface
    within 1e-6 relative depth, t within 1e-6 relative; and a 480x270 x 32 spp
    image against the reference-numerics oracle (FP64 hits, libm sin/cos,
    independent passes): under 1 % of pixels outside 3 sigma of the per-pixel
    Monte Carlo noise, |z| < 5 for the image mean;
  - config 3 (the 32-scene batched workload, 10.6 M triangles, 166.6 M paths)
    in one batched launch with the autotuned shape: every job's accumulator
    bit-identical to prt_hip_render on the same scene, camera and passes, and
    the first pass of an 8-row band of every job bit-identical to the
    oracle's single-ray renderer (kernel numerics);
  - config 5 (the 15.96 M-triangle city): 230,400 primary rays against FP64 as
    above; a 480x270 x 32 spp image within 3 sigma of the reference-numerics
    oracle; a 960x540 x 2 spp batched job equal to the plain render; the
    per-triangle albedo adjoint of a 256x144 x 4 spp frame against the
    oracle's exact FP64 adjoint over the same paths, 1e-4 relative (norm and
    every entry above 1e-3 of the largest).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
FLT_MAX = np.float32(3.402823466e38)


def _cam(orc, scenes, name, w, h):
    v = scenes.CONFIGS[name]["view"]
    return orc.camera_from_view(v["pos"], v["angle_x"], v["angle_y"], v["fov_y"], w / h)


def _extent(mesh):
    return float(np.linalg.norm(mesh.pos.max(0).astype(np.float64) - mesh.pos.min(0)))


def _check_fp64(sc, rays, hits, extent):
    """primId and t of GPU hits against the FP64 intersector; returns the
    number of (explained) primId differences."""
    ref = sc.intersect_fp64(rays)
    gp = hits[:, 1].view(np.int32)
    tol = 1e-6 * extent
    mism = gp != ref["prim"]
    if mism.any():
        idx = np.nonzero(mism)[0]
        near_ref = ref["edge"][idx] <= tol
        near_gpu = sc.ray_edge_distance(rays[idx], gp[idx]) <= tol
        t64 = ref["t"][idx]
        ambiguous = (ref["prim2"][idx] >= 0) & (np.abs(ref["t2"][idx] - t64) <= 1e-6 * np.abs(t64))
        unexplained = ~(near_ref | near_gpu | ambiguous)
        assert not unexplained.any(), (
            f"{int(unexplained.sum())} of {len(rays)} rays hit a different triangle than FP64 "
            f"away from any edge; first ray {idx[unexplained][0]}")
    both = (~mism) & (gp >= 0)
    rel = np.abs(hits[both, 0].astype(np.float64) - ref["t"][both]) / np.abs(ref["t"][both])
    assert rel.max(initial=0.0) <= 1e-6, f"t differs from FP64 by {rel.max():.3g} relative"
    assert both.sum() > 0.05 * len(rays)
    return int(mism.sum())


def _three_sigma(dev, sc, cam, w, h, spp, depth, albedo):
    """GPU image (spp passes from pass 0) vs the reference-numerics oracle on
    passes 1000.. (independent samples); per-pixel sample variance from the
    GPU's per-pass increments and the oracle's sum of squares."""
    dev.set_camera(cam)
    dev.init_frame(w, h)
    prev = np.zeros((h, w), np.float64)
    s1 = np.zeros((h, w))
    s2 = np.zeros((h, w))
    for _ in range(spp):
        dev.render("diffuseHip", max_depth=depth, albedo=albedo)
        cur = dev.accum()[:, :, 0].astype(np.float64)
        x = cur - prev
        prev = cur
        s1 += x
        s2 += x * x
    ref = sc.render(cam, w, h, spp=spp, max_depth=depth, albedo=albedo, renderer=1, numerics=1, pass0=1000)
    mg, mr = s1 / spp, ref["accum"][:, :, 0].astype(np.float64) / spp
    vg = np.maximum(s2 / spp - mg * mg, 0.0)
    vr = np.maximum(ref["sumsq"].astype(np.float64) / spp - mr * mr, 0.0)
    sigma = np.maximum(np.sqrt((vg + vr) / spp), 1e-3)
    outside = np.abs(mg - mr) > 3.0 * sigma
    assert outside.mean() < 0.01, f"{outside.mean():.4f} of pixels outside 3 sigma"
    z = (mg.sum() - mr.sum()) / math.sqrt(((vg + vr) / spp).sum())
    assert abs(z) < 5.0, f"image means differ: z = {z:.2f}"


# ---------------------------------------------------------------- config 1
@pytest.mark.parametrize("pipeline", ["fused", "wavefront"])
def test_config1_full_size_bit_exact(dev, orc, scenes, pipeline):
    cfg = scenes.CONFIGS["heightfield_box_512"]
    w, h, spp, depth, albedo = cfg["width"], cfg["height"], cfg["spp"], cfg["max_depth"], cfg["albedo"]
    mesh = scenes.build_scene("heightfield_box_512")
    assert mesh.tri_count == 8202
    dev.upload_scene(mesh.idx, mesh.pos)
    dev.build()
    cam = _cam(orc, scenes, "heightfield_box_512", w, h)
    dev.set_camera(cam)
    dev.init_frame(w, h)
    dev.set_pipeline(pipeline)
    try:
        rays = sum(dev.render("diffuseHip", max_depth=depth, albedo=albedo)["ray"] for _ in range(spp))
    finally:
        dev.set_pipeline("fused")
    got = dev.accum()
    ref = orc.Scene(mesh.idx, mesh.pos).render(cam, w, h, spp=spp, max_depth=depth, albedo=albedo, renderer=1,
                                               numerics=0)
    assert rays == ref["rays"]
    assert np.array_equal(got, ref["accum"]), f"{int((got != ref['accum']).any(-1).sum())} pixels differ"


# ---------------------------------------------------------------- config 2
@pytest.fixture(scope="module")
def spheres(scenes, orc):
    mesh = scenes.build_scene("spheres_1080p")
    return mesh, orc.Scene(mesh.idx, mesh.pos)


def test_config2_all_primary_hits_vs_fp64(dev, orc, scenes, spheres):
    mesh, sc = spheres
    cfg = scenes.CONFIGS["spheres_1080p"]
    w, h = cfg["width"], cfg["height"]
    dev.upload_scene(mesh.idx, mesh.pos)
    dev.build()
    dev.set_camera(_cam(orc, scenes, "spheres_1080p", w, h))
    dev.init_frame(w, h)
    rays, _ = dev.debug_raygen(0)
    assert rays.shape[0] == 2_073_600
    hits = dev.intersect(rays)
    _check_fp64(sc, rays, hits, _extent(mesh))


def test_config2_image_within_3_sigma_of_reference_numerics(dev, orc, scenes, spheres):
    mesh, sc = spheres
    cfg = scenes.CONFIGS["spheres_1080p"]
    dev.upload_scene(mesh.idx, mesh.pos)
    dev.build()
    w, h = 480, 270
    _three_sigma(dev, sc, _cam(orc, scenes, "spheres_1080p", w, h), w, h, 32, cfg["max_depth"], cfg["albedo"])


# ---------------------------------------------------------------- config 3
def test_config3_batched_workload_equals_plain_and_oracle(prt, orc, scenes):
    specs = scenes.batch_workload()
    assert len(specs) == 32
    meshes = [scenes.scene_mix(sp["seed"], sp["target_tris"]) for sp in specs]
    jobs = []
    for sp in specs:
        v = sp["view"]
        cam = prt.camera_from_view(v["pos"], v["angle_x"], v["angle_y"], v["fov_y"], sp["width"] / sp["height"])
        jobs.append({"scene": len(jobs), "width": sp["width"], "height": sp["height"], "spp": sp["spp"],
                     "max_depth": sp["max_depth"], "pass0": 0, "albedo": sp["albedo"], "camera": cam})
    with prt.HipDevice(0) as bdev:
        for m in meshes:
            bdev.batch_add_scene(m.idx, m.pos)
        bdev.batch_set_jobs(jobs)
        bdev.batch_autotune()
        tuning = bdev.batch_get_tuning()
        st = bdev.batch_render()
        batched = [bdev.batch_accum(j) for j in range(len(jobs))]
    assert sum(m.tri_count for m in meshes) > 10_000_000
    assert st["paths"] == sum(j["spp"] * j["width"] * j["height"] for j in jobs) > 160_000_000
    print(f"config 3: {st['paths']} paths, {st['ray']} rays, tuning {tuning}", flush=True)
    total = 0
    with prt.HipDevice(0) as d:
        for j, (job, m) in enumerate(zip(jobs, meshes)):
            w, h = job["width"], job["height"]
            d.upload_scene(m.idx, m.pos)
            d.build()
            d.set_camera(job["camera"])
            d.init_frame(w, h)
            total += d.render("diffuseHip", max_depth=job["max_depth"], albedo=job["albedo"])["ray"]
            first = d.accum()
            for _ in range(job["spp"] - 1):
                total += d.render("diffuseHip", max_depth=job["max_depth"], albedo=job["albedo"])["ray"]
            got = d.accum()
            assert np.array_equal(batched[j], got), \
                f"job {j} ({w}x{h}, {job['spp']} spp, {m.tri_count} tris): batched differs from plain"
            r0 = (h // 2) & ~7
            ref = orc.Scene(m.idx, m.pos).render(job["camera"], w, h, spp=1, max_depth=job["max_depth"],
                                                 albedo=job["albedo"], renderer=1, numerics=0, rows=(r0, r0 + 8))
            assert np.array_equal(first[r0:r0 + 8], ref["accum"][r0:r0 + 8]), f"job {j}: first pass rows differ"
    assert st["ray"] == total


# ---------------------------------------------------------------- config 5
@pytest.fixture(scope="module")
def city(scenes, orc):
    mesh = scenes.build_scene("city_4k")
    return mesh, orc.Scene(mesh.idx, mesh.pos)


def test_city_primary_hits_vs_fp64(dev, orc, scenes, city):
    mesh, sc = city
    dev.upload_scene(mesh.idx, mesh.pos)
    dev.build()
    w, h = 640, 360
    dev.set_camera(_cam(orc, scenes, "city_4k", w, h))
    dev.init_frame(w, h)
    rays, _ = dev.debug_raygen(3)
    assert rays.shape[0] >= 200_000
    hits = dev.intersect(rays)
    _check_fp64(sc, rays, hits, _extent(mesh))


def test_city_image_within_3_sigma_of_reference_numerics(dev, orc, scenes, city):
    mesh, sc = city
    cfg = scenes.CONFIGS["city_4k"]
    dev.upload_scene(mesh.idx, mesh.pos)
    dev.build()
    w, h = 480, 270
    _three_sigma(dev, sc, _cam(orc, scenes, "city_4k", w, h), w, h, 32, cfg["max_depth"], cfg["albedo"])


def test_city_batched_job_equals_plain(prt, dev, orc, scenes, city):
    mesh, _ = city
    cfg = scenes.CONFIGS["city_4k"]
    w, h, spp = 960, 540, 2
    cam = _cam(orc, scenes, "city_4k", w, h)
    with prt.HipDevice(0) as bdev:
        bdev.batch_add_scene(mesh.idx, mesh.pos)
        bdev.batch_set_jobs([{"scene": 0, "width": w, "height": h, "spp": spp, "max_depth": cfg["max_depth"],
                              "pass0": 0, "albedo": cfg["albedo"], "camera": cam}])
        st = bdev.batch_render()
        got = bdev.batch_accum(0)
    dev.upload_scene(mesh.idx, mesh.pos)
    dev.build()
    dev.set_camera(cam)
    dev.init_frame(w, h)
    rays = sum(dev.render("diffuseHip", max_depth=cfg["max_depth"], albedo=cfg["albedo"])["ray"]
               for _ in range(spp))
    assert st["ray"] == rays
    assert np.array_equal(got, dev.accum())


def test_city_adjoint_crop_vs_fp64_adjoint(dev, orc, scenes, city):
    mesh, sc = city
    cfg = scenes.CONFIGS["city_4k"]
    w, h, spp, depth = 256, 144, 4, cfg["max_depth"]
    cam = _cam(orc, scenes, "city_4k", w, h)
    dev.upload_scene(mesh.idx, mesh.pos)
    dev.build()
    dev.set_camera(cam)
    dev.init_frame(w, h)
    a = scenes.albedo_set(mesh.tri_count, cfg["albedo_seed"])
    dev.set_albedo(a)
    img, st = dev.render_albedo_image(pass0=0, spp=spp, max_depth=depth)
    tape = orc.PathTape(sc, cam, w, h, pass0=0, spp=spp, max_depth=depth)
    a64 = a.astype(np.float64)
    assert st["ray"] == tape.rays
    assert np.allclose(img, tape.image(a64), rtol=2e-6, atol=1e-7)
    dldi = np.random.default_rng(8).normal(size=(h, w, 3)).astype(np.float32)
    grad, ast = dev.adjoint_albedo(pass0=0, spp=spp, max_depth=depth, dldi=dldi)
    gref = tape.adjoint(a64, dldi.astype(np.float64))
    assert (np.abs(gref).sum(1) > 0).sum() > 1000
    assert np.linalg.norm(grad - gref) <= 1e-4 * np.linalg.norm(gref)
    big = np.abs(gref) > 1e-3 * np.abs(gref).max()
    assert np.all(np.abs(grad[big] - gref[big]) <= 1e-4 * np.abs(gref[big]) + 1e-7)
