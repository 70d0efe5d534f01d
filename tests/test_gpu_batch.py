"""Batched form: many scenes, cameras and spp counts in one persistent launch.

Bar: every job's accumulator is bit-identical to rendering the same scene
with prt_hip_render one sample at a time (same passes), the ray count equals
the plain renders' sum, replays of the captured hipGraph equal stream
launches, and a second batch render continues with the next passes.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _jobs(prt, scenes):
    meshes = [scenes.scene_heightfield_box(1001, cells=16), scenes.scene_convex_sphere(2),
              scenes.scene_mix(5, 3000)]
    v = scenes.BATCH_VIEW
    specs = [(0, 64, 48, 3, 4, 0), (1, 40, 24, 1, 3, 7), (2, 96, 64, 2, 6, 100), (0, 72, 56, 4, 2, 3),
             (2, 33, 17, 5, 6, 11)]
    jobs = []
    for scene, w, h, spp, depth, pass0 in specs:
        cam = prt.camera_from_view(v["pos"], v["angle_x"], v["angle_y"], v["fov_y"], w / h)
        jobs.append({"scene": scene, "width": w, "height": h, "spp": spp, "max_depth": depth, "pass0": pass0,
                     "albedo": 0.8, "camera": cam})
    return meshes, jobs


def _plain(prt, mesh, job, passes):
    with prt.HipDevice(0) as d:
        d.upload_scene(mesh.idx, mesh.pos)
        d.build()
        d.set_camera(job["camera"])
        d.init_frame(job["width"], job["height"])
        d.set_pass(job["pass0"])
        rays = sum(d.render("diffuseHip", max_depth=job["max_depth"], albedo=job["albedo"])["ray"]
                   for _ in range(passes))
        return d.accum(), rays


def test_batch_equals_plain(prt, scenes):
    meshes, jobs = _jobs(prt, scenes)
    with prt.HipDevice(0) as dev:
        for m in meshes:
            dev.batch_add_scene(m.idx, m.pos)
        dev.batch_set_jobs(jobs)
        st1 = dev.batch_render()
        first = [dev.batch_accum(j) for j in range(len(jobs))]
        st2 = dev.batch_render()
        second = [dev.batch_accum(j) for j in range(len(jobs))]
    total1 = 0
    for j, job in enumerate(jobs):
        ref, rays = _plain(prt, meshes[job["scene"]], job, job["spp"])
        total1 += rays
        assert np.array_equal(first[j], ref), f"job {j} differs from plain rendering"
        ref2, _ = _plain(prt, meshes[job["scene"]], job, 2 * job["spp"])
        assert np.array_equal(second[j], ref2), f"job {j}: second batch render is not the next passes"
    assert st1["ray"] == total1
    assert st1["paths"] == sum(j["spp"] * j["width"] * j["height"] for j in jobs)
    assert st2["paths"] == st1["paths"]


def test_batch_graph_equals_stream_launches(prt, scenes):
    meshes, jobs = _jobs(prt, scenes)
    out = {}
    for graph in (True, False):
        with prt.HipDevice(0) as dev:
            for m in meshes:
                dev.batch_add_scene(m.idx, m.pos)
            dev.batch_set_graph(graph)
            dev.batch_set_jobs(jobs)
            for _ in range(3):  # replays of the same graph, eager reads between them
                dev.batch_render()
                out.setdefault(graph, []).append([dev.batch_accum(j).copy() for j in range(len(jobs))])
    for a, b in zip(out[True], out[False]):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_batch_autotune_and_variants(prt, scenes):
    meshes, jobs = _jobs(prt, scenes)
    with prt.HipDevice(0) as dev:
        for m in meshes:
            dev.batch_add_scene(m.idx, m.pos)
        dev.batch_set_jobs(jobs)
        results = dev.batch_autotune()
        assert len(results) >= 8 and all(r["ms"] > 0 for r in results)
        assert {r["variant"] for r in results} == {0, 1, 2, 3}
        chosen = dev.batch_get_tuning()
        best = min(results, key=lambda r: r["ms"])
        assert (chosen["variant"], chosen["blocks_per_cu"], chosen["refill"]) == \
            (best["variant"], best["blocks_per_cu"], best["refill"])
        # every launch shape renders the same image
        imgs = []
        for variant in range(4):
            dev.batch_set_jobs(jobs)
            dev.batch_set_tuning(variant, 0, 16 + 8 * variant)
            dev.batch_render()
            imgs.append([dev.batch_accum(j) for j in range(len(jobs))])
        for other in imgs[1:]:
            for x, y in zip(imgs[0], other):
                assert np.array_equal(x, y)


def test_batch_rejects_bad_jobs(prt, scenes):
    mesh = scenes.scene_convex_sphere(1)
    cam = prt.camera_from_view([0, 0, 3], 0, 0, math.radians(45), 1.0)
    with prt.HipDevice(0) as dev:
        dev.batch_add_scene(mesh.idx, mesh.pos)
        for bad in ({"scene": 1}, {"spp": 0}, {"max_depth": 1025}, {"max_depth": -1}, {"width": 0}):
            job = {"scene": 0, "width": 8, "height": 8, "spp": 1, "max_depth": 2, "camera": cam}
            job.update(bad)
            with pytest.raises(RuntimeError):
                dev.batch_set_jobs([job])


def test_batch_graph_after_deeper_plain_build(prt, scenes):
    """A captured batch graph holds the traversal spill buffer and its per-lane
    stride.  Building a deeper plain scene on the same context afterwards
    replaces both; the next batch render must replay a graph captured with
    the new ones and still equal plain stream launches."""
    meshes, jobs = _jobs(prt, scenes)
    deep = scenes.scene_spheres(1002, count=64, subdiv=4)
    out = {}
    depths = {}
    for graph in (True, False):
        with prt.HipDevice(0) as dev:
            batch_depth = max(dev.batch_add_scene(m.idx, m.pos)[1]["depth"] for m in meshes)
            dev.batch_set_graph(graph)
            dev.batch_set_jobs(jobs)
            dev.batch_render()
            dev.upload_scene(deep.idx, deep.pos)
            depths[graph] = (batch_depth, dev.build()["depth"])
            dev.batch_render()
            out[graph] = [dev.batch_accum(j).copy() for j in range(len(jobs))]
    batch_depth, plain_depth = depths[True]
    assert plain_depth > batch_depth, "the plain scene must need a deeper traversal stack"
    for j, (x, y) in enumerate(zip(out[True], out[False])):
        assert np.array_equal(x, y), f"job {j}: graph replay after the rebuild differs"
