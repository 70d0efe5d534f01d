#!/usr/bin/env python3
"""Benchmark of the diffuseHip path (ProtoRay `render -r diffuse` on the HIP device).

Workload: 256 icospheres (subdivision 4, 1,310,722 fp32 triangles with the
ground quad), 1920x1080, max depth 8, albedo 0.7.  One step is one frame: one
sample per pixel through raygen -> (closest hit -> shade) x (max depth + 1),
exactly one call of Device::render in the reference
(render/device_cuda.cpp:95-109).  The metric is the reference's Mrays/s: every
ray handed to the closest-hit query, primary and secondary, divided by the
timed seconds (render/diffuse_renderer_cuda.cu:186).

Prints ONE JSON line.  `value` is wall-clock Mrays/s over the K timed frames
with the scene, BVH and path state resident in HBM.  The default pipeline is
the fused persistent path kernel (raygen, closest hit, cosine bounce and
accumulation in one launch per frame), which is the dominant and only kernel
of a frame, with the frame lookahead on (the library's default: while a
frame's last paths run, idle lanes start the next pass's paths, which the next
call keeps; each call's accumulator and ray count stay those of its own frame).
The same frames with the lookahead off are timed beside it
(`no_lookahead_mrays`).  `roofline` is for the path kernel: algorithmic bytes per frame
rays * (80 * nodes/ray + 36 * tris/ray) + 32 * paths (BVH8 nodes and triangle
records the traversal fetches, the accumulator read + write per path), with
nodes/tris per ray from an instrumented replay of the same frames, divided by
the kernel's device time (HIP events on the library's stream).  The
per-depth wavefront pipeline is timed beside it and its algorithmic bytes
(ray and hit queues in HBM) are reported too.  `cpu_baseline` is the oracle's
restatement of the CPU stream renderer (render/diffuse_renderer_stream.h) on
the host cores.  Further legs: the batched form (one persistent launch over
many scenes, cameras and spp), the per-triangle albedo adjoint, and the
15.96M-triangle city scene.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = "embree-benchmark-protoray_amd"
METRIC = "Mrays/s diffuse PT 1080p 1.3M tris: plain vs batched; adjoint Mrays/s; %HBM"
HBM_PEAK_GBS = 8000.0
WORKLOAD = "spheres_1080p"


def measured_traffic() -> dict | None:
    """HBM-side bytes per fused frame and the counters around them, from the
    committed rocprofv3 --pmc summary of this workload (FETCH_SIZE and
    WRITE_SIZE scaled by a calibration copy of known size)."""
    path = os.path.join(ROOT, "profiles", "k_trace_fused_spheres_1080p.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r'\{"traffic_bytes_per_launch": \d+[^}]*\}', text)
    return json.loads(m.group(0)) if m else None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def wavefront_bytes(rays: int, cont: int, paths: int, n_node: float, n_tri: float) -> float:
    """Algorithmic HBM bytes of one frame of the wavefront pipeline (per-depth
    launches with ray queues in HBM): traversal 48 + 80 n_node + 36 n_tri per
    ray, shading reads 56 per ray, 88 per continuing hit (index, vertices,
    next-queue write), 32 per finished path (accumulator read + write) and 40
    per primary ray written by raygen."""
    return rays * (48 + 56 + 80 * n_node + 36 * n_tri) + cont * 88 + paths * 32 + paths * 40


def fused_bytes(rays: int, paths: int, n_node: float, n_tri: float) -> float:
    """Algorithmic HBM bytes of one frame of the fused pipeline, whose rays,
    hits and path state never leave the lanes: BVH8 nodes (80 B) and triangle
    records (36 B) fetched by the traversal, and the accumulator read + write
    (32 B) of every path."""
    return rays * (80 * n_node + 36 * n_tri) + paths * 32


def host_threads() -> tuple[int, int]:
    """Threads the CPU baseline runs on, and the host's configured CPU count.
    The reference's Tasking runs one worker per hardware thread and reports
    sysconf(_SC_NPROCESSORS_CONF) (sys/tasking.h:33-52, sys/sysinfo.cpp:50-61).
    Here the workers are the CPUs this process may run on, capped by the
    OMP_NUM_THREADS share the host gives one GPU's job when it sets one."""
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        avail = min(avail, int(share))
    try:
        conf = int(os.sysconf("SC_NPROCESSORS_CONF"))
    except (AttributeError, ValueError, OSError):
        conf = os.cpu_count() or avail
    return max(1, avail), conf


def cpu_stream(mesh, cfg, cam, frames: int, threads: int) -> dict:
    """The oracle's restatement of the CPU stream renderer
    (render/diffuse_renderer_stream.h:82-225) on `frames` full frames."""
    sys.path.insert(0, ROOT)
    from oracle import pyoracle as orc
    sc = orc.Scene(mesh.idx, mesh.pos)
    w, h = cfg["width"], cfg["height"]
    r = sc.stream_bench(cam, w, h, 0, h, frames, cfg["max_depth"], cfg["albedo"], threads=threads)
    sc.close()
    return r


def cpu_tree() -> str:
    sys.path.insert(0, ROOT)
    from oracle import pyoracle as orc
    return ("an 8-wide binned-SAH BVH with AVX2 child tests" if orc.cpu_has_avx2()
            else "a scalar binned-SAH BVH2")


def cpu_baseline(scenes, mesh, cfg, cam, frames: int) -> dict:
    threads, conf = host_threads()
    tree = cpu_tree()
    r = cpu_stream(mesh, cfg, cam, frames, threads)
    w, h = cfg["width"], cfg["height"]
    return {"value": round(r["rays"] / r["seconds"] / 1e6, 3), "unit": "Mrays/s", "cores": threads,
            "kind": "port", "host_cpus_configured": conf,
            "sample": f"{frames} full {w}x{h} frames (passes 0-{frames - 1}), max depth {cfg['max_depth']}, "
                      f"restated stream renderer over {tree} (not Embree), "
                      f"{r['rays']} rays in {r['seconds']:.2f} s"}


def bench_config1(pkg, scenes, args) -> dict:
    """The reference's CPU-runnable case: 64x64-cell heightfield in an open
    box (8,202 triangles), 512x512, 4 spp, max depth 4, R = 0.8, on the GPU
    (plain fused path, the 4 frames timed) and on the CPU stream baseline
    (64 frames: passes 0-63, about a second of work)."""
    cfg = scenes.CONFIGS["heightfield_box_512"]
    w, h, spp, depth, albedo = cfg["width"], cfg["height"], cfg["spp"], cfg["max_depth"], cfg["albedo"]
    mesh = scenes.build_scene("heightfield_box_512")
    v = cfg["view"]
    cam = pkg.camera_from_view(v["pos"], v["angle_x"], v["angle_y"], v["fov_y"], w / h)
    with pkg.HipDevice(0) as dev:
        dev.upload_scene(mesh.idx, mesh.pos)
        dev.build()
        dev.set_camera(cam)
        dev.init_frame(w, h)
        for _ in range(2 * spp):  # warm-up: the same passes' chain, then cleared
            dev.render("diffuseHip", max_depth=depth, albedo=albedo)
        best = None
        for _ in range(5):
            dev.clear_frame()
            t0 = time.perf_counter()
            rays = sum(dev.render("diffuseHip", max_depth=depth, albedo=albedo)["ray"] for _ in range(spp))
            wall = time.perf_counter() - t0
            best = wall if best is None or wall < best else best
    out = {"workload": f"heightfield_box_512: {mesh.tri_count} tris, {w}x{h}, {spp} spp, depth {depth}, R {albedo}",
           "gpu_mrays": round(rays / best / 1e6, 2), "gpu_ms_per_frame": round(best * 1e3 / spp, 4),
           "rays_per_image": rays}
    if not args.no_cpu_baseline:
        threads, conf = host_threads()
        frames = 16 * spp  # about a second of CPU work
        r = cpu_stream(mesh, cfg, cam, frames, threads)
        out["cpu_baseline"] = {"value": round(r["rays"] / r["seconds"] / 1e6, 3), "unit": "Mrays/s",
                               "cores": threads, "kind": "port", "host_cpus_configured": conf,
                               "sample": f"{frames} full {w}x{h} frames (passes 0-{frames - 1}), restated stream "
                                         f"renderer over {cpu_tree()} (not Embree), {r['rays']} rays "
                                         f"in {r['seconds']:.2f} s"}
        out["gpu_over_cpu"] = round(out["gpu_mrays"] / out["cpu_baseline"]["value"], 1)
    log(f"config 1 (heightfield_box_512): GPU {out['gpu_mrays']} Mrays/s"
        + (f", CPU {out['cpu_baseline']['value']} Mrays/s on {out['cpu_baseline']['cores']} threads"
           if "cpu_baseline" in out else ""))
    return out


def profile_traffic(name: str) -> dict | None:
    """HBM-side bytes per launch of a kernel and workload from its committed
    rocprofv3 --pmc summary (profiles/<name>.md), or None."""
    try:
        text = open(os.path.join(ROOT, "profiles", name + ".md")).read()
    except OSError:
        return None
    m = re.search(r'\{"traffic_bytes_per_launch": \d+[^}]*\}', text)
    return json.loads(m.group(0)) if m else None


def roofline_entry(kernel: str, algo_bytes: float, kernel_ms: float, rays: int, traffic: dict | None,
                   source: str) -> dict:
    """roofline object of one kernel; `traffic` (the committed --pmc summary)
    was measured on a launch of `rays_per_launch` rays, so it is scaled to
    this launch's ray count."""
    achieved = algo_bytes / (kernel_ms * 1e-3) / 1e9
    hbm = None
    if traffic:
        hbm = traffic["traffic_bytes_per_launch"]
        if traffic.get("rays_per_launch"):
            hbm = int(round(hbm * rays / traffic["rays_per_launch"]))
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": hbm,
            "traffic_source": f"profiles/{source}.md" if traffic else None,
            "l2_hit_rate": traffic.get("l2_hit_rate") if traffic else None,
            "valu_busy": traffic.get("valu_busy") if traffic else None,
            "valu_lane_utilization": traffic.get("valu_lane_utilization") if traffic else None,
            "kernel": kernel, "bytes_per_ray": round(algo_bytes / max(1, rays), 1),
            "avg_launch_ms": round(kernel_ms, 4)}


def batch_roofline(dev, timed: list, kernel: str, source: str) -> dict:
    """Roofline of the batched path kernel: per launch rays * (80 * nodes/ray
    + 36 * tris/ray) of BVH8 node and triangle-record fetches plus 4 B per
    path (its sample-plane write), with nodes/tris per ray from one
    instrumented render of the same jobs (the next passes), over the device
    time of the replayed graph (HIP events around it; the graph's finalize
    and counter launches are inside, so this is a lower bound)."""
    dev.set_profiling(counters=True)
    c = dev.batch_render()
    dev.set_profiling()
    n_node = c["nodes"] / max(1, c["ray"])
    n_tri = c["tris"] / max(1, c["ray"])
    rays = timed[-1]["ray"]
    ms = min(x["renderMs"] for x in timed)
    algo = rays * (80 * n_node + 36 * n_tri) + 4 * timed[-1]["paths"]
    out = roofline_entry(kernel, algo, ms, rays, profile_traffic(source), source)
    out.update({"nodes_per_ray": round(n_node, 3), "tris_per_ray": round(n_tri, 3)})
    return out


def bench_batched(pkg, scenes, dev, mesh, cam, cfg, args, plain_mrays: float, wavefront_mrays: float,
                  no_look_mrays: float) -> dict:
    """Batched form.  (1) spheres_1080p as one job of spp = steps: the same
    frames as the plain run (passes warmup..warmup+steps-1) in one persistent
    launch per render; (2) the 32-scene workload (triangles log-uniform
    1k-2M, resolutions {256^2, 640x480, 1280x720, 1920x1080}, spp 1..16,
    depth 6) in one launch, launch shape from the autotuner."""
    w, h, depth, albedo = cfg["width"], cfg["height"], cfg["max_depth"], cfg["albedo"]
    dev.batch_reset()
    sid, _ = dev.batch_add_scene(mesh.idx, mesh.pos)
    job = {"scene": sid, "width": w, "height": h, "spp": args.steps, "max_depth": depth,
           "pass0": args.warmup, "albedo": albedo, "camera": cam}
    dev.batch_set_jobs([job])
    tune2 = dev.batch_autotune()
    chosen2 = dev.batch_get_tuning()
    dev.batch_set_jobs([job])
    dev.batch_render()  # warm-up (also captures the graph)
    reps = 3
    t0 = time.perf_counter()
    st = [dev.batch_render() for _ in range(reps)]
    wall = time.perf_counter() - t0
    rays = sum(x["ray"] for x in st)
    c2 = rays / wall / 1e6
    roof2 = batch_roofline(dev, st, "k_trace<batch> (persistent batched path kernel), spheres_1080p job",
                           "k_trace_batch_spheres_1080p")
    log(f"batched spheres_1080p ({args.steps} spp per launch): {c2:.1f} Mrays/s "
        f"({c2 / plain_mrays:.2f}x plain fused, {c2 / wavefront_mrays:.2f}x plain wavefront), tuning {chosen2}")

    dev.batch_reset()
    specs = scenes.batch_workload(args.batch_scenes)
    tris = 0
    bjobs = []
    for sp in specs:
        m = scenes.scene_mix(sp["seed"], sp["target_tris"])
        tris += m.tri_count
        sid, _ = dev.batch_add_scene(m.idx, m.pos)
        v = sp["view"]
        bcam = pkg.camera_from_view(v["pos"], v["angle_x"], v["angle_y"], v["fov_y"], sp["width"] / sp["height"])
        bjobs.append({"scene": sid, "width": sp["width"], "height": sp["height"], "spp": sp["spp"],
                      "max_depth": sp["max_depth"], "albedo": sp["albedo"], "camera": bcam})
    dev.batch_set_jobs(bjobs)
    tune3 = dev.batch_autotune()
    chosen3 = dev.batch_get_tuning()
    dev.batch_render()
    t0 = time.perf_counter()
    st3 = [dev.batch_render() for _ in range(2)]
    wall3 = time.perf_counter() - t0
    rays3 = sum(x["ray"] for x in st3)
    c3 = rays3 / wall3 / 1e6
    roof3 = batch_roofline(dev, st3, "k_trace<batch> (persistent batched path kernel), 32-scene workload",
                           "k_trace_batch_scenes32")
    log(f"batched {len(specs)} scenes ({tris} triangles, {st3[0]['paths']} paths per launch): {c3:.1f} Mrays/s, "
        f"tuning {chosen3}")
    dev.batch_reset()
    return {"spheres_1080p_mrays": round(c2, 2), "spheres_1080p_spp_per_launch": args.steps,
            "ratio_vs_plain_fused": round(c2 / plain_mrays, 3),
            "ratio_vs_plain_fused_no_lookahead": round(c2 / no_look_mrays, 3),
            "ratio_vs_plain_wavefront": round(c2 / wavefront_mrays, 3),
            "spheres_1080p_tuning": chosen2,
            "scenes32_mrays": round(c3, 2), "scenes32_triangles": tris, "scenes32_paths": st3[0]["paths"],
            "scenes32_rays_per_launch": st3[0]["ray"], "scenes32_tuning": chosen3,
            "autotune_candidates": len(tune3), "roofline": roof2, "scenes32_roofline": roof3}


def bench_adjoint(pkg, scenes, args) -> dict:
    """Differentiable workload: 128x128 heightfield + 64 boxes (33,536
    triangles), 512x512, 8 spp, depth 6, per-triangle RGB albedo U[0.1, 0.9];
    L2 loss against a target rendered from a second albedo set.  Times the
    forward image (one launch over all 8 samples) and the path-replay adjoint
    (one launch over all 8 samples)."""
    cfg = scenes.CONFIGS["heightfield_boxes_512"]
    w, h, spp, depth = cfg["width"], cfg["height"], cfg["spp"], cfg["max_depth"]
    mesh = scenes.build_scene("heightfield_boxes_512")
    v = cfg["view"]
    cam = pkg.camera_from_view(v["pos"], v["angle_x"], v["angle_y"], v["fov_y"], w / h)
    a = scenes.albedo_set(mesh.tri_count, cfg["albedo_seed"])
    b = scenes.albedo_set(mesh.tri_count, cfg["target_seed"])
    with pkg.HipDevice(0) as dev:
        dev.upload_scene(mesh.idx, mesh.pos)
        dev.build()
        dev.set_camera(cam)
        dev.init_frame(w, h)
        dev.set_albedo(b)
        target, _ = dev.render_albedo_image(0, spp, depth)
        dev.set_albedo(a)
        img, _ = dev.render_albedo_image(0, spp, depth)
        dldi = 2.0 * (img - target)
        dev.adjoint_albedo(0, spp, depth, dldi)  # warm-up
        reps = 5
        fwd = [dev.render_albedo_image(0, spp, depth)[1] for _ in range(reps)]
        adj = [dev.adjoint_albedo(0, spp, depth, dldi)[1] for _ in range(reps)]
        dev.set_profiling(counters=True)
        cadj = dev.adjoint_albedo(0, spp, depth, dldi)[1]
        dev.set_profiling()
    f_ms = min(s["renderMs"] for s in fwd)
    a_ms = min(s["renderMs"] for s in adj)
    a_rays = adj[0]["ray"]
    # adjoint kernel: both traces' node and triangle fetches, one dL/dI read
    # (12 B) per path and at most one 12 B gradient atomic per replayed ray
    n_node = cadj["nodes"] / max(1, cadj["ray"])
    n_tri = cadj["tris"] / max(1, cadj["ray"])
    replay_rays = a_rays - fwd[0]["ray"]
    a_bytes = a_rays * (80 * n_node + 36 * n_tri) + 12 * replay_rays + 12 * adj[0]["paths"]
    a_roof = roofline_entry("k_trace<adjoint> (path-replay albedo adjoint, one launch over all samples)", a_bytes,
                            a_ms, a_rays, profile_traffic("k_trace_adjoint_heightfield_boxes_512"),
                            "k_trace_adjoint_heightfield_boxes_512")
    a_roof.update({"nodes_per_ray": round(n_node, 3), "tris_per_ray": round(n_tri, 3), "replay_rays": replay_rays})
    out = {"workload": "heightfield_boxes_512: 33,536 tris, 512x512, 8 spp, depth 6, per-triangle RGB albedo",
           "forward_ms": round(f_ms, 3), "forward_mrays": round(fwd[0]["ray"] / f_ms / 1e3, 2),
           "adjoint_ms": round(a_ms, 3), "adjoint_mrays": round(a_rays / a_ms / 1e3, 2),
           "adjoint_over_forward": round(a_ms / f_ms, 3), "adjoint_rays": a_rays, "forward_rays": fwd[0]["ray"],
           "loss": float(((img.astype(np.float64) - target) ** 2).sum()), "roofline": a_roof}
    log(f"adjoint: forward {f_ms:.2f} ms, adjoint {a_ms:.2f} ms ({a_ms / f_ms:.2f}x), "
        f"{out['adjoint_mrays']} Mrays/s")
    return out


def bench_city(pkg, scenes, args) -> dict:
    """Large stress workload: 1.33 M boxes (15.96 M triangles), 3840x2160,
    max depth 8: LBVH->BVH8 build rate, the batched forward (all spp in one
    launch) and the per-triangle albedo adjoint (16 M x 3 fp32 gradient)."""
    cfg = scenes.CONFIGS["city_4k"]
    w, h, depth = cfg["width"], cfg["height"], cfg["max_depth"]
    spp = args.city_spp
    mesh = scenes.build_scene("city_4k")
    v = cfg["view"]
    cam = pkg.camera_from_view(v["pos"], v["angle_x"], v["angle_y"], v["fov_y"], w / h)
    out = {"workload": f"city_4k: {mesh.tri_count} tris, {w}x{h}, {spp} spp, depth {depth}", "spp": spp}
    with pkg.HipDevice(0) as dev:
        sid, bst = dev.batch_add_scene(mesh.idx, mesh.pos)
        out["buildMs"] = round(bst["buildMs"], 2)
        out["buildMprim"] = round(bst["buildMprim"], 1)
        out["bvh_depth"] = bst["depth"]
        job = {"scene": sid, "width": w, "height": h, "spp": spp, "max_depth": depth, "albedo": cfg["albedo"],
               "camera": cam}
        dev.batch_set_jobs([job])
        dev.batch_render()
        st = dev.batch_render()
        out["batched_mrays"] = round(st["ray"] / st["renderMs"] / 1e3, 2)
        out["batched_ms"] = round(st["renderMs"], 2)
        out["rays_per_launch"] = st["ray"]
        log(f"city: build {bst['buildMs']:.1f} ms ({bst['buildMprim']:.0f} Mprim/s), batched forward "
            f"{out['batched_mrays']} Mrays/s ({st['ray']} rays)")
        dev.batch_reset()
        dev.upload_scene(mesh.idx, mesh.pos)
        dev.build()
        dev.set_camera(cam)
        dev.init_frame(w, h)
        dev.set_albedo(scenes.albedo_set(mesh.tri_count, cfg["albedo_seed"]))
        aspp = args.city_adjoint_spp
        img, fst = dev.render_albedo_image(0, aspp, depth)
        grad, ast = dev.adjoint_albedo(0, aspp, depth, img)  # loss = 0.5 * sum(image^2)
        out.update({"adjoint_spp": aspp, "forward_ms": round(fst["renderMs"], 2),
                    "adjoint_ms": round(ast["renderMs"], 2), "adjoint_mrays": round(ast["ray"] / ast["renderMs"] / 1e3, 2),
                    "adjoint_over_forward": round(ast["renderMs"] / fst["renderMs"], 3),
                    "grad_nonzero_triangles": int((np.abs(grad).sum(1) > 0).sum())})
        log(f"city adjoint ({aspp} spp): forward {fst['renderMs']:.1f} ms, adjoint {ast['renderMs']:.1f} ms")
    return out


COPY_FLOATS = 1 << 27   # calibration copy: 512 MiB read + 512 MiB written


def pmc_pass(pkg, mesh, cam, cfg, args) -> int:
    """Short run for the rocprofv3 --pmc passes: `--steps` frames of the fused
    path kernel (the timed kernel), 3 batched launches of spheres_1080p x 16
    spp and 2 of the 32-scene workload, 3 albedo adjoint launches of the
    differentiable workload, then a plain elementwise copy of
    COPY_FLOATS floats (torch, `dst = src * 1`), whose FETCH_SIZE / WRITE_SIZE
    against its known 512 MiB read and written scale the counters to bytes."""
    import torch
    w, h, depth, albedo = cfg["width"], cfg["height"], cfg["max_depth"], cfg["albedo"]
    dev = pkg.HipDevice(0)
    dev.upload_scene(mesh.idx, mesh.pos)
    dev.build()
    dev.set_camera(cam)
    dev.init_frame(w, h)
    for _ in range(args.warmup):
        dev.render("diffuseHip", max_depth=depth, albedo=albedo)
    rays = sum(dev.render("diffuseHip", max_depth=depth, albedo=albedo)["ray"] for _ in range(args.steps))
    log(f"pmc pass: {args.steps} fused frames, {rays} rays")
    # the batched kernel: spheres_1080p as one job of 16 spp (graph replays),
    # then the 32-scene workload, with the launch shapes bench_batched uses
    dev.batch_add_scene(mesh.idx, mesh.pos)
    dev.batch_set_jobs([{"scene": 0, "width": w, "height": h, "spp": 16, "max_depth": depth, "pass0": 0,
                         "albedo": albedo, "camera": cam}])
    dev.batch_set_tuning(0, 0, 32)
    for _ in range(3):
        dev.batch_render()
    log("pmc pass: 3 batched launches of spheres_1080p x 16 spp")
    dev.batch_reset()
    scenes = importlib.import_module(PKG + ".scenes")
    jobs = []
    for sp in scenes.batch_workload(args.batch_scenes):
        m = scenes.scene_mix(sp["seed"], sp["target_tris"])
        sid, _ = dev.batch_add_scene(m.idx, m.pos)
        v = sp["view"]
        jobs.append({"scene": sid, "width": sp["width"], "height": sp["height"], "spp": sp["spp"],
                     "max_depth": sp["max_depth"], "albedo": sp["albedo"],
                     "camera": pkg.camera_from_view(v["pos"], v["angle_x"], v["angle_y"], v["fov_y"],
                                                    sp["width"] / sp["height"])})
    dev.batch_set_jobs(jobs)
    dev.batch_set_tuning(2, 0, 32)
    for _ in range(2):
        dev.batch_render()
    log("pmc pass: 2 batched launches of the 32-scene workload")
    dev.close()
    # the albedo forward and its path-replay adjoint on the differentiable workload
    acfg = scenes.CONFIGS["heightfield_boxes_512"]
    amesh = scenes.build_scene("heightfield_boxes_512")
    av = acfg["view"]
    aw, ah = acfg["width"], acfg["height"]
    with pkg.HipDevice(0) as adev:
        adev.upload_scene(amesh.idx, amesh.pos)
        adev.build()
        adev.set_camera(pkg.camera_from_view(av["pos"], av["angle_x"], av["angle_y"], av["fov_y"], aw / ah))
        adev.init_frame(aw, ah)
        adev.set_albedo(scenes.albedo_set(amesh.tri_count, acfg["target_seed"]))
        target, _ = adev.render_albedo_image(0, acfg["spp"], acfg["max_depth"])
        adev.set_albedo(scenes.albedo_set(amesh.tri_count, acfg["albedo_seed"]))
        img, _ = adev.render_albedo_image(0, acfg["spp"], acfg["max_depth"])
        for _ in range(3):
            adev.adjoint_albedo(0, acfg["spp"], acfg["max_depth"], 2.0 * (img - target))
    log("pmc pass: 3 adjoint launches of heightfield_boxes_512 (8 spp)")
    src = torch.rand(COPY_FLOATS, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    log(f"pmc pass: 3 copies of {COPY_FLOATS * 4} bytes (read) + {COPY_FLOATS * 4} bytes (written)")
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-frames", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-city", action="store_true")
    ap.add_argument("--city-spp", type=int, default=64)
    ap.add_argument("--city-adjoint-spp", type=int, default=64)
    ap.add_argument("--batch-scenes", type=int, default=32)
    ap.add_argument("--pmc-pass", action="store_true",
                    help="short run for rocprofv3 --pmc: the fused frames and a calibration copy, no JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the path is built for one GPU "
              "(image tiles or whole scenes would shard across GPUs, which is not built)", file=sys.stderr)
        return 2

    sys.path.insert(0, ROOT)
    pkg = importlib.import_module(PKG)
    scenes = importlib.import_module(PKG + ".scenes")
    cfg = scenes.CONFIGS[WORKLOAD]
    w, h, depth, albedo = cfg["width"], cfg["height"], cfg["max_depth"], cfg["albedo"]
    mesh = scenes.build_scene(WORKLOAD)
    v = cfg["view"]
    cam = pkg.camera_from_view(v["pos"], v["angle_x"], v["angle_y"], v["fov_y"], w / h)

    if args.pmc_pass:
        return pmc_pass(pkg, mesh, cam, cfg, args)

    dev = pkg.HipDevice(0)
    log(f"device: {dev.info()}; workload {WORKLOAD}: {mesh.tri_count} triangles, {w}x{h}, depth {depth}")
    dev.upload_scene(mesh.idx, mesh.pos)
    build = dev.build()
    log(f"BVH8 build: {build['buildMs']:.2f} ms, {build['buildMprim']:.1f} Mprim/s, "
        f"{build['nodes']} nodes, depth {build['depth']}")
    dev.set_camera(cam)
    dev.init_frame(w, h)

    for _ in range(args.warmup):
        dev.render("diffuseHip", max_depth=depth, albedo=albedo)
    log(f"warmup done ({args.warmup} frames)")

    t0 = time.perf_counter()
    stats = [dev.render("diffuseHip", max_depth=depth, albedo=albedo) for _ in range(args.steps)]
    t1 = time.perf_counter()
    wall = t1 - t0
    rays = sum(s["ray"] for s in stats)
    value = rays / wall / 1e6
    dev_ms = sum(s["renderMs"] for s in stats)
    log(f"timed {args.steps} frames: {rays} rays in {wall * 1e3:.1f} ms wall ({dev_ms:.1f} ms device) "
        f"-> {value:.1f} Mrays/s")

    # the same frames with the frame lookahead off (each call starts its frame
    # from nothing and ends with its own tail)
    dev.set_lookahead(False)
    dev.set_pass(args.warmup)
    t2 = time.perf_counter()
    nstats = [dev.render("diffuseHip", max_depth=depth, albedo=albedo) for _ in range(args.steps)]
    no_look_wall = time.perf_counter() - t2
    dev.set_lookahead(True)
    assert sum(s["ray"] for s in nstats) == rays, "lookahead changed the ray count"
    no_look = rays / no_look_wall / 1e6
    log(f"same frames without the frame lookahead: {no_look:.1f} Mrays/s")

    # replay of the same passes: per-launch device time, then traversal counters
    def replay(timing: bool, counters: bool) -> list:
        dev.set_pass(args.warmup)
        dev.set_profiling(timing=timing, counters=counters)
        r = [dev.render("diffuseHip", max_depth=depth, albedo=albedo) for _ in range(args.steps)]
        dev.set_profiling()
        return r

    tstats = replay(True, False)
    cstats = replay(False, True)
    kernel_ms = sum(s["traceMs"] for s in tstats)
    crays = max(1, sum(s["ray"] for s in cstats))
    n_node = sum(s["nodes"] for s in cstats) / crays
    n_tri = sum(s["tris"] for s in cstats) / crays
    paths = sum(s["paths"] for s in stats)
    cont = rays - paths
    kbytes = fused_bytes(rays, paths, n_node, n_tri)
    achieved = kbytes / (kernel_ms * 1e-3) / 1e9
    wavefront_formula_bytes = wavefront_bytes(rays, cont, paths, n_node, n_tri)
    depth_rays = [sum(s["queue"][k] for s in cstats) // args.steps for k in range(depth + 1)]
    wsteps = max(1, sum(s["waveSteps"] for s in cstats))
    lsteps = sum(s["laneSteps"] for s in cstats)
    occupancy = {"lanes_per_step": round(lsteps / wsteps, 2), "steps_per_ray": round(lsteps / crays, 3),
                 "shade_passes_per_frame": sum(s["waveShades"] for s in cstats) // args.steps,
                 "spills_per_frame": sum(s["spills"] for s in cstats) // args.steps,
                 "drain_step_share": round(sum(s["drainWaveSteps"] for s in cstats) / wsteps, 3),
                 "drain_lanes_per_step": round(sum(s["drainLaneSteps"] for s in cstats)
                                               / max(1, sum(s["drainWaveSteps"] for s in cstats)), 2),
                 "drain_ms_per_frame": round(sum(s["drainMs"] for s in cstats) / args.steps, 4)}
    log(f"traversal loop: {occupancy['lanes_per_step']} of 64 lanes per step, "
        f"{occupancy['steps_per_ray']} steps per ray, {occupancy['shade_passes_per_frame']} shading passes "
        f"and {occupancy['spills_per_frame']} stack spills per frame; {occupancy['drain_step_share']} of the "
        f"steps after the work ran out, at {occupancy['drain_lanes_per_step']} lanes per step, "
        f"drain {occupancy['drain_ms_per_frame']} ms per frame")
    log(f"fused path kernel: {n_node:.2f} nodes/ray, {n_tri:.2f} tris/ray, {kernel_ms:.1f} ms over "
        f"{args.steps} launches ({achieved:.0f} GB/s algorithmic); rays per depth {depth_rays}")

    # the reference-structured wavefront pipeline on the same frames, for comparison
    dev.set_pipeline("wavefront")
    dev.set_pass(args.warmup)
    t2 = time.perf_counter()
    wstats = [dev.render("diffuseHip", max_depth=depth, albedo=albedo) for _ in range(args.steps)]
    wwall = time.perf_counter() - t2
    dev.set_pipeline("fused")
    wrays = sum(s["ray"] for s in wstats)
    assert wrays == rays, "wavefront and fused pipelines traced different ray counts"
    log(f"wavefront pipeline: {wrays / wwall / 1e6:.1f} Mrays/s")

    # batched form, config 2: the 16 spp of the same frame sequence in one launch
    batched = None
    if not args.no_batched:
        batched = bench_batched(pkg, scenes, dev, mesh, cam, cfg, args, value, wrays / wwall / 1e6, no_look)

    traffic = measured_traffic()
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "spheres_1080p: 256 icospheres (subdiv 4) + ground quad, 1,310,722 tris, "
                               "1920x1080, 1 spp per step, max depth 8, albedo 0.7 (diffuseHip, plain)",
                   "triangles": mesh.tri_count, "width": w, "height": h, "max_depth": depth, "albedo": albedo,
                   "pipeline": "fused"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic["traffic_bytes_per_launch"] if traffic else None,
                     "traffic_source": ("profiles/k_trace_fused_spheres_1080p.md: rocprofv3 --pmc FETCH_SIZE x "
                                        f"{traffic['fetch_scale']} + WRITE_SIZE x {traffic['write_scale']} per "
                                        "frame launch") if traffic else None,
                     "l2_hit_rate": traffic["l2_hit_rate"] if traffic else None,
                     "valu_busy": traffic["valu_busy"] if traffic else None,
                     "valu_lane_utilization": traffic["valu_lane_utilization"] if traffic else None,
                     "kernel": "k_trace<fused diffuse> (persistent path kernel: BVH8 closest hit + shading)",
                     "bytes_per_ray": round(kbytes / rays, 1),
                     "avg_launch_ms": round(kernel_ms / args.steps, 4)},
        "device_mrays": round(rays / (dev_ms * 1e-3) / 1e6, 2),
        "no_lookahead_mrays": round(no_look, 2),
        "wavefront_mrays": round(wrays / wwall / 1e6, 2),
        "frame_hbm_frac_wavefront_formula": round(wavefront_formula_bytes / (dev_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "nodes_per_ray": round(n_node, 3), "tris_per_ray": round(n_tri, 3), "traversal_occupancy": occupancy,
        "rays_per_frame": rays // args.steps, "rays_per_depth": depth_rays,
        "build": {"buildMs": round(build["buildMs"], 3), "buildMprim": round(build["buildMprim"], 2)},
        "batched": batched,
        "cpu_baseline": None,
    }
    dev.close()
    out["config1"] = bench_config1(pkg, scenes, args)
    if not args.no_adjoint:
        out["adjoint"] = bench_adjoint(pkg, scenes, args)
    if not args.no_city:
        out["city"] = bench_city(pkg, scenes, args)
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(scenes, mesh, cfg, cam, args.cpu_frames)
        log(f"cpu baseline: {out['cpu_baseline']['value']} Mrays/s on {out['cpu_baseline']['cores']} threads")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
