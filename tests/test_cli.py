"""The `protoray render` command (main/render/render.cpp): option handling and
failure exits on the CPU; a full benchmark-mode render on the GPU."""
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "embree-benchmark-protoray_amd", "protoray")


def run(args, cwd):
    return subprocess.run([EXE] + args, cwd=cwd, capture_output=True, text=True, timeout=300)


@pytest.mark.parametrize("args,msg", [
    (["render", "s.mesh", "-size", "100,60"], "Resolution must be multiple of 8"),
    (["render", "s.mesh", "-dev", "cuda"], "Invalid device: cuda"),
    (["render", "-spp", "2"], "Scene not specified"),
    (["render", "s.mesh", "-view", "12"], "Invalid view"),
])
def test_cli_rejects_bad_options(tmp_path, args, msg):
    r = run(args, tmp_path)
    assert r.returncode == 1 and msg in r.stderr


def test_cli_usage(tmp_path):
    r = run([], tmp_path)
    assert r.returncode == 0 and "Usage" in r.stdout


@pytest.mark.gpu
def test_cli_benchmark_render(tmp_path, scenes, prt, orc):
    mesh = scenes.scene_heightfield_box(1001, cells=32)
    scenes.save_mesh(str(tmp_path / "hf.mesh"), mesh)
    v = scenes.CONFIGS["heightfield_box_512"]["view"]
    (tmp_path / "hf.view").write_bytes(scenes.views_to_bytes([dict(v, radius=0.0, focus=1.0)]))
    r = run(["render", "hf.mesh", "-dev", "hip", "-r", "diffuseHip", "-size", "64,48", "-spp", "6",
             "-maxDepth", "4", "-do", "-benchmark", "out"], tmp_path)
    assert r.returncode == 0, r.stderr
    stats = dict(line.strip().split("=", 1) for line in (tmp_path / "out.txt").read_text().splitlines())
    for key in ("stats_buildMs", "stats_buildMprim", "stats_mray", "stats_ray", "stats_renderMs", "stats_fps"):
        assert key in stats and float(stats[key]) >= 0.0
    # 6 spp with warm-up 6/6 = 1 -> 5 recorded frames in the CSV
    rows = (tmp_path / "out.csv").read_text().strip().splitlines()
    assert len(rows) == 1 + 5
    ppm = (tmp_path / "out.ppm").read_bytes()
    assert ppm.startswith(b"P6\n64 48\n255\n") and len(ppm) == len(b"P6\n64 48\n255\n") + 64 * 48 * 3
    # the image equals the oracle's accumulation, tonemapped (gamma 1/2.2, B,G,R bytes)
    cam = orc.camera_from_view(v["pos"], v["angle_x"], v["angle_y"], v["fov_y"], 64 / 48)
    ref = orc.Scene(mesh.idx, mesh.pos).render(cam, 64, 48, spp=6, max_depth=4, albedo=0.8, numerics=0)
    c = ref["accum"][:, :, 0] / ref["accum"][:, :, 3]
    expect = np.clip((np.power(c.astype(np.float32), np.float32(1 / 2.2)) * 255).astype(np.int32), 0, 255)
    got = np.frombuffer(ppm[len(b"P6\n64 48\n255\n"):], np.uint8).reshape(48, 64, 3)
    assert np.abs(got[:, :, 2].astype(int) - expect).max() <= 1
    assert (got[:, :, 0] == got[:, :, 2]).all()


def _parse_queries(text):
    out, cur = [], None
    for line in text.splitlines():
        if line.startswith("query "):
            cur = {}
        elif line == "end":
            out.append(cur)
            cur = None
        elif cur is not None and "=" in line:
            k, v = line.split("=", 1)
            cur[k] = v
    return out


def _nums(s):
    return np.array([float(x) for x in s.split(",")])


def _query_mesh(scenes, attributes):
    rng = np.random.default_rng(31)
    if attributes:
        idx, pos = scenes.heightfield(8, rng, 0.4)
        normals = scenes.vertex_normals(idx, pos)
        # a non-affine mapping, so dp/du differs from triangle to triangle
        tex = np.stack([(pos[:, 0] + 1.0) * 0.5 + 0.1 * pos[:, 2] ** 2, (pos[:, 2] + 1.0) * 0.5], 1)
        mats = (np.arange(len(idx)) % 2).astype(np.int32)
        return scenes.Mesh(idx, pos, mats, ["ground", "rock"], normals.astype(np.float32), tex.astype(np.float32))
    m = scenes.scene_convex_sphere(2)
    return scenes.Mesh(m.idx, m.pos, np.zeros(len(m.idx), np.int32), ["default"])


@pytest.mark.gpu
@pytest.mark.parametrize("attributes", [True, False])
def test_cli_query_pixel_and_ray(tmp_path, scenes, prt, orc, attributes):
    """DeviceHip::queryPixel / queryRay (render/renderer_single.h:137-163,
    render/renderer.h:57-81) through `protoray query`: on every pixel of a
    small frame, the ShadingContext fields against the oracle's FP64 hit and
    shading frame; misses give an empty result."""
    mesh = _query_mesh(scenes, attributes)
    scenes.save_mesh(str(tmp_path / "q.mesh"), mesh)
    view = {"pos": [0.1, 1.6, 2.2], "angle_x": 0.6, "angle_y": 0.1, "fov_y": math.radians(55.0)}
    (tmp_path / "q.view").write_bytes(scenes.views_to_bytes([dict(view, radius=0.0, focus=1.0)]))
    w, h = 32, 24
    pixels = [(x, y) for y in range(h) for x in range(w)]
    rng = np.random.default_rng(5)
    free = np.zeros((40, 8), np.float32)
    free[:, 0:3] = rng.uniform(-0.5, 0.5, (40, 3)) + np.array([0.0, 1.5, 0.0])
    dvec = rng.normal(size=(40, 3)) * np.array([0.3, 1.0, 0.3]) - np.array([0.0, 1.0, 0.0])
    free[:, 4:7] = dvec / np.linalg.norm(dvec, axis=1, keepdims=True)
    free[:, 7] = np.float32(3.402823466e38)
    args = ["query", "q.mesh", "-size", f"{w},{h}"]
    for x, y in pixels:
        args += ["-pixel", f"{x},{y}"]
    for r in free:
        args += ["-ray", ",".join("%.9g" % c for c in (*r[0:3], *r[4:7]))]
    res = run(args, tmp_path)
    assert res.returncode == 0, res.stderr
    got = _parse_queries(res.stdout)
    assert len(got) == len(pixels) + len(free)

    # the rays DeviceHip traces: pinhole camera of the view, pixel centres
    cam = np.asarray(prt.camera_from_view(view["pos"], view["angle_x"], view["angle_y"], view["fov_y"], w / h),
                     np.float32)
    org, io, dx, dy = cam[0:3], cam[3:6], cam[6:9], cam[9:12]

    def pixel_dir(ix, iy):
        d = (io + np.float32(ix) * dx) + np.float32(iy) * dy
        inv = np.float32(1.0) / np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
        return d * inv

    rays = np.zeros((len(pixels), 8), np.float32)
    for i, (x, y) in enumerate(pixels):
        ix = (np.float32(x) + np.float32(0.5)) / np.float32(w)
        iy = (np.float32(y) + np.float32(0.5)) / np.float32(h)
        rays[i, 0:3] = org
        rays[i, 4:7] = pixel_dir(ix, iy)
        rays[i, 7] = np.float32(3.402823466e38)
    centre = pixel_dir(np.float32(0.5), np.float32(0.5)).astype(np.float64)
    all_rays = np.concatenate([rays, free])
    sc = orc.Scene(mesh.idx, mesh.pos)
    if attributes:
        sc.set_attributes(mesh.normals, mesh.texcoords)
    ref = sc.intersect_fp64(all_rays)
    n_hit = n_checked = 0
    for i, q in enumerate(got):
        if ref["prim"][i] < 0:
            if q:  # only a ray grazing a silhouette edge may hit in fp32 and miss in FP64
                d_edge = sc.ray_edge_distance(all_rays[i:i + 1], np.array([int(q["primId"])], np.int32))[0]
                assert d_edge <= 1e-4, f"query {i}: a miss must give an empty result, got {q}"
            continue
        n_hit += 1
        if ref["edge"][i] <= 1e-4:  # too close to an edge for the FP64 primId to decide
            continue
        n_checked += 1
        prim = int(ref["prim"][i])
        assert int(q["primId"]) == prim, f"query {i}"
        t64, u64, v64 = ref["t"][i], ref["u"][i], ref["v"][i]
        dist = float(q["dist"])
        assert abs(dist - t64) <= 1e-6 * t64, f"query {i}: dist {dist} vs {t64}"
        o = all_rays[i, 0:3].astype(np.float64)
        dirv = all_rays[i, 4:7].astype(np.float64)
        p = _nums(q["p"])
        assert np.abs(p - (o + t64 * dirv)).max() <= 1e-5
        tri = mesh.pos[mesh.idx[prim]].astype(np.float64)
        ng = np.cross(tri[1] - tri[0], tri[2] - tri[0])
        ng /= np.linalg.norm(ng)
        side = -1.0 if ng @ dirv > 0 else 1.0
        assert np.abs(_nums(q["Ng"]) - side * ng).max() <= 1e-5
        frame = sc.shading_frame(prim, u64, v64, all_rays[i, 4:7])
        assert np.abs(_nums(q["N"]) - frame[2]).max() <= 2e-5, f"query {i}: N"
        uv = _nums(q["uv"])
        assert uv.shape == (2,)
        if attributes:
            b = np.array([1.0 - u64 - v64, u64, v64])
            uv_ref = b @ mesh.texcoords[mesh.idx[prim]].astype(np.float64)
        else:
            uv_ref = np.array([u64, v64])
        assert np.abs(uv - uv_ref).max() <= 1e-5, f"query {i}: uv"
        eps_ref = max(t64, np.abs(o + t64 * dirv).max()) * float.fromhex("0x1.fp-18")
        assert abs(float(q["eps"]) - eps_ref) <= 1e-5 * eps_ref
        mat = int(mesh.material_ids[prim])
        assert int(q["matId"]) == mat and q["mat"] == mesh.material_names[mat]
        if i < len(pixels):
            assert abs(float(q["depth"]) - t64 * float(dirv @ centre)) <= 1e-5 * t64
        else:
            assert "depth" not in q
    assert n_checked >= 100 and n_hit < len(got), (n_checked, n_hit)
