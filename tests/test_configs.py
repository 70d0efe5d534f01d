"""The recorded workloads: configs.json and the .view files match what the
generators produce today (scene bytes by SHA-256, parameters, cameras)."""
import hashlib
import json
import math
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CFG_DIR = os.path.join(ROOT, "embree-benchmark-protoray_amd", "configs")


@pytest.fixture(scope="module")
def record():
    with open(os.path.join(CFG_DIR, "configs.json")) as f:
        return json.load(f)


def _sha(scenes, mesh):
    return hashlib.sha256(scenes.mesh_to_bytes(mesh)).hexdigest()


def test_parameters_match_code(record, scenes):
    assert set(record["workloads"]) == set(scenes.CONFIGS)
    for name, cfg in scenes.CONFIGS.items():
        rec = record["workloads"][name]
        for k, v in cfg.items():
            assert rec[k] == v, (name, k)
    jobs = scenes.batch_workload()
    assert len(record["batch"]["jobs"]) == len(jobs) == 32
    for rj, j in zip(record["batch"]["jobs"], jobs):
        for k in ("seed", "target_tris", "width", "height", "spp", "max_depth", "albedo"):
            assert rj[k] == j[k]
        assert 1 <= rj["spp"] <= 16 and (rj["width"], rj["height"]) in scenes.BATCH_RESOLUTIONS
        assert 1e3 <= rj["target_tris"] <= 2e6


@pytest.mark.parametrize("name", ["heightfield_box_512", "heightfield_boxes_512", "spheres_1080p", "city_4k"])
def test_scene_bytes_match_record(record, scenes, name):
    rec = record["workloads"][name]
    mesh = scenes.build_scene(name)
    assert mesh.tri_count == rec["triangles"]
    assert _sha(scenes, mesh) == rec["mesh_sha256"]


def test_headline_triangle_counts(record):
    w = record["workloads"]
    assert w["heightfield_box_512"]["triangles"] == 8202        # 64x64 cells + 5 quads
    assert w["spheres_1080p"]["triangles"] == 1310722           # 256 x 5120 + 2
    assert w["heightfield_boxes_512"]["triangles"] == 33536     # 128x128 cells + 64 boxes
    assert w["city_4k"]["triangles"] == 15960000                # 1.33 M boxes x 12


def test_batch_scene_bytes_match_record(record, scenes):
    for rj in record["batch"]["jobs"][:8]:
        mesh = scenes.scene_mix(rj["seed"], rj["target_tris"])
        assert mesh.tri_count == rj["triangles"]
        assert _sha(scenes, mesh) == rj["mesh_sha256"]


def test_view_files(record, scenes):
    for name, cfg in scenes.CONFIGS.items():
        with open(os.path.join(CFG_DIR, record["workloads"][name]["view_file"]), "rb") as f:
            v = scenes.views_from_bytes(f.read())[0]
        assert all(math.isclose(a, b, rel_tol=1e-6, abs_tol=1e-7) for a, b in zip(v["pos"], cfg["view"]["pos"]))
        for k in ("angle_x", "angle_y", "fov_y"):
            assert math.isclose(v[k], cfg["view"][k], rel_tol=1e-6, abs_tol=1e-7)
