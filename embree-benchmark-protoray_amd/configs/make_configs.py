#!/usr/bin/env python3
"""Writes the workload record: configs.json and one ProtoRay .view per workload.

configs.json holds, for every workload of scenes.CONFIGS and for the batched
workload, the generator, its seed, the camera view, image size, spp, max depth
and albedo, plus the SHA-256 of the generated scene as a .mesh file, so that a
change to a generator (or to the numpy RNG underneath it) is caught by
tests/test_configs.py instead of silently changing a measured workload.
The .view files use the reference's ViewSet layout (main/render/view.h:25-40)
and can be passed to `protoray render -view`.

Run from the repository root:  python embree-benchmark-protoray_amd/configs/make_configs.py
"""
from __future__ import annotations

import hashlib
import importlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def mesh_sha256(scenes, mesh) -> str:
    return hashlib.sha256(scenes.mesh_to_bytes(mesh)).hexdigest()


def record(scenes) -> dict:
    out = {"workloads": {}, "batch": {}}
    for name, cfg in scenes.CONFIGS.items():
        mesh = scenes.build_scene(name)
        entry = dict(cfg)
        entry["triangles"] = mesh.tri_count
        entry["vertices"] = int(mesh.pos.shape[0])
        entry["mesh_sha256"] = mesh_sha256(scenes, mesh)
        entry["view_file"] = f"{name}.view"
        out["workloads"][name] = entry
    jobs = scenes.batch_workload()
    batch = []
    for job in jobs:
        mesh = scenes.scene_mix(job["seed"], job["target_tris"])
        j = {k: v for k, v in job.items() if k != "view"}
        j["triangles"] = mesh.tri_count
        j["mesh_sha256"] = mesh_sha256(scenes, mesh)
        batch.append(j)
    out["batch"] = {"view": scenes.BATCH_VIEW, "view_file": "batch.view", "jobs": batch}
    return out


def main() -> int:
    sys.path.insert(0, ROOT)
    scenes = importlib.import_module("embree-benchmark-protoray_amd.scenes")
    rec = record(scenes)
    with open(os.path.join(HERE, "configs.json"), "w") as f:
        json.dump(rec, f, indent=1, sort_keys=True)
        f.write("\n")
    for name, cfg in scenes.CONFIGS.items():
        with open(os.path.join(HERE, f"{name}.view"), "wb") as f:
            f.write(scenes.views_to_bytes([cfg["view"]]))
    with open(os.path.join(HERE, "batch.view"), "wb") as f:
        f.write(scenes.views_to_bytes([scenes.BATCH_VIEW]))
    print(f"wrote configs.json and {len(scenes.CONFIGS) + 1} view files to {HERE}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
