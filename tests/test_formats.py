"""ProtoRay file formats: .mesh blob, .view set, P6 PPM."""
import os
import struct

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_mesh_fixture_reads_field_by_field(scenes):
    with open(os.path.join(GOLDEN, "two_tris.mesh"), "rb") as f:
        data = f.read()
    m = scenes.mesh_from_bytes(data)
    assert m.idx.tolist() == [[0, 1, 2], [0, 2, 3]]
    assert m.pos.tolist() == [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0.5]]
    assert m.material_names == ["default"] and m.material_ids.tolist() == [0, 0]
    assert scenes.mesh_to_bytes(m) == data


def test_mesh_roundtrip_with_attributes(scenes, tmp_path):
    rng = np.random.default_rng(0)
    m = scenes.Mesh(rng.integers(0, 5, (7, 3)).astype(np.int32), rng.random((5, 3)).astype(np.float32),
                    np.arange(7, dtype=np.int32), ["a", "bb"], rng.random((5, 3)).astype(np.float32),
                    rng.random((5, 2)).astype(np.float32))
    p = tmp_path / "m.mesh"
    scenes.save_mesh(str(p), m)
    r = scenes.load_mesh(str(p))
    assert np.array_equal(r.idx, m.idx) and np.array_equal(r.pos, m.pos)
    assert np.array_equal(r.normals, m.normals) and np.array_equal(r.texcoords, m.texcoords)
    assert r.material_names == ["a", "bb"]


def test_mesh_errors(scenes):
    with open(os.path.join(GOLDEN, "two_tris.mesh"), "rb") as f:
        data = f.read()
    with pytest.raises(RuntimeError, match="wrong blob ID"):
        scenes.mesh_from_bytes(b"\0\0\0\0" + data[4:])
    with pytest.raises(RuntimeError, match="unexpected end of stream"):
        scenes.mesh_from_bytes(data[:60])


def test_view_roundtrip(scenes):
    v = [{"pos": [1.0, 2.0, 3.0], "angle_x": 0.5, "angle_y": -0.25, "fov_y": 0.75, "radius": 0.0, "focus": 2.0}]
    b = scenes.views_to_bytes(v)
    assert len(b) == 320
    r = scenes.views_from_bytes(b)
    assert r[0] == v[0] and r[9] == v[0]


@pytest.mark.skipif(not os.path.isdir("/root/reference/benchmark"), reason="reference checkout not present")
def test_reference_view_files_parse(scenes):
    d = "/root/reference/benchmark"
    names = [n for n in os.listdir(d) if n.endswith(".view")]
    assert names
    for n in names:
        with open(os.path.join(d, n), "rb") as f:
            views = scenes.views_from_bytes(f.read())
        for v in views:
            assert all(np.isfinite(v["pos"])) and 0.0 < v["fov_y"] < np.pi


def test_ppm_bytes(scenes):
    px = np.array([[1 | (2 << 8) | (3 << 16), 255]], dtype=np.int32)
    b = scenes.ppm_bytes(px)
    assert b == b"P6\n2 1\n255\n" + bytes([1, 2, 3, 255, 0, 0])


def test_generated_scene_sizes(scenes):
    assert scenes.scene_heightfield_box(1001).tri_count == 8202
    f, v = scenes.icosphere(4)
    assert f.shape == (5120, 3) and v.shape == (2562, 3)
    assert np.allclose(np.linalg.norm(v, axis=1), 1.0)
    # every edge is shared by exactly two faces (closed surface)
    e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
    _, cnt = np.unique(e, axis=0, return_counts=True)
    assert (cnt == 2).all()
