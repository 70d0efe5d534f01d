"""`protoray build`: OBJ -> .mesh (main/build/build.cpp:31-130, main/build/obj_loader.cpp,
geometry/triangle_mesh_builder.cpp).  CPU only.

Checked against what the OBJ file states, not against the builder's code:
the triangle set (as vertex positions) equals the file's polygons split as
fans, shared vertices are merged, indices follow OBJ rules (1-based,
negative = relative), materials get ids in order of first use after the
reserved background 0 ("Default" for faces before any usemtl), missing
normals of a mesh with normals become the face's geometric normal, and the
layout's top-level split separates the vertices at the median of the longest
axis.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "embree-benchmark-protoray_amd", "protoray")

CUBE = """# unit cube, quads
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 0 0 1
v 1 0 1
v 1 1 1
v 0 1 1
f 1 4 3 2
f 5 6 7 8
f 1 2 6 5
f 2 3 7 6
f 3 4 8 7
f 4 1 5 8
"""


def build(tmp_path, text, name="in.obj", extra=()):
    src = tmp_path / name
    src.write_text(text)
    out = tmp_path / "out.mesh"
    r = subprocess.run([CLI, "build", str(src), "-o", str(out), *extra], capture_output=True, text=True, timeout=60)
    return r, out


def tri_positions(mesh):
    return np.sort(mesh.pos[mesh.idx].reshape(-1, 9).astype(np.float64), axis=0)


def fan(verts, faces):
    out = []
    for f in faces:
        for i in range(len(f) - 2):
            out.append(np.concatenate([verts[f[0] - 1], verts[f[i + 1] - 1], verts[f[i + 2] - 1]]))
    return np.sort(np.array(out, np.float64), axis=0)


@pytest.fixture(scope="module", autouse=True)
def need_cli():
    if not os.path.exists(CLI):
        pytest.skip("protoray CLI not built")


def test_cube_quads(tmp_path, scenes):
    r, out = build(tmp_path, CUBE)
    assert r.returncode == 0, r.stderr
    m = scenes.load_mesh(str(out))
    assert m.tri_count == 12 and m.pos.shape[0] == 8
    verts = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    faces = [[1, 4, 3, 2], [5, 6, 7, 8], [1, 2, 6, 5], [2, 3, 7, 6], [3, 4, 8, 7], [4, 1, 5, 8]]
    assert np.array_equal(tri_positions(m), fan(verts, faces))
    assert np.array_equal(m.bounds(), np.array([0, 0, 0, 1, 1, 1], np.float32))
    # faces before any usemtl use "Default" (id 1; 0 is the background)
    assert m.material_names == ["", "Default"] and (m.material_ids == 1).all()
    assert m.normals is None and m.texcoords is None
    # every triangle keeps its winding: its normal points out of the cube
    p = m.pos[m.idx].astype(np.float64)
    n = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    c = p.mean(axis=1) - 0.5
    assert (np.sum(n * c, axis=1) > 0).all()


def test_negative_indices_equal_positive(tmp_path, scenes):
    # all 8 vertices precede the faces, so index i is i - 9 counted from the end
    body = [ln for ln in CUBE.splitlines() if not ln.startswith("f ")]
    faces = ["f " + " ".join(str(int(t) - 9) for t in ln.split()[1:])
             for ln in CUBE.splitlines() if ln.startswith("f ")]
    neg = "\n".join(body + faces) + "\n"
    r1, o1 = build(tmp_path, CUBE, "a.obj")
    m1 = scenes.load_mesh(str(o1))
    r2, o2 = build(tmp_path, neg, "b.obj")
    assert r2.returncode == 0, r2.stderr
    m2 = scenes.load_mesh(str(o2))
    assert np.array_equal(tri_positions(m1), tri_positions(m2))


def test_normals_texcoords_and_materials(tmp_path, scenes):
    text = """v 0 0 0
v 2 0 0
v 0 2 0
v 2 2 0
vn 0 0 1
vt 0.25 0.5
vt 1 0
usemtl red
f 1/1/1 2/2/1 3/1/1
usemtl blue
f 2//1 4//1 3//1
usemtl red
f 1 3 4
"""
    r, out = build(tmp_path, text)
    assert r.returncode == 0, r.stderr
    m = scenes.load_mesh(str(out))
    assert m.material_names == ["", "red", "blue"]
    assert m.tri_count == 3
    pos = m.pos[m.idx]
    by_first = {tuple(np.round(t.astype(np.float64).mean(0), 4)): k for k, t in enumerate(pos)}
    # materials follow the faces
    t_red1 = np.array([[0, 0, 0], [2, 0, 0], [0, 2, 0]], float).mean(0)
    t_blue = np.array([[2, 0, 0], [2, 2, 0], [0, 2, 0]], float).mean(0)
    t_red2 = np.array([[0, 0, 0], [0, 2, 0], [2, 2, 0]], float).mean(0)
    for c, mat in ((t_red1, 1), (t_blue, 2), (t_red2, 1)):
        assert m.material_ids[by_first[tuple(np.round(c, 4))]] == mat
    assert m.normals is not None and m.texcoords is not None
    # the face without normals gets its geometric normal cross(p1-p0, p2-p0):
    # (0,2,0)-(0,0,0) x (2,2,0)-(0,0,0) = (0,0,-4)
    k = by_first[tuple(np.round(t_red2, 4))]
    nrm = m.normals[m.idx[k]]
    assert np.array_equal(nrm, np.tile([0, 0, -4], (3, 1)).astype(np.float32))
    # distinct (position, normal, texcoord) combinations are distinct vertices
    combos = {tuple(np.concatenate([m.pos[i], m.normals[i], m.texcoords[i]])) for i in range(m.pos.shape[0])}
    assert len(combos) == m.pos.shape[0]


def test_degenerate_dropped_and_errors(tmp_path, scenes):
    r, out = build(tmp_path, "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\nf 1 1 2\n")
    assert r.returncode == 0 and "Removed 1 degenerate" in r.stdout
    assert scenes.load_mesh(str(out)).tri_count == 1
    r, _ = build(tmp_path, "v 0 0 0\nv 1 0 0\nf 1 2\n")
    assert r.returncode == 1 and "Invalid face" in r.stderr
    r, _ = build(tmp_path, "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 9\n")
    assert r.returncode == 1 and "Invalid vertex reference" in r.stderr
    r, _ = build(tmp_path, "v 0 0\n")
    assert r.returncode == 1 and "Invalid vertex position" in r.stderr
    r, _ = build(tmp_path, "v 0 0 0\n", name="in.ply")
    assert r.returncode == 1 and "Unsupported input format" in r.stderr


def test_layout_median_split(tmp_path, scenes):
    rng = np.random.default_rng(5)
    n = 400
    pts = rng.uniform([-3, -1, -0.5], [3, 1, 0.5], size=(n, 3))   # longest axis: x
    lines = [f"v {x:.6f} {y:.6f} {z:.6f}" for x, y, z in pts]
    faces = rng.integers(1, n + 1, size=(600, 3))
    faces = faces[(faces[:, 0] != faces[:, 1]) & (faces[:, 0] != faces[:, 2]) & (faces[:, 1] != faces[:, 2])]
    lines += [f"f {a} {b} {c}" for a, b, c in faces]
    r, out = build(tmp_path, "\n".join(lines) + "\n")
    assert r.returncode == 0, r.stderr
    m = scenes.load_mesh(str(out))
    # only vertices some face uses enter the mesh
    used = len(np.unique(faces))
    assert m.pos.shape[0] == used and m.tri_count == len(faces)
    printed = np.array([[float(f"{c:.6f}") for c in p] for p in pts], np.float32).astype(np.float64)
    assert np.array_equal(tri_positions(m), fan(printed, faces.tolist()))
    # top-level split at the median vertex along x: it is the smallest of the
    # second half (the halves are reordered further inside themselves)
    x = m.pos[:, 0]
    split = x[used // 2:].min()
    assert x[: used // 2].max() <= split
    centres = m.pos[m.idx].astype(np.float32).sum(axis=1) * np.float32(1.0 / 3.0)
    below = centres[:, 0] < split
    k = int(below.sum())
    assert below[:k].all() and not below[k:].any()


def test_mesh_input_round_trip(tmp_path, scenes):
    mesh = scenes.scene_heightfield_box(1001, cells=4)
    src = tmp_path / "in.mesh"
    scenes.save_mesh(str(src), mesh)
    out = tmp_path / "copy.mesh"
    r = subprocess.run([CLI, "build", str(src), "-o", str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert out.read_bytes() == src.read_bytes()
