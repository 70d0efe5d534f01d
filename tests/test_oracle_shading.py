"""The oracle's full shading context (TriangleMesh::postIntersect with a
ShadingContext, geometry/triangle_mesh.cpp:59-153) pinned by geometry, not
by restating its formulas:

- a hit exactly on vertex k has the normal of vertex k (pins which
  barycentric weighs which vertex), and a hit on the midpoint of an edge the
  normalised sum of that edge's two normals;
- on an icosphere with exact vertex normals the interpolated normal is
  closer to the true sphere normal at the hit point than the facet normal;
- the normal is turned to the side of the geometric normal, and both face
  against the ray;
- with texture coordinates U is the direction of increasing u in the
  surface (finite differences of the planar mapping, in double), V = N x U,
  and the frame is orthonormal; a degenerate mapping and a mesh without
  attributes give makeFrame(N);
- the "Fast" variant ignores the attributes, and rendering with them only
  changes an image through the frame.
"""
import math

import numpy as np
import pytest


def _tri_scene(orc, p, n=None, t=None):
    s = orc.Scene(np.array([[0, 1, 2]], np.int32), np.asarray(p, np.float32))
    s.set_attributes(None if n is None else np.asarray(n, np.float32), None if t is None else np.asarray(t, np.float32))
    return s


P = [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]]          # Ng = +z
N3 = [[0.3, 0.0, 1.0], [0.0, 0.4, 1.0], [-0.2, -0.2, 1.0]]
DOWN = [0.1, 0.2, -1.0]                                             # ray towards -z hits the front face


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("k,uv", [(0, (0.0, 0.0)), (1, (1.0, 0.0)), (2, (0.0, 1.0))])
def test_vertex_hit_takes_that_vertex_normal(orc, k, uv):
    s = _tri_scene(orc, P, N3)
    f = s.shading_frame(0, uv[0], uv[1], DOWN)
    assert np.allclose(f[2], _unit(N3[k]), atol=1e-6)


def test_edge_midpoint_normal(orc):
    s = _tri_scene(orc, P, N3)
    for (a, b), uv in (((0, 1), (0.5, 0.0)), ((1, 2), (0.5, 0.5)), ((0, 2), (0.0, 0.5))):
        f = s.shading_frame(0, uv[0], uv[1], DOWN)
        assert np.allclose(f[2], _unit(np.add(N3[a], N3[b])), atol=1e-6)


def test_normal_sides(orc):
    # vertex normals against Ng: turned to the Ng side, then towards the ray
    s = _tri_scene(orc, P, [[0.2, 0.1, -1.0]] * 3)
    assert np.allclose(s.shading_frame(0, 0.2, 0.3, DOWN)[2], _unit([-0.2, -0.1, 1.0]), atol=1e-6)
    assert np.allclose(s.shading_frame(0, 0.2, 0.3, [0.0, 0.0, 1.0])[2], _unit([0.2, 0.1, -1.0]), atol=1e-6)
    # no attributes: the ray-facing geometric normal
    s = _tri_scene(orc, P)
    assert np.array_equal(s.shading_frame(0, 0.2, 0.3, DOWN)[2], [0, 0, 1])
    assert np.array_equal(s.shading_frame(0, 0.2, 0.3, [0, 0, 1])[2], [0, 0, -1])


def test_sphere_normals_beat_facets(orc, scenes):
    f, v = scenes.icosphere(2)
    pos = v.astype(np.float32)
    sc = orc.Scene(f, pos)
    sc.set_attributes(pos.copy(), None)           # exact unit-sphere normals
    flat = orc.Scene(f, pos)
    rng = np.random.default_rng(4)
    n = 4000
    org = rng.normal(size=(n, 3))
    org = 3.0 * org / np.linalg.norm(org, axis=1, keepdims=True)
    tgt = rng.normal(size=(n, 3)) * 0.3
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 4:7], rays[:, 7] = org, d, 3.4e38
    hits = sc.intersect_fp32(rays)
    ok = hits[:, 1].view(np.int32) >= 0
    assert ok.mean() > 0.9
    err_s, err_f = [], []
    for i in np.nonzero(ok)[0][:1500]:
        prim = int(hits[i, 1].view(np.int32))
        p = org[i] + float(hits[i, 0]) * d[i]
        true_n = p / np.linalg.norm(p)                     # outward; the ray comes from outside
        ns = sc.shading_frame(prim, hits[i, 2], hits[i, 3], rays[i, 4:7])[2]
        nf = flat.shading_frame(prim, hits[i, 2], hits[i, 3], rays[i, 4:7])[2]
        err_s.append(math.acos(min(1.0, float(np.dot(ns, true_n)))))
        err_f.append(math.acos(min(1.0, float(np.dot(nf, true_n)))))
    err_s, err_f = np.array(err_s), np.array(err_f)
    # facets of a subdivision-2 icosphere are ~0.16 rad across
    assert err_s.max() < 0.02 and err_f.mean() > 4 * err_s.mean()


def _planar_dpdu(p, t, uv_point):
    """Direction of increasing u at fixed v on the plane of the triangle, from
    the affine map (u, v) -> position fitted through the three corners."""
    p, t = np.asarray(p, np.float64), np.asarray(t, np.float64)
    A = np.column_stack([t, np.ones(3)])               # [u v 1] @ M = p
    M = np.linalg.solve(A, p)
    h = 1e-3
    q0 = np.array([uv_point[0], uv_point[1], 1.0]) @ M
    q1 = np.array([uv_point[0] + h, uv_point[1], 1.0]) @ M
    return _unit(q1 - q0)


@pytest.mark.parametrize("tex", [[[0, 0], [1, 0], [0, 1]],            # u = x
                                 [[0, 0], [0, 1], [-1, 0]],           # u = y
                                 [[0.2, 0.1], [2.2, 0.7], [-0.3, 1.6]]])  # a skewed, scaled mapping
def test_texcoord_frame(orc, tex):
    s = _tri_scene(orc, P, None, tex)
    for uv in ((0.2, 0.3), (0.6, 0.1)):
        f = s.shading_frame(0, uv[0], uv[1], DOWN).astype(np.float64)
        b = np.array([1 - uv[0] - uv[1], uv[0], uv[1]])
        tpt = b @ np.asarray(tex, np.float64)
        assert np.allclose(f[0], _planar_dpdu(P, tex, tpt), atol=1e-6)   # N = Ng here, so U = dp/du
        assert np.allclose(f[2], [0, 0, 1])
        assert np.allclose(f[1], np.cross(f[2], f[0]), atol=1e-6)
        assert np.allclose(f @ f.T, np.eye(3), atol=1e-6)


def test_texcoord_frame_with_normals_is_projected(orc):
    tex = [[0.2, 0.1], [2.2, 0.7], [-0.3, 1.6]]
    s = _tri_scene(orc, P, N3, tex)
    for uv in ((0.2, 0.3), (0.5, 0.25)):
        f = s.shading_frame(0, uv[0], uv[1], DOWN).astype(np.float64)
        b = np.array([1 - uv[0] - uv[1], uv[0], uv[1]])
        dpdu = _planar_dpdu(P, tex, b @ np.asarray(tex, np.float64))
        n = f[2]
        assert np.allclose(f[0], _unit(dpdu - np.dot(dpdu, n) * n), atol=1e-6)
        assert np.allclose(f @ f.T, np.eye(3), atol=1e-6)


def test_degenerate_mapping_and_fast(orc):
    plain = _tri_scene(orc, P, N3)                     # makeFrame(N) without texcoords
    same = _tri_scene(orc, P, N3, [[0.3, 0.3]] * 3)    # det = 0
    line = _tri_scene(orc, P, N3, [[0, 0], [1, 1], [2, 2]])  # det = 0 too
    for uv in ((0.2, 0.3), (0.7, 0.1)):
        ref = plain.shading_frame(0, *uv, DOWN)
        assert np.array_equal(same.shading_frame(0, *uv, DOWN), ref)
        assert np.array_equal(line.shading_frame(0, *uv, DOWN), ref)
        # makeFrame: orthonormal, N unchanged
        assert np.allclose(ref @ ref.T.astype(np.float64), np.eye(3), atol=1e-6)
    bare = _tri_scene(orc, P)
    for s in (plain, same):
        assert np.array_equal(s.shading_frame(0, 0.2, 0.3, DOWN, fast=True), bare.shading_frame(0, 0.2, 0.3, DOWN))


def test_render_with_attributes(orc, scenes):
    m = scenes.scene_smooth(cells=8, spheres=6, subdiv=2)
    v = scenes.SMOOTH_VIEW
    w, h = 48, 32
    cam = orc.camera_from_view(v["pos"], v["angle_x"], v["angle_y"], v["fov_y"], w / h)
    bare = orc.Scene(m.idx, m.pos)
    sm = orc.Scene(m.idx, m.pos)
    sm.set_attributes(m.normals, m.texcoords)
    for renderer in (0, 1, 2):
        a = sm.render(cam, w, h, spp=2, max_depth=4, albedo=0.8, renderer=renderer)
        b = bare.render(cam, w, h, spp=2, max_depth=4, albedo=0.8, renderer=renderer)
        if renderer == 0:
            # primary rays do not depend on the frame: same rays, other normals,
            # still unit vectors (one sample per pixel)
            assert a["rays"] == b["rays"] and not np.array_equal(a["accum"], b["accum"])
            one = sm.render(cam, w, h, spp=1, max_depth=0, albedo=0.8, renderer=0)["accum"]
            hit = (one[:, :, :3] != np.float32(0.05)).any(axis=2)
            assert hit.mean() > 0.2
            n = one[:, :, :3] * 2 - 1
            assert np.allclose(np.linalg.norm(n, axis=2)[hit], 1.0, atol=1e-5)
        else:
            assert not np.array_equal(a["accum"], b["accum"])
        # Fast ignores the attributes and keeps only the last pass
        fa = sm.render(cam, w, h, spp=2, max_depth=4, albedo=0.8, renderer=renderer | 0x100)
        last = bare.render(cam, w, h, spp=1, max_depth=4, albedo=0.8, renderer=renderer, pass0=1)
        assert np.array_equal(fa["accum"], last["accum"])
