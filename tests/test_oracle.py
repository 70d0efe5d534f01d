"""Pins the CPU oracle against things it does not compute itself: published
MurmurHash3 vectors, the LCG's closed-form jump-ahead, IEEE conversions in
numpy, float64 trigonometry, a float64 brute-force intersector written here in
numpy, and scenes whose images are known exactly."""
import json
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
M32 = 0xFFFFFFFF


# --- MurmurHash3_x86_32 (Austin Appleby, public domain), restated here in Python
def _rotl(x, r):
    return ((x << r) | (x >> (32 - r))) & M32


def py_mix(h, k):
    k = (k * 0xCC9E2D51) & M32
    k = _rotl(k, 15)
    k = (k * 0x1B873593) & M32
    h ^= k
    return (_rotl(h, 13) * 5 + 0xE6546B64) & M32


def py_fmix(h):
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & M32
    h ^= h >> 13
    h = (h * 0xC2B2AE35) & M32
    h ^= h >> 16
    return h


def murmur3_32(data: bytes, seed: int) -> int:
    h = seed & M32
    nblocks = len(data) // 4
    for i in range(nblocks):
        h = py_mix(h, int.from_bytes(data[4 * i:4 * i + 4], "little"))
    tail = data[4 * nblocks:]
    k = 0
    for j in reversed(range(len(tail))):
        k = (k << 8) | tail[j]
    if tail:
        k = (k * 0xCC9E2D51) & M32
        k = _rotl(k, 15)
        k = (k * 0x1B873593) & M32
        h ^= k
    h ^= len(data)
    return py_fmix(h)


@pytest.mark.parametrize("data,seed,expect", [
    (b"", 0, 0x00000000),
    (b"", 1, 0x514E28B7),
    (b"", 0xFFFFFFFF, 0x81F16F39),
    (b"\0\0\0\0", 0, 0x2362F9DE),
    (b"test", 0, 0xBA6BD213),
    (b"aaaa", 0x9747B28C, 0x5A97808A),
    (b"abcd", 0x9747B28C, 0xF0478627),
    (b"Hello, world!", 0x9747B28C, 0x24884CBA),
    (b"The quick brown fox jumps over the lazy dog", 0x9747B28C, 0x2FA826CD),
])
def test_murmur3_published_vectors(data, seed, expect):
    assert murmur3_32(data, seed) == expect


def test_sampler_seed_is_murmur_mix_of_pixel_then_pass(orc):
    rng = np.random.default_rng(0)
    for pixel, p in zip(rng.integers(0, 1 << 31, 64), rng.integers(0, 1 << 20, 64)):
        expect = py_fmix(py_mix(py_mix(0, int(pixel)), int(p)))
        assert orc.sampler_seed(int(p), int(pixel)) == expect
    # negative ints wrap as uint32 (murmurHash3Mix<int>, math/math.h:138-146)
    assert orc.sampler_seed(-1, 5) == py_fmix(py_mix(py_mix(0, 5), M32))


def test_lcg_matches_jump_ahead_closed_form(orc):
    a, c = 1664525, 1013904223
    for p, pixel in [(0, 0), (3, 12345), (999, 2 ** 20 + 7)]:
        s0 = orc.sampler_seed(p, pixel)
        states, vals = orc.sampler_sequence(p, pixel, 40)
        for n in (1, 2, 7, 40):
            an = pow(a, n)
            expect = (an * s0 + c * (an - 1) // (a - 1)) & M32
            assert int(states[n - 1]) == expect
        # u32 -> float: round-to-nearest conversion times 2^-32 (math/math.h:179)
        expect_vals = states.astype(np.float32) * np.float32(2.0 ** -32)
        assert np.array_equal(vals.view(np.uint32), expect_vals.view(np.uint32))


def test_sampler_golden_table(orc):
    with open(os.path.join(GOLDEN, "sampler_states.json")) as f:
        table = json.load(f)
    for row in table["rows"]:
        states, _ = orc.sampler_sequence(row["pass"], row["pixel"], len(row["states"]))
        assert [int(s) for s in states] == row["states"]


def test_unorm_can_reach_one(orc):
    # round-to-nearest maps the top 128 states to exactly 1.0f
    assert np.float32(np.uint32(0xFFFFFFFF)) * np.float32(2.0 ** -32) == np.float32(1.0)


def test_sincos_2pi_accuracy(orc):
    v = np.concatenate([np.linspace(0, 1, 100001, dtype=np.float32),
                        np.random.default_rng(1).random(100000, dtype=np.float32)])
    s, c = orc.sincos_2pi(v)
    ref_s = np.sin(2 * np.pi * v.astype(np.float64))
    ref_c = np.cos(2 * np.pi * v.astype(np.float64))
    assert np.abs(s - ref_s).max() < 4e-7
    assert np.abs(c - ref_c).max() < 4e-7
    assert s[0] == 0.0 and c[0] == 1.0


def _rot(vec, axis, a):
    return vec * math.cos(a) + axis * np.dot(axis, vec) * (1 - math.cos(a)) + np.cross(vec, axis) * math.sin(a)


@pytest.mark.parametrize("ax,ay,fov,aspect", [(0.0, 0.0, 0.785, 1.0), (1.3, 0.0, 0.785, 1.0),
                                              (0.25, 0.7, 1.1, 16 / 9), (-0.4, -2.0, 0.3, 0.5)])
def test_camera_closed_form(orc, ax, ay, fov, aspect):
    cam = orc.camera_from_view([0.5, 1.0, 2.0], ax, ay, fov, aspect).astype(np.float64)
    U, V, N = np.eye(3)
    N1 = _rot(N, V, ay)
    U1 = np.cross(V, N1)
    V2 = _rot(V, U1, ax)
    N2 = np.cross(U1, V2)
    hh = math.tan(0.5 * fov)
    hw = hh * aspect
    B = np.stack([U1, V2, N2], axis=1)
    assert np.allclose(cam[0:3], [0.5, 1.0, 2.0])
    assert np.allclose(cam[3:6], B @ [-hw, hh, -1.0], atol=2e-6)
    assert np.allclose(cam[6:9], B @ [2 * hw, 0, 0], atol=2e-6)
    assert np.allclose(cam[9:12], B @ [0, -2 * hh, 0], atol=2e-6)
    # the image centre looks along -N
    centre = cam[3:6] + 0.5 * cam[6:9] + 0.5 * cam[9:12]
    assert np.allclose(centre / np.linalg.norm(centre), -N2, atol=2e-6)


def test_raygen_closed_form(orc):
    cam = orc.camera_from_view([0, 0, 0], 0.0, 0.0, math.radians(60), 2.0)
    w, h = 32, 16
    rays, states = orc.raygen(cam, w, h, 3)
    for pid in (0, 17, w * h - 1):
        x, y = pid % w, pid // w
        _, vals = orc.sampler_sequence(3, pid, 4)
        ix, iy = (x + float(vals[0])) / w, (y + float(vals[1])) / h
        d = cam[3:6].astype(np.float64) + ix * cam[6:9] + iy * cam[9:12]
        d /= np.linalg.norm(d)
        assert np.allclose(rays[pid, 4:7], d, atol=1e-6)
        st, _ = orc.sampler_sequence(3, pid, 4)
        assert states[pid] == st[3]
    assert np.allclose(np.linalg.norm(rays[:, 4:7], axis=1), 1.0, atol=1e-6)


def brute_force_f64(tris, org, d):
    """Moller-Trumbore in float64 over every triangle; closest (t, id)."""
    p0, p1, p2 = tris[:, 0], tris[:, 1], tris[:, 2]
    e1, e2 = p1 - p0, p2 - p0
    pv = np.cross(d, e2)
    det = np.einsum("ij,j->i", e1, pv) if pv.ndim == 1 else np.sum(e1 * pv, 1)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / det
        tv = org - p0
        u = np.sum(tv * pv, 1) * inv
        qv = np.cross(tv, e1)
        v = (qv @ d) * inv
        t = np.sum(e2 * qv, 1) * inv
    ok = (det != 0) & (u >= 0) & (v >= 0) & (u + v <= 1) & (t >= 0)
    if not ok.any():
        return -1, -1.0
    tt = np.where(ok, t, np.inf)
    i = int(np.argmin(tt))
    return i, float(tt[i])


def test_fp64_intersector_matches_brute_force(orc, scenes):
    rng = np.random.default_rng(11)
    T = 300
    c = rng.uniform(-1, 1, size=(T, 1, 3))
    tris = (c + 0.15 * rng.normal(size=(T, 3, 3))).astype(np.float32)
    idx = np.arange(3 * T, dtype=np.int32).reshape(T, 3)
    pos = tris.reshape(-1, 3)
    sc = orc.Scene(idx, pos)
    n = 3000
    org = rng.uniform(-2, 2, size=(n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = org
    rays[:, 4:7] = d
    rays[:, 7] = np.float32(3.4e38)
    r = sc.intersect_fp64(rays)
    t64 = tris.astype(np.float64)
    hits = 0
    for i in range(n):
        bi, bt = brute_force_f64(t64, rays[i, 0:3].astype(np.float64), rays[i, 4:7].astype(np.float64))
        if bi != r["prim"][i]:
            assert r["edge"][i] < 1e-9 or (bi >= 0 and r["prim"][i] >= 0 and abs(bt - r["t"][i]) < 1e-9), i
        elif bi >= 0:
            hits += 1
            assert abs(bt - r["t"][i]) <= 1e-12 * max(1.0, bt)
    assert hits > 100
    # the fp32 kernel-numerics intersector agrees with FP64 away from edges
    h32 = sc.intersect_fp32(rays)
    p32 = h32[:, 1].view(np.int32)
    diff = p32 != r["prim"]
    assert (r["edge"][diff] < 1e-5).all()


def test_oracle_watertight_and_tie(orc, scenes):
    mesh = scenes.scene_convex_sphere(3)
    sc = orc.Scene(mesh.idx, mesh.pos)
    pos = mesh.pos.astype(np.float64)
    tgt = np.concatenate([pos] + [0.5 * (pos[mesh.idx[:, a]] + pos[mesh.idx[:, b]]) for a, b in [(0, 1), (1, 2), (2, 0)]])
    # directions that enter the sphere at the target (not supporting lines)
    g = np.random.default_rng(4).normal(size=tgt.shape)
    nrm = tgt / np.linalg.norm(tgt, axis=1, keepdims=True)
    d = -nrm + 0.6 * g / np.linalg.norm(g, axis=1, keepdims=True)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((len(tgt), 8), np.float32)
    rays[:, 0:3] = tgt - 3 * d
    rays[:, 4:7] = d
    rays[:, 7] = np.float32(3.4e38)
    h32 = sc.intersect_fp32(rays)
    assert (h32[:, 1].view(np.int32) >= 0).all()
    assert (sc.intersect_fp64(rays)["prim"] >= 0).all()
    # the BVH walk of the fp32 query equals testing every triangle, even on
    # rays through shared vertices and edges
    assert np.array_equal(h32.view(np.uint32), sc.intersect_fp32_brute(rays).view(np.uint32))
    # duplicate triangles: smaller primId wins in both numerics
    pos2 = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    sc2 = orc.Scene(np.array([[0, 1, 2], [0, 2, 1], [0, 1, 2]], np.int32), pos2)
    ray = np.array([[0.2, 0.2, 1, 0, 0, 0, -1, 3.4e38]], np.float32)
    assert sc2.intersect_fp32(ray)[0, 1].view(np.int32) == 0
    assert sc2.intersect_fp64(ray)["prim"][0] == 0


@pytest.mark.parametrize("numerics", [0, 1])
def test_oracle_known_answers(orc, scenes, numerics):
    w = h = 48
    sph = scenes.scene_convex_sphere(2)
    sc = orc.Scene(sph.idx, sph.pos)
    cam = orc.camera_from_view([0, 0, 3], 0, 0, math.radians(45), 1.0)
    r = sc.render(cam, w, h, spp=1, max_depth=2, albedo=0.8, numerics=numerics)
    assert set(np.unique(r["accum"][:, :, 0]).tolist()) == {np.float32(0.8), 1.0}
    r0 = sc.render(cam, w, h, spp=1, max_depth=0, albedo=0.8, numerics=numerics)
    assert set(np.unique(r0["accum"][:, :, 0]).tolist()) == {0.0, 1.0}
    box = scenes.scene_closed_box()
    sb = orc.Scene(box.idx, box.pos)
    cam2 = orc.camera_from_view([0, 0, 0.5], 0.3, 0.2, math.radians(60), 1.0)
    rb = sb.render(cam2, w, h, spp=2, max_depth=5, albedo=0.8, numerics=numerics)
    # Paths cannot escape a closed box, except through the reference's own
    # self-intersection gap: the CPU renderer moves the new origin eps along the
    # new direction with tnear 0 (core/ray.h:47-52), so a grazing bounce off a
    # hit point that rounded outside a wall can restart on the wall plane.
    # Such leaks are rare and carry albedo^k.
    a = rb["accum"][:, :, 0]
    assert (a == 0).mean() > 0.995 and (rb["accum"][:, :, 3] == 2).all()
    leaked = a[a != 0]
    assert np.isin(leaked, [np.float32(0.8) ** k for k in range(1, 6)]).all()
    assert rb["rays"] >= 2 * w * h * 5


@pytest.mark.parametrize("numerics", [0, 1])
def test_cosine_hemisphere_bounce(orc, numerics):
    """Bounce directions: unit length, in the hemisphere of the normal,
    E[cos] = 2/3, E[cos^2] = 1/2, azimuth uniform (chi^2 over 16 bins), and
    the origin offset eps = max(t, max|p|) * 0x1.fp-18 along the new direction."""
    rng = np.random.default_rng(5)
    n = 200000
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    org = rng.uniform(-1, 1, size=(n, 3))
    d = -nrm + 0.3 * rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    t = rng.uniform(0.1, 2.0, size=n)
    s = rng.random((n, 2))
    inp = np.concatenate([org, d, t[:, None], nrm, s], axis=1).astype(np.float32)
    out = orc.bounce(inp, numerics).astype(np.float64)
    nd = out[:, 3:6]
    nn = inp[:, 7:10].astype(np.float64)
    assert np.allclose(np.linalg.norm(nd, axis=1), 1.0, atol=2e-6)
    cos_t = np.sum(nd * nn, axis=1)
    assert (cos_t >= -1e-6).all()
    assert abs(cos_t.mean() - 2 / 3) < 4e-3
    assert abs((cos_t ** 2).mean() - 0.5) < 4e-3
    # azimuth around the normal relative to the frame's U axis of math/basis3.h:128-145
    u0 = np.stack([np.zeros(n), nn[:, 2], -nn[:, 1]], 1)
    u1 = np.stack([-nn[:, 2], np.zeros(n), nn[:, 0]], 1)
    U = np.where((np.sum(u0 * u0, 1) > np.sum(u1 * u1, 1))[:, None], u0, u1)
    U /= np.linalg.norm(U, axis=1, keepdims=True)
    V = np.cross(nn, U)
    phi = np.arctan2(np.sum(nd * V, 1), np.sum(nd * U, 1)) % (2 * np.pi)
    expect_phi = 2 * np.pi * inp[:, 11].astype(np.float64)
    dphi = np.abs(((phi - expect_phi) + np.pi) % (2 * np.pi) - np.pi)
    ok = cos_t < 0.999  # azimuth is undefined at the pole
    assert dphi[ok].max() < 1e-3
    counts, _ = np.histogram(phi, bins=16, range=(0, 2 * np.pi))
    chi2 = ((counts - n / 16) ** 2 / (n / 16)).sum()
    assert chi2 < 50.0  # 15 dof; p < 1e-5
    p = inp[:, 0:3].astype(np.float64) + inp[:, 6:7] * inp[:, 3:6].astype(np.float64)
    eps = np.maximum(inp[:, 6], np.abs(p).max(1)) * (0x1F / 2.0 ** 22)
    assert np.allclose(out[:, 0:3], p + eps[:, None] * nd, atol=3e-6)


def test_nearest_of_close_parallel_triangles(orc):
    """Candidates are ordered by the plane distance computed in double: two
    parallel triangles 1e-5 apart, far from the origin of coordinates
    (|x| ~ 1000, where fp32 T/det has an absolute error near 1e-4), resolve to
    the truly nearer one for every ray, and the reported t is the float64
    plane distance to within float rounding."""
    rng = np.random.default_rng(9)
    base = np.array([1000.0, 1000.0, 1000.0])
    # triangle k lies in the plane z = 1000 + 1e-3 + k * 1e-5 (k = 0 nearer)
    tris = []
    for k in range(2):
        z = 1000.0 + 1e-3 + k * 1e-5
        tris.append([[base[0] - 5, base[1] - 5, z], [base[0] + 5, base[1] - 5, z], [base[0], base[1] + 5, z]])
    pos = np.array(tris, np.float32).reshape(-1, 3)
    z_planes = pos[[0, 3], 2].astype(np.float64)
    assert z_planes[0] < z_planes[1]
    idx = np.array([[3, 4, 5], [0, 1, 2]], np.int32)   # the nearer plane gets the larger primId
    sc = orc.Scene(idx, pos)
    n = 2000
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:2] = base[:2] + rng.uniform(-1, 1, size=(n, 2))
    rays[:, 2] = 1000.0
    d = np.stack([rng.uniform(-0.3, 0.3, n), rng.uniform(-0.3, 0.3, n), np.ones(n)], 1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 4:7] = d
    rays[:, 7] = np.float32(3.4e38)
    for hits in (sc.intersect_fp32(rays), sc.intersect_fp32_brute(rays)):
        prim = hits[:, 1].view(np.int32)
        assert (prim == 1).all(), f"{int((prim != 1).sum())} rays picked the farther plane"
        t64 = (z_planes[0] - rays[:, 2].astype(np.float64)) / rays[:, 6].astype(np.float64)
        assert np.allclose(hits[:, 0], t64, rtol=2 ** -22, atol=0)


def test_stream_baseline_renders_the_oracle_image(orc, scenes):
    """The CPU timing baseline (restated DiffuseRendererStream over a float
    BVH2 or its 8-wide AVX2 form, render/diffuse_renderer_stream.h:82-225)
    renders the same image as
    the single-ray diffuse oracle (render/diffuse_integrator_single.h:41-69)
    for the same passes.  Its float slab test may cull a box that a ray only
    grazes within rounding (the oracle's walk is conservative), so a few
    paths may differ; every other pixel is bit-identical."""
    mesh = scenes.scene_heightfield_box(1001, cells=16)
    sc = orc.Scene(mesh.idx, mesh.pos)
    v = scenes.CONFIGS["heightfield_box_512"]["view"]
    w, h, frames = 96, 64, 3
    cam = orc.camera_from_view(v["pos"], v["angle_x"], v["angle_y"], v["fov_y"], w / h)
    ref = sc.render(cam, w, h, spp=frames, max_depth=4, albedo=0.8, renderer=1, numerics=0, threads=4)
    for simd in (False, True):  # the scalar BVH2 walk and the 8-wide AVX2 tree (scalar where AVX2 is absent)
        st = sc.stream_bench(cam, w, h, 0, h, frames, max_depth=4, albedo=0.8, threads=4, simd=simd)
        diff = (st["accum"] != ref["accum"]).any(axis=-1)
        assert diff.mean() <= 1e-3, f"simd={simd}: {int(diff.sum())} of {w * h} pixels differ"
        assert (st["accum"][:, :, 3] == frames).all()
        assert abs(st["rays"] - ref["rays"]) <= 4 * int(diff.sum())


def test_render_rows_equal_full_frame_rows(orc, scenes):
    """A row band of the single-ray renderer is those rows of the full frame
    (same pixel ids, samples and rays); the other rows stay empty."""
    mesh = scenes.scene_heightfield_box(1001, cells=16)
    sc = orc.Scene(mesh.idx, mesh.pos)
    v = scenes.CONFIGS["heightfield_box_512"]["view"]
    w, h = 40, 33
    cam = orc.camera_from_view(v["pos"], v["angle_x"], v["angle_y"], v["fov_y"], w / h)
    full = sc.render(cam, w, h, spp=2, max_depth=4, albedo=0.8, pass0=3)
    band = sc.render(cam, w, h, spp=2, max_depth=4, albedo=0.8, pass0=3, rows=(13, 22))
    assert np.array_equal(band["accum"][13:22], full["accum"][13:22])
    assert not band["accum"][:13].any() and not band["accum"][22:].any()
    assert 0 < band["rays"] < full["rays"]
