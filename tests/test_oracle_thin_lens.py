"""Oracle: the thin-lens camera (ThinLensCamera::getRay, camera/thin_lens_camera.h:38-44).

Pins, from the geometry rather than the oracle's formula:
  - every camera ray starts on the lens disk: in the plane spanned by basisU
    and basisV through the camera origin, within lensRadius of it;
  - every ray passes through its focal point origin + f (imageO + ix dx + iy dy),
    the pinhole ray's point at distance f along it, whatever the lens sample;
  - the lens samples cover the disk uniformly: E[r^2] = R^2 / 2, mean offset 0,
    equal counts in the four quadrants;
  - lens radius 0 gives exactly the pinhole rays.
"""
import math

import numpy as np


def _cam(orc, radius, focus, w=64, h=48):
    cam = orc.camera_from_view([0.2, 0.9, 2.6], 0.25, 0.1, math.radians(45), w / h)
    return orc.camera20(cam, radius, focus), cam


def test_thin_lens_geometry(orc):
    w, h = 64, 48
    R, f = 0.05, 2.5
    cam, _ = _cam(orc, R, f, w, h)
    rays, states = orc.raygen(cam, w, h, 3)
    o = cam[0:3].astype(np.float64)
    U, V = cam[12:15].astype(np.float64), cam[15:18].astype(np.float64)
    N = np.cross(U, V)
    begin = rays[:, 0:3].astype(np.float64)
    off = begin - o
    assert np.abs(off @ N).max() < 1e-6                       # on the lens plane
    assert (np.linalg.norm(off, axis=1) <= R * (1 + 1e-5)).all()
    # the focal point of the pixel sample, origin + f (imageO + ix dx + iy dy)
    # with (ix, iy) from the pixel's first two sampler draws, lies on the ray
    io, dx, dy = (cam[3:6].astype(np.float64), cam[6:9].astype(np.float64), cam[9:12].astype(np.float64))
    pick = np.random.default_rng(1).choice(w * h, size=300, replace=False)
    d = rays[:, 4:7].astype(np.float64)
    assert np.abs(np.linalg.norm(d, axis=1) - 1).max() < 1e-6
    for pid in pick:
        _, u = orc.sampler_sequence(3, int(pid), 4)
        x, y = pid % w, pid // w
        img = io + ((x + float(u[0])) / w) * dx + ((y + float(u[1])) / h) * dy
        focal = o + f * img
        t = float((focal - begin[pid]) @ d[pid])
        assert np.linalg.norm(begin[pid] + t * d[pid] - focal) < 1e-5 * f


def test_thin_lens_disk_is_uniform(orc):
    w, h = 128, 128
    R = 0.1
    cam, _ = _cam(orc, R, 3.0, w, h)
    rays, _ = orc.raygen(cam, w, h, 0)
    o = cam[0:3].astype(np.float64)
    U, V = cam[12:15].astype(np.float64), cam[15:18].astype(np.float64)
    off = rays[:, 0:3].astype(np.float64) - o
    a, b = off @ U, off @ V
    r2 = (a * a + b * b) / (R * R)
    n = r2.size
    assert abs(r2.mean() - 0.5) < 4 * math.sqrt(1 / 12 / n)       # r^2 ~ U[0,1]
    assert abs(a.mean()) < 4 * R * 0.5 / math.sqrt(n) and abs(b.mean()) < 4 * R * 0.5 / math.sqrt(n)
    q = np.histogram(np.arctan2(b, a), bins=4, range=(-math.pi, math.pi))[0]
    assert (np.abs(q - n / 4) < 4 * math.sqrt(n * 0.25 * 0.75)).all()


def test_zero_radius_is_the_pinhole(orc):
    w, h = 40, 30
    _, base = _cam(orc, 0.0, 1.0, w, h)
    a, sa = orc.raygen(orc.camera20(base), w, h, 5)
    b, sb = orc.raygen(orc.camera20(base, 0.0, 7.0), w, h, 5)   # focal distance plays no role
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32)) and np.array_equal(sa, sb)
    o = base[0:3]
    assert (a[:, 0:3] == o).all()


def test_thin_lens_convex_known_answer(orc, scenes):
    """A convex mesh under a white sky still gives exactly {R, 1} per pixel
    with a lens: every hit path escapes after one bounce."""
    sph = scenes.scene_convex_sphere(2)
    sc = orc.Scene(sph.idx, sph.pos)
    cam = orc.camera20(orc.camera_from_view([0, 0, 3], 0, 0, math.radians(45), 1.0), 0.2, 3.0)
    for numerics in (0, 1):
        r = sc.render(cam, 48, 48, spp=2, max_depth=2, albedo=0.8, numerics=numerics)
        vals = set(np.unique(r["accum"][:, :, 0]).tolist())
        assert vals <= {0.0, np.float32(0.8), np.float32(1.6), 1.0, np.float32(1.8), 2.0}
        assert len(vals) > 2   # defocus mixes hit and miss samples at the silhouette


def test_thin_lens_orientation(orc):
    """The lens offset of each ray is uniformSampleDisk(draws 3, 4) * R in the
    (basisU, basisV) frame (sampling/shape_sampler.h:115-124, math/basis3.h:120-123),
    evaluated here in float64 with numpy's sin/cos."""
    w, h = 32, 24
    R = 0.07
    cam, _ = _cam(orc, R, 2.0, w, h)
    rays, _ = orc.raygen(cam, w, h, 9)
    o = cam[0:3].astype(np.float64)
    U, V = cam[12:15].astype(np.float64), cam[15:18].astype(np.float64)
    for pid in range(0, w * h, 7):
        _, u = orc.sampler_sequence(9, pid, 4)
        r = math.sqrt(float(u[2]))
        lx, ly = r * math.cos(2 * math.pi * float(u[3])) * R, r * math.sin(2 * math.pi * float(u[3])) * R
        expect = o + lx * U + ly * V
        assert np.linalg.norm(rays[pid, 0:3] - expect) < 1e-6
