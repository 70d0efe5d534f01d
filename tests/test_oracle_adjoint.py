"""Pins the oracle's per-triangle albedo model: its exact adjoint against
central finite differences of its own forward image (double precision, the
image is a polynomial in the albedos), and closed forms on a convex mesh."""
import math

import numpy as np


def _scene(scenes, orc):
    mesh = scenes.scene_heightfield_boxes(7, cells=8, boxes=4)
    sc = orc.Scene(mesh.idx, mesh.pos)
    v = scenes.CONFIGS["heightfield_boxes_512"]["view"]
    cam = orc.camera_from_view(v["pos"], v["angle_x"], v["angle_y"], v["fov_y"], 1.0)
    return mesh, sc, cam


def test_adjoint_equals_central_differences(scenes, orc):
    mesh, sc, cam = _scene(scenes, orc)
    tape = orc.PathTape(sc, cam, 32, 32, pass0=3, spp=2, max_depth=4)
    rng = np.random.default_rng(0)
    a = rng.uniform(0.1, 0.9, (mesh.tri_count, 3))
    target = tape.image(rng.uniform(0.1, 0.9, (mesh.tri_count, 3)))
    img = tape.image(a)
    dldi = 2.0 * (img - target)
    grad = tape.adjoint(a, dldi)

    def loss(x):
        return float(((tape.image(x) - target) ** 2).sum())

    hit = np.nonzero(np.abs(grad).sum(1) > 0)[0]
    assert len(hit) > 20
    h = 1e-4
    for t in rng.choice(hit, size=min(40, len(hit)), replace=False):
        for c in range(3):
            ap, am = a.copy(), a.copy()
            ap[t, c] += h
            am[t, c] -= h
            fd = (loss(ap) - loss(am)) / (2 * h)
            assert abs(fd - grad[t, c]) <= 1e-6 * max(1.0, abs(fd)), (t, c, fd, grad[t, c])


def test_convex_mesh_closed_forms(scenes, orc):
    """Every path that hits a convex mesh escapes after one bounce: the pixel
    is the mean albedo of the triangles its samples hit, and dI/da(t) counts
    the samples that hit t."""
    mesh = scenes.scene_convex_sphere(2)
    sc = orc.Scene(mesh.idx, mesh.pos)
    cam = orc.camera_from_view([0, 0, 3], 0, 0, math.radians(45), 1.0)
    w = h = 24
    tape = orc.PathTape(sc, cam, w, h, pass0=0, spp=3, max_depth=2)
    a = np.random.default_rng(1).uniform(0.1, 0.9, (mesh.tri_count, 3))
    img = tape.image(a)
    grey = sc.render(cam, w, h, spp=3, max_depth=2, albedo=0.5, numerics=0)
    # same paths: where the grey render saw a hit (value R per sample), the
    # tape's image is a mean of albedos; where it saw sky (1.0), the image is 1
    n_hit = (3.0 - grey["accum"][:, :, 0]) / 0.5  # samples that hit, per pixel
    sky = n_hit == 0
    assert np.allclose(img[sky], 1.0)
    ones = np.ones((h, w, 3))
    grad = tape.adjoint(a, ones)
    # sum over triangles of dI/da(t) = (samples that hit) / spp, per channel
    assert np.isclose(grad[:, 0].sum(), n_hit.sum() / 3.0)
