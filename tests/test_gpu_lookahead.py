"""Frame lookahead of the fused diffuse renderer, on the GPU.

With lookahead on, a render call starts paths of the next pass while its own
frame finishes, and the next call keeps that work when it renders that pass
with the same inputs.  Bar: after every call the accumulator is bit-identical
to rendering the same calls with lookahead off, and so is each call's ray
count; a chain of calls also matches the kernel-numerics oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cam(orc, scenes, name, w, h, shift=0.0):
    v = scenes.CONFIGS[name]["view"]
    pos = [v["pos"][0] + shift, v["pos"][1], v["pos"][2]]
    return orc.camera_from_view(pos, v["angle_x"], v["angle_y"], v["fov_y"], w / h)


def _run(dev, steps, lookahead):
    """Runs the calls in `steps` from a cleared frame; returns the accumulator
    and the ray count after each diffuse call."""
    dev.set_lookahead(lookahead)
    out = []
    try:
        for step in steps:
            r = step(dev)
            if r is not None:
                out.append((dev.accum(), r["ray"]))
    finally:
        dev.set_lookahead(True)
    return out


def _same(a, b):
    assert len(a) == len(b)
    for k, ((ia, ra), (ib, rb)) in enumerate(zip(a, b)):
        assert np.array_equal(ia, ib), f"call {k}: {int((ia != ib).any(-1).sum())} pixels differ"
        assert ra == rb, f"call {k}: ray count {ra} with lookahead, {rb} without"


@pytest.fixture(scope="module")
def spheres(dev, scenes):
    mesh = scenes.scene_spheres(1002, count=96, subdiv=3)
    dev.upload_scene(mesh.idx, mesh.pos)
    dev.build()
    return mesh


def test_lookahead_chain_bit_exact_small_frame(dev, orc, scenes, spheres):
    """480x270 (fewer paths than lanes of the launch): every call starts the
    whole next pass and suspends most of it; 6 calls in a chain."""
    w, h = 480, 270
    cam = _cam(orc, scenes, "spheres_1080p", w, h)

    def setup(d):
        d.set_camera(cam)
        d.init_frame(w, h)

    steps = [setup] + [lambda d: d.render("diffuseHip", max_depth=8, albedo=0.7)] * 6
    on = _run(dev, steps, True)
    off = _run(dev, steps, False)
    _same(on, off)
    ref = orc.Scene(spheres.idx, spheres.pos).render(cam, w, h, spp=6, max_depth=8, albedo=0.7, renderer=1,
                                                     numerics=0)
    assert np.array_equal(on[-1][0], ref["accum"])
    assert sum(r for _, r in on) == ref["rays"]


def test_lookahead_chain_bit_exact_1080p(dev, orc, scenes, spheres):
    """1920x1080 (more paths than lanes): 4 calls in a chain."""
    w, h = 1920, 1080
    cam = _cam(orc, scenes, "spheres_1080p", w, h)

    def setup(d):
        d.set_camera(cam)
        d.init_frame(w, h)

    steps = [setup] + [lambda d: d.render("diffuseHip", max_depth=8, albedo=0.7)] * 4
    _same(_run(dev, steps, True), _run(dev, steps, False))


def test_lookahead_dropped_when_inputs_change(dev, orc, scenes, spheres):
    """Calls that change the camera, the pass, max depth, albedo, the renderer,
    the profiling mode or the pipeline in between: each call still equals its
    lookahead-off counterpart (work started for a pass that is not rendered
    next, or with other inputs, is dropped)."""
    w, h = 320, 200
    cam_a = _cam(orc, scenes, "spheres_1080p", w, h)
    cam_b = _cam(orc, scenes, "spheres_1080p", w, h, shift=0.5)

    def diffuse(depth=8, albedo=0.7):
        return lambda d: d.render("diffuseHip", max_depth=depth, albedo=albedo)

    def setup(d):
        d.set_camera(cam_a)
        d.init_frame(w, h)

    def cam(c):
        def f(d):
            d.set_camera(c)
        return f

    def set_pass(p):
        def f(d):
            d.set_pass(p)
        return f

    def primary(d):
        d.render("primaryHip")

    def fast(d):
        d.render("diffuseFastHip", max_depth=8, albedo=0.7)

    def counters(on):
        def f(d):
            d.set_profiling(counters=on)
        return f

    def clear(d):
        d.clear_frame()

    def pipeline(name):
        def f(d):
            d.set_pipeline(name)
        return f

    def diffuse2(d):
        d.render("diffuse2Hip", max_depth=4, albedo=0.7)

    steps = [setup, diffuse(), diffuse(), cam(cam_b), diffuse(), diffuse(), set_pass(7), diffuse(),
             diffuse(depth=5), diffuse(depth=5), diffuse(albedo=0.5), primary, diffuse(), fast, diffuse(),
             counters(True), diffuse(), counters(False), diffuse(), diffuse(), clear, diffuse(), diffuse(),
             diffuse2, diffuse(), diffuse(), pipeline("wavefront"), diffuse(), pipeline("fused"), diffuse(),
             diffuse(), diffuse()]
    _same(_run(dev, steps, True), _run(dev, steps, False))


@pytest.mark.parametrize("w,h,depth", [(13, 7, 8), (64, 1, 0), (257, 161, 1), (33, 65, 14)])
def test_lookahead_chain_edge_sizes(dev, orc, scenes, spheres, w, h, depth):
    """Frames smaller than one wave, one pixel high, with partial 8x8 tiles,
    depth 0 (paths end at their first hit) and the deepest depth: 5 calls in
    a chain equal the lookahead-off calls."""
    cam = _cam(orc, scenes, "spheres_1080p", w, h)

    def setup(d):
        d.set_camera(cam)
        d.init_frame(w, h)

    steps = [setup] + [lambda d: d.render("diffuseHip", max_depth=depth, albedo=0.7)] * 5
    _same(_run(dev, steps, True), _run(dev, steps, False))


def test_lookahead_dropped_after_scene_rebuild(dev, orc, scenes):
    """A rebuilt scene between two calls of consecutive passes: the second call
    drops the lookahead work of the first (different geometry)."""
    w, h = 256, 160
    m1 = scenes.scene_spheres(1002, count=96, subdiv=3)
    m2 = scenes.scene_spheres(1003, count=64, subdiv=3)
    cam = _cam(orc, scenes, "spheres_1080p", w, h)

    def load(m):
        def f(d):
            d.upload_scene(m.idx, m.pos)
            d.build()
        return f

    def setup(d):
        d.set_camera(cam)
        d.init_frame(w, h)

    render = lambda d: d.render("diffuseHip", max_depth=8, albedo=0.7)  # noqa: E731
    steps = [load(m1), setup, render, render, load(m2), render, render]
    _same(_run(dev, steps, True), _run(dev, steps, False))
