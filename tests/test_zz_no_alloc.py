"""prt_hip_intersect_closest_device_noalloc: the closest-hit query that
allocates nothing (its scratch is the library's fixed per-GPU array).

Bar: through the C ABI, its hit records are bit-identical to those of
prt_hip_intersect_closest_device (the allocating entry point) on the same
rays, for two calls in a row on one context, the second with fewer rays;
sizes beyond the limit are rejected with an error.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rays(rng, n, lo, hi):
    org = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = org
    rays[:, 4:7] = d.astype(np.float32)
    rays[:, 7] = np.float32(3.4e38)
    rays[::97, 7] = -1.0  # dead rays report a miss
    return rays


def test_noalloc_equals_allocating_query(prt, scenes):
    import torch

    mesh = scenes.scene_spheres(1002, count=64, subdiv=3)
    rng = np.random.default_rng(77)
    lib = prt.load_library()
    fn = lib.prt_hip_intersect_closest_device_noalloc
    ref_fn = lib.prt_hip_intersect_closest_device
    with prt.HipDevice(0) as d:
        d.upload_scene(mesh.idx, mesh.pos)
        d.build()
        for n in (300_000, 70_001):
            rays = torch.from_numpy(_rays(rng, n, -0.2, 1.2)).cuda()
            got = torch.full((n, 4), 7.0, dtype=torch.float32, device="cuda")
            ref = torch.full((n, 4), 9.0, dtype=torch.float32, device="cuda")
            torch.cuda.synchronize()
            print(f"noalloc query: {n} rays", flush=True)
            assert fn(d.ctx, rays.data_ptr(), got.data_ptr(), n) == 0, d.lib.prt_hip_last_error(d.ctx)
            assert ref_fn(d.ctx, rays.data_ptr(), ref.data_ptr(), n) == 0, d.lib.prt_hip_last_error(d.ctx)
            torch.cuda.synchronize()
            a = got.cpu().numpy()
            b = ref.cpu().numpy()
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"{int((a != b).any(1).sum())} hits differ"
            assert (a[:, 0] >= 0).mean() > 0.05  # some rays hit
            assert (a[::97, 1].view(np.int32) == -1).all()
        # limits: no work, only the checks
        assert fn(d.ctx, rays.data_ptr(), got.data_ptr(), 0) == 0
        assert fn(d.ctx, rays.data_ptr(), got.data_ptr(), 2 ** 31) != 0
        assert b"bad ray count" in d.lib.prt_hip_last_error(d.ctx)
        assert fn(d.ctx, None, got.data_ptr(), 5) != 0
        assert b"null array" in d.lib.prt_hip_last_error(d.ctx)
