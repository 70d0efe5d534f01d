"""ctypes wrapper of oracle/liboracle.so -- the CPU checker.

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product path never loads it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_LIB = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P, i64, f, i = ctypes.c_void_p, ctypes.c_int64, ctypes.c_float, ctypes.c_int
        L.orc_sampler_seed.argtypes = [i, i]
        L.orc_sampler_seed.restype = ctypes.c_uint32
        L.orc_sampler_sequence.argtypes = [i, i, i, P, P]
        L.orc_sincos_2pi.argtypes = [P, i64, P, P]
        L.orc_camera_from_view.argtypes = [P, f, f, f, f, P]
        L.orc_raygen.argtypes = [P, i, i, i, P, P]
        L.orc_scene_create.argtypes = [P, i64, P, i64]
        L.orc_scene_create.restype = P
        L.orc_scene_destroy.argtypes = [P]
        L.orc_scene_attributes.argtypes = [P, P, P]
        L.orc_shading_frame.argtypes = [P, i, f, f, P, i, P]
        L.orc_intersect_fp64.argtypes = [P, P, i64, P, P, P, P, P, P, P, i]
        L.orc_ray_edge_distance.argtypes = [P, P, P, i64, P]
        L.orc_intersect_fp32.argtypes = [P, P, i64, P, i]
        L.orc_intersect_fp32_brute.argtypes = [P, P, i64, P]
        L.orc_occluded_fp32.argtypes = [P, P, i64, P, i]
        L.orc_render.argtypes = [P, P, i, i, i, i, i, f, i, i, i, P, P, P]
        L.orc_render.restype = i
        L.orc_render_rows.argtypes = [P, P, i, i, i, i, i, i, i, f, i, i, i, P, P, P]
        L.orc_render_rows.restype = i
        L.orc_stream_bench.argtypes = [P, P, i, i, i, i, i, i, i, f, i, i, P, P, P]
        L.orc_stream_bench.restype = i
        L.orc_bounce.argtypes = [P, i64, i, P]
        L.orc_thread_count.restype = i
        L.orc_cpu_has_avx2.restype = i
        L.orc_paths_record.argtypes = [P, P, i, i, i, i, i, i]
        L.orc_paths_record.restype = P
        L.orc_paths_destroy.argtypes = [P]
        L.orc_paths_rays.argtypes = [P]
        L.orc_paths_rays.restype = i64
        L.orc_paths_image.argtypes = [P, P, P]
        L.orc_paths_adjoint.argtypes = [P, P, P, P]
        _LIB = L
    return _LIB


def _p(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def sampler_seed(pass_index: int, pixel: int) -> int:
    return int(lib().orc_sampler_seed(pass_index, pixel))


def sampler_sequence(pass_index: int, pixel: int, n: int) -> tuple[np.ndarray, np.ndarray]:
    st = np.zeros(n, dtype=np.uint32)
    vals = np.zeros(n, dtype=np.float32)
    lib().orc_sampler_sequence(pass_index, pixel, n, _p(st), _p(vals))
    return st, vals


def sincos_2pi(v: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    v = np.ascontiguousarray(v, dtype=np.float32)
    s = np.zeros_like(v)
    c = np.zeros_like(v)
    lib().orc_sincos_2pi(_p(v), v.size, _p(s), _p(c))
    return s, c


def camera20(cam, lens_radius: float | None = None, focal_distance: float | None = None) -> np.ndarray:
    """Camera as the 20 floats the oracle reads: origin, imageO, imageDx,
    imageDy, basisU, basisV, lensRadius, focalDistance.  An 18-float (or
    shorter) array is a pinhole camera (lens radius 0, focal distance 1)."""
    c = np.asarray(cam, dtype=np.float32).ravel()
    out = np.zeros(20, dtype=np.float32)
    n = min(c.size, 20)
    out[:n] = c[:n]
    if c.size < 20:
        out[18], out[19] = 0.0, 1.0
    if lens_radius is not None:
        out[18] = lens_radius
    if focal_distance is not None:
        out[19] = focal_distance
    return out


def camera_from_view(pos, angle_x: float, angle_y: float, fov_y: float, aspect: float) -> np.ndarray:
    p = np.asarray(pos, dtype=np.float32)
    out = np.zeros(20, dtype=np.float32)
    lib().orc_camera_from_view(_p(p), angle_x, angle_y, fov_y, aspect, _p(out))
    return out


def raygen(cam: np.ndarray, w: int, h: int, pass_index: int) -> tuple[np.ndarray, np.ndarray]:
    cam = camera20(cam)
    rays = np.zeros((w * h, 8), dtype=np.float32)
    st = np.zeros(w * h, dtype=np.uint32)
    lib().orc_raygen(_p(cam), w, h, pass_index, _p(rays), _p(st))
    return rays, st


class Scene:
    """BVH2 over the triangles (binned SAH), for the FP64 and fp32 queries."""

    def __init__(self, idx: np.ndarray, pos: np.ndarray):
        self.idx = np.ascontiguousarray(idx, dtype=np.int32).reshape(-1, 3)
        self.pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        self.h = lib().orc_scene_create(_p(self.idx), self.idx.shape[0], _p(self.pos), self.pos.shape[0])

    def close(self) -> None:
        if self.h:
            lib().orc_scene_destroy(self.h)
            self.h = None

    def set_attributes(self, normals: np.ndarray | None = None, texcoords: np.ndarray | None = None) -> None:
        """Per-vertex normals (V x 3) and texture coordinates (V x 2); None removes one."""
        self.normals = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        self.texcoords = None if texcoords is None else np.ascontiguousarray(texcoords, np.float32).reshape(-1, 2)
        for a in (self.normals, self.texcoords):
            if a is not None and a.shape[0] != self.pos.shape[0]:
                raise ValueError("one attribute per vertex expected")
        lib().orc_scene_attributes(self.h, None if self.normals is None else _p(self.normals),
                                   None if self.texcoords is None else _p(self.texcoords))

    def shading_frame(self, prim: int, u: float, v: float, direction, fast: bool = False) -> np.ndarray:
        """Shading frame (rows U, V, N) of hit (prim, u, v) for a ray along direction."""
        d = np.ascontiguousarray(direction, np.float32)
        out = np.zeros((3, 3), np.float32)
        lib().orc_shading_frame(self.h, int(prim), float(u), float(v), _p(d), int(fast), _p(out))
        return out

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def intersect_fp64(self, rays: np.ndarray, threads: int = 0) -> dict:
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = rays.shape[0]
        out = {"prim": np.zeros(n, np.int32), "t": np.zeros(n), "u": np.zeros(n), "v": np.zeros(n),
               "edge": np.zeros(n), "prim2": np.zeros(n, np.int32), "t2": np.zeros(n)}
        lib().orc_intersect_fp64(self.h, _p(rays), n, _p(out["prim"]), _p(out["t"]), _p(out["u"]), _p(out["v"]),
                                 _p(out["edge"]), _p(out["prim2"]), _p(out["t2"]), threads)
        return out

    def ray_edge_distance(self, rays: np.ndarray, prims: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        prims = np.ascontiguousarray(prims, dtype=np.int32)
        out = np.zeros(rays.shape[0])
        lib().orc_ray_edge_distance(self.h, _p(rays), _p(prims), rays.shape[0], _p(out))
        return out

    def intersect_fp32(self, rays: np.ndarray, threads: int = 0) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        hits = np.zeros((rays.shape[0], 4), dtype=np.float32)
        lib().orc_intersect_fp32(self.h, _p(rays), rays.shape[0], _p(hits), threads)
        return hits

    def intersect_fp32_brute(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        hits = np.zeros((rays.shape[0], 4), dtype=np.float32)
        lib().orc_intersect_fp32_brute(self.h, _p(rays), rays.shape[0], _p(hits))
        return hits

    def occluded_fp32(self, rays: np.ndarray, brute: bool = False) -> np.ndarray:
        """Any-hit query: distance of a blocking triangle per ray, or -1."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        out = np.zeros(rays.shape[0], dtype=np.float32)
        lib().orc_occluded_fp32(self.h, _p(rays), rays.shape[0], _p(out), int(brute))
        return out

    def render(self, cam: np.ndarray, w: int, h: int, spp: int, max_depth: int, albedo: float,
               renderer: int = 1, numerics: int = 0, pass0: int = 0, threads: int = 0,
               rows: tuple[int, int] | None = None) -> dict:
        """rows=(r0, r1): only image rows r0..r1-1 of the w x h frame (the
        others stay 0)."""
        cam = camera20(cam)
        accum = np.zeros((h, w, 4), dtype=np.float32)
        sumsq = np.zeros((h, w), dtype=np.float32)
        rays = ctypes.c_int64(0)
        r0, r1 = rows if rows is not None else (0, h)
        lib().orc_render_rows(self.h, _p(cam), w, h, r0, r1, pass0, spp, max_depth, albedo, renderer, numerics,
                              threads, _p(accum), _p(sumsq), ctypes.byref(rays))
        return {"accum": accum, "sumsq": sumsq, "rays": rays.value}

    def stream_bench(self, cam: np.ndarray, w: int, h: int, row0: int, row1: int, frames: int,
                     max_depth: int, albedo: float, threads: int = 0, pass0: int = 0, simd: bool = True) -> dict:
        """simd: the 8-wide AVX2 traversal when the host has AVX2, else the scalar BVH2 walk."""
        cam = camera20(cam)
        accum = np.zeros((h, w, 4), dtype=np.float32)
        rays = ctypes.c_int64(0)
        secs = ctypes.c_double(0.0)
        lib().orc_stream_bench(self.h, _p(cam), w, h, row0, row1, pass0, frames, max_depth, albedo, threads,
                               1 if simd else 0, _p(accum), ctypes.byref(rays), ctypes.byref(secs))
        return {"accum": accum, "rays": rays.value, "seconds": secs.value}


class PathTape:
    """Hit lists of every path (pixel, pass) of a render, traced once with the
    kernel numerics; the per-triangle albedo model is then evaluated in double."""

    def __init__(self, scene: "Scene", cam: np.ndarray, w: int, h: int, pass0: int, spp: int, max_depth: int,
                 threads: int = 0):
        cam = camera20(cam)
        self.w, self.h, self.tri_count = w, h, scene.idx.shape[0]
        self.h_ = lib().orc_paths_record(scene.h, _p(cam), w, h, pass0, spp, max_depth, threads)

    def close(self) -> None:
        if self.h_:
            lib().orc_paths_destroy(self.h_)
            self.h_ = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def rays(self) -> int:
        return int(lib().orc_paths_rays(self.h_))

    def image(self, albedo: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(albedo, dtype=np.float64).reshape(-1)
        out = np.zeros((self.h, self.w, 3), dtype=np.float64)
        lib().orc_paths_image(self.h_, _p(a), _p(out))
        return out

    def adjoint(self, albedo: np.ndarray, dldi: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(albedo, dtype=np.float64).reshape(-1)
        g = np.ascontiguousarray(dldi, dtype=np.float64).reshape(-1)
        out = np.zeros((self.tri_count, 3), dtype=np.float64)
        lib().orc_paths_adjoint(self.h_, _p(a), _p(g), _p(out))
        return out


def bounce(inp: np.ndarray, numerics: int = 0) -> np.ndarray:
    """inp: (n, 12) org, dir, t, normal, s1, s2 -> (n, 6) new org, new dir."""
    inp = np.ascontiguousarray(inp, dtype=np.float32).reshape(-1, 12)
    out = np.zeros((inp.shape[0], 6), dtype=np.float32)
    lib().orc_bounce(_p(inp), inp.shape[0], numerics, _p(out))
    return out


def cpu_has_avx2() -> bool:
    return bool(lib().orc_cpu_has_avx2())


def thread_count() -> int:
    return int(lib().orc_thread_count())
