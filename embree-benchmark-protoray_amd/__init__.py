"""protoray_hip: Python binding of libprt_hip.so, the MI355X wavefront diffuse
path tracer behind ProtoRay's Device API (render/device.h:29-54).

The binding is a thin ctypes layer over the C ABI in include/prt_hip.h.  It
raises RuntimeError with the library's message on every nonzero status, the
way the reference's Device methods throw (render/device_cuda.cpp:63,92).
There is no CPU fallback: if the HIP library is missing or no GPU is present,
constructing HipDevice fails.
"""
from __future__ import annotations

import ctypes
import importlib.util
import os
import re

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "libprt_hip.so")
HEADER_PATH = os.path.join(REPO_DIR, "include", "prt_hip.h")

RENDERER_PRIMARY = 0
RENDERER_DIFFUSE = 1
RENDERER_DIFFUSE2 = 2
RENDERER_FAST = 0x100   # modifier: SimpleShadingContext, each pass overwrites the accumulator
RENDERER_IDS = {"primaryHip": RENDERER_PRIMARY, "diffuseHip": RENDERER_DIFFUSE, "diffuse2Hip": RENDERER_DIFFUSE2,
                "primaryFastHip": RENDERER_PRIMARY | RENDERER_FAST,
                "diffuseFastHip": RENDERER_DIFFUSE | RENDERER_FAST,
                "diffuse2FastHip": RENDERER_DIFFUSE2 | RENDERER_FAST}

RAY_DTYPE = np.dtype([("org", "<f4", 3), ("tnear", "<f4"), ("dir", "<f4", 3), ("tfar", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("prim_id", "<i4"), ("u", "<f4"), ("v", "<f4")])
NODE_BYTES = 80


class PrtCamera(ctypes.Structure):
    _fields_ = [("origin", ctypes.c_float * 3), ("image_o", ctypes.c_float * 3),
                ("image_dx", ctypes.c_float * 3), ("image_dy", ctypes.c_float * 3),
                ("basis_u", ctypes.c_float * 3), ("basis_v", ctypes.c_float * 3),
                ("lens_radius", ctypes.c_float), ("focal_distance", ctypes.c_float)]


class PrtBuildStats(ctypes.Structure):
    _fields_ = [("build_ms", ctypes.c_double), ("build_mprim", ctypes.c_double),
                ("node_count", ctypes.c_int64), ("tri_count", ctypes.c_int64),
                ("depth", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class PrtRenderStats(ctypes.Structure):
    _fields_ = [("rays", ctypes.c_int64), ("ms", ctypes.c_double), ("mray", ctypes.c_double),
                ("paths", ctypes.c_int64), ("queue", ctypes.c_int64 * 16), ("trace_ms", ctypes.c_double),
                ("shade_ms", ctypes.c_double), ("nodes", ctypes.c_int64), ("tris", ctypes.c_int64),
                ("spills", ctypes.c_int64), ("wave_steps", ctypes.c_int64), ("lane_steps", ctypes.c_int64),
                ("wave_shades", ctypes.c_int64), ("drain_wave_steps", ctypes.c_int64),
                ("drain_lane_steps", ctypes.c_int64), ("drain_ms", ctypes.c_double)]


class PrtTraceCounters(ctypes.Structure):
    _fields_ = [("rays", ctypes.c_int64), ("nodes", ctypes.c_int64), ("tris", ctypes.c_int64),
                ("spills", ctypes.c_int64)]


class PrtBatchJob(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_int32), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("spp", ctypes.c_int32), ("max_depth", ctypes.c_int32), ("pass0", ctypes.c_int32),
                ("albedo", ctypes.c_float), ("camera", PrtCamera)]


class PrtBatchStats(ctypes.Structure):
    _fields_ = [("rays", ctypes.c_int64), ("paths", ctypes.c_int64), ("ms", ctypes.c_double),
                ("mray", ctypes.c_double), ("nodes", ctypes.c_int64), ("tris", ctypes.c_int64)]


class PrtBatchTuning(ctypes.Structure):
    _fields_ = [("variant", ctypes.c_int32), ("block", ctypes.c_int32), ("stack", ctypes.c_int32),
                ("blocks_per_cu", ctypes.c_int32), ("refill", ctypes.c_int32), ("ms", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


def declared_symbols(header: str = HEADER_PATH) -> list[str]:
    """Every function name declared in include/prt_hip.h."""
    text = open(header).read()
    return sorted(set(re.findall(r"\b(prt_hip_\w+)\s*\(", text)))


_LIB = None


def share_torch_runtime() -> None:
    """One HIP runtime per process.  PyTorch's ROCm wheel carries its own
    libamdhip64 (same soname as /opt/rocm's); a process that loads both cannot
    use the second one -- torch reports no GPU if the library came first.  When
    torch is installed it is therefore imported before the library, which then
    binds to the runtime torch already holds, so device pointers and streams
    are shared with torch tensors."""
    if importlib.util.find_spec("torch") is not None:
        import torch  # noqa: F401  (loads torch's HIP runtime)


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libprt_hip.so (built in-tree by `make -C embree-benchmark-protoray_amd`)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library not built: {path} (run __graft_entry__.build())")
    share_torch_runtime()
    lib = ctypes.CDLL(path)
    P = ctypes.c_void_p
    i64 = ctypes.c_int64
    sigs = {
        "prt_hip_abi_version": ([], ctypes.c_int),
        "prt_hip_create": ([ctypes.c_int, ctypes.POINTER(P)], ctypes.c_int),
        "prt_hip_destroy": ([P], ctypes.c_int),
        "prt_hip_last_error": ([P], ctypes.c_char_p),
        "prt_hip_device_name": ([P, ctypes.c_char_p, ctypes.c_int], ctypes.c_int),
        "prt_hip_scene_upload": ([P, P, i64, P, i64], ctypes.c_int),
        "prt_hip_scene_attributes": ([P, P, P], ctypes.c_int),
        "prt_hip_batch_scene_attributes": ([P, ctypes.c_int32, P, P], ctypes.c_int),
        "prt_hip_bvh_build": ([P, ctypes.POINTER(PrtBuildStats)], ctypes.c_int),
        "prt_hip_intersect_closest": ([P, P, P, i64], ctypes.c_int),
        "prt_hip_intersect_closest_device": ([P, P, P, i64], ctypes.c_int),
        "prt_hip_intersect_closest_device_noalloc": ([P, P, P, i64], ctypes.c_int),
        "prt_hip_intersect_counters": ([P, P, P, i64, ctypes.POINTER(PrtTraceCounters)], ctypes.c_int),
        "prt_hip_occluded": ([P, P, P, i64], ctypes.c_int),
        "prt_hip_occluded_device": ([P, P, P, i64], ctypes.c_int),
        "prt_hip_camera_from_view": ([ctypes.POINTER(ctypes.c_float * 3), ctypes.c_float, ctypes.c_float,
                                      ctypes.c_float, ctypes.c_float, ctypes.POINTER(PrtCamera)], ctypes.c_int),
        "prt_hip_camera_set": ([P, ctypes.POINTER(PrtCamera)], ctypes.c_int),
        "prt_hip_frame_init": ([P, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "prt_hip_frame_clear": ([P], ctypes.c_int),
        "prt_hip_render": ([P, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.POINTER(PrtRenderStats)],
                           ctypes.c_int),
        "prt_hip_set_pass": ([P, ctypes.c_int], ctypes.c_int),
        "prt_hip_set_profiling": ([P, ctypes.c_int], ctypes.c_int),
        "prt_hip_set_pipeline": ([P, ctypes.c_int], ctypes.c_int),
        "prt_hip_set_bvh_collapse": ([P, ctypes.c_int, ctypes.c_float], ctypes.c_int),
        "prt_hip_set_tail": ([P, ctypes.c_int], ctypes.c_int),
        "prt_hip_set_lookahead": ([P, ctypes.c_int], ctypes.c_int),
        "prt_hip_accum_read": ([P, P], ctypes.c_int),
        "prt_hip_frame_read_bgr8": ([P, P], ctypes.c_int),
        "prt_hip_debug_raygen": ([P, ctypes.c_int, P, P], ctypes.c_int),
        "prt_hip_bvh_info": ([P, ctypes.POINTER(i64), ctypes.POINTER(i64), ctypes.POINTER(ctypes.c_int32)],
                             ctypes.c_int),
        "prt_hip_bvh_read": ([P, P, P, P], ctypes.c_int),
        "prt_hip_set_albedo_host": ([P, P], ctypes.c_int),
        "prt_hip_set_albedo_device": ([P, P], ctypes.c_int),
        "prt_hip_render_albedo_image": ([P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P,
                                         ctypes.POINTER(PrtRenderStats)], ctypes.c_int),
        "prt_hip_render_albedo_image_host": ([P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P,
                                              ctypes.POINTER(PrtRenderStats)], ctypes.c_int),
        "prt_hip_adjoint_albedo": ([P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P,
                                    ctypes.POINTER(PrtRenderStats)], ctypes.c_int),
        "prt_hip_adjoint_albedo_host": ([P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P,
                                         ctypes.POINTER(PrtRenderStats)], ctypes.c_int),
        "prt_hip_view_basis": ([ctypes.c_float, ctypes.c_float, P], ctypes.c_int),
        "prt_hip_camera_from_basis": ([P, P, ctypes.c_float, ctypes.c_float, ctypes.POINTER(PrtCamera)],
                                      ctypes.c_int),
        "prt_hip_batch_scene_add": ([P, P, i64, P, i64, ctypes.POINTER(PrtBuildStats),
                                     ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
        "prt_hip_batch_reset": ([P], ctypes.c_int),
        "prt_hip_batch_set_jobs": ([P, P, ctypes.c_int], ctypes.c_int),
        "prt_hip_batch_render": ([P, ctypes.POINTER(PrtBatchStats)], ctypes.c_int),
        "prt_hip_batch_accum_read": ([P, ctypes.c_int, P], ctypes.c_int),
        "prt_hip_batch_set_tuning": ([P, ctypes.POINTER(PrtBatchTuning)], ctypes.c_int),
        "prt_hip_batch_get_tuning": ([P, ctypes.POINTER(PrtBatchTuning)], ctypes.c_int),
        "prt_hip_batch_autotune": ([P, P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "prt_hip_batch_set_graph": ([P, ctypes.c_int], ctypes.c_int),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _LIB = lib
    return lib


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def _ptr_or_null(a: np.ndarray | None) -> ctypes.c_void_p | None:
    return None if a is None else _ptr(a)


def _attributes(normals, texcoords, vertex_count: int):
    """Contiguous float32 (V x 3) normals and (V x 2) texture coordinates, or None."""
    n = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
    t = None if texcoords is None else np.ascontiguousarray(texcoords, dtype=np.float32).reshape(-1, 2)
    for a in (n, t):
        if a is not None and a.shape[0] != vertex_count:
            raise ValueError("one attribute per vertex expected")
    return n, t


def camera_struct(cam: np.ndarray) -> PrtCamera:
    """PrtCamera from 20 floats: origin, imageO, imageDx, imageDy, basisU,
    basisV, lensRadius, focalDistance.  18 floats are a pinhole camera."""
    cam = np.asarray(cam, dtype=np.float32).ravel()
    c = PrtCamera()
    for k, name in enumerate(["origin", "image_o", "image_dx", "image_dy", "basis_u", "basis_v"]):
        if 3 * k + 3 <= cam.size:
            getattr(c, name)[:] = [float(x) for x in cam[3 * k:3 * k + 3]]
    c.lens_radius = float(cam[18]) if cam.size >= 20 else 0.0
    c.focal_distance = float(cam[19]) if cam.size >= 20 else 1.0
    return c


def camera_from_view(position, angle_x: float, angle_y: float, fov_y: float, aspect: float,
                     lens_radius: float = 0.0, focal_distance: float = 1.0) -> np.ndarray:
    """20 floats (origin, imageO, imageDx, imageDy, basisU, basisV, lensRadius,
    focalDistance) of a view's camera: pinhole when lens_radius is 0, thin lens
    otherwise (RenderWindow::makeCamera, main/render/render_window.cpp:307-325)."""
    lib = load_library()
    pos = (ctypes.c_float * 3)(*[float(x) for x in position])
    cam = PrtCamera()
    if lib.prt_hip_camera_from_view(ctypes.byref(pos), angle_x, angle_y, fov_y, aspect, ctypes.byref(cam)) != 0:
        raise RuntimeError("camera_from_view failed")
    out = []
    for name in ["origin", "image_o", "image_dx", "image_dy", "basis_u", "basis_v"]:
        out.extend(getattr(cam, name))
    out += [lens_radius, focal_distance]
    return np.array(out, dtype=np.float32)


class HipDevice:
    """One prt_hip_ctx: scene, BVH8, camera, frame and wavefront queues on one GPU."""

    def __init__(self, device_index: int = 0):
        self.lib = load_library()
        ctx = ctypes.c_void_p()
        rc = self.lib.prt_hip_create(device_index, ctypes.byref(ctx))
        if rc != 0:
            raise RuntimeError(f"prt_hip_create failed with status {rc} (no usable GPU?)")
        self.ctx = ctx
        self.width = self.height = 0
        self.tri_count = 0

    # -- plumbing
    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            msg = self.lib.prt_hip_last_error(self.ctx).decode()
            raise RuntimeError(f"{what}: {msg}")

    def close(self) -> None:
        if getattr(self, "ctx", None):
            self.lib.prt_hip_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> str:
        buf = ctypes.create_string_buffer(256)
        self._check(self.lib.prt_hip_device_name(self.ctx, buf, 256), "device_name")
        return buf.value.decode()

    # -- scene / BVH
    def upload_scene(self, idx: np.ndarray, pos: np.ndarray) -> None:
        idx = np.ascontiguousarray(idx, dtype=np.int32).reshape(-1, 3)
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        self._keep = (idx, pos)
        self._check(self.lib.prt_hip_scene_upload(self.ctx, _ptr(idx), idx.shape[0], _ptr(pos), pos.shape[0]),
                    "scene_upload")
        self.tri_count = idx.shape[0]
        self.vertex_count = pos.shape[0]

    def set_attributes(self, normals: np.ndarray | None = None, texcoords: np.ndarray | None = None) -> None:
        """Per-vertex normals (V x 3) and texture coordinates (V x 2) of the uploaded
        scene; the primary/diffuse/diffuse2 renderers then shade with the full
        ShadingContext.  None removes an attribute."""
        n, t = _attributes(normals, texcoords, self.vertex_count)
        self._check(self.lib.prt_hip_scene_attributes(self.ctx, _ptr_or_null(n), _ptr_or_null(t)),
                    "scene_attributes")

    def set_bvh_collapse(self, method: str = "sah", prim_cost: float = 1.0) -> None:
        """BVH8 collapse of later builds: 'sah' (surface-area cost optimum) or 'greedy'."""
        ids = {"greedy": 0, "sah": 1}
        if method not in ids:
            raise ValueError("method must be 'sah' or 'greedy'")
        self._check(self.lib.prt_hip_set_bvh_collapse(self.ctx, ids[method], prim_cost), "set_bvh_collapse")

    def build(self) -> dict:
        st = PrtBuildStats()
        self._check(self.lib.prt_hip_bvh_build(self.ctx, ctypes.byref(st)), "bvh_build")
        return {"buildMs": st.build_ms, "buildMprim": st.build_mprim, "nodes": st.node_count,
                "tris": st.tri_count, "depth": st.depth}

    def bvh(self) -> tuple[np.ndarray, np.ndarray, np.ndarray, int]:
        nn, nt, dp = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
        self._check(self.lib.prt_hip_bvh_info(self.ctx, ctypes.byref(nn), ctypes.byref(nt), ctypes.byref(dp)),
                    "bvh_info")
        nodes = np.zeros((nn.value, NODE_BYTES), dtype=np.uint8)
        tris = np.zeros((nt.value, 9), dtype=np.float32)
        prims = np.zeros(nt.value, dtype=np.int32)
        self._check(self.lib.prt_hip_bvh_read(self.ctx, _ptr(nodes), _ptr(tris), _ptr(prims)), "bvh_read")
        return nodes, tris, prims, dp.value

    # -- closest hit
    def intersect(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        hits = np.zeros((rays.shape[0], 4), dtype=np.float32)
        self._check(self.lib.prt_hip_intersect_closest(self.ctx, _ptr(rays), _ptr(hits), rays.shape[0]),
                    "intersect_closest")
        return hits

    def occluded(self, rays: np.ndarray) -> np.ndarray:
        """Any-hit query (IntersectorStreamCuda::occluded): per ray, the distance
        of some blocking triangle in [tnear, tfar], or -1."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        out = np.zeros(rays.shape[0], dtype=np.float32)
        self._check(self.lib.prt_hip_occluded(self.ctx, _ptr(rays), _ptr(out), rays.shape[0]), "occluded")
        return out

    def intersect_counters(self, rays: np.ndarray) -> tuple[np.ndarray, dict]:
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        hits = np.zeros((rays.shape[0], 4), dtype=np.float32)
        c = PrtTraceCounters()
        self._check(self.lib.prt_hip_intersect_counters(self.ctx, _ptr(rays), _ptr(hits), rays.shape[0],
                                                        ctypes.byref(c)), "intersect_counters")
        return hits, {"rays": c.rays, "nodes": c.nodes, "tris": c.tris, "spills": c.spills}

    # -- camera / frame / render
    def set_camera(self, cam: np.ndarray) -> None:
        c = camera_struct(cam)
        self._check(self.lib.prt_hip_camera_set(self.ctx, ctypes.byref(c)), "camera_set")

    def init_frame(self, width: int, height: int) -> None:
        self._check(self.lib.prt_hip_frame_init(self.ctx, width, height), "frame_init")
        self.width, self.height = width, height

    def clear_frame(self) -> None:
        self._check(self.lib.prt_hip_frame_clear(self.ctx), "frame_clear")

    def set_pass(self, p: int) -> None:
        self._check(self.lib.prt_hip_set_pass(self.ctx, p), "set_pass")

    def render(self, renderer: int | str = RENDERER_DIFFUSE, max_depth: int = 6, albedo: float = 0.8) -> dict:
        if isinstance(renderer, str):
            if renderer not in RENDERER_IDS:
                raise ValueError("invalid renderer type")
            renderer = RENDERER_IDS[renderer]
        st = PrtRenderStats()
        self._check(self.lib.prt_hip_render(self.ctx, renderer, max_depth, albedo, ctypes.byref(st)), "render")
        return {"ray": st.rays, "renderMs": st.ms, "mray": st.mray, "paths": st.paths,
                "queue": [st.queue[k] for k in range(16)], "traceMs": st.trace_ms, "shadeMs": st.shade_ms,
                "nodes": st.nodes, "tris": st.tris, "spills": st.spills, "waveSteps": st.wave_steps,
                "laneSteps": st.lane_steps, "waveShades": st.wave_shades,
                "drainWaveSteps": st.drain_wave_steps, "drainLaneSteps": st.drain_lane_steps, "drainMs": st.drain_ms}

    def set_tail(self, mode: str = "cooperative") -> None:
        """Tail of the fused diffuse frame: 'cooperative' (a wave with at most 8
        paths left spreads each over 8 lanes) or 'lanes'."""
        ids = {"lanes": 0, "cooperative": 1}
        if mode not in ids:
            raise ValueError("mode must be 'cooperative' or 'lanes'")
        self._check(self.lib.prt_hip_set_tail(self.ctx, ids[mode]), "set_tail")

    def set_lookahead(self, enable: bool = True) -> None:
        """Frame lookahead of the fused diffuse renderer (default on): in a
        call that continues the previous one (next pass, same inputs), idle
        lanes at the end of the frame start the next two passes' paths, which
        the next call keeps if it continues this one.  Images and ray counts per
        call do not change."""
        self._check(self.lib.prt_hip_set_lookahead(self.ctx, 1 if enable else 0), "set_lookahead")

    def set_pipeline(self, pipeline: str) -> None:
        """'fused' (one persistent path launch per frame) or 'wavefront' (per-depth launches)."""
        ids = {"fused": 0, "wavefront": 1}
        if pipeline not in ids:
            raise ValueError("pipeline must be 'fused' or 'wavefront'")
        self._check(self.lib.prt_hip_set_pipeline(self.ctx, ids[pipeline]), "set_pipeline")

    def set_profiling(self, timing: bool = False, counters: bool = False) -> None:
        flags = (1 if timing else 0) | (2 if counters else 0)
        self._check(self.lib.prt_hip_set_profiling(self.ctx, flags), "set_profiling")

    def accum(self) -> np.ndarray:
        out = np.zeros((self.height, self.width, 4), dtype=np.float32)
        self._check(self.lib.prt_hip_accum_read(self.ctx, _ptr(out)), "accum_read")
        return out

    def frame_bgr8(self) -> np.ndarray:
        out = np.zeros((self.height, self.width), dtype=np.int32)
        self._check(self.lib.prt_hip_frame_read_bgr8(self.ctx, _ptr(out)), "frame_read_bgr8")
        return out

    # -- per-triangle RGB albedo and its adjoint
    def set_albedo(self, rgb: np.ndarray) -> None:
        rgb = np.ascontiguousarray(rgb, dtype=np.float32).reshape(-1, 3)
        if rgb.shape[0] != self.tri_count:
            raise ValueError("albedo needs 3 floats per triangle")
        self._check(self.lib.prt_hip_set_albedo_host(self.ctx, _ptr(rgb)), "set_albedo")

    def set_albedo_device(self, ptr: int) -> None:
        self._check(self.lib.prt_hip_set_albedo_device(self.ctx, ctypes.c_void_p(ptr)), "set_albedo_device")

    @staticmethod
    def _stats(st: PrtRenderStats) -> dict:
        return {"ray": st.rays, "paths": st.paths, "renderMs": st.ms, "mray": st.mray, "nodes": st.nodes,
                "tris": st.tris}

    def render_albedo_image(self, pass0: int, spp: int, max_depth: int) -> tuple[np.ndarray, dict]:
        img = np.zeros((self.height, self.width, 3), dtype=np.float32)
        st = PrtRenderStats()
        self._check(self.lib.prt_hip_render_albedo_image_host(self.ctx, pass0, spp, max_depth, _ptr(img),
                                                              ctypes.byref(st)), "render_albedo_image")
        return img, self._stats(st)

    def render_albedo_image_device(self, pass0: int, spp: int, max_depth: int, image_ptr: int) -> dict:
        st = PrtRenderStats()
        self._check(self.lib.prt_hip_render_albedo_image(self.ctx, pass0, spp, max_depth, ctypes.c_void_p(image_ptr),
                                                         ctypes.byref(st)), "render_albedo_image")
        return self._stats(st)

    def adjoint_albedo(self, pass0: int, spp: int, max_depth: int, dldi: np.ndarray) -> tuple[np.ndarray, dict]:
        dldi = np.ascontiguousarray(dldi, dtype=np.float32).reshape(self.height, self.width, 3)
        grad = np.zeros((self.tri_count, 3), dtype=np.float32)
        st = PrtRenderStats()
        self._check(self.lib.prt_hip_adjoint_albedo_host(self.ctx, pass0, spp, max_depth, _ptr(dldi), _ptr(grad),
                                                         ctypes.byref(st)), "adjoint_albedo")
        return grad, self._stats(st)

    def adjoint_albedo_device(self, pass0: int, spp: int, max_depth: int, dldi_ptr: int, grad_ptr: int) -> dict:
        st = PrtRenderStats()
        self._check(self.lib.prt_hip_adjoint_albedo(self.ctx, pass0, spp, max_depth, ctypes.c_void_p(dldi_ptr),
                                                    ctypes.c_void_p(grad_ptr), ctypes.byref(st)),
                    "adjoint_albedo")
        return self._stats(st)

    # -- batched form
    def batch_add_scene(self, idx: np.ndarray, pos: np.ndarray) -> tuple[int, dict]:
        idx = np.ascontiguousarray(idx, dtype=np.int32).reshape(-1, 3)
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        st = PrtBuildStats()
        sid = ctypes.c_int32()
        self._check(self.lib.prt_hip_batch_scene_add(self.ctx, _ptr(idx), idx.shape[0], _ptr(pos), pos.shape[0],
                                                     ctypes.byref(st), ctypes.byref(sid)), "batch_scene_add")
        if not hasattr(self, "_batch_vertex_counts"):
            self._batch_vertex_counts = {}
        self._batch_vertex_counts[sid.value] = pos.shape[0]
        return sid.value, {"buildMs": st.build_ms, "buildMprim": st.build_mprim, "nodes": st.node_count,
                           "depth": st.depth}

    def batch_set_attributes(self, scene: int, normals: np.ndarray | None = None,
                             texcoords: np.ndarray | None = None) -> None:
        """Shading attributes of batch scene `scene` (as set_attributes)."""
        n, t = _attributes(normals, texcoords, self._batch_vertex_counts[scene])
        self._check(self.lib.prt_hip_batch_scene_attributes(self.ctx, scene, _ptr_or_null(n), _ptr_or_null(t)),
                    "batch_scene_attributes")

    def batch_reset(self) -> None:
        self._check(self.lib.prt_hip_batch_reset(self.ctx), "batch_reset")

    def batch_set_jobs(self, jobs: list[dict]) -> None:
        arr = (PrtBatchJob * len(jobs))()
        for a, j in zip(arr, jobs):
            a.scene, a.width, a.height = j["scene"], j["width"], j["height"]
            a.spp, a.max_depth, a.pass0 = j["spp"], j["max_depth"], j.get("pass0", 0)
            a.albedo = j.get("albedo", 0.8)
            a.camera = camera_struct(j["camera"])
        self._batch_sizes = [(j["height"], j["width"]) for j in jobs]
        self._check(self.lib.prt_hip_batch_set_jobs(self.ctx, ctypes.cast(arr, ctypes.c_void_p), len(jobs)),
                    "batch_set_jobs")

    def batch_render(self) -> dict:
        st = PrtBatchStats()
        self._check(self.lib.prt_hip_batch_render(self.ctx, ctypes.byref(st)), "batch_render")
        return {"ray": st.rays, "paths": st.paths, "renderMs": st.ms, "mray": st.mray, "nodes": st.nodes,
                "tris": st.tris}

    def batch_accum(self, job: int) -> np.ndarray:
        h, w = self._batch_sizes[job]
        out = np.zeros((h, w, 4), dtype=np.float32)
        self._check(self.lib.prt_hip_batch_accum_read(self.ctx, job, _ptr(out)), "batch_accum_read")
        return out

    def batch_get_tuning(self) -> dict:
        t = PrtBatchTuning()
        self._check(self.lib.prt_hip_batch_get_tuning(self.ctx, ctypes.byref(t)), "batch_get_tuning")
        return t.as_dict()

    def batch_set_tuning(self, variant: int, blocks_per_cu: int = 0, refill: int = 24) -> None:
        t = PrtBatchTuning(variant=variant, blocks_per_cu=blocks_per_cu, refill=refill)
        self._check(self.lib.prt_hip_batch_set_tuning(self.ctx, ctypes.byref(t)), "batch_set_tuning")

    def batch_autotune(self) -> list[dict]:
        res = (PrtBatchTuning * 64)()
        n = ctypes.c_int()
        self._check(self.lib.prt_hip_batch_autotune(self.ctx, ctypes.cast(res, ctypes.c_void_p), 64,
                                                    ctypes.byref(n)), "batch_autotune")
        return [res[i].as_dict() for i in range(n.value)]

    def batch_set_graph(self, enable: bool) -> None:
        self._check(self.lib.prt_hip_batch_set_graph(self.ctx, 1 if enable else 0), "batch_set_graph")

    def debug_raygen(self, pass_index: int) -> tuple[np.ndarray, np.ndarray]:
        rays = np.zeros((self.height * self.width, 8), dtype=np.float32)
        states = np.zeros(self.height * self.width, dtype=np.uint32)
        self._check(self.lib.prt_hip_debug_raygen(self.ctx, pass_index, _ptr(rays), _ptr(states)), "debug_raygen")
        return rays, states
