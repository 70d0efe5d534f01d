"""Procedural workloads and ProtoRay's file formats.

Formats (byte-for-byte the reference's):
  .mesh  blob id 0x04df8666 (geometry/triangle_mesh.h:34), then Box3f bounds,
         int T, int V, bool hasNormals, bool hasTexcoords, int32[3T] indices,
         int32[T] materialIds, float[V*stride] vertex attributes, and the
         material names as int count + (int length + bytes) each
         (geometry/triangle_mesh.cpp:305-337, sys/blob.h:26-50,
         sys/array.h:357-393, sys/string.cpp:22-45).
  .view  ViewSet: 10 x View{Vec3f pos; float angleX, angleY, fovY, radius,
         focus} = 320 bytes, raw (main/render/view.h:25-40, main/render/view.cpp:23-46).
  .ppm   P6 with the B, G, R bytes of the packed pixel in that order
         (image/image_io.cpp:50-80).

Workloads: the generators below build the synthetic scenes the benchmark and
the tests run on.  Their seeds and camera views are fixed in CONFIGS and never
change once measured.
"""
from __future__ import annotations

import math
import struct
from dataclasses import dataclass, field

import numpy as np

MESH_BLOB_ID = 0x04DF8666


# ------------------------------------------------------------------ formats
@dataclass
class Mesh:
    idx: np.ndarray                      # (T, 3) int32
    pos: np.ndarray                      # (V, 3) float32
    material_ids: np.ndarray | None = None
    material_names: list = field(default_factory=list)
    normals: np.ndarray | None = None    # (V, 3) float32, optional
    texcoords: np.ndarray | None = None  # (V, 2) float32, optional

    @property
    def tri_count(self) -> int:
        return int(self.idx.shape[0])

    def bounds(self) -> np.ndarray:
        if self.pos.shape[0] == 0:
            return np.array([np.inf] * 3 + [-np.inf] * 3, dtype=np.float32)
        return np.concatenate([self.pos.min(axis=0), self.pos.max(axis=0)]).astype(np.float32)


def mesh_to_bytes(m: Mesh) -> bytes:
    T, V = m.idx.shape[0], m.pos.shape[0]
    has_n, has_uv = m.normals is not None, m.texcoords is not None
    attribs = [m.pos.astype(np.float32)]
    if has_n:
        attribs.append(m.normals.astype(np.float32))
    if has_uv:
        attribs.append(m.texcoords.astype(np.float32))
    verts = np.concatenate(attribs, axis=1) if len(attribs) > 1 else attribs[0]
    mat = m.material_ids if m.material_ids is not None else np.zeros(T, dtype=np.int32)
    names = m.material_names if m.material_names else ["default"]
    out = [struct.pack("<i", MESH_BLOB_ID), m.bounds().astype("<f4").tobytes(),
           struct.pack("<iiBB", T, V, int(has_n), int(has_uv)),
           np.ascontiguousarray(m.idx, dtype="<i4").tobytes(),
           np.ascontiguousarray(mat, dtype="<i4").tobytes(),
           np.ascontiguousarray(verts, dtype="<f4").tobytes(),
           struct.pack("<i", len(names))]
    for nm in names:
        b = nm.encode()
        out.append(struct.pack("<i", len(b)) + b)
    return b"".join(out)


def mesh_from_bytes(data: bytes) -> Mesh:
    off = 0

    def take(n: int) -> bytes:
        nonlocal off
        if off + n > len(data):
            raise RuntimeError("unexpected end of stream")
        b = data[off:off + n]
        off += n
        return b

    (blob,) = struct.unpack("<i", take(4))
    if blob != MESH_BLOB_ID:
        raise RuntimeError("file has wrong blob ID")
    take(24)  # bounds
    T, V, has_n, has_uv = struct.unpack("<iiBB", take(10))
    stride = 3 + (3 if has_n else 0) + (2 if has_uv else 0)
    idx = np.frombuffer(take(12 * T), dtype="<i4").reshape(T, 3).astype(np.int32)
    mat = np.frombuffer(take(4 * T), dtype="<i4").astype(np.int32)
    verts = np.frombuffer(take(4 * V * stride), dtype="<f4").reshape(V, stride).astype(np.float32)
    (n,) = struct.unpack("<i", take(4))
    names = []
    for _ in range(n):
        (ln,) = struct.unpack("<i", take(4))
        names.append(take(ln).decode())
    normals = verts[:, 3:6].copy() if has_n else None
    tc0 = 6 if has_n else 3
    texcoords = verts[:, tc0:tc0 + 2].copy() if has_uv else None
    return Mesh(idx, np.ascontiguousarray(verts[:, :3]), mat, names, normals, texcoords)


def save_mesh(path: str, m: Mesh) -> None:
    with open(path, "wb") as f:
        f.write(mesh_to_bytes(m))


def load_mesh(path: str) -> Mesh:
    with open(path, "rb") as f:
        return mesh_from_bytes(f.read())


VIEW_FMT = "<3f5f"   # pos xyz, angleX, angleY, fovY, radius, focus
VIEWSET_SIZE = 10


def views_to_bytes(views: list[dict]) -> bytes:
    vs = list(views) + [views[0]] * (VIEWSET_SIZE - len(views))
    return b"".join(struct.pack(VIEW_FMT, *v["pos"], v["angle_x"], v["angle_y"], v["fov_y"],
                                v.get("radius", 0.0), v.get("focus", 1.0)) for v in vs[:VIEWSET_SIZE])


def views_from_bytes(data: bytes) -> list[dict]:
    if len(data) != struct.calcsize(VIEW_FMT) * VIEWSET_SIZE:
        raise RuntimeError("view set must be 320 bytes")
    out = []
    for k in range(VIEWSET_SIZE):
        f = struct.unpack_from(VIEW_FMT, data, 32 * k)
        out.append({"pos": list(f[0:3]), "angle_x": f[3], "angle_y": f[4], "fov_y": f[5],
                    "radius": f[6], "focus": f[7]})
    return out


def ppm_bytes(bgr8: np.ndarray) -> bytes:
    """P6 image of packed b | g<<8 | r<<16 pixels, bytes written in memory order."""
    h, w = bgr8.shape
    px = np.ascontiguousarray(bgr8, dtype="<i4").view(np.uint8).reshape(h, w, 4)[:, :, :3]
    return b"P6\n%d %d\n255\n" % (w, h) + px.tobytes()


# --------------------------------------------------------------- geometry
def heightfield(cells: int, rng: np.random.Generator, hmax: float, lo=-1.0, hi=1.0) -> tuple:
    n = cells + 1
    xs = np.linspace(lo, hi, n, dtype=np.float64)
    gx, gz = np.meshgrid(xs, xs, indexing="xy")
    gy = rng.uniform(0.0, hmax, size=(n, n))
    pos = np.stack([gx, gy, gz], axis=-1).reshape(-1, 3).astype(np.float32)
    i = np.arange(cells)
    a = (i[None, :] + n * i[:, None]).ravel()
    b, c, d = a + 1, a + n, a + n + 1
    idx = np.concatenate([np.stack([a, c, b], 1), np.stack([b, c, d], 1)], axis=0).astype(np.int32)
    return idx, pos


def quad(p0, p1, p2, p3) -> tuple:
    pos = np.array([p0, p1, p2, p3], dtype=np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int32)
    return idx, pos


def icosphere(subdiv: int) -> tuple:
    """Unit icosphere: 20 * 4^subdiv triangles, 10 * 4^subdiv + 2 vertices."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
             (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    v = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in verts]
    for _ in range(subdiv):
        cache = {}

        def mid(i, j):
            key = (min(i, j), max(i, j))
            if key not in cache:
                m = v[i] + v[j]
                v.append(m / np.linalg.norm(m))
                cache[key] = len(v) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.array(faces, dtype=np.int32), np.array(v, dtype=np.float64)


def box_tris(lo, hi) -> tuple:
    """Closed axis-aligned box: 8 vertices, 12 triangles."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    pos = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                    [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], dtype=np.float32)
    idx = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                    [3, 7, 6], [3, 6, 2], [0, 4, 7], [0, 7, 3], [1, 2, 6], [1, 6, 5]], dtype=np.int32)
    return idx, pos


def merge(parts: list[tuple]) -> Mesh:
    idxs, poss, off = [], [], 0
    for idx, pos in parts:
        idxs.append(np.asarray(idx, dtype=np.int64) + off)
        poss.append(np.asarray(pos, dtype=np.float32))
        off += len(pos)
    return Mesh(np.concatenate(idxs).astype(np.int32), np.concatenate(poss).astype(np.float32))


# -------------------------------------------------------------- workloads
def scene_heightfield_box(seed: int = 1001, cells: int = 64) -> Mesh:
    """64x64-cell heightfield (heights U[0, 0.2]) on [-1,1]^2 inside the open
    5-quad box [-1,1] x [0,2] x [-1,1] (floor and four walls, no lid)."""
    rng = np.random.default_rng(seed)
    hf = heightfield(cells, rng, 0.2)
    parts = [hf,
             quad((-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1)),
             quad((-1, 0, -1), (-1, 2, -1), (1, 2, -1), (1, 0, -1)),
             quad((-1, 0, 1), (1, 0, 1), (1, 2, 1), (-1, 2, 1)),
             quad((-1, 0, -1), (-1, 0, 1), (-1, 2, 1), (-1, 2, -1)),
             quad((1, 0, -1), (1, 2, -1), (1, 2, 1), (1, 0, 1))]
    return merge(parts)


def scene_spheres(seed: int = 1002, count: int = 256, subdiv: int = 4) -> Mesh:
    """`count` icospheres (centres U[0,1]^3, radii log-uniform [0.01, 0.1]) over
    a two-triangle ground quad at y = 0 spanning [-1, 2]^2."""
    rng = np.random.default_rng(seed)
    f, v = icosphere(subdiv)
    centres = rng.uniform(0.0, 1.0, size=(count, 3))
    radii = np.exp(rng.uniform(math.log(0.01), math.log(0.1), size=count))
    parts = [(f, (v * r + c).astype(np.float32)) for c, r in zip(centres, radii)]
    parts.append(quad((-1, 0, -1), (-1, 0, 2), (2, 0, 2), (2, 0, -1)))
    return merge(parts)


def vertex_normals(idx: np.ndarray, pos: np.ndarray) -> np.ndarray:
    """Smooth-shading vertex normals: the normalised sum of the area-weighted
    normals of the triangles around each vertex."""
    p = pos.astype(np.float64)
    fn = np.cross(p[idx[:, 1]] - p[idx[:, 0]], p[idx[:, 2]] - p[idx[:, 0]])
    vn = np.zeros_like(p)
    for k in range(3):
        np.add.at(vn, idx[:, k], fn)
    ln = np.linalg.norm(vn, axis=1, keepdims=True)
    return (vn / np.where(ln > 0, ln, 1.0)).astype(np.float32)


def scene_smooth(seed: int = 1006, cells: int = 32, spheres: int = 24, subdiv: int = 3) -> Mesh:
    """A mesh with per-vertex normals and texture coordinates for the full
    shading context: a heightfield (heights U[0, 0.3]) on [-1,1]^2 with
    smooth vertex normals and planar texture coordinates (x, z); icospheres
    above it with their exact normals and latitude/longitude texture
    coordinates; a wall quad at z = 1 whose vertex normals point against its
    geometric normal and whose texture coordinates are all equal (a
    degenerate mapping)."""
    rng = np.random.default_rng(seed)
    hidx, hpos = heightfield(cells, rng, 0.3)
    parts = [(hidx, hpos, vertex_normals(hidx, hpos), np.stack([hpos[:, 0], hpos[:, 2]], 1))]
    f, v = icosphere(subdiv)
    lon = (np.arctan2(v[:, 2], v[:, 0]) / (2 * math.pi) + 0.5).astype(np.float32)
    lat = (np.arccos(np.clip(v[:, 1], -1.0, 1.0)) / math.pi).astype(np.float32)
    for _ in range(spheres):
        c = rng.uniform([-0.8, 0.35, -0.8], [0.8, 0.9, 0.8])
        r = rng.uniform(0.08, 0.25)
        parts.append((f, (v * r + c).astype(np.float32), v.astype(np.float32), np.stack([lon, lat], 1)))
    widx, wpos = quad((-1, 0, 1), (1, 0, 1), (1, 1.2, 1), (-1, 1.2, 1))
    wn = np.cross(wpos[1] - wpos[0], wpos[2] - wpos[0])
    parts.append((widx, wpos, np.tile(-wn / np.linalg.norm(wn), (4, 1)).astype(np.float32),
                  np.full((4, 2), 0.25, np.float32)))
    m = merge([(i, p) for i, p, _, _ in parts])
    m.normals = np.concatenate([n for _, _, n, _ in parts]).astype(np.float32)
    m.texcoords = np.concatenate([t for _, _, _, t in parts]).astype(np.float32)
    return m


# a view of scene_smooth from z = -2.6 towards +z: spheres, heightfield, wall
SMOOTH_VIEW = {"pos": [0.0, 1.6, -2.6], "angle_x": 0.45, "angle_y": math.pi, "fov_y": 0.9}


def scene_convex_sphere(subdiv: int = 3) -> Mesh:
    f, v = icosphere(subdiv)
    return Mesh(f, v.astype(np.float32))


def scene_closed_box() -> Mesh:
    idx, pos = box_tris((-1, -1, -1), (1, 1, 1))
    return Mesh(idx, pos)


# Workload table: scene builder, view (ProtoRay .view semantics), image size,
# spp, max depth and albedo.
CONFIGS = {
    "heightfield_box_512": {
        "scene": "heightfield_box", "seed": 1001,
        "view": {"pos": [0.0, 4.5, 1.2], "angle_x": 1.3, "angle_y": 0.0, "fov_y": math.radians(45.0)},
        "width": 512, "height": 512, "spp": 4, "max_depth": 4, "albedo": 0.8,
    },
    "heightfield_boxes_512": {
        "scene": "heightfield_boxes", "seed": 1004, "albedo_seed": 3004, "target_seed": 4004,
        "view": {"pos": [0.0, 1.9, 2.1], "angle_x": 0.75, "angle_y": 0.0, "fov_y": math.radians(50.0)},
        "width": 512, "height": 512, "spp": 8, "max_depth": 6, "albedo": None,
    },
    "city_4k": {
        "scene": "city", "seed": 1005, "albedo_seed": 3005,
        "view": {"pos": [0.5, 6.0, 2.6], "angle_x": 1.17, "angle_y": 0.0, "fov_y": math.radians(40.0)},
        "width": 3840, "height": 2160, "spp": 64, "max_depth": 8, "albedo": 0.8,
    },
    "spheres_1080p": {
        "scene": "spheres", "seed": 1002,
        "view": {"pos": [0.5, 0.9, 2.6], "angle_x": 0.25, "angle_y": 0.0, "fov_y": math.radians(45.0)},
        "width": 1920, "height": 1080, "spp": 16, "max_depth": 8, "albedo": 0.7,
    },
}


def scene_mix(seed: int, target_tris: int) -> Mesh:
    """A heightfield on [-1,1]^2 plus icospheres, about `target_tris` triangles
    split at a random fraction between the two (the batch workload)."""
    rng = np.random.default_rng(seed)
    frac = rng.uniform(0.2, 0.8)
    hf_tris = max(8, int(frac * target_tris))
    cells = max(2, int(round(math.sqrt(hf_tris / 2.0))))
    parts = [heightfield(cells, rng, rng.uniform(0.05, 0.3))]
    rest = max(80, target_tris - 2 * cells * cells)
    subdiv = int(np.clip(math.floor(math.log(max(rest / 160.0, 1.0), 4)), 0, 4))
    f, v = icosphere(subdiv)
    count = max(1, int(round(rest / len(f))))
    centres = np.stack([rng.uniform(-1, 1, count), rng.uniform(0.15, 0.9, count), rng.uniform(-1, 1, count)], 1)
    radii = np.exp(rng.uniform(math.log(0.02), math.log(0.15), count))
    parts += [(f, (v * r + c).astype(np.float32)) for c, r in zip(centres, radii)]
    return merge(parts)


def scene_heightfield_boxes(seed: int = 1004, cells: int = 128, boxes: int = 64) -> Mesh:
    """The differentiable workload: a `cells` x `cells` heightfield (heights
    U[0, 0.2]) on [-1,1]^2 plus `boxes` closed boxes (12 triangles each) standing
    on it, footprints U[0.05, 0.25], heights U[0.1, 0.6]."""
    rng = np.random.default_rng(seed)
    parts = [heightfield(cells, rng, 0.2)]
    for _ in range(boxes):
        cx, cz = rng.uniform(-0.9, 0.9, 2)
        sx, sz = rng.uniform(0.05, 0.25, 2)
        hgt = rng.uniform(0.1, 0.6)
        parts.append(box_tris((cx - sx / 2, 0.0, cz - sz / 2), (cx + sx / 2, hgt, cz + sz / 2)))
    return merge(parts)


def scene_city(seed: int = 1005, boxes: int = 1_330_000) -> Mesh:
    """Large stress workload: `boxes` closed axis-aligned boxes (12 triangles,
    8 vertices each; 1.33 M boxes = 15.96 M triangles) with footprints
    U[0.002, 0.01] per axis, centres U[0,1]^2 on the ground y = 0, and heights
    log-normal(mu=0, sigma=1): a dense field of thin towers.  Overlapping
    boxes produce coincident faces, which exercise the (t, primId) tie rule."""
    rng = np.random.default_rng(seed)
    cx = rng.uniform(0.0, 1.0, boxes)
    cz = rng.uniform(0.0, 1.0, boxes)
    sx = rng.uniform(0.002, 0.01, boxes)
    sz = rng.uniform(0.002, 0.01, boxes)
    hy = np.exp(rng.normal(0.0, 1.0, boxes))
    lo = np.stack([cx - sx / 2, np.zeros(boxes), cz - sz / 2], 1).astype(np.float32)
    hi = np.stack([cx + sx / 2, hy, cz + sz / 2], 1).astype(np.float32)
    bi, bp = box_tris((0, 0, 0), (1, 1, 1))          # unit-box topology
    sel = bp.astype(bool)                            # which corner takes hi per axis
    pos = np.where(sel[None, :, :], hi[:, None, :], lo[:, None, :]).reshape(-1, 3)
    idx = (bi[None, :, :] + 8 * np.arange(boxes, dtype=np.int64)[:, None, None]).reshape(-1, 3)
    return Mesh(idx.astype(np.int32), np.ascontiguousarray(pos, dtype=np.float32))


def albedo_set(tri_count: int, seed: int, lo: float = 0.1, hi: float = 0.9) -> np.ndarray:
    """Per-triangle RGB albedo, uniform in [lo, hi], shape (tri_count, 3)."""
    return np.random.default_rng(seed).uniform(lo, hi, size=(tri_count, 3)).astype(np.float32)


BATCH_RESOLUTIONS = [(256, 256), (640, 480), (1280, 720), (1920, 1080)]
BATCH_VIEW = {"pos": [0.0, 1.7, 2.7], "angle_x": 0.55, "angle_y": 0.0, "fov_y": math.radians(50.0)}


def batch_workload(n: int = 32, seed0: int = 2000, min_tris: float = 1e3, max_tris: float = 2e6,
                   max_depth: int = 6) -> list[dict]:
    """The batched workload: n scenes with triangle counts log-uniform in
    [min_tris, max_tris], resolutions from BATCH_RESOLUTIONS, spp uniform in
    1..16, depth 6; scene i uses seed seed0 + i."""
    jobs = []
    for i in range(n):
        rng = np.random.default_rng(seed0 + i)
        tris = int(math.exp(rng.uniform(math.log(min_tris), math.log(max_tris))))
        w, h = BATCH_RESOLUTIONS[int(rng.integers(0, len(BATCH_RESOLUTIONS)))]
        spp = int(rng.integers(1, 17))
        jobs.append({"seed": seed0 + i, "target_tris": tris, "width": w, "height": h, "spp": spp,
                     "max_depth": max_depth, "albedo": 0.8, "view": BATCH_VIEW})
    return jobs


def build_scene(name: str) -> Mesh:
    cfg = CONFIGS[name]
    if cfg["scene"] == "heightfield_box":
        return scene_heightfield_box(cfg["seed"])
    if cfg["scene"] == "spheres":
        return scene_spheres(cfg["seed"])
    if cfg["scene"] == "heightfield_boxes":
        return scene_heightfield_boxes(cfg["seed"])
    if cfg["scene"] == "city":
        return scene_city(cfg["seed"])
    raise KeyError(name)
