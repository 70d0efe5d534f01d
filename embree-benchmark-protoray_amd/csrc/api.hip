// api.hip -- the extern "C" entry points declared in include/prt_hip.h.
//
// The host control flow of one prt_hip_render call mirrors
// DiffuseRendererCuda::render (render/diffuse_renderer_cuda.cu:159-212):
// raygen, then intersect + shade per depth up to max_depth.  Unlike the
// reference, the queue sizes stay on the device: every kernel reads its count
// from device memory, so a frame issues no host round trip until the final
// read of the per-depth counters for the ray statistics.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "prt_internal.h"

namespace prt {

int set_error(prt_hip_ctx* ctx, const std::string& msg)
{
    if (ctx) ctx->err = msg;
    return 1;
}

int check_hip(prt_hip_ctx* ctx, hipError_t e, const char* what)
{
    if (e == hipSuccess) return 0;
    return set_error(ctx, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace prt

using namespace prt;

namespace {

int use_device(prt_hip_ctx* ctx)
{
    if (!ctx) return 1;
    PRT_TRY(ctx, hipSetDevice(ctx->device));
    return 0;
}

void free_ptr(void* p)
{
    if (p) (void)hipFree(p);
}

int ensure_queues(prt_hip_ctx* ctx, int64_t entries)
{
    if (entries <= ctx->queue_cap) return 0;
    for (int b = 0; b < 2; ++b) {
        free_ptr(ctx->d_rays[b]);
        free_ptr(ctx->d_pix[b]);
        free_ptr(ctx->d_rng[b]);
        ctx->d_rays[b] = nullptr;
        ctx->d_pix[b] = nullptr;
        ctx->d_rng[b] = nullptr;
    }
    free_ptr(ctx->d_hits);
    ctx->d_hits = nullptr;
    ctx->queue_cap = 0;
    for (int b = 0; b < 2; ++b) {
        PRT_TRY(ctx, hipMalloc(&ctx->d_rays[b], sizeof(float4) * 2 * entries));
        PRT_TRY(ctx, hipMalloc(&ctx->d_pix[b], sizeof(int32_t) * entries));
        PRT_TRY(ctx, hipMalloc(&ctx->d_rng[b], sizeof(uint32_t) * entries));
    }
    PRT_TRY(ctx, hipMalloc(&ctx->d_hits, sizeof(float4) * entries));
    ctx->queue_cap = entries;
    return 0;
}

int64_t frame_entries(const prt_hip_ctx* ctx)
{
    return (int64_t)((ctx->width + 7) / 8) * ((ctx->height + 7) / 8) * 64;
}

int check_trace_error(prt_hip_ctx* ctx)
{
    uint32_t e = 0;
    PRT_TRY(ctx, hipMemcpyAsync(&e, ctx->d_counters + 32, 4, hipMemcpyDeviceToHost, ctx->stream));
    PRT_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (e) return set_error(ctx, "trace: traversal stack overflow");
    return 0;
}

int ensure_io(prt_hip_ctx* ctx, int64_t n)
{
    if (n <= ctx->io_cap) return 0;
    free_ptr(ctx->d_io_rays);
    free_ptr(ctx->d_io_hits);
    ctx->d_io_rays = nullptr;
    ctx->d_io_hits = nullptr;
    ctx->io_cap = 0;
    PRT_TRY(ctx, hipMalloc(&ctx->d_io_rays, sizeof(prt_ray) * n));
    PRT_TRY(ctx, hipMalloc(&ctx->d_io_hits, sizeof(prt_hit) * n));
    ctx->io_cap = n;
    return 0;
}

int intersect_impl(prt_hip_ctx* ctx, const prt_ray* d_rays, prt_hit* d_hits, int64_t n,
                   unsigned long long* d_stats)
{
    if (!ctx->bvh_ready) return set_error(ctx, "intersect: no BVH (call prt_hip_bvh_build)");
    if (n < 0 || n > (int64_t)UINT32_MAX / 2) return set_error(ctx, "intersect: bad ray count");
    if (ensure_trace_resources(ctx)) return 1;
    PRT_TRY(ctx, launch_zero_u32(ctx->d_counters, 33, ctx->stream));
    TraceLaunch t{};
    t.rays = reinterpret_cast<const float4*>(d_rays);
    t.hits = reinterpret_cast<float4*>(d_hits);
    t.count_ptr = nullptr;
    t.count = (uint32_t)n;
    t.fetch_ptr = ctx->d_counters + 16;
    t.stats = d_stats;
    PRT_TRY(ctx, launch_trace(ctx, t, d_stats != nullptr));
    return 0;
}

}  // namespace

extern "C" {

int prt_hip_abi_version(void) { return PRT_HIP_ABI_VERSION; }

int prt_hip_create(int device_index, prt_hip_ctx** out_ctx)
{
    if (!out_ctx) return 1;
    *out_ctx = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return 2;
    if (device_index < 0 || device_index >= n) return 3;
    prt_hip_ctx* ctx = new (std::nothrow) prt_hip_ctx();
    if (!ctx) return 4;
    ctx->device = device_index;
    hipDeviceProp_t prop;
    if (hipSetDevice(device_index) != hipSuccess ||
        hipGetDeviceProperties(&prop, device_index) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
        hipMalloc(&ctx->d_counters, 64 * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&ctx->d_stats, 32 * sizeof(unsigned long long)) != hipSuccess) {
        prt_hip_destroy(ctx);
        return 5;
    }
    ctx->num_cus = prop.multiProcessorCount;
    ctx->device_name = prop.name[0] ? std::string(prop.name) : std::string("AMD ") + prop.gcnArchName;
    if (hipMemsetAsync(ctx->d_counters, 0, 64 * sizeof(uint32_t), ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
        prt_hip_destroy(ctx);
        return 5;
    }
    *out_ctx = ctx;
    return 0;
}

int prt_hip_destroy(prt_hip_ctx* ctx)
{
    if (!ctx) return 0;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    void* ptrs[] = {ctx->d_idx, ctx->d_pos, ctx->d_nodes, ctx->d_tris, ctx->d_prim_ids, ctx->d_spill,
                    ctx->d_accum, ctx->d_bgr8, ctx->d_rays[0], ctx->d_rays[1], ctx->d_hits,
                    ctx->d_pix[0], ctx->d_pix[1], ctx->d_rng[0], ctx->d_rng[1], ctx->d_counters,
                    ctx->d_stats, ctx->d_io_rays, ctx->d_io_hits};
    for (void* p : ptrs) free_ptr(p);
    for (hipEvent_t e : ctx->kev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return 0;
}

const char* prt_hip_last_error(prt_hip_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int prt_hip_device_name(prt_hip_ctx* ctx, char* buf, int buf_len)
{
    if (!ctx || !buf || buf_len <= 0) return 1;
    std::strncpy(buf, ctx->device_name.c_str(), (size_t)buf_len - 1);
    buf[buf_len - 1] = '\0';
    return 0;
}

int prt_hip_scene_upload(prt_hip_ctx* ctx, const int32_t* idx3, int64_t tri_count, const float* pos3,
                         int64_t vertex_count)
{
    if (use_device(ctx)) return 1;
    if (tri_count < 0 || vertex_count < 0 || (tri_count > 0 && (!idx3 || !pos3)))
        return set_error(ctx, "scene_upload: bad arguments");
    if (tri_count >= ((int64_t)1 << 31) || vertex_count >= ((int64_t)1 << 31))
        return set_error(ctx, "scene_upload: more than 2^31-1 triangles or vertices");
    for (int64_t i = 0; i < 3 * tri_count; ++i)
        if (idx3[i] < 0 || idx3[i] >= vertex_count)
            return set_error(ctx, "scene_upload: vertex index out of range");
    free_ptr(ctx->d_idx);
    free_ptr(ctx->d_pos);
    ctx->d_idx = nullptr;
    ctx->d_pos = nullptr;
    ctx->bvh_ready = false;
    ctx->tri_count = tri_count;
    ctx->vertex_count = vertex_count;
    PRT_TRY(ctx, hipMalloc(&ctx->d_idx, std::max<int64_t>(16, 12 * tri_count)));
    PRT_TRY(ctx, hipMalloc(&ctx->d_pos, std::max<int64_t>(16, 12 * vertex_count)));
    if (tri_count) {
        PRT_TRY(ctx, hipMemcpyAsync(ctx->d_idx, idx3, 12 * tri_count, hipMemcpyHostToDevice, ctx->stream));
        PRT_TRY(ctx, hipMemcpyAsync(ctx->d_pos, pos3, 12 * vertex_count, hipMemcpyHostToDevice, ctx->stream));
    }
    PRT_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int prt_hip_bvh_build(prt_hip_ctx* ctx, prt_build_stats* out_stats)
{
    if (use_device(ctx)) return 1;
    if (!ctx->d_idx) return set_error(ctx, "bvh_build: no scene uploaded");
    if (build_bvh8(ctx, out_stats)) return 1;
    return ensure_trace_resources(ctx);
}

int prt_hip_intersect_closest_device(prt_hip_ctx* ctx, const prt_ray* d_rays, prt_hit* d_hits, int64_t n)
{
    if (use_device(ctx)) return 1;
    return intersect_impl(ctx, d_rays, d_hits, n, nullptr);
}

int prt_hip_intersect_closest(prt_hip_ctx* ctx, const prt_ray* rays, prt_hit* hits, int64_t n)
{
    if (use_device(ctx)) return 1;
    if (n == 0) return 0;
    if (!rays || !hits) return set_error(ctx, "intersect: null array");
    if (ensure_io(ctx, n)) return 1;
    PRT_TRY(ctx, hipMemcpyAsync(ctx->d_io_rays, rays, sizeof(prt_ray) * n, hipMemcpyHostToDevice, ctx->stream));
    if (intersect_impl(ctx, reinterpret_cast<prt_ray*>(ctx->d_io_rays),
                       reinterpret_cast<prt_hit*>(ctx->d_io_hits), n, nullptr))
        return 1;
    PRT_TRY(ctx, hipMemcpyAsync(hits, ctx->d_io_hits, sizeof(prt_hit) * n, hipMemcpyDeviceToHost, ctx->stream));
    PRT_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return check_trace_error(ctx);
}

int prt_hip_intersect_counters(prt_hip_ctx* ctx, const prt_ray* rays, prt_hit* hits, int64_t n,
                               prt_trace_counters* out)
{
    if (use_device(ctx)) return 1;
    if (!rays || !hits || !out) return set_error(ctx, "intersect_counters: null argument");
    if (n == 0) { std::memset(out, 0, sizeof(*out)); return 0; }
    if (ensure_io(ctx, n)) return 1;
    PRT_TRY(ctx, hipMemsetAsync(ctx->d_stats, 0, 8 * sizeof(unsigned long long), ctx->stream));
    PRT_TRY(ctx, hipMemcpyAsync(ctx->d_io_rays, rays, sizeof(prt_ray) * n, hipMemcpyHostToDevice, ctx->stream));
    if (intersect_impl(ctx, reinterpret_cast<prt_ray*>(ctx->d_io_rays),
                       reinterpret_cast<prt_hit*>(ctx->d_io_hits), n, ctx->d_stats))
        return 1;
    unsigned long long st[8];
    PRT_TRY(ctx, hipMemcpyAsync(hits, ctx->d_io_hits, sizeof(prt_hit) * n, hipMemcpyDeviceToHost, ctx->stream));
    PRT_TRY(ctx, hipMemcpyAsync(st, ctx->d_stats, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
    PRT_TRY(ctx, hipStreamSynchronize(ctx->stream));
    out->rays = n;
    out->nodes = (int64_t)st[0];
    out->tris = (int64_t)st[1];
    out->spills = (int64_t)st[2];
    return check_trace_error(ctx);
}

int prt_hip_view_basis(float angle_x, float angle_y, float out_basis[9])
{
    if (!out_basis) return 1;
    // rotate(vec, axis, a) of math/vec3.h:409-412
    auto rot = [](f3 vec, f3 axis, float a) {
        float c = std::cos(a), s = std::sin(a);
        f3 r = add3(mul3(vec, c), mul3(mul3(axis, dot3(axis, vec)), 1.0f - c));
        return add3(r, mul3(cross3(vec, axis), s));
    };
    const f3 V = mk3(0.0f, 1.0f, 0.0f), N = mk3(0.0f, 0.0f, 1.0f);
    // Basis3f(one).rotateV(angleY).rotateU(angleX) (math/basis3.h:63-79)
    const f3 n1 = rot(N, V, angle_y);
    const f3 u1 = cross3(V, n1);
    const f3 v2 = rot(V, u1, angle_x);
    const f3 n2 = cross3(u1, v2);
    const f3 b[3] = {u1, v2, n2};
    for (int k = 0; k < 3; ++k) {
        out_basis[3 * k + 0] = b[k].x;
        out_basis[3 * k + 1] = b[k].y;
        out_basis[3 * k + 2] = b[k].z;
    }
    return 0;
}

int prt_hip_camera_from_basis(const float position[3], const float basis[9], float fov_y, float aspect_ratio,
                              prt_camera* out)
{
    if (!position || !basis || !out) return 1;
    const f3 u = mk3(basis[0], basis[1], basis[2]);
    const f3 v = mk3(basis[3], basis[4], basis[5]);
    const f3 n = mk3(basis[6], basis[7], basis[8]);
    const float hh = std::tan(0.5f * fov_y);
    const float hw = hh * aspect_ratio;
    auto to_global = [&](f3 a) { return add3(add3(mul3(u, a.x), mul3(v, a.y)), mul3(n, a.z)); };
    const f3 io = to_global(mk3(-hw, hh, -1.0f));
    const f3 dx = to_global(mk3(2.0f * hw, 0.0f, 0.0f));
    const f3 dy = to_global(mk3(0.0f, (float)(-2.0 * (double)hh), 0.0f));
    const f3 vals[6] = {mk3(position[0], position[1], position[2]), io, dx, dy, u, v};
    float* dst[6] = {out->origin, out->image_o, out->image_dx, out->image_dy, out->basis_u, out->basis_v};
    for (int k = 0; k < 6; ++k) {
        dst[k][0] = vals[k].x;
        dst[k][1] = vals[k].y;
        dst[k][2] = vals[k].z;
    }
    out->lens_radius = 0.0f;
    out->focal_distance = 1.0f;
    return 0;
}

int prt_hip_camera_from_view(const float position[3], float angle_x, float angle_y, float fov_y,
                             float aspect_ratio, prt_camera* out)
{
    float basis[9];
    if (!position || !out || prt_hip_view_basis(angle_x, angle_y, basis)) return 1;
    return prt_hip_camera_from_basis(position, basis, fov_y, aspect_ratio, out);
}

int prt_hip_camera_set(prt_hip_ctx* ctx, const prt_camera* camera)
{
    if (use_device(ctx)) return 1;
    if (!camera) return set_error(ctx, "camera_set: null camera");
    if (camera->lens_radius != 0.0f) return set_error(ctx, "camera_set: only the pinhole camera is supported");
    ctx->cam = *camera;
    ctx->cam_ready = true;
    return 0;
}

int prt_hip_frame_init(prt_hip_ctx* ctx, int width, int height)
{
    if (use_device(ctx)) return 1;
    if (width <= 0 || height <= 0 || (int64_t)width * height > ((int64_t)1 << 28))
        return set_error(ctx, "frame_init: bad size");
    free_ptr(ctx->d_accum);
    free_ptr(ctx->d_bgr8);
    ctx->d_accum = nullptr;
    ctx->d_bgr8 = nullptr;
    ctx->width = width;
    ctx->height = height;
    PRT_TRY(ctx, hipMalloc(&ctx->d_accum, sizeof(float4) * (int64_t)width * height));
    PRT_TRY(ctx, hipMalloc(&ctx->d_bgr8, sizeof(int32_t) * (int64_t)width * height));
    if (ensure_queues(ctx, frame_entries(ctx))) return 1;
    return prt_hip_frame_clear(ctx);
}

int prt_hip_frame_clear(prt_hip_ctx* ctx)
{
    if (use_device(ctx)) return 1;
    if (!ctx->d_accum) return set_error(ctx, "frame_clear: no frame");
    PRT_TRY(ctx, hipMemsetAsync(ctx->d_accum, 0, sizeof(float4) * (int64_t)ctx->width * ctx->height, ctx->stream));
    PRT_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->pass = 0;
    return 0;
}

int prt_hip_set_pass(prt_hip_ctx* ctx, int pass)
{
    if (!ctx) return 1;
    ctx->pass = pass;
    return 0;
}

int prt_hip_set_profiling(prt_hip_ctx* ctx, int flags)
{
    if (use_device(ctx)) return 1;
    if (flags & PRT_PROFILE_TIMING) {
        for (hipEvent_t& e : ctx->kev)
            if (!e) PRT_TRY(ctx, hipEventCreate(&e));
    }
    ctx->profiling = flags;
    return 0;
}

int prt_hip_set_pipeline(prt_hip_ctx* ctx, int pipeline)
{
    if (!ctx) return 1;
    if (pipeline != PRT_PIPELINE_FUSED && pipeline != PRT_PIPELINE_WAVEFRONT)
        return set_error(ctx, "set_pipeline: unknown pipeline");
    ctx->pipeline = pipeline;
    return 0;
}

int prt_hip_render(prt_hip_ctx* ctx, int renderer, int max_depth, float albedo, prt_render_stats* out)
{
    if (use_device(ctx)) return 1;
    if (!ctx->bvh_ready) return set_error(ctx, "render: no BVH");
    if (!ctx->cam_ready) return set_error(ctx, "render: no camera");
    if (!ctx->d_accum) return set_error(ctx, "render: no frame");
    if (renderer != PRT_RENDERER_PRIMARY && renderer != PRT_RENDERER_DIFFUSE)
        return set_error(ctx, "render: invalid renderer type");
    if (max_depth < 0 || max_depth + 1 >= kMaxQueueDepth)
        return set_error(ctx, "render: max_depth must be in [0, 14]");
    if (ensure_trace_resources(ctx)) return 1;

    const bool timing = (ctx->profiling & PRT_PROFILE_TIMING) != 0;
    const bool counting = (ctx->profiling & PRT_PROFILE_COUNTERS) != 0;
    const bool fused = ctx->pipeline == PRT_PIPELINE_FUSED;
    const int last = renderer == PRT_RENDERER_PRIMARY ? 0 : max_depth;
    const int tiles_x = (ctx->width + 7) / 8;
    const uint32_t entries = (uint32_t)frame_entries(ctx);

    PRT_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    if (fused) {
        PRT_TRY(ctx, launch_frame_begin(ctx, entries));
        PathLaunch p{};
        p.cam = to_camera(ctx->cam);
        p.width = ctx->width;
        p.height = ctx->height;
        p.tiles_x = tiles_x;
        p.entries = entries;
        p.pass = ctx->pass;
        p.inv_w = 1.0f / (float)ctx->width;
        p.inv_h = 1.0f / (float)ctx->height;
        p.accum = ctx->d_accum;
        p.max_depth = last;
        p.albedo = albedo;
        p.path_counter = ctx->d_counters + 16;
        p.depth_rays = ctx->d_stats + 16;
        p.stats = ctx->d_stats;
        if (timing) PRT_TRY(ctx, hipEventRecord(ctx->kev[0], ctx->stream));
        PRT_TRY(ctx, launch_paths(ctx, p, renderer == PRT_RENDERER_PRIMARY, counting));
        if (timing) PRT_TRY(ctx, hipEventRecord(ctx->kev[1], ctx->stream));
    } else {
        PRT_TRY(ctx, launch_raygen(ctx, ctx->pass, ctx->d_rays[0], ctx->d_pix[0], ctx->d_rng[0]));
        float throughput = 1.0f;
        for (int depth = 0; depth <= last; ++depth) {
            TraceLaunch t{};
            t.rays = ctx->d_rays[depth & 1];
            t.hits = ctx->d_hits;
            t.count_ptr = ctx->d_counters + depth;
            t.count = 0;
            t.fetch_ptr = ctx->d_counters + 16 + depth;
            t.stats = counting ? ctx->d_stats : nullptr;
            if (timing) PRT_TRY(ctx, hipEventRecord(ctx->kev[4 * depth + 0], ctx->stream));
            PRT_TRY(ctx, launch_trace(ctx, t, counting));
            if (timing) PRT_TRY(ctx, hipEventRecord(ctx->kev[4 * depth + 1], ctx->stream));
            if (renderer == PRT_RENDERER_PRIMARY) {
                PRT_TRY(ctx, launch_shade_primary(ctx));
            } else {
                PRT_TRY(ctx, launch_shade_diffuse(ctx, depth, max_depth, throughput, albedo));
                throughput *= albedo;
            }
            if (timing) PRT_TRY(ctx, hipEventRecord(ctx->kev[4 * depth + 2], ctx->stream));
        }
    }
    PRT_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    uint32_t counters[33];
    unsigned long long st[32];
    PRT_TRY(ctx, hipMemcpyAsync(counters, ctx->d_counters, sizeof(counters), hipMemcpyDeviceToHost, ctx->stream));
    PRT_TRY(ctx, hipMemcpyAsync(st, ctx->d_stats, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
    PRT_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (counters[32]) return set_error(ctx, "render: traversal stack overflow");
    float ms = 0.0f;
    PRT_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->pass += 1;
    if (out) {
        std::memset(out, 0, sizeof(*out));
        const int64_t live = (int64_t)ctx->width * ctx->height;
        int64_t rays = 0;
        for (int k = 0; k <= last; ++k) {
            const int64_t q = fused ? (int64_t)st[16 + k] : (k == 0 ? live : (int64_t)counters[k]);
            out->queue[k] = q;
            rays += q;
        }
        if (fused) rays = (int64_t)st[3];
        out->rays = rays;
        out->paths = live;
        out->ms = ms;
        out->mray = ms > 0.0f ? (double)rays / 1000.0 / ms : 0.0;
        if (timing) {
            if (fused) {
                float a = 0.0f;
                PRT_TRY(ctx, hipEventElapsedTime(&a, ctx->kev[0], ctx->kev[1]));
                out->trace_ms = a;
            } else {
                for (int depth = 0; depth <= last; ++depth) {
                    float a = 0.0f, b = 0.0f;
                    PRT_TRY(ctx, hipEventElapsedTime(&a, ctx->kev[4 * depth + 0], ctx->kev[4 * depth + 1]));
                    PRT_TRY(ctx, hipEventElapsedTime(&b, ctx->kev[4 * depth + 1], ctx->kev[4 * depth + 2]));
                    out->trace_ms += a;
                    out->shade_ms += b;
                }
            }
        }
        if (counting) {
            out->nodes = (int64_t)st[0];
            out->tris = (int64_t)st[1];
        }
    }
    return 0;
}

int prt_hip_accum_read(prt_hip_ctx* ctx, float* dst)
{
    if (use_device(ctx)) return 1;
    if (!ctx->d_accum || !dst) return set_error(ctx, "accum_read: no frame");
    PRT_TRY(ctx, hipMemcpyAsync(dst, ctx->d_accum, sizeof(float4) * (int64_t)ctx->width * ctx->height,
                                hipMemcpyDeviceToHost, ctx->stream));
    PRT_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int prt_hip_frame_read_bgr8(prt_hip_ctx* ctx, int32_t* dst)
{
    if (use_device(ctx)) return 1;
    if (!ctx->d_accum || !dst) return set_error(ctx, "frame_read: no frame");
    PRT_TRY(ctx, launch_tonemap(ctx));
    PRT_TRY(ctx, hipMemcpyAsync(dst, ctx->d_bgr8, sizeof(int32_t) * (int64_t)ctx->width * ctx->height,
                                hipMemcpyDeviceToHost, ctx->stream));
    PRT_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int prt_hip_debug_raygen(prt_hip_ctx* ctx, int pass, prt_ray* rays, uint32_t* states)
{
    if (use_device(ctx)) return 1;
    if (!ctx->cam_ready || !ctx->d_accum) return set_error(ctx, "debug_raygen: no camera or frame");
    if (!rays) return set_error(ctx, "debug_raygen: null output");
    const int64_t n = frame_entries(ctx);
    const int64_t np = (int64_t)ctx->width * ctx->height;
    PRT_TRY(ctx, launch_raygen(ctx, pass, ctx->d_rays[0], ctx->d_pix[0], ctx->d_rng[0]));
    prt_ray* d_out = nullptr;
    uint32_t* d_states = nullptr;
    PRT_TRY(ctx, hipMalloc(&d_out, sizeof(prt_ray) * np));
    PRT_TRY(ctx, hipMalloc(&d_states, sizeof(uint32_t) * np));
    hipError_t e = launch_unpack_rays(ctx->d_rays[0], ctx->d_pix[0], ctx->d_rng[0], n, ctx->width,
                                      ctx->height, d_out, d_states, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(rays, d_out, sizeof(prt_ray) * np, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && states)
        e = hipMemcpyAsync(states, d_states, sizeof(uint32_t) * np, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    free_ptr(d_out);
    free_ptr(d_states);
    PRT_TRY(ctx, e);
    return 0;
}

int prt_hip_bvh_info(prt_hip_ctx* ctx, int64_t* node_count, int64_t* tri_count, int32_t* depth)
{
    if (!ctx) return 1;
    if (!ctx->bvh_ready) return set_error(ctx, "bvh_info: no BVH");
    if (node_count) *node_count = ctx->node_count;
    if (tri_count) *tri_count = ctx->tri_count;
    if (depth) *depth = ctx->bvh_depth;
    return 0;
}

int prt_hip_bvh_read(prt_hip_ctx* ctx, void* nodes, float* tri_records, int32_t* prim_ids)
{
    if (use_device(ctx)) return 1;
    if (!ctx->bvh_ready) return set_error(ctx, "bvh_read: no BVH");
    if (nodes)
        PRT_TRY(ctx, hipMemcpyAsync(nodes, ctx->d_nodes, sizeof(Bvh8Node) * ctx->node_count,
                                    hipMemcpyDeviceToHost, ctx->stream));
    if (tri_records)
        PRT_TRY(ctx, hipMemcpyAsync(tri_records, ctx->d_tris, sizeof(float) * 9 * ctx->tri_count,
                                    hipMemcpyDeviceToHost, ctx->stream));
    if (prim_ids)
        PRT_TRY(ctx, hipMemcpyAsync(prim_ids, ctx->d_prim_ids, sizeof(int32_t) * ctx->tri_count,
                                    hipMemcpyDeviceToHost, ctx->stream));
    PRT_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

}  // extern "C"
