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
#include <cstddef>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "prt_internal.h"

// Fixed scratch of prt_hip_intersect_closest_device_noalloc: a 256-byte
// header (u32 [0] work counter, [1] error flag), then the global part of each
// persistent lane's traversal stack.  Every GPU the library runs on holds its
// own copy until the process ends.  In host code the name is only a host-side
// stand-in for the symbol: kernels and copies take the address that
// hipGetSymbolAddress returns for the current device (scratch_address).
static __device__ __align__(256) unsigned char scratch[PRT_HIP_NOALLOC_SCRATCH_BYTES];

namespace prt {

void SceneBuffers::release()
{
    void* ptrs[] = {d_idx, d_pos, d_nodes, d_tris, d_prim_ids, d_normals, d_texcoords};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    d_normals = nullptr;
    d_texcoords = nullptr;
    d_idx = nullptr;
    d_pos = nullptr;
    d_nodes = nullptr;
    d_tris = nullptr;
    d_prim_ids = nullptr;
    bvh_ready = false;
    node_count = 0;
    bvh_depth = 0;
}

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

// Counters and statistics share one device allocation (64 u32, then 32 u64);
// a call reads both back with one copy into the context's pinned buffer.
struct ReadBack {
    uint32_t counters[64];
    unsigned long long st[32];
};
static_assert(sizeof(ReadBack) == 64 * sizeof(uint32_t) + 32 * sizeof(unsigned long long), "read-back layout");

int read_back(prt_hip_ctx* ctx, const ReadBack** out)
{
    PRT_TRY(ctx, hipMemcpyAsync(ctx->h_readback, ctx->d_counters, sizeof(ReadBack), hipMemcpyDeviceToHost,
                                ctx->stream));
    PRT_TRY(ctx, hipStreamSynchronize(ctx->stream));
    *out = static_cast<const ReadBack*>(ctx->h_readback);
    return 0;
}

constexpr size_t kScratchHeader = 256;
constexpr int kScratchMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchMaxDevices] = {};

// Device address of `scratch` on device_index, which must be the current
// device; looked up once per device and kept.
hipError_t scratch_address(int device_index, void** out)
{
    *out = nullptr;
    if (device_index < 0 || device_index >= kScratchMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device_index]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[device_index] = p;
    }
    *out = g_scratch_ptr[device_index];
    return hipSuccess;
}

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

// Replaces the shading attributes of a scene: 3 floats per vertex of normals
// and 2 of texture coordinates; a null pointer removes that attribute.
int set_attributes(prt_hip_ctx* ctx, SceneBuffers& sb, const float* normals3, const float* texcoords2)
{
    free_ptr(sb.d_normals);
    free_ptr(sb.d_texcoords);
    sb.d_normals = nullptr;
    sb.d_texcoords = nullptr;
    const int64_t nv = sb.vertex_count;
    if (normals3 && nv > 0) {
        PRT_TRY(ctx, hipMalloc(&sb.d_normals, 12 * nv));
        PRT_TRY(ctx, hipMemcpyAsync(sb.d_normals, normals3, 12 * nv, hipMemcpyHostToDevice, ctx->stream));
    }
    if (texcoords2 && nv > 0) {
        PRT_TRY(ctx, hipMalloc(&sb.d_texcoords, 8 * nv));
        PRT_TRY(ctx, hipMemcpyAsync(sb.d_texcoords, texcoords2, 8 * nv, hipMemcpyHostToDevice, ctx->stream));
    }
    PRT_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
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

// Adds the trace and shade times of the wavefront launches whose events sit
// in the first n of the 16 per-depth event sets.
int add_wave_timing(prt_hip_ctx* ctx, int n, double& trace_ms, double& shade_ms)
{
    for (int k = 0; k < n; ++k) {
        float a = 0.0f, b = 0.0f;
        PRT_TRY(ctx, hipEventElapsedTime(&a, ctx->kev[4 * k + 0], ctx->kev[4 * k + 1]));
        PRT_TRY(ctx, hipEventElapsedTime(&b, ctx->kev[4 * k + 1], ctx->kev[4 * k + 2]));
        trace_ms += a;
        shade_ms += b;
    }
    return 0;
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

// Rate of the device's constant wall clock (wall_clock64), in Hz.
double wall_clock_hz(prt_hip_ctx* ctx)
{
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device) != hipSuccess || khz <= 0)
        return 1e8;
    return (double)khz * 1e3;
}

int intersect_impl(prt_hip_ctx* ctx, const prt_ray* d_rays, prt_hit* d_hits, int64_t n,
                   unsigned long long* d_stats)
{
    if (!ctx->scene.bvh_ready) return set_error(ctx, "intersect: no BVH (call prt_hip_bvh_build)");
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

int occluded_impl(prt_hip_ctx* ctx, const prt_ray* d_rays, float* d_t, int64_t n)
{
    if (!ctx->scene.bvh_ready) return set_error(ctx, "occluded: no BVH (call prt_hip_bvh_build)");
    if (n < 0 || n > (int64_t)UINT32_MAX / 2) return set_error(ctx, "occluded: bad ray count");
    if (ensure_trace_resources(ctx)) return 1;
    PRT_TRY(ctx, launch_zero_u32(ctx->d_counters, 33, ctx->stream));
    TraceLaunch t{};
    t.rays = reinterpret_cast<const float4*>(d_rays);
    t.any_t = d_t;
    t.count_ptr = nullptr;
    t.count = (uint32_t)n;
    t.fetch_ptr = ctx->d_counters + 16;
    PRT_TRY(ctx, launch_occluded(ctx, t, false));
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
        // counters ([48] batch work counter) and statistics share one
        // allocation, so a frame reads both back with one copy
        hipMalloc(&ctx->d_counters, 64 * sizeof(uint32_t) + 32 * sizeof(unsigned long long)) != hipSuccess ||
        // pinned, so the per-call read-back is one direct copy
        hipHostMalloc(&ctx->h_readback, 64 * sizeof(uint32_t) + 32 * sizeof(unsigned long long),
                      hipHostMallocDefault) != hipSuccess) {
        prt_hip_destroy(ctx);
        return 5;
    }
    // the no-allocation query's scratch on this device (the device is current)
    if (scratch_address(device_index, &ctx->d_scratch) != hipSuccess) {
        prt_hip_destroy(ctx);
        return 5;
    }
    ctx->d_stats = reinterpret_cast<unsigned long long*>(ctx->d_counters + 64);
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
    ctx->scene.release();
    prt_hip_batch_reset(ctx);
    free_ptr(ctx->d_albedo_own);
    free_ptr(ctx->d_rgb_planes);
    void* ptrs[] = {ctx->d_spill,
                    ctx->d_accum, ctx->d_bgr8, ctx->d_rays[0], ctx->d_rays[1], ctx->d_hits,
                    ctx->d_pix[0], ctx->d_pix[1], ctx->d_rng[0], ctx->d_rng[1], ctx->d_counters,
                    ctx->d_io_rays, ctx->d_io_hits, ctx->d_look, ctx->d_look_side, ctx->d_look_flags};
    for (void* p : ptrs) free_ptr(p);
    for (auto& level : ctx->d_look_q)
        for (uint4* q : level) free_ptr(q);
    for (hipEvent_t e : ctx->kev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->h_readback) (void)hipHostFree(ctx->h_readback);
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
    ctx->scene_gen += 1;
    free_ptr(ctx->scene.d_idx);
    free_ptr(ctx->scene.d_pos);
    free_ptr(ctx->scene.d_normals);
    free_ptr(ctx->scene.d_texcoords);
    ctx->scene.d_idx = nullptr;
    ctx->scene.d_pos = nullptr;
    ctx->scene.d_normals = nullptr;
    ctx->scene.d_texcoords = nullptr;
    ctx->scene.bvh_ready = false;
    ctx->scene.tri_count = tri_count;
    ctx->scene.vertex_count = vertex_count;
    free_ptr(ctx->d_albedo_own);
    ctx->d_albedo_own = nullptr;
    ctx->d_albedo = nullptr;
    PRT_TRY(ctx, hipMalloc(&ctx->scene.d_idx, std::max<int64_t>(16, 12 * tri_count)));
    PRT_TRY(ctx, hipMalloc(&ctx->scene.d_pos, std::max<int64_t>(16, 12 * vertex_count)));
    if (tri_count) {
        PRT_TRY(ctx, hipMemcpyAsync(ctx->scene.d_idx, idx3, 12 * tri_count, hipMemcpyHostToDevice, ctx->stream));
        PRT_TRY(ctx, hipMemcpyAsync(ctx->scene.d_pos, pos3, 12 * vertex_count, hipMemcpyHostToDevice, ctx->stream));
    }
    PRT_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int prt_hip_scene_attributes(prt_hip_ctx* ctx, const float* normals3, const float* texcoords2)
{
    if (use_device(ctx)) return 1;
    if (!ctx->scene.d_idx) return set_error(ctx, "scene_attributes: no scene uploaded");
    ctx->scene_gen += 1;
    return set_attributes(ctx, ctx->scene, normals3, texcoords2);
}

int prt_hip_set_bvh_collapse(prt_hip_ctx* ctx, int method, float prim_cost)
{
    if (!ctx) return 1;
    if ((method != PRT_COLLAPSE_GREEDY && method != PRT_COLLAPSE_SAH) || !(prim_cost > 0.0f && prim_cost < 100.0f))
        return set_error(ctx, "set_bvh_collapse: unknown method or prim_cost outside (0, 100)");
    ctx->bvh_collapse = method;
    ctx->bvh_prim_cost = prim_cost;
    return 0;
}

int prt_hip_bvh_build(prt_hip_ctx* ctx, prt_build_stats* out_stats)
{
    if (use_device(ctx)) return 1;
    if (!ctx->scene.d_idx) return set_error(ctx, "bvh_build: no scene uploaded");
    ctx->scene_gen += 1;
    if (build_bvh8(ctx, ctx->scene, out_stats)) return 1;
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

int prt_hip_intersect_closest_device_noalloc(prt_hip_ctx* ctx, const prt_ray* d_rays, prt_hit* d_hits, int64_t n)
{
    if (use_device(ctx)) return 1;
    if (!ctx->scene.bvh_ready || ctx->trace_per_cu <= 0)
        return set_error(ctx, "intersect_noalloc: no BVH (call prt_hip_bvh_build)");
    if (n < 0 || n > (int64_t)UINT32_MAX / 2) return set_error(ctx, "intersect_noalloc: bad ray count");
    if (n == 0) return 0;
    if (!d_rays || !d_hits) return set_error(ctx, "intersect_noalloc: null array");
    if (!ctx->d_scratch) return set_error(ctx, "intersect_noalloc: no scratch on this device");
    // the same worst-case stack as the allocating query; the launch keeps as
    // many persistent workgroups as the fixed scratch holds stacks for
    int need = 7 * (ctx->scene.bvh_depth + 1) + 8;
    if (need < 16) need = 16;
    const size_t per_block = (size_t)kTraceBlock * (size_t)need * sizeof(uint2);
    const size_t blocks_fit = ((size_t)PRT_HIP_NOALLOC_SCRATCH_BYTES - kScratchHeader) / per_block;
    if (blocks_fit < 1)
        return set_error(ctx, "intersect_noalloc: a BVH of depth " + std::to_string(ctx->scene.bvh_depth) +
                                  " needs " + std::to_string(per_block) + " bytes of stack per workgroup, more than the " +
                                  std::to_string((size_t)PRT_HIP_NOALLOC_SCRATCH_BYTES) + "-byte scratch holds");
    int grid = ctx->trace_per_cu * ctx->num_cus;
    if ((size_t)grid > blocks_fit) grid = (int)blocks_fit;
    uint32_t* hdr = static_cast<uint32_t*>(ctx->d_scratch);
    PRT_TRY(ctx, launch_zero_u32(hdr, 2, ctx->stream));  // a kernel, not a memset: safe under graph capture
    TraceLaunch t{};
    t.rays = reinterpret_cast<const float4*>(d_rays);
    t.hits = reinterpret_cast<float4*>(d_hits);
    t.count_ptr = nullptr;
    t.count = (uint32_t)n;
    t.fetch_ptr = hdr;
    t.stats = nullptr;
    PRT_TRY(ctx, launch_trace_scratch(ctx, t, reinterpret_cast<uint2*>(static_cast<char*>(ctx->d_scratch) + kScratchHeader),
                                      need, hdr + 1, grid));
    return 0;
}

int prt_hip_occluded_device(prt_hip_ctx* ctx, const prt_ray* d_rays, float* d_out_t, int64_t n)
{
    if (use_device(ctx)) return 1;
    if (n > 0 && (!d_rays || !d_out_t)) return set_error(ctx, "occluded: null array");
    return occluded_impl(ctx, d_rays, d_out_t, n);
}

int prt_hip_occluded(prt_hip_ctx* ctx, const prt_ray* rays, float* out_t, int64_t n)
{
    if (use_device(ctx)) return 1;
    if (n == 0) return 0;
    if (!rays || !out_t) return set_error(ctx, "occluded: null array");
    if (ensure_io(ctx, n)) return 1;
    float* d_t = reinterpret_cast<float*>(ctx->d_io_hits);  // n floats fit in the n-hit buffer
    PRT_TRY(ctx, hipMemcpyAsync(ctx->d_io_rays, rays, sizeof(prt_ray) * n, hipMemcpyHostToDevice, ctx->stream));
    if (occluded_impl(ctx, reinterpret_cast<prt_ray*>(ctx->d_io_rays), d_t, n)) return 1;
    PRT_TRY(ctx, hipMemcpyAsync(out_t, d_t, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
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
    PRT_TRY(ctx, hipMemsetAsync(ctx->d_stats, 0, 16 * sizeof(unsigned long long), ctx->stream));
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
    if (!(camera->lens_radius >= 0.0f) || !(camera->focal_distance > 0.0f))
        return set_error(ctx, "camera_set: lens radius must be >= 0 and focal distance > 0");
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

int prt_hip_set_tail(prt_hip_ctx* ctx, int mode)
{
    if (!ctx) return 1;
    if (mode != PRT_TAIL_LANES && mode != PRT_TAIL_COOPERATIVE) return set_error(ctx, "set_tail: unknown mode");
    ctx->tail_coop = mode == PRT_TAIL_COOPERATIVE ? 1 : 0;
    return 0;
}

int prt_hip_set_lookahead(prt_hip_ctx* ctx, int enable)
{
    if (!ctx) return 1;
    ctx->lookahead = enable ? 1 : 0;
    ctx->look_prev_on = false;
    return 0;
}

}  // extern "C"

namespace {

// Buffers of the frame lookahead: the state, kLookLevels side entries per
// pixel and, per level, two queues (read by this call, written for the next)
// of one path per lane.
int ensure_look_resources(prt_hip_ctx* ctx)
{
    const int64_t pixels = (int64_t)ctx->width * ctx->height;
    const int64_t side_len = pixels * kLookLevels;
    const uint32_t cap = (uint32_t)ctx->trace_grid * (uint32_t)kTraceBlock;
    bool fresh = false;
    if (!ctx->d_look) {
        PRT_TRY(ctx, hipMalloc(&ctx->d_look, sizeof(LookState)));
        PRT_TRY(ctx, hipMemsetAsync(ctx->d_look, 0, sizeof(LookState), ctx->stream));
        fresh = true;
    }
    if (side_len != ctx->look_side_len) {
        free_ptr(ctx->d_look_side);
        ctx->d_look_side = nullptr;
        ctx->look_side_len = 0;
        PRT_TRY(ctx, hipMalloc(&ctx->d_look_side, sizeof(uint2) * side_len));
        PRT_TRY(ctx, hipMemsetAsync(ctx->d_look_side, 0, sizeof(uint2) * side_len, ctx->stream));
        ctx->look_side_len = side_len;
        fresh = true;
    }
    if (cap > ctx->look_q_cap) {
        for (auto& level : ctx->d_look_q)
            for (uint4*& q : level) {
                free_ptr(q);
                q = nullptr;
            }
        free_ptr(ctx->d_look_flags);
        ctx->d_look_flags = nullptr;
        ctx->look_q_cap = 0;
        for (auto& level : ctx->d_look_q)
            for (uint4*& q : level) PRT_TRY(ctx, hipMalloc(&q, 3 * sizeof(uint4) * (size_t)cap));
        const size_t flag_bytes = sizeof(uint32_t) * kLookFlagStride * (size_t)ctx->trace_grid;
        PRT_TRY(ctx, hipMalloc(&ctx->d_look_flags, flag_bytes));
        PRT_TRY(ctx, hipMemsetAsync(ctx->d_look_flags, 0, flag_bytes, ctx->stream));
        ctx->look_q_cap = cap;
        fresh = true;
    }
    if (fresh) ctx->look_prev_spec = false;  // nothing the previous call left is kept
    return 0;
}

// The previous call's lookahead work belongs to this call: the same inputs and
// the next pass.
bool look_follows(const prt_hip_ctx::LookKey& prev, const prt_hip_ctx::LookKey& cur)
{
    return prev.scene_gen == cur.scene_gen && std::memcmp(&prev.cam, &cur.cam, sizeof(prt_camera)) == 0 &&
           prev.width == cur.width && prev.height == cur.height && prev.max_depth == cur.max_depth &&
           prev.grid == cur.grid && prev.albedo_bits == cur.albedo_bits && prev.pass + 1 == cur.pass;
}

}  // namespace

extern "C" {

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
    if (!ctx->scene.bvh_ready) return set_error(ctx, "render: no BVH");
    if (!ctx->cam_ready) return set_error(ctx, "render: no camera");
    if (!ctx->d_accum) return set_error(ctx, "render: no frame");
    const bool fast = (renderer & PRT_RENDERER_FAST) != 0;
    renderer &= ~PRT_RENDERER_FAST;
    if (renderer != PRT_RENDERER_PRIMARY && renderer != PRT_RENDERER_DIFFUSE && renderer != PRT_RENDERER_DIFFUSE2)
        return set_error(ctx, "render: invalid renderer type");
    if (renderer == PRT_RENDERER_DIFFUSE2 && ctx->pipeline != PRT_PIPELINE_FUSED)
        return set_error(ctx, "render: diffuse2 runs on the fused pipeline only");
    if (max_depth < 0 || max_depth > kMaxDepth)
        return set_error(ctx, "render: max_depth must be in [0, " + std::to_string(kMaxDepth) + "]");
    if (ensure_trace_resources(ctx)) return 1;

    const bool timing = (ctx->profiling & PRT_PROFILE_TIMING) != 0;
    const bool counting = (ctx->profiling & PRT_PROFILE_COUNTERS) != 0;
    const bool fused = ctx->pipeline == PRT_PIPELINE_FUSED;
    const int last = renderer == PRT_RENDERER_PRIMARY ? 0 : max_depth;
    int wave_last = last;  // wavefront: the last depth whose queue counter is meaningful
    double wave_trace_ms = 0.0, wave_shade_ms = 0.0;  // wavefront timing of the depths already read
    int pending_timed = 0;  // ... and the number of depths whose events are still to be read
    const int tiles_x = (ctx->width + 7) / 8;
    const uint32_t entries = (uint32_t)frame_entries(ctx);

    // Lookahead: only in a call that continues the previous call's inputs
    // (the next pass, same scene, camera, size, depth and albedo), so a
    // camera that moves every frame renders each frame exactly as without
    // it; the work is kept only if that previous call ran the lookahead too.
    const bool look_on = fused && renderer == PRT_RENDERER_DIFFUSE && !fast && !counting && ctx->lookahead != 0 &&
                         ctx->pass < INT32_MAX - kLookLevels;  // the passes started ahead must not overflow
    prt_hip_ctx::LookKey key{};
    if (look_on) {
        key.scene_gen = ctx->scene_gen;
        key.cam = ctx->cam;
        key.width = ctx->width;
        key.height = ctx->height;
        key.max_depth = last;
        key.pass = ctx->pass;
        key.grid = ctx->diffuse_grid;
        std::memcpy(&key.albedo_bits, &albedo, sizeof(float));
    }
    const bool look_run = look_on && ctx->look_prev_on && look_follows(ctx->look_prev, key);
    ctx->look_prev_on = false;  // set again once this call has succeeded
    if (look_run && ensure_look_resources(ctx)) return 1;
    const bool look_valid = look_run && ctx->look_prev_spec;

    PRT_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    if (fused) {
        // a lookahead call clears the counters in its hand-over launch
        if (!look_run) PRT_TRY(ctx, launch_frame_begin(ctx, entries));
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
        p.fast = fast;
        p.coop = ctx->tail_coop;
        if (look_run) {
            // a call that keeps no earlier work skips kLookLevels tags, so no
            // side sample left by an earlier chain of calls carries its tag
            // or a later one's
            ctx->look_gen += look_valid ? 1u : (uint32_t)kLookLevels + 1u;
            PRT_TRY(ctx, launch_look_call_begin(ctx, look_valid, entries));
            p.look = ctx->d_look;
            p.side = ctx->d_look_side;
            for (int l = 0; l < kLookLevels; ++l) {
                p.q_in[l] = ctx->d_look_q[l][(ctx->look_gen - 1) & 1];
                p.q_out[l] = ctx->d_look_q[l][ctx->look_gen & 1];
            }
            p.q_cap = ctx->look_q_cap;
            p.look_gen = ctx->look_gen;
            p.look_flags = ctx->d_look_flags;
            p.path_counter = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(ctx->d_look) +
                                                         offsetof(LookState, counter));
        }
        if (timing) PRT_TRY(ctx, hipEventRecord(ctx->kev[0], ctx->stream));
        PRT_TRY(ctx, launch_paths(ctx, p, renderer, counting));
        if (timing) PRT_TRY(ctx, hipEventRecord(ctx->kev[1], ctx->stream));
    } else {
        PRT_TRY(ctx, launch_raygen(ctx, ctx->pass, ctx->d_rays[0], ctx->d_pix[0], ctx->d_rng[0]));
        float throughput = 1.0f;
        int traced = -1;  // deepest depth launched
        bool stopped = false;  // left the loop early: no path left
        for (int depth = 0; depth <= last; ++depth) {
            traced = depth;
            // queue size and fetch counter of depth k live in slot k mod 16
            const int slot = depth & (kMaxQueueDepth - 1);
            TraceLaunch t{};
            t.rays = ctx->d_rays[depth & 1];
            t.hits = ctx->d_hits;
            t.count_ptr = ctx->d_counters + slot;
            t.count = 0;
            t.fetch_ptr = ctx->d_counters + 16 + slot;
            t.stats = counting ? ctx->d_stats : nullptr;
            const int ev = slot;  // timing events of depth k: set k mod 16, read every 16 depths
            if (timing) PRT_TRY(ctx, hipEventRecord(ctx->kev[4 * ev + 0], ctx->stream));
            PRT_TRY(ctx, launch_trace(ctx, t, counting));
            if (timing) PRT_TRY(ctx, hipEventRecord(ctx->kev[4 * ev + 1], ctx->stream));
            if (renderer == PRT_RENDERER_PRIMARY) {
                PRT_TRY(ctx, launch_shade_primary(ctx, fast));
            } else {
                // the next depth's slot was last used by depth + 1 - 16: move
                // its count to the statistics and clear it first
                if (depth + 1 >= kMaxQueueDepth && depth < last)
                    PRT_TRY(ctx, launch_queue_recycle(ctx, depth + 1 - kMaxQueueDepth));
                PRT_TRY(ctx, launch_shade_diffuse(ctx, depth, max_depth, throughput, fast));
                throughput *= albedo;
            }
            if (timing) PRT_TRY(ctx, hipEventRecord(ctx->kev[4 * ev + 2], ctx->stream));
            // deep frames: stop once no path is left (as the reference's
            // while (count > 0)), checked every 16 depths, where the timing
            // events of those 16 depths are read before they are reused
            if (renderer != PRT_RENDERER_PRIMARY && depth < last && slot == kMaxQueueDepth - 1) {
                uint32_t left = 0;
                PRT_TRY(ctx, hipMemcpyAsync(&left, ctx->d_counters + ((depth + 1) & (kMaxQueueDepth - 1)), 4,
                                            hipMemcpyDeviceToHost, ctx->stream));
                PRT_TRY(ctx, hipStreamSynchronize(ctx->stream));
                if (timing && add_wave_timing(ctx, kMaxQueueDepth, wave_trace_ms, wave_shade_ms)) return 1;
                if (left == 0) {
                    wave_last = depth + 1;  // the remaining depths trace nothing
                    stopped = true;
                    break;
                }
            }
        }
        // events not yet read: those since the last 16-depth check (none after a stop)
        if (timing && !stopped) pending_timed = (traced & (kMaxQueueDepth - 1)) + 1;
    }
    PRT_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    const ReadBack* rb = nullptr;
    if (read_back(ctx, &rb)) return 1;
    const uint32_t* counters = rb->counters;
    const unsigned long long* st = rb->st;
    if (counters[32])
        return set_error(ctx, (counters[32] & 2u) ? "render: lookahead queue overflow"
                                                  : "render: traversal stack overflow");
    float ms = 0.0f;
    PRT_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (look_on) {
        ctx->look_prev = key;
        ctx->look_prev_on = true;
        ctx->look_prev_spec = look_run;
    }
    ctx->pass += 1;
    if (out) {
        std::memset(out, 0, sizeof(*out));
        const int64_t live = (int64_t)ctx->width * ctx->height;
        int64_t rays = 0;
        if (fused) {
            // per-depth bins filled by the instrumented kernel (last bin: depths >= 15)
            for (int k = 0; k < kMaxQueueDepth; ++k) out->queue[k] = (int64_t)st[16 + k];
            rays = (int64_t)st[3];
        } else {
            // depth 0: every pixel; depths whose ring slot was reused later
            // were moved to the statistics bins by the recycle launches
            const int end = std::min(wave_last, last);
            for (int k = 0; k <= end; ++k) {
                const bool recycled = k + kMaxQueueDepth <= end;
                const int64_t q = k == 0 ? live : (recycled ? 0 : (int64_t)counters[k & (kMaxQueueDepth - 1)]);
                out->queue[std::min(k, kMaxQueueDepth - 1)] += q;
            }
            for (int b = 1; b < kMaxQueueDepth; ++b) out->queue[b] += (int64_t)st[16 + b];
            for (int b = 0; b < kMaxQueueDepth; ++b) rays += out->queue[b];
        }
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
                if (add_wave_timing(ctx, pending_timed, wave_trace_ms, wave_shade_ms)) return 1;
                out->trace_ms = wave_trace_ms;
                out->shade_ms = wave_shade_ms;
            }
        }
        if (counting) {
            out->nodes = (int64_t)st[0];
            out->tris = (int64_t)st[1];
            out->spills = (int64_t)st[2];
            out->wave_steps = (int64_t)st[4];
            out->lane_steps = (int64_t)st[5];
            out->wave_shades = (int64_t)st[6];
            out->drain_wave_steps = (int64_t)st[7];
            out->drain_lane_steps = (int64_t)st[8];
            out->drain_ms = st[10] > st[9] ? (double)(st[10] - st[9]) * 1e3 / wall_clock_hz(ctx) : 0.0;
        }
    }
    return 0;
}

}  // extern "C"

namespace {

void batch_drop_graph(prt_hip_ctx* ctx)
{
    if (ctx->batch_graph_exec) (void)hipGraphExecDestroy(ctx->batch_graph_exec);
    if (ctx->batch_graph) (void)hipGraphDestroy(ctx->batch_graph);
    ctx->batch_graph_exec = nullptr;
    ctx->batch_graph = nullptr;
}

void batch_free_jobs(prt_hip_ctx* ctx)
{
    batch_drop_graph(ctx);
    void* ptrs[] = {ctx->d_batch_jobs, ctx->d_batch_work_begin, ctx->d_batch_pixel_begin, ctx->d_batch_samples,
                    ctx->d_batch_accum};
    for (void* p : ptrs) free_ptr(p);
    ctx->d_batch_jobs = nullptr;
    ctx->d_batch_work_begin = nullptr;
    ctx->d_batch_pixel_begin = nullptr;
    ctx->d_batch_samples = nullptr;
    ctx->d_batch_accum = nullptr;
    ctx->batch_jobs.clear();
    ctx->batch_samples_len = ctx->batch_pixels = 0;
    ctx->batch_total_work = 0;
}

BatchLaunch batch_launch_args(prt_hip_ctx* ctx)
{
    BatchLaunch bl{};
    bl.jobs = ctx->d_batch_jobs;
    bl.work_begin = ctx->d_batch_work_begin;
    bl.n_jobs = (int)ctx->batch_jobs.size();
    bl.scenes = ctx->d_batch_scenes;
    bl.samples = ctx->d_batch_samples;
    bl.total_work = ctx->batch_total_work;
    bl.work_counter = ctx->d_counters + 48;
    bl.stats = ctx->d_stats;
    return bl;
}

int batch_resolve_tuning(prt_hip_ctx* ctx, BatchTuning& tu)
{
    if (tu.variant < 0 || tu.variant >= kBatchVariantCount) return set_error(ctx, "batch: bad kernel variant");
    int max_per_cu = 0;
    PRT_TRY(ctx, launch_batch(ctx, BatchLaunch{}, tu, false, &max_per_cu));
    const int lane_cap = 2048 / kBatchVariants[tu.variant].block;
    if (max_per_cu > lane_cap) max_per_cu = lane_cap;
    if (max_per_cu < 1) max_per_cu = 1;
    if (tu.blocks_per_cu <= 0 || tu.blocks_per_cu > max_per_cu) tu.blocks_per_cu = max_per_cu;
    if (tu.refill < 1) tu.refill = 1;
    if (tu.refill > 64) tu.refill = 64;
    return 0;
}

int batch_enqueue(prt_hip_ctx* ctx, const BatchTuning& tu, bool instrumented, bool full)
{
    PRT_TRY(ctx, launch_batch_begin(ctx));
    PRT_TRY(ctx, launch_batch(ctx, batch_launch_args(ctx), tu, instrumented, nullptr));
    if (full) {
        PRT_TRY(ctx, launch_batch_finalize(ctx));
        PRT_TRY(ctx, launch_batch_advance(ctx));
    }
    return 0;
}

}  // namespace

extern "C" {

int prt_hip_batch_scene_add(prt_hip_ctx* ctx, const int32_t* idx3, int64_t tri_count, const float* pos3,
                            int64_t vertex_count, prt_build_stats* out_stats, int32_t* out_scene)
{
    if (use_device(ctx)) return 1;
    if (tri_count <= 0 || vertex_count <= 0 || !idx3 || !pos3 || !out_scene)
        return set_error(ctx, "batch_scene_add: bad arguments");
    if (tri_count >= ((int64_t)1 << 31) || vertex_count >= ((int64_t)1 << 31))
        return set_error(ctx, "batch_scene_add: more than 2^31-1 triangles or vertices");
    for (int64_t i = 0; i < 3 * tri_count; ++i)
        if (idx3[i] < 0 || idx3[i] >= vertex_count) return set_error(ctx, "batch_scene_add: vertex index out of range");
    SceneBuffers sb;
    sb.tri_count = tri_count;
    sb.vertex_count = vertex_count;
    hipError_t e = hipMalloc(&sb.d_idx, 12 * tri_count);
    if (e == hipSuccess) e = hipMalloc(&sb.d_pos, 12 * vertex_count);
    if (e == hipSuccess) e = hipMemcpyAsync(sb.d_idx, idx3, 12 * tri_count, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(sb.d_pos, pos3, 12 * vertex_count, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        sb.release();
        return check_hip(ctx, e, "batch_scene_add");
    }
    if (build_bvh8(ctx, sb, out_stats)) {
        sb.release();
        return 1;
    }
    batch_free_jobs(ctx);  // jobs refer to the scene table, which is rebuilt
    ctx->batch_scenes.push_back(sb);
    free_ptr(ctx->d_batch_scenes);
    ctx->d_batch_scenes = nullptr;
    std::vector<SceneDevice> table;
    for (const SceneBuffers& b : ctx->batch_scenes) table.push_back(b.device());
    PRT_TRY(ctx, hipMalloc(&ctx->d_batch_scenes, sizeof(SceneDevice) * table.size()));
    PRT_TRY(ctx, hipMemcpy(ctx->d_batch_scenes, table.data(), sizeof(SceneDevice) * table.size(),
                           hipMemcpyHostToDevice));
    *out_scene = (int32_t)ctx->batch_scenes.size() - 1;
    return ensure_trace_resources(ctx);
}

int prt_hip_batch_scene_attributes(prt_hip_ctx* ctx, int32_t scene, const float* normals3,
                                   const float* texcoords2)
{
    if (use_device(ctx)) return 1;
    if (scene < 0 || scene >= (int32_t)ctx->batch_scenes.size())
        return set_error(ctx, "batch_scene_attributes: no such scene");
    if (set_attributes(ctx, ctx->batch_scenes[scene], normals3, texcoords2)) return 1;
    // the table keeps its address (a captured launch graph refers to it)
    const SceneDevice d = ctx->batch_scenes[scene].device();
    PRT_TRY(ctx, hipMemcpy(ctx->d_batch_scenes + scene, &d, sizeof(SceneDevice), hipMemcpyHostToDevice));
    return 0;
}

int prt_hip_batch_reset(prt_hip_ctx* ctx)
{
    if (!ctx) return 1;
    (void)hipSetDevice(ctx->device);
    batch_free_jobs(ctx);
    for (SceneBuffers& b : ctx->batch_scenes) b.release();
    ctx->batch_scenes.clear();
    free_ptr(ctx->d_batch_scenes);
    ctx->d_batch_scenes = nullptr;
    return 0;
}

int prt_hip_batch_set_jobs(prt_hip_ctx* ctx, const prt_batch_job* jobs, int n_jobs)
{
    if (use_device(ctx)) return 1;
    if (!jobs || n_jobs <= 0) return set_error(ctx, "batch_set_jobs: no jobs");
    std::vector<BatchJobDevice> dj(n_jobs);
    std::vector<uint32_t> work_begin(n_jobs + 1);
    std::vector<uint64_t> pixel_begin(n_jobs + 1);
    uint64_t work = 0, samples = 0, pixels = 0;
    for (int j = 0; j < n_jobs; ++j) {
        const prt_batch_job& in = jobs[j];
        if (in.scene < 0 || in.scene >= (int)ctx->batch_scenes.size())
            return set_error(ctx, "batch_set_jobs: unknown scene");
        if (in.width <= 0 || in.height <= 0 || (int64_t)in.width * in.height > ((int64_t)1 << 28))
            return set_error(ctx, "batch_set_jobs: bad image size");
        if (in.spp < 1 || in.max_depth < 0 || in.max_depth > kMaxDepth)
            return set_error(ctx, "batch_set_jobs: spp must be >= 1 and max_depth in [0, " +
                                      std::to_string(kMaxDepth) + "]");
        if (!(in.camera.lens_radius >= 0.0f) || !(in.camera.focal_distance > 0.0f))
            return set_error(ctx, "batch_set_jobs: lens radius must be >= 0 and focal distance > 0");
        BatchJobDevice& d = dj[j];
        d.cam = to_camera(in.camera);
        d.scene = in.scene;
        d.width = in.width;
        d.height = in.height;
        d.tiles_x = (in.width + 7) / 8;
        d.entries = (uint32_t)((int64_t)d.tiles_x * ((in.height + 7) / 8) * 64);
        d.spp = in.spp;
        d.pass0 = in.pass0;
        d.max_depth = in.max_depth;
        d.albedo = in.albedo;
        d.inv_w = 1.0f / (float)in.width;
        d.inv_h = 1.0f / (float)in.height;
        d.work_begin = (uint32_t)work;
        d.sample_base = samples;
        d.pixel_base = pixels;
        work_begin[j] = (uint32_t)work;
        pixel_begin[j] = pixels;
        work += (uint64_t)d.entries * (uint64_t)in.spp;
        samples += (uint64_t)in.spp * (uint64_t)in.width * (uint64_t)in.height;
        pixels += (uint64_t)in.width * (uint64_t)in.height;
        if (work >= (uint64_t)UINT32_MAX - (1u << 20))
            return set_error(ctx, "batch_set_jobs: more than 2^32 work items; split the batch");
    }
    work_begin[n_jobs] = (uint32_t)work;
    pixel_begin[n_jobs] = pixels;
    batch_free_jobs(ctx);
    PRT_TRY(ctx, hipMalloc(&ctx->d_batch_jobs, sizeof(BatchJobDevice) * n_jobs));
    PRT_TRY(ctx, hipMalloc(&ctx->d_batch_work_begin, sizeof(uint32_t) * (n_jobs + 1)));
    PRT_TRY(ctx, hipMalloc(&ctx->d_batch_pixel_begin, sizeof(uint64_t) * (n_jobs + 1)));
    PRT_TRY(ctx, hipMalloc(&ctx->d_batch_samples, sizeof(float) * samples));
    PRT_TRY(ctx, hipMalloc(&ctx->d_batch_accum, sizeof(float4) * pixels));
    PRT_TRY(ctx, hipMemcpy(ctx->d_batch_jobs, dj.data(), sizeof(BatchJobDevice) * n_jobs, hipMemcpyHostToDevice));
    PRT_TRY(ctx, hipMemcpy(ctx->d_batch_work_begin, work_begin.data(), sizeof(uint32_t) * (n_jobs + 1),
                           hipMemcpyHostToDevice));
    PRT_TRY(ctx, hipMemcpy(ctx->d_batch_pixel_begin, pixel_begin.data(), sizeof(uint64_t) * (n_jobs + 1),
                           hipMemcpyHostToDevice));
    PRT_TRY(ctx, hipMemset(ctx->d_batch_accum, 0, sizeof(float4) * pixels));
    ctx->batch_jobs = dj;
    ctx->batch_samples_len = samples;
    ctx->batch_pixels = pixels;
    ctx->batch_total_work = (uint32_t)work;
    return batch_resolve_tuning(ctx, ctx->batch_tuning);
}

int prt_hip_batch_render(prt_hip_ctx* ctx, prt_batch_stats* out)
{
    if (use_device(ctx)) return 1;
    if (ctx->batch_jobs.empty()) return set_error(ctx, "batch_render: no jobs");
    if (ensure_trace_resources(ctx)) return 1;
    const bool counting = (ctx->profiling & PRT_PROFILE_COUNTERS) != 0;
    const bool graph = ctx->batch_use_graph && !counting;
    // the graph holds the spill buffer and its per-lane stride as kernel
    // arguments: a later build of a deeper scene replaces both
    if (ctx->batch_graph_exec && ctx->batch_graph_spill_gen != ctx->spill_gen) batch_drop_graph(ctx);
    if (graph && !ctx->batch_graph_exec) {
        PRT_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const int rc = batch_enqueue(ctx, ctx->batch_tuning, false, true);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        PRT_TRY(ctx, e);
        ctx->batch_graph = g;
        PRT_TRY(ctx, hipGraphInstantiate(&ctx->batch_graph_exec, g, nullptr, nullptr, 0));
        ctx->batch_graph_spill_gen = ctx->spill_gen;
    }
    PRT_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    if (graph) {
        PRT_TRY(ctx, hipGraphLaunch(ctx->batch_graph_exec, ctx->stream));
    } else if (batch_enqueue(ctx, ctx->batch_tuning, counting, true)) {
        return 1;
    }
    PRT_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    const ReadBack* rb = nullptr;
    if (read_back(ctx, &rb)) return 1;
    const unsigned long long* st = rb->st;
    if (rb->counters[32]) return set_error(ctx, "batch_render: traversal stack overflow");
    float ms = 0.0f;
    PRT_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    int64_t paths = 0;
    for (BatchJobDevice& j : ctx->batch_jobs) {
        paths += (int64_t)j.spp * j.width * j.height;
        j.pass0 += j.spp;
    }
    if (out) {
        std::memset(out, 0, sizeof(*out));
        out->rays = (int64_t)st[3];
        out->paths = paths;
        out->ms = ms;
        out->mray = ms > 0.0f ? (double)st[3] / 1000.0 / ms : 0.0;
        if (counting) {
            out->nodes = (int64_t)st[0];
            out->tris = (int64_t)st[1];
        }
    }
    return 0;
}

int prt_hip_batch_accum_read(prt_hip_ctx* ctx, int job, float* dst)
{
    if (use_device(ctx)) return 1;
    if (job < 0 || job >= (int)ctx->batch_jobs.size() || !dst) return set_error(ctx, "batch_accum_read: bad job");
    const BatchJobDevice& j = ctx->batch_jobs[job];
    PRT_TRY(ctx, hipMemcpy(dst, ctx->d_batch_accum + j.pixel_base, sizeof(float4) * (size_t)j.width * j.height,
                           hipMemcpyDeviceToHost));
    return 0;
}

int prt_hip_batch_set_tuning(prt_hip_ctx* ctx, const prt_batch_tuning* t)
{
    if (use_device(ctx)) return 1;
    if (!t) return set_error(ctx, "batch_set_tuning: null");
    BatchTuning tu{t->variant, t->blocks_per_cu, t->refill};
    if (batch_resolve_tuning(ctx, tu)) return 1;
    ctx->batch_tuning = tu;
    batch_drop_graph(ctx);
    return 0;
}

int prt_hip_batch_get_tuning(prt_hip_ctx* ctx, prt_batch_tuning* t)
{
    if (!ctx || !t) return 1;
    std::memset(t, 0, sizeof(*t));
    t->variant = ctx->batch_tuning.variant;
    t->block = kBatchVariants[t->variant].block;
    t->stack = kBatchVariants[t->variant].stack;
    t->blocks_per_cu = ctx->batch_tuning.blocks_per_cu;
    t->refill = ctx->batch_tuning.refill;
    return 0;
}

int prt_hip_batch_autotune(prt_hip_ctx* ctx, prt_batch_tuning* results, int max_results, int* n_results)
{
    if (use_device(ctx)) return 1;
    if (ctx->batch_jobs.empty()) return set_error(ctx, "batch_autotune: no jobs");
    if (ensure_trace_resources(ctx)) return 1;
    std::vector<prt_batch_tuning> all;
    BatchTuning best{};
    double best_ms = 1e300;
    const int refills[2] = {16, 32};
    for (int v = 0; v < kBatchVariantCount; ++v) {
        BatchTuning probe{v, 0, 24};
        if (batch_resolve_tuning(ctx, probe)) return 1;
        const int max_bpc = probe.blocks_per_cu;
        const int bpcs[2] = {max_bpc, max_bpc > 1 ? (max_bpc + 1) / 2 : 0};
        for (int b = 0; b < 2; ++b) {
            if (bpcs[b] <= 0 || (b == 1 && bpcs[1] == bpcs[0])) continue;
            for (int r = 0; r < 2; ++r) {
                BatchTuning tu{v, bpcs[b], refills[r]};
                double t_best = 1e300;
                for (int rep = 0; rep < 2; ++rep) {
                    PRT_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
                    if (batch_enqueue(ctx, tu, false, false)) return 1;
                    PRT_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
                    PRT_TRY(ctx, hipEventSynchronize(ctx->ev1));
                    float ms = 0.0f;
                    PRT_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
                    t_best = ms < t_best ? ms : t_best;
                }
                prt_batch_tuning rt{};
                rt.variant = v;
                rt.block = kBatchVariants[v].block;
                rt.stack = kBatchVariants[v].stack;
                rt.blocks_per_cu = tu.blocks_per_cu;
                rt.refill = tu.refill;
                rt.ms = t_best;
                all.push_back(rt);
                if (t_best < best_ms) { best_ms = t_best; best = tu; }
            }
        }
    }
    uint32_t err = 0;
    PRT_TRY(ctx, hipMemcpy(&err, ctx->d_counters + 32, 4, hipMemcpyDeviceToHost));
    if (err) return set_error(ctx, "batch_autotune: traversal stack overflow");
    ctx->batch_tuning = best;
    batch_drop_graph(ctx);
    if (n_results) *n_results = (int)all.size();
    if (results)
        for (int i = 0; i < (int)all.size() && i < max_results; ++i) results[i] = all[i];
    return 0;
}

int prt_hip_batch_set_graph(prt_hip_ctx* ctx, int enable)
{
    if (!ctx) return 1;
    ctx->batch_use_graph = enable != 0;
    batch_drop_graph(ctx);
    return 0;
}

}  // extern "C"

namespace {

PathLaunch frame_path_launch(prt_hip_ctx* ctx, int pass, int max_depth)
{
    PathLaunch p{};
    p.cam = to_camera(ctx->cam);
    p.width = ctx->width;
    p.height = ctx->height;
    p.tiles_x = (ctx->width + 7) / 8;
    p.entries = (uint32_t)frame_entries(ctx);
    p.pass = pass;
    p.inv_w = 1.0f / (float)ctx->width;
    p.inv_h = 1.0f / (float)ctx->height;
    p.accum = ctx->d_accum;
    p.max_depth = max_depth;
    p.albedo = 0.0f;
    p.path_counter = ctx->d_counters + 16;
    p.depth_rays = ctx->d_stats + 16;
    p.stats = ctx->d_stats;
    p.albedo_rgb = ctx->d_albedo;
    p.spp = 1;
    p.inv_spp = 1.0f;
    return p;
}

int albedo_ready(prt_hip_ctx* ctx, int spp, int max_depth, const char* what)
{
    if (!ctx->scene.bvh_ready) return set_error(ctx, std::string(what) + ": no BVH");
    if (!ctx->cam_ready || !ctx->d_accum) return set_error(ctx, std::string(what) + ": no camera or frame");
    if (!ctx->d_albedo) return set_error(ctx, std::string(what) + ": no per-triangle albedo");
    if (spp < 1 || max_depth < 0 || max_depth > kMaxDepth)
        return set_error(ctx, std::string(what) + ": spp must be >= 1 and max_depth in [0, " +
                                  std::to_string(kMaxDepth) + "]");
    return ensure_trace_resources(ctx);
}

int read_path_stats(prt_hip_ctx* ctx, prt_render_stats* out, int64_t paths)
{
    const ReadBack* rb = nullptr;
    if (read_back(ctx, &rb)) return 1;
    const unsigned long long* st = rb->st;
    if (rb->counters[32]) return set_error(ctx, "traversal stack overflow");
    float ms = 0.0f;
    PRT_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (out) {
        std::memset(out, 0, sizeof(*out));
        out->rays = (int64_t)st[3];
        out->paths = paths;
        out->ms = ms;
        out->mray = ms > 0.0f ? (double)st[3] / 1000.0 / ms : 0.0;
        out->trace_ms = ms;
        if (ctx->profiling & PRT_PROFILE_COUNTERS) {
            out->nodes = (int64_t)st[0];
            out->tris = (int64_t)st[1];
            out->spills = (int64_t)st[2];
            out->wave_steps = (int64_t)st[4];
            out->lane_steps = (int64_t)st[5];
            out->wave_shades = (int64_t)st[6];
            out->drain_wave_steps = (int64_t)st[7];
            out->drain_lane_steps = (int64_t)st[8];
            out->drain_ms = st[10] > st[9] ? (double)(st[10] - st[9]) * 1e3 / wall_clock_hz(ctx) : 0.0;
        }
    }
    return 0;
}

}  // namespace

extern "C" {

int prt_hip_set_albedo_host(prt_hip_ctx* ctx, const float* rgb)
{
    if (use_device(ctx)) return 1;
    if (!rgb || ctx->scene.tri_count <= 0) return set_error(ctx, "set_albedo: no scene or null array");
    free_ptr(ctx->d_albedo_own);
    ctx->d_albedo_own = nullptr;
    PRT_TRY(ctx, hipMalloc(&ctx->d_albedo_own, sizeof(float) * 3 * ctx->scene.tri_count));
    PRT_TRY(ctx, hipMemcpy(ctx->d_albedo_own, rgb, sizeof(float) * 3 * ctx->scene.tri_count, hipMemcpyHostToDevice));
    ctx->d_albedo = ctx->d_albedo_own;
    return 0;
}

int prt_hip_set_albedo_device(prt_hip_ctx* ctx, const float* d_rgb)
{
    if (use_device(ctx)) return 1;
    if (!d_rgb) return set_error(ctx, "set_albedo_device: null pointer");
    ctx->d_albedo = d_rgb;
    return 0;
}

int prt_hip_render_albedo_image(prt_hip_ctx* ctx, int pass0, int spp, int max_depth, float* d_image,
                                prt_render_stats* out)
{
    if (use_device(ctx)) return 1;
    if (albedo_ready(ctx, spp, max_depth, "render_albedo_image")) return 1;
    if (!d_image) return set_error(ctx, "render_albedo_image: null image");
    const int64_t pixels = (int64_t)ctx->width * ctx->height;
    // The samples run in one launch per chunk of passes (work items (sample,
    // path), so the chunk pays one tail): each path leaves its RGB in its
    // sample's plane, and the planes are added to the accumulator in pass
    // order, which keeps the sums of one launch per pass.  Planes take at most
    // kRgbPlanesMaxBytes.
    constexpr int64_t kRgbPlanesMaxBytes = (int64_t)1 << 30;
    const int64_t plane_bytes = (int64_t)sizeof(float4) * pixels;
    const int chunk = (int)std::max<int64_t>(1, std::min<int64_t>(spp, kRgbPlanesMaxBytes / plane_bytes));
    if (ctx->rgb_planes_cap < (int64_t)chunk * pixels) {
        free_ptr(ctx->d_rgb_planes);
        ctx->d_rgb_planes = nullptr;
        ctx->rgb_planes_cap = 0;
        PRT_TRY(ctx, hipMalloc(&ctx->d_rgb_planes, (size_t)chunk * plane_bytes));
        ctx->rgb_planes_cap = (int64_t)chunk * pixels;
    }
    PRT_TRY(ctx, hipMemsetAsync(ctx->d_accum, 0, sizeof(float4) * pixels, ctx->stream));
    PRT_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    PRT_TRY(ctx, launch_frame_begin(ctx, (uint32_t)frame_entries(ctx)));
    for (int s0 = 0; s0 < spp; s0 += chunk) {
        const int n = std::min(chunk, spp - s0);
        PathLaunch p = frame_path_launch(ctx, pass0 + s0, max_depth);
        p.spp = n;
        p.rgb_planes = ctx->d_rgb_planes;
        PRT_TRY(ctx, launch_zero_u32(ctx->d_counters + 16, 1, ctx->stream));
        PRT_TRY(ctx, launch_paths_rgb(ctx, p, (ctx->profiling & PRT_PROFILE_COUNTERS) != 0));
        PRT_TRY(ctx, launch_rgb_planes_add(ctx, ctx->d_rgb_planes, n));
    }
    PRT_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    PRT_TRY(ctx, launch_accum_mean_rgb(ctx, d_image));
    return read_path_stats(ctx, out, pixels * spp);
}

int prt_hip_render_albedo_image_host(prt_hip_ctx* ctx, int pass0, int spp, int max_depth, float* image,
                                     prt_render_stats* out)
{
    if (use_device(ctx)) return 1;
    if (!image) return set_error(ctx, "render_albedo_image: null image");
    const int64_t n = (int64_t)ctx->width * ctx->height * 3;
    float* d = nullptr;
    PRT_TRY(ctx, hipMalloc(&d, sizeof(float) * n));
    int rc = prt_hip_render_albedo_image(ctx, pass0, spp, max_depth, d, out);
    if (rc == 0) {
        hipError_t e = hipMemcpy(image, d, sizeof(float) * n, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = check_hip(ctx, e, "render_albedo_image");
    }
    free_ptr(d);
    return rc;
}

int prt_hip_adjoint_albedo(prt_hip_ctx* ctx, int pass0, int spp, int max_depth, const float* d_dldi, float* d_grad,
                           prt_render_stats* out)
{
    if (use_device(ctx)) return 1;
    if (albedo_ready(ctx, spp, max_depth, "adjoint_albedo")) return 1;
    if (!d_dldi || !d_grad) return set_error(ctx, "adjoint_albedo: null array");
    if ((uint64_t)frame_entries(ctx) * (uint64_t)spp >= (uint64_t)UINT32_MAX - (1u << 20))
        return set_error(ctx, "adjoint_albedo: too many paths for one launch");
    PathLaunch p = frame_path_launch(ctx, pass0, max_depth);
    p.spp = spp;
    p.inv_spp = 1.0f / (float)spp;
    p.dldi = d_dldi;
    p.grad = d_grad;
    PRT_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    PRT_TRY(ctx, launch_frame_begin(ctx, (uint32_t)frame_entries(ctx)));
    PRT_TRY(ctx, launch_adjoint(ctx, p, (ctx->profiling & PRT_PROFILE_COUNTERS) != 0));
    PRT_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    return read_path_stats(ctx, out, (int64_t)ctx->width * ctx->height * spp);
}

int prt_hip_adjoint_albedo_host(prt_hip_ctx* ctx, int pass0, int spp, int max_depth, const float* dldi, float* grad,
                                prt_render_stats* out)
{
    if (use_device(ctx)) return 1;
    if (!dldi || !grad) return set_error(ctx, "adjoint_albedo: null array");
    const int64_t ni = (int64_t)ctx->width * ctx->height * 3, ng = ctx->scene.tri_count * 3;
    float *di = nullptr, *dg = nullptr;
    hipError_t e = hipMalloc(&di, sizeof(float) * ni);
    if (e == hipSuccess) e = hipMalloc(&dg, sizeof(float) * ng);
    if (e == hipSuccess) e = hipMemcpy(di, dldi, sizeof(float) * ni, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dg, grad, sizeof(float) * ng, hipMemcpyHostToDevice);
    int rc = e == hipSuccess ? prt_hip_adjoint_albedo(ctx, pass0, spp, max_depth, di, dg, out)
                             : check_hip(ctx, e, "adjoint_albedo");
    if (rc == 0) {
        e = hipMemcpy(grad, dg, sizeof(float) * ng, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = check_hip(ctx, e, "adjoint_albedo");
    }
    free_ptr(di);
    free_ptr(dg);
    return rc;
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
    if (!ctx->scene.bvh_ready) return set_error(ctx, "bvh_info: no BVH");
    if (node_count) *node_count = ctx->scene.node_count;
    if (tri_count) *tri_count = ctx->scene.tri_count;
    if (depth) *depth = ctx->scene.bvh_depth;
    return 0;
}

int prt_hip_bvh_read(prt_hip_ctx* ctx, void* nodes, float* tri_records, int32_t* prim_ids)
{
    if (use_device(ctx)) return 1;
    if (!ctx->scene.bvh_ready) return set_error(ctx, "bvh_read: no BVH");
    if (nodes)
        PRT_TRY(ctx, hipMemcpyAsync(nodes, ctx->scene.d_nodes, sizeof(Bvh8Node) * ctx->scene.node_count,
                                    hipMemcpyDeviceToHost, ctx->stream));
    if (tri_records)
        PRT_TRY(ctx, hipMemcpyAsync(tri_records, ctx->scene.d_tris, sizeof(float) * 9 * ctx->scene.tri_count,
                                    hipMemcpyDeviceToHost, ctx->stream));
    if (prim_ids)
        PRT_TRY(ctx, hipMemcpyAsync(prim_ids, ctx->scene.d_prim_ids, sizeof(int32_t) * ctx->scene.tri_count,
                                    hipMemcpyDeviceToHost, ctx->stream));
    PRT_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

}  // extern "C"
