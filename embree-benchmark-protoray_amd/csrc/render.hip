// render.hip -- camera rays, diffuse / primary shading with wave-ballot queue
// compaction, accumulation and tonemapping.
//
// Counterparts in the reference (written fresh, not translated):
//   generateRaysKernel  render/diffuse_renderer_cuda.cu:27-59
//   shadeRaysKernel     render/diffuse_renderer_cuda.cu:61-120,
//                       render/primary_renderer_cuda.cu:59-96
//   accumBufferUpdateKernel  image/accum_buffer_cuda.cu:22-44
// Semantics follow the CPU `diffuse` renderer, the oracle of this path
// (render/renderer_single.h:102-134, render/diffuse_integrator_single.h:41-69):
// sampler draws in the order pixel 2D, lens 2D, bounce k 2D; secondary rays
// start at p + eps*dir with tnear 0.
#include "prt_internal.h"

namespace prt {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ uint64_t lanes_below(int lane)
{
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Queue entry i = tile * 64 + lane; a wave covers one 8x8 tile in the Morton
// order of image/morton.cpp:22 (x from lane bits 0,2,4; y from bits 1,3,5).
// Lanes outside a partial tile emit a dead ray (tfar = -1, pixel -1).
__global__ __launch_bounds__(kBlock) void k_raygen(Camera cam, int width, int height, int tiles_x,
                                                   int tiles, int pass, float inv_w, float inv_h,
                                                   float4* rays, int32_t* pix, uint32_t* rng)
{
    const int gid = blockIdx.x * kBlock + threadIdx.x;
    const int tile = gid >> 6;
    if (tile >= tiles) return;
    const int lane = gid & 63;
    const int mx = (lane & 1) | ((lane >> 1) & 2) | ((lane >> 2) & 4);
    const int my = ((lane >> 1) & 1) | ((lane >> 2) & 2) | ((lane >> 3) & 4);
    const int x = (tile % tiles_x) * 8 + mx;
    const int y = (tile / tiles_x) * 8 + my;
    if (x >= width || y >= height) {
        rays[2 * (size_t)gid] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        rays[2 * (size_t)gid + 1] = make_float4(0.0f, 0.0f, 1.0f, -1.0f);
        pix[gid] = -1;
        rng[gid] = 0u;
        return;
    }
    const int pixel_id = x + y * width;
    uint32_t s = sampler_seed(pass, pixel_id);
    float u1 = sampler_next(s);
    float u2 = sampler_next(s);
    const float l1 = sampler_next(s);  // lens 2D, drawn for the pinhole too
    const float l2 = sampler_next(s);
    float ix = ((float)x + u1) * inv_w;
    float iy = ((float)y + u2) * inv_h;
    f3 org, d;
    camera_ray(cam, ix, iy, l1, l2, org, d);
    rays[2 * (size_t)gid] = make_float4(org.x, org.y, org.z, 0.0f);
    rays[2 * (size_t)gid + 1] = make_float4(d.x, d.y, d.z, 3.402823466e38f);
    pix[gid] = pixel_id;
    rng[gid] = s;
}

__global__ __launch_bounds__(kBlock) void k_frame_begin(uint32_t* counters, uint32_t n0,
                                                        unsigned long long* stats)
{
    if (threadIdx.x < 33) counters[threadIdx.x] = threadIdx.x == 0 ? n0 : 0u;
    if (threadIdx.x < 32) stats[threadIdx.x] = threadIdx.x == 9 ? ~0ull : 0ull;  // [9]: a minimum
}

__device__ __forceinline__ void tri_vertices(const int32_t* idx, const float* pos, int prim, f3& p0,
                                             f3& p1, f3& p2)
{
    const int a = idx[3 * prim], b = idx[3 * prim + 1], c = idx[3 * prim + 2];
    p0 = mk3(pos[3 * a], pos[3 * a + 1], pos[3 * a + 2]);
    p1 = mk3(pos[3 * b], pos[3 * b + 1], pos[3 * b + 2]);
    p2 = mk3(pos[3 * c], pos[3 * c + 1], pos[3 * c + 2]);
}

// Adds a sample to a pixel, or (the "Fast" renderers) replaces the pixel.
__device__ __forceinline__ void accum_add(float4* accum, int pixel, float c, bool overwrite)
{
    if (overwrite) {
        accum[pixel] = make_float4(c, c, c, 1.0f);
        return;
    }
    float4 a = accum[pixel];
    accum[pixel] = make_float4(a.x + c, a.y + c, a.z + c, a.w + 1.0f);
}

// Set-up of a lookahead call, in one launch before the path kernel.
//  - Block 0 clears the counters and statistics (as k_frame_begin), then
//    hands the state that the previous call left to this call (LookState):
//    level l of the previous call is level l - 1 here.  valid: the previous
//    call rendered the previous pass with the same inputs and lookahead on;
//    else every path it started for later frames is dropped.  The previous
//    call took each level's items in path order from its share of
//    spec_counter, so the items of a frame started so far are a prefix.
//  - With a valid hand-over, every block adds the samples of this frame that
//    earlier calls finished (side slot tag % kLookLevels, tagged `tag`) to
//    their pixels' accumulators; those pixels get no other sample in this
//    call.
__global__ __launch_bounds__(kBlock) void k_look_call_begin(uint32_t* counters, uint32_t n0,
                                                            unsigned long long* stats, LookState* ls,
                                                            uint32_t entries, int valid, uint32_t q_cap,
                                                            uint32_t waves, const uint2* side, float4* accum, int n,
                                                            uint32_t tag)
{
    if (blockIdx.x == 0) {
        if (threadIdx.x < 33) counters[threadIdx.x] = threadIdx.x == 0 ? n0 : 0u;
        if (threadIdx.x < 32) stats[threadIdx.x] = threadIdx.x == 9 ? ~0ull : 0ull;  // [9]: a minimum
        __syncthreads();
        if (threadIdx.x == 0) {
            LookState nx{};
            unsigned long long carried = 0;
            if (valid) {
                const uint32_t claimed = ls->spec_counter;
                for (int l = 1; l <= kLookLevels; ++l) {
                    const uint32_t begin = ls->spec_begin[l], end = ls->spec_begin[l + 1];
                    const uint32_t got = claimed > begin ? min(claimed, end) - begin : 0u;
                    nx.started[l - 1] = ls->started[l] + got;
                    nx.resume_in[l - 1] = min(ls->resume_out[l], q_cap);
                    if (l < kLookLevels) nx.rays[l] = ls->rays[l + 1];
                }
                carried = ls->rays[1];
            }
            // this frame's space: its items not started yet, then the
            // suspended paths level by level
            uint32_t items = entries - nx.started[0];
            for (int l = 0; l < kLookLevels; ++l) items += nx.resume_in[l];
            nx.target = items;
            for (int l = 1; l <= kLookLevels; ++l) nx.spec_begin[l + 1] = nx.spec_begin[l] + (entries - nx.started[l]);
            nx.live = waves;
            *ls = nx;
            stats[3] = carried;
        }
    }
    if (valid) {
        const int i = blockIdx.x * kBlock + threadIdx.x;
        if (i < n) {
            const uint2 e = side[(size_t)(tag % (uint32_t)kLookLevels) * (size_t)n + i];
            if (e.y == tag) accum_add(accum, i, __uint_as_float(e.x), false);
        }
    }
}

// One bounce of the diffuse integrator for every queued ray: misses add the
// path throughput, hits at the final depth add 0, other hits append the
// bounced ray to the next queue (one atomic per wave, slots by lane prefix).
__global__ __launch_bounds__(kBlock) void k_shade_diffuse(const float4* rays_in, const float4* hits,
                                                          const int32_t* pix_in, const uint32_t* rng_in,
                                                          const uint32_t* count_ptr, float4* rays_out,
                                                          int32_t* pix_out, uint32_t* rng_out,
                                                          uint32_t* out_count, float4* accum,
                                                          const int32_t* idx, const float* pos,
                                                          const float* normals, const float* texcoords,
                                                          float throughput, int final_depth, bool overwrite)
{
    const uint32_t count = *count_ptr;
    const int lane = threadIdx.x & 63;
    const uint64_t lt = lanes_below(lane);
    const uint32_t wave = (blockIdx.x * kBlock + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * kBlock) >> 6;
    for (uint32_t base = wave * 64; base < count; base += nwaves * 64) {
        const uint32_t i = base + lane;
        bool cont = false;
        float4 r0, r1;
        f3 no, nd;
        uint32_t s = 0;
        int32_t p = -1;
        if (i < count) {
            p = pix_in[i];
            if (p >= 0) {
                float4 h = hits[i];
                if (h.x >= 0.0f) {
                    if (final_depth) {
                        accum_add(accum, p, 0.0f, overwrite);
                    } else {
                        r0 = rays_in[2 * (size_t)i];
                        r1 = rays_in[2 * (size_t)i + 1];
                        f3 org = mk3(r0.x, r0.y, r0.z), dir = mk3(r1.x, r1.y, r1.z);
                        f3 p0, p1, p2;
                        const int prim = __float_as_int(h.y);
                        tri_vertices(idx, pos, prim, p0, p1, p2);
                        const Frame f = normals || texcoords
                                            ? attribute_frame(idx, normals, texcoords, prim, p0, p1, p2, dir, h.z, h.w)
                                            : make_frame(face_normal(p0, p1, p2, dir));
                        s = rng_in[i];
                        float s1 = sampler_next(s);
                        float s2 = sampler_next(s);
                        diffuse_bounce(org, dir, h.x, f, s1, s2, no, nd);
                        cont = true;
                    }
                } else {
                    accum_add(accum, p, throughput, overwrite);
                }
            }
        }
        const uint64_t m = __ballot(cont);
        if (m) {
            const int leader = __ffsll((unsigned long long)m) - 1;
            uint32_t ob = 0;
            if (lane == leader) ob = atomicAdd(out_count, (uint32_t)__popcll(m));
            ob = __shfl(ob, leader, 64);
            if (cont) {
                const uint32_t o = ob + (uint32_t)__popcll(m & lt);
                rays_out[2 * (size_t)o] = make_float4(no.x, no.y, no.z, 0.0f);
                rays_out[2 * (size_t)o + 1] = make_float4(nd.x, nd.y, nd.z, 3.402823466e38f);
                pix_out[o] = p;
                rng_out[o] = s;
            }
        }
    }
}

// Primary visibility: (N + 1) / 2 with N the shading normal (the ray-facing
// geometric normal of a mesh without normals), 0.05 on a miss
// (render/primary_integrator_single.h:34-47, render/primary_renderer_cuda.cu:59-90).
__global__ __launch_bounds__(kBlock) void k_shade_primary(const float4* rays_in, const float4* hits,
                                                          const int32_t* pix_in, const uint32_t* count_ptr,
                                                          float4* accum, const int32_t* idx,
                                                          const float* pos, const float* normals,
                                                          const float* texcoords, bool overwrite)
{
    const uint32_t count = *count_ptr;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < count; i += gridDim.x * kBlock) {
        const int32_t p = pix_in[i];
        if (p < 0) continue;
        const float4 h = hits[i];
        f3 c = mk3(0.05f, 0.05f, 0.05f);
        if (h.x >= 0.0f) {
            float4 r1 = rays_in[2 * (size_t)i + 1];
            const f3 dir = mk3(r1.x, r1.y, r1.z);
            f3 p0, p1, p2;
            const int prim = __float_as_int(h.y);
            tri_vertices(idx, pos, prim, p0, p1, p2);
            const f3 n = normals || texcoords
                             ? attribute_frame(idx, normals, texcoords, prim, p0, p1, p2, dir, h.z, h.w).n
                             : face_normal(p0, p1, p2, dir);
            c = mk3((n.x + 1.0f) * 0.5f, (n.y + 1.0f) * 0.5f, (n.z + 1.0f) * 0.5f);
        }
        if (overwrite) {
            accum[p] = make_float4(c.x, c.y, c.z, 1.0f);
        } else {
            float4 a = accum[p];
            accum[p] = make_float4(a.x + c.x, a.y + c.y, a.z + c.z, a.w + 1.0f);
        }
    }
}

__device__ __forceinline__ int encode_channel(float v)
{
    int c = (int)(powf(v, 1.0f / 2.2f) * 255.0f);
    return c < 0 ? 0 : (c > 255 ? 255 : c);
}

__global__ __launch_bounds__(kBlock) void k_tonemap(const float4* accum, int32_t* out, int n)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 a = accum[i];
    const float w = 1.0f / a.w;
    const int r = encode_channel(a.x * w), g = encode_channel(a.y * w), b = encode_channel(a.z * w);
    out[i] = b | (g << 8) | (r << 16);
}

__global__ __launch_bounds__(kBlock) void k_unpack_rays(const float4* rays, const int32_t* pix,
                                                        const uint32_t* rng, int64_t n, prt_ray* out,
                                                        uint32_t* states)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int32_t p = pix[i];
    if (p < 0) return;
    const float4 r0 = rays[2 * i], r1 = rays[2 * i + 1];
    prt_ray r;
    r.org[0] = r0.x; r.org[1] = r0.y; r.org[2] = r0.z; r.tnear = r0.w;
    r.dir[0] = r1.x; r.dir[1] = r1.y; r.dir[2] = r1.z; r.tfar = r1.w;
    out[p] = r;
    if (states) states[p] = rng[i];
}

__global__ __launch_bounds__(kBlock) void k_zero_u32(uint32_t* p, int n)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Wavefront frames deeper than 16: the queue-size and fetch-counter slot of
// depth k is k mod 16.  Before depth d (d >= 16) reuses its slot, the count
// of depth d - 16 moves to its statistics bin (the last bin holds depths >=
// 15; depth 0's count is the pixel count, known on the host) and both
// counters are cleared.
__global__ void k_queue_recycle(uint32_t* counters, unsigned long long* stats, int old_depth)
{
    if (threadIdx.x != 0) return;
    const int slot = old_depth & (kMaxQueueDepth - 1);
    if (old_depth > 0) stats[16 + min(old_depth, kMaxQueueDepth - 1)] += counters[slot];
    counters[slot] = 0u;
    counters[16 + slot] = 0u;
}

// ---- batched form
__global__ __launch_bounds__(64) void k_batch_begin(uint32_t* counters, unsigned long long* stats)
{
    if (threadIdx.x == 0) {
        counters[32] = 0u;
        counters[48] = 0u;
    }
    if (threadIdx.x < 16) stats[threadIdx.x] = 0ull;
}

// accum[job pixel] += the job's samples in pass order, so the result is the
// same float sequence as one prt_hip_render call per sample.
__global__ __launch_bounds__(kBlock) void k_batch_finalize(const BatchJobDevice* jobs,
                                                           const uint64_t* pixel_begin, int n_jobs,
                                                           const float* samples, float4* accum,
                                                           uint64_t total_pixels)
{
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= total_pixels) return;
    int lo = 0, hi = n_jobs;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pixel_begin[mid] <= i) lo = mid; else hi = mid;
    }
    const BatchJobDevice& job = jobs[lo];
    const uint64_t p = i - pixel_begin[lo];
    const uint64_t plane = (uint64_t)job.width * (uint64_t)job.height;
    float4 a = accum[i];
    for (int k = 0; k < job.spp; ++k) {
        const float v = samples[job.sample_base + (uint64_t)k * plane + p];
        a = make_float4(a.x + v, a.y + v, a.z + v, a.w + 1.0f);
    }
    accum[i] = a;
}

// next render of the batch continues with the following passes
__global__ __launch_bounds__(64) void k_batch_advance(BatchJobDevice* jobs, int n_jobs)
{
    for (int j = threadIdx.x; j < n_jobs; j += 64) jobs[j].pass0 += jobs[j].spp;
}

inline int persistent_blocks(prt_hip_ctx* ctx) { return ctx->num_cus * 8; }

}  // namespace

hipError_t launch_zero_u32(uint32_t* p, int n, hipStream_t s)
{
    k_zero_u32<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(p, n);
    return hipGetLastError();
}

Camera to_camera(const prt_camera& c)
{
    Camera k;
    k.origin = mk3(c.origin[0], c.origin[1], c.origin[2]);
    k.image_o = mk3(c.image_o[0], c.image_o[1], c.image_o[2]);
    k.image_dx = mk3(c.image_dx[0], c.image_dx[1], c.image_dx[2]);
    k.image_dy = mk3(c.image_dy[0], c.image_dy[1], c.image_dy[2]);
    k.basis_u = mk3(c.basis_u[0], c.basis_u[1], c.basis_u[2]);
    k.basis_v = mk3(c.basis_v[0], c.basis_v[1], c.basis_v[2]);
    k.lens_radius = c.lens_radius;
    k.focal_distance = c.focal_distance;
    return k;
}

hipError_t launch_raygen(prt_hip_ctx* ctx, int pass, float4* rays, int32_t* pix, uint32_t* rng)
{
    const int tiles_x = (ctx->width + 7) / 8, tiles_y = (ctx->height + 7) / 8;
    const int tiles = tiles_x * tiles_y;
    const uint32_t entries = (uint32_t)tiles * 64u;
    k_frame_begin<<<1, 64, 0, ctx->stream>>>(ctx->d_counters, entries, ctx->d_stats);
    const float inv_w = 1.0f / (float)ctx->width, inv_h = 1.0f / (float)ctx->height;
    k_raygen<<<(tiles * 64 + kBlock - 1) / kBlock, kBlock, 0, ctx->stream>>>(
        to_camera(ctx->cam), ctx->width, ctx->height, tiles_x, tiles, pass, inv_w, inv_h, rays, pix, rng);
    return hipGetLastError();
}

hipError_t launch_frame_begin(prt_hip_ctx* ctx, uint32_t entries)
{
    k_frame_begin<<<1, 64, 0, ctx->stream>>>(ctx->d_counters, entries, ctx->d_stats);
    return hipGetLastError();
}

hipError_t launch_look_call_begin(prt_hip_ctx* ctx, bool valid, uint32_t entries)
{
    const int n = ctx->width * ctx->height;
    const int grid = valid ? (n + kBlock - 1) / kBlock : 1;
    k_look_call_begin<<<grid, kBlock, 0, ctx->stream>>>(
        ctx->d_counters, entries, ctx->d_stats, ctx->d_look, entries, valid ? 1 : 0, ctx->look_q_cap,
        (uint32_t)ctx->diffuse_grid * (uint32_t)(kTraceBlock / 64), ctx->d_look_side, ctx->d_accum, n,
        ctx->look_gen);
    return hipGetLastError();
}

static __global__ __launch_bounds__(kBlock) void k_accum_mean_rgb(const float4* accum, float* image, int n)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 a = accum[i];
    const float inv = a.w > 0.0f ? 1.0f / a.w : 0.0f;
    image[3 * i + 0] = a.x * inv;
    image[3 * i + 1] = a.y * inv;
    image[3 * i + 2] = a.z * inv;
}

static __global__ __launch_bounds__(kBlock) void k_rgb_planes_add(const float4* planes, int spp, int n,
                                                                 float4* accum)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float4 a = accum[i];
    for (int s = 0; s < spp; ++s) {
        const float4 c = planes[(size_t)s * n + i];
        a = make_float4(a.x + c.x, a.y + c.y, a.z + c.z, a.w + 1.0f);
    }
    accum[i] = a;
}

hipError_t launch_rgb_planes_add(prt_hip_ctx* ctx, const float4* planes, int spp)
{
    const int n = ctx->width * ctx->height;
    k_rgb_planes_add<<<(n + kBlock - 1) / kBlock, kBlock, 0, ctx->stream>>>(planes, spp, n, ctx->d_accum);
    return hipGetLastError();
}

hipError_t launch_accum_mean_rgb(prt_hip_ctx* ctx, float* image)
{
    const int n = ctx->width * ctx->height;
    k_accum_mean_rgb<<<(n + kBlock - 1) / kBlock, kBlock, 0, ctx->stream>>>(ctx->d_accum, image, n);
    return hipGetLastError();
}

hipError_t launch_batch_begin(prt_hip_ctx* ctx)
{
    k_batch_begin<<<1, 64, 0, ctx->stream>>>(ctx->d_counters, ctx->d_stats);
    return hipGetLastError();
}

hipError_t launch_batch_finalize(prt_hip_ctx* ctx)
{
    const uint64_t n = ctx->batch_pixels;
    k_batch_finalize<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, ctx->stream>>>(
        ctx->d_batch_jobs, ctx->d_batch_pixel_begin, (int)ctx->batch_jobs.size(), ctx->d_batch_samples,
        ctx->d_batch_accum, n);
    return hipGetLastError();
}

hipError_t launch_batch_advance(prt_hip_ctx* ctx)
{
    k_batch_advance<<<1, 64, 0, ctx->stream>>>(ctx->d_batch_jobs, (int)ctx->batch_jobs.size());
    return hipGetLastError();
}

hipError_t launch_queue_recycle(prt_hip_ctx* ctx, int old_depth)
{
    k_queue_recycle<<<1, 64, 0, ctx->stream>>>(ctx->d_counters, ctx->d_stats, old_depth);
    return hipGetLastError();
}

hipError_t launch_shade_diffuse(prt_hip_ctx* ctx, int depth, int max_depth, float throughput, bool fast)
{
    const int cur = depth & 1;
    const int slot = depth & (kMaxQueueDepth - 1), next = (depth + 1) & (kMaxQueueDepth - 1);
    k_shade_diffuse<<<persistent_blocks(ctx), kBlock, 0, ctx->stream>>>(
        ctx->d_rays[cur], ctx->d_hits, ctx->d_pix[cur], ctx->d_rng[cur], ctx->d_counters + slot,
        ctx->d_rays[1 - cur], ctx->d_pix[1 - cur], ctx->d_rng[1 - cur], ctx->d_counters + next,
        ctx->d_accum, ctx->scene.d_idx, ctx->scene.d_pos, fast ? nullptr : ctx->scene.d_normals,
        fast ? nullptr : ctx->scene.d_texcoords, throughput, depth == max_depth ? 1 : 0, fast);
    return hipGetLastError();
}

hipError_t launch_shade_primary(prt_hip_ctx* ctx, bool fast)
{
    k_shade_primary<<<persistent_blocks(ctx), kBlock, 0, ctx->stream>>>(
        ctx->d_rays[0], ctx->d_hits, ctx->d_pix[0], ctx->d_counters, ctx->d_accum, ctx->scene.d_idx,
        ctx->scene.d_pos, fast ? nullptr : ctx->scene.d_normals, fast ? nullptr : ctx->scene.d_texcoords, fast);
    return hipGetLastError();
}

hipError_t launch_tonemap(prt_hip_ctx* ctx)
{
    const int n = ctx->width * ctx->height;
    k_tonemap<<<(n + kBlock - 1) / kBlock, kBlock, 0, ctx->stream>>>(ctx->d_accum, ctx->d_bgr8, n);
    return hipGetLastError();
}

hipError_t launch_unpack_rays(const float4* rays, const int32_t* pix, const uint32_t* rng, int64_t n,
                              int, int, prt_ray* out_rays, uint32_t* out_states, hipStream_t s)
{
    k_unpack_rays<<<(int)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(rays, pix, rng, n, out_rays,
                                                                      out_states);
    return hipGetLastError();
}

}  // namespace prt
