// trace.hip -- BVH8 closest-hit traversal, alone or fused with the diffuse
// path it serves.
//
// Replaces IntersectorStreamCuda::intersect (core/intersector_stream_cuda.h:27-34),
// which the reference forwards to OptiX Prime's closest query
// (accel/optix/optix_intersector_stream_cuda.cpp:77-82), and, in the fused
// modes, the raygen / shade kernels and the per-bounce host loop of
// DiffuseRendererCuda::render (render/diffuse_renderer_cuda.cu:27-120,159-212).
//
// One ray per lane.  The kernel is persistent: each wave takes work from a
// global counter; when at least kRefill of its 64 lanes are idle, the wave
// ballots them, takes that many work items with one atomic and hands them out
// by lane prefix, so lanes stay busy on incoherent bounces.  The traversal
// stack is a per-lane ring of kStackLds entries in LDS; deeper entries spill
// to a per-lane slice of global memory.  Each entry keeps the entry distance
// of its box, so nodes behind the closest hit are dropped when popped.
//
// Modes:
//   kModeTrace    work item = a ray of the input array; writes its hit record.
//   kModeDiffuse  work item = a path (one pixel sample).  The lane generates
//                 the camera ray, and when a ray finishes it shades it in
//                 place: a miss adds the path throughput, a hit at max depth
//                 adds 0, any other hit bounces (cosine hemisphere) and the
//                 lane traces the next ray of the same path.  Paths of every
//                 depth share the launch, so a frame pays the latency of its
//                 slowest path once instead of once per bounce, and no ray or
//                 hit record goes through HBM.
//   kModePrimary  as kModeDiffuse with the primary shading ((N+1)/2, 0.05).
//   kModeDiffuse2 as kModeDiffuse with next-event shadow rays: at each hit the
//                 lane traces an any-hit shadow ray, then the path's next ray.
//   kModeOccluded work item = a ray of the input array; any-hit query, writes
//                 the distance of a blocking triangle or -1.
//
// Result of a query: the hit with the smallest (t, primId) among triangles
// that the watertight test says the ray crosses with tnear <= t <= tfar, t
// being the plane distance computed in double -- the same answer as testing
// every triangle, whatever the visiting order.
#include "prt_internal.h"

namespace prt {
namespace {

constexpr int kStackLds = 16;      // default LDS entries per lane (power of two)
constexpr int kRefill = 32;        // default: refill when this many lanes are idle

enum {
    kModeTrace = 0, kModeDiffuse = 1, kModePrimary = 2, kModeBatch = 3, kModeRgb = 4, kModeAdjoint = 5,
    kModeDiffuse2 = 6, kModeOccluded = 7
};

__device__ __forceinline__ float byte_f(uint32_t w, int k)
{
    return (float)((w >> (8 * k)) & 0xffu);
}

__device__ __forceinline__ float as_f(uint32_t u) { return __uint_as_float(u); }

// kLook: the fused diffuse frame with the frame lookahead compiled in (and the
// cooperative tail, which it never uses, left out); the plain variant is the
// other way round, so each carries only the code it runs.
template <bool kCount, int kMode, int kBlock = kTraceBlock, int kStack = kStackLds, bool kLook = false>
__global__ __launch_bounds__(kBlock) void k_trace(SceneDevice sc, TraceLaunch tl, PathLaunch pl, BatchLaunch bl,
                                                  uint2* spill, int stack_entries, int refill, uint32_t* err)
{
    static_assert((kStack & (kStack - 1)) == 0, "LDS stack depth must be a power of two");
    constexpr bool kRayQuery = kMode == kModeTrace || kMode == kModeOccluded;
    __shared__ uint32_t st_node[kStack * kBlock];
    __shared__ float st_t[kStack * kBlock];
    // rays per depth: an LDS histogram in the instrumented variant only, so
    // the timed variant keeps 32 KiB of LDS per block (5 blocks per CU)
    uint32_t* depth_hist = nullptr;
    if constexpr (kCount && !kRayQuery && kMode != kModeBatch) {
        __shared__ uint32_t hist[kMaxQueueDepth];
        depth_hist = hist;
    }
    uint32_t n_rays = 0;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    // LDS ring entry q of this lane: byte offset ((q mod kStack) * kBlock + tid) * 4
    constexpr int kRingShift = kBlock == 256 ? 10 : (kBlock == 128 ? 9 : 8);
    static_assert(kBlock == 256 || kBlock == 128 || kBlock == 64, "ring layout");
    constexpr uint32_t kRingMask = (uint32_t)(kStack - 1) << kRingShift;
    const uint32_t tid4 = (uint32_t)tid * 4u;
    auto ring_off = [&](int q) -> uint32_t { return (((uint32_t)q << kRingShift) & kRingMask) | tid4; };
    auto ring_node = [&](uint32_t off) -> uint32_t& { return *reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(st_node) + off); };
    auto ring_t = [&](uint32_t off) -> float& { return *reinterpret_cast<float*>(reinterpret_cast<char*>(st_t) + off); };
    const uint64_t lane_lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint32_t count = kRayQuery ? (tl.count_ptr ? *tl.count_ptr : tl.count)
                         : (kMode == kModeBatch ? bl.total_work
                                                : (kMode == kModeAdjoint || (kMode == kModeRgb && pl.rgb_planes)
                                                       ? pl.entries * (uint32_t)pl.spp
                                                       : pl.entries));
    uint2* my_spill = spill + ((size_t)blockIdx.x * kBlock + tid) * (size_t)stack_entries;
    const int spill_cap = stack_entries;
    if constexpr (kCount && !kRayQuery && kMode != kModeBatch) {
        if (tid < kMaxQueueDepth) depth_hist[tid] = 0;
        __syncthreads();
    }

    bool active = false;   // the lane holds a ray or path
    bool tracing = false;  // ... and is traversing (else its ray waits for shading)
    uint32_t ray_idx = 0;
    f3 o{0, 0, 0}, d{0, 0, 0};
    float tmin = 0.0f, tbest = 0.0f;
    float inv_x = 0, inv_y = 0, inv_z = 0;
    bool neg_x = false, neg_y = false, neg_z = false;
    WatertightRay wr{};
    int32_t best_prim = -1;
    uint32_t best_rec = 0;
    float bu = 0.0f, bv = 0.0f;
    int sp = 0, bottom = 0;
    int32_t node = -1;
    float node_t = 0.0f;        // entry distance of `node`
    uint32_t pend = 0, pend_base = 0;  // triangle records still to test: pend_base + bit index
    uint32_t pend2 = 0, pend2_base = 0;  // a second leaf group, found while the first is tested
    // path state (fused modes)
    int32_t pixel = 0;
    uint32_t rng = 0;
    int depth = 0;
    float lw = 1.0f;
    int max_depth = pl.max_depth;
    float albedo = pl.albedo;
    // scene of the lane's current ray (per lane in the batch mode)
    const PRT_GLOBAL Bvh8Node* nodes = global_view(sc.nodes);
    const PRT_GLOBAL float* tris = global_view(sc.tris);
    const PRT_GLOBAL int32_t* prim_ids = global_view(sc.prim_ids);
    int32_t scene_id = 0;  // batch mode: the lane's scene (for its shading attributes)
    float* sample_out = nullptr;
    // RGB throughput (RGB and adjoint modes); adjoint: per-channel product of
    // the path's nonzero albedos and count of zero albedos from the primal
    // pass, and the path's weight dLoss/dI / spp
    float lg = 1.0f, lb = 1.0f;
    float pr = 1.0f, pg = 1.0f, pb = 1.0f;
    uint32_t zeros = 0;  // 8 bits per channel
    float gw_r = 0.0f, gw_g = 0.0f, gw_b = 0.0f;
    bool replay = false;
    int path_pass = 0;
    int path_x = 0, path_y = 0;
    // diffuse2: the current ray is a shadow ray; the path's next ray waits in
    // (eo, ed); radiance gathered so far
    bool shadow = false;
    f3 eo{0, 0, 0}, ed{0, 0, 0};
    float lsum = 0.0f;
    unsigned long long n_nodes = 0, n_tris = 0, n_spill = 0;
    unsigned long long n_wsteps = 0, n_lsteps = 0, n_wshade = 0;  // instrumented SIMT occupancy
    unsigned long long n_dwsteps = 0, n_dlsteps = 0;              // ... after the work ran out
    unsigned long long t_exhausted = 0;                           // wall clock when it ran out
    bool exhausted = false;  // wave-uniform: the work counter has run past count
    // Tail of a diffuse frame (kModeDiffuse): once out of work with at most
    // 8 paths left, the wave traces them cooperatively, 8 lanes per path
    // (see the cooperative phase after the main loop)
    constexpr bool kCoop = kMode == kModeDiffuse && !kLook;
    const bool coop_on = kCoop && pl.coop != 0;
    bool cont = false;  // the path's next ray waits in (o, d) for the cooperative phase

    // Frame lookahead (kModeDiffuse, LookState); a path of frame pass + l is
    // at level l.  This frame's space, from LookState::counter: [0, F) are
    // its items S_0..entries-1 (F = entries - S_0), then the paths the
    // previous call suspended, level by level: queue l holds lk_res[l + 1] -
    // lk_res[l] of them from lk_res[l] on (they are part way through, so
    // they go last).
    //  - A wave that has seen all of them handed out takes items of the
    //    later frames from LookState::spec_counter (level l: path S_l + j
    //    for the j-th item of its share), but only while it holds no path of
    //    this frame: waves that do take nothing more, stay sparse and issue
    //    first (s_setprio).
    //  - Every wave starts registered in LookState::live and leaves it once,
    //    at that same point (all handed out, no level-0 path held); the last
    //    to leave writes this call's tag into each block's stop flag.
    //  - After the stop no more items are taken, and paths of later frames
    //    leave for the queue of their level: at a bounce, or in the middle
    //    of a ray, which the next call then traces again from its origin.
    //  - The cooperative tail is off under the lookahead.
    constexpr bool kLookMode = kMode == kModeDiffuse && kLook;
    LookState* const look = kLookMode ? pl.look : nullptr;
    uint32_t lk_F = 0, lk_fresh_end = 0, lk_target = 0;
    uint32_t lk_S[kLookLevels + 1] = {};         // S_l
    uint32_t lk_res[kLookLevels + 1] = {};       // this frame's space: queue l from lk_res[l]
    uint32_t lk_spec[kLookLevels + 2] = {};      // later frames' space: level l from lk_spec[l]
    if (kLookMode && look) {
#pragma unroll
        for (int l = 0; l <= kLookLevels; ++l) lk_S[l] = look->started[l];
        lk_F = count - lk_S[0];
        lk_target = look->target;
        lk_res[0] = lk_F;
#pragma unroll
        for (int l = 0; l < kLookLevels; ++l) lk_res[l + 1] = lk_res[l] + look->resume_in[l];
        lk_fresh_end = lk_res[kLookLevels];
#pragma unroll
        for (int l = 1; l <= kLookLevels + 1; ++l) lk_spec[l] = look->spec_begin[l];
    }
    bool stop = kLookMode && look && lk_target == 0;  // wave-uniform: this frame is finished
    bool fresh_out = false;  // wave-uniform: every item of this frame's space is handed out
    bool live = kLookMode && look;  // wave-uniform: registered in LookState::live
    int level = 0;           // the lane's path belongs to frame pass + level
    bool suspend = false;    // the lane's later-frame path waits in (o, d) for the queue
    uint32_t poll_steps = 0;
    uint32_t* const stop_flag = kLookMode && look ? pl.look_flags + (size_t)blockIdx.x * kLookFlagStride : nullptr;
    // leaves LookState::live; the last wave to leave raises every stop flag
    auto look_leave = [&]() {
        uint32_t old = 0;
        if (lane == 0) old = atomicSub(&look->live, 1u);
        old = __shfl(old, 0, 64);
        if (old == 1u)
            for (uint32_t b = (uint32_t)lane; b < gridDim.x; b += 64u)
                __hip_atomic_store(pl.look_flags + (size_t)b * kLookFlagStride, pl.look_gen, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        live = false;
    };
    unsigned long long n_ahead[kLookLevels] = {};  // rays of frame pass + l + 1
    if (stop) exhausted = true;

    auto start_ray = [&](f3 org, f3 dir, float near, float far) {
        o = org;
        d = dir;
        tmin = near;
        tbest = far;
        const float dx = fabsf(d.x) < 1e-20f ? copysignf(1e-20f, d.x) : d.x;
        const float dy = fabsf(d.y) < 1e-20f ? copysignf(1e-20f, d.y) : d.y;
        const float dz = fabsf(d.z) < 1e-20f ? copysignf(1e-20f, d.z) : d.z;
        inv_x = __builtin_amdgcn_rcpf(dx);
        inv_y = __builtin_amdgcn_rcpf(dy);
        inv_z = __builtin_amdgcn_rcpf(dz);
        neg_x = dx < 0.0f;
        neg_y = dy < 0.0f;
        neg_z = dz < 0.0f;
        wr = watertight_setup(d);
        best_prim = -1;
        bu = bv = 0.0f;
        sp = bottom = 0;
        node = 0;
        node_t = near;
        pend = 0;
        pend2 = 0;
        tracing = true;
    };

    // Shading frame of the current ray's closest hit (triangle p0, p1, p2):
    // the full ShadingContext when the scene has per-vertex attributes (the
    // hit coordinates are recomputed from the winner's edge functions), else
    // the frame of the ray-facing geometric normal.
    auto hit_frame = [&](f3 p0, f3 p1, f3 p2) -> Frame {
        const int32_t* aidx = sc.idx;
        const float* an = pl.fast ? nullptr : sc.normals;
        const float* at = pl.fast ? nullptr : sc.texcoords;
        if (kMode == kModeBatch) {
            const SceneDevice& js = bl.scenes[scene_id];
            aidx = js.idx;
            an = js.normals;
            at = js.texcoords;
        }
        if (an != nullptr || at != nullptr) {
            float hu, hv;
            tri_barycentrics(wr, o, p0, p1, p2, hu, hv);
            return attribute_frame(aidx, an, at, best_prim, p0, p1, p2, d, hu, hv);
        }
        return make_frame(face_normal(p0, p1, p2, d));
    };

    for (;;) {
        // lookahead: a wave that holds paths of this frame takes no items of
        // later frames once this frame's are handed out, and issues ahead of
        // the waves that run later frames' paths, so the frame's last paths
        // keep the pace they would have alone
        bool hold = false;
        if (kLookMode && look) {
            hold = fresh_out && __ballot(active && level == 0) != 0;
            if (live && fresh_out && !hold) look_leave();
            if (!live && !stop) {
                const uint32_t f = __hip_atomic_load(stop_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                stop = (uint32_t)__builtin_amdgcn_readfirstlane(f) == pl.look_gen;
                if (stop) exhausted = true;
            }
            // after the stop only later frames' paths are left: those in the
            // middle of a ray leave now and trace that ray again next call
            // (its origin, direction and path state are what a bounce leaves)
            if (stop && tracing) {
                suspend = true;
                tracing = false;
                active = false;
            }
            if (hold) __builtin_amdgcn_s_setprio(3);
            else __builtin_amdgcn_s_setprio(0);
        }
        // the wave goes cooperative once out of work, with at most 8 paths
        // left and all of them between rays
        bool go_coop = false;
        if (coop_on && exhausted) {
            const uint64_t a = __ballot(active);
            go_coop = a != 0 && __popcll(a) <= 8 && __ballot(active && tracing) == 0;
        }
        // ---- shade the lanes whose ray has finished.  The traversal loop
        // hands them over in groups, so this code runs for many lanes at once
        // instead of stalling every traversal step of the wave for one lane.
        if (kCount && __ballot(active && !tracing) != 0 && lane == 0) ++n_wshade;
        if (active && !tracing) do {
            if (kMode == kModeOccluded) {
                tl.any_t[ray_idx] = best_prim >= 0 ? tbest : -1.0f;
                active = false;
                break;
            }
            if (kMode == kModeDiffuse2 && shadow) {
                // shadow ray done: light from the sky unless blocked (MIS weight 1/2)
                if constexpr (kCount) atomicAdd(&depth_hist[min(depth, kMaxQueueDepth - 1)], 1u);
                ++n_rays;
                if (best_prim < 0) lsum += lw * 0.5f;
                shadow = false;
                start_ray(eo, ed, 0.0f, 3.402823466e38f);
                break;
            }
            float t = -1.0f;
            f3 p0{}, p1{}, p2{};
            if (best_prim >= 0) {
                const PRT_GLOBAL float* tp = tris + 9 * (size_t)best_rec;
                p0 = mk3(tp[0], tp[1], tp[2]);
                p1 = mk3(tp[3], tp[4], tp[5]);
                p2 = mk3(tp[6], tp[7], tp[8]);
                t = tbest;  // plane distance of the winner (tri_test)
            }
            if (kMode == kModeTrace) {
                tl.hits[ray_idx] = best_prim >= 0
                                       ? make_float4(t, __int_as_float(best_prim), bu, bv)
                                       : make_float4(-1.0f, __int_as_float(-1), 0.0f, 0.0f);
                active = false;
                break;
            }
            if constexpr (kCount && kMode != kModeBatch) atomicAdd(&depth_hist[min(depth, kMaxQueueDepth - 1)], 1u);
            if (kLookMode && level > 0) {
#pragma unroll
                for (int l = 1; l <= kLookLevels; ++l)
                    if (level == l) ++n_ahead[l - 1];
            } else {
                ++n_rays;
            }
            float4* acc = kMode == kModeBatch ? nullptr : pl.accum + pixel;
            if (kMode == kModeDiffuse2) {
                // extension ray done (render/diffuse2_integrator_single.h:52-85)
                bool done = false;
                if (best_prim < 0) {
                    lsum += lw * (depth > 0 ? 0.5f : 1.0f);
                    done = true;
                } else if (depth == max_depth) {
                    done = true;
                }
                if (done) {
                    const float4 a = pl.fast ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : *acc;
                    *acc = make_float4(a.x + lsum, a.y + lsum, a.z + lsum, a.w + 1.0f);
                    active = false;
                    break;
                }
                lw *= albedo;
                const Frame fr = hit_frame(p0, p1, p2);
                const float s1 = sampler_next(rng);
                const float s2 = sampler_next(rng);
                const float s3 = sampler_next(rng);
                const float s4 = sampler_next(rng);
                f3 so, sd;
                diffuse_bounce(o, d, t, fr, s1, s2, so, sd);
                diffuse_bounce(o, d, t, fr, s3, s4, eo, ed);
                ++depth;
                shadow = true;
                start_ray(so, sd, 0.0f, 3.402823466e38f);
                break;
            }
            if (kMode == kModePrimary) {
                f3 c = mk3(0.05f, 0.05f, 0.05f);
                if (best_prim >= 0) {
                    const f3 n = hit_frame(p0, p1, p2).n;
                    c = mk3((n.x + 1.0f) * 0.5f, (n.y + 1.0f) * 0.5f, (n.z + 1.0f) * 0.5f);
                }
                const float4 a = pl.fast ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : *acc;
                *acc = make_float4(a.x + c.x, a.y + c.y, a.z + c.z, a.w + 1.0f);
                active = false;
                break;
            }
            if (best_prim < 0 || depth == max_depth) {
                if (kMode == kModeAdjoint) {
                    if (!replay && best_prim < 0) {
                        // the path escaped: replay it from its seed to
                        // hand each hit its share of the gradient
                        replay = true;
                        rng = sampler_seed(path_pass, pixel);
                        const float u1 = sampler_next(rng);
                        const float u2 = sampler_next(rng);
                        const float l1 = sampler_next(rng);
                        const float l2 = sampler_next(rng);
                        f3 org, dir;
                        camera_ray(pl.cam, ((float)path_x + u1) * pl.inv_w, ((float)path_y + u2) * pl.inv_h,
                                   l1, l2, org, dir);
                        depth = 0;
                        start_ray(org, dir, 0.0f, 3.402823466e38f);
                        continue;
                    }
                    active = false;
                    break;
                }
                const float c = best_prim < 0 ? lw : 0.0f;
                if (kMode == kModeBatch) {
                    *sample_out = c;
                } else if (kMode == kModeRgb) {
                    const float cg = best_prim < 0 ? lg : 0.0f, cb = best_prim < 0 ? lb : 0.0f;
                    if (pl.rgb_planes) {
                        // one plane per sample; they are added in pass order afterwards
                        pl.rgb_planes[(size_t)(path_pass - pl.pass) * (size_t)(pl.width * pl.height) + pixel] =
                            make_float4(c, cg, cb, 0.0f);
                    } else {
                        const float4 a = *acc;
                        *acc = make_float4(a.x + c, a.y + cg, a.z + cb, a.w + 1.0f);
                    }
                } else if (kLookMode && level > 0) {
                    // a later frame's sample, for the call whose tag it carries
                    const uint32_t tag = pl.look_gen + (uint32_t)level;
                    pl.side[(size_t)(tag % (uint32_t)kLookLevels) * (size_t)(pl.width * pl.height) + pixel] =
                        make_uint2(__float_as_uint(c), tag);
                } else {
                    const float4 a = pl.fast ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : *acc;
                    *acc = make_float4(a.x + c, a.y + c, a.z + c, a.w + 1.0f);
                }
                active = false;
                break;
            }
            // diffuse bounce: the next ray of this path
            if (kMode == kModeRgb || kMode == kModeAdjoint) {
                const float* a = pl.albedo_rgb + 3 * (size_t)best_prim;
                const float ar = a[0], ag = a[1], ab = a[2];
                if (kMode == kModeAdjoint) {
                    if (!replay) {
                        // primal: products of the nonzero albedos, zero counts
                        if (ar != 0.0f) pr *= ar; else zeros += 1u;
                        if (ag != 0.0f) pg *= ag; else zeros += 1u << 8;
                        if (ab != 0.0f) pb *= ab; else zeros += 1u << 16;
                    } else {
                        // dL_c/da_c(t) = product of the path's other albedos
                        float* g = pl.grad + 3 * (size_t)best_prim;
                        const uint32_t zr = zeros & 0xffu, zg = (zeros >> 8) & 0xffu, zb = zeros >> 16;
                        const float dr = zr == 0 ? pr / ar : (zr == 1 && ar == 0.0f ? pr : 0.0f);
                        const float dg = zg == 0 ? pg / ag : (zg == 1 && ag == 0.0f ? pg : 0.0f);
                        const float db = zb == 0 ? pb / ab : (zb == 1 && ab == 0.0f ? pb : 0.0f);
                        if (gw_r * dr != 0.0f) atomicAdd(g + 0, gw_r * dr);
                        if (gw_g * dg != 0.0f) atomicAdd(g + 1, gw_g * dg);
                        if (gw_b * db != 0.0f) atomicAdd(g + 2, gw_b * db);
                    }
                } else {
                    lw *= ar;
                    lg *= ag;
                    lb *= ab;
                }
            } else {
                lw *= albedo;
            }
            const Frame fr = hit_frame(p0, p1, p2);
            const float s1 = sampler_next(rng);
            const float s2 = sampler_next(rng);
            f3 no, nd;
            diffuse_bounce(o, d, t, fr, s1, s2, no, nd);
            ++depth;
            if (kLookMode && level > 0 && stop) {
                o = no;
                d = nd;
                suspend = true;
                active = false;
                break;
            }
            if (kCoop && go_coop) {
                o = no;
                d = nd;
                cont = true;
                break;
            }
            start_ray(no, nd, 0.0f, 3.402823466e38f);
        } while (0);
        if (kCoop && go_coop) break;

        // ---- lookahead: queue the later frames' paths that stopped, by level
        if (kLookMode && look && __ballot(suspend) != 0) {
#pragma unroll
            for (int l = 1; l <= kLookLevels; ++l) {
                const bool mine = suspend && level == l;
                const uint64_t sm = __ballot(mine);
                if (sm == 0) continue;
                const int leader = __ffsll((unsigned long long)sm) - 1;
                uint32_t qb = 0;
                if (lane == leader) qb = atomicAdd(&look->resume_out[l], (uint32_t)__popcll(sm));
                qb = __shfl(qb, leader, 64);
                if (mine) {
                    const uint32_t slot = qb + (uint32_t)__popcll(sm & lane_lt);
                    if (slot < pl.q_cap) {
                        uint4* q = pl.q_out[l - 1] + 3 * (size_t)slot;
                        q[0] = make_uint4(__float_as_uint(o.x), __float_as_uint(o.y), __float_as_uint(o.z),
                                          (uint32_t)pixel);
                        q[1] = make_uint4(__float_as_uint(d.x), __float_as_uint(d.y), __float_as_uint(d.z), rng);
                        q[2] = make_uint4((uint32_t)depth, __float_as_uint(lw), 0u, 0u);
                    } else {
                        atomicOr(err, 2u);
                    }
                }
            }
            suspend = false;
        }

        // ---- refill idle lanes (wave-uniform branch)
        const uint64_t idle = __ballot(!active);
        const int n_idle = __popcll(idle);
        if (!exhausted && !hold && n_idle >= refill) {
            const int leader = __ffsll((unsigned long long)idle) - 1;
            uint32_t* ctr = kRayQuery ? tl.fetch_ptr : (kMode == kModeBatch ? bl.work_counter : pl.path_counter);
            uint32_t lim = count;
            const bool spec_claim = kLookMode && look && fresh_out;  // items of later frames
            if (kLookMode && look) lim = spec_claim ? lk_spec[kLookLevels + 1] : lk_fresh_end;
            if (spec_claim) ctr = &look->spec_counter;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(ctr, (uint32_t)n_idle);
            base = __shfl(base, leader, 64);
            if (!active) {
                const uint32_t my = base + (uint32_t)__popcll(idle & lane_lt);
                if (my < lim) {
                    if (kRayQuery) {
                        ray_idx = my;
                        const float4 r0 = tl.rays[2 * (size_t)my];
                        const float4 r1 = tl.rays[2 * (size_t)my + 1];
                        if (r1.w < 0.0f) {
                            if (kMode == kModeOccluded) tl.any_t[my] = -1.0f;
                            else tl.hits[my] = make_float4(-1.0f, __int_as_float(-1), 0.0f, 0.0f);
                        } else {
                            start_ray(mk3(r0.x, r0.y, r0.z), mk3(r1.x, r1.y, r1.z), r0.w, r1.w);
                            active = true;
                        }
                    } else if (kMode == kModeBatch) {
                        // work item -> job (binary search over the job prefix),
                        // then sample and path inside the job
                        int lo = 0, hi = bl.n_jobs;
                        while (hi - lo > 1) {
                            const int mid = (lo + hi) >> 1;
                            if (bl.work_begin[mid] <= my) lo = mid; else hi = mid;
                        }
                        const BatchJobDevice& job = bl.jobs[lo];
                        const uint32_t local = my - bl.work_begin[lo];
                        const uint32_t sample = local / job.entries;
                        const uint32_t path = local - sample * job.entries;
                        const uint32_t tile = path >> 6, tl_lane = path & 63u;
                        const int mx = (tl_lane & 1) | ((tl_lane >> 1) & 2) | ((tl_lane >> 2) & 4);
                        const int my_ = ((tl_lane >> 1) & 1) | ((tl_lane >> 2) & 2) | ((tl_lane >> 3) & 4);
                        const int x = (int)(tile % (uint32_t)job.tiles_x) * 8 + mx;
                        const int y = (int)(tile / (uint32_t)job.tiles_x) * 8 + my_;
                        if (x < job.width && y < job.height) {
                            pixel = x + y * job.width;
                            const SceneDevice jsc = bl.scenes[job.scene];
                            scene_id = job.scene;
                            nodes = global_view(jsc.nodes);
                            tris = global_view(jsc.tris);
                            prim_ids = global_view(jsc.prim_ids);
                            sample_out = bl.samples + job.sample_base +
                                         (size_t)sample * (size_t)(job.width * job.height) + pixel;
                            max_depth = job.max_depth;
                            albedo = job.albedo;
                            rng = sampler_seed(job.pass0 + (int)sample, pixel);
                            const float u1 = sampler_next(rng);
                            const float u2 = sampler_next(rng);
                            const float l1 = sampler_next(rng);  // lens 2D
                            const float l2 = sampler_next(rng);
                            f3 org, dir;
                            camera_ray(job.cam, ((float)x + u1) * job.inv_w, ((float)y + u2) * job.inv_h, l1, l2,
                                       org, dir);
                            start_ray(org, dir, 0.0f, 3.402823466e38f);
                            depth = 0;
                            lw = 1.0f;
                            active = true;
                        }
                    } else if (kLookMode && look && !spec_claim && my >= lk_F) {
                        // a path that the previous call stopped at a bounce
                        // (or in the middle of a ray): its next ray
                        const uint4* q = pl.q_in[0] + 3 * (size_t)(my - lk_res[0]);
                        level = 0;
#pragma unroll
                        for (int l = 1; l < kLookLevels; ++l)
                            if (my >= lk_res[l]) {
                                q = pl.q_in[l] + 3 * (size_t)(my - lk_res[l]);
                                level = l;
                            }
                        const uint4 e0 = q[0], e1 = q[1], e2 = q[2];
                        pixel = (int32_t)e0.w;
                        rng = e1.w;
                        depth = (int)e2.x;
                        lw = __uint_as_float(e2.y);
                        start_ray(mk3(as_f(e0.x), as_f(e0.y), as_f(e0.z)), mk3(as_f(e1.x), as_f(e1.y), as_f(e1.z)),
                                  0.0f, 3.402823466e38f);
                        active = true;
                    } else {
                        // camera ray of path `my`: 8x8 tiles, Morton order inside
                        // a tile (image/morton.cpp:22), as the wavefront raygen
                        uint32_t path = my;
                        path_pass = pl.pass;
                        if (kLookMode && look) {
                            level = 0;
                            path = lk_S[0] + my;
                            if (spec_claim) {
#pragma unroll
                                for (int l = 1; l <= kLookLevels; ++l)
                                    if (my >= lk_spec[l]) {
                                        level = l;
                                        path = lk_S[l] + (my - lk_spec[l]);
                                    }
                            }
                            path_pass = pl.pass + level;
                        }
                        if (kMode == kModeAdjoint || (kMode == kModeRgb && pl.rgb_planes)) {
                            const uint32_t sample = my / pl.entries;
                            path = my - sample * pl.entries;
                            path_pass = pl.pass + (int)sample;
                        }
                        const uint32_t tile = path >> 6, tl_lane = path & 63u;
                        const int mx = (tl_lane & 1) | ((tl_lane >> 1) & 2) | ((tl_lane >> 2) & 4);
                        const int my_ = ((tl_lane >> 1) & 1) | ((tl_lane >> 2) & 2) | ((tl_lane >> 3) & 4);
                        const int x = (int)(tile % (uint32_t)pl.tiles_x) * 8 + mx;
                        const int y = (int)(tile / (uint32_t)pl.tiles_x) * 8 + my_;
                        if (x < pl.width && y < pl.height) {
                            pixel = x + y * pl.width;
                            path_x = x;
                            path_y = y;
                            rng = sampler_seed(path_pass, pixel);
                            const float u1 = sampler_next(rng);
                            const float u2 = sampler_next(rng);
                            const float l1 = sampler_next(rng);  // lens 2D
                            const float l2 = sampler_next(rng);
                            f3 org, dir;
                            camera_ray(pl.cam, ((float)x + u1) * pl.inv_w, ((float)y + u2) * pl.inv_h, l1, l2, org,
                                       dir);
                            start_ray(org, dir, 0.0f, 3.402823466e38f);
                            depth = 0;
                            lw = lg = lb = 1.0f;
                            lsum = 0.0f;
                            shadow = false;
                            active = true;
                            if (kMode == kModeAdjoint) {
                                replay = false;
                                pr = pg = pb = 1.0f;
                                zeros = 0;
                                const float* g = pl.dldi + 3 * (size_t)pixel;
                                gw_r = g[0] * pl.inv_spp;
                                gw_g = g[1] * pl.inv_spp;
                                gw_b = g[2] * pl.inv_spp;
                                if (gw_r == 0.0f && gw_g == 0.0f && gw_b == 0.0f) active = tracing = false;
                            }
                        }
                    }
                }
            }
            if ((uint64_t)base + (uint64_t)n_idle >= lim) {
                if (kLookMode && look && !spec_claim) {
                    fresh_out = true;
                } else {
                    exhausted = true;
                    if (kCount) t_exhausted = wall_clock64();
                }
            }
        }

        const uint64_t busy = __ballot(active);
        if (busy == 0) {
            if (exhausted) break;
            continue;
        }

        // ---- traverse until enough lanes wait for shading or for new work:
        // refill lanes once no more than 64 - refill trace; after the work
        // counter ran out, shade once half of the remaining lanes wait
        const int n_act = __popcll(busy);
        const int keep = !(exhausted || hold) ? 64 - refill
                                    : (coop_on && n_act <= 8 ? 0 : n_act - max(1, min(refill, n_act >> 1)));
        while (tracing) {
            if (kCount) {
                const bool lead = lane == __ffsll((unsigned long long)__ballot(1)) - 1;
                ++n_lsteps;
                if (lead) ++n_wsteps;
                if (exhausted) {
                    ++n_dlsteps;
                    if (lead) ++n_dwsteps;
                }
            }
            // A step tests one pending triangle and visits one node: a lane
            // keeps descending while its leaf triangles are tested, with room
            // for one more leaf group.  The test comes first, so the node's
            // culling sees its result.
            if (pend != 0) {
                // ---- one triangle of the pending group per step, so lanes
                // with many candidate triangles do not hold up the wave
                const uint32_t rec = pend_base + (uint32_t)__builtin_ctz(pend);
                pend &= pend - 1;
                const PRT_GLOBAL float* tp = tris + 9 * (size_t)rec;
                const f3 p0 = mk3(tp[0], tp[1], tp[2]);
                const f3 p1 = mk3(tp[3], tp[4], tp[5]);
                const f3 p2 = mk3(tp[6], tp[7], tp[8]);
                if (kCount) ++n_tris;
                TriCandidate c;
                if (tri_test(wr, o, d, p0, p1, p2, tmin, tbest, c)) {
                    if (kMode == kModeOccluded || (kMode == kModeDiffuse2 && shadow)) {
                        // any hit ends the query: drop the stack and the pending work
                        tbest = c.t;
                        best_prim = 0;
                        sp = bottom = 0;
                        pend = pend2 = 0;
                        node = -1;
                    } else {
                        const int32_t prim = prim_ids[rec];
                        if (c.t < tbest || best_prim < 0 || prim < best_prim) {
                            tbest = c.t;
                            best_prim = prim;
                            best_rec = rec;
                            bu = c.v_num / c.det;
                            bv = c.w_num / c.det;
                        }
                    }
                }
                if (pend == 0 && pend2 != 0) {
                    pend = pend2;
                    pend_base = pend2_base;
                    pend2 = 0;
                }
            }

            if (pend2 == 0) {
                // ---- node step: the next node that can still hold a closer hit
                if (node >= 0 && node_t > tbest) node = -1;
                if (node < 0) {
                    while (sp > 0) {
                        --sp;
                        uint32_t nid;
                        float nt;
                        if (sp < bottom) {
                            const uint2 e = my_spill[sp];
                            bottom = sp;
                            nid = e.x;
                            nt = as_f(e.y);
                        } else {
                            const uint32_t off = ring_off(sp);
                            nid = ring_node(off);
                            nt = ring_t(off);
                        }
                        if (nt <= tbest) { node = (int32_t)nid; break; }
                    }
                }
            }
            if (pend2 == 0 && node < 0) {
                if (pend == 0) {
                    tracing = false;  // the ray is finished: shade it in the outer loop
                    break;
                }
            } else if (pend2 == 0) {

                // ---- visit one BVH8 node
                const PRT_GLOBAL uint4* np = reinterpret_cast<const PRT_GLOBAL uint4*>(nodes + node);
                const uint4 w0 = np[0], w1 = np[1], w2 = np[2], w3 = np[3], w4 = np[4];
                if (kCount) ++n_nodes;
                const int ex = (int)(int8_t)(w0.w & 0xffu);
                const int ey = (int)(int8_t)((w0.w >> 8) & 0xffu);
                const int ez = (int)(int8_t)((w0.w >> 16) & 0xffu);
                const uint32_t imask = w0.w >> 24;
                const float ax = __uint_as_float((uint32_t)(ex + 127) << 23) * inv_x;
                const float ay = __uint_as_float((uint32_t)(ey + 127) << 23) * inv_y;
                const float az = __uint_as_float((uint32_t)(ez + 127) << 23) * inv_z;
                const float bx = (as_f(w0.x) - o.x) * inv_x;
                const float by = (as_f(w0.y) - o.y) * inv_y;
                const float bz = (as_f(w0.z) - o.z) * inv_z;
                // A plane byte q enters as the half 1024 + q (two per v_perm,
                // read by v_fma_mix_f32 without a conversion), so the distance
                // is fma(1024 + q, a, b' - 1024 a), b' = b -/+ slack.  Slack
                // 2^-21 (512|a| + 2|b|) covers the rounding of b, of b' and of
                // b' - 1024 a and of the fma, whose operands now reach
                // 1279|a| + |b'| + 1024|a|, with room to spare.
                const float slx = 4.76837158203125e-7f * (512.0f * fabsf(ax) + 2.0f * fabsf(bx));
                const float sly = 4.76837158203125e-7f * (512.0f * fabsf(ay) + 2.0f * fabsf(by));
                const float slz = 4.76837158203125e-7f * (512.0f * fabsf(az) + 2.0f * fabsf(bz));
                const float bxn = __builtin_fmaf(-1024.0f, ax, bx - slx), bxf = __builtin_fmaf(-1024.0f, ax, bx + slx);
                const float byn = __builtin_fmaf(-1024.0f, ay, by - sly), byf = __builtin_fmaf(-1024.0f, ay, by + sly);
                const float bzn = __builtin_fmaf(-1024.0f, az, bz - slz), bzf = __builtin_fmaf(-1024.0f, az, bz + slz);
                // near/far planes per axis from the direction signs
                const uint32_t nx0 = neg_x ? w3.z : w2.x, nx1 = neg_x ? w3.w : w2.y;
                const uint32_t fx0 = neg_x ? w2.x : w3.z, fx1 = neg_x ? w2.y : w3.w;
                const uint32_t ny0 = neg_y ? w4.x : w2.z, ny1 = neg_y ? w4.y : w2.w;
                const uint32_t fy0 = neg_y ? w2.z : w4.x, fy1 = neg_y ? w2.w : w4.y;
                const uint32_t nz0 = neg_z ? w4.z : w3.x, nz1 = neg_z ? w4.w : w3.y;
                const uint32_t fz0 = neg_z ? w3.x : w4.z, fz1 = neg_z ? w3.y : w4.w;
                const uint32_t meta0 = w1.z, meta1 = w1.w;
                // the halves 1024 + q of plane word w: bytes 0, 1 (lo) and 2, 3 (hi)
                auto halves = [](uint32_t w) {
                    return make_uint2(__builtin_amdgcn_perm(0x64646464u, w, 0x04010400u),
                                      __builtin_amdgcn_perm(0x64646464u, w, 0x04030402u));
                };
                const uint2 hnx0 = halves(nx0), hnx1 = halves(nx1), hfx0 = halves(fx0), hfx1 = halves(fx1);
                const uint2 hny0 = halves(ny0), hny1 = halves(ny1), hfy0 = halves(fy0), hfy1 = halves(fy1);
                const uint2 hnz0 = halves(nz0), hnz1 = halves(nz1), hfz0 = halves(fz0), hfz1 = halves(fz1);
                // 1024 + byte k of the plane word pair (h0: bytes 0-3, h1: bytes 4-7)
                auto qf = [](uint2 h0, uint2 h1, int k) -> float {
                    const uint2 h = k < 4 ? h0 : h1;
                    const uint32_t w = (k & 2) ? h.y : h.x;
                    const unsigned short bits = (unsigned short)((k & 1) ? (w >> 16) : (w & 0xffffu));
                    return (float)__builtin_bit_cast(_Float16, bits);
                };

                // branchless child tests; empty slots carry inverted boxes, and a
                // leaf slot with meta 0 holds no triangle, so no leaf mask is needed
                uint32_t hits = 0;
                float tk[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const float tnx = __builtin_fmaf(qf(hnx0, hnx1, k), ax, bxn);
                    const float tny = __builtin_fmaf(qf(hny0, hny1, k), ay, byn);
                    const float tnz = __builtin_fmaf(qf(hnz0, hnz1, k), az, bzn);
                    const float tfx = __builtin_fmaf(qf(hfx0, hfx1, k), ax, bxf);
                    const float tfy = __builtin_fmaf(qf(hfy0, hfy1, k), ay, byf);
                    const float tfz = __builtin_fmaf(qf(hfz0, hfz1, k), az, bzf);
                    const float tn = fmaxf(fmaxf(fmaxf(tnx, tny), tnz), tmin);
                    const float tf = fminf(fminf(fminf(tfx, tfy), tfz), tbest);
                    tk[k] = tn;
                    hits |= (tn <= tf ? 1u : 0u) << k;
                }
                const uint32_t hit_int = hits & imask;
                uint32_t hit_leaf = hits & ~imask;

                // ---- leaf slots: one pending mask over the node's triangle
                // records (a node owns at most 8 x kLeafMax <= 32 of them),
                // the first group if it is free, else the second
                uint32_t newp = 0;
                while (hit_leaf) {
                    const int k = __builtin_ctz(hit_leaf);
                    hit_leaf &= hit_leaf - 1;
                    const uint32_t mb = ((k < 4 ? meta0 : meta1) >> (8 * (k & 3))) & 0xffu;
                    newp |= ((1u << (mb >> 5)) - 1u) << (mb & 31u);
                }
                if (newp != 0) {
                    if (pend == 0) {
                        pend = newp;
                        pend_base = w1.y;
                    } else {
                        pend2 = newp;
                        pend2_base = w1.y;
                    }
                }

                // ---- internal slots: the nearest is visited next, the others
                // are pushed with the second nearest on top.  Unrolled over the
                // 8 slots with per-slot predicates, so a wave pays one pass
                // whatever the number of hit children per lane.
                node = -1;
                if (hit_int) {
                    // nearest and second-nearest hit internal child by integer
                    // minimum over keys (entry distance bits with the slot in
                    // the low 3 bits; the distances are >= 0, and a -0 loses
                    // its sign bit).  Ties and distances within 8 ulps order by
                    // slot; node_t is the key's distance, at most 7 ulps below
                    // the true one, so culling with it stays conservative.
                    uint32_t key[8];
                    const uint32_t not_int = ~hit_int;
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        // all ones unless bit k of hit_int is set (a 1-bit signed field)
                        const uint32_t miss = (uint32_t)__builtin_amdgcn_sbfe((int)not_int, (unsigned)k, 1u);
                        key[k] = (__float_as_uint(tk[k]) & 0x7ffffff8u) | (uint32_t)k | miss;
                    }
                    const uint32_t m1 = min(min(min(key[0], key[1]), min(key[2], key[3])),
                                            min(min(key[4], key[5]), min(key[6], key[7])));
                    // second smallest: no other key equals m1, so key - m1 - 1
                    // (mod 2^32) maps m1 to the largest value and keeps the
                    // order of the keys above it
                    uint32_t e[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k) e[k] = key[k] + ~m1;
                    const uint32_t m2 = min(min(min(e[0], e[1]), min(e[2], e[3])), min(min(e[4], e[5]), min(e[6], e[7]))) +
                                        m1 + 1u;
                    const int k1 = (int)(m1 & 7u);
                    const int k2 = (int)(m2 & 7u);
                    const uint32_t child_base = w1.x;
                    node = (int32_t)(child_base + __popc(imask & ((1u << k1) - 1u)));
                    node_t = __uint_as_float(m1 & ~7u);
                    const uint32_t pushm = hit_int & ~(1u << k1);
                    if (pushm) {
                        const int npush = __popc(pushm);
                        // make room in the LDS ring: move its oldest entries to
                        // the lane's global slice
                        bool ok = true;
                        while (sp + npush - bottom > kStack) {
                            if (bottom >= spill_cap) { atomicOr(err, 1u); ok = false; break; }
                            const uint32_t boff = ring_off(bottom);
                            my_spill[bottom] = make_uint2(ring_node(boff), __float_as_uint(ring_t(boff)));
                            ++bottom;
                            if (kCount) ++n_spill;
                        }
                        if (ok) {
                            // branch-free: slot k goes to sp + (pushed slots
                            // below k, k2 left out).  A slot that is not pushed
                            // (or is k2) lands on the position of the next
                            // pushed slot above it, written later, or on the
                            // top, written last with k2's entry; so every
                            // write stays inside [sp, sp + npush).  k2 is in
                            // pushm: it is the second hit internal child.
                            const uint32_t pm2 = pushm & ~(1u << k2);
#pragma unroll
                            for (int k = 0; k < 8; ++k) {
                                const int pos = __popc(pm2 & ((1u << k) - 1u));
                                const uint32_t off = ring_off(sp + pos);
                                ring_node(off) = child_base + __popc(imask & ((1u << k) - 1u));
                                ring_t(off) = tk[k];
                            }
                            // k2 on top, with its key's distance (at most 7
                            // ulps below its entry distance: conservative)
                            const uint32_t off = ring_off(sp + npush - 1);
                            ring_node(off) = child_base + __popc(imask & ((1u << k2) - 1u));
                            ring_t(off) = __uint_as_float(m2 & ~7u);
                            sp += npush;
                        }
                    }
                }
            }

            if ((int)__popcll(__ballot(1)) <= keep) break;
            // lookahead: look at the stop flag every 8 steps (its load is
            // waited for with the next node's)
            if (kLookMode && look && !live && !stop && (++poll_steps & 7u) == 0) {
                const uint32_t f = __hip_atomic_load(stop_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((uint32_t)__builtin_amdgcn_readfirstlane(f) == pl.look_gen) {
                    stop = true;
                    exhausted = true;
                    break;
                }
            }
        }
    }

    if (kLookMode && look && live) look_leave();

    // ---- cooperative tail (kModeDiffuse).  Each of the wave's remaining
    // paths (at most 8) moves to a group of G = 8 lanes that hold its ray
    // state alike.  Per step, lane j of a group tests child j of the visited
    // node, or one of G pending triangles, and the group's choices (nearest children, closest hit) come
    // from shuffles inside the group.  The closest hit is the same (t, primId)
    // minimum as lane by lane, whatever the order of the tests, so images do
    // not change; a path's tail of dependent steps gets shorter.  The group's
    // stack is its lanes' LDS rings joined (G x kStack entries), spilling to
    // its first lane's slice.
    if constexpr (kCoop) {
        const uint64_t cm = __ballot(cont);
        if (coop_on && cm != 0) {
            const int npaths = __popcll(cm);
            constexpr int lg = 3;                 // log2 G: 8 lanes per path
            constexpr int G = 1 << lg;
            const int gj = lane & (G - 1);        // this lane's slot in its group
            const int gbase = lane & ~(G - 1);    // first lane of the group
            const int grp = lane >> lg;
            constexpr uint32_t gmask = (1u << G) - 1u;
            constexpr int kGStack = G * kStack;
            uint64_t mm = cm;
            for (int k = 0; k < grp; ++k) mm &= mm - 1;
            const int src = mm ? __ffsll((unsigned long long)mm) - 1 : 0;
            const bool gact = grp < npaths;
            o = mk3(__shfl(o.x, src, 64), __shfl(o.y, src, 64), __shfl(o.z, src, 64));
            d = mk3(__shfl(d.x, src, 64), __shfl(d.y, src, 64), __shfl(d.z, src, 64));
            pixel = __shfl(pixel, src, 64);
            rng = __shfl(rng, src, 64);
            depth = __shfl(depth, src, 64);
            lw = __shfl(lw, src, 64);
            cont = false;
            active = gact;
            tracing = false;
            if (gact) start_ray(o, d, 0.0f, 3.402823466e38f);
            uint2* g_spill = spill + ((size_t)blockIdx.x * kBlock + (tid & ~(G - 1))) * (size_t)stack_entries;
            auto gslot = [&](int q) { return ((q >> lg) & (kStack - 1)) * kBlock + (tid & ~(G - 1)) + (q & (G - 1)); };
            for (;;) {
                while (__ballot(tracing) != 0) {
                    if (kCount) {
                        const bool lead = lane == __ffsll((unsigned long long)__ballot(1)) - 1;
                        if (tracing) {
                            ++n_lsteps;
                            ++n_dlsteps;
                        }
                        if (lead) {
                            ++n_wsteps;
                            ++n_dwsteps;
                        }
                    }
                    if (tracing && pend == 0) {
                        if (node >= 0 && node_t > tbest) node = -1;
                        if (node < 0) {
                            while (sp > 0) {
                                --sp;
                                uint32_t nid;
                                float nt;
                                if (sp < bottom) {
                                    const uint2 e = g_spill[sp];
                                    bottom = sp;
                                    nid = e.x;
                                    nt = as_f(e.y);
                                } else {
                                    const int slot = gslot(sp & (kGStack - 1));
                                    nid = st_node[slot];
                                    nt = st_t[slot];
                                }
                                if (nt <= tbest) { node = (int32_t)nid; break; }
                            }
                            if (node < 0) tracing = false;
                        }
                    }
                    const bool node_step = tracing && pend == 0;
                    // this lane's children: slots gj + c G, c < 8 / G
                    float tkc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                    bool hitc[4] = {false, false, false, false};
                    uint4 w0{}, w1{};
                    if (node_step) {
                        const PRT_GLOBAL uint4* np = reinterpret_cast<const PRT_GLOBAL uint4*>(nodes + node);
                        w0 = np[0];
                        w1 = np[1];
                        const uint4 w2 = np[2], w3 = np[3], w4 = np[4];
                        if (kCount && gj == 0) ++n_nodes;
                        const int ex = (int)(int8_t)(w0.w & 0xffu);
                        const int ey = (int)(int8_t)((w0.w >> 8) & 0xffu);
                        const int ez = (int)(int8_t)((w0.w >> 16) & 0xffu);
                        const float ax = __uint_as_float((uint32_t)(ex + 127) << 23) * inv_x;
                        const float ay = __uint_as_float((uint32_t)(ey + 127) << 23) * inv_y;
                        const float az = __uint_as_float((uint32_t)(ez + 127) << 23) * inv_z;
                        const float bx = (as_f(w0.x) - o.x) * inv_x;
                        const float by = (as_f(w0.y) - o.y) * inv_y;
                        const float bz = (as_f(w0.z) - o.z) * inv_z;
                        const float slx = 4.76837158203125e-7f * (255.0f * fabsf(ax) + fabsf(bx));
                        const float sly = 4.76837158203125e-7f * (255.0f * fabsf(ay) + fabsf(by));
                        const float slz = 4.76837158203125e-7f * (255.0f * fabsf(az) + fabsf(bz));
#pragma unroll
                        for (int c = 0; c < 4; ++c) {
                            const int k = gj + (c << lg);
                            if (k < 8) {
                                // child k: byte k & 3 of the words of its half of the node
                                const bool hi = k >= 4;
                                const int kb = k & 3;
                                const uint32_t qlx = hi ? w2.y : w2.x, qhx = hi ? w3.w : w3.z;
                                const uint32_t qly = hi ? w2.w : w2.z, qhy = hi ? w4.y : w4.x;
                                const uint32_t qlz = hi ? w3.y : w3.x, qhz = hi ? w4.w : w4.z;
                                const float lox = byte_f(qlx, kb), hix = byte_f(qhx, kb);
                                const float loy = byte_f(qly, kb), hiy = byte_f(qhy, kb);
                                const float loz = byte_f(qlz, kb), hiz = byte_f(qhz, kb);
                                const float tnx = __builtin_fmaf(neg_x ? hix : lox, ax, bx - slx);
                                const float tny = __builtin_fmaf(neg_y ? hiy : loy, ay, by - sly);
                                const float tnz = __builtin_fmaf(neg_z ? hiz : loz, az, bz - slz);
                                const float tfx = __builtin_fmaf(neg_x ? lox : hix, ax, bx + slx);
                                const float tfy = __builtin_fmaf(neg_y ? loy : hiy, ay, by + sly);
                                const float tfz = __builtin_fmaf(neg_z ? loz : hiz, az, bz + slz);
                                tkc[c] = fmaxf(fmaxf(fmaxf(tnx, tny), tnz), tmin);
                                const float tf = fminf(fminf(fminf(tfx, tfy), tfz), tbest);
                                hitc[c] = tkc[c] <= tf;
                            }
                        }
                    }
                    uint32_t ghits = 0;
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        ghits |= ((uint32_t)(__ballot(hitc[c]) >> gbase) & gmask) << (c << lg);
                    ghits &= 0xffu;
                    if (node_step) {
                        const uint32_t imask = w0.w >> 24;
                        const uint32_t hit_int = ghits & imask;
                        uint32_t hit_leaf = ghits & ~imask;
                        while (hit_leaf) {
                            const int k = __builtin_ctz(hit_leaf);
                            hit_leaf &= hit_leaf - 1;
                            const uint32_t mb = ((k < 4 ? w1.z : w1.w) >> (8 * (k & 3))) & 0xffu;
                            pend |= ((1u << (mb >> 5)) - 1u) << (mb & 31u);
                        }
                        pend_base = w1.y;
                        node = -1;
                        if (hit_int) {
                            // nearest and second nearest internal child over the group
                            float t1 = __builtin_inff(), t2 = __builtin_inff();
                            int k1 = 8, k2 = 8;
#pragma unroll
                            for (int c = 0; c < 4; ++c) {
                                const int k = gj + (c << lg);
                                if (k < 8 && ((hit_int >> k) & 1u) && (tkc[c] < t1 || (tkc[c] == t1 && k < k1))) {
                                    t1 = tkc[c];
                                    k1 = k;
                                }
                            }
                            for (int x = 1; x < G; x <<= 1) {
                                const float ot = __shfl_xor(t1, x, 64);
                                const int ok = __shfl_xor(k1, x, 64);
                                if (ot < t1 || (ot == t1 && ok < k1)) { t1 = ot; k1 = ok; }
                            }
#pragma unroll
                            for (int c = 0; c < 4; ++c) {
                                const int k = gj + (c << lg);
                                if (k < 8 && k != k1 && ((hit_int >> k) & 1u) &&
                                    (tkc[c] < t2 || (tkc[c] == t2 && k < k2))) {
                                    t2 = tkc[c];
                                    k2 = k;
                                }
                            }
                            for (int x = 1; x < G; x <<= 1) {
                                const float ot = __shfl_xor(t2, x, 64);
                                const int ok = __shfl_xor(k2, x, 64);
                                if (ot < t2 || (ot == t2 && ok < k2)) { t2 = ot; k2 = ok; }
                            }
                            const uint32_t child_base = w1.x;
                            node = (int32_t)(child_base + __popc(imask & ((1u << k1) - 1u)));
                            node_t = t1;
                            const uint32_t pushm = hit_int & ~(1u << k1);
                            if (pushm) {
                                const int npush = __popc(pushm);
                                bool ok = true;
                                while (sp + npush - bottom > kGStack) {
                                    if (bottom >= spill_cap) { atomicOr(err, 1u); ok = false; break; }
                                    const int bslot = gslot(bottom & (kGStack - 1));
                                    g_spill[bottom] = make_uint2(st_node[bslot], __float_as_uint(st_t[bslot]));
                                    ++bottom;
                                    if (kCount && gj == 0) ++n_spill;
                                }
                                if (ok) {
                                    const int klast = 31 - __builtin_clz(pushm);
                                    if (k2 > 7 || !((pushm >> k2) & 1u)) k2 = klast;
                                    const int pos_k2 = __popc(pushm & ((1u << k2) - 1u));
#pragma unroll
                                    for (int c = 0; c < 4; ++c) {
                                        const int k = gj + (c << lg);
                                        if (k < 8 && ((pushm >> k) & 1u)) {
                                            int pos = __popc(pushm & ((1u << k) - 1u));
                                            pos = k == k2 ? npush - 1 : (k == klast ? pos_k2 : pos);
                                            const int slot = gslot((sp + pos) & (kGStack - 1));
                                            st_node[slot] = child_base + __popc(imask & ((1u << k) - 1u));
                                            st_t[slot] = tkc[c];
                                        }
                                    }
                                    sp += npush;
                                }
                            }
                        }
                    }
                    if (tracing && pend != 0 && !node_step) {
                        // up to G pending triangles at once: lane j takes the
                        // j-th lowest pending record
                        uint32_t mj = pend;
                        for (int k = 0; k < gj; ++k) mj &= mj - 1;
                        for (int k = 0; k < G; ++k) pend &= pend - 1;
                        float ct = __builtin_inff();
                        int32_t cp = INT32_MAX;
                        uint32_t crec = 0;
                        if (mj != 0) {
                            const uint32_t rec = pend_base + (uint32_t)__builtin_ctz(mj);
                            const PRT_GLOBAL float* tp = tris + 9 * (size_t)rec;
                            const f3 p0 = mk3(tp[0], tp[1], tp[2]);
                            const f3 p1 = mk3(tp[3], tp[4], tp[5]);
                            const f3 p2 = mk3(tp[6], tp[7], tp[8]);
                            if (kCount) ++n_tris;
                            TriCandidate c;
                            if (tri_test(wr, o, d, p0, p1, p2, tmin, tbest, c)) {
                                ct = c.t;
                                cp = prim_ids[rec];
                                crec = rec;
                            }
                        }
                        for (int x = 1; x < G; x <<= 1) {
                            const float ot = __shfl_xor(ct, x, 64);
                            const int32_t op = __shfl_xor(cp, x, 64);
                            const uint32_t orr = __shfl_xor(crec, x, 64);
                            if (ot < ct || (ot == ct && op < cp)) { ct = ot; cp = op; crec = orr; }
                        }
                        if (cp != INT32_MAX && (ct < tbest || best_prim < 0 || cp < best_prim)) {
                            tbest = ct;
                            best_prim = cp;
                            best_rec = crec;
                        }
                    }
                }
                // ---- shade each group's finished ray (the same in its G lanes;
                // its first lane counts the ray and writes the pixel)
                if (active) {
                    if constexpr (kCount) {
                        if (gj == 0) atomicAdd(&depth_hist[min(depth, kMaxQueueDepth - 1)], 1u);
                    }
                    if (gj == 0) ++n_rays;
                    float4* acc = pl.accum + pixel;
                    if (best_prim < 0 || depth == max_depth) {
                        const float c = best_prim < 0 ? lw : 0.0f;
                        if (gj == 0) {
                            const float4 a = pl.fast ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : *acc;
                            *acc = make_float4(a.x + c, a.y + c, a.z + c, a.w + 1.0f);
                        }
                        active = false;
                    } else {
                        const PRT_GLOBAL float* tp = tris + 9 * (size_t)best_rec;
                        const f3 p0 = mk3(tp[0], tp[1], tp[2]);
                        const f3 p1 = mk3(tp[3], tp[4], tp[5]);
                        const f3 p2 = mk3(tp[6], tp[7], tp[8]);
                        lw *= albedo;
                        const Frame fr = hit_frame(p0, p1, p2);
                        const float s1 = sampler_next(rng);
                        const float s2 = sampler_next(rng);
                        f3 no, nd;
                        diffuse_bounce(o, d, tbest, fr, s1, s2, no, nd);
                        ++depth;
                        start_ray(no, nd, 0.0f, 3.402823466e38f);
                    }
                }
                if (__ballot(active) == 0) break;
            }
        }
    }

    if (kCount) {
        for (int off = 32; off > 0; off >>= 1) {
            n_nodes += __shfl_down(n_nodes, off, 64);
            n_tris += __shfl_down(n_tris, off, 64);
            n_spill += __shfl_down(n_spill, off, 64);
            n_wsteps += __shfl_down(n_wsteps, off, 64);
            n_lsteps += __shfl_down(n_lsteps, off, 64);
            n_wshade += __shfl_down(n_wshade, off, 64);
            n_dwsteps += __shfl_down(n_dwsteps, off, 64);
            n_dlsteps += __shfl_down(n_dlsteps, off, 64);
        }
        if (lane == 0) {
            unsigned long long* st = kRayQuery ? tl.stats : (kMode == kModeBatch ? bl.stats : pl.stats);
            atomicAdd(&st[0], n_nodes);
            atomicAdd(&st[1], n_tris);
            atomicAdd(&st[2], n_spill);
            atomicAdd(&st[4], n_wsteps);
            atomicAdd(&st[5], n_lsteps);
            atomicAdd(&st[6], n_wshade);
            atomicAdd(&st[7], n_dwsteps);
            atomicAdd(&st[8], n_dlsteps);
            // drain window: earliest exhaustion, latest exit (wall clock ticks)
            if (t_exhausted) atomicMin(&st[9], t_exhausted);
            atomicMax(&st[10], (unsigned long long)wall_clock64());
        }
    }
    if constexpr (kLookMode) {
        if (look) {
#pragma unroll
            for (int l = 0; l < kLookLevels; ++l) {
                unsigned long long r = n_ahead[l];
                for (int off = 32; off > 0; off >>= 1) r += __shfl_down(r, off, 64);
                if (lane == 0 && r) atomicAdd(&look->rays[l + 1], r);
            }
        }
    }
    if constexpr (!kRayQuery) {
        unsigned long long r = n_rays;
        for (int off = 32; off > 0; off >>= 1) r += __shfl_down(r, off, 64);
        if (lane == 0 && r) atomicAdd(&(kMode == kModeBatch ? bl.stats : pl.stats)[3], r);
        if constexpr (kCount && kMode != kModeBatch) {
            __syncthreads();
            if (tid < kMaxQueueDepth && depth_hist[tid])
                atomicAdd(&pl.depth_rays[tid], (unsigned long long)depth_hist[tid]);
        }
    }
}

template <int kMode>
hipError_t launch_mode(prt_hip_ctx* ctx, const TraceLaunch& t, const PathLaunch& p, bool instrumented)
{
    const SceneDevice sc = ctx->scene.device();
    uint32_t* err = ctx->d_counters + 32;
    const BatchLaunch bl{};
    const int grid = kMode == kModeDiffuse ? ctx->diffuse_grid : ctx->trace_grid;
    if (instrumented) {
        k_trace<true, kMode><<<grid, kTraceBlock, 0, ctx->stream>>>(
            sc, t, p, bl, ctx->d_spill, ctx->stack_entries, kRefill, err);
        return hipGetLastError();
    }
    if constexpr (kMode == kModeDiffuse) {
        if (p.look != nullptr) {
            k_trace<false, kMode, kTraceBlock, kStackLds, true><<<grid, kTraceBlock, 0, ctx->stream>>>(
                sc, t, p, bl, ctx->d_spill, ctx->stack_entries, kRefill, err);
            return hipGetLastError();
        }
    }
    k_trace<false, kMode><<<grid, kTraceBlock, 0, ctx->stream>>>(
        sc, t, p, bl, ctx->d_spill, ctx->stack_entries, kRefill, err);
    return hipGetLastError();
}

// Batch kernel variants the autotuner chooses from: (block size, LDS stack).
template <int kBlock, int kStack>
hipError_t launch_batch_variant(prt_hip_ctx* ctx, const BatchLaunch& bl, int blocks_per_cu, int refill,
                                bool instrumented, int* max_blocks_per_cu)
{
    if (max_blocks_per_cu) {
        int per_cu = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, k_trace<false, kModeBatch, kBlock, kStack>, kBlock, 0);
        *max_blocks_per_cu = per_cu;
        return e;
    }
    const int grid = blocks_per_cu * ctx->num_cus;
    uint32_t* err = ctx->d_counters + 32;
    const SceneDevice sc{};
    const TraceLaunch t{};
    const PathLaunch p{};
    if (instrumented)
        k_trace<true, kModeBatch, kBlock, kStack><<<grid, kBlock, 0, ctx->stream>>>(
            sc, t, p, bl, ctx->d_spill, ctx->stack_entries, refill, err);
    else
        k_trace<false, kModeBatch, kBlock, kStack><<<grid, kBlock, 0, ctx->stream>>>(
            sc, t, p, bl, ctx->d_spill, ctx->stack_entries, refill, err);
    return hipGetLastError();
}

}  // namespace

int ensure_trace_resources(prt_hip_ctx* ctx)
{
    if (ctx->trace_per_cu == 0) {
        // the lookahead variant's workgroups must all be resident at once: a
        // workgroup that starts only when another ends would hold the frame's
        // stop (every wave leaves the live count) until then
        int per_cu = 0, per_cu_look = 0;
        PRT_TRY(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace<false, kModeDiffuse>,
                                                                  kTraceBlock, 0));
        PRT_TRY(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                         &per_cu_look, k_trace<false, kModeDiffuse, kTraceBlock, kStackLds, true>, kTraceBlock, 0));
        if (per_cu_look < per_cu) per_cu = per_cu_look;
        if (per_cu * kTraceBlock > 2048) per_cu = 2048 / kTraceBlock;
        if (per_cu < 1) per_cu = 1;
        ctx->trace_per_cu = per_cu;
    }
    ctx->trace_grid = ctx->trace_per_cu * ctx->num_cus;
    // Diffuse render launches on a scene whose nodes and triangles outgrow the
    // memory-side cache (256 MiB) run one workgroup per CU fewer.  Paths on such
    // a scene take long, so a frame's tail (its last paths, in sparse waves)
    // is long, and fewer resident waves shorten it by more than they slow the
    // rest; the albedo and trace launches keep the full occupancy, which is
    // faster for them
    const SceneBuffers& s = ctx->scene;
    const int64_t scene_bytes =
        s.node_count * (int64_t)sizeof(Bvh8Node) + s.tri_count * (int64_t)(9 * sizeof(float) + sizeof(int32_t));
    int per_cu = ctx->trace_per_cu;
    if (scene_bytes > kLargeSceneBytes && per_cu > 1) per_cu -= 1;
    ctx->diffuse_grid = per_cu * ctx->num_cus;
    // worst case stack: 7 pushes per level of the tree, plus slack; sized for
    // the deepest scene and for every lane a CU can hold (2048), so any
    // persistent launch shape fits
    int depth = ctx->scene.bvh_depth;
    for (const SceneBuffers& b : ctx->batch_scenes) depth = depth > b.bvh_depth ? depth : b.bvh_depth;
    // The per-lane stride only grows, so a shallower scene keeps the buffer;
    // a captured launch graph holds the buffer and stride as kernel
    // arguments, so every change bumps spill_gen, which the graph's owner
    // compares before a replay.
    int need = 7 * (depth + 1) + 8;
    if (need < kStackLds) need = kStackLds;
    if (need < ctx->stack_entries) need = ctx->stack_entries;
    const int64_t lanes = (int64_t)ctx->num_cus * 2048;
    const int64_t bytes = lanes * need * (int64_t)sizeof(uint2);
    if (bytes > ctx->spill_bytes || need != ctx->stack_entries) {
        ctx->spill_gen += 1;
        if (ctx->d_spill) { (void)hipFree(ctx->d_spill); ctx->d_spill = nullptr; }
        ctx->spill_bytes = 0;
        ctx->stack_entries = 0;
        PRT_TRY(ctx, hipMalloc(&ctx->d_spill, bytes));
        ctx->spill_bytes = bytes;
        ctx->stack_entries = need;
    }
    return 0;
}

hipError_t launch_trace(prt_hip_ctx* ctx, const TraceLaunch& t, bool instrumented)
{
    return launch_mode<kModeTrace>(ctx, t, PathLaunch{}, instrumented);
}

hipError_t launch_trace_scratch(prt_hip_ctx* ctx, const TraceLaunch& t, uint2* spill, int stack_entries,
                                uint32_t* err, int grid)
{
    k_trace<false, kModeTrace><<<grid, kTraceBlock, 0, ctx->stream>>>(ctx->scene.device(), t, PathLaunch{},
                                                                     BatchLaunch{}, spill, stack_entries, kRefill,
                                                                     err);
    return hipGetLastError();
}

hipError_t launch_paths_rgb(prt_hip_ctx* ctx, const PathLaunch& p, bool instrumented)
{
    return launch_mode<kModeRgb>(ctx, TraceLaunch{}, p, instrumented);
}

hipError_t launch_adjoint(prt_hip_ctx* ctx, const PathLaunch& p, bool instrumented)
{
    return launch_mode<kModeAdjoint>(ctx, TraceLaunch{}, p, instrumented);
}

hipError_t launch_occluded(prt_hip_ctx* ctx, const TraceLaunch& t, bool instrumented)
{
    return launch_mode<kModeOccluded>(ctx, t, PathLaunch{}, instrumented);
}

hipError_t launch_paths(prt_hip_ctx* ctx, const PathLaunch& p, int renderer, bool instrumented)
{
    TraceLaunch t{};
    switch (renderer) {
    case PRT_RENDERER_PRIMARY: return launch_mode<kModePrimary>(ctx, t, p, instrumented);
    case PRT_RENDERER_DIFFUSE: return launch_mode<kModeDiffuse>(ctx, t, p, instrumented);
    case PRT_RENDERER_DIFFUSE2: return launch_mode<kModeDiffuse2>(ctx, t, p, instrumented);
    default: return hipErrorInvalidValue;
    }
}

const BatchKernelConfig kBatchVariants[kBatchVariantCount] = {
    {256, 16}, {256, 8}, {128, 16}, {128, 8},
};

hipError_t launch_batch(prt_hip_ctx* ctx, const BatchLaunch& bl, const BatchTuning& tu, bool instrumented,
                        int* max_blocks_per_cu)
{
    switch (tu.variant) {
    case 0: return launch_batch_variant<256, 16>(ctx, bl, tu.blocks_per_cu, tu.refill, instrumented, max_blocks_per_cu);
    case 1: return launch_batch_variant<256, 8>(ctx, bl, tu.blocks_per_cu, tu.refill, instrumented, max_blocks_per_cu);
    case 2: return launch_batch_variant<128, 16>(ctx, bl, tu.blocks_per_cu, tu.refill, instrumented, max_blocks_per_cu);
    case 3: return launch_batch_variant<128, 8>(ctx, bl, tu.blocks_per_cu, tu.refill, instrumented, max_blocks_per_cu);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace prt
