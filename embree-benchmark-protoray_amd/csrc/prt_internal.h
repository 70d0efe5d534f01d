// prt_internal.h -- device data layout and the context behind the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "prt_math.h"
#include "../../include/prt_hip.h"

namespace prt {

// BVH8 node, 80 B, quantised child boxes (S. Ylitie, T. Karras, S. Laine,
// "Efficient Incoherent Ray Traversal on GPUs Through Compressed Wide BVHs",
// HPG 2017).  Child k's box is p + 2^e * [qlo_k, qhi_k] per axis.  Internal
// children are stored contiguously from child_base in slot order (their index
// is child_base + popcount(imask below k)); a leaf slot k has meta[k] =
// count << 5 | offset and owns triangle records tri_base + offset ..
// + count - 1.  An empty slot has meta 0, imask bit 0 and an inverted box.
struct alignas(16) Bvh8Node {
    float px, py, pz;
    uint8_t ex, ey, ez;   // int8 exponents stored as two's complement bytes
    uint8_t imask;
    uint32_t child_base;
    uint32_t tri_base;
    uint8_t meta[8];
    uint8_t qlo_x[8], qlo_y[8], qlo_z[8];
    uint8_t qhi_x[8], qhi_y[8], qhi_z[8];
};
static_assert(sizeof(Bvh8Node) == 80, "BVH8 node must be 80 bytes");

constexpr int kLeafMax = 3;           // triangles per leaf slot
constexpr int kTraceBlock = 256;      // threads per trace workgroup
constexpr int64_t kLargeSceneBytes = int64_t(256) << 20;  // nodes + triangles beyond MI355X's 256 MiB Infinity Cache
constexpr int kMaxQueueDepth = 16;    // per-depth ray statistics bins; the last bin holds depths >= 15
constexpr int kMaxDepth = 1024;       // largest max_depth accepted (an R^k throughput below 1 is 0 in fp32 long before)

// Triangle record: the three fp32 vertices (36 B), in BVH8 leaf order.
struct TriRecord {
    float v[9];
};

struct SceneDevice {
    const Bvh8Node* nodes;
    const float* tris;        // 9 floats per record
    const int32_t* prim_ids;  // original triangle index per record
    int64_t tri_count;
    // shading attributes (nullptr when the mesh has none): per-vertex normals
    // (3 floats) and texture coordinates (2 floats), reached through the
    // uploaded vertex indices (3 per original triangle)
    const int32_t* idx;
    const float* normals;
    const float* texcoords;
};

// Global-address-space view of a device pointer.  Loads through it compile to
// global_load; through a generic pointer that was itself read from memory (a
// scene table entry) they compile to flat loads, which also wait on the LDS
// counter.  (The host pass of the compiler sees plain pointers.)
#if defined(__HIP_DEVICE_COMPILE__)
#define PRT_GLOBAL __attribute__((address_space(1)))
#else
#define PRT_GLOBAL
#endif
template <class T>
__device__ __forceinline__ const PRT_GLOBAL T* global_view(const T* p)
{
    return (const PRT_GLOBAL T*)p;
}

// Shading frame of a hit on original triangle `prim` (vertices p0, p1, p2,
// hit coordinates (bu, bv)) for a ray along dir, from the per-vertex
// attributes that are present: the full ShadingContext.
__device__ __forceinline__ Frame attribute_frame(const int32_t* idx, const float* normals,
                                                 const float* texcoords, int prim, f3 p0, f3 p1, f3 p2,
                                                 f3 dir, float bu, float bv)
{
    CornerAttributes a{};
    a.has_normals = normals != nullptr;
    a.has_texcoords = texcoords != nullptr;
    const PRT_GLOBAL int32_t* gi = global_view(idx);
    const int i0 = gi[3 * (size_t)prim], i1 = gi[3 * (size_t)prim + 1], i2 = gi[3 * (size_t)prim + 2];
    if (a.has_normals) {
        const PRT_GLOBAL float* gn = global_view(normals);
        a.n0 = mk3(gn[3 * (size_t)i0], gn[3 * (size_t)i0 + 1], gn[3 * (size_t)i0 + 2]);
        a.n1 = mk3(gn[3 * (size_t)i1], gn[3 * (size_t)i1 + 1], gn[3 * (size_t)i1 + 2]);
        a.n2 = mk3(gn[3 * (size_t)i2], gn[3 * (size_t)i2 + 1], gn[3 * (size_t)i2 + 2]);
    }
    if (a.has_texcoords) {
        const PRT_GLOBAL float* gt = global_view(texcoords);
        a.u0 = gt[2 * (size_t)i0];
        a.v0 = gt[2 * (size_t)i0 + 1];
        a.u1 = gt[2 * (size_t)i1];
        a.v1 = gt[2 * (size_t)i1 + 1];
        a.u2 = gt[2 * (size_t)i2];
        a.v2 = gt[2 * (size_t)i2 + 1];
    }
    return shading_frame(p0, p1, p2, dir, bu, bv, a);
}

// One scene on the device: the uploaded mesh and its BVH8.
struct SceneBuffers {
    int64_t tri_count = 0, vertex_count = 0;
    int32_t* d_idx = nullptr;
    float* d_pos = nullptr;
    Bvh8Node* d_nodes = nullptr;
    float* d_tris = nullptr;
    int32_t* d_prim_ids = nullptr;
    float* d_normals = nullptr;     // 3 floats per vertex, or nullptr
    float* d_texcoords = nullptr;   // 2 floats per vertex, or nullptr
    int64_t node_count = 0;
    int32_t bvh_depth = 0;
    bool bvh_ready = false;
    void release();
    SceneDevice device() const
    {
        return SceneDevice{d_nodes, d_tris, d_prim_ids, tri_count, d_idx, d_normals, d_texcoords};
    }
};

// One job of a batch: a scene, a camera, an image and a sample count.
struct BatchJobDevice {
    Camera cam;
    int32_t scene, width, height, tiles_x;
    uint32_t entries;       // 8x8 tiles * 64 path slots per sample
    int32_t spp, pass0, max_depth;
    float albedo, inv_w, inv_h;
    uint32_t work_begin;    // first work item (sample-major, then path)
    uint64_t sample_base;   // float offset of the job's sample 0 in the sample planes
    uint64_t pixel_base;    // float4 offset of the job's accumulator
};

struct BatchLaunch {
    const BatchJobDevice* jobs;
    const uint32_t* work_begin;  // n_jobs + 1 prefix of work items
    int n_jobs;
    const SceneDevice* scenes;
    float* samples;              // path radiance per (job, sample, pixel)
    uint32_t total_work;
    uint32_t* work_counter;      // zero before launch
    unsigned long long* stats;   // [0..2] instrumented counters, [3] rays
};

struct BatchKernelConfig {
    int block, stack;
};
constexpr int kBatchVariantCount = 4;
extern const BatchKernelConfig kBatchVariants[kBatchVariantCount];

// Launch shape of the batch kernel: compiled variant (block size, LDS stack
// depth), resident workgroups per CU (paths in flight) and refill threshold.
struct BatchTuning {
    int variant = 0;
    int blocks_per_cu = 0;   // 0: the occupancy maximum
    int refill = 32;
};

struct LookState;
constexpr int kLookLevels = 2;  // frames past its own that a diffuse call may start

}  // namespace prt

struct prt_hip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int num_cus = 0;
    std::string device_name;

    // the scene of the plain (Device) path
    prt::SceneBuffers scene;

    // batched form: scenes, jobs and their buffers
    std::vector<prt::SceneBuffers> batch_scenes;
    std::vector<prt::BatchJobDevice> batch_jobs;    // host mirror (pass0 advanced per render)
    prt::SceneDevice* d_batch_scenes = nullptr;
    prt::BatchJobDevice* d_batch_jobs = nullptr;
    uint32_t* d_batch_work_begin = nullptr;          // n+1
    uint64_t* d_batch_pixel_begin = nullptr;         // n+1
    float* d_batch_samples = nullptr;
    float4* d_batch_accum = nullptr;
    uint64_t batch_samples_len = 0, batch_pixels = 0;
    uint32_t batch_total_work = 0;
    prt::BatchTuning batch_tuning;
    hipGraph_t batch_graph = nullptr;
    hipGraphExec_t batch_graph_exec = nullptr;
    uint64_t batch_graph_spill_gen = 0;  // spill_gen the graph was captured with (its kernel arguments)
    bool batch_use_graph = true;

    // traversal spill stack (global part of the short stack)
    uint2* d_spill = nullptr;
    int64_t spill_bytes = 0;
    int trace_grid = 0;        // persistent workgroups per trace launch
    int diffuse_grid = 0;      // persistent workgroups per diffuse render launch (for the current scene)
    int trace_per_cu = 0;      // occupancy limit of the persistent trace kernel, workgroups per CU
    int stack_entries = 0;     // max stack depth per lane (LDS + spill); only grows
    uint64_t spill_gen = 0;    // bumped whenever d_spill or stack_entries change

    // frame
    int width = 0, height = 0;
    float4* d_accum = nullptr;
    int32_t* d_bgr8 = nullptr;
    int pass = 0;
    prt_camera cam{};
    bool cam_ready = false;

    // wavefront queues
    int64_t queue_cap = 0;
    float4* d_rays[2] = {nullptr, nullptr};   // 2 float4 per ray
    float4* d_hits = nullptr;                 // 1 float4 per hit
    int32_t* d_pix[2] = {nullptr, nullptr};
    uint32_t* d_rng[2] = {nullptr, nullptr};
    uint32_t* d_counters = nullptr;           // [0..15] queue sizes, [16..31] fetch, [32] error
    void* h_readback = nullptr;               // pinned host copy of d_counters + d_stats
    unsigned long long* d_stats = nullptr;    // instrumented counters

    // host-array intersect staging
    float4* d_io_rays = nullptr;
    float4* d_io_hits = nullptr;
    int64_t io_cap = 0;

    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int profiling = 0;
    int pipeline = PRT_PIPELINE_FUSED;

    // per-triangle RGB albedo: owned copy or an external device pointer
    float* d_albedo_own = nullptr;
    const float* d_albedo = nullptr;
    float4* d_rgb_planes = nullptr;  // per-sample RGB planes of the albedo forward
    int64_t rgb_planes_cap = 0;      // ... in float4
    hipEvent_t kev[4 * 16] = {};    // per-launch timing events (trace/shade x begin/end)

    // cooperative tail of the fused diffuse frame (prt_hip_set_tail)
    int tail_coop = 1;

    // frame lookahead of the fused diffuse frame (prt_hip_set_lookahead)
    int lookahead = 1;
    uint64_t scene_gen = 0;             // bumped by every change of the plain scene
    prt::LookState* d_look = nullptr;
    uint2* d_look_side = nullptr;       // kLookLevels slots of one entry per pixel
    int64_t look_side_len = 0;
    uint4* d_look_q[prt::kLookLevels][2] = {};  // per level, read by this call / written for the next
    uint32_t look_q_cap = 0;
    uint32_t* d_look_flags = nullptr;   // one stop flag per block of the launch
    uint32_t look_gen = 0;
    struct LookKey {
        uint64_t scene_gen;
        prt_camera cam;
        int width, height, max_depth, pass, grid;
        uint32_t albedo_bits;
    } look_prev{};
    bool look_prev_on = false;    // look_prev holds the inputs of the last successful diffuse call
    bool look_prev_spec = false;  // ... and that call ran the lookahead (its state is meaningful)

    // device address of the library's fixed scratch array on this context's
    // GPU (prt_hip_intersect_closest_device_noalloc)
    void* d_scratch = nullptr;

    // BVH8 collapse (prt_hip_set_bvh_collapse)
    int bvh_collapse = PRT_COLLAPSE_SAH;
    float bvh_prim_cost = 1.0f;     // triangle test cost relative to a wide-node visit
};

namespace prt {

// bvh_build.hip
int build_bvh8(prt_hip_ctx* ctx, SceneBuffers& sb, prt_build_stats* stats);

// trace.hip
struct TraceLaunch {
    const float4* rays;
    float4* hits;               // closest-hit records
    float* any_t;               // any-hit query: distance of a blocking triangle, or -1
    const uint32_t* count_ptr;  // device count of rays (nullptr: use count)
    uint32_t count;
    uint32_t* fetch_ptr;        // device fetch counter, zero before launch
    unsigned long long* stats;  // nullptr unless instrumented
};
// Fused path launch (one frame, one sample per pixel)
// Frame lookahead of the fused diffuse renderer: device state that one render
// call hands to the next.  A call renders frame p; a path of frame p + l is
// "at level l".  Work items, in the order they are handed out:
//  - this frame's space: the items of frame p that earlier calls did not
//    start (S_0 of them were), then the paths that the previous call
//    suspended, level by level: R_0 of frame p, R_1 of frame p + 1, ...;
//  - once those are all handed out, the later frames' space: the items of
//    frame p + 1 not started yet, then those of p + 2, ... up to
//    p + kLookLevels (each level's items in path order, from S_l on).
// When the last path of frame p finishes, paths of later frames stop at their
// next bounce (or in the middle of a ray, which is traced again) and are
// queued for the next call, where level l becomes level l - 1; those that
// finished wait in a per-pixel side buffer, tagged with the call that adds
// them to the accumulator.
struct LookState {
    uint32_t counter;       // work counter over this frame's space
    uint32_t spec_counter;  // work counter over the later frames' space
    uint32_t target;        // items of this frame's space (0: the frame is finished)
    uint32_t live;          // waves that hold, or can still take, paths of this frame
    uint32_t started[kLookLevels + 1];     // S_l: items of frame p + l that earlier calls started (S_K = 0)
    uint32_t resume_in[kLookLevels + 1];   // R_l: paths of frame p + l in input queue l (R_K = 0)
    uint32_t spec_begin[kLookLevels + 2];  // level l >= 1 owns [spec_begin[l], spec_begin[l + 1]) of spec_counter
    uint32_t resume_out[kLookLevels + 1];  // paths of frame p + l (l >= 1) that this call suspended
    unsigned long long rays[kLookLevels + 1];  // rays of frame p + l (l >= 1) traced by this call and earlier ones
};
constexpr int kLookFlagStride = 16;  // u32 per block stop flag (one 64 B line each)

struct PathLaunch {
    Camera cam;
    int width, height, tiles_x;
    uint32_t entries;                // tiles * 64 path slots (partial tiles hold dead slots)
    int pass;
    float inv_w, inv_h;
    float4* accum;
    int max_depth;
    float albedo;
    uint32_t* path_counter;          // zero before launch
    unsigned long long* depth_rays;  // [16] rays traced per depth, zero before launch
    unsigned long long* stats;       // instrumented counters; [3] rays
    // per-triangle RGB albedo (RGB forward and adjoint modes), original triangle order
    const float* albedo_rgb;
    // adjoint: work items are (sample, path) for passes pass .. pass + spp - 1
    int spp;
    float inv_spp;
    const float* dldi;               // dLoss/dImage, 3 floats per pixel
    float* grad;                     // dLoss/dAlbedo, 3 floats per triangle (accumulated)
    // RGB forward: when set, work items are (sample, path) for passes pass ..
    // pass + spp - 1, and each path writes its RGB to plane sample (W*H
    // float4 each) instead of adding to the accumulator
    float4* rgb_planes;
    // the "Fast" renderers: frame of the geometric normal whatever attributes
    // the mesh has, and each pass overwrites the accumulator
    bool fast;
    // diffuse frames: a wave with at most 8 paths left traces them
    // cooperatively (0: never)
    int coop;
    // diffuse frames: frame lookahead (nullptr: off)
    LookState* look;
    uint2* side;          // kLookLevels slots of W*H: (sample bits, tag of the call that adds it)
    const uint4* q_in[kLookLevels];   // suspended paths of frame p + l, 3 uint4 each
    uint4* q_out[kLookLevels];        // suspended paths of frame p + l + 1 (the next call's level l)
    uint32_t q_cap;       // capacity of each queue in paths (>= lanes of the launch)
    uint32_t look_gen;    // this call's tag
    uint32_t* look_flags; // per block (kLookFlagStride apart): look_gen once this frame is finished
};
int ensure_trace_resources(prt_hip_ctx* ctx);
hipError_t launch_trace(prt_hip_ctx* ctx, const TraceLaunch& t, bool instrumented);
// closest-hit launch over `grid` persistent workgroups whose stack spill
// (stack_entries per lane) and error flag the caller provides
hipError_t launch_trace_scratch(prt_hip_ctx* ctx, const TraceLaunch& t, uint2* spill, int stack_entries,
                                uint32_t* err, int grid);
hipError_t launch_occluded(prt_hip_ctx* ctx, const TraceLaunch& t, bool instrumented);
// renderer: PRT_RENDERER_PRIMARY, _DIFFUSE or _DIFFUSE2
hipError_t launch_paths(prt_hip_ctx* ctx, const PathLaunch& p, int renderer, bool instrumented);
// frame lookahead, in one launch: clear the counters, roll the state to a new
// call (valid: the previous call rendered the previous pass with the same
// inputs and lookahead on) and add the side samples of the previous call to
// the accumulator
hipError_t launch_look_call_begin(prt_hip_ctx* ctx, bool valid, uint32_t entries);
hipError_t launch_paths_rgb(prt_hip_ctx* ctx, const PathLaunch& p, bool instrumented);
hipError_t launch_adjoint(prt_hip_ctx* ctx, const PathLaunch& p, bool instrumented);
// max_blocks_per_cu != nullptr: only query the variant's occupancy limit
hipError_t launch_batch(prt_hip_ctx* ctx, const BatchLaunch& bl, const BatchTuning& tu, bool instrumented,
                        int* max_blocks_per_cu);

// render.hip
hipError_t launch_zero_u32(uint32_t* p, int n, hipStream_t s);
hipError_t launch_raygen(prt_hip_ctx* ctx, int pass, float4* rays, int32_t* pix, uint32_t* rng);
hipError_t launch_frame_begin(prt_hip_ctx* ctx, uint32_t entries);
Camera to_camera(const prt_camera& c);
// fast: the "Fast" renderers (geometric-normal frame, accumulator overwritten)
hipError_t launch_shade_diffuse(prt_hip_ctx* ctx, int depth, int max_depth, float throughput, bool fast);
// wavefront frames deeper than 16: frees the counter slot of old_depth (its
// count moves to the per-depth statistics) for depth old_depth + 16
hipError_t launch_queue_recycle(prt_hip_ctx* ctx, int old_depth);
hipError_t launch_shade_primary(prt_hip_ctx* ctx, bool fast);
hipError_t launch_tonemap(prt_hip_ctx* ctx);
hipError_t launch_batch_begin(prt_hip_ctx* ctx);
hipError_t launch_accum_mean_rgb(prt_hip_ctx* ctx, float* image);
// adds planes 0..spp-1 (W*H float4 each) to the accumulator, in that order
hipError_t launch_rgb_planes_add(prt_hip_ctx* ctx, const float4* planes, int spp);
hipError_t launch_batch_finalize(prt_hip_ctx* ctx);
hipError_t launch_batch_advance(prt_hip_ctx* ctx);
hipError_t launch_unpack_rays(const float4* rays, const int32_t* pix, const uint32_t* rng,
                              int64_t n, int width, int height, prt_ray* out_rays,
                              uint32_t* out_states, hipStream_t s);

int set_error(prt_hip_ctx* ctx, const std::string& msg);
int check_hip(prt_hip_ctx* ctx, hipError_t e, const char* what);

}  // namespace prt

#define PRT_TRY(ctx, expr)                                              \
    do {                                                                \
        hipError_t _e = (expr);                                         \
        if (_e != hipSuccess) return prt::check_hip((ctx), _e, #expr);  \
    } while (0)
