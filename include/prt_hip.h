/*
 * prt_hip.h -- C ABI of libprt_hip.so, the MI355X (gfx950) wavefront diffuse
 * path tracer that sits behind ProtoRay's Device API as "-dev hip".
 *
 * Conventions
 *   - Every function returns int status: 0 = ok, nonzero = error.  The text of
 *     the last error of a context is prt_hip_last_error(ctx).  No C++
 *     exception crosses this boundary.
 *   - The caller owns host arrays; the context owns every device allocation it
 *     makes and frees them in prt_hip_destroy.  One HIP stream per context.
 *     A context is not thread-safe; several contexts may coexist.
 *   - Pointers named d_* are device pointers (HBM), everything else is host.
 *
 * Reference interfaces replaced (paths are in embree-benchmark-protoray):
 *   render/device.h:29-54                     class Device (plugin point)
 *   render/device_cuda.cpp:121-168            DeviceCuda::initScene (mesh upload)
 *   accel/optix/optix_intersector_stream_cuda.cpp:23-75   BVH build (+ buildMs/buildMprim)
 *   core/intersector_stream_cuda.h:27-34      IntersectorStreamCuda::intersect
 *   render/device_cuda.cpp:175-200            DeviceCuda::initCamera
 *   image/frame_buffer_cuda.cpp:22-50         FrameBufferCuda (init/clear/map)
 *   render/diffuse_renderer_cuda.cu:159-212   DiffuseRendererCuda::render
 *   render/primary_renderer_cuda.cu:124-153   PrimaryRendererCuda::render
 *   image/accum_buffer_cuda.cu:22-44          accumBufferUpdateKernel (tonemap)
 */
#ifndef PRT_HIP_H
#define PRT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PRT_HIP_ABI_VERSION 4

typedef struct prt_hip_ctx prt_hip_ctx;

/* Ray record, 32 B.  Same field order as RayCuda (core/ray_cuda.cuh:46-52).
 * A hit is accepted for tnear <= t <= tfar.  tfar < 0 marks a dead ray that
 * is not traced (its hit reports a miss). */
typedef struct prt_ray {
    float org[3];
    float tnear;
    float dir[3];
    float tfar;
} prt_ray;

/* Hit record, 16 B.  Same layout and miss encoding as HitCuda
 * (core/ray_cuda.cuh:23-34): miss <=> t < 0 (t = -1, prim_id = -1).
 * u, v are barycentrics with p = (1-u-v) p0 + u p1 + v p2
 * (geometry/triangle_mesh.cpp:69-71). */
typedef struct prt_hit {
    float t;
    int32_t prim_id;
    float u;
    float v;
} prt_hit;

/* Camera, the data of ThinLensCameraCuda (camera/thin_lens_camera_cuda.h:23-30).
 * lens_radius == 0 is the pinhole camera (camera/pinhole_camera_cuda.cuh:24-28);
 * lens_radius > 0 is the thin lens (camera/thin_lens_camera_cuda.cuh:25-31):
 * the origin moves to a uniform disk sample of that radius in the basis_u /
 * basis_v plane and the ray passes through the point of the pinhole ray at
 * focal_distance. */
typedef struct prt_camera {
    float origin[3];
    float image_o[3];
    float image_dx[3];
    float image_dy[3];
    float basis_u[3];
    float basis_v[3];
    float lens_radius;
    float focal_distance;
} prt_camera;

/* Stats of prt_hip_bvh_build (the buildMs / buildMprim keys of
 * accel/optix/optix_intersector_stream_cuda.cpp:61-70). */
typedef struct prt_build_stats {
    double build_ms;      /* device time of one build                      */
    double build_mprim;   /* triangles / build time, millions per second   */
    int64_t node_count;   /* BVH8 nodes                                    */
    int64_t tri_count;    /* triangle records (== triangles)               */
    int32_t depth;        /* BVH8 levels                                   */
    int32_t reserved;
} prt_build_stats;

/* Stats of one prt_hip_render call: the mray / ray keys of
 * render/device_cuda.cpp:99-107.  rays counts every ray handed to the
 * closest-hit query, primary and secondary (render/diffuse_renderer_cuda.cu:186). */
typedef struct prt_render_stats {
    int64_t rays;
    double ms;
    double mray;
    int64_t paths;        /* live primary rays = W*H                       */
    int64_t queue[16];    /* rays traced at depth k (k <= max_depth), the
                             last entry summing depths >= 15; with the fused
                             pipeline only when PRT_PROFILE_COUNTERS is on,
                             else 0                                         */
    /* filled only when prt_hip_set_profiling enables them:                  */
    double trace_ms;      /* summed device time of the closest-hit launches */
    double shade_ms;      /* summed device time of the shading launches     */
    int64_t nodes;        /* BVH8 nodes visited, all rays of the frame      */
    int64_t tris;         /* triangle tests, all rays of the frame          */
    /* fused pipeline with PRT_PROFILE_COUNTERS: SIMT occupancy of the path
     * kernel's traversal loop                                               */
    int64_t spills;       /* stack entries moved from LDS to global memory  */
    int64_t wave_steps;   /* traversal-loop iterations, counted per wave    */
    int64_t lane_steps;   /* traversal-loop iterations, counted per lane    */
    int64_t wave_shades;  /* shading passes, counted per wave               */
    int64_t drain_wave_steps;  /* wave_steps / lane_steps taken after the   */
    int64_t drain_lane_steps;  /* frame's work counter ran out              */
    double drain_ms;      /* from the first wave finding the work counter
                             run out to the last wave's exit (wall clock)   */
} prt_render_stats;

/* Per-ray traversal counters (the PROFILE_MODE nodes/prims of
 * render/renderer_single.h:93-96), filled by the instrumented replay. */
typedef struct prt_trace_counters {
    int64_t rays;
    int64_t nodes;        /* BVH8 nodes visited                             */
    int64_t tris;         /* triangle tests                                 */
    int64_t spills;       /* short-stack entries moved to the global spill  */
} prt_trace_counters;

enum {
    PRT_RENDERER_PRIMARY = 0,   /* "primaryHip": (N+1)/2 on hit, 0.05 on miss */
    PRT_RENDERER_DIFFUSE = 1,   /* "diffuseHip": diffuse path tracing         */
    PRT_RENDERER_DIFFUSE2 = 2,  /* "diffuse2Hip": with next-event shadow rays  */
    /* modifier bit of the "Fast" renderers ("primaryFast", "diffuseFast",
     * "diffuse2Fast", render/device_cuda.cpp:79-90): SimpleShadingContext
     * (frame of the geometric normal, attributes ignored) and each pass
     * overwrites the accumulator instead of adding to it (isAccum = false,
     * render/diffuse_renderer_cuda.cu:109-118) */
    PRT_RENDERER_FAST = 0x100
};

/* --- context ------------------------------------------------------------ */
int         prt_hip_abi_version(void);
int         prt_hip_create(int device_index, prt_hip_ctx** out_ctx);
int         prt_hip_destroy(prt_hip_ctx* ctx);
const char* prt_hip_last_error(prt_hip_ctx* ctx);
/* Device name (DeviceCuda::getInfo, render/device_cuda.cpp:40-48). */
int         prt_hip_device_name(prt_hip_ctx* ctx, char* buf, int buf_len);

/* --- scene (DeviceCuda::initScene) -------------------------------------- */
/* idx3: 3*T int32 vertex indices; pos3: 3*V float positions (float3 repack of
 * the strided .mesh vertex attributes, render/device_cuda.cpp:135-142). */
int prt_hip_scene_upload(prt_hip_ctx* ctx, const int32_t* idx3, int64_t tri_count,
                         const float* pos3, int64_t vertex_count);
/* Per-vertex shading attributes of the uploaded scene (the optional normals
 * and texcoords that DeviceCuda::initScene uploads, render/device_cuda.cpp:
 * 144-160): normals3 = 3 floats per vertex, texcoords2 = 2 floats per vertex,
 * in the vertex order of the upload; NULL leaves the mesh without that
 * attribute.  prt_hip_scene_upload removes both.  With them the primary,
 * diffuse and diffuse2 renderers shade with the full ShadingContext
 * (geometry/triangle_mesh_cuda.cuh:41-130): interpolated vertex normal and a
 * frame aligned with dp/du; the PRT_RENDERER_FAST renderers ignore them. */
int prt_hip_scene_attributes(prt_hip_ctx* ctx, const float* normals3, const float* texcoords2);

/* --- acceleration structure (OptixIntersectorStreamCuda ctor) ------------ */
/* LBVH on the GPU: Morton codes, LSD radix sort, Karras radix tree, bottom-up
 * boxes, collapse to a BVH8 with 8-bit quantised child boxes. */
int prt_hip_bvh_build(prt_hip_ctx* ctx, prt_build_stats* out_stats);

/* Collapse of the binary LBVH into the BVH8 (no counterpart in the reference,
 * whose build is inside OptiX).  PRT_COLLAPSE_SAH (default): the slots of each
 * wide node follow the least surface-area cost, computed bottom-up during the
 * refit (Ylitie, Karras, Laine 2017, section 4.1), with a triangle test
 * costing prim_cost wide-node visits.  PRT_COLLAPSE_GREEDY: the widest-area
 * internal child is opened until 8 slots are used.  Applies to later builds;
 * hits and images do not depend on it. */
#define PRT_COLLAPSE_GREEDY 0
#define PRT_COLLAPSE_SAH    1
int prt_hip_set_bvh_collapse(prt_hip_ctx* ctx, int method, float prim_cost);

/* --- closest hit (IntersectorStreamCuda::intersect) ---------------------- */
/* Host arrays; copies in, traces, copies out (parity entry). */
int prt_hip_intersect_closest(prt_hip_ctx* ctx, const prt_ray* rays, prt_hit* hits, int64_t n);
/* Device arrays, on the context's stream; asynchronous. */
int prt_hip_intersect_closest_device(prt_hip_ctx* ctx, const prt_ray* d_rays, prt_hit* d_hits,
                                     int64_t n);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  The same query and the same hit records as
 * prt_hip_intersect_closest_device, on the context's stream, asynchronous,
 * but it allocates nothing: the traversal's scratch (work counter, error
 * flag, the global part of each lane's stack) is a fixed device array of
 * PRT_HIP_NOALLOC_SCRATCH_BYTES that the library holds on each GPU it runs
 * on until the process ends.  The launch keeps as many persistent
 * workgroups as that array holds worst-case stacks for (BVH depth D needs
 * 256 * (7 (D + 1) + 8) * 8 bytes per workgroup), and a BVH too deep for one
 * workgroup is rejected with an error.  n <= 2^31 - 1 rays.  Every call on
 * one GPU shares that GPU's array, whatever the context: calls that overlap
 * in time on one GPU (other streams, other threads) race, so the caller
 * orders them. */
#define PRT_HIP_NOALLOC_SCRATCH_BYTES (128u << 20)
int prt_hip_intersect_closest_device_noalloc(prt_hip_ctx* ctx, const prt_ray* d_rays, prt_hit* d_hits,
                                             int64_t n);
/* Same query on the instrumented kernel; adds per-ray node/triangle counts. */
int prt_hip_intersect_counters(prt_hip_ctx* ctx, const prt_ray* rays, prt_hit* hits, int64_t n,
                               prt_trace_counters* out_counters);

/* --- any hit (IntersectorStreamCuda::occluded, core/intersector_stream_cuda.h:33;
 * OptixIntersectorStreamCuda::occluded, accel/optix/optix_intersector_stream_cuda.cpp:84-89,
 * hit format HIT_T = AnyHitCuda{float t}, core/ray_cuda.cuh:36-44) ----------
 * out_t[i] >= 0 (the distance of some blocking triangle in [tnear, tfar]) if
 * ray i is occluded, -1 otherwise; a ray with tfar < 0 gives -1.  Which
 * blocking triangle's distance is reported is unspecified; only the sign is
 * part of the contract, as with the reference's isHit().  Host arrays. */
int prt_hip_occluded(prt_hip_ctx* ctx, const prt_ray* rays, float* out_t, int64_t n);
/* Device arrays, on the context's stream; asynchronous. */
int prt_hip_occluded_device(prt_hip_ctx* ctx, const prt_ray* d_rays, float* d_out_t, int64_t n);

/* --- camera / frame (DeviceCuda::initCamera / initFrame / clearFrame) ---- */
/* Pinhole camera of a view: RenderWindow::makeCamera
 * (main/render/render_window.cpp:307-325: basis = Basis3(one).rotateV(angleY)
 * .rotateU(angleX)) followed by the PinholeCamera constructor
 * (camera/pinhole_camera.h:34-48).  Host-only; needs no context. */
int prt_hip_camera_from_view(const float position[3], float angle_x, float angle_y, float fov_y,
                             float aspect_ratio, prt_camera* out_camera);
/* The two halves of prt_hip_camera_from_view: the view basis U, V, N
 * (9 floats) of RenderWindow::makeCamera, and the PinholeCamera of a
 * position + basis (camera/pinhole_camera.h:34-48).  Host-only. */
int prt_hip_view_basis(float angle_x, float angle_y, float out_basis[9]);
int prt_hip_camera_from_basis(const float position[3], const float basis[9], float fov_y,
                              float aspect_ratio, prt_camera* out_camera);
int prt_hip_camera_set(prt_hip_ctx* ctx, const prt_camera* camera);
int prt_hip_frame_init(prt_hip_ctx* ctx, int width, int height);
int prt_hip_frame_clear(prt_hip_ctx* ctx);    /* also resets the pass counter */

/* --- rendering (DiffuseRendererCuda::render / PrimaryRendererCuda::render) */
/* One sample per pixel per call, pass = number of calls since the last clear
 * (render/diffuse_renderer_cuda.cu:210).  albedo is the grey reflectance R
 * (the reference hard-codes 0.8, render/diffuse_integrator_single.h:43).
 * max_depth in [0, 1024] (the reference takes any maxDepth prop,
 * render/device_cuda.cpp:68; a throughput of R^k below 1 is 0 in fp32 long
 * before depth 1024); the wavefront pipeline stops early once no path is left,
 * as the reference's while (count > 0). */
int prt_hip_render(prt_hip_ctx* ctx, int renderer, int max_depth, float albedo,
                   prt_render_stats* out_stats);
/* Sets the pass index the next prt_hip_render call uses. */
int prt_hip_set_pass(prt_hip_ctx* ctx, int pass);
/* Profiling of prt_hip_render: PRT_PROFILE_TIMING records HIP events around
 * every trace and shade launch (trace_ms, shade_ms); PRT_PROFILE_COUNTERS runs
 * the instrumented traversal (nodes, tris).  0 turns both off. */
#define PRT_PROFILE_TIMING   1
#define PRT_PROFILE_COUNTERS 2
int prt_hip_set_profiling(prt_hip_ctx* ctx, int flags);
/* Frame pipeline of prt_hip_render.  PRT_PIPELINE_FUSED (default): one
 * persistent launch per frame in which every lane generates, traces, shades
 * and bounces whole paths.  PRT_PIPELINE_WAVEFRONT: the reference's structure
 * (render/diffuse_renderer_cuda.cu:159-212), raygen then one closest-hit and
 * one shading launch per depth with compacted ray queues in HBM (queue sizes
 * stay on the device).  Both give bit-identical images. */
#define PRT_PIPELINE_FUSED     0
#define PRT_PIPELINE_WAVEFRONT 1
int prt_hip_set_pipeline(prt_hip_ctx* ctx, int pipeline);
/* Tail of the fused diffuse frame (no counterpart in the reference): once the
 * frame's paths are all handed out and a wave holds at most 8 of them,
 * PRT_TAIL_COOPERATIVE (default) traces each on 8 lanes (one BVH8 child, or
 * one of up to 8 pending triangles, per lane and step); PRT_TAIL_LANES keeps
 * one lane per path.  Images do not change. */
#define PRT_TAIL_LANES       0
#define PRT_TAIL_COOPERATIVE 1
int prt_hip_set_tail(prt_hip_ctx* ctx, int mode);
/* Frame lookahead of the fused diffuse renderer (no counterpart in the
 * reference; default on).  In a call that continues the previous call (pass
 * p + 1 after pass p, with the same scene, camera, frame size, max depth and
 * albedo), idle lanes start paths of the next two passes while the last
 * paths of this one run; the call returns when its own frame is complete, and the
 * next call keeps that work if it continues this one (else drops it).  The
 * first call after any change renders as without the lookahead.  The
 * accumulator after each call and its ray count are those of rendering the
 * frames one by one.  Off for the primary, diffuse2 and Fast renderers, the
 * wavefront pipeline and under PRT_PROFILE_COUNTERS. */
int prt_hip_set_lookahead(prt_hip_ctx* ctx, int enable);

/* --- frame buffer read-back -------------------------------------------- */
/* Raw accumulation buffer, float4 (r, g, b, weight) per pixel, row-major. */
int prt_hip_accum_read(prt_hip_ctx* ctx, float* dst_rgba);
/* Tonemapped 32-bit pixels b | g<<8 | r<<16 (accumBufferUpdateKernel). */
int prt_hip_frame_read_bgr8(prt_hip_ctx* ctx, int32_t* dst);

/* --- per-triangle RGB albedo and its adjoint (no counterpart in the
 * reference, which hard-codes R = 0.8, render/diffuse_integrator_single.h:43) */
/* Per-triangle RGB albedo, 3 floats per triangle in the order of the uploaded
 * indices.  _host copies it; _device uses the caller's device array (not
 * copied; it must stay alive and unchanged while the context renders). */
int prt_hip_set_albedo_host(prt_hip_ctx* ctx, const float* rgb);
int prt_hip_set_albedo_device(prt_hip_ctx* ctx, const float* d_rgb);
/* Mean RGB image over passes pass0 .. pass0+spp-1 with the per-triangle
 * albedo: a path that escapes after hitting t_0..t_{K-1} carries
 * prod_k albedo(t_k) per channel, one hitting at max_depth carries 0.
 * image: 3 floats per pixel, row-major.  Resets the accumulation buffer.
 * The samples run in one launch per chunk, through per-sample RGB planes
 * that the context keeps (W*H*16 bytes per sample, at most 1 GiB); the
 * result equals one launch per pass bit for bit. */
int prt_hip_render_albedo_image(prt_hip_ctx* ctx, int pass0, int spp, int max_depth, float* d_image,
                                prt_render_stats* out_stats);
int prt_hip_render_albedo_image_host(prt_hip_ctx* ctx, int pass0, int spp, int max_depth, float* image,
                                     prt_render_stats* out_stats);
/* Path-replay adjoint of prt_hip_render_albedo_image: grad += dLoss/dalbedo
 * for dLoss/dImage = dldi (3 floats per pixel).  Each path is traced once to
 * get its per-channel albedo product, then re-traced from its sampler seed
 * (pass, pixelId) -- nothing about the path is stored -- and every hit
 * atomically adds dldi * (product of the path's other albedos) / spp.
 * rays in out_stats counts both traces. */
int prt_hip_adjoint_albedo(prt_hip_ctx* ctx, int pass0, int spp, int max_depth, const float* d_dldi,
                           float* d_grad, prt_render_stats* out_stats);
int prt_hip_adjoint_albedo_host(prt_hip_ctx* ctx, int pass0, int spp, int max_depth, const float* dldi,
                                float* grad, prt_render_stats* out_stats);

/* --- batched form (no counterpart in the reference) ---------------------- */
/* Many scenes, cameras and sample counts of uneven size in ONE persistent
 * launch: a work item is (job, sample, path); waves steal work items from one
 * counter across all jobs.  Each path writes its radiance to a per-sample
 * plane and a finalize kernel adds the samples to the job's accumulator in
 * pass order, so a job's image is bit-identical to rendering it with
 * prt_hip_render one sample at a time.  The launch sequence (reset, batch
 * kernel, finalize, pass advance) is captured once as a hipGraph and replayed
 * by every prt_hip_batch_render; the bounce loop itself runs on the device
 * inside the batch kernel.  The kernel's block size, LDS stack depth and
 * in-flight path count come from prt_hip_batch_autotune. */
typedef struct prt_batch_job {
    int32_t scene;        /* index returned by prt_hip_batch_scene_add          */
    int32_t width, height;
    int32_t spp;          /* samples per pixel per prt_hip_batch_render call    */
    int32_t max_depth;    /* 0..1024                                            */
    int32_t pass0;        /* sampler pass of the first sample                   */
    float albedo;
    prt_camera camera;    /* pinhole                                            */
} prt_batch_job;

typedef struct prt_batch_stats {
    int64_t rays;         /* rays traced by the whole batch                      */
    int64_t paths;        /* live paths = sum of spp * width * height            */
    double ms;            /* device time of the batch render (graph launch)      */
    double mray;
    int64_t nodes, tris;  /* BVH8 nodes / triangle tests with PRT_PROFILE_COUNTERS */
} prt_batch_stats;

typedef struct prt_batch_tuning {
    int32_t variant;       /* 0..3: (block, LDS stack) = (256,16) (256,8) (128,16) (128,8) */
    int32_t block, stack;  /* filled in from the variant                          */
    int32_t blocks_per_cu; /* resident workgroups per CU (paths in flight); 0 = occupancy maximum */
    int32_t refill;        /* idle lanes of a wave that trigger a refill (1..64)  */
    double ms;             /* autotuner: best measured time of this configuration */
} prt_batch_tuning;

int prt_hip_batch_scene_add(prt_hip_ctx* ctx, const int32_t* idx3, int64_t tri_count, const float* pos3,
                            int64_t vertex_count, prt_build_stats* out_stats, int32_t* out_scene);
/* Shading attributes of batch scene `scene`, as prt_hip_scene_attributes. */
int prt_hip_batch_scene_attributes(prt_hip_ctx* ctx, int32_t scene, const float* normals3,
                                   const float* texcoords2);
int prt_hip_batch_reset(prt_hip_ctx* ctx);
/* Allocates the accumulators (zeroed) and sample planes of the jobs. */
int prt_hip_batch_set_jobs(prt_hip_ctx* ctx, const prt_batch_job* jobs, int n_jobs);
int prt_hip_batch_render(prt_hip_ctx* ctx, prt_batch_stats* out_stats);
int prt_hip_batch_accum_read(prt_hip_ctx* ctx, int job, float* dst_rgba);
int prt_hip_batch_set_tuning(prt_hip_ctx* ctx, const prt_batch_tuning* tuning);
int prt_hip_batch_get_tuning(prt_hip_ctx* ctx, prt_batch_tuning* tuning);
/* Times every candidate launch shape on the current jobs (accumulators are
 * left untouched), keeps the fastest and reports all of them. */
int prt_hip_batch_autotune(prt_hip_ctx* ctx, prt_batch_tuning* results, int max_results, int* n_results);
/* 1 (default): replay a captured hipGraph; 0: plain stream launches. */
int prt_hip_batch_set_graph(prt_hip_ctx* ctx, int enable);

/* --- inspection (tests and the roofline model) -------------------------- */
/* Primary rays of a pass exactly as the diffuse renderer generates them,
 * one per pixel in row-major pixel order, plus their sampler states. */
int prt_hip_debug_raygen(prt_hip_ctx* ctx, int pass, prt_ray* rays, uint32_t* states);
/* Sizes of the built BVH8 (node and triangle-record arrays). */
int prt_hip_bvh_info(prt_hip_ctx* ctx, int64_t* node_count, int64_t* tri_count, int32_t* depth);
/* Copies the BVH8: nodes (80 B each), triangle records (9 floats each) and
 * the original triangle id of each record. */
int prt_hip_bvh_read(prt_hip_ctx* ctx, void* nodes, float* tri_records, int32_t* prim_ids);

#ifdef __cplusplus
}
#endif

#endif /* PRT_HIP_H */
