// device_hip.cpp -- DeviceHip over the C ABI.  Mirrors DeviceCuda
// (render/device_cuda.cpp:40-228): the same props, stats keys and exceptions.
#include "device_hip.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "mesh_io.h"

namespace prt {

DeviceHip::DeviceHip(int device_index)
{
    const int rc = prt_hip_create(device_index, &ctx_);
    if (rc != 0) throw std::runtime_error("HIP device " + std::to_string(device_index) + " unavailable (status " +
                                          std::to_string(rc) + ")");
}

DeviceHip::~DeviceHip() { prt_hip_destroy(ctx_); }

void DeviceHip::check(int rc, const char* what)
{
    if (rc != 0) throw std::runtime_error(std::string(what) + ": " + prt_hip_last_error(ctx_));
}

std::string DeviceHip::getInfo()
{
    char name[256];
    check(prt_hip_device_name(ctx_, name, sizeof(name)), "getInfo");
    return name;
}

void DeviceHip::initScene(const std::string& path, const Props&)
{
    TriangleMesh mesh = load_mesh(path);
    indices_ = mesh.indices;
    positions_ = mesh.positions();
    bounds_ = mesh.bounds;
    material_ids_ = mesh.material_ids;
    material_names_ = mesh.material_names;
    check(prt_hip_scene_upload(ctx_, indices_.data(), mesh.triangle_count(), positions_.data(),
                               mesh.vertex_count()),
          "initScene");
    // normals follow the position, texture coordinates follow both
    // (geometry/triangle_mesh.cpp:28-56); repacked as DeviceCuda::initScene
    // does (render/device_cuda.cpp:144-160)
    normals_.clear();
    texcoords_.clear();
    if (mesh.has_normals || mesh.has_texcoords) {
        const int64_t nv = mesh.vertex_count();
        const int stride = mesh.vertex_stride;
        std::vector<float>& normals = normals_;
        std::vector<float>& texcoords = texcoords_;
        if (mesh.has_normals) {
            normals.resize(3 * (size_t)nv);
            for (int64_t i = 0; i < nv; ++i)
                for (int k = 0; k < 3; ++k) normals[3 * i + k] = mesh.vertex_attribs[(size_t)i * stride + 3 + k];
        }
        if (mesh.has_texcoords) {
            const int off = mesh.has_normals ? 6 : 3;
            texcoords.resize(2 * (size_t)nv);
            for (int64_t i = 0; i < nv; ++i)
                for (int k = 0; k < 2; ++k)
                    texcoords[2 * i + k] = mesh.vertex_attribs[(size_t)i * stride + off + k];
        }
        check(prt_hip_scene_attributes(ctx_, mesh.has_normals ? normals.data() : nullptr,
                                       mesh.has_texcoords ? texcoords.data() : nullptr),
              "initScene");
    }
}

void DeviceHip::initRenderer(const Props& props, Props& stats)
{
    // acceleration structure: the project's own LBVH -> BVH8 ("optix" is the
    // reference's CUDA default and has no meaning here)
    const std::string accel = props.get("accel", "bvh8");
    if (accel != "bvh8" && accel != "lbvh") throw std::invalid_argument("invalid acceleration structure type");

    const std::string type = props.get<std::string>("type");
    // the reference's CUDA names (render/device_cuda.cpp:79-92) and *Hip
    if (type == "primaryHip" || type == "primary") renderer_ = PRT_RENDERER_PRIMARY;
    else if (type == "diffuseHip" || type == "diffuse") renderer_ = PRT_RENDERER_DIFFUSE;
    else if (type == "diffuse2Hip" || type == "diffuse2") renderer_ = PRT_RENDERER_DIFFUSE2;
    else if (type == "primaryFastHip" || type == "primaryFast") renderer_ = PRT_RENDERER_PRIMARY | PRT_RENDERER_FAST;
    else if (type == "diffuseFastHip" || type == "diffuseFast") renderer_ = PRT_RENDERER_DIFFUSE | PRT_RENDERER_FAST;
    else if (type == "diffuse2FastHip" || type == "diffuse2Fast")
        renderer_ = PRT_RENDERER_DIFFUSE2 | PRT_RENDERER_FAST;
    else throw std::invalid_argument("invalid renderer type");

    std::string sampler = props.get("sampler", "default");
    if (sampler == "default") sampler = "random";
    if (sampler != "random") throw std::invalid_argument("invalid sampler type");

    max_depth_ = props.get("maxDepth", 6);
    albedo_ = props.get("albedo", 0.8f);

    // build timing as accel/optix/optix_intersector_stream_cuda.cpp:25-70
    const bool bench = props.exists("benchmark");
    const int build_count = props.get("buildCount", bench ? 12 : 1);
    const int build_warmup = props.get("buildWarmup", build_count / 6);
    double sum_ms = 0.0;
    for (int i = 0; i < build_count; ++i) {
        prt_build_stats bs{};
        check(prt_hip_bvh_build(ctx_, &bs), "initRenderer");
        if (build_count == 1 || i >= build_warmup) sum_ms += bs.build_ms;
    }
    const int counted = build_count == 1 ? 1 : build_count - build_warmup;
    const double ms = counted > 0 ? sum_ms / counted : 0.0;
    stats.set("buildMs", ms);
    stats.set("buildMprim", ms > 0.0 ? (double)(indices_.size() / 3) / 1000.0 / ms : 0.0);
}

void DeviceHip::render(Props& stats)
{
    prt_render_stats rs{};
    check(prt_hip_render(ctx_, renderer_, max_depth_, albedo_, &rs), "render");
    stats.set("mray", rs.mray);
    stats.set("ray", (int64_t)rs.rays);
}

namespace {

struct V3f {
    float x, y, z;
};
V3f sub(V3f a, V3f b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3f crs(V3f a, V3f b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
float dt(V3f a, V3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3f nrm(V3f a)
{
    const float l = std::sqrt(dt(a, a));
    return l > 0.0f ? V3f{a.x / l, a.y / l, a.z / l} : a;
}
V3f ng3(V3f a) { return {-a.x, -a.y, -a.z}; }
V3f at3(const std::vector<float>& v, int i) { return {v[3 * (size_t)i], v[3 * (size_t)i + 1], v[3 * (size_t)i + 2]}; }

}  // namespace

Props DeviceHip::queryRay(const Ray& r)
{
    // RendererSingle::queryRay (render/renderer_single.h:137-163): closest hit
    // on the GPU, then the ShadingContext of TriangleMesh::postIntersect
    // (geometry/triangle_mesh.cpp:59-153) on the host: material, primId,
    // distance, hit point, ray-facing Ng, shading normal N, uv and eps
    prt_ray ray{{r.org.x, r.org.y, r.org.z}, 0.0f, {r.dir.x, r.dir.y, r.dir.z}, r.far};
    prt_hit hit{};
    check(prt_hip_intersect_closest(ctx_, &ray, &hit, 1), "queryRay");
    Props res;
    if (hit.t < 0.0f) return res;
    const int32_t* t = &indices_[3 * (size_t)hit.prim_id];
    const V3f p0 = at3(positions_, t[0]), p1 = at3(positions_, t[1]), p2 = at3(positions_, t[2]);
    const V3f dir{r.dir.x, r.dir.y, r.dir.z};
    const float b1 = hit.u, b2 = hit.v, b0 = 1.0f - b1 - b2;
    V3f ng = nrm(crs(sub(p1, p0), sub(p2, p0)));
    V3f n = ng;
    if (!normals_.empty()) {
        const V3f n0 = at3(normals_, t[0]), n1 = at3(normals_, t[1]), n2 = at3(normals_, t[2]);
        n = nrm({b0 * n0.x + b1 * n1.x + b2 * n2.x, b0 * n0.y + b1 * n1.y + b2 * n2.y,
                 b0 * n0.z + b1 * n1.z + b2 * n2.z});
        if (dt(n, ng) < 0.0f) n = ng3(n);
    }
    if (dt(ng, dir) > 0.0f) {
        ng = ng3(ng);
        n = ng3(n);
    }
    // ctx.uv is a Vec2f: the interpolated texture coordinates, or the hit's
    // (u, v) when the mesh has none (geometry/triangle_mesh.cpp:115,150)
    Vec2f uv{hit.u, hit.v};
    if (!texcoords_.empty()) {
        const float* a = &texcoords_[2 * (size_t)t[0]];
        const float* b = &texcoords_[2 * (size_t)t[1]];
        const float* c = &texcoords_[2 * (size_t)t[2]];
        uv = Vec2f{b0 * a[0] + b1 * b[0] + b2 * c[0], b0 * a[1] + b1 * b[1] + b2 * c[1]};
    }
    const Vec3f p{r.org.x + hit.t * r.dir.x, r.org.y + hit.t * r.dir.y, r.org.z + hit.t * r.dir.z};
    const float eps = std::max(hit.t, std::max(std::fabs(p.x), std::max(std::fabs(p.y), std::fabs(p.z)))) *
                      0x1.fp-18f;
    const int mat = hit.prim_id < (int32_t)material_ids_.size() ? material_ids_[hit.prim_id] : 0;
    res.set("mat", mat >= 0 && mat < (int)material_names_.size() ? material_names_[mat] : std::string());
    res.set("matId", mat);
    res.set("primId", (int)hit.prim_id);
    res.set("dist", hit.t);
    res.set("p", p);
    res.set("Ng", Vec3f{ng.x, ng.y, ng.z});
    res.set("N", Vec3f{n.x, n.y, n.z});
    res.set("uv", uv);
    res.set("eps", eps);
    return res;
}

Ray DeviceHip::cameraRay(float ix, float iy) const
{
    // Camera::getRay with the lens sample 0 (render/renderer.h:59-71):
    // PinholeCamera::getRay (camera/pinhole_camera.h:50-54), or for a lens
    // radius > 0 ThinLensCamera::getRay (camera/thin_lens_camera.h:38-44),
    // whose disk sample of 0 is the lens centre, so the ray starts at the
    // origin and passes through the focal-plane point of the image sample
    float d[3];
    for (int k = 0; k < 3; ++k) d[k] = camera_.image_o[k] + ix * camera_.image_dx[k] + iy * camera_.image_dy[k];
    if (camera_.lens_radius != 0.0f)
        for (int k = 0; k < 3; ++k)
            d[k] = (camera_.origin[k] + camera_.focal_distance * d[k]) - camera_.origin[k];
    const float inv = 1.0f / std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    Ray r;
    r.org = Vec3f{camera_.origin[0], camera_.origin[1], camera_.origin[2]};
    r.dir = Vec3f{d[0] * inv, d[1] * inv, d[2] * inv};
    return r;
}

Props DeviceHip::queryPixel(int x, int y)
{
    // Renderer::queryPixel (render/renderer.h:57-81): the ray through the
    // pixel centre; depth is the hit distance along the view direction (the
    // ray through the image centre)
    if (size_.x <= 0 || x < 0 || y < 0 || x >= size_.x || y >= size_.y) return Props();
    const Ray centre = cameraRay(0.5f, 0.5f);
    const Ray r = cameraRay(((float)x + 0.5f) / (float)size_.x, ((float)y + 0.5f) / (float)size_.y);
    Props res = queryRay(r);
    if (!res.empty()) {
        const float cosv = r.dir.x * centre.dir.x + r.dir.y * centre.dir.y + r.dir.z * centre.dir.z;
        res.set("depth", res.get<float>("dist") * cosv);
    }
    return res;
}

void DeviceHip::initCamera(const Props& props)
{
    // PinholeCamera (camera/pinhole_camera.h:34-48) plus the thin-lens
    // parameters of ThinLensCamera (camera/thin_lens_camera.h:31-35); a lens
    // radius of 0 renders as the pinhole, as DeviceCuda::initCamera does
    // (render/device_cuda.cpp:175-200)
    const float lens = props.get("lensRadius", 0.0f);
    const float focus = props.get("focalDistance", 1.0f);
    const Vec3f pos = props.get<Vec3f>("position");
    const Basis3f b = props.get<Basis3f>("basis");
    const float position[3] = {pos.x, pos.y, pos.z};
    const float basis[9] = {b.U.x, b.U.y, b.U.z, b.V.x, b.V.y, b.V.z, b.N.x, b.N.y, b.N.z};
    check(prt_hip_camera_from_basis(position, basis, props.get<float>("fov"), props.get<float>("aspectRatio"),
                                    &camera_),
          "initCamera");
    camera_.lens_radius = lens;
    camera_.focal_distance = focus;
    check(prt_hip_camera_set(ctx_, &camera_), "initCamera");
}

void DeviceHip::initFrame(const Vec2i& size, const Props&)
{
    check(prt_hip_frame_init(ctx_, size.x, size.y), "initFrame");
    size_ = size;
}

void DeviceHip::initToneMapper(const Props& props)
{
    // only "none" exists (render/device_cpu_impl.cpp:105-116)
    if (props.get("type", "none") != "none") throw std::invalid_argument("invalid tone mapper type");
}

void DeviceHip::clearFrame() { check(prt_hip_frame_clear(ctx_), "clearFrame"); }

void DeviceHip::blitFrame(Surface& dest)
{
    std::vector<int32_t> px((size_t)size_.x * size_.y);
    check(prt_hip_frame_read_bgr8(ctx_, px.data()), "blitFrame");
    for (int y = 0; y < dest.height && y < size_.y; ++y)
        std::memcpy((char*)dest.data + (ptrdiff_t)y * dest.pitch, px.data() + (size_t)y * size_.x,
                    sizeof(int32_t) * (size_t)std::min(dest.width, size_.x));
}

}  // namespace prt
