// device_hip.cpp -- DeviceHip over the C ABI.  Mirrors DeviceCuda
// (render/device_cuda.cpp:40-228): the same props, stats keys and exceptions.
#include "device_hip.h"

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
    check(prt_hip_scene_upload(ctx_, indices_.data(), mesh.triangle_count(), positions_.data(),
                               mesh.vertex_count()),
          "initScene");
}

void DeviceHip::initRenderer(const Props& props, Props& stats)
{
    // acceleration structure: the project's own LBVH -> BVH8 ("optix" is the
    // reference's CUDA default and has no meaning here)
    const std::string accel = props.get("accel", "bvh8");
    if (accel != "bvh8" && accel != "lbvh") throw std::invalid_argument("invalid acceleration structure type");

    const std::string type = props.get<std::string>("type");
    if (type == "primaryHip" || type == "primary") renderer_ = PRT_RENDERER_PRIMARY;
    else if (type == "diffuseHip" || type == "diffuse") renderer_ = PRT_RENDERER_DIFFUSE;
    else if (type == "diffuse2Hip" || type == "diffuse2") renderer_ = PRT_RENDERER_DIFFUSE2;
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

Props DeviceHip::queryRay(const Ray& r)
{
    // RendererSingle::queryRay (render/renderer_single.h:137-163): primId,
    // dist, hit point and the ray-facing geometric normal
    prt_ray ray{{r.org.x, r.org.y, r.org.z}, 0.0f, {r.dir.x, r.dir.y, r.dir.z}, r.far};
    prt_hit hit{};
    check(prt_hip_intersect_closest(ctx_, &ray, &hit, 1), "queryRay");
    Props res;
    if (hit.t < 0.0f) return res;
    const int32_t* t = &indices_[3 * (size_t)hit.prim_id];
    float p[3][3];
    for (int v = 0; v < 3; ++v)
        for (int k = 0; k < 3; ++k) p[v][k] = positions_[3 * (size_t)t[v] + k];
    float e1[3], e2[3];
    for (int k = 0; k < 3; ++k) { e1[k] = p[1][k] - p[0][k]; e2[k] = p[2][k] - p[0][k]; }
    float n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const float len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (len > 0.0f) for (float& c : n) c /= len;
    if (n[0] * r.dir.x + n[1] * r.dir.y + n[2] * r.dir.z > 0.0f) for (float& c : n) c = -c;
    res.set("primId", (int64_t)hit.prim_id);
    res.set("dist", hit.t);
    res.set("p", Vec3f{r.org.x + hit.t * r.dir.x, r.org.y + hit.t * r.dir.y, r.org.z + hit.t * r.dir.z});
    res.set("Ng", Vec3f{n[0], n[1], n[2]});
    res.set("uv", Vec3f{hit.u, hit.v, 0.0f});
    return res;
}

Props DeviceHip::queryPixel(int x, int y)
{
    if (size_.x <= 0 || x < 0 || y < 0 || x >= size_.x || y >= size_.y) return Props();
    const float ix = ((float)x + 0.5f) / (float)size_.x, iy = ((float)y + 0.5f) / (float)size_.y;
    float d[3];
    for (int k = 0; k < 3; ++k) d[k] = camera_.image_o[k] + ix * camera_.image_dx[k] + iy * camera_.image_dy[k];
    const float inv = 1.0f / std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    Ray r;
    r.org = Vec3f{camera_.origin[0], camera_.origin[1], camera_.origin[2]};
    r.dir = Vec3f{d[0] * inv, d[1] * inv, d[2] * inv};
    Props res = queryRay(r);
    if (!res.empty()) res.set("depth", res.get<float>("dist"));
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
