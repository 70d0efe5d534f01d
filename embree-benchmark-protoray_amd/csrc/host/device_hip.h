// device_hip.h -- DeviceHip, the "-dev hip" Device (render/device.h:29-54)
// beside the reference's DeviceCpu / DeviceCuda (render/device_cuda.h).
// All GPU work goes through the C ABI of libprt_hip.so (include/prt_hip.h).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/prt_hip.h"
#include "device.h"

namespace prt {

class DeviceHip : public Device {
public:
    explicit DeviceHip(int device_index = 0);  // throws std::runtime_error without a GPU
    ~DeviceHip() override;
    DeviceHip(const DeviceHip&) = delete;
    DeviceHip& operator=(const DeviceHip&) = delete;

    std::string getInfo() override;
    void initScene(const std::string& path, const Props& props) override;
    Box3f getSceneBounds() override { return bounds_; }
    void initRenderer(const Props& props, Props& stats) override;
    void render(Props& stats) override;
    Props queryPixel(int x, int y) override;
    Props queryRay(const Ray& ray) override;
    void initCamera(const Props& props) override;
    void initFrame(const Vec2i& size, const Props& props) override;
    void initToneMapper(const Props& props) override;
    void clearFrame() override;
    void blitFrame(Surface& dest) override;

    prt_hip_ctx* context() { return ctx_; }

private:
    void check(int rc, const char* what);
    Ray cameraRay(float ix, float iy) const;  // camera ray of image point (ix, iy), lens sample 0

    prt_hip_ctx* ctx_ = nullptr;
    Box3f bounds_;
    std::vector<int32_t> indices_;
    std::vector<float> positions_;
    std::vector<float> normals_;     // 3 per vertex, empty without normals
    std::vector<float> texcoords_;   // 2 per vertex, empty without texcoords
    std::vector<int32_t> material_ids_;
    std::vector<std::string> material_names_;
    int renderer_ = PRT_RENDERER_DIFFUSE;
    int max_depth_ = 6;
    float albedo_ = 0.8f;
    Vec2i size_;
    prt_camera camera_{};
};

}  // namespace prt
