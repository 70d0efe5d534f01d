// device.h -- ProtoRay's Device plugin interface (render/device.h:29-54),
// restated with this build's Props.  DeviceHip implements it over the C ABI.
#pragma once

#include <string>

#include "props.h"

namespace prt {

// RGBA8 surface of image/surface.h:24-35
struct Surface {
    void* data = nullptr;
    int width = 0;
    int height = 0;
    int pitch = 0;  // bytes
};

// Ray of core/ray.h:26-64 (far = hit distance, +inf-ish when missed)
struct Ray {
    Vec3f org, dir;
    float far = 3.402823466e38f;
};

class Device {
public:
    virtual ~Device() {}

    virtual std::string getInfo() = 0;

    // Scene
    virtual void initScene(const std::string& path, const Props& props) = 0;
    virtual Box3f getSceneBounds() = 0;

    // Renderer (initialize the scene first)
    virtual void initRenderer(const Props& props, Props& stats) = 0;
    virtual void render(Props& stats) = 0;
    virtual Props queryPixel(int x, int y) = 0;
    virtual Props queryRay(const Ray& ray) = 0;

    // Camera
    virtual void initCamera(const Props& props) = 0;

    // Framebuffer
    virtual void initFrame(const Vec2i& size, const Props& props) = 0;
    virtual void initToneMapper(const Props& props) = 0;
    virtual void clearFrame() = 0;
    virtual void blitFrame(Surface& dest) = 0;
};

}  // namespace prt
