// mesh_io.h -- the .mesh blob (geometry/triangle_mesh.cpp:305-337,
// sys/blob.h:26-50, sys/array.h:357-393, sys/string.cpp:22-45) and the
// .view set (main/render/view.h:25-40, main/render/view.cpp:23-46).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "props.h"

namespace prt {

constexpr int kMeshBlobId = 0x04df8666;

struct TriangleMesh {
    Box3f bounds;
    int vertex_stride = 3;
    bool has_normals = false, has_texcoords = false;
    std::vector<int32_t> indices;       // 3 per triangle
    std::vector<int32_t> material_ids;  // 1 per triangle
    std::vector<float> vertex_attribs;  // vertex_stride per vertex
    std::vector<std::string> material_names;

    int64_t triangle_count() const { return (int64_t)material_ids.size(); }
    int64_t vertex_count() const { return (int64_t)(vertex_attribs.size() / vertex_stride); }
    // positions repacked to 3 floats per vertex (render/device_cuda.cpp:135-142)
    std::vector<float> positions() const;
};

// Throws std::runtime_error("file has wrong blob ID: ...") or
// ("unexpected end of stream") like loadBlob / Stream::readFull.
TriangleMesh load_mesh(const std::string& path);
void save_mesh(const std::string& path, const TriangleMesh& mesh);

struct View {
    Vec3f pos;
    float angle_x = 0, angle_y = 0, fov_y = 0, radius = 0, focus = 1;
};
constexpr int kViewSetSize = 10;

// Returns false if the file does not exist (silent, like loadViewSet).
bool load_view_set(const std::string& path, View (&views)[kViewSetSize]);
void save_view_set(const std::string& path, const View (&views)[kViewSetSize]);

}  // namespace prt
