// mesh_io.cpp -- .mesh and .view readers/writers; see mesh_io.h.
#include "mesh_io.h"

#include <cstdio>
#include <cstring>
#include <stdexcept>

namespace prt {

namespace {

struct Reader {
    FILE* f;
    void read(void* dst, size_t n)
    {
        if (n && fread(dst, 1, n, f) != n) throw std::runtime_error("unexpected end of stream");
    }
    template <class T> T pod()
    {
        T v;
        read(&v, sizeof(T));
        return v;
    }
};

struct Writer {
    FILE* f;
    void write(const void* src, size_t n)
    {
        if (n && fwrite(src, 1, n, f) != n) throw std::runtime_error("write failed");
    }
    template <class T> void pod(const T& v) { write(&v, sizeof(T)); }
};

}  // namespace

std::vector<float> TriangleMesh::positions() const
{
    const int64_t n = vertex_count();
    std::vector<float> p(3 * (size_t)n);
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) p[3 * i + k] = vertex_attribs[(size_t)i * vertex_stride + k];
    return p;
}

TriangleMesh load_mesh(const std::string& path)
{
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("could not open file: " + path);
    TriangleMesh m;
    try {
        Reader r{f};
        if (r.pod<int32_t>() != kMeshBlobId) throw std::runtime_error("file has wrong blob ID: " + path);
        float b[6];
        r.read(b, sizeof(b));
        m.bounds = Box3f{{b[0], b[1], b[2]}, {b[3], b[4], b[5]}};
        const int32_t tc = r.pod<int32_t>(), vc = r.pod<int32_t>();
        m.has_normals = r.pod<uint8_t>() != 0;
        m.has_texcoords = r.pod<uint8_t>() != 0;
        if (tc < 0 || vc < 0) throw std::runtime_error("corrupt mesh header: " + path);
        m.vertex_stride = 3 + (m.has_normals ? 3 : 0) + (m.has_texcoords ? 2 : 0);
        m.indices.resize(3 * (size_t)tc);
        m.material_ids.resize(tc);
        m.vertex_attribs.resize((size_t)vc * m.vertex_stride);
        r.read(m.indices.data(), m.indices.size() * 4);
        r.read(m.material_ids.data(), m.material_ids.size() * 4);
        r.read(m.vertex_attribs.data(), m.vertex_attribs.size() * 4);
        const int32_t names = r.pod<int32_t>();
        for (int32_t i = 0; i < names; ++i) {
            const int32_t len = r.pod<int32_t>();
            if (len < 0) throw std::runtime_error("corrupt material name");
            std::string s((size_t)len, '\0');
            r.read(&s[0], (size_t)len);
            m.material_names.push_back(s);
        }
    } catch (...) {
        fclose(f);
        throw;
    }
    fclose(f);
    return m;
}

void save_mesh(const std::string& path, const TriangleMesh& m)
{
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("could not create file: " + path);
    Writer w{f};
    w.pod<int32_t>(kMeshBlobId);
    const float b[6] = {m.bounds.low.x, m.bounds.low.y, m.bounds.low.z,
                        m.bounds.high.x, m.bounds.high.y, m.bounds.high.z};
    w.write(b, sizeof(b));
    w.pod<int32_t>((int32_t)m.triangle_count());
    w.pod<int32_t>((int32_t)m.vertex_count());
    w.pod<uint8_t>(m.has_normals ? 1 : 0);
    w.pod<uint8_t>(m.has_texcoords ? 1 : 0);
    w.write(m.indices.data(), m.indices.size() * 4);
    w.write(m.material_ids.data(), m.material_ids.size() * 4);
    w.write(m.vertex_attribs.data(), m.vertex_attribs.size() * 4);
    w.pod<int32_t>((int32_t)m.material_names.size());
    for (const std::string& s : m.material_names) {
        w.pod<int32_t>((int32_t)s.size());
        w.write(s.data(), s.size());
    }
    fclose(f);
}

bool load_view_set(const std::string& path, View (&views)[kViewSetSize])
{
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return false;
    float buf[8 * kViewSetSize];
    const size_t n = fread(buf, 1, sizeof(buf), f);
    fclose(f);
    if (n != sizeof(buf)) return false;
    for (int i = 0; i < kViewSetSize; ++i) {
        const float* v = buf + 8 * i;
        views[i] = View{{v[0], v[1], v[2]}, v[3], v[4], v[5], v[6], v[7]};
    }
    return true;
}

void save_view_set(const std::string& path, const View (&views)[kViewSetSize])
{
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("could not create file: " + path);
    float buf[8 * kViewSetSize];
    for (int i = 0; i < kViewSetSize; ++i) {
        const View& v = views[i];
        const float vals[8] = {v.pos.x, v.pos.y, v.pos.z, v.angle_x, v.angle_y, v.fov_y, v.radius, v.focus};
        std::memcpy(buf + 8 * i, vals, sizeof(vals));
    }
    fwrite(buf, 1, sizeof(buf), f);
    fclose(f);
}

}  // namespace prt
