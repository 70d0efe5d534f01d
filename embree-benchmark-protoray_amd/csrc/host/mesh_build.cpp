// mesh_build.cpp -- OBJ loading and mesh building for `protoray build`.
#include "mesh_build.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <unordered_map>

namespace prt {
namespace {

struct ObjRef {
    int p = 0, t = 0, n = 0;  // 1-based, 0 = absent
};

// negative OBJ indices count back from the end of the list read so far
// (ObjLoader::convertIndex, main/build/obj_loader.cpp:242-247)
int absolute_index(int i, size_t size) { return i >= 0 ? i : i + (int)size + 1; }

// one face vertex token: p/t/n, p//n, p/t or p, tried in that order
// (ObjLoader::parseFace, main/build/obj_loader.cpp:160-185)
bool parse_ref(const char* tok, ObjRef& r)
{
    if (std::sscanf(tok, "%d/%d/%d", &r.p, &r.t, &r.n) == 3) return true;
    if (std::sscanf(tok, "%d//%d", &r.p, &r.n) == 2) { r.t = 0; return true; }
    if (std::sscanf(tok, "%d/%d", &r.p, &r.t) == 2) { r.n = 0; return true; }
    if (std::sscanf(tok, "%d", &r.p) == 1) { r.t = r.n = 0; return true; }
    return false;
}

struct ObjTri {
    ObjRef v[3];
    int material;
};

uint32_t fnv1a(const float* f, int n, uint32_t h)
{
    for (int k = 0; k < n; ++k) {
        uint32_t b;
        std::memcpy(&b, f + k, 4);
        h = (h ^ b) * 16777619u;
    }
    return h;
}

// vertex identity of the builder's hash map: FNV-1a over the float bits
// (geometry/vertex.h:55-78), equality by float comparison
struct VertexHash {
    size_t operator()(const SoupVertex& v) const
    {
        return fnv1a(v.tex, 2, fnv1a(v.normal, 3, fnv1a(v.pos, 3, 2166136261u)));
    }
};
struct VertexEq {
    bool operator()(const SoupVertex& a, const SoupVertex& b) const
    {
        for (int k = 0; k < 3; ++k)
            if (a.pos[k] != b.pos[k] || a.normal[k] != b.normal[k]) return false;
        return a.tex[0] == b.tex[0] && a.tex[1] == b.tex[1];
    }
};

struct Fragment {
    float pos[3];
    int id;
};

int longest_axis(const float lo[3], const float hi[3])
{
    const float sx = hi[0] - lo[0], sy = hi[1] - lo[1], sz = hi[2] - lo[2];
    const float s[3] = {sx, sy, sz};
    int axis = sx > sy ? 0 : 1;  // selectMax (math/vec3.h:444-449)
    return s[axis] > sz ? axis : 2;
}

// Cache-oblivious layout (TriangleMeshBuilder::reorderFragments,
// geometry/triangle_mesh_builder.cpp:152-210): split the vertices at their
// median along the longest axis of their bounds (quickselect with the last
// element as pivot), split the triangles by the same plane through their
// centres, and recurse on both halves.
void reorder(Fragment* vb, Fragment* ve, Fragment* tb, Fragment* te)
{
    if (vb + 1 >= ve) return;
    float lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = std::numeric_limits<float>::infinity();
        hi[k] = -std::numeric_limits<float>::infinity();
    }
    for (Fragment* f = vb; f < ve; ++f)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], f->pos[k]);
            hi[k] = std::max(hi[k], f->pos[k]);
        }
    const int axis = longest_axis(lo, hi);
    Fragment* vm = vb + (ve - vb) / 2;
    Fragment* first = vb;
    Fragment* last = ve;
    for (;;) {
        const float pivot = (last - 1)->pos[axis];
        Fragment* p = std::partition(first, last - 1, [=](const Fragment& a) { return a.pos[axis] < pivot; });
        std::iter_swap(last - 1, p);
        if (p == vm) break;
        if (p > vm) last = p;
        else first = p + 1;
    }
    const float split = vm->pos[axis];
    Fragment* tm = tb;
    for (Fragment* f = tb; f < te; ++f) {
        if (f->pos[axis] < split) {
            std::swap(*f, *tm);
            ++tm;
        }
    }
    reorder(vb, vm, tb, tm);
    reorder(vm, ve, tm, te);
}

}  // namespace

TriangleSoup load_obj(const std::string& path, int* degenerate)
{
    FILE* file = std::fopen(path.c_str(), "rt");
    if (!file) throw std::runtime_error("Could not open OBJ file: " + path);
    std::vector<float> positions, normals, texcoords;
    std::vector<ObjTri> tris;
    std::vector<std::string> materials{""};  // 0 is reserved for background
    int current = -1;
    int dropped = 0;
    std::vector<char> line(256 * 1024);
    char head[256];
    auto fail = [&](const char* msg) {
        std::fclose(file);
        throw std::runtime_error(msg);
    };
    while (std::fscanf(file, "%255s", head) == 1) {
        line[0] = '\0';
        if (!std::fgets(line.data(), (int)line.size(), file)) line[0] = '\0';
        const char* b = line.data();
        if (std::strcmp(head, "v") == 0 || std::strcmp(head, "vn") == 0) {
            float x, y, z;
            if (std::sscanf(b, "%f %f %f", &x, &y, &z) != 3)
                fail(head[1] ? "Invalid vertex normal" : "Invalid vertex position");
            std::vector<float>& dst = head[1] ? normals : positions;
            dst.insert(dst.end(), {x, y, z});
        } else if (std::strcmp(head, "vt") == 0) {
            float u, v;
            if (std::sscanf(b, "%f %f", &u, &v) != 2) fail("Invalid vertex texture coordinates");
            texcoords.insert(texcoords.end(), {u, v});
        } else if (std::strcmp(head, "f") == 0) {
            std::vector<ObjRef> refs;
            bool ok = true;
            for (char* tok = std::strtok(line.data(), " \t\n\r"); tok; tok = std::strtok(nullptr, " \t\n\r")) {
                ObjRef r;
                if (!parse_ref(tok, r)) ok = false;
                r.p = absolute_index(r.p, positions.size() / 3);
                r.t = absolute_index(r.t, texcoords.size() / 2);
                r.n = absolute_index(r.n, normals.size() / 3);
                refs.push_back(r);
            }
            if (!ok || refs.size() < 3) fail("Invalid face");
            if (current < 0) {
                materials.push_back("Default");
                current = (int)materials.size() - 1;
            }
            for (size_t i = 0; i + 2 < refs.size(); ++i) {
                ObjTri t{{refs[0], refs[i + 1], refs[i + 2]}, current};
                if (t.v[0].p == t.v[1].p || t.v[0].p == t.v[2].p || t.v[1].p == t.v[2].p) ++dropped;
                else tris.push_back(t);
            }
        } else if (std::strcmp(head, "usemtl") == 0) {
            char name[256];
            if (std::sscanf(b, "%255s", name) != 1) fail("Invalid material");
            auto it = std::find(materials.begin(), materials.end(), std::string(name));
            if (it == materials.end()) {
                materials.push_back(name);
                current = (int)materials.size() - 1;
            } else {
                current = (int)(it - materials.begin());
            }
        }
    }
    std::fclose(file);
    if (degenerate) *degenerate = dropped;

    // soup (ObjLoader::makeSoup, main/build/obj_loader.cpp:249-316)
    TriangleSoup soup;
    soup.triangles.resize(tris.size());
    const int np = (int)(positions.size() / 3), nn = (int)(normals.size() / 3), nt = (int)(texcoords.size() / 2);
    for (size_t i = 0; i < tris.size(); ++i) {
        for (int k = 0; k < 3; ++k) {
            const ObjRef& r = tris[i].v[k];
            SoupVertex& v = soup.triangles[i].v[k];
            if (r.p < 1 || r.p > np || r.n < 0 || r.n > nn || r.t < 0 || r.t > nt)
                throw std::runtime_error("Invalid vertex reference");
            std::memcpy(v.pos, &positions[3 * (size_t)(r.p - 1)], 12);
            if (r.n > 0) std::memcpy(v.normal, &normals[3 * (size_t)(r.n - 1)], 12);
            if (r.t > 0) std::memcpy(v.tex, &texcoords[2 * (size_t)(r.t - 1)], 8);
        }
        soup.triangles[i].material = tris[i].material;
    }
    soup.material_names = materials;
    return soup;
}

TriangleMesh build_mesh(const TriangleSoup& soup, bool do_reorder)
{
    const size_t T = soup.triangles.size();
    bool has_normals = false, has_tex = false, has_materials = false;
    for (const SoupTriangle& t : soup.triangles) {
        for (const SoupVertex& v : t.v) {
            if (v.normal[0] != 0.0f || v.normal[1] != 0.0f || v.normal[2] != 0.0f) has_normals = true;
            if (v.tex[0] != 0.0f || v.tex[1] != 0.0f) has_tex = true;
        }
        if (t.material != 0) has_materials = true;
    }

    std::unordered_map<SoupVertex, int, VertexHash, VertexEq> ids;
    ids.reserve(T * 3);
    std::vector<SoupVertex> vertices;
    std::vector<Fragment> vfrag, tfrag(T);
    std::vector<int32_t> tri_ids(3 * T);
    float lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = std::numeric_limits<float>::infinity();
        hi[k] = -std::numeric_limits<float>::infinity();
    }
    for (size_t i = 0; i < T; ++i) {
        const SoupTriangle& t = soup.triangles[i];
        // cross(p1 - p0, p2 - p0), the triangle's unnormalised normal
        float e1[3], e2[3];
        for (int k = 0; k < 3; ++k) {
            e1[k] = t.v[1].pos[k] - t.v[0].pos[k];
            e2[k] = t.v[2].pos[k] - t.v[0].pos[k];
        }
        const float gn[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                             e1[0] * e2[1] - e1[1] * e2[0]};
        for (int k = 0; k < 3; ++k) {
            SoupVertex v = t.v[k];
            if (has_normals && v.normal[0] == 0.0f && v.normal[1] == 0.0f && v.normal[2] == 0.0f)
                std::memcpy(v.normal, gn, 12);
            auto ins = ids.emplace(v, (int)vertices.size());
            if (ins.second) {
                vertices.push_back(v);
                Fragment f;
                std::memcpy(f.pos, v.pos, 12);
                f.id = ins.first->second;
                vfrag.push_back(f);
                for (int a = 0; a < 3; ++a) {
                    lo[a] = std::min(lo[a], v.pos[a]);
                    hi[a] = std::max(hi[a], v.pos[a]);
                }
            }
            tri_ids[3 * i + k] = ins.first->second;
        }
        // centre (p0 + p1 + p2) * (1/3)
        for (int a = 0; a < 3; ++a)
            tfrag[i].pos[a] = ((t.v[0].pos[a] + t.v[1].pos[a]) + t.v[2].pos[a]) * (1.0f / 3.0f);
        tfrag[i].id = (int)i;
    }
    if (do_reorder) reorder(vfrag.data(), vfrag.data() + vfrag.size(), tfrag.data(), tfrag.data() + tfrag.size());

    TriangleMesh mesh;
    mesh.has_normals = has_normals;
    mesh.has_texcoords = has_tex;
    mesh.vertex_stride = 3 + (has_normals ? 3 : 0) + (has_tex ? 2 : 0);
    const size_t V = vertices.size();
    mesh.vertex_attribs.resize(V * mesh.vertex_stride);
    std::vector<int32_t> new_id(V);
    for (size_t i = 0; i < V; ++i) {
        const SoupVertex& v = vertices[vfrag[i].id];
        new_id[vfrag[i].id] = (int32_t)i;
        float* a = &mesh.vertex_attribs[i * mesh.vertex_stride];
        std::memcpy(a, v.pos, 12);
        int o = 3;
        if (has_normals) { std::memcpy(a + o, v.normal, 12); o += 3; }
        if (has_tex) std::memcpy(a + o, v.tex, 8);
    }
    mesh.indices.resize(3 * T);
    mesh.material_ids.resize(T);
    for (size_t i = 0; i < T; ++i) {
        const int old = tfrag[i].id;
        for (int k = 0; k < 3; ++k) mesh.indices[3 * i + k] = new_id[tri_ids[3 * (size_t)old + k]];
        mesh.material_ids[i] = has_materials ? soup.triangles[old].material : 0;
    }
    mesh.material_names = soup.material_names;
    mesh.bounds.low = Vec3f{lo[0], lo[1], lo[2]};
    mesh.bounds.high = Vec3f{hi[0], hi[1], hi[2]};
    return mesh;
}

}  // namespace prt
