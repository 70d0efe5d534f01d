// mesh_build.h -- `protoray build`: a Wavefront OBJ file to a ProtoRay .mesh.
//
// Restates the reference's build command (main/build/build.cpp:31-130): the
// OBJ loader (main/build/obj_loader.cpp:37-329) makes a triangle soup, and
// TriangleMeshBuilder (geometry/triangle_mesh_builder.cpp:23-246) merges
// equal vertices and reorders vertices and triangles for locality.
#pragma once

#include <string>
#include <vector>

#include "mesh_io.h"

namespace prt {

struct SoupVertex {
    float pos[3] = {0, 0, 0};
    float normal[3] = {0, 0, 0};
    float tex[2] = {0, 0};
};

struct SoupTriangle {
    SoupVertex v[3];
    int material = 0;
};

struct TriangleSoup {
    std::vector<SoupTriangle> triangles;
    std::vector<std::string> material_names;  // [0] = "" (background)
};

// Loads an OBJ file: v, vn, vt, f (p, p/t, p//n, p/t/n; negative indices
// relative to the lists so far; polygons split as fans around the first
// vertex) and usemtl; other statements are skipped.  Triangles repeating a
// position index are dropped and counted in *degenerate.  Throws
// std::runtime_error with the reference loader's messages ("Could not open
// OBJ file: ...", "Invalid vertex position", "Invalid face",
// "Invalid vertex reference", ...).
TriangleSoup load_obj(const std::string& path, int* degenerate);

// Indexed mesh of a soup: vertices equal in position, normal and texture
// coordinates are merged; zero normals of a mesh that has normals get their
// triangle's (unnormalised) geometric normal; with `reorder`, vertices and
// triangles are laid out by recursive median splits of the vertices along
// the longest axis of their bounds.
TriangleMesh build_mesh(const TriangleSoup& soup, bool reorder = true);

}  // namespace prt
