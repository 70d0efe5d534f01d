"""Writes two_tris.mesh: a 2-triangle mesh assembled field by field with
struct, following the .mesh layout of geometry/triangle_mesh.cpp:305-337
(blob id, Box3f, int T, int V, bool hasNormals, bool hasTexcoords,
int32[3T], int32[T], float[V*3], names)."""
import os
import struct

pos = [(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (1.0, 1.0, 0.0), (0.0, 1.0, 0.5)]
tris = [(0, 1, 2), (0, 2, 3)]
out = struct.pack("<i", 0x04DF8666)
out += struct.pack("<6f", 0.0, 0.0, 0.0, 1.0, 1.0, 0.5)
out += struct.pack("<ii??", 2, 4, False, False)
for t in tris:
    out += struct.pack("<3i", *t)
out += struct.pack("<2i", 0, 0)
for p in pos:
    out += struct.pack("<3f", *p)
out += struct.pack("<i", 1) + struct.pack("<i", 7) + b"default"
path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "two_tris.mesh")
with open(path, "wb") as f:
    f.write(out)
