// prt_math.h -- scalar numerics of the diffuse path, shared by every kernel.
//
// All functions here are written so that their fp32 results do not depend on
// the compiler: floating-point contraction is off, fused multiply-adds are
// spelled out with fmaf, and only correctly rounded IEEE operations are used
// (add, sub, mul, div, sqrt, fma).  A CPU implementation that performs the
// same operations in the same order therefore reproduces every bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PRT_HD __host__ __device__ __forceinline__

#pragma clang fp contract(off)

namespace prt {

struct f3 {
    float x, y, z;
};

PRT_HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
PRT_HD f3 add3(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
PRT_HD f3 sub3(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
PRT_HD f3 mul3(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
PRT_HD f3 neg3(f3 a) { return f3{-a.x, -a.y, -a.z}; }
PRT_HD float dot3(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
PRT_HD f3 cross3(f3 a, f3 b)
{
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// normalize(a) = a * (1 / sqrt(a.a)); the reference uses an approximate
// rsqrt (math/math_avx.h:30-34, math/vec3.h:395-398), this build the exact one.
PRT_HD f3 normalize3(f3 a)
{
    float inv = 1.0f / sqrtf(dot3(a, a));
    return mul3(a, inv);
}
PRT_HD float max3f(float a, float b, float c) { return fmaxf(fmaxf(a, b), c); }

// ---------------------------------------------------------------------------
// Sampler: murmur3 x86_32 mix/finalize of (seed=0, pixelId, pass), then an LCG
// per 1D draw.  sampling/random_sampler.h:44-73, math/hash.h:23-72.
// ---------------------------------------------------------------------------
PRT_HD uint32_t rotl32(uint32_t x, uint32_t r) { return (x << r) | (x >> (32u - r)); }

PRT_HD uint32_t murmur_mix(uint32_t h, uint32_t k)
{
    k *= 0xcc9e2d51u;
    k = rotl32(k, 15u);
    k *= 0x1b873593u;
    h ^= k;
    h = rotl32(h, 13u) * 5u + 0xe6546b64u;
    return h;
}

PRT_HD uint32_t murmur_fmix(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

PRT_HD uint32_t sampler_seed(int pass, int pixel_id)
{
    uint32_t h = 0u;
    h = murmur_mix(h, (uint32_t)pixel_id);
    h = murmur_mix(h, (uint32_t)pass);
    return murmur_fmix(h);
}

PRT_HD uint32_t lcg_next(uint32_t s) { return s * 1664525u + 1013904223u; }

// u32 -> [0,1]: round-to-nearest conversion times 2^-32, the CPU scalar form
// (math/math.h:179).  It can return exactly 1.0f.
PRT_HD float to_unorm(uint32_t s) { return (float)s * 0x1.0p-32f; }

PRT_HD float sampler_next(uint32_t& s)
{
    s = lcg_next(s);
    return to_unorm(s);
}

// ---------------------------------------------------------------------------
// sin(2 pi v), cos(2 pi v) for v in [0, 1]: quadrant reduction by k = round(4v)
// (exact: 4v is a power-of-two scaling and 4v - k is exact by Sterbenz), then
// Taylor polynomials of sin and cos on |x| <= pi/4 evaluated by Horner with
// fmaf.  Truncation error < 3e-8 absolute, below fp32 rounding at this range.
// ---------------------------------------------------------------------------
PRT_HD void sincos_2pi(float v, float& s_out, float& c_out)
{
    float w = v * 4.0f;
    int k = (int)(w + 0.5f);
    float f = w - (float)k;
    float x = f * 1.57079632679489662f;  // pi/2
    float x2 = x * x;
    // sin x = x (1 - x^2/3! + x^4/5! - x^6/7! + x^8/9!)
    float ps = 2.75573192239858907e-6f;          // 1/9!
    ps = fmaf(ps, x2, -1.98412698412698413e-4f); // -1/7!
    ps = fmaf(ps, x2, 8.33333333333333333e-3f);  // 1/5!
    ps = fmaf(ps, x2, -1.66666666666666667e-1f); // -1/3!
    float sx = fmaf(ps * x2, x, x);
    // cos x = 1 - x^2/2! + x^4/4! - x^6/6! + x^8/8! - x^10/10!
    float pc = -2.75573192239858907e-7f;         // -1/10!
    pc = fmaf(pc, x2, 2.48015873015873016e-5f);  // 1/8!
    pc = fmaf(pc, x2, -1.38888888888888889e-3f); // -1/6!
    pc = fmaf(pc, x2, 4.16666666666666667e-2f);  // 1/4!
    pc = fmaf(pc, x2, -0.5f);
    float cx = fmaf(pc, x2, 1.0f);
    switch (k & 3) {
    case 0: s_out = sx; c_out = cx; break;
    case 1: s_out = cx; c_out = -sx; break;
    case 2: s_out = -sx; c_out = -cx; break;
    default: s_out = -cx; c_out = sx; break;
    }
}

// ---------------------------------------------------------------------------
// Pinhole camera ray (camera/pinhole_camera.h:51-55): image = (pixel + jitter)
// * invSize; dir = normalize(imageO + image.x imageDx + image.y imageDy).
// ---------------------------------------------------------------------------
struct Camera {
    f3 origin, image_o, image_dx, image_dy;
    f3 basis_u, basis_v;
    float lens_radius, focal_distance;
};

PRT_HD f3 camera_dir(const Camera& c, float ix, float iy)
{
    f3 d;
    d.x = (c.image_o.x + ix * c.image_dx.x) + iy * c.image_dy.x;
    d.y = (c.image_o.y + ix * c.image_dx.y) + iy * c.image_dy.y;
    d.z = (c.image_o.z + ix * c.image_dx.z) + iy * c.image_dy.z;
    return normalize3(d);
}

// Camera ray of image position (ix, iy) and lens sample (l1, l2).  Pinhole
// (lens_radius 0): the origin and normalize(imageO + ix dx + iy dy)
// (camera/pinhole_camera.h:50-54).  Thin lens: ThinLensCamera::getRay
// (camera/thin_lens_camera.h:38-44) with uniformSampleDisk
// (sampling/shape_sampler.h:115-124): lens = (r cos 2 pi l2, r sin 2 pi l2) R,
// r = sqrt(l1); begin = origin + (lens.x U + lens.y V); end = origin + f
// (imageO + ix dx + iy dy); direction normalize(end - begin).
PRT_HD void camera_ray(const Camera& c, float ix, float iy, float l1, float l2, f3& org, f3& dir)
{
    if (c.lens_radius == 0.0f) {
        org = c.origin;
        dir = camera_dir(c, ix, iy);
        return;
    }
    const float r = sqrtf(l1);
    float sn, cs;
    sincos_2pi(l2, sn, cs);
    const float lx = (r * cs) * c.lens_radius, ly = (r * sn) * c.lens_radius;
    f3 b;
    b.x = c.origin.x + (lx * c.basis_u.x + ly * c.basis_v.x);
    b.y = c.origin.y + (lx * c.basis_u.y + ly * c.basis_v.y);
    b.z = c.origin.z + (lx * c.basis_u.z + ly * c.basis_v.z);
    f3 e;
    e.x = c.origin.x + c.focal_distance * ((c.image_o.x + ix * c.image_dx.x) + iy * c.image_dy.x);
    e.y = c.origin.y + c.focal_distance * ((c.image_o.y + ix * c.image_dx.y) + iy * c.image_dy.y);
    e.z = c.origin.z + c.focal_distance * ((c.image_o.z + ix * c.image_dx.z) + iy * c.image_dy.z);
    org = b;
    dir = normalize3(mk3(e.x - b.x, e.y - b.y, e.z - b.z));
}

// ---------------------------------------------------------------------------
// Shading frame and diffuse bounce (render/diffuse_integrator_single.h:55-62
// with geometry/triangle_mesh.cpp:59-153,248-273, core/ray.h:47-52,71-76,
// math/basis3.h:128-145, sampling/shape_sampler.h:50-74).
// ---------------------------------------------------------------------------
struct Frame {
    f3 u, v, n;
};

// makeFrame: U from the larger of (0, N.z, -N.y) and (-N.z, 0, N.x), then
// V = normalize(N x U).
PRT_HD Frame make_frame(f3 n)
{
    f3 u0 = mk3(0.0f, n.z, -n.y);
    f3 u1 = mk3(-n.z, 0.0f, n.x);
    Frame f;
    f.u = normalize3(dot3(u0, u0) > dot3(u1, u1) ? u0 : u1);
    f.v = normalize3(cross3(n, f.u));
    f.n = n;
    return f;
}

// Geometric normal of the hit triangle, flipped to face against the ray.
PRT_HD f3 face_normal(f3 p0, f3 p1, f3 p2, f3 dir)
{
    f3 ng = normalize3(cross3(sub3(p1, p0), sub3(p2, p0)));
    if (dot3(ng, dir) > 0.0f) ng = neg3(ng);
    return ng;
}

// Per-vertex shading attributes of the three corners of a hit triangle.
struct CornerAttributes {
    bool has_normals, has_texcoords;
    f3 n0, n1, n2;
    float u0, v0, u1, v1, u2, v2;  // texture coordinates
};

// Full shading frame of a hit (TriangleMesh::postIntersect with a
// ShadingContext, geometry/triangle_mesh.cpp:59-153 and
// geometry/triangle_mesh_cuda.cuh:41-130).  (bu, bv) are the hit's
// barycentric coordinates of p1 and p2, so p = (1 - bu - bv) p0 + bu p1 + bv p2.
//   N: the interpolated vertex normal, turned to the side of the geometric
//      normal Ng (Ng itself without normals); both are flipped when Ng faces
//      along the ray.
//   U, V: with texture coordinates, U = normalize(dp/du) and V = N x U
//      normalised, then U = V x N (makeFrame when the texture mapping or
//      N x U is degenerate); without them, makeFrame(N).
// 1/det is the correctly rounded quotient (the reference's rcp).
PRT_HD Frame shading_frame(f3 p0, f3 p1, f3 p2, f3 dir, float bu, float bv, const CornerAttributes& a)
{
    const float b1 = bu, b2 = bv;
    const float b0 = 1.0f - b1 - b2;
    const f3 dp1 = sub3(p1, p0), dp2 = sub3(p2, p0);
    f3 ng = normalize3(cross3(dp1, dp2));
    f3 n = ng;
    if (a.has_normals) {
        f3 s;
        s.x = (b0 * a.n0.x + b1 * a.n1.x) + b2 * a.n2.x;
        s.y = (b0 * a.n0.y + b1 * a.n1.y) + b2 * a.n2.y;
        s.z = (b0 * a.n0.z + b1 * a.n1.z) + b2 * a.n2.z;
        n = normalize3(s);
        if (dot3(n, ng) < 0.0f) n = neg3(n);
    }
    if (dot3(ng, dir) > 0.0f) n = neg3(n);
    if (a.has_texcoords) {
        const float du1 = a.u1 - a.u0, du2 = a.u2 - a.u0;
        const float dv1 = a.v1 - a.v0, dv2 = a.v2 - a.v0;
        const float det = du1 * dv2 - dv1 * du2;
        if (det != 0.0f) {
            const float inv_det = 1.0f / det;
            f3 dpdu;
            dpdu.x = (dv2 * dp1.x - dv1 * dp2.x) * inv_det;
            dpdu.y = (dv2 * dp1.y - dv1 * dp2.y) * inv_det;
            dpdu.z = (dv2 * dp1.z - dv1 * dp2.z) * inv_det;
            Frame f;
            f.n = n;
            f.u = normalize3(dpdu);
            f.v = cross3(n, f.u);
            if (dot3(f.v, f.v) > 0.0f) {
                f.v = normalize3(f.v);
                f.u = cross3(f.v, n);
                return f;
            }
        }
    }
    return make_frame(n);
}

// New ray of a diffuse bounce at hit distance t along (org, dir): cosine-
// weighted direction in frame f, origin offset by eps along it, tnear 0.
PRT_HD void diffuse_bounce(f3 org, f3 dir, float t, const Frame& f, float s1, float s2, f3& org_o, f3& dir_o)
{
    f3 p;
    p.x = org.x + t * dir.x;
    p.y = org.y + t * dir.y;
    p.z = org.z + t * dir.z;
    float eps = fmaxf(t, max3f(fabsf(p.x), fabsf(p.y), fabsf(p.z))) * 0x1.fp-18f;
    // cosineSampleHemisphere
    float cos_t = sqrtf(s1);
    float sin_t = sqrtf(1.0f - s1);
    float sp, cp;
    sincos_2pi(s2, sp, cp);
    float lx = cp * sin_t;
    float ly = sp * sin_t;
    float lz = cos_t;
    f3 d;
    d.x = (lx * f.u.x + ly * f.v.x) + lz * f.n.x;
    d.y = (lx * f.u.y + ly * f.v.y) + lz * f.n.y;
    d.z = (lx * f.u.z + ly * f.v.z) + lz * f.n.z;
    dir_o = d;
    org_o.x = p.x + eps * d.x;
    org_o.y = p.y + eps * d.y;
    org_o.z = p.z + eps * d.z;
}

// The same with the frame of the (ray-facing) geometric normal n, the
// SimpleShadingContext of a mesh without shading attributes.
PRT_HD void diffuse_bounce(f3 org, f3 dir, float t, f3 n, float s1, float s2, f3& org_o, f3& dir_o)
{
    diffuse_bounce(org, dir, t, make_frame(n), s1, s2, org_o, dir_o);
}

// ---------------------------------------------------------------------------
// Watertight ray/triangle test (S. Woop, C. Benthin, I. Wald, "Watertight
// Ray/Triangle Intersection", JCGT 2(1), 2013).  Per-ray set-up below; the
// per-triangle part is tri_test.  U, V, W are the scaled barycentrics of
// p0, p1, p2; a zero edge function is recomputed in double.
// ---------------------------------------------------------------------------
struct WatertightRay {
    int kx, ky, kz;
    float sx, sy, sz;
};

PRT_HD float comp(f3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

PRT_HD WatertightRay watertight_setup(f3 d)
{
    WatertightRay w;
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = 0;
    float m = ax;
    if (ay > m) { kz = 1; m = ay; }
    if (az > m) { kz = 2; }
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    float dz = comp(d, kz);
    if (dz < 0.0f) { int t = kx; kx = ky; ky = t; }
    w.kx = kx; w.ky = ky; w.kz = kz;
    w.sx = comp(d, kx) / dz;
    w.sy = comp(d, ky) / dz;
    w.sz = 1.0f / dz;
    return w;
}

// Distance to a triangle's plane in double, rounded to float:
// t = ((p0 - o) . n) / (d . n) with n = (p1 - p0) x (p2 - p0).  Differences of
// floats and their products are exact in double; each sum rounds once, so t
// is the plane distance to within float rounding even for tiny t, where the
// fp32 T/det of the watertight test loses most of its digits.  False if the
// plane is parallel to the ray.
PRT_HD bool plane_t(f3 p0, f3 p1, f3 p2, f3 o, f3 d, float& t_out)
{
    double e1x = (double)p1.x - (double)p0.x, e1y = (double)p1.y - (double)p0.y,
           e1z = (double)p1.z - (double)p0.z;
    double e2x = (double)p2.x - (double)p0.x, e2y = (double)p2.y - (double)p0.y,
           e2z = (double)p2.z - (double)p0.z;
    double nx = e1y * e2z - e1z * e2y;
    double ny = e1z * e2x - e1x * e2z;
    double nz = e1x * e2y - e1y * e2x;
    double wx = (double)p0.x - (double)o.x, wy = (double)p0.y - (double)o.y,
           wz = (double)p0.z - (double)o.z;
    double num = (wx * nx + wy * ny) + wz * nz;
    double den = ((double)d.x * nx + (double)d.y * ny) + (double)d.z * nz;
    if (den == 0.0) return false;
    t_out = (float)(num / den);
    return true;
}

// Candidate intersection.  The watertight test decides whether the ray's line
// crosses the triangle (fp32 edge functions, a zero one recomputed in
// double); the hit distance is the plane distance of plane_t.  Returns false
// for a miss, a plane parallel to the ray or a distance outside [tmin, tmax];
// (V, W, det) are kept so that the caller computes u = V/det, v = W/det only
// for accepted hits.  Ordering candidates by this t rather than by the fp32
// T/det makes the choice of the nearest triangle follow the true distances,
// and keeps the traversal's conservative box culling (slack relative to the
// box distances) consistent with every candidate's t.
struct TriCandidate {
    float t, v_num, w_num, det;
};

// Scaled barycentrics U, V, W of p0, p1, p2 (the edge functions in the
// ray's shear space); a zero one sends all three to double.
PRT_HD void tri_edges(const WatertightRay& w, f3 org, f3 p0, f3 p1, f3 p2, float& U, float& V, float& W)
{
    f3 a = sub3(p0, org), b = sub3(p1, org), cc = sub3(p2, org);
    float az_ = comp(a, w.kz), bz_ = comp(b, w.kz), cz_ = comp(cc, w.kz);
    float ax = comp(a, w.kx) - w.sx * az_;
    float ay = comp(a, w.ky) - w.sy * az_;
    float bx = comp(b, w.kx) - w.sx * bz_;
    float by = comp(b, w.ky) - w.sy * bz_;
    float cx = comp(cc, w.kx) - w.sx * cz_;
    float cy = comp(cc, w.ky) - w.sy * cz_;
    U = cx * by - cy * bx;
    V = ax * cy - ay * cx;
    W = bx * ay - by * ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)cx * (double)by - (double)cy * (double)bx);
        V = (float)((double)ax * (double)cy - (double)ay * (double)cx);
        W = (float)((double)bx * (double)ay - (double)by * (double)ax);
    }
}

PRT_HD bool tri_test(const WatertightRay& w, f3 org, f3 dir, f3 p0, f3 p1, f3 p2, float tmin, float tmax,
                     TriCandidate& c)
{
    float U, V, W;
    tri_edges(w, org, p0, p1, p2, U, V, W);
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = (U + V) + W;
    if (det == 0.0f) return false;
    float t;
    if (!plane_t(p0, p1, p2, org, dir, t)) return false;
    if (!(t >= tmin && t <= tmax)) return false;
    c.t = t;
    c.v_num = V;
    c.w_num = W;
    c.det = det;
    return true;
}

// Hit coordinates (u, v) = (V/det, W/det) of a triangle that tri_test
// accepted for this ray, recomputed from the same edge functions (so equal to
// what tri_test's caller derives from the candidate).
PRT_HD void tri_barycentrics(const WatertightRay& w, f3 org, f3 p0, f3 p1, f3 p2, float& u, float& v)
{
    float U, V, W;
    tri_edges(w, org, p0, p1, p2, U, V, W);
    const float det = (U + V) + W;
    u = V / det;
    v = W / det;
}

}  // namespace prt
