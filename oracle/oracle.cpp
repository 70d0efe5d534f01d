// oracle.cpp -- CPU restatement of ProtoRay's diffuse path, used ONLY as the
// checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
// It is test infrastructure: nothing in the shipped library calls it.
//
// What it restates (reference paths in embree-benchmark-protoray):
//   sampler          sampling/random_sampler.h:44-73, math/hash.h:23-72, math/math.h:179
//   camera           camera/pinhole_camera.h:34-55, main/render/render_window.cpp:307-325,
//                    math/basis3.h:53-79, math/vec3.h:409-412
//   primary rays     render/renderer_single.h:102-123
//   diffuse path     render/diffuse_integrator_single.h:41-69, core/ray.h:47-52,71-76,
//                    geometry/triangle_mesh.cpp:248-273, math/basis3.h:128-145,
//                    sampling/shape_sampler.h:50-74
//   primary shading  render/primary_integrator_single.h:34-47
//   accumulation     image/accum_buffer.h:124-147
//   stream renderer  render/diffuse_renderer_stream.h:82-225, render/renderer_stream.h:69-88
//   tasking          sys/tasking.h:33-52 (atomic tile counter over threads)
//
// Intersection: the reference delegates it to Embree 3 / OptiX Prime, which are
// absent here.  Two restatements:
//   - FP64: watertight test (Woop, Benthin, Wald, JCGT 2013) in double over a
//     BVH2, closest (t, primId) pair; the reference-semantics parity check.
//   - kernel numerics: the same test in fp32 with the double fallback for zero
//     edge functions and a (t, primId) lexicographic minimum, evaluated over
//     every triangle the double-precision box test admits; the GPU kernel must
//     reproduce it bit for bit.
// Numerics: compile with -ffp-contract=off (no fused operations unless spelled
// fmaf), IEEE scalar SSE arithmetic.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include <immintrin.h>

namespace {

// ----------------------------------------------------------------- sampler
uint32_t rotl(uint32_t x, uint32_t r) { return (x << r) | (x >> (32u - r)); }

uint32_t mmix(uint32_t h, uint32_t k)
{
    k *= 0xcc9e2d51u;
    k = rotl(k, 15u);
    k *= 0x1b873593u;
    h ^= k;
    return rotl(h, 13u) * 5u + 0xe6546b64u;
}

uint32_t mfin(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

uint32_t seed_of(int pass, int pixel) { return mfin(mmix(mmix(0u, (uint32_t)pixel), (uint32_t)pass)); }

struct Sampler {
    uint32_t s;
    float next()
    {
        s = s * 1664525u + 1013904223u;
        return (float)s * 0x1.0p-32f;
    }
};

// ----------------------------------------------------------------- vectors
struct V3 {
    float x, y, z;
};
V3 v3(float x, float y, float z) { return V3{x, y, z}; }
V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
V3 cross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
V3 normalized(V3 a)
{
    float inv = 1.0f / std::sqrt(dot(a, a));
    return a * inv;
}
float at(V3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

struct D3 {
    double x, y, z;
};
D3 d3(V3 a) { return D3{a.x, a.y, a.z}; }
D3 operator-(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
D3 operator+(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
D3 operator*(D3 a, double s) { return D3{a.x * s, a.y * s, a.z * s}; }
double dotd(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 crossd(D3 a, D3 b) { return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double atd(D3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// ----------------------------------------------------------------- sin/cos
// sin(2 pi v), cos(2 pi v): quadrant k = (int)(4v + 0.5), x = (4v - k) pi/2,
// Taylor polynomials (x^9 for sin, x^10 for cos) by Horner with fmaf.
void sincos_2pi_kernel(float v, float& so, float& co)
{
    float w = v * 4.0f;
    int k = (int)(w + 0.5f);
    float f = w - (float)k;
    float x = f * 1.57079632679489662f;
    float x2 = x * x;
    float ps = 2.75573192239858907e-6f;
    ps = std::fma(ps, x2, -1.98412698412698413e-4f);
    ps = std::fma(ps, x2, 8.33333333333333333e-3f);
    ps = std::fma(ps, x2, -1.66666666666666667e-1f);
    float sx = std::fma(ps * x2, x, x);
    float pc = -2.75573192239858907e-7f;
    pc = std::fma(pc, x2, 2.48015873015873016e-5f);
    pc = std::fma(pc, x2, -1.38888888888888889e-3f);
    pc = std::fma(pc, x2, 4.16666666666666667e-2f);
    pc = std::fma(pc, x2, -0.5f);
    float cx = std::fma(pc, x2, 1.0f);
    switch (k & 3) {
    case 0: so = sx; co = cx; break;
    case 1: so = cx; co = -sx; break;
    case 2: so = -sx; co = -cx; break;
    default: so = -cx; co = sx; break;
    }
}

// ----------------------------------------------------------------- scene
struct Tri {
    V3 p0, p1, p2;
};

struct Node {      // BVH2 node over float boxes
    float lo[3], hi[3];
    int left;      // internal: index of left child (right = left + 1); leaf: -1
    int first, count;
};

struct Scene {
    std::vector<Tri> tris;
    std::vector<int> order;   // leaf order -> triangle id
    std::vector<Node> nodes;
    // 8-wide version of the same tree for the CPU timing baseline (built on
    // first use, see build_bvh8)
    struct WideTree;
    WideTree* wide = nullptr;
    Scene() = default;
    Scene(const Scene&) = delete;
    Scene& operator=(const Scene&) = delete;
    ~Scene();
    // vertex indices and the optional per-vertex shading attributes
    std::vector<int32_t> idx;     // 3 per triangle
    std::vector<V3> normals;      // empty: no normals
    std::vector<float> texcoords; // 2 per vertex; empty: none
    int64_t vertex_count = 0;
};

void tri_bounds(const Tri& t, float lo[3], float hi[3])
{
    for (int k = 0; k < 3; ++k) {
        float a = at(t.p0, k), b = at(t.p1, k), c = at(t.p2, k);
        lo[k] = std::min(std::min(a, b), c);
        hi[k] = std::max(std::max(a, b), c);
    }
}

// binned SAH build, 16 bins, leaves of up to 4 triangles
void build_bvh(Scene& sc)
{
    const int n = (int)sc.tris.size();
    sc.order.resize(n);
    for (int i = 0; i < n; ++i) sc.order[i] = i;
    std::vector<float> blo(3 * (size_t)n), bhi(3 * (size_t)n), cen(3 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        tri_bounds(sc.tris[i], &blo[3 * i], &bhi[3 * i]);
        for (int k = 0; k < 3; ++k) cen[3 * i + k] = 0.5f * (blo[3 * i + k] + bhi[3 * i + k]);
    }
    sc.nodes.clear();
    sc.nodes.reserve(2 * (size_t)std::max(n, 1));
    struct Task { int node, first, count; };
    std::vector<Task> stack;
    sc.nodes.push_back(Node{});
    stack.push_back({0, 0, n});
    while (!stack.empty()) {
        Task t = stack.back();
        stack.pop_back();
        Node nd{};
        for (int k = 0; k < 3; ++k) { nd.lo[k] = INFINITY; nd.hi[k] = -INFINITY; }
        float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int i = t.first; i < t.first + t.count; ++i) {
            int id = sc.order[i];
            for (int k = 0; k < 3; ++k) {
                nd.lo[k] = std::min(nd.lo[k], blo[3 * id + k]);
                nd.hi[k] = std::max(nd.hi[k], bhi[3 * id + k]);
                clo[k] = std::min(clo[k], cen[3 * id + k]);
                chi[k] = std::max(chi[k], cen[3 * id + k]);
            }
        }
        nd.left = -1;
        nd.first = t.first;
        nd.count = t.count;
        int split = -1;
        if (t.count > 4) {
            // pick the axis/bin with the lowest SAH cost
            const int B = 16;
            double best = INFINITY;
            int best_axis = -1, best_bin = -1;
            for (int k = 0; k < 3; ++k) {
                float ext = chi[k] - clo[k];
                if (!(ext > 0.0f)) continue;
                int cnt[B] = {0};
                float bl[B][3], bh[B][3];
                for (int b = 0; b < B; ++b)
                    for (int a = 0; a < 3; ++a) { bl[b][a] = INFINITY; bh[b][a] = -INFINITY; }
                for (int i = t.first; i < t.first + t.count; ++i) {
                    int id = sc.order[i];
                    int b = std::min(B - 1, (int)((cen[3 * id + k] - clo[k]) / ext * B));
                    cnt[b]++;
                    for (int a = 0; a < 3; ++a) {
                        bl[b][a] = std::min(bl[b][a], blo[3 * id + a]);
                        bh[b][a] = std::max(bh[b][a], bhi[3 * id + a]);
                    }
                }
                auto area = [](const float l[3], const float h[3]) {
                    double dx = h[0] - l[0], dy = h[1] - l[1], dz = h[2] - l[2];
                    return dx * dy + dy * dz + dz * dx;
                };
                double right_area[B];
                int right_cnt[B];
                float rl[3] = {INFINITY, INFINITY, INFINITY}, rh[3] = {-INFINITY, -INFINITY, -INFINITY};
                int rc = 0;
                for (int b = B - 1; b > 0; --b) {
                    rc += cnt[b];
                    for (int a = 0; a < 3; ++a) { rl[a] = std::min(rl[a], bl[b][a]); rh[a] = std::max(rh[a], bh[b][a]); }
                    right_area[b] = rc ? area(rl, rh) : 0.0;
                    right_cnt[b] = rc;
                }
                float ll[3] = {INFINITY, INFINITY, INFINITY}, lh[3] = {-INFINITY, -INFINITY, -INFINITY};
                int lc = 0;
                for (int b = 0; b < B - 1; ++b) {
                    lc += cnt[b];
                    for (int a = 0; a < 3; ++a) { ll[a] = std::min(ll[a], bl[b][a]); lh[a] = std::max(lh[a], bh[b][a]); }
                    if (lc == 0 || right_cnt[b + 1] == 0) continue;
                    double c = (lc ? area(ll, lh) * lc : 0.0) + right_area[b + 1] * right_cnt[b + 1];
                    if (c < best) { best = c; best_axis = k; best_bin = b; }
                }
            }
            int mid;
            if (best_axis >= 0) {
                int k = best_axis;
                float ext = chi[k] - clo[k];
                auto it = std::partition(sc.order.begin() + t.first, sc.order.begin() + t.first + t.count,
                                         [&](int id) {
                                             int b = std::min(B - 1, (int)((cen[3 * id + k] - clo[k]) / ext * B));
                                             return b <= best_bin;
                                         });
                mid = (int)(it - sc.order.begin());
            } else {
                mid = t.first + t.count / 2;  // all centroids equal: split by count
            }
            if (mid == t.first || mid == t.first + t.count) mid = t.first + t.count / 2;
            split = mid;
        }
        if (split >= 0) {
            nd.left = (int)sc.nodes.size();
            sc.nodes.push_back(Node{});
            sc.nodes.push_back(Node{});
            stack.push_back({nd.left, t.first, split - t.first});
            stack.push_back({nd.left + 1, split, t.first + t.count - split});
        }
        sc.nodes[t.node] = nd;
    }
}

// Slab test in double against the box padded by 1e-5 (1 + max |coordinate|
// of box and origin).  The padding exceeds the lateral error of the fp32
// watertight test (about 2^-23 times the distance), so every triangle that
// test could accept is visited: the walk is equivalent to testing all
// triangles, for both the fp32 and the FP64 queries.
bool box_hit_d(const Node& nd, D3 o, D3 inv, double tmin, double tmax, double& tnear)
{
    double m = std::max(std::max(std::fabs(o.x), std::fabs(o.y)), std::fabs(o.z));
    for (int k = 0; k < 3; ++k) m = std::max(m, std::max(std::fabs((double)nd.lo[k]), std::fabs((double)nd.hi[k])));
    const double pad = 1e-5 * (1.0 + m);
    double t0 = tmin, t1 = tmax;
    for (int k = 0; k < 3; ++k) {
        double a = ((nd.lo[k] - pad) - atd(o, k)) * atd(inv, k);
        double b = ((nd.hi[k] + pad) - atd(o, k)) * atd(inv, k);
        if (a > b) std::swap(a, b);
        double sl = 1e-12 * (std::fabs(a) + std::fabs(b)) + 1e-300;
        t0 = std::max(t0, a - sl);
        t1 = std::min(t1, b + sl);
    }
    tnear = t0;
    return t0 <= t1;
}

D3 safe_inv(D3 d)
{
    auto f = [](double x) { return std::fabs(x) < 1e-300 ? (x < 0 ? -1e300 : 1e300) : 1.0 / x; };
    return D3{f(d.x), f(d.y), f(d.z)};
}

// Visit every leaf whose box the ray can reach within [tmin, *tmax];
// *tmax may shrink during the walk (closest-first order is not required).
template <class F>
void walk(const Scene& sc, D3 o, D3 d, double tmin, const double* tmax, F&& leaf)
{
    if (sc.nodes.empty() || sc.tris.empty()) return;
    D3 inv = safe_inv(d);
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const Node& nd = sc.nodes[stack[--sp]];
        double tn;
        if (!box_hit_d(nd, o, inv, tmin, *tmax, tn)) continue;
        if (nd.left < 0) {
            for (int i = nd.first; i < nd.first + nd.count; ++i) leaf(sc.order[i]);
        } else {
            stack[sp++] = nd.left + 1;
            stack[sp++] = nd.left;
        }
    }
}

// ----------------------------------------------------------- FP64 triangle
struct Hit64 {
    bool hit;
    double t, u, v;
};

Hit64 tri_test_d(D3 o, D3 d, const Tri& tr, double tmin, double tmax)
{
    Hit64 h{false, 0, 0, 0};
    double ax_ = std::fabs(d.x), ay_ = std::fabs(d.y), az_ = std::fabs(d.z);
    int kz = 0;
    double m = ax_;
    if (ay_ > m) { kz = 1; m = ay_; }
    if (az_ > m) kz = 2;
    int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if (atd(d, kz) < 0.0) std::swap(kx, ky);
    double sx = atd(d, kx) / atd(d, kz), sy = atd(d, ky) / atd(d, kz), sz = 1.0 / atd(d, kz);
    D3 a = d3(tr.p0) - o, b = d3(tr.p1) - o, c = d3(tr.p2) - o;
    double Ax = atd(a, kx) - sx * atd(a, kz), Ay = atd(a, ky) - sy * atd(a, kz);
    double Bx = atd(b, kx) - sx * atd(b, kz), By = atd(b, ky) - sy * atd(b, kz);
    double Cx = atd(c, kx) - sx * atd(c, kz), Cy = atd(c, ky) - sy * atd(c, kz);
    double U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if ((U < 0 || V < 0 || W < 0) && (U > 0 || V > 0 || W > 0)) return h;
    double det = U + V + W;
    if (det == 0.0) return h;
    double T = sz * (U * atd(a, kz) + V * atd(b, kz) + W * atd(c, kz));
    double t = T / det;
    if (!(t >= tmin && t <= tmax)) return h;
    h.hit = true;
    h.t = t;
    h.u = V / det;
    h.v = W / det;
    return h;
}

// distance from the hit point (barycentrics u, v) to the nearest edge
double edge_distance(const Tri& tr, double u, double v)
{
    D3 p0 = d3(tr.p0), p1 = d3(tr.p1), p2 = d3(tr.p2);
    D3 n = crossd(p1 - p0, p2 - p0);
    double a2 = std::sqrt(dotd(n, n));  // twice the area
    double w = 1.0 - u - v;
    auto len = [](D3 x) { return std::sqrt(dotd(x, x)); };
    double h0 = a2 / std::max(len(p2 - p1), 1e-300);  // opposite p0, weight w
    double h1 = a2 / std::max(len(p0 - p2), 1e-300);  // opposite p1, weight u
    double h2 = a2 / std::max(len(p1 - p0), 1e-300);  // opposite p2, weight v
    return std::min(std::min(std::fabs(w) * h0, std::fabs(u) * h1), std::fabs(v) * h2);
}

// shortest distance between the ray's line and the triangle's edges
double ray_edge_distance(D3 o, D3 d, const Tri& tr)
{
    D3 p[3] = {d3(tr.p0), d3(tr.p1), d3(tr.p2)};
    double best = INFINITY;
    for (int e = 0; e < 3; ++e) {
        D3 a = p[e], b = p[(e + 1) % 3];
        D3 u = b - a;
        D3 w0 = o - a;
        double A = dotd(d, d), B = dotd(d, u), C = dotd(u, u), D = dotd(d, w0), E = dotd(u, w0);
        double den = A * C - B * B;
        double sc_, tc;
        if (den < 1e-300) { sc_ = 0.0; tc = C > 0 ? E / C : 0.0; }
        else { sc_ = (B * E - C * D) / den; tc = (A * E - B * D) / den; }
        tc = std::min(1.0, std::max(0.0, tc));
        sc_ = std::max(0.0, (B * tc - D) / std::max(A, 1e-300));
        D3 q = (o + d * sc_) - (a + u * tc);
        best = std::min(best, std::sqrt(dotd(q, q)));
    }
    return best;
}

struct Closest64 {
    int prim = -1;
    double t = INFINITY, u = 0, v = 0;
    int prim2 = -1;        // runner-up distinct triangle
    double t2 = INFINITY;
};

Closest64 closest_fp64(const Scene& sc, V3 of, V3 df, float tminf, float tmaxf, bool want_second)
{
    D3 o = d3(of), d = d3(df);
    Closest64 r;
    double bound = tmaxf;
    walk(sc, o, d, tminf, &bound, [&](int id) {
        Hit64 h = tri_test_d(o, d, sc.tris[id], tminf, tmaxf);
        if (!h.hit) return;
        bool better = h.t < r.t || (h.t == r.t && id < r.prim);
        if (better) {
            if (r.prim >= 0) { r.prim2 = r.prim; r.t2 = r.t; }
            r.prim = id; r.t = h.t; r.u = h.u; r.v = h.v;
        } else if (h.t < r.t2 || (h.t == r.t2 && id < r.prim2)) {
            r.prim2 = id; r.t2 = h.t;
        }
        bound = want_second ? std::min((double)tmaxf, r.t2) : std::min((double)tmaxf, r.t);
    });
    return r;
}

// ----------------------------------------------------- kernel-numerics fp32
struct WRay {
    int kx, ky, kz;
    float sx, sy, sz;
};

WRay wsetup(V3 d)
{
    WRay w;
    float ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    int kz = 0;
    float m = ax;
    if (ay > m) { kz = 1; m = ay; }
    if (az > m) kz = 2;
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    float dz = at(d, kz);
    if (dz < 0.0f) std::swap(kx, ky);
    w.kx = kx; w.ky = ky; w.kz = kz;
    w.sx = at(d, kx) / dz;
    w.sy = at(d, ky) / dz;
    w.sz = 1.0f / dz;
    return w;
}

// Plane distance in double, rounded to float: t = ((p0 - o) . n) / (d . n),
// n = (p1 - p0) x (p2 - p0).  Differences of floats and their products are
// exact in double; each sum rounds once.  False if the plane is parallel to
// the ray.
bool plane_t(const Tri& tr, V3 o, V3 d, float& t_out)
{
    double e1x = (double)tr.p1.x - (double)tr.p0.x, e1y = (double)tr.p1.y - (double)tr.p0.y,
           e1z = (double)tr.p1.z - (double)tr.p0.z;
    double e2x = (double)tr.p2.x - (double)tr.p0.x, e2y = (double)tr.p2.y - (double)tr.p0.y,
           e2z = (double)tr.p2.z - (double)tr.p0.z;
    double nx = e1y * e2z - e1z * e2y;
    double ny = e1z * e2x - e1x * e2z;
    double nz = e1x * e2y - e1y * e2x;
    double wx = (double)tr.p0.x - (double)o.x, wy = (double)tr.p0.y - (double)o.y,
           wz = (double)tr.p0.z - (double)o.z;
    double num = (wx * nx + wy * ny) + wz * nz;
    double den = ((double)d.x * nx + (double)d.y * ny) + (double)d.z * nz;
    if (den == 0.0) return false;
    t_out = (float)(num / den);
    return true;
}

// Kernel numerics: the watertight test (Woop, Benthin, Wald 2013) in fp32,
// a zero edge function recomputed in double, decides whether the ray's line
// crosses the triangle; the hit distance is the double plane distance, and
// the hit counts if tmin <= t <= tmax.
bool tri_test_f(const WRay& w, V3 org, V3 dir, const Tri& tr, float tmin, float tmax, float& t_out, float& vn,
                float& wn, float& det_out)
{
    V3 a = tr.p0 - org, b = tr.p1 - org, c = tr.p2 - org;
    float az_ = at(a, w.kz), bz_ = at(b, w.kz), cz_ = at(c, w.kz);
    float ax = at(a, w.kx) - w.sx * az_, ay = at(a, w.ky) - w.sy * az_;
    float bx = at(b, w.kx) - w.sx * bz_, by = at(b, w.ky) - w.sy * bz_;
    float cx = at(c, w.kx) - w.sx * cz_, cy = at(c, w.ky) - w.sy * cz_;
    float U = cx * by - cy * bx, V = ax * cy - ay * cx, W = bx * ay - by * ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)cx * (double)by - (double)cy * (double)bx);
        V = (float)((double)ax * (double)cy - (double)ay * (double)cx);
        W = (float)((double)bx * (double)ay - (double)by * (double)ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = (U + V) + W;
    if (det == 0.0f) return false;
    float t;
    if (!plane_t(tr, org, dir, t)) return false;
    if (!(t >= tmin && t <= tmax)) return false;
    t_out = t; vn = V; wn = W; det_out = det;
    return true;
}

struct Closest32 {
    int prim = -1;
    float t = 0, u = 0, v = 0;
};

Closest32 closest_fp32(const Scene& sc, V3 o, V3 d, float tmin, float tmax)
{
    Closest32 r;
    r.t = tmax;
    WRay w = wsetup(d);
    double bound = tmax;
    walk(sc, d3(o), d3(d), tmin, &bound, [&](int id) {
        float t, vn, wn, det;
        if (!tri_test_f(w, o, d, sc.tris[id], tmin, r.t, t, vn, wn, det)) return;
        if (t < r.t || r.prim < 0 || (t == r.t && id < r.prim)) {
            r.t = t; r.prim = id; r.u = vn / det; r.v = wn / det;
            bound = std::max((double)t * (1.0 + 1e-6), (double)t);
        }
    });
    return r;
}

// Any-hit query in kernel fp32 numerics: the distance of a triangle that the
// watertight test accepts with tmin <= t <= tmax, or -1.  Whether a blocker
// exists does not depend on the visiting order; which one is reported does.
float occluded_fp32(const Scene& sc, V3 o, V3 d, float tmin, float tmax)
{
    WRay w = wsetup(d);
    double bound = tmax;
    float found = -1.0f;
    walk(sc, d3(o), d3(d), tmin, &bound, [&](int id) {
        if (found >= 0.0f) return;
        float t, vn, wn, det;
        if (!tri_test_f(w, o, d, sc.tris[id], tmin, tmax, t, vn, wn, det)) return;
        found = t;
        bound = -1.0;  // no further box can be entered
    });
    return found;
}

// Ordered fp32 traversal for the CPU timing baseline: float slab tests with
// the far bound widened by 2^-20 (robust against rounding), near child first,
// culling by the current closest hit.  Same triangle test and (t, primId)
// rule as closest_fp32; its answer can differ only for rays grazing a box
// within float rounding.
Closest32 closest_fp32_fast(const Scene& sc, V3 o, V3 d)
{
    Closest32 r;
    r.t = 3.402823466e38f;
    if (sc.nodes.empty() || sc.tris.empty()) return r;
    WRay w = wsetup(d);
    auto sinv = [](float x) { return 1.0f / (std::fabs(x) < 1e-20f ? (x < 0 ? -1e-20f : 1e-20f) : x); };
    const float ix = sinv(d.x), iy = sinv(d.y), iz = sinv(d.z);
    const float ox = o.x, oy = o.y, oz = o.z;
    auto slab = [&](const Node& nd, float& tn) {
        float ax = (nd.lo[0] - ox) * ix, bx = (nd.hi[0] - ox) * ix;
        float ay = (nd.lo[1] - oy) * iy, by = (nd.hi[1] - oy) * iy;
        float az = (nd.lo[2] - oz) * iz, bz = (nd.hi[2] - oz) * iz;
        float t0 = std::max(std::max(std::min(ax, bx), std::min(ay, by)), std::max(std::min(az, bz), 0.0f));
        float t1 = std::min(std::min(std::max(ax, bx), std::max(ay, by)), std::max(az, bz));
        tn = t0;
        return t0 <= t1 * 1.00000095367431641f && t0 <= r.t;
    };
    struct Ent { int node; float t; };
    Ent stack[128];
    int sp = 0;
    float tn0;
    if (!slab(sc.nodes[0], tn0)) return r;
    stack[sp++] = {0, tn0};
    while (sp) {
        Ent e = stack[--sp];
        if (e.t > r.t) continue;
        const Node& nd = sc.nodes[e.node];
        if (nd.left < 0) {
            for (int i = nd.first; i < nd.first + nd.count; ++i) {
                int id = sc.order[i];
                float t, vn, wn, det;
                if (!tri_test_f(w, o, d, sc.tris[id], 0.0f, r.t, t, vn, wn, det)) continue;
                if (t < r.t || r.prim < 0 || (t == r.t && id < r.prim)) {
                    r.t = t; r.prim = id; r.u = vn / det; r.v = wn / det;
                }
            }
            continue;
        }
        float ta, tb;
        bool ha = slab(sc.nodes[nd.left], ta), hb = slab(sc.nodes[nd.left + 1], tb);
        if (ha && hb) {
            if (ta <= tb) { stack[sp++] = {nd.left + 1, tb}; stack[sp++] = {nd.left, ta}; }
            else { stack[sp++] = {nd.left, ta}; stack[sp++] = {nd.left + 1, tb}; }
        } else if (ha) {
            stack[sp++] = {nd.left, ta};
        } else if (hb) {
            stack[sp++] = {nd.left + 1, tb};
        }
    }
    return r;
}

// ----------------------------------------------------------------- camera
struct Cam {
    V3 origin, image_o, image_dx, image_dy, basis_u, basis_v;
    float lens_radius, focal_distance;
};

// 20 floats: origin, imageO, imageDx, imageDy, basisU, basisV, lensRadius,
// focalDistance (the prt_camera layout)
Cam cam_from(const float* c)
{
    return Cam{v3(c[0], c[1], c[2]), v3(c[3], c[4], c[5]), v3(c[6], c[7], c[8]), v3(c[9], c[10], c[11]),
               v3(c[12], c[13], c[14]), v3(c[15], c[16], c[17]), c[18], c[19]};
}

V3 cam_dir(const Cam& c, float ix, float iy)
{
    V3 d;
    d.x = (c.image_o.x + ix * c.image_dx.x) + iy * c.image_dy.x;
    d.y = (c.image_o.y + ix * c.image_dx.y) + iy * c.image_dy.y;
    d.z = (c.image_o.z + ix * c.image_dx.z) + iy * c.image_dy.z;
    return normalized(d);
}

// Camera ray: PinholeCamera::getRay (camera/pinhole_camera.h:50-54) when the
// lens radius is 0, else ThinLensCamera::getRay (camera/thin_lens_camera.h:38-44):
// lens = uniformSampleDisk(l1, l2) * R (sampling/shape_sampler.h:115-124),
// begin = origin + basis * (lens, 0), end = origin + f * (imageO + ix dx + iy dy).
// numerics 0: the kernel's sin/cos of 2 pi l2; 1: libm sinf/cosf.
void cam_ray(const Cam& c, float ix, float iy, float l1, float l2, int numerics, V3& org, V3& dir)
{
    if (c.lens_radius == 0.0f) {
        org = c.origin;
        dir = cam_dir(c, ix, iy);
        return;
    }
    float r = std::sqrt(l1), sn, cs;
    if (numerics == 0) {
        sincos_2pi_kernel(l2, sn, cs);
    } else {
        float theta = 2.0f * 3.14159265358979323846f * l2;
        sn = std::sin(theta);
        cs = std::cos(theta);
    }
    float lx = (r * cs) * c.lens_radius, ly = (r * sn) * c.lens_radius;
    V3 b = v3(c.origin.x + (lx * c.basis_u.x + ly * c.basis_v.x), c.origin.y + (lx * c.basis_u.y + ly * c.basis_v.y),
              c.origin.z + (lx * c.basis_u.z + ly * c.basis_v.z));
    V3 e = v3(c.origin.x + c.focal_distance * ((c.image_o.x + ix * c.image_dx.x) + iy * c.image_dy.x),
              c.origin.y + c.focal_distance * ((c.image_o.y + ix * c.image_dx.y) + iy * c.image_dy.y),
              c.origin.z + c.focal_distance * ((c.image_o.z + ix * c.image_dx.z) + iy * c.image_dy.z));
    org = b;
    dir = normalized(v3(e.x - b.x, e.y - b.y, e.z - b.z));
}

// ----------------------------------------------------------------- shading
V3 facing_normal(const Tri& tr, V3 dir)
{
    V3 ng = normalized(cross(tr.p1 - tr.p0, tr.p2 - tr.p0));
    if (dot(ng, dir) > 0.0f) ng = v3(-ng.x, -ng.y, -ng.z);
    return ng;
}

V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }

struct Frame {
    V3 u, v, n;
};

// makeFrame (math/basis3.h:128-136)
Frame make_frame(V3 n)
{
    V3 u0 = v3(0.0f, n.z, -n.y), u1 = v3(-n.z, 0.0f, n.x);
    Frame f;
    f.u = normalized(dot(u0, u0) > dot(u1, u1) ? u0 : u1);
    f.v = normalized(cross(n, f.u));
    f.n = n;
    return f;
}

// Shading frame of a hit on triangle `prim` at hit coordinates (hu, hv)
// (p = (1-hu-hv) p0 + hu p1 + hv p2) for a ray along dir.  fast or a mesh
// without attributes: SimpleShadingContext, makeFrame of the ray-facing
// geometric normal (geometry/triangle_mesh.cpp:248-273).  Otherwise the
// ShadingContext of TriangleMesh::postIntersect (geometry/triangle_mesh.cpp:
// 59-153): N = normalize(b0 n0 + b1 n1 + b2 n2) turned to the side of Ng
// (Ng without normals), both flipped if Ng faces along the ray; with texture
// coordinates U = normalize(dp/du), V = normalize(N x U), U = V x N, unless
// the mapping's determinant is 0 or N x U vanishes (then makeFrame(N)).
Frame hit_frame(const Scene& sc, int prim, float hu, float hv, V3 dir, bool fast)
{
    const Tri& tr = sc.tris[prim];
    const bool hn = !sc.normals.empty() && !fast, ht = !sc.texcoords.empty() && !fast;
    if (!hn && !ht) return make_frame(facing_normal(tr, dir));
    const int i0 = sc.idx[3 * (size_t)prim], i1 = sc.idx[3 * (size_t)prim + 1], i2 = sc.idx[3 * (size_t)prim + 2];
    const float b1 = hu, b2 = hv;
    const float b0 = 1.0f - b1 - b2;
    const V3 dp1 = tr.p1 - tr.p0, dp2 = tr.p2 - tr.p0;
    const V3 ng = normalized(cross(dp1, dp2));
    V3 n = ng;
    if (hn) {
        const V3 n0 = sc.normals[i0], n1 = sc.normals[i1], n2 = sc.normals[i2];
        n = normalized(v3((b0 * n0.x + b1 * n1.x) + b2 * n2.x, (b0 * n0.y + b1 * n1.y) + b2 * n2.y,
                          (b0 * n0.z + b1 * n1.z) + b2 * n2.z));
        if (dot(n, ng) < 0.0f) n = neg(n);
    }
    if (dot(ng, dir) > 0.0f) n = neg(n);
    if (ht) {
        const float* t0 = &sc.texcoords[2 * (size_t)i0];
        const float* t1 = &sc.texcoords[2 * (size_t)i1];
        const float* t2 = &sc.texcoords[2 * (size_t)i2];
        const float du1 = t1[0] - t0[0], du2 = t2[0] - t0[0];
        const float dv1 = t1[1] - t0[1], dv2 = t2[1] - t0[1];
        const float det = du1 * dv2 - dv1 * du2;
        if (det != 0.0f) {
            const float inv = 1.0f / det;
            V3 dpdu = v3((dv2 * dp1.x - dv1 * dp2.x) * inv, (dv2 * dp1.y - dv1 * dp2.y) * inv,
                         (dv2 * dp1.z - dv1 * dp2.z) * inv);
            Frame f;
            f.n = n;
            f.u = normalized(dpdu);
            f.v = cross(n, f.u);
            if (dot(f.v, f.v) > 0.0f) {
                f.v = normalized(f.v);
                f.u = cross(f.v, n);
                return f;
            }
        }
    }
    return make_frame(n);
}

// Diffuse bounce in frame f (render/diffuse_integrator_single.h:55-62,
// core/ray.h:71-76, sampling/shape_sampler.h:50-74).
// numerics 0: kernel fp32 (polynomial sin/cos); 1: reference (libm sinf/cosf)
void bounce(V3 org, V3 dir, float t, const Frame& f, float s1, float s2, int numerics, V3& no, V3& nd)
{
    V3 p = v3(org.x + t * dir.x, org.y + t * dir.y, org.z + t * dir.z);
    float eps = std::max(t, std::max(std::max(std::fabs(p.x), std::fabs(p.y)), std::fabs(p.z))) * 0x1.fp-18f;
    float ct = std::sqrt(s1), st = std::sqrt(1.0f - s1);
    float sp, cp;
    if (numerics == 0) {
        sincos_2pi_kernel(s2, sp, cp);
    } else {
        float phi = 2.0f * 3.14159265358979323846f * s2;
        sp = std::sin(phi);
        cp = std::cos(phi);
    }
    float lx = cp * st, ly = sp * st, lz = ct;
    const V3 u = f.u, v = f.v, n = f.n;
    nd = v3((lx * u.x + ly * v.x) + lz * n.x, (lx * u.y + ly * v.y) + lz * n.y, (lx * u.z + ly * v.z) + lz * n.z);
    no = v3(p.x + eps * nd.x, p.y + eps * nd.y, p.z + eps * nd.z);
}

void bounce(V3 org, V3 dir, float t, V3 n, float s1, float s2, int numerics, V3& no, V3& nd)
{
    bounce(org, dir, t, make_frame(n), s1, s2, numerics, no, nd);
}

struct HitAny {
    int prim;
    float t, u, v;
};

HitAny closest_any(const Scene& sc, V3 o, V3 d, int numerics)
{
    if (numerics == 0) {
        Closest32 c = closest_fp32(sc, o, d, 0.0f, 3.402823466e38f);
        return HitAny{c.prim, c.t, c.u, c.v};
    }
    Closest64 c = closest_fp64(sc, o, d, 0.0f, 3.402823466e38f, false);
    return HitAny{c.prim, (float)c.t, (float)c.u, (float)c.v};
}

bool occluded_any(const Scene& sc, V3 o, V3 d, int numerics)
{
    if (numerics == 0) return occluded_fp32(sc, o, d, 0.0f, 3.402823466e38f) >= 0.0f;
    return closest_fp64(sc, o, d, 0.0f, 3.402823466e38f, false).prim >= 0;
}

template <class F>
void parallel_tiles(int ntiles, int threads, F&& f)
{
    std::atomic<int> next{0};
    auto worker = [&]() {
        for (;;) {
            int t = next.fetch_add(1);
            if (t >= ntiles) break;
            f(t);
        }
    };
    if (threads <= 1) { worker(); return; }
    std::vector<std::thread> pool;
    for (int i = 0; i < threads; ++i) pool.emplace_back(worker);
    for (auto& th : pool) th.join();
}

int default_threads(int threads)
{
    if (threads > 0) return threads;
    unsigned h = std::thread::hardware_concurrency();
    return h ? (int)h : 1;
}

}  // namespace

extern "C" {

uint32_t orc_sampler_seed(int pass, int pixel) { return seed_of(pass, pixel); }

// n draws after the seed: states[i] and the unorm value of each
void orc_sampler_sequence(int pass, int pixel, int n, uint32_t* states, float* values)
{
    Sampler s{seed_of(pass, pixel)};
    for (int i = 0; i < n; ++i) {
        float v = s.next();
        if (states) states[i] = s.s;
        if (values) values[i] = v;
    }
}

void orc_sincos_2pi(const float* v, int64_t n, float* s, float* c)
{
    for (int64_t i = 0; i < n; ++i) sincos_2pi_kernel(v[i], s[i], c[i]);
}

// Pinhole camera of a view (render_window.cpp:307-325 + pinhole_camera.h:34-48):
// basis = Basis3(one).rotateV(angleY).rotateU(angleX); out = origin, imageO,
// imageDx, imageDy, basisU, basisV (18 floats).
void orc_camera_from_view(const float* pos, float angle_x, float angle_y, float fov_y, float aspect, float* out)
{
    auto rot = [](V3 vec, V3 axis, float ang) {
        float c = std::cos(ang), s = std::sin(ang);
        return (vec * c + (axis * dot(axis, vec)) * (1.0f - c)) + cross(vec, axis) * s;
    };
    V3 V = v3(0, 1, 0), N = v3(0, 0, 1);
    // rotateV(angleY): N' = rotate(N, V, a), U' = cross(V, N')
    V3 N1 = rot(N, V, angle_y);
    V3 U1 = cross(V, N1);
    V3 V1 = V;
    // rotateU(angleX): V'' = rotate(V1, U1, a), N'' = cross(U1, V'')
    V3 V2 = rot(V1, U1, angle_x);
    V3 N2 = cross(U1, V2);
    V3 U2 = U1;
    float hh = std::tan(0.5f * fov_y);
    float hw = hh * aspect;
    auto tog = [&](V3 a) { return U2 * a.x + V2 * a.y + N2 * a.z; };
    V3 io = tog(v3(-hw, hh, -1.0f));
    V3 idx = tog(v3(2.0f * hw, 0.0f, 0.0f));
    V3 idy = tog(v3(0.0f, (float)(-2.0 * hh), 0.0f));
    float vals[20] = {pos[0], pos[1], pos[2], io.x, io.y, io.z, idx.x, idx.y, idx.z,
                      idy.x, idy.y, idy.z, U2.x, U2.y, U2.z, V2.x, V2.y, V2.z, 0.0f, 1.0f};
    std::memcpy(out, vals, sizeof(vals));
}

// Primary rays of one pass, row-major pixel order, 8 floats per ray
// (org, tnear, dir, tfar), plus the sampler state after the 4 camera draws.
void orc_raygen(const float* cam, int w, int h, int pass, float* rays, uint32_t* states)
{
    Cam c = cam_from(cam);
    float inv_w = 1.0f / (float)w, inv_h = 1.0f / (float)h;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            int pid = x + y * w;
            Sampler s{seed_of(pass, pid)};
            float u1 = s.next(), u2 = s.next();
            float l1 = s.next(), l2 = s.next();
            V3 o, d;
            cam_ray(c, ((float)x + u1) * inv_w, ((float)y + u2) * inv_h, l1, l2, 0, o, d);
            float* r = rays + 8 * (size_t)pid;
            r[0] = o.x; r[1] = o.y; r[2] = o.z; r[3] = 0.0f;
            r[4] = d.x; r[5] = d.y; r[6] = d.z; r[7] = 3.402823466e38f;
            if (states) states[pid] = s.s;
        }
}

void* orc_scene_create(const int32_t* idx, int64_t tri_count, const float* pos, int64_t vertex_count)
{
    Scene* sc = new Scene();
    sc->tris.resize(tri_count);
    for (int64_t i = 0; i < tri_count; ++i) {
        const float* a = pos + 3 * (size_t)idx[3 * i];
        const float* b = pos + 3 * (size_t)idx[3 * i + 1];
        const float* c = pos + 3 * (size_t)idx[3 * i + 2];
        sc->tris[i] = Tri{v3(a[0], a[1], a[2]), v3(b[0], b[1], b[2]), v3(c[0], c[1], c[2])};
    }
    sc->idx.assign(idx, idx + 3 * tri_count);
    sc->vertex_count = vertex_count;
    build_bvh(*sc);
    return sc;
}

// Per-vertex shading attributes: normals (3 floats per vertex) and texture
// coordinates (2 floats per vertex); null removes the attribute.
void orc_scene_attributes(void* scp, const float* normals, const float* texcoords)
{
    Scene& sc = *static_cast<Scene*>(scp);
    sc.normals.clear();
    sc.texcoords.clear();
    if (normals)
        for (int64_t i = 0; i < sc.vertex_count; ++i)
            sc.normals.push_back(v3(normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]));
    if (texcoords) sc.texcoords.assign(texcoords, texcoords + 2 * sc.vertex_count);
}

// Shading frame of hit (prim, u, v) for a ray along dir: out = U, V, N.
void orc_shading_frame(void* scp, int prim, float u, float v, const float* dir, int fast, float* out)
{
    const Frame f = hit_frame(*static_cast<Scene*>(scp), prim, u, v, v3(dir[0], dir[1], dir[2]), fast != 0);
    const V3 c[3] = {f.u, f.v, f.n};
    for (int k = 0; k < 3; ++k) {
        out[3 * k] = c[k].x;
        out[3 * k + 1] = c[k].y;
        out[3 * k + 2] = c[k].z;
    }
}

void orc_scene_destroy(void* sc) { delete static_cast<Scene*>(sc); }

// FP64 closest hit.  rays: 8 floats each; out per ray: prim, t (double),
// u, v, distance of the hit point to the nearest edge, runner-up prim and t.
void orc_intersect_fp64(void* scp, const float* rays, int64_t n, int32_t* prim, double* t, double* u,
                        double* v, double* edge_dist, int32_t* prim2, double* t2, int threads)
{
    const Scene& sc = *static_cast<Scene*>(scp);
    int nt = default_threads(threads);
    const int64_t chunk = 1024;
    int ntiles = (int)((n + chunk - 1) / chunk);
    parallel_tiles(ntiles, nt, [&](int tile) {
        for (int64_t i = tile * chunk; i < std::min(n, (tile + 1) * chunk); ++i) {
            const float* r = rays + 8 * (size_t)i;
            if (r[7] < 0.0f) {
                prim[i] = -1;
                t[i] = -1;
                u[i] = v[i] = 0;
                edge_dist[i] = INFINITY;
                if (prim2) prim2[i] = -1;
                if (t2) t2[i] = INFINITY;
                continue;
            }
            Closest64 c = closest_fp64(sc, v3(r[0], r[1], r[2]), v3(r[4], r[5], r[6]), r[3], r[7], true);
            prim[i] = c.prim;
            t[i] = c.prim >= 0 ? c.t : -1.0;
            u[i] = c.u;
            v[i] = c.v;
            edge_dist[i] = c.prim >= 0 ? edge_distance(sc.tris[c.prim], c.u, c.v) : INFINITY;
            if (prim2) prim2[i] = c.prim2;
            if (t2) t2[i] = c.t2;
        }
    });
}

// distance between ray i's line and the edges of triangle prim[i] (FP64)
void orc_ray_edge_distance(void* scp, const float* rays, const int32_t* prims, int64_t n, double* out)
{
    const Scene& sc = *static_cast<Scene*>(scp);
    for (int64_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * (size_t)i;
        out[i] = prims[i] >= 0 ? ray_edge_distance(D3{r[0], r[1], r[2]}, D3{r[4], r[5], r[6]}, sc.tris[prims[i]])
                               : INFINITY;
    }
}

// kernel-numerics fp32 closest hit by testing every triangle (no BVH);
// hits as orc_intersect_fp32.  For cross-checking the BVH walk on small scenes.
void orc_intersect_fp32_brute(void* scp, const float* rays, int64_t n, float* hits)
{
    const Scene& sc = *static_cast<Scene*>(scp);
    for (int64_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * (size_t)i;
        float* h = hits + 4 * (size_t)i;
        V3 o = v3(r[0], r[1], r[2]), d = v3(r[4], r[5], r[6]);
        WRay w = wsetup(d);
        Closest32 c;
        c.t = r[7];
        for (int id = 0; id < (int)sc.tris.size() && r[7] >= 0.0f; ++id) {
            float t, vn, wn, det;
            if (!tri_test_f(w, o, d, sc.tris[id], r[3], c.t, t, vn, wn, det)) continue;
            if (t < c.t || c.prim < 0 || (t == c.t && id < c.prim)) {
                c.t = t; c.prim = id; c.u = vn / det; c.v = wn / det;
            }
        }
        int32_t pid = c.prim;
        if (pid >= 0) { h[0] = c.t; h[2] = c.u; h[3] = c.v; }
        else { h[0] = -1.0f; h[2] = 0.0f; h[3] = 0.0f; }
        std::memcpy(&h[1], &pid, 4);
    }
}

// kernel-numerics fp32 any-hit query over the BVH (brute = 1: every
// triangle); out_t: distance of a blocking triangle or -1 per ray
void orc_occluded_fp32(void* scp, const float* rays, int64_t n, float* out_t, int brute)
{
    const Scene& sc = *static_cast<Scene*>(scp);
    for (int64_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * (size_t)i;
        V3 o = v3(r[0], r[1], r[2]), d = v3(r[4], r[5], r[6]);
        float found = -1.0f;
        if (r[7] >= 0.0f) {
            if (!brute) {
                found = occluded_fp32(sc, o, d, r[3], r[7]);
            } else {
                WRay w = wsetup(d);
                for (int id = 0; id < (int)sc.tris.size() && found < 0.0f; ++id) {
                    float t, vn, wn, det;
                    if (tri_test_f(w, o, d, sc.tris[id], r[3], r[7], t, vn, wn, det)) found = t;
                }
            }
        }
        out_t[i] = found;
    }
}

// kernel-numerics fp32 closest hit; hits: 4 floats (t, prim bits, u, v) each
void orc_intersect_fp32(void* scp, const float* rays, int64_t n, float* hits, int threads)
{
    const Scene& sc = *static_cast<Scene*>(scp);
    int nt = default_threads(threads);
    const int64_t chunk = 1024;
    int ntiles = (int)((n + chunk - 1) / chunk);
    parallel_tiles(ntiles, nt, [&](int tile) {
        for (int64_t i = tile * chunk; i < std::min(n, (tile + 1) * chunk); ++i) {
            const float* r = rays + 8 * (size_t)i;
            float* h = hits + 4 * (size_t)i;
            int32_t pid = -1;
            if (r[7] >= 0.0f) {
                Closest32 c = closest_fp32(sc, v3(r[0], r[1], r[2]), v3(r[4], r[5], r[6]), r[3], r[7]);
                if (c.prim >= 0) {
                    h[0] = c.t;
                    pid = c.prim;
                    h[2] = c.u;
                    h[3] = c.v;
                }
            }
            if (pid < 0) { h[0] = -1.0f; h[2] = 0.0f; h[3] = 0.0f; }
            std::memcpy(&h[1], &pid, 4);
        }
    });
}

// Single-ray renderer (RendererSingle + DiffuseIntegratorSingle /
// PrimaryIntegratorSingle), passes pass0 .. pass0+spp-1, 1 sample per pass.
// accum: 4 floats per pixel (sum r, g, b, weight); sumsq: per-pixel sum of the
// squared grey value (for the Monte Carlo variance); rays_out: intersect calls.
// renderer 0 = primary, 1 = diffuse, 2 = diffuse2 (next-event shadow rays),
// plus 0x100 for the "Fast" variant (SimpleShadingContext, each pass
// overwrites the pixel: render/device_cuda.cpp:79-90,
// render/diffuse_renderer_cuda.cu:109-118);
// numerics 0 = kernel fp32, 1 = reference fp64.
int orc_render_rows(void* scp, const float* cam, int w, int h, int row0, int row1, int pass0, int spp,
                    int max_depth, float albedo, int renderer, int numerics, int threads, float* accum,
                    float* sumsq, int64_t* rays_out);

int orc_render(void* scp, const float* cam, int w, int h, int pass0, int spp, int max_depth, float albedo,
               int renderer, int numerics, int threads, float* accum, float* sumsq, int64_t* rays_out)
{
    return orc_render_rows(scp, cam, w, h, 0, h, pass0, spp, max_depth, albedo, renderer, numerics, threads, accum,
                           sumsq, rays_out);
}

// orc_render restricted to image rows [row0, row1) of the w x h frame (same
// pixel ids, samples and rays as the full frame; other pixels untouched).
int orc_render_rows(void* scp, const float* cam, int w, int h, int row0, int row1, int pass0, int spp,
                    int max_depth, float albedo, int renderer, int numerics, int threads, float* accum,
                    float* sumsq, int64_t* rays_out)
{
    row0 = std::max(0, row0);
    row1 = std::min(h, row1);
    const Scene& sc = *static_cast<Scene*>(scp);
    const bool fast = (renderer & 0x100) != 0;
    renderer &= 0xff;
    Cam c = cam_from(cam);
    float inv_w = 1.0f / (float)w, inv_h = 1.0f / (float)h;
    int nt = default_threads(threads);
    std::atomic<int64_t> total_rays{0};
    int rows_per_tile = 8;
    int ntiles = row1 > row0 ? (row1 - row0 + rows_per_tile - 1) / rows_per_tile : 0;
    for (int pass = pass0; pass < pass0 + spp; ++pass) {
        parallel_tiles(ntiles, nt, [&](int tile) {
            int64_t local_rays = 0;
            for (int y = row0 + tile * rows_per_tile; y < std::min(row1, row0 + (tile + 1) * rows_per_tile); ++y)
                for (int x = 0; x < w; ++x) {
                    int pid = x + y * w;
                    Sampler s{seed_of(pass, pid)};
                    float u1 = s.next(), u2 = s.next();
                    float l1 = s.next(), l2 = s.next();
                    V3 org, dir;
                    cam_ray(c, ((float)x + u1) * inv_w, ((float)y + u2) * inv_h, l1, l2, numerics, org, dir);
                    float col[3];
                    if (renderer == 0) {
                        HitAny hit = closest_any(sc, org, dir, numerics);
                        ++local_rays;
                        if (hit.prim >= 0) {
                            V3 n = hit_frame(sc, hit.prim, hit.u, hit.v, dir, fast).n;
                            col[0] = (n.x + 1.0f) * 0.5f;
                            col[1] = (n.y + 1.0f) * 0.5f;
                            col[2] = (n.z + 1.0f) * 0.5f;
                        } else {
                            col[0] = col[1] = col[2] = 0.05f;
                        }
                    } else if (renderer == 2) {
                        // Diffuse2IntegratorSingle::getColor (render/diffuse2_integrator_single.h:42-87)
                        float lw = 1.0f, L = 0.0f;
                        int depth = 0;
                        for (;;) {
                            HitAny hit = closest_any(sc, org, dir, numerics);
                            ++local_rays;
                            if (hit.prim < 0) {
                                L += lw * (depth > 0 ? 0.5f : 1.0f);
                                break;
                            }
                            if (depth == max_depth) break;
                            lw *= albedo;
                            Frame fr = hit_frame(sc, hit.prim, hit.u, hit.v, dir, fast);
                            float s1 = s.next(), s2 = s.next();
                            V3 so, sd;
                            bounce(org, dir, hit.t, fr, s1, s2, numerics, so, sd);
                            ++local_rays;
                            if (!occluded_any(sc, so, sd, numerics)) L += lw * 0.5f;
                            float s3 = s.next(), s4 = s.next();
                            V3 no, nd;
                            bounce(org, dir, hit.t, fr, s3, s4, numerics, no, nd);
                            org = no;
                            dir = nd;
                            ++depth;
                        }
                        col[0] = col[1] = col[2] = L;
                    } else {
                        float lw = 1.0f;
                        int depth = 0;
                        HitAny hit;
                        for (;;) {
                            hit = closest_any(sc, org, dir, numerics);
                            ++local_rays;
                            if (hit.prim < 0 || depth == max_depth) break;
                            lw *= albedo;
                            Frame fr = hit_frame(sc, hit.prim, hit.u, hit.v, dir, fast);
                            float s1 = s.next(), s2 = s.next();
                            V3 no, nd;
                            bounce(org, dir, hit.t, fr, s1, s2, numerics, no, nd);
                            org = no;
                            dir = nd;
                            ++depth;
                        }
                        float L = hit.prim >= 0 ? 0.0f : lw;
                        col[0] = col[1] = col[2] = L;
                    }
                    float* a = accum + 4 * (size_t)pid;
                    if (fast) a[0] = a[1] = a[2] = a[3] = 0.0f;  // isAccum = false: overwrite
                    a[0] = a[0] + col[0];
                    a[1] = a[1] + col[1];
                    a[2] = a[2] + col[2];
                    a[3] = a[3] + 1.0f;
                    if (sumsq) sumsq[pid] += col[0] * col[0];
                }
            total_rays += local_rays;
        });
    }
    if (rays_out) *rays_out = total_rays.load();
    return 0;
}

// ---------------------------------------------------------------- BVH8 (CPU baseline)
// The CPU timing baseline traverses an 8-wide tree with AVX2, as a stand-in
// for Embree's 8-wide stream traversal (rtcIntersectNp).  Each wide node's
// children are BVH2 nodes of the binned-SAH tree: a wide node opens its
// largest-area internal child until it has 8 children, so every child box is
// a BVH2 node box and the slab test is the one of closest_fp32_fast, per lane.
struct WideNode {
    alignas(32) float lo_x[8], lo_y[8], lo_z[8], hi_x[8], hi_y[8], hi_z[8];
    int32_t child[8];  // >= 0: wide node; < 0: ~(first << 3 | count), the leaf's triangles in WideTree::tris
    int32_t n;         // children in slots 0 .. n-1
};

// The tree's triangles in leaf order (a leaf's <= 4 triangles contiguous),
// with their original ids.
struct WideTri {
    Tri tri;
    int32_t id;
};

struct Scene::WideTree {
    std::vector<WideNode> nodes;
    std::vector<WideTri> tris;
};

void build_bvh8(Scene& sc)
{
    auto* wt = new Scene::WideTree();
    sc.wide = wt;
    if (sc.nodes.empty() || sc.tris.empty()) return;
    wt->tris.reserve(sc.tris.size());
    auto area = [&](int i) {
        const Node& nd = sc.nodes[i];
        double dx = nd.hi[0] - nd.lo[0], dy = nd.hi[1] - nd.lo[1], dz = nd.hi[2] - nd.lo[2];
        return dx * dy + dy * dz + dz * dx;
    };
    // (wide node index, BVH2 node it stands for)
    std::vector<std::pair<int, int>> todo{{0, 0}};
    wt->nodes.push_back(WideNode{});
    while (!todo.empty()) {
        auto [wi, b] = todo.back();
        todo.pop_back();
        int kids[8];
        int n = 0;
        if (sc.nodes[b].left < 0) {
            kids[n++] = b;  // a leaf root
        } else {
            kids[n++] = sc.nodes[b].left;
            kids[n++] = sc.nodes[b].left + 1;
            while (n < 8) {
                int best = -1;
                double ba = -1.0;
                for (int k = 0; k < n; ++k)
                    if (sc.nodes[kids[k]].left >= 0 && area(kids[k]) > ba) { ba = area(kids[k]); best = k; }
                if (best < 0) break;
                const int c = kids[best];
                kids[best] = sc.nodes[c].left;
                kids[n++] = sc.nodes[c].left + 1;
            }
        }
        WideNode w{};
        w.n = n;
        for (int k = 0; k < n; ++k) {
            const Node& nd = sc.nodes[kids[k]];
            w.lo_x[k] = nd.lo[0]; w.lo_y[k] = nd.lo[1]; w.lo_z[k] = nd.lo[2];
            w.hi_x[k] = nd.hi[0]; w.hi_y[k] = nd.hi[1]; w.hi_z[k] = nd.hi[2];
            if (nd.left < 0) {
                const int first = (int)wt->tris.size();
                for (int i = nd.first; i < nd.first + nd.count; ++i)
                    wt->tris.push_back(WideTri{sc.tris[sc.order[i]], sc.order[i]});
                w.child[k] = ~((first << 3) | nd.count);
            } else {
                w.child[k] = (int)wt->nodes.size();
                wt->nodes.push_back(WideNode{});
                todo.push_back({w.child[k], kids[k]});
            }
        }
        wt->nodes[wi] = w;
    }
}

// closest_fp32_fast over the 8-wide tree: per visited wide node, the 8 child
// slab tests in one AVX2 pass (the same sub/mul/min/max per lane as the
// scalar test), hit children ordered nearest first; leaves test their
// triangles with the scalar watertight test and the (t, primId) rule.
__attribute__((target("avx2"))) Closest32 closest_fp32_wide(const Scene& sc, V3 o, V3 d)
{
    Closest32 r;
    r.t = 3.402823466e38f;
    const std::vector<WideNode>& wide = sc.wide->nodes;
    const WideTri* wtris = sc.wide->tris.data();
    if (wide.empty() || sc.tris.empty()) return r;
    WRay w = wsetup(d);
    auto sinv = [](float x) { return 1.0f / (std::fabs(x) < 1e-20f ? (x < 0 ? -1e-20f : 1e-20f) : x); };
    const __m256 ix = _mm256_set1_ps(sinv(d.x)), iy = _mm256_set1_ps(sinv(d.y)), iz = _mm256_set1_ps(sinv(d.z));
    const __m256 ox = _mm256_set1_ps(o.x), oy = _mm256_set1_ps(o.y), oz = _mm256_set1_ps(o.z);
    const __m256 zero = _mm256_setzero_ps(), widen = _mm256_set1_ps(1.00000095367431641f);
    const __m256i lane = _mm256_setr_epi32(0, 1, 2, 3, 4, 5, 6, 7);
    struct Ent { int ref; float t; };
    Ent stack[512];
    int sp = 0;
    stack[sp++] = {0, 0.0f};
    while (sp) {
        const Ent e = stack[--sp];
        if (e.t > r.t) continue;
        if (e.ref < 0) {
            const int first = (~e.ref) >> 3, count = (~e.ref) & 7;
            for (int i = first; i < first + count; ++i) {
                const int id = wtris[i].id;
                float t, vn, wn, det;
                if (!tri_test_f(w, o, d, wtris[i].tri, 0.0f, r.t, t, vn, wn, det)) continue;
                if (t < r.t || r.prim < 0 || (t == r.t && id < r.prim)) {
                    r.t = t; r.prim = id; r.u = vn / det; r.v = wn / det;
                }
            }
            continue;
        }
        const WideNode& nd = wide[e.ref];
        const __m256 ax = _mm256_mul_ps(_mm256_sub_ps(_mm256_load_ps(nd.lo_x), ox), ix);
        const __m256 bx = _mm256_mul_ps(_mm256_sub_ps(_mm256_load_ps(nd.hi_x), ox), ix);
        const __m256 ay = _mm256_mul_ps(_mm256_sub_ps(_mm256_load_ps(nd.lo_y), oy), iy);
        const __m256 by = _mm256_mul_ps(_mm256_sub_ps(_mm256_load_ps(nd.hi_y), oy), iy);
        const __m256 az = _mm256_mul_ps(_mm256_sub_ps(_mm256_load_ps(nd.lo_z), oz), iz);
        const __m256 bz = _mm256_mul_ps(_mm256_sub_ps(_mm256_load_ps(nd.hi_z), oz), iz);
        const __m256 t0 = _mm256_max_ps(_mm256_max_ps(_mm256_min_ps(ax, bx), _mm256_min_ps(ay, by)),
                                        _mm256_max_ps(_mm256_min_ps(az, bz), zero));
        const __m256 t1 = _mm256_min_ps(_mm256_min_ps(_mm256_max_ps(ax, bx), _mm256_max_ps(ay, by)),
                                        _mm256_max_ps(az, bz));
        const __m256 hit = _mm256_and_ps(_mm256_cmp_ps(t0, _mm256_mul_ps(t1, widen), _CMP_LE_OQ),
                                         _mm256_cmp_ps(t0, _mm256_set1_ps(r.t), _CMP_LE_OQ));
        const __m256i valid = _mm256_cmpgt_epi32(_mm256_set1_epi32(nd.n), lane);
        unsigned mask = (unsigned)_mm256_movemask_ps(_mm256_and_ps(hit, _mm256_castsi256_ps(valid)));
        if (!mask) continue;
        alignas(32) float tn[8];
        _mm256_store_ps(tn, t0);
        // hit children, farthest first onto the stack (insertion by distance)
        Ent hits[8];
        int nh = 0;
        while (mask) {
            const int k = __builtin_ctz(mask);
            mask &= mask - 1;
            Ent c{nd.child[k], tn[k]};
            int j = nh++;
            while (j > 0 && hits[j - 1].t < c.t) { hits[j] = hits[j - 1]; --j; }
            hits[j] = c;
        }
        for (int k = 0; k < nh; ++k) stack[sp++] = hits[k];
    }
    return r;
}

Scene::~Scene() { delete wide; }

bool cpu_has_avx2()
{
    static const bool yes = __builtin_cpu_supports("avx2");
    return yes;
}

// CPU timing baseline: the stream renderer (DiffuseRendererStream) over the
// oracle's binned-SAH tree, fp32 kernel numerics.  Each thread takes 64x32
// tiles from an atomic counter, generates the tile's rays as one stream,
// intersects the whole stream, partitions misses before hits, accumulates the
// misses and shades the hits into the next stream.  Renders `frames` passes
// of rows [row0, row1); returns rays traced and wall seconds of the frames.
// simd != 0 (and an AVX2 host): the 8-wide tree with AVX2 child tests
// (closest_fp32_wide); else the scalar BVH2 walk (closest_fp32_fast).
int orc_stream_bench(void* scp, const float* cam, int w, int h, int row0, int row1, int pass0, int frames,
                     int max_depth, float albedo, int threads, int simd, float* accum, int64_t* rays_out,
                     double* seconds_out)
{
    const Scene& sc = *static_cast<Scene*>(scp);
    Cam c = cam_from(cam);
    float inv_w = 1.0f / (float)w, inv_h = 1.0f / (float)h;
    int nt = default_threads(threads);
    const int TX = 64, TY = 32;
    row0 = std::max(0, row0);
    row1 = std::min(h, row1);
    int tiles_x = (w + TX - 1) / TX, tiles_y = (row1 - row0 + TY - 1) / TY;
    std::atomic<int64_t> total{0};
    // the 8-wide tree of the AVX2 traversal, built once per scene, outside the timed region
    const bool wide = cpu_has_avx2() && simd != 0;
    if (wide && !sc.wide) build_bvh8(const_cast<Scene&>(sc));
    auto t0 = std::chrono::steady_clock::now();
    for (int f = 0; f < frames; ++f) {
        int pass = pass0 + f;
        parallel_tiles(tiles_x * tiles_y, nt, [&](int tile) {
            struct RayS { V3 o, d; float t; int prim; int pix; uint32_t rng; };
            std::vector<RayS> in, out, hits;
            in.reserve(TX * TY);
            int tx = tile % tiles_x, ty = tile / tiles_x;
            for (int y = row0 + ty * TY; y < std::min(row1, row0 + (ty + 1) * TY); ++y)
                for (int x = tx * TX; x < std::min(w, (tx + 1) * TX); ++x) {
                    int pid = x + y * w;
                    Sampler s{seed_of(pass, pid)};
                    float u1 = s.next(), u2 = s.next();
                    float l1 = s.next(), l2 = s.next();
                    V3 o, d;
                    cam_ray(c, ((float)x + u1) * inv_w, ((float)y + u2) * inv_h, l1, l2, 0, o, d);
                    in.push_back(RayS{o, d, 0.0f, -1, pid, s.s});
                }
            int64_t local = 0;
            float lw = 1.0f;
            for (int depth = 0;; ++depth) {
                for (RayS& r : in) {
                    Closest32 hc = wide ? closest_fp32_wide(sc, r.o, r.d) : closest_fp32_fast(sc, r.o, r.d);
                    r.prim = hc.prim;
                    r.t = hc.t;
                }
                local += (int64_t)in.size();
                // misses first (rayStreamSort), then hits
                auto mid = std::stable_partition(in.begin(), in.end(), [](const RayS& r) { return r.prim < 0; });
                for (auto it = in.begin(); it != mid; ++it) {
                    float* a = accum + 4 * (size_t)it->pix;
                    a[0] += lw; a[1] += lw; a[2] += lw; a[3] += 1.0f;
                }
                if (mid == in.end()) break;
                if (depth == max_depth) {
                    for (auto it = mid; it != in.end(); ++it) accum[4 * (size_t)it->pix + 3] += 1.0f;
                    break;
                }
                out.clear();
                for (auto it = mid; it != in.end(); ++it) {
                    Sampler s{it->rng};
                    float s1 = s.next(), s2 = s.next();
                    V3 n = facing_normal(sc.tris[it->prim], it->d);
                    V3 no, nd;
                    bounce(it->o, it->d, it->t, n, s1, s2, 0, no, nd);
                    out.push_back(RayS{no, nd, 0.0f, -1, it->pix, s.s});
                }
                lw *= albedo;
                std::swap(in, out);
            }
            total += local;
        });
    }
    auto t1 = std::chrono::steady_clock::now();
    if (rays_out) *rays_out = total.load();
    if (seconds_out) *seconds_out = std::chrono::duration<double>(t1 - t0).count();
    return 0;
}

// Diffuse bounce of n rays: in[i] = org(3), dir(3), t, normal(3), s1, s2
// (12 floats); out[i] = new org(3), new dir(3).
void orc_bounce(const float* in, int64_t n, int numerics, float* out)
{
    for (int64_t i = 0; i < n; ++i) {
        const float* a = in + 12 * (size_t)i;
        V3 no, nd;
        bounce(v3(a[0], a[1], a[2]), v3(a[3], a[4], a[5]), a[6], v3(a[7], a[8], a[9]), a[10], a[11], numerics,
               no, nd);
        float* o = out + 6 * (size_t)i;
        o[0] = no.x; o[1] = no.y; o[2] = no.z;
        o[3] = nd.x; o[4] = nd.y; o[5] = nd.z;
    }
}

// ---------------------------------------------------------------- albedo
// Path tape for the per-triangle albedo model: the geometry of a path does
// not depend on the albedo, so each path (pixel, pass) is traced once with
// the kernel numerics and its hit triangles are recorded.  The image, the
// exact adjoint and finite differences are then evaluated in double over the
// tape.  (The GPU adjoint re-traces paths instead of storing them.)
struct PathTape {
    int w, h, spp;
    std::vector<int32_t> pixel;      // per path
    std::vector<uint8_t> escaped;    // per path
    std::vector<int64_t> start;      // per path + 1, into hits
    std::vector<int32_t> hits;
    int64_t rays = 0;
};

void* orc_paths_record(void* scp, const float* cam, int w, int h, int pass0, int spp, int max_depth, int threads)
{
    const Scene& sc = *static_cast<Scene*>(scp);
    Cam c = cam_from(cam);
    const float inv_w = 1.0f / (float)w, inv_h = 1.0f / (float)h;
    const int64_t npx = (int64_t)w * h;
    PathTape* tape = new PathTape();
    tape->w = w;
    tape->h = h;
    tape->spp = spp;
    const int64_t npaths = npx * spp;
    std::vector<std::vector<int32_t>> path_hits(npaths);
    tape->pixel.resize(npaths);
    tape->escaped.resize(npaths);
    std::atomic<int64_t> rays{0};
    const int64_t chunk = 4096;
    parallel_tiles((int)((npaths + chunk - 1) / chunk), default_threads(threads), [&](int tile) {
        int64_t local = 0;
        for (int64_t i = tile * chunk; i < std::min(npaths, (int64_t)(tile + 1) * chunk); ++i) {
            const int sample = (int)(i / npx);
            const int pid = (int)(i % npx);
            const int x = pid % w, y = pid / w;
            Sampler s{seed_of(pass0 + sample, pid)};
            float u1 = s.next(), u2 = s.next();
            float l1 = s.next(), l2 = s.next();
            V3 org, dir;
            cam_ray(c, ((float)x + u1) * inv_w, ((float)y + u2) * inv_h, l1, l2, 0, org, dir);
            std::vector<int32_t>& hl = path_hits[i];
            int depth = 0;
            HitAny hit;
            for (;;) {
                hit = closest_any(sc, org, dir, 0);
                ++local;
                if (hit.prim < 0 || depth == max_depth) break;
                hl.push_back(hit.prim);
                Frame fr = hit_frame(sc, hit.prim, hit.u, hit.v, dir, false);
                float s1 = s.next(), s2 = s.next();
                V3 no, nd;
                bounce(org, dir, hit.t, fr, s1, s2, 0, no, nd);
                org = no;
                dir = nd;
                ++depth;
            }
            tape->pixel[i] = pid;
            tape->escaped[i] = hit.prim < 0 ? 1 : 0;
        }
        rays += local;
    });
    tape->start.resize(npaths + 1);
    tape->start[0] = 0;
    for (int64_t i = 0; i < npaths; ++i) tape->start[i + 1] = tape->start[i] + (int64_t)path_hits[i].size();
    tape->hits.resize(tape->start[npaths]);
    for (int64_t i = 0; i < npaths; ++i)
        std::copy(path_hits[i].begin(), path_hits[i].end(), tape->hits.begin() + tape->start[i]);
    tape->rays = rays.load();
    return tape;
}

void orc_paths_destroy(void* t) { delete static_cast<PathTape*>(t); }

int64_t orc_paths_rays(void* t) { return static_cast<PathTape*>(t)->rays; }

// mean RGB image over the tape's samples; albedo: 3 doubles per triangle
void orc_paths_image(void* tp, const double* albedo, double* image)
{
    const PathTape& t = *static_cast<PathTape*>(tp);
    const int64_t npx = (int64_t)t.w * t.h;
    std::fill(image, image + 3 * npx, 0.0);
    const int64_t npaths = (int64_t)t.pixel.size();
    for (int64_t i = 0; i < npaths; ++i) {
        if (!t.escaped[i]) continue;
        double l[3] = {1.0, 1.0, 1.0};
        for (int64_t k = t.start[i]; k < t.start[i + 1]; ++k)
            for (int c = 0; c < 3; ++c) l[c] *= albedo[3 * (size_t)t.hits[k] + c];
        for (int c = 0; c < 3; ++c) image[3 * (size_t)t.pixel[i] + c] += l[c];
    }
    for (int64_t i = 0; i < 3 * npx; ++i) image[i] /= (double)t.spp;
}

// grad (3 doubles per triangle) += dLoss/dalbedo for dLoss/dImage = dldi
void orc_paths_adjoint(void* tp, const double* albedo, const double* dldi, double* grad)
{
    const PathTape& t = *static_cast<PathTape*>(tp);
    const int64_t npaths = (int64_t)t.pixel.size();
    for (int64_t i = 0; i < npaths; ++i) {
        if (!t.escaped[i]) continue;
        const int64_t b = t.start[i], e = t.start[i + 1];
        for (int c = 0; c < 3; ++c) {
            const double wgt = dldi[3 * (size_t)t.pixel[i] + c] / (double)t.spp;
            if (wgt == 0.0) continue;
            for (int64_t k = b; k < e; ++k) {
                double others = 1.0;  // product of the other hits' albedos
                for (int64_t j = b; j < e; ++j)
                    if (j != k) others *= albedo[3 * (size_t)t.hits[j] + c];
                grad[3 * (size_t)t.hits[k] + c] += wgt * others;
            }
        }
    }
}

int orc_thread_count(void) { return default_threads(0); }
int orc_cpu_has_avx2(void) { return cpu_has_avx2() ? 1 : 0; }

}  // extern "C"
