// bvh_build.hip -- LBVH build on the GPU, collapsed to a quantised BVH8.
//
// Replaces the closed OptiX Prime build of
// accel/optix/optix_intersector_stream_cuda.cpp:23-75 (timed buildMs / buildMprim).
//   1. centroid bounds (two-pass reduction)
//   2. 63-bit Morton codes of triangle box centres
//   3. LSD radix sort of (key, triangle id), 8 passes of 8 bits
//   4. radix tree (T. Karras, "Maximizing Parallelism in the Construction of
//      BVHs, Octrees, and k-d Trees", HPG 2012), equal keys split by index
//   5. bottom-up boxes with one arrival counter per internal node: inside
//      4096-leaf tiles in LDS, across tiles with device-scope atomics
//   6. top-down collapse to BVH8 one level per launch.  Which binary nodes
//      become the (at most 8) slots of a wide node follows a surface-area
//      cost optimum computed bottom-up during the refit (S. Ylitie, T.
//      Karras, S. Laine, "Efficient Incoherent Ray Traversal on GPUs Through
//      Compressed Wide BVHs", HPG 2017, section 4.1); subtrees of at most
//      kLeafMax triangles become leaf slots.  (The greedy alternative opens
//      the widest-area internal child until 8 slots are used.)  Node and
//      record offsets come from exclusive scans, so the layout is
//      deterministic.
#include <algorithm>
#include <cmath>
#include <vector>

#include "prt_internal.h"

namespace prt {
namespace {

constexpr int kBlock = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kBlock * kSortItems;
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;
constexpr int32_t kEmptySlot = INT32_MIN;

__device__ __forceinline__ float3 vtx(const float* pos, int i)
{
    return make_float3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
}

struct Box {
    float lo[3], hi[3];
};

__device__ __forceinline__ Box tri_box(const int32_t* idx, const float* pos, int t)
{
    float3 a = vtx(pos, idx[3 * t]), b = vtx(pos, idx[3 * t + 1]), c = vtx(pos, idx[3 * t + 2]);
    Box r;
    r.lo[0] = fminf(fminf(a.x, b.x), c.x);
    r.lo[1] = fminf(fminf(a.y, b.y), c.y);
    r.lo[2] = fminf(fminf(a.z, b.z), c.z);
    r.hi[0] = fmaxf(fmaxf(a.x, b.x), c.x);
    r.hi[1] = fmaxf(fmaxf(a.y, b.y), c.y);
    r.hi[2] = fmaxf(fmaxf(a.z, b.z), c.z);
    return r;
}

__device__ __forceinline__ float box_half_area(const Box& b)
{
    const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    return dx * dy + dy * dz + dz * dx;
}

// ---------------------------------------------------------------- collapse cost
// cost(n, i), i = 1..7: the least surface-area cost of covering the subtree of
// binary node n with at most i slots of the parent wide node.  A slot that is
// a subtree of at most kLeafMax triangles is a leaf (area x triangles x
// c_prim); a larger one is a wide node (area x c_node plus the best split of
// its 8 slots between its two children).  Per internal node the array holds
// the 7 costs and one word of choices: bits 0-2 the left child's share of
// the wide node's 8 slots, bits 3(i-1)..3(i-1)+2 for i = 2..7 the left share
// of i slots (0: cost(n, i-1) is no worse, use fewer).
constexpr int kDpWords = 8;

struct DpCost {
    float c[7];
};

__device__ __forceinline__ DpCost dp_leaf(float area, float c_prim)
{
    DpCost r;
#pragma unroll
    for (int i = 0; i < 7; ++i) r.c[i] = area * c_prim;
    return r;
}

__device__ __forceinline__ DpCost dp_combine(const DpCost& l, const DpCost& r, float area, int size, float c_node,
                                             float c_prim, uint32_t& choice)
{
    float dist[9];
    int dk[9];
#pragma unroll
    for (int j = 2; j <= 8; ++j) {
        float best = __builtin_inff();
        int bk = 1;
#pragma unroll
        for (int k = 1; k < j; ++k) {
            if (k > 7 || j - k > 7) continue;
            const float v = l.c[k - 1] + r.c[j - k - 1];
            if (v < best) {
                best = v;
                bk = k;
            }
        }
        dist[j] = best;
        dk[j] = bk;
    }
    DpCost out;
    out.c[0] = size <= kLeafMax ? area * (float)size * c_prim : area * c_node + dist[8];
    choice = (uint32_t)dk[8];
#pragma unroll
    for (int i = 2; i <= 7; ++i) {
        if (dist[i] < out.c[i - 2]) {
            out.c[i - 1] = dist[i];
            choice |= (uint32_t)dk[i] << (3 * (i - 1));
        } else {
            out.c[i - 1] = out.c[i - 2];
        }
    }
    return out;
}

// Written and read as agent-scope atomics: in the refit's first phase a
// sibling's costs come from another thread of the workgroup (ordered by the
// LDS arrival counter and workgroup fences), in the second from any
// workgroup (device-scope fences).
__device__ __forceinline__ void dp_store(uint32_t* dp, int32_t node, const DpCost& c, uint32_t choice)
{
    uint32_t* d = dp + (int64_t)kDpWords * node;
#pragma unroll
    for (int i = 0; i < 7; ++i)
        __hip_atomic_store(d + i, __float_as_uint(c.c[i]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(d + 7, choice, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------- bounds
__global__ __launch_bounds__(kBlock) void k_centroid_partial(const int32_t* idx, const float* pos,
                                                             int64_t n, float* partial)
{
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        Box b = tri_box(idx, pos, (int)i);
        for (int k = 0; k < 3; ++k) {
            float c = 0.5f * (b.lo[k] + b.hi[k]);
            lo[k] = fminf(lo[k], c);
            hi[k] = fmaxf(hi[k], c);
        }
    }
    __shared__ float red[6][kBlock];
    for (int k = 0; k < 3; ++k) {
        red[k][threadIdx.x] = lo[k];
        red[3 + k][threadIdx.x] = hi[k];
    }
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            for (int k = 0; k < 3; ++k) {
                red[k][threadIdx.x] = fminf(red[k][threadIdx.x], red[k][threadIdx.x + s]);
                red[3 + k][threadIdx.x] = fmaxf(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + s]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(kBlock) void k_centroid_final(const float* partial, int nparts, float* out)
{
    __shared__ float red[6][kBlock];
    float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int i = threadIdx.x; i < nparts; i += kBlock)
        for (int k = 0; k < 6; ++k)
            v[k] = k < 3 ? fminf(v[k], partial[i * 6 + k]) : fmaxf(v[k], partial[i * 6 + k]);
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < 6; ++k)
                red[k][threadIdx.x] = k < 3 ? fminf(red[k][threadIdx.x], red[k][threadIdx.x + s])
                                            : fmaxf(red[k][threadIdx.x], red[k][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x < 6) out[threadIdx.x] = red[threadIdx.x][0];
}

// ---------------------------------------------------------------- Morton
__device__ __forceinline__ uint64_t spread3_21(uint32_t v)
{
    uint64_t x = v & 0x1fffffu;
    x = (x | x << 32) & 0x001f00000000ffffull;
    x = (x | x << 16) & 0x001f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

__global__ __launch_bounds__(kBlock) void k_morton(const int32_t* idx, const float* pos, int64_t n,
                                                   const float* bounds, uint64_t* keys, uint32_t* vals)
{
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    Box b = tri_box(idx, pos, (int)i);
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        float c = 0.5f * (b.lo[k] + b.hi[k]);
        float ext = bounds[3 + k] - bounds[k];
        float u = ext > 0.0f ? (c - bounds[k]) / ext : 0.0f;
        float s = u * 2097152.0f;  // 2^21
        uint32_t qi = s <= 0.0f ? 0u : (s >= 2097151.0f ? 2097151u : (uint32_t)s);
        q[k] = qi;
    }
    keys[i] = (spread3_21(q[0]) << 2) | (spread3_21(q[1]) << 1) | spread3_21(q[2]);
    vals[i] = (uint32_t)i;
}

// ---------------------------------------------------------------- scan
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const int lane = threadIdx.x & 63;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// exclusive scan of kBlock values held one per thread; returns the block total
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t& total)
{
    __shared__ uint32_t wsum[kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
        if (w < wave) off += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    total = tot;
    return off + inc - v;
}

__global__ __launch_bounds__(kBlock) void k_scan_reduce(const uint32_t* in, int64_t n, uint32_t* sums)
{
    int64_t base = (int64_t)blockIdx.x * kScanTile;
    uint32_t s = 0;
    for (int j = 0; j < kScanItems; ++j) {
        int64_t i = base + (int64_t)j * kBlock + threadIdx.x;
        if (i < n) s += in[i];
    }
    uint32_t total;
    block_excl_scan(s, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kBlock) void k_scan_apply(const uint32_t* in, uint32_t* out, int64_t n,
                                                       const uint32_t* tile_off)
{
    int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
    for (int j = 0; j < kScanItems; ++j) {
        int64_t i = base + j;
        v[j] = i < n ? in[i] : 0u;
        s += v[j];
    }
    uint32_t total;
    uint32_t off = block_excl_scan(s, total) + (tile_off ? tile_off[blockIdx.x] : 0u);
    for (int j = 0; j < kScanItems; ++j) {
        int64_t i = base + j;
        if (i < n) out[i] = off;
        off += v[j];
    }
}

// ---------------------------------------------------------------- radix sort
__global__ __launch_bounds__(kBlock) void k_radix_hist(const uint64_t* keys, int64_t n, int shift,
                                                       uint32_t* counts, int nblocks)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    int64_t base = (int64_t)blockIdx.x * kSortTile;
    for (int j = 0; j < kSortItems; ++j) {
        int64_t i = base + (int64_t)j * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    counts[(int64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: keys are ranked round by round (256 consecutive keys per
// round); inside a wave, lanes with equal digits are found with 8 ballots.
__global__ __launch_bounds__(kBlock) void k_radix_scatter(const uint64_t* kin, const uint32_t* vin,
                                                          uint64_t* kout, uint32_t* vout, int64_t n,
                                                          int shift, const uint32_t* offsets, int nblocks)
{
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[kBlock / 64][256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    run[tid] = offsets[(int64_t)tid * nblocks + blockIdx.x];
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int64_t base = (int64_t)blockIdx.x * kSortTile;
    for (int r = 0; r < kSortItems; ++r) {
        int64_t i = base + (int64_t)r * kBlock + tid;
        bool valid = i < n;
        uint64_t key = valid ? kin[i] : 0ull;
        uint32_t val = valid ? vin[i] : 0u;
        uint32_t d = (uint32_t)(key >> shift) & 255u;
        for (int w = 0; w < kBlock / 64; ++w) wcnt[w][tid] = 0;
        __syncthreads();
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < 8; ++b) {
            bool bit = (d >> b) & 1u;
            uint64_t bb = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[d] + rank;
            for (int w = 0; w < wave; ++w) pos += wcnt[w][d];
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        uint32_t add = 0;
        for (int w = 0; w < kBlock / 64; ++w) add += wcnt[w][tid];
        run[tid] += add;
        __syncthreads();
    }
}

// ---------------------------------------------------------------- radix tree
__device__ __forceinline__ int common_prefix(const uint64_t* k, int64_t n, int64_t i, int64_t j)
{
    if (j < 0 || j >= n) return -1;
    uint64_t a = k[i], b = k[j];
    if (a == b) return 64 + __clz((uint32_t)(i ^ j));
    return __clzll(a ^ b);
}

// child encoding: >= 0 internal node index, < 0 leaf ~index (sorted order)
__global__ __launch_bounds__(kBlock) void k_radix_tree(const uint64_t* keys, int64_t n, int32_t* left,
                                                       int32_t* right, int32_t* first, int32_t* last,
                                                       int32_t* iparent, int32_t* lparent)
{
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n - 1) return;
    int d = (common_prefix(keys, n, i, i + 1) - common_prefix(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = common_prefix(keys, n, i, i - d);
    int64_t lmax = 2;
    while (common_prefix(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int64_t l = 0;
    for (int64_t t = lmax >> 1; t >= 1; t >>= 1)
        if (common_prefix(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int64_t j = i + l * d;
    int dnode = common_prefix(keys, n, i, j);
    int64_t s = 0;
    int64_t t = l;
    do {
        t = (t + 1) >> 1;
        if (common_prefix(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
    int64_t lo = i < j ? i : j, hi = i < j ? j : i;
    int32_t lc = (lo == gamma) ? ~(int32_t)gamma : (int32_t)gamma;
    int32_t rc = (hi == gamma + 1) ? ~(int32_t)(gamma + 1) : (int32_t)(gamma + 1);
    left[i] = lc;
    right[i] = rc;
    first[i] = (int32_t)lo;
    last[i] = (int32_t)hi;
    if (lc >= 0) iparent[lc] = (int32_t)i; else lparent[~lc] = (int32_t)i;
    if (rc >= 0) iparent[rc] = (int32_t)i; else lparent[~rc] = (int32_t)i;
    if (i == 0) iparent[0] = -1;
}

__device__ __forceinline__ float ld_coherent(const float* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ Box child_box(int32_t c, const uint32_t* sorted, const int32_t* idx,
                                         const float* pos, const float* ibox)
{
    if (c < 0) return tri_box(idx, pos, (int)sorted[~c]);
    Box b;
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = ld_coherent(ibox + 6 * (int64_t)c + k);
        b.hi[k] = ld_coherent(ibox + 6 * (int64_t)c + 3 + k);
    }
    return b;
}

__device__ __forceinline__ DpCost child_cost(int32_t c, const Box& cb, const uint32_t* dp, float c_prim)
{
    if (c < 0) return dp_leaf(box_half_area(cb), c_prim);
    DpCost r;
#pragma unroll
    for (int i = 0; i < 7; ++i)
        r.c[i] = __uint_as_float(__hip_atomic_load(dp + (int64_t)kDpWords * c + i, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT));
    return r;
}

// Two-phase refit.  Phase A: one workgroup per tile of kRefitTile sorted
// leaves finishes every internal node whose leaf range lies inside the tile.
// A Karras node's index is one end of its range, so those nodes have indices
// inside the tile too and their arrival flags and boxes live in LDS; only
// workgroup-scope ordering is needed.  A walk that reaches a node crossing
// the tile border appends it to `pending`.  Phase B runs the device-scope
// arrival protocol above for the crossing nodes only.
constexpr int kRefitTile = 4096;
constexpr int kRefitThreads = 1024;

__global__ __launch_bounds__(kRefitThreads) void k_refit_local(
    int64_t n, const uint32_t* sorted, const int32_t* idx, const float* pos, const int32_t* left,
    const int32_t* right, const int32_t* first, const int32_t* last, const int32_t* iparent,
    const int32_t* lparent, float* ibox, int32_t* pending, uint32_t* pending_count, uint32_t* dp, float c_node,
    float c_prim)
{
    __shared__ uint32_t flag[kRefitTile];
    __shared__ float lbox[kRefitTile][6];
    const int64_t lo = (int64_t)blockIdx.x * kRefitTile;
    const int64_t hi = (lo + kRefitTile < n ? lo + kRefitTile : n) - 1;
    for (int i = threadIdx.x; i < kRefitTile; i += kRefitThreads) flag[i] = 0u;
    __syncthreads();
    for (int k = 0; k < kRefitTile / kRefitThreads; ++k) {
        const int64_t j = lo + (int64_t)k * kRefitThreads + threadIdx.x;
        if (j > hi) break;
        Box b = tri_box(idx, pos, (int)sorted[j]);
        int32_t cur = ~(int32_t)j;
        DpCost cur_cost = dp_leaf(box_half_area(b), c_prim);
        int32_t p = lparent[j];
        while (p >= 0) {
            if (first[p] < lo || last[p] > hi) {
                const uint32_t slot = atomicAdd(pending_count, 1u);
                pending[slot] = p;
                break;
            }
            const int s = (int)(p - lo);
            __threadfence_block();
            if (atomicAdd(&flag[s], 1u) == 0u) break;
            __threadfence_block();
            const bool cur_left = left[p] == cur;
            const int32_t sib = cur_left ? right[p] : left[p];
            Box o;
            if (sib < 0) {
                o = tri_box(idx, pos, (int)sorted[~sib]);
            } else {
                for (int a = 0; a < 3; ++a) {
                    o.lo[a] = lbox[sib - lo][a];
                    o.hi[a] = lbox[sib - lo][3 + a];
                }
            }
            for (int a = 0; a < 3; ++a) {
                b.lo[a] = fminf(b.lo[a], o.lo[a]);
                b.hi[a] = fmaxf(b.hi[a], o.hi[a]);
                lbox[s][a] = b.lo[a];
                lbox[s][3 + a] = b.hi[a];
                ibox[6 * (int64_t)p + a] = b.lo[a];
                ibox[6 * (int64_t)p + 3 + a] = b.hi[a];
            }
            if (dp) {
                const DpCost sib_cost = child_cost(sib, o, dp, c_prim);
                uint32_t choice;
                cur_cost = cur_left ? dp_combine(cur_cost, sib_cost, box_half_area(b), last[p] - first[p] + 1,
                                                 c_node, c_prim, choice)
                                    : dp_combine(sib_cost, cur_cost, box_half_area(b), last[p] - first[p] + 1,
                                                 c_node, c_prim, choice);
                dp_store(dp, p, cur_cost, choice);
            }
            cur = p;
            p = iparent[p];
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_refit_crossing(const int32_t* pending, const uint32_t* pending_count,
                                                           const uint32_t* sorted, const int32_t* idx,
                                                           const float* pos, const int32_t* left,
                                                           const int32_t* right, const int32_t* first,
                                                           const int32_t* last, const int32_t* iparent,
                                                           float* ibox, uint32_t* flags, uint32_t* dp, float c_node,
                                                           float c_prim)
{
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= *pending_count) return;
    int32_t p = pending[k];
    while (p >= 0) {
        __threadfence();
        if (atomicAdd(&flags[p], 1u) == 0u) return;
        __threadfence();
        Box a = child_box(left[p], sorted, idx, pos, ibox);
        Box b = child_box(right[p], sorted, idx, pos, ibox);
        Box u;
        for (int c = 0; c < 3; ++c) {
            u.lo[c] = fminf(a.lo[c], b.lo[c]);
            u.hi[c] = fmaxf(a.hi[c], b.hi[c]);
            ibox[6 * (int64_t)p + c] = u.lo[c];
            ibox[6 * (int64_t)p + 3 + c] = u.hi[c];
        }
        if (dp) {
            uint32_t choice;
            const DpCost c = dp_combine(child_cost(left[p], a, dp, c_prim), child_cost(right[p], b, dp, c_prim),
                                        box_half_area(u), last[p] - first[p] + 1, c_node, c_prim, choice);
            dp_store(dp, p, c, choice);
        }
        p = iparent[p];
    }
}

// ---------------------------------------------------------------- collapse
__device__ __forceinline__ float half_area(const float* b)
{
    float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    return dx * dy + dy * dz + dz * dx;
}

__global__ __launch_bounds__(kBlock) void k_collapse_plan(const int32_t* tasks, int ntasks,
                                                          const int32_t* left, const int32_t* right,
                                                          const int32_t* first, const int32_t* last,
                                                          const float* ibox, const uint32_t* dp, int32_t* slots,
                                                          uint32_t* n_int, uint32_t* n_tri)
{
    int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= ntasks) return;
    int32_t b = tasks[t];
    int32_t s[8];
    int ns = 0;
    if (dp) {
        // follow the cost choices: (node, slots) pairs; the slot counts on the
        // stack plus the slots taken always sum to at most 8
        int32_t st_n[8];
        int st_j[8];
        const uint32_t k8 = dp[(int64_t)kDpWords * b + 7] & 7u;
        st_n[0] = right[b];
        st_j[0] = 8 - (int)k8;
        st_n[1] = left[b];
        st_j[1] = (int)k8;
        int sp = 2;
        while (sp > 0) {
            --sp;
            const int32_t c = st_n[sp];
            int j = st_j[sp];
            if (c < 0) {
                s[ns++] = c;
                continue;
            }
            const uint32_t ch = dp[(int64_t)kDpWords * c + 7];
            while (j > 1 && ((ch >> (3 * (j - 1))) & 7u) == 0u) --j;
            if (j == 1) {
                s[ns++] = c;
                continue;
            }
            const int k = (int)((ch >> (3 * (j - 1))) & 7u);
            st_n[sp] = right[c];
            st_j[sp] = j - k;
            st_n[sp + 1] = left[c];
            st_j[sp + 1] = k;
            sp += 2;
        }
    } else {
        s[0] = left[b];
        s[1] = right[b];
        ns = 2;
    }
    while (!dp && ns < 8) {
        int best = -1;
        float best_a = -1.0f;
        for (int k = 0; k < ns; ++k) {
            if (s[k] >= 0) {
                float a = half_area(ibox + 6 * (int64_t)s[k]);
                if (a > best_a) { best_a = a; best = k; }
            }
        }
        if (best < 0) break;
        int32_t c = s[best];
        s[best] = left[c];
        s[ns++] = right[c];
    }
    uint32_t ni = 0, nt = 0;
    for (int k = 0; k < 8; ++k) {
        int32_t c = k < ns ? s[k] : kEmptySlot;
        slots[8 * (int64_t)t + k] = c;
        if (c == kEmptySlot) continue;
        if (c < 0) { nt += 1; continue; }
        int size = last[c] - first[c] + 1;
        if (size <= kLeafMax) nt += size; else ni += 1;
    }
    n_int[t] = ni;
    n_tri[t] = nt;
}

__device__ __forceinline__ void write_record(int64_t rec, int tri, const int32_t* idx, const float* pos,
                                             float* tris, int32_t* prim_ids)
{
    for (int v = 0; v < 3; ++v) {
        int vi = idx[3 * tri + v];
        tris[9 * rec + 3 * v + 0] = pos[3 * vi + 0];
        tris[9 * rec + 3 * v + 1] = pos[3 * vi + 1];
        tris[9 * rec + 3 * v + 2] = pos[3 * vi + 2];
    }
    prim_ids[rec] = tri;
}

// smallest e with 255 * 2^e >= extent (e >= -126); exact in double
__device__ __forceinline__ int quant_exponent(double extent)
{
    if (!(extent > 0.0)) return -126;
    int e = ilogb(extent) - 7;
    while (ldexp(255.0, e) < extent) ++e;
    while (e > -126 && ldexp(255.0, e - 1) >= extent) --e;
    return e < -126 ? -126 : e;
}

__global__ __launch_bounds__(kBlock) void k_collapse_emit(
    const int32_t* tasks, int ntasks, const int32_t* slots, const uint32_t* scan_int,
    const uint32_t* scan_tri, int64_t node_base, int64_t next_base, int64_t tri_base_level,
    const int32_t* first, const int32_t* last, const float* ibox, const uint32_t* sorted,
    const int32_t* idx, const float* pos, Bvh8Node* nodes, float* tris, int32_t* prim_ids,
    int32_t* next_tasks)
{
    int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= ntasks) return;
    int32_t b = tasks[t];
    const float* nb = ibox + 6 * (int64_t)b;
    Bvh8Node node;
    node.px = nb[0];
    node.py = nb[1];
    node.pz = nb[2];
    int e[3];
    double inv_scale[3];
    for (int k = 0; k < 3; ++k) {
        e[k] = quant_exponent((double)nb[3 + k] - (double)nb[k]);
        inv_scale[k] = ldexp(1.0, -e[k]);
    }
    node.ex = (uint8_t)(int8_t)e[0];
    node.ey = (uint8_t)(int8_t)e[1];
    node.ez = (uint8_t)(int8_t)e[2];
    uint32_t child_base = (uint32_t)(next_base + scan_int[t]);
    uint32_t tri_base = (uint32_t)(tri_base_level + scan_tri[t]);
    node.child_base = child_base;
    node.tri_base = tri_base;
    uint8_t imask = 0;
    uint32_t ci = 0, toff = 0;
    uint8_t* qlo[3] = {node.qlo_x, node.qlo_y, node.qlo_z};
    uint8_t* qhi[3] = {node.qhi_x, node.qhi_y, node.qhi_z};
    const double p[3] = {(double)nb[0], (double)nb[1], (double)nb[2]};
    for (int k = 0; k < 8; ++k) {
        int32_t c = slots[8 * (int64_t)t + k];
        node.meta[k] = 0;
        if (c == kEmptySlot) {
            for (int a = 0; a < 3; ++a) { qlo[a][k] = 255; qhi[a][k] = 0; }
            continue;
        }
        Box cb;
        if (c < 0) cb = tri_box(idx, pos, (int)sorted[~c]);
        else {
            for (int a = 0; a < 3; ++a) {
                cb.lo[a] = ibox[6 * (int64_t)c + a];
                cb.hi[a] = ibox[6 * (int64_t)c + 3 + a];
            }
        }
        for (int a = 0; a < 3; ++a) {
            double lo = floor(((double)cb.lo[a] - p[a]) * inv_scale[a]);
            double hi = ceil(((double)cb.hi[a] - p[a]) * inv_scale[a]);
            lo = lo < 0.0 ? 0.0 : (lo > 255.0 ? 255.0 : lo);
            hi = hi < 0.0 ? 0.0 : (hi > 255.0 ? 255.0 : hi);
            qlo[a][k] = (uint8_t)lo;
            qhi[a][k] = (uint8_t)hi;
        }
        if (c < 0) {
            write_record((int64_t)tri_base + toff, (int)sorted[~c], idx, pos, tris, prim_ids);
            node.meta[k] = (uint8_t)((1u << 5) | toff);
            toff += 1;
        } else {
            int size = last[c] - first[c] + 1;
            if (size <= kLeafMax) {
                for (int q = 0; q < size; ++q)
                    write_record((int64_t)tri_base + toff + q, (int)sorted[first[c] + q], idx, pos,
                                 tris, prim_ids);
                node.meta[k] = (uint8_t)(((uint32_t)size << 5) | toff);
                toff += size;
            } else {
                imask |= (uint8_t)(1u << k);
                next_tasks[scan_int[t] + ci] = c;
                ++ci;
            }
        }
    }
    node.imask = imask;
    const uint4* src = reinterpret_cast<const uint4*>(&node);
    uint4* dst = reinterpret_cast<uint4*>(nodes + node_base + t);
    for (int w = 0; w < 5; ++w) dst[w] = src[w];
}

// ---------------------------------------------------------------- host side
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    template <class T> T* as() { return static_cast<T*>(p); }
};

hipError_t dev_alloc(DevBuf& b, size_t bytes)
{
    return hipMalloc(&b.p, bytes < 16 ? 16 : bytes);
}

inline int blocks_for(int64_t n, int per) { return (int)((n + per - 1) / per); }

// Exclusive scan of n uint32 values; scratch holds the per-level tile sums.
hipError_t scan_exclusive(const uint32_t* in, uint32_t* out, int64_t n, hipStream_t s,
                          std::vector<DevBuf>& scratch, int level = 0)
{
    int nt = blocks_for(n, kScanTile);
    if (n == 0) return hipSuccess;
    if (nt == 1) {
        k_scan_apply<<<1, kBlock, 0, s>>>(in, out, n, nullptr);
        return hipGetLastError();
    }
    if ((int)scratch.size() < 2 * (level + 1)) scratch.resize(2 * (level + 1));
    DevBuf& sums = scratch[2 * level];
    DevBuf& offs = scratch[2 * level + 1];
    if (!sums.p) {
        hipError_t e = dev_alloc(sums, sizeof(uint32_t) * nt);
        if (e != hipSuccess) return e;
        e = dev_alloc(offs, sizeof(uint32_t) * nt);
        if (e != hipSuccess) return e;
    }
    k_scan_reduce<<<nt, kBlock, 0, s>>>(in, n, sums.as<uint32_t>());
    hipError_t e = scan_exclusive(sums.as<uint32_t>(), offs.as<uint32_t>(), nt, s, scratch, level + 1);
    if (e != hipSuccess) return e;
    k_scan_apply<<<nt, kBlock, 0, s>>>(in, out, n, offs.as<uint32_t>());
    return hipGetLastError();
}

}  // namespace

int build_bvh8(prt_hip_ctx* ctx, SceneBuffers& sb, prt_build_stats* stats)
{
    const int64_t n = sb.tri_count;
    hipStream_t s = ctx->stream;
    if (n <= 0) return set_error(ctx, "bvh_build: scene has no triangles");
    if (n >= (int64_t)1 << 31) return set_error(ctx, "bvh_build: more than 2^31-1 triangles");

    if (sb.d_nodes) { (void)hipFree(sb.d_nodes); sb.d_nodes = nullptr; }
    if (sb.d_tris) { (void)hipFree(sb.d_tris); sb.d_tris = nullptr; }
    if (sb.d_prim_ids) { (void)hipFree(sb.d_prim_ids); sb.d_prim_ids = nullptr; }
    sb.bvh_ready = false;

    const int64_t max_nodes = std::max<int64_t>(1, n - 1);
    PRT_TRY(ctx, hipMalloc(&sb.d_nodes, sizeof(Bvh8Node) * max_nodes));
    PRT_TRY(ctx, hipMalloc(&sb.d_tris, sizeof(float) * 9 * n));
    PRT_TRY(ctx, hipMalloc(&sb.d_prim_ids, sizeof(int32_t) * n));

    const int nsort = blocks_for(n, kSortTile);
    const int nred = std::min(blocks_for(n, kBlock), 1024);
    const int64_t ni = std::max<int64_t>(1, n - 1);
    DevBuf keys[2], vals[2], counts, offs, partial, bounds, left, right, first, last, iparent, lparent,
        ibox, flags, tasks[2], slots, n_int, n_tri, scan_int, scan_tri;
    std::vector<DevBuf> scratch;
    PRT_TRY(ctx, dev_alloc(keys[0], 8 * n));
    PRT_TRY(ctx, dev_alloc(keys[1], 8 * n));
    PRT_TRY(ctx, dev_alloc(vals[0], 4 * n));
    PRT_TRY(ctx, dev_alloc(vals[1], 4 * n));
    PRT_TRY(ctx, dev_alloc(counts, 4 * 256 * (int64_t)nsort));
    PRT_TRY(ctx, dev_alloc(offs, 4 * 256 * (int64_t)nsort));
    PRT_TRY(ctx, dev_alloc(partial, 4 * 6 * nred));
    PRT_TRY(ctx, dev_alloc(bounds, 4 * 6));
    PRT_TRY(ctx, dev_alloc(left, 4 * ni));
    PRT_TRY(ctx, dev_alloc(right, 4 * ni));
    PRT_TRY(ctx, dev_alloc(first, 4 * ni));
    PRT_TRY(ctx, dev_alloc(last, 4 * ni));
    PRT_TRY(ctx, dev_alloc(iparent, 4 * ni));
    PRT_TRY(ctx, dev_alloc(lparent, 4 * n));
    PRT_TRY(ctx, dev_alloc(ibox, 4 * 6 * ni));
    PRT_TRY(ctx, dev_alloc(flags, 4 * (ni + 1)));
    DevBuf pending;
    PRT_TRY(ctx, dev_alloc(pending, 4 * 2 * ni));
    PRT_TRY(ctx, dev_alloc(tasks[0], 4 * ni));
    PRT_TRY(ctx, dev_alloc(tasks[1], 4 * ni));
    PRT_TRY(ctx, dev_alloc(slots, 4 * 8 * ni));
    PRT_TRY(ctx, dev_alloc(n_int, 4 * ni));
    PRT_TRY(ctx, dev_alloc(n_tri, 4 * ni));
    PRT_TRY(ctx, dev_alloc(scan_int, 4 * ni));
    PRT_TRY(ctx, dev_alloc(scan_tri, 4 * ni));
    DevBuf dp;
    if (ctx->bvh_collapse == PRT_COLLAPSE_SAH) PRT_TRY(ctx, dev_alloc(dp, 4 * kDpWords * ni));

    PRT_TRY(ctx, hipEventRecord(ctx->ev0, s));

    if (n == 1) {
        // one triangle: a root whose single leaf slot holds it
        Bvh8Node root{};
        std::vector<int32_t> idx(3);
        std::vector<float> pos;
        PRT_TRY(ctx, hipMemcpyAsync(idx.data(), sb.d_idx, 12, hipMemcpyDeviceToHost, s));
        PRT_TRY(ctx, hipStreamSynchronize(s));
        pos.resize(9);
        for (int v = 0; v < 3; ++v)
            PRT_TRY(ctx, hipMemcpyAsync(&pos[3 * v], sb.d_pos + 3 * (int64_t)idx[v], 12,
                                        hipMemcpyDeviceToHost, s));
        PRT_TRY(ctx, hipStreamSynchronize(s));
        float lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(std::min(pos[a], pos[3 + a]), pos[6 + a]);
            hi[a] = std::max(std::max(pos[a], pos[3 + a]), pos[6 + a]);
        }
        root.px = lo[0]; root.py = lo[1]; root.pz = lo[2];
        uint8_t* ql[3] = {root.qlo_x, root.qlo_y, root.qlo_z};
        uint8_t* qh[3] = {root.qhi_x, root.qhi_y, root.qhi_z};
        uint8_t* ex[3] = {&root.ex, &root.ey, &root.ez};
        for (int a = 0; a < 3; ++a) {
            double extent = (double)hi[a] - (double)lo[a];
            int e = -126;
            if (extent > 0.0) {
                e = std::ilogb(extent) - 7;
                while (std::ldexp(255.0, e) < extent) ++e;
                while (e > -126 && std::ldexp(255.0, e - 1) >= extent) --e;
            }
            *ex[a] = (uint8_t)(int8_t)e;
            for (int k = 0; k < 8; ++k) { ql[a][k] = 255; qh[a][k] = 0; }
            ql[a][0] = 0;
            qh[a][0] = (uint8_t)std::min(255.0, std::ceil(extent * std::ldexp(1.0, -e)));
        }
        root.meta[0] = (uint8_t)(1u << 5);
        root.child_base = 1;
        root.tri_base = 0;
        int32_t zero = 0;
        PRT_TRY(ctx, hipMemcpyAsync(sb.d_nodes, &root, sizeof(root), hipMemcpyHostToDevice, s));
        PRT_TRY(ctx, hipMemcpyAsync(sb.d_tris, pos.data(), 36, hipMemcpyHostToDevice, s));
        PRT_TRY(ctx, hipMemcpyAsync(sb.d_prim_ids, &zero, 4, hipMemcpyHostToDevice, s));
        sb.node_count = 1;
        sb.bvh_depth = 1;
    } else {
        k_centroid_partial<<<nred, kBlock, 0, s>>>(sb.d_idx, sb.d_pos, n, partial.as<float>());
        k_centroid_final<<<1, kBlock, 0, s>>>(partial.as<float>(), nred, bounds.as<float>());
        k_morton<<<blocks_for(n, kBlock), kBlock, 0, s>>>(sb.d_idx, sb.d_pos, n, bounds.as<float>(),
                                                          keys[0].as<uint64_t>(), vals[0].as<uint32_t>());
        PRT_TRY(ctx, hipGetLastError());
        int cur = 0;
        for (int shift = 0; shift < 64; shift += 8) {
            k_radix_hist<<<nsort, kBlock, 0, s>>>(keys[cur].as<uint64_t>(), n, shift,
                                                  counts.as<uint32_t>(), nsort);
            PRT_TRY(ctx, hipGetLastError());
            PRT_TRY(ctx, scan_exclusive(counts.as<uint32_t>(), offs.as<uint32_t>(), 256 * (int64_t)nsort,
                                        s, scratch));
            k_radix_scatter<<<nsort, kBlock, 0, s>>>(keys[cur].as<uint64_t>(), vals[cur].as<uint32_t>(),
                                                     keys[1 - cur].as<uint64_t>(),
                                                     vals[1 - cur].as<uint32_t>(), n, shift,
                                                     offs.as<uint32_t>(), nsort);
            PRT_TRY(ctx, hipGetLastError());
            cur = 1 - cur;
        }
        const uint64_t* sk = keys[cur].as<uint64_t>();
        const uint32_t* sv = vals[cur].as<uint32_t>();
        k_radix_tree<<<blocks_for(n - 1, kBlock), kBlock, 0, s>>>(
            sk, n, left.as<int32_t>(), right.as<int32_t>(), first.as<int32_t>(), last.as<int32_t>(),
            iparent.as<int32_t>(), lparent.as<int32_t>());
        PRT_TRY(ctx, hipMemsetAsync(flags.p, 0, 4 * (ni + 1), s));
        uint32_t* pending_count = flags.as<uint32_t>() + ni;
        const int tiles = blocks_for(n, kRefitTile);
        uint32_t* dpw = ctx->bvh_collapse == PRT_COLLAPSE_SAH ? dp.as<uint32_t>() : nullptr;
        const float c_node = 1.0f, c_prim = ctx->bvh_prim_cost;
        k_refit_local<<<tiles, kRefitThreads, 0, s>>>(n, sv, sb.d_idx, sb.d_pos, left.as<int32_t>(),
                                                      right.as<int32_t>(), first.as<int32_t>(), last.as<int32_t>(),
                                                      iparent.as<int32_t>(), lparent.as<int32_t>(), ibox.as<float>(),
                                                      pending.as<int32_t>(), pending_count, dpw, c_node, c_prim);
        PRT_TRY(ctx, hipGetLastError());
        // at most two arrivals per crossing node; bounded by the internal node count
        k_refit_crossing<<<blocks_for(2 * ni, kBlock), kBlock, 0, s>>>(pending.as<int32_t>(), pending_count, sv,
                                                                   sb.d_idx, sb.d_pos, left.as<int32_t>(),
                                                                   right.as<int32_t>(), first.as<int32_t>(),
                                                                   last.as<int32_t>(), iparent.as<int32_t>(),
                                                                   ibox.as<float>(), flags.as<uint32_t>(), dpw,
                                                                   c_node, c_prim);
        PRT_TRY(ctx, hipGetLastError());

        // collapse, one level per iteration
        int32_t root = 0;
        PRT_TRY(ctx, hipMemcpyAsync(tasks[0].p, &root, 4, hipMemcpyHostToDevice, s));
        int64_t ntasks = 1, node_base = 0, tri_base = 0;
        int depth = 0, tc = 0;
        while (ntasks > 0) {
            int nb = blocks_for(ntasks, kBlock);
            k_collapse_plan<<<nb, kBlock, 0, s>>>(tasks[tc].as<int32_t>(), (int)ntasks, left.as<int32_t>(),
                                                  right.as<int32_t>(), first.as<int32_t>(),
                                                  last.as<int32_t>(), ibox.as<float>(), dpw,
                                                  slots.as<int32_t>(), n_int.as<uint32_t>(),
                                                  n_tri.as<uint32_t>());
            PRT_TRY(ctx, hipGetLastError());
            PRT_TRY(ctx, scan_exclusive(n_int.as<uint32_t>(), scan_int.as<uint32_t>(), ntasks, s, scratch));
            PRT_TRY(ctx, scan_exclusive(n_tri.as<uint32_t>(), scan_tri.as<uint32_t>(), ntasks, s, scratch));
            uint32_t tail[4];
            PRT_TRY(ctx, hipMemcpyAsync(&tail[0], scan_int.as<uint32_t>() + ntasks - 1, 4,
                                        hipMemcpyDeviceToHost, s));
            PRT_TRY(ctx, hipMemcpyAsync(&tail[1], n_int.as<uint32_t>() + ntasks - 1, 4,
                                        hipMemcpyDeviceToHost, s));
            PRT_TRY(ctx, hipMemcpyAsync(&tail[2], scan_tri.as<uint32_t>() + ntasks - 1, 4,
                                        hipMemcpyDeviceToHost, s));
            PRT_TRY(ctx, hipMemcpyAsync(&tail[3], n_tri.as<uint32_t>() + ntasks - 1, 4,
                                        hipMemcpyDeviceToHost, s));
            PRT_TRY(ctx, hipStreamSynchronize(s));
            int64_t next_tasks = (int64_t)tail[0] + tail[1];
            int64_t level_tris = (int64_t)tail[2] + tail[3];
            int64_t next_base = node_base + ntasks;
            if (next_base + next_tasks > max_nodes)
                return set_error(ctx, "bvh_build: node budget exceeded");
            k_collapse_emit<<<nb, kBlock, 0, s>>>(
                tasks[tc].as<int32_t>(), (int)ntasks, slots.as<int32_t>(), scan_int.as<uint32_t>(),
                scan_tri.as<uint32_t>(), node_base, next_base, tri_base, first.as<int32_t>(),
                last.as<int32_t>(), ibox.as<float>(), sv, sb.d_idx, sb.d_pos, sb.d_nodes,
                sb.d_tris, sb.d_prim_ids, tasks[1 - tc].as<int32_t>());
            PRT_TRY(ctx, hipGetLastError());
            node_base = next_base;
            tri_base += level_tris;
            ntasks = next_tasks;
            tc = 1 - tc;
            ++depth;
        }
        if (tri_base != n) return set_error(ctx, "bvh_build: triangle records do not cover the scene");
        sb.node_count = node_base;
        sb.bvh_depth = depth;
    }

    PRT_TRY(ctx, hipEventRecord(ctx->ev1, s));
    PRT_TRY(ctx, hipEventSynchronize(ctx->ev1));
    float ms = 0.0f;
    PRT_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    sb.bvh_ready = true;
    if (stats) {
        stats->build_ms = ms;
        stats->build_mprim = ms > 0.0f ? (double)n / 1000.0 / ms : 0.0;
        stats->node_count = sb.node_count;
        stats->tri_count = n;
        stats->depth = sb.bvh_depth;
        stats->reserved = 0;
    }
    return 0;
}

}  // namespace prt
