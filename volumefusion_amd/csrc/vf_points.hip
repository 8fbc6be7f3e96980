// vf_points.hip -- the mesh-point path: knn of points to motion-graph
// nodes, Gaussian weights with cutoff, and the blend.
//
// The reference warps mesh points every Gauss-Newton iteration
// (src/VolumeFusion.cu:423, 439, 687) with:
//   af::getKnnData<K>          host nanoflann search (include/af/GraphBuilder.h:20-32,
//                              include/af/MotionGraph.h:50-70)
//   af::runComputeWarpWeightsKernel  (src/VertexManipulation.cu:105-145)
//   af::runWarpKernel                (src/ComputeCorrespondence.cu:5-43)
// Here all three run on the device, separately (drop-in) or fused in one
// kernel.  The knn is exact: a uniform grid over the nodes, built on the
// device for each call, is searched in rings of cells around each point until
// every cell left is provably farther than the K-th distance; the K best are
// kept in registers by (distance, node index), and a node whose squared
// distance is safely above the current K-th distance is rejected before the
// square root.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <mutex>

#include "vf_host.hpp"

namespace vf {
namespace {

constexpr int kPtThreads = 256;

// Sorted insertion of `key` into keys[0..KK) (ascending), dropping the largest.
template <int KK>
__device__ __forceinline__ void insert_key(unsigned long long* keys, unsigned long long key) {
#pragma unroll
    for (int s = 0; s < KK; ++s) {
        const bool lt = key < keys[s];
        const unsigned long long lo = lt ? key : keys[s];
        const unsigned long long hi = lt ? keys[s] : key;
        keys[s] = lo;
        key = hi;
    }
}

// Squared distance below which a node can still enter a full list: every s
// above it has RN(sqrt(s)) > dK.  dK^2 (1 + 2^-20) rounds to at least
// dK^2 (1 + 2^-21); sqrt of more than that exceeds dK by more than an ulp.
__device__ __forceinline__ float reject_above(unsigned long long kth) {
    if (kth == ~0ull) return __builtin_inff();
    const float dk = __uint_as_float((uint32_t)(kth >> 32));
    const float t = dk * dk;
    return t + t * 0x1p-20f;
}

// ---- uniform grid over the nodes --------------------------------------------
// Built on the device for each call (no host round trip): D cells per axis
// over the finite nodes' bounding box, D fixed by the node count.  Nodes with
// a non-finite coordinate go to an extra list that every point tests.  A
// point visits its cell, then rings of cells around it, and stops once every
// unvisited cell lies farther than its K-th distance with margin, so the
// result equals the brute-force search over all nodes.
struct NodeGrid {
    float lo[3];   // grid origin (finite nodes' minimum)
    float h[3];    // cell size
    float inv[3];  // 1 / h
    float tol[3];  // margin for the rounding of binning and cell planes
    int32_t D;
    uint32_t nextra;
};

constexpr int kGridThreads = 1024;

__device__ __forceinline__ bool finite3(float x, float y, float z) {
    return __builtin_isfinite(x) && __builtin_isfinite(y) && __builtin_isfinite(z);
}

__device__ __forceinline__ int32_t grid_coord(float v, float lo, float inv, int32_t D) {
    const float f = (v - lo) * inv;
    int32_t c = cvt_i32_rz(f);  // truncation; negative or NaN inputs clamp below
    c = c < 0 ? 0 : c;
    return c > D - 1 ? D - 1 : c;
}

// One workgroup: the bounding box, the cell sizes, and zeroed counters.
__global__ __launch_bounds__(kGridThreads) void k_grid_init(const float* __restrict__ nodes, uint32_t G, int32_t D,
                                                            NodeGrid* __restrict__ hdr, uint32_t* __restrict__ count,
                                                            uint32_t* __restrict__ fill) {
    __shared__ float s_min[3][kGridThreads / 64], s_max[3][kGridThreads / 64];
    float mn[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
    float mx[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
    for (uint32_t j = threadIdx.x; j < G; j += blockDim.x) {
        const float x = nodes[3ull * j], y = nodes[3ull * j + 1], z = nodes[3ull * j + 2];
        if (!finite3(x, y, z)) continue;
        mn[0] = fminf(mn[0], x), mn[1] = fminf(mn[1], y), mn[2] = fminf(mn[2], z);
        mx[0] = fmaxf(mx[0], x), mx[1] = fmaxf(mx[1], y), mx[2] = fmaxf(mx[2], z);
    }
    const uint32_t cells = (uint32_t)D * (uint32_t)D * (uint32_t)D;
    for (uint32_t c = threadIdx.x; c <= cells; c += blockDim.x) {  // fill[cells]: the extra list's counter
        if (c < cells) count[c] = 0u;
        fill[c] = 0u;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        for (int o = 32; o > 0; o >>= 1) {
            mn[a] = fminf(mn[a], __shfl_xor(mn[a], o));
            mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], o));
        }
    }
    const int wave = threadIdx.x >> 6;
    if (__lane_id() == 0) {
#pragma unroll
        for (int a = 0; a < 3; ++a) s_min[a][wave] = mn[a], s_max[a][wave] = mx[a];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        NodeGrid g;
        for (int a = 0; a < 3; ++a) {
            float lo = __builtin_inff(), hi = -__builtin_inff();
            for (int w = 0; w < kGridThreads / 64; ++w) lo = fminf(lo, s_min[a][w]), hi = fmaxf(hi, s_max[a][w]);
            if (!(lo <= hi)) lo = hi = 0.0f;  // no finite node
            float ext = hi - lo;
            if (!(ext > 0.0f)) ext = 1.0f;
            g.lo[a] = lo;
            g.h[a] = ext / (float)D;
            if (!(g.h[a] > 0.0f) || !__builtin_isfinite(g.h[a])) g.h[a] = 1.0f;
            g.inv[a] = 1.0f / g.h[a];
            // a cell coordinate and a plane lo + k h are each within a few
            // ulps of the coordinates involved; 2^-20 of their magnitude
            // and 10^-3 of a cell cover both
            g.tol[a] = 0x1p-20f * (fabsf(lo) + (float)D * g.h[a]) + 1e-3f * g.h[a];
        }
        g.D = D;
        g.nextra = 0u;
        *hdr = g;
    }
}

__global__ __launch_bounds__(256) void k_grid_count(const float* __restrict__ nodes, uint32_t G,
                                                    NodeGrid* __restrict__ hdr, uint32_t* __restrict__ count) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= G) return;
    const NodeGrid g = *hdr;
    const float x = nodes[3ull * j], y = nodes[3ull * j + 1], z = nodes[3ull * j + 2];
    if (!finite3(x, y, z)) {
        atomicAdd(&hdr->nextra, 1u);
        return;
    }
    const int32_t cx = grid_coord(x, g.lo[0], g.inv[0], g.D), cy = grid_coord(y, g.lo[1], g.inv[1], g.D),
                  cz = grid_coord(z, g.lo[2], g.inv[2], g.D);
    atomicAdd(count + ((uint32_t)cz * g.D + (uint32_t)cy) * g.D + (uint32_t)cx, 1u);
}

// One workgroup: exclusive scan of the cell counts into start[0..cells].
__global__ __launch_bounds__(kGridThreads) void k_grid_scan(const uint32_t* __restrict__ count, uint32_t cells,
                                                            uint32_t* __restrict__ start) {
    __shared__ uint32_t s_sum[kGridThreads];
    const uint32_t per = (cells + kGridThreads - 1) / kGridThreads;
    const uint32_t b = threadIdx.x * per, e = b + per < cells ? b + per : cells;
    uint32_t t = 0;
    for (uint32_t c = b; c < e; ++c) t += count[c];
    s_sum[threadIdx.x] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int i = 0; i < kGridThreads; ++i) {
            const uint32_t v = s_sum[i];
            s_sum[i] = run;
            run += v;
        }
        start[cells] = run;
    }
    __syncthreads();
    uint32_t run = s_sum[threadIdx.x];
    for (uint32_t c = b; c < e; ++c) {
        start[c] = run;
        run += count[c];
    }
}

// Nodes by cell as (x, y, z, index bits); the extra list after them.
__global__ __launch_bounds__(256) void k_grid_fill(const float* __restrict__ nodes, uint32_t G,
                                                   const NodeGrid* __restrict__ hdr,
                                                   const uint32_t* __restrict__ start, uint32_t* __restrict__ fill,
                                                   float4* __restrict__ sorted) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= G) return;
    const NodeGrid g = *hdr;
    const float x = nodes[3ull * j], y = nodes[3ull * j + 1], z = nodes[3ull * j + 2];
    const uint32_t cells = (uint32_t)g.D * (uint32_t)g.D * (uint32_t)g.D;
    uint32_t slot;
    if (!finite3(x, y, z)) {
        slot = start[cells] + atomicAdd(fill + cells, 1u);
    } else {
        const int32_t cx = grid_coord(x, g.lo[0], g.inv[0], g.D), cy = grid_coord(y, g.lo[1], g.inv[1], g.D),
                      cz = grid_coord(z, g.lo[2], g.inv[2], g.D);
        const uint32_t c = ((uint32_t)cz * g.D + (uint32_t)cy) * g.D + (uint32_t)cx;
        slot = start[c] + atomicAdd(fill + c, 1u);
    }
    sorted[slot] = make_float4(x, y, z, __uint_as_float(j));
}

// One node against the K(+1) best so far (the brute force's test).
template <int KK>
__device__ __forceinline__ void knn_offer(const float p[3], const float4 nd, unsigned long long keys[KK], float& thr) {
    const float dx = p[0] - nd.x, dy = p[1] - nd.y, dz = p[2] - nd.z;
    const float s2 = (dx * dx + dy * dy) + dz * dz;
    if (s2 > thr) return;
    const float d = __builtin_sqrtf(s2);
    const unsigned long long key = ((unsigned long long)__float_as_uint(d) << 32) | __float_as_uint(nd.w);
    if (key < keys[KK - 1]) {
        insert_key<KK>(keys, key);
        thr = reject_above(keys[KK - 1]);
    }
}

// The K(+1) nearest nodes of p by rings of grid cells.  Stops when every cell
// not yet visited is farther than the K-th distance: its lower bound L (the
// distance from p to the visited block's faces, less the grid's rounding
// margin and 2^-20 of |p|) exceeds dK (1 + 2^-18), while a computed distance
// is within 2^-21 (relative) of the true one.  A point still unsettled after
// kMaxRing rings (far from every node) starts over with one pass over all
// nodes.  The walk hands out one node range at a time (the extra list, then
// cell after cell).
struct RingWalk {
    int32_t c[3];
    int32_t r, D;
    int32_t x, y, z;   // current cell
    int32_t xe, step;  // last x of the current row and the stride to it
    int32_t b0[3], b1[3];  // the ring's clamped block
    bool finite;
};

__device__ __forceinline__ void walk_row(RingWalk& w) {
    const bool face = !w.finite || abs(w.z - w.c[2]) == w.r || abs(w.y - w.c[1]) == w.r;
    if (face) {
        w.x = w.b0[0];
        w.xe = w.b1[0];
        w.step = 1;
    } else {
        const bool lo_ok = w.c[0] - w.r >= 0, hi_ok = w.c[0] + w.r <= w.D - 1;
        w.x = lo_ok ? w.c[0] - w.r : (hi_ok ? w.c[0] + w.r : 1);
        w.xe = hi_ok ? w.c[0] + w.r : (lo_ok ? w.c[0] - w.r : 0);
        w.step = 2 * w.r;
    }
}

__device__ __forceinline__ void walk_ring(RingWalk& w, int32_t r) {
    w.r = r;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        w.b0[a] = max(w.c[a] - r, 0);
        w.b1[a] = min(w.c[a] + r, w.D - 1);
    }
    w.z = w.b0[2];
    w.y = w.b0[1];
    walk_row(w);
}

// Moves to the next cell of the ring; false when the ring is done.
__device__ __forceinline__ bool walk_next(RingWalk& w) {
    w.x += w.step;
    while (w.x > w.xe) {
        if (++w.y > w.b1[1]) {
            w.y = w.b0[1];
            if (++w.z > w.b1[2]) return false;
        }
        walk_row(w);
    }
    return true;
}

constexpr int32_t kMaxRing = 2;  // rings searched before a point falls back to every node

template <int KK>
__device__ __forceinline__ void grid_knn(const float p[3], const NodeGrid& g, const uint32_t* __restrict__ start,
                                         const float4* __restrict__ sorted, unsigned long long keys[KK]) {
#pragma unroll
    for (int s = 0; s < KK; ++s) keys[s] = ~0ull;
    float thr = __builtin_inff();
    RingWalk w;
    w.D = g.D;
    w.finite = finite3(p[0], p[1], p[2]);
#pragma unroll
    for (int a = 0; a < 3; ++a) w.c[a] = w.finite ? grid_coord(p[a], g.lo[a], g.inv[a], g.D) : 0;
    // a non-finite point visits every cell in one ring
    walk_ring(w, w.finite ? 0 : g.D);
    const uint32_t D = (uint32_t)g.D;
    const uint32_t cells = D * D * D;
    // first range: the extra list (nodes with a non-finite coordinate)
    uint32_t e = start[cells], e1 = e + g.nextra;
    bool cell_pending = w.x <= w.xe;  // the ring's first cell is still to visit
    for (;;) {
        for (; e < e1; ++e) knn_offer<KK>(p, sorted[e], keys, thr);
        if (!cell_pending) {
            // the ring is done: stop, or start the next one
            if (!w.finite) break;
            if (w.b0[0] == 0 && w.b0[1] == 0 && w.b0[2] == 0 && w.b1[0] == w.D - 1 && w.b1[1] == w.D - 1 &&
                w.b1[2] == w.D - 1)
                break;  // every cell visited
            if (keys[KK - 1] != ~0ull) {
                float L = __builtin_inff();
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const float tol = g.tol[a] + 0x1p-20f * fabsf(p[a]);
                    if (w.b0[a] > 0) L = fminf(L, p[a] - (g.lo[a] + (float)w.b0[a] * g.h[a]) - tol);
                    if (w.b1[a] < w.D - 1) L = fminf(L, (g.lo[a] + (float)(w.b1[a] + 1) * g.h[a]) - p[a] - tol);
                }
                const float dk = __uint_as_float((uint32_t)(keys[KK - 1] >> 32));
                if (L > dk + dk * 0x1p-18f) break;
            }
            if (w.r >= kMaxRing) {
                // not settled near the point (it lies far from the nodes):
                // start over with one pass over every node
#pragma unroll
                for (int s = 0; s < KK; ++s) keys[s] = ~0ull;
                thr = __builtin_inff();
                const uint32_t all = start[cells] + g.nextra;
                for (uint32_t j = 0; j < all; ++j) knn_offer<KK>(p, sorted[j], keys, thr);
                break;
            }
            walk_ring(w, w.r + 1);
            cell_pending = w.x <= w.xe;
            if (!cell_pending) cell_pending = walk_next(w);
            if (!cell_pending) continue;
        }
        const uint32_t cell = ((uint32_t)w.z * D + (uint32_t)w.y) * D + (uint32_t)w.x;
        e = start[cell];
        e1 = start[cell + 1];
        cell_pending = walk_next(w);
    }
}

// Each node's dual quaternion (DQB mode).
__global__ __launch_bounds__(256) void k_node_dq(const float* __restrict__ T, uint32_t G, float4* __restrict__ qr,
                                                 float4* __restrict__ qd) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < G) mat4_to_dq(T + 16ull * j, qr + j, qd + j);
}

template <int K, bool EXCL>
__global__ __launch_bounds__(kPtThreads) void k_mesh_knn(const float* __restrict__ pts, uint64_t n,
                                                         const NodeGrid* __restrict__ hdr,
                                                         const uint32_t* __restrict__ start,
                                                         const float4* __restrict__ sorted,
                                                         uint32_t* __restrict__ idx, float* __restrict__ dist) {
    constexpr int KK = K + (EXCL ? 1 : 0);
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const NodeGrid g = *hdr;
    const float p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    unsigned long long keys[KK];
    grid_knn<KK>(p, g, start, sorted, keys);
#pragma unroll
    for (int s = 0; s < K; ++s) {
        const unsigned long long k = keys[s + (EXCL ? 1 : 0)];
        const bool have = k != ~0ull;
        idx[(uint64_t)K * i + s] = have ? (uint32_t)k : 0xffffffffu;
        dist[(uint64_t)K * i + s] = have ? __uint_as_float((uint32_t)(k >> 32)) : FLT_MAX;
    }
}

// computeWarpWeightsKernel body for one point (src/VertexManipulation.cu:105-136).
template <int K>
__device__ __forceinline__ void point_weights(const uint32_t* id, const float* d, const float* __restrict__ radii,
                                              uint32_t G, const double* tab, float w[K]) {
    float sum = 0.0f;
#pragma unroll
    for (int s = 0; s < K; ++s) {
        float x = 0.0f;
        if (d[s] < FLT_MAX && (G == 0 || id[s] < G)) {
            const float r = radii[id[s]];
            if (!(d[s] >= r)) {
                const float r2 = r * r;
                const float den = 2.0f * r2;
                x = expw(-(d[s] * d[s]) / den, tab);
            }
        }
        w[s] = x;
        sum = sum + x;
    }
    if (sum != 0.0f) {
        const float rs = 1.0f / sum;
#pragma unroll
        for (int s = 0; s < K; ++s) w[s] = div_rcp(w[s], sum, rs);
    }
}

// warpKernel body for one point (src/ComputeCorrespondence.cu:5-30): Eigen's
// four-term row reduction (a0 + a1) + (a2 + a3) of the materialised Vec4f.
// A slot of weight exactly 0 adds nothing (for a finite transform w t is a
// signed zero and the sum, never -0, is unchanged), so a missing neighbour's
// index is never read.
template <int K>
__device__ __forceinline__ void point_lbs(const float p[3], const uint32_t* id, const float* w,
                                          const float* __restrict__ T, uint32_t G, float o[3]) {
    bool zero = true;
#pragma unroll
    for (int s = 0; s < K; ++s) zero = zero && (__builtin_fabsf(w[s]) <= 1e-5f);
    if (zero) {
        o[0] = p[0];
        o[1] = p[1];
        o[2] = p[2];
        return;
    }
    float acc[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < K; ++s) {
        if (w[s] == 0.0f) continue;  // w t would be a signed zero: the sum is unchanged
        float t[3] = {0.0f, 0.0f, 0.0f};
        if (G == 0 || id[s] < G) {
            const float* M = T + 16ull * id[s];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float a0 = M[0 * 4 + a] * p[0];
                const float a1 = M[1 * 4 + a] * p[1];
                const float a2 = M[2 * 4 + a] * p[2];
                const float a3 = M[3 * 4 + a] * 1.0f;
                t[a] = (a0 + a1) + (a2 + a3);
            }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float wt = w[s] * t[a];
            acc[a] = acc[a] + wt;
        }
    }
    o[0] = acc[0];
    o[1] = acc[1];
    o[2] = acc[2];
}

template <int K>
__global__ __launch_bounds__(256) void k_warp_weights(float* __restrict__ w_out, const float* __restrict__ radii,
                                                      const uint32_t* __restrict__ idx,
                                                      const float* __restrict__ dist, uint64_t n, uint32_t G) {
    __shared__ double s_exp2[32];
    if (threadIdx.x < 32) s_exp2[threadIdx.x] = kExp2Table[threadIdx.x];
    __syncthreads();
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t id[K];
    float d[K], w[K];
#pragma unroll
    for (int s = 0; s < K; ++s) {
        id[s] = idx[(uint64_t)K * i + s];
        d[s] = dist[(uint64_t)K * i + s];
    }
    point_weights<K>(id, d, radii, G, s_exp2, w);
#pragma unroll
    for (int s = 0; s < K; ++s) w_out[(uint64_t)K * i + s] = w[s];
}

template <int K>
__global__ __launch_bounds__(256) void k_warp_mesh(float* out, const float* pts,
                                                   const float* __restrict__ T, const uint32_t* __restrict__ idx,
                                                   const float* __restrict__ wts, uint64_t n, uint32_t G) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t id[K];
    float w[K];
#pragma unroll
    for (int s = 0; s < K; ++s) {
        id[s] = idx[(uint64_t)K * i + s];
        w[s] = wts[(uint64_t)K * i + s];
    }
    const float p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    float o[3];
    point_lbs<K>(p, id, w, T, G, o);
    out[3 * i] = o[0];
    out[3 * i + 1] = o[1];
    out[3 * i + 2] = o[2];
}

// knn + weights + blend for each point.  out may alias pts.
template <int K, int BLEND>
__global__ __launch_bounds__(kPtThreads) void k_warp_points(float* out, const float* pts, uint64_t n,
                                                            const NodeGrid* __restrict__ hdr,
                                                            const uint32_t* __restrict__ start,
                                                            const float4* __restrict__ sorted,
                                                            const float* __restrict__ radii,
                                                            const float* __restrict__ T,
                                                            const float4* __restrict__ qr,
                                                            const float4* __restrict__ qd, uint32_t G) {
    __shared__ double s_exp2[32];
    if (threadIdx.x < 32) s_exp2[threadIdx.x] = kExp2Table[threadIdx.x];
    __syncthreads();
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const NodeGrid g = *hdr;
    const float p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    unsigned long long keys[K];
    grid_knn<K>(p, g, start, sorted, keys);
    uint32_t id[K];
    float d[K], w[K];
#pragma unroll
    for (int s = 0; s < K; ++s) {
        const bool have = keys[s] != ~0ull;
        id[s] = have ? (uint32_t)keys[s] : 0xffffffffu;
        d[s] = have ? __uint_as_float((uint32_t)(keys[s] >> 32)) : FLT_MAX;
    }
    point_weights<K>(id, d, radii, G, s_exp2, w);
    float o[3];
    if (BLEND == VF_BLEND_LBS) {
        point_lbs<K>(p, id, w, T, G, o);
    } else {
        bool zero = true;
#pragma unroll
        for (int s = 0; s < K; ++s) zero = zero && (__builtin_fabsf(w[s]) <= 1e-5f);
        if (zero) {
            o[0] = p[0];
            o[1] = p[1];
            o[2] = p[2];
        } else {
            // slots in order, missing neighbours skipped; the first present
            // slot's rotation is the hemisphere reference
            float4 br = make_float4(0.0f, 0.0f, 0.0f, 0.0f), bd = br;
            const float4 f = qr[id[0] < G ? id[0] : 0];
#pragma unroll
            for (int s = 0; s < K; ++s)
                if (id[s] < G) dqb_accumulate(qr[id[s]], qd[id[s]], f, w[s], &br, &bd);
            dqb_transform(br, bd, p, o);
        }
    }
    out[3 * i] = o[0];
    out[3 * i + 1] = o[1];
    out[3 * i + 2] = o[2];
}

inline unsigned grid_for(uint64_t n, int block) { return (unsigned)((n + block - 1) / block); }

// The static scratch of vf_warp_points_static: 4 MiB per device holds the
// dual quaternions and node grid of graphs up to 65536 nodes.
constexpr uint64_t kStaticScratchBytes = 4ull << 20;
constexpr uint32_t kStaticMaxNodes = 65536;
constexpr int kStaticMaxDevices = 64;

// Cells per axis for G nodes: about two nodes per cell of a full grid
// (nodes on a surface fill fewer cells, more densely).
inline int32_t grid_dim(uint32_t G) {
    int32_t D = (int32_t)std::lround(std::cbrt((double)G / 2.0));
    return D < 1 ? 1 : (D > 64 ? 64 : D);
}

struct GridBufs {
    NodeGrid* hdr;
    uint32_t* count;
    uint32_t* fill;
    uint32_t* start;
    float4* sorted;
    uint64_t bytes;
};

inline uint64_t align256(uint64_t b) { return (b + 255) & ~255ull; }

// Scratch layout of the grid (after `base` bytes of other scratch).
inline GridBufs grid_layout(char* base, uint32_t G) {
    const uint32_t D = (uint32_t)grid_dim(G);
    const uint64_t cells = (uint64_t)D * D * D;
    GridBufs b;
    uint64_t off = 0;
    b.hdr = (NodeGrid*)(base + off);
    off += align256(sizeof(NodeGrid));
    b.count = (uint32_t*)(base + off);
    off += align256(4 * cells);
    b.fill = (uint32_t*)(base + off);
    off += align256(4 * (cells + 1));
    b.start = (uint32_t*)(base + off);
    off += align256(4 * (cells + 1));
    b.sorted = (float4*)(base + off);
    off += align256(sizeof(float4) * (uint64_t)std::max<uint32_t>(G, 1));
    b.bytes = off;
    return b;
}

int build_grid(const GridBufs& b, const float* nodes, uint32_t G, hipStream_t stream) {
    const int32_t D = grid_dim(G);
    const uint32_t cells = (uint32_t)(D * D * D);
    hipLaunchKernelGGL(k_grid_init, dim3(1), dim3(kGridThreads), 0, stream, nodes, G, D, b.hdr, b.count, b.fill);
    VF_LAUNCH_CHECK("k_grid_init");
    if (G) {
        hipLaunchKernelGGL(k_grid_count, dim3(grid_for(G, 256)), dim3(256), 0, stream, nodes, G, b.hdr, b.count);
        VF_LAUNCH_CHECK("k_grid_count");
    }
    hipLaunchKernelGGL(k_grid_scan, dim3(1), dim3(kGridThreads), 0, stream, b.count, cells, b.start);
    VF_LAUNCH_CHECK("k_grid_scan");
    if (G) {
        hipLaunchKernelGGL(k_grid_fill, dim3(grid_for(G, 256)), dim3(256), 0, stream, nodes, G, b.hdr, b.start,
                           b.fill, b.sorted);
        VF_LAUNCH_CHECK("k_grid_fill");
    }
    return VF_OK;
}

}  // namespace
}  // namespace vf

using namespace vf;

#define VF_PT_K_SWITCH(k, MACRO) \
    switch (k) {                 \
        case 1: MACRO(1); break; \
        case 2: MACRO(2); break; \
        case 3: MACRO(3); break; \
        case 4: MACRO(4); break; \
        case 5: MACRO(5); break; \
        case 6: MACRO(6); break; \
        case 7: MACRO(7); break; \
        case 8: MACRO(8); break; \
    }

// Scratch of vf_warp_points_static: a file-scope device array, one copy per
// GPU, so that the call allocates nothing.  Its address on each device is
// taken once (hipGetSymbolAddress) and kept; every call on a device shares
// that device's copy, so calls that overlap on one device race.  The kernels
// re-initialise every byte they read (grid counters, header, node lists, dual
// quaternions) on each call.
static __device__ __align__(256) unsigned char scratch[vf::kStaticScratchBytes];

namespace vf {
namespace {

std::mutex g_static_mutex;
void* g_static_ptr[kStaticMaxDevices] = {};

// The current device's address of `scratch` (taken on the first call there).
int static_scratch(void** out) {
    int dev = 0;
    VF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kStaticMaxDevices)
        return fail(VF_ERR_INVALID_ARGUMENT, "device %d: the static scratch serves devices 0..%d", dev,
                    kStaticMaxDevices - 1);
    std::lock_guard<std::mutex> lock(g_static_mutex);
    if (!g_static_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(VF_ERR_HIP, "hipGetSymbolAddress of the static scratch on device %d: %s", dev,
                        hipGetErrorString(e));
        g_static_ptr[dev] = p;
    }
    *out = g_static_ptr[dev];
    return VF_OK;
}

int check_warp_points(int k, int blend, vf_float3* warped, const vf_float3* points, size_t n, const vf_float3* nodes,
                      const float* radii, const vf_mat4* transforms, size_t graph_size) {
    int st = check_k(k);
    if (st) return st;
    if (blend != VF_BLEND_LBS && blend != VF_BLEND_DQB)
        return fail(VF_ERR_INVALID_ARGUMENT, "blend mode %d: supported VF_BLEND_LBS (0), VF_BLEND_DQB (1)", blend);
    if (!n) return VF_OK;
    if (!warped || !points || (graph_size && (!nodes || !radii || !transforms)))
        return fail(VF_ERR_INVALID_ARGUMENT, "vf_warp_points: NULL argument");
    if (graph_size > 0xffffffffull) return fail(VF_ERR_INVALID_ARGUMENT, "graph_size exceeds 2^32-1");
    return VF_OK;
}

// Bytes of scratch the point warp uses for G nodes: the nodes' dual
// quaternions, then the node grid.
uint64_t warp_points_scratch_bytes(uint32_t G) {
    const uint64_t g1 = std::max<uint32_t>(G, 1);
    return align256(sizeof(float4) * g1 * 2) + grid_layout(nullptr, G).bytes;
}

int warp_points_on(void* mem, int k, int blend, vf_float3* warped, const vf_float3* points, size_t n,
                   const vf_float3* nodes, const float* radii, const vf_mat4* transforms, uint32_t G,
                   hipStream_t stream);

}  // namespace
}  // namespace vf

extern "C" {

int vf_compute_warp_weights(int k, float* knn_weights, const float* graph_radii, const uint32_t* knn_idxs,
                            const float* knn_dists, size_t mesh_size, size_t graph_size, vf_stream stream) {
    int st = check_k(k);
    if (st) return st;
    if (!mesh_size) return VF_OK;
    if (!knn_weights || !graph_radii || !knn_idxs || !knn_dists)
        return fail(VF_ERR_INVALID_ARGUMENT, "vf_compute_warp_weights: NULL buffer");
    if (graph_size > 0xffffffffull) return fail(VF_ERR_INVALID_ARGUMENT, "graph_size exceeds 2^32-1");
#define VF_WW(KK)                                                                                                 \
    hipLaunchKernelGGL(k_warp_weights<KK>, dim3(grid_for(mesh_size, 256)), dim3(256), 0, (hipStream_t)stream,   \
                       knn_weights, graph_radii, knn_idxs, knn_dists, (uint64_t)mesh_size, (uint32_t)graph_size)
    VF_PT_K_SWITCH(k, VF_WW)
#undef VF_WW
    VF_LAUNCH_CHECK("k_warp_weights");
    return VF_OK;
}

int vf_warp_mesh(int k, vf_float3* warped_mesh, const vf_float3* mesh, const vf_mat4* motion_graph,
                 const uint32_t* knn_idxs, const float* knn_weights, size_t mesh_size, size_t graph_size,
                 vf_stream stream) {
    int st = check_k(k);
    if (st) return st;
    if (!mesh_size) return VF_OK;
    if (!warped_mesh || !mesh || !motion_graph || !knn_idxs || !knn_weights)
        return fail(VF_ERR_INVALID_ARGUMENT, "vf_warp_mesh: NULL buffer");
    if (graph_size > 0xffffffffull) return fail(VF_ERR_INVALID_ARGUMENT, "graph_size exceeds 2^32-1");
#define VF_WM(KK)                                                                                                 \
    hipLaunchKernelGGL(k_warp_mesh<KK>, dim3(grid_for(mesh_size, 256)), dim3(256), 0, (hipStream_t)stream,      \
                       (float*)warped_mesh, (const float*)mesh, (const float*)motion_graph, knn_idxs, knn_weights, \
                       (uint64_t)mesh_size, (uint32_t)graph_size)
    VF_PT_K_SWITCH(k, VF_WM)
#undef VF_WM
    VF_LAUNCH_CHECK("k_warp_mesh");
    return VF_OK;
}

int vf_mesh_knn(vf_workspace* ws, int k, uint32_t* knn_idxs, float* knn_dists, const vf_float3* points, size_t n,
                const vf_float3* nodes, size_t graph_size, int exclude_closest, vf_stream stream) {
    int st = check_k(k);
    if (st) return st;
    if (!n) return VF_OK;
    if (!knn_idxs || !knn_dists || !points || (graph_size && !nodes))
        return fail(VF_ERR_INVALID_ARGUMENT, "vf_mesh_knn: NULL argument");
    if (graph_size > 0xffffffffull) return fail(VF_ERR_INVALID_ARGUMENT, "graph_size exceeds 2^32-1");
    const uint32_t G = (uint32_t)graph_size;
    void* mem = nullptr;
    const GridBufs probe = grid_layout(nullptr, G);
    st = workspace_scratch(ws, probe.bytes, &mem, (hipStream_t)stream);
    if (st) return st;
    const GridBufs gb = grid_layout((char*)mem, G);
    st = build_grid(gb, (const float*)nodes, G, (hipStream_t)stream);
    if (st) return st;
#define VF_KNN(KK)                                                                                                   \
    if (exclude_closest)                                                                                             \
        hipLaunchKernelGGL((k_mesh_knn<KK, true>), dim3(grid_for(n, kPtThreads)), dim3(kPtThreads), 0,              \
                           (hipStream_t)stream, (const float*)points, (uint64_t)n, gb.hdr, gb.start, gb.sorted,      \
                           knn_idxs, knn_dists);                                                                     \
    else                                                                                                             \
        hipLaunchKernelGGL((k_mesh_knn<KK, false>), dim3(grid_for(n, kPtThreads)), dim3(kPtThreads), 0,             \
                           (hipStream_t)stream, (const float*)points, (uint64_t)n, gb.hdr, gb.start, gb.sorted,      \
                           knn_idxs, knn_dists)
    VF_PT_K_SWITCH(k, VF_KNN)
#undef VF_KNN
    VF_LAUNCH_CHECK("k_mesh_knn");
    return VF_OK;
}

int vf_warp_points(vf_workspace* ws, int k, int blend, vf_float3* warped, const vf_float3* points, size_t n,
                   const vf_float3* nodes, const float* radii, const vf_mat4* transforms, size_t graph_size,
                   vf_stream stream) {
    int st = check_warp_points(k, blend, warped, points, n, nodes, radii, transforms, graph_size);
    if (st || !n) return st;
    void* mem = nullptr;
    st = workspace_scratch(ws, warp_points_scratch_bytes((uint32_t)graph_size), &mem, (hipStream_t)stream);
    if (st) return st;
    return warp_points_on(mem, k, blend, warped, points, n, nodes, radii, transforms, (uint32_t)graph_size,
                          (hipStream_t)stream);
}

int vf_warp_points_static(int k, int blend, vf_float3* warped, const vf_float3* points, size_t n,
                          const vf_float3* nodes, const float* radii, const vf_mat4* transforms, size_t graph_size,
                          vf_stream stream) {
    int st = check_warp_points(k, blend, warped, points, n, nodes, radii, transforms, graph_size);
    if (st || !n) return st;
    const uint64_t need = warp_points_scratch_bytes((uint32_t)graph_size);
    if (graph_size > kStaticMaxNodes || need > kStaticScratchBytes)
        return fail(VF_ERR_CAPACITY,
                    "vf_warp_points_static: a graph of %llu nodes needs %llu B of scratch; the library's static "
                    "scratch holds %llu B (graphs up to %u nodes)",
                    (unsigned long long)graph_size, (unsigned long long)need, (unsigned long long)kStaticScratchBytes,
                    kStaticMaxNodes);
    void* mem = nullptr;
    st = static_scratch(&mem);
    if (st) return st;
    return warp_points_on(mem, k, blend, warped, points, n, nodes, radii, transforms, (uint32_t)graph_size,
                          (hipStream_t)stream);
}

}  // extern "C"

namespace vf {
namespace {

int warp_points_on(void* mem, int k, int blend, vf_float3* warped, const vf_float3* points, size_t n,
                   const vf_float3* nodes, const float* radii, const vf_mat4* transforms, uint32_t G,
                   hipStream_t stream) {
    const uint64_t g1 = std::max<uint32_t>(G, 1);
    const uint64_t dq_bytes = align256(sizeof(float4) * g1 * 2);
    float4* qr = (float4*)mem;
    float4* qd = qr + g1;
    const GridBufs gb = grid_layout((char*)mem + dq_bytes, G);
    int st = build_grid(gb, (const float*)nodes, G, stream);
    if (st) return st;
    if (G && blend == VF_BLEND_DQB) {
        hipLaunchKernelGGL(k_node_dq, dim3(grid_for(G, 256)), dim3(256), 0, stream, (const float*)transforms, G, qr,
                           qd);
        VF_LAUNCH_CHECK("k_node_dq");
    }
#define VF_WP(KK)                                                                                                 \
    if (blend == VF_BLEND_DQB)                                                                                    \
        hipLaunchKernelGGL((k_warp_points<KK, VF_BLEND_DQB>), dim3(grid_for(n, kPtThreads)), dim3(kPtThreads), 0, \
                           stream, (float*)warped, (const float*)points, (uint64_t)n, gb.hdr, gb.start, gb.sorted, \
                           radii, (const float*)transforms, qr, qd, G);                                           \
    else                                                                                                          \
        hipLaunchKernelGGL((k_warp_points<KK, VF_BLEND_LBS>), dim3(grid_for(n, kPtThreads)), dim3(kPtThreads), 0, \
                           stream, (float*)warped, (const float*)points, (uint64_t)n, gb.hdr, gb.start, gb.sorted, \
                           radii, (const float*)transforms, qr, qd, G)
    VF_PT_K_SWITCH(k, VF_WP)
#undef VF_WP
    VF_LAUNCH_CHECK("k_warp_points");
    return VF_OK;
}

}  // namespace
}  // namespace vf
