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
// kernel.  The knn is exact: every node is visited (nodes stream through LDS
// in tiles), the K best kept in registers by (distance, node index); a node
// whose squared distance is safely above the current K-th distance is
// rejected before the square root.
#include <cfloat>

#include "vf_host.hpp"

namespace vf {
namespace {

constexpr int kPtThreads = 256;
constexpr int kTile = 1024;  // nodes per LDS tile (16 KB)

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

// The K(+1 when excluding the closest) nearest nodes of point p, sorted.
template <int KK>
__device__ __forceinline__ void point_knn(const float p[3], bool valid, const float4* __restrict__ nodes, uint32_t G,
                                          float4* s_nodes, unsigned long long keys[KK]) {
#pragma unroll
    for (int s = 0; s < KK; ++s) keys[s] = ~0ull;
    float thr = __builtin_inff();
    for (uint32_t base = 0; base < G; base += kTile) {
        const uint32_t m = G - base < (uint32_t)kTile ? G - base : (uint32_t)kTile;
        __syncthreads();
        for (uint32_t e = threadIdx.x; e < m; e += blockDim.x) s_nodes[e] = nodes[base + e];
        __syncthreads();
        if (!valid) continue;
        for (uint32_t e = 0; e < m; ++e) {
            const float4 nd = s_nodes[e];
            const float dx = p[0] - nd.x, dy = p[1] - nd.y, dz = p[2] - nd.z;
            // nanoflann's L2 order, ((0 + dx^2) + dy^2) + dz^2
            const float s2 = (dx * dx + dy * dy) + dz * dz;
            if (s2 > thr) continue;
            const float d = __builtin_sqrtf(s2);
            const unsigned long long key = ((unsigned long long)__float_as_uint(d) << 32) | (base + e);
            if (key < keys[KK - 1]) {
                insert_key<KK>(keys, key);
                thr = reject_above(keys[KK - 1]);
            }
        }
    }
}

// Node positions as float4 (x, y, z, 0) for aligned LDS staging.
__global__ __launch_bounds__(256) void k_pack_nodes(const float* __restrict__ nodes, uint32_t G,
                                                    float4* __restrict__ out) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < G) out[j] = make_float4(nodes[3ull * j], nodes[3ull * j + 1], nodes[3ull * j + 2], 0.0f);
}

// Each node's dual quaternion (DQB mode).
__global__ __launch_bounds__(256) void k_node_dq(const float* __restrict__ T, uint32_t G, float4* __restrict__ qr,
                                                 float4* __restrict__ qd) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < G) mat4_to_dq(T + 16ull * j, qr + j, qd + j);
}

template <int K, bool EXCL>
__global__ __launch_bounds__(kPtThreads) void k_mesh_knn(const float* __restrict__ pts, uint64_t n,
                                                         const float4* __restrict__ nodes, uint32_t G,
                                                         uint32_t* __restrict__ idx, float* __restrict__ dist) {
    constexpr int KK = K + (EXCL ? 1 : 0);
    __shared__ float4 s_nodes[kTile];
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    float p[3] = {0.0f, 0.0f, 0.0f};
    if (valid) {
        p[0] = pts[3 * i];
        p[1] = pts[3 * i + 1];
        p[2] = pts[3 * i + 2];
    }
    unsigned long long keys[KK];
    point_knn<KK>(p, valid, nodes, G, s_nodes, keys);
    if (!valid) return;
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
                                                            const float4* __restrict__ nodes,
                                                            const float* __restrict__ radii,
                                                            const float* __restrict__ T,
                                                            const float4* __restrict__ qr,
                                                            const float4* __restrict__ qd, uint32_t G) {
    __shared__ float4 s_nodes[kTile];
    __shared__ double s_exp2[32];
    if (threadIdx.x < 32) s_exp2[threadIdx.x] = kExp2Table[threadIdx.x];
    __syncthreads();
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    float p[3] = {0.0f, 0.0f, 0.0f};
    if (valid) {
        p[0] = pts[3 * i];
        p[1] = pts[3 * i + 1];
        p[2] = pts[3 * i + 2];
    }
    unsigned long long keys[K];
    point_knn<K>(p, valid, nodes, G, s_nodes, keys);
    if (!valid) return;
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
    void* scratch = nullptr;
    st = workspace_scratch(ws, sizeof(float4) * (uint64_t)std::max<uint32_t>(G, 1), &scratch, (hipStream_t)stream);
    if (st) return st;
    float4* packed = (float4*)scratch;
    if (G) {
        hipLaunchKernelGGL(k_pack_nodes, dim3(grid_for(G, 256)), dim3(256), 0, (hipStream_t)stream, (const float*)nodes,
                           G, packed);
        VF_LAUNCH_CHECK("k_pack_nodes");
    }
#define VF_KNN(KK)                                                                                                  \
    if (exclude_closest)                                                                                            \
        hipLaunchKernelGGL((k_mesh_knn<KK, true>), dim3(grid_for(n, kPtThreads)), dim3(kPtThreads), 0,             \
                           (hipStream_t)stream, (const float*)points, (uint64_t)n, packed, G, knn_idxs, knn_dists); \
    else                                                                                                            \
        hipLaunchKernelGGL((k_mesh_knn<KK, false>), dim3(grid_for(n, kPtThreads)), dim3(kPtThreads), 0,            \
                           (hipStream_t)stream, (const float*)points, (uint64_t)n, packed, G, knn_idxs, knn_dists)
    VF_PT_K_SWITCH(k, VF_KNN)
#undef VF_KNN
    VF_LAUNCH_CHECK("k_mesh_knn");
    return VF_OK;
}

int vf_warp_points(vf_workspace* ws, int k, int blend, vf_float3* warped, const vf_float3* points, size_t n,
                   const vf_float3* nodes, const float* radii, const vf_mat4* transforms, size_t graph_size,
                   vf_stream stream) {
    int st = check_k(k);
    if (st) return st;
    if (blend != VF_BLEND_LBS && blend != VF_BLEND_DQB)
        return fail(VF_ERR_INVALID_ARGUMENT, "blend mode %d: supported VF_BLEND_LBS (0), VF_BLEND_DQB (1)", blend);
    if (!n) return VF_OK;
    if (!warped || !points || (graph_size && (!nodes || !radii || !transforms)))
        return fail(VF_ERR_INVALID_ARGUMENT, "vf_warp_points: NULL argument");
    if (graph_size > 0xffffffffull) return fail(VF_ERR_INVALID_ARGUMENT, "graph_size exceeds 2^32-1");
    const uint32_t G = (uint32_t)graph_size;
    void* scratch = nullptr;
    const uint64_t g1 = std::max<uint32_t>(G, 1);
    st = workspace_scratch(ws, sizeof(float4) * g1 * 3, &scratch, (hipStream_t)stream);
    if (st) return st;
    float4* packed = (float4*)scratch;
    float4* qr = packed + g1;
    float4* qd = qr + g1;
    if (G) {
        hipLaunchKernelGGL(k_pack_nodes, dim3(grid_for(G, 256)), dim3(256), 0, (hipStream_t)stream, (const float*)nodes,
                           G, packed);
        VF_LAUNCH_CHECK("k_pack_nodes");
        if (blend == VF_BLEND_DQB) {
            hipLaunchKernelGGL(k_node_dq, dim3(grid_for(G, 256)), dim3(256), 0, (hipStream_t)stream,
                               (const float*)transforms, G, qr, qd);
            VF_LAUNCH_CHECK("k_node_dq");
        }
    }
#define VF_WP(KK)                                                                                                 \
    if (blend == VF_BLEND_DQB)                                                                                    \
        hipLaunchKernelGGL((k_warp_points<KK, VF_BLEND_DQB>), dim3(grid_for(n, kPtThreads)), dim3(kPtThreads), 0, \
                           (hipStream_t)stream, (float*)warped, (const float*)points, (uint64_t)n, packed, radii,  \
                           (const float*)transforms, qr, qd, G);                                                  \
    else                                                                                                          \
        hipLaunchKernelGGL((k_warp_points<KK, VF_BLEND_LBS>), dim3(grid_for(n, kPtThreads)), dim3(kPtThreads), 0, \
                           (hipStream_t)stream, (float*)warped, (const float*)points, (uint64_t)n, packed, radii,  \
                           (const float*)transforms, qr, qd, G)
    VF_PT_K_SWITCH(k, VF_WP)
#undef VF_WP
    VF_LAUNCH_CHECK("k_warp_points");
    return VF_OK;
}

}  // extern "C"
