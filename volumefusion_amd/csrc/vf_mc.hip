// vf_mc.hip -- marching cubes over the TSDF (af::runMarchingCubesMeshKernel
// and af::runMarchingCubesFullKernel, src/MarchingCubes.cu:422-702).
//
// The reference appends every vertex (and face) with an atomicAdd on one
// counter, so its output order is a race and the counter a hot spot.  Here
// the output is compacted in a canonical order (cubes by linear index, edges
// by number, triangles in table order) with two sweeps:
//   k_mc_count   per cube: case and output count; per 256-cube block: sum
//   (scan)       exclusive scan of the block sums (rocPRIM)
//   k_mc_emit    per cube again: block-local prefix, then the writes
//   k_mc_finish  advances the caller's counters
// Each sweep reads the 8 corners of every cube (tsdf and weight), so the
// pair is HBM-bound: about 2 x 8 B per voxel.
#include <hipcub/hipcub.hpp>

#include "vf_host.hpp"
#include "vf_mc_tables.h"

namespace vf {
namespace {

constexpr int kMcBlock = 256;

__constant__ unsigned long long kMcTri[256] = VF_MC_TRI_TABLE;
__constant__ int kMcEdge[12][2] = VF_MC_EDGE_CORNERS;
__constant__ int kMcCorner[8][3] = VF_MC_CORNER_OFFSETS;

struct McGeom {
    int32_t X, Y, Z;
    int32_t cx, cy;  // cubes per row (X - 2), per column (Y - 2)
    float vs[3];
    float halfs[3];
    float iso;
    uint64_t ncubes;
};

// interpolate, src/MarchingCubes.cu:294-312: ratio in float, widened,
// clamped, lerp in binary64 rounded to float.
__device__ __forceinline__ void mc_interp(float t0, float t1, const float p0[3], const float p1[3], float iso,
                                          float o[3]) {
    if (__builtin_fabs((double)(iso - t0)) < 1e-7) {
        o[0] = p0[0], o[1] = p0[1], o[2] = p0[2];
        return;
    }
    if (__builtin_fabs((double)(iso - t1)) < 1e-7) {
        o[0] = p1[0], o[1] = p1[1], o[2] = p1[2];
        return;
    }
    if (__builtin_fabs((double)(t0 - t1)) < 1e-7) {
        o[0] = p0[0], o[1] = p0[1], o[2] = p0[2];
        return;
    }
    double mu = (double)((iso - t0) / (t1 - t0));
    if (mu > 1.0)
        mu = 1.0;
    else if (mu < 0)
        mu = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) o[a] = (float)((double)p0[a] + mu * (double)(p1[a] - p0[a]));
}

__device__ __forceinline__ void mc_world(const McGeom& g, int32_t i, int32_t j, int32_t k, float p[3]) {
    p[0] = (float)i * g.vs[0] - g.halfs[0];
    p[1] = (float)j * g.vs[1] - g.halfs[1];
    p[2] = (float)k * g.vs[2] - g.halfs[2];
}

struct Cube {
    int32_t i, j, k;
    int cs;          // case, 0 when any corner weight is 0
    float t[8];      // corner tsdf
};

__device__ __forceinline__ void load_cube(const McGeom& g, uint64_t c, const float* __restrict__ tsdf,
                                          const float* __restrict__ wt, Cube* cb) {
    cb->i = (int32_t)(c % (uint64_t)g.cx);
    cb->j = (int32_t)((c / (uint64_t)g.cx) % (uint64_t)g.cy);
    cb->k = (int32_t)(c / ((uint64_t)g.cx * (uint64_t)g.cy));
    bool zero = false;
    uint64_t off[8];
#pragma unroll
    for (int v = 0; v < 8; ++v)
        off[v] = ((uint64_t)(cb->k + kMcCorner[v][2]) * (uint64_t)g.Y + (uint64_t)(cb->j + kMcCorner[v][1])) *
                     (uint64_t)g.X +
                 (uint64_t)(cb->i + kMcCorner[v][0]);
#pragma unroll
    for (int v = 0; v < 8; ++v) zero = zero || (wt[off[v]] == 0.0f);
    int cs = 0;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        cb->t[v] = tsdf[off[v]];
        if (cb->t[v] > g.iso) cs |= 1 << v;
    }
    cb->cs = zero ? 0 : cs;
}

__device__ __forceinline__ int edge_mask(int cs) {
    int m = 0;
#pragma unroll
    for (int e = 0; e < 12; ++e)
        if (((cs >> kMcEdge[e][0]) ^ (cs >> kMcEdge[e][1])) & 1) m |= 1 << e;
    return m;
}

__device__ __forceinline__ void edge_point(const McGeom& g, const Cube& cb, int e, float o[3]) {
    const int a = kMcEdge[e][0], b = kMcEdge[e][1];
    float pa[3], pb[3];
    mc_world(g, cb.i + kMcCorner[a][0], cb.j + kMcCorner[a][1], cb.k + kMcCorner[a][2], pa);
    mc_world(g, cb.i + kMcCorner[b][0], cb.j + kMcCorner[b][1], cb.k + kMcCorner[b][2], pb);
    mc_interp(cb.t[a], cb.t[b], pa, pb, g.iso, o);
}

__device__ __forceinline__ bool same3(const float* a, const float* b) {
    return a[0] == b[0] && a[1] == b[1] && a[2] == b[2];
}

// Output vertex count of a cube; with FULL, its edge points are left in ep.
template <bool FULL>
__device__ __forceinline__ int cube_outputs(const McGeom& g, const Cube& cb, float ep[12][3], unsigned& keep) {
    keep = 0;
    if (cb.cs == 0 || cb.cs == 255) return 0;
    const int mask = edge_mask(cb.cs);
    if (!FULL) return __popc(mask);
#pragma unroll
    for (int e = 0; e < 12; ++e)
        if (mask & (1 << e)) edge_point(g, cb, e, ep[e]);
    const unsigned long long t = kMcTri[cb.cs];
    const int nt = (int)(t & 15);
    int n = 0;
    for (int q = 0; q < nt; ++q) {
        const float* p0 = ep[(t >> (4 + 12 * q)) & 15];
        const float* p1 = ep[(t >> (8 + 12 * q)) & 15];
        const float* p2 = ep[(t >> (12 + 12 * q)) & 15];
        if (!same3(p0, p1) && !same3(p0, p2) && !same3(p1, p2)) {
            keep |= 1u << q;
            n += 3;
        }
    }
    return n;
}

template <bool FULL>
__global__ __launch_bounds__(kMcBlock) void k_mc_count(McGeom g, const float* __restrict__ tsdf,
                                                        const float* __restrict__ wt, uint32_t* __restrict__ sums) {
    typedef hipcub::BlockReduce<uint32_t, kMcBlock> Reduce;
    __shared__ typename Reduce::TempStorage tmp;
    const uint64_t c = (uint64_t)blockIdx.x * kMcBlock + threadIdx.x;
    uint32_t n = 0;
    if (c < g.ncubes) {
        Cube cb;
        load_cube(g, c, tsdf, wt, &cb);
        float ep[12][3];
        unsigned keep;
        n = (uint32_t)cube_outputs<FULL>(g, cb, ep, keep);
    }
    const uint32_t total = Reduce(tmp).Sum(n);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

template <bool FULL>
__global__ __launch_bounds__(kMcBlock) void k_mc_emit(McGeom g, const float* __restrict__ tsdf,
                                                       const float* __restrict__ wt,
                                                       const uint32_t* __restrict__ offsets,
                                                       const uint32_t* __restrict__ mesh_size_in,
                                                       const uint32_t* __restrict__ faces_size_in,
                                                       float* __restrict__ mesh, uint64_t mesh_cap,
                                                       int32_t* __restrict__ faces, uint64_t faces_cap) {
    typedef hipcub::BlockScan<uint32_t, kMcBlock> Scan;
    __shared__ typename Scan::TempStorage tmp;
    const uint64_t c = (uint64_t)blockIdx.x * kMcBlock + threadIdx.x;
    Cube cb;
    cb.cs = 0;
    float ep[12][3];
    unsigned keep = 0;
    uint32_t n = 0;
    if (c < g.ncubes) {
        load_cube(g, c, tsdf, wt, &cb);
        n = (uint32_t)cube_outputs<FULL>(g, cb, ep, keep);
    }
    uint32_t prefix;
    Scan(tmp).ExclusiveSum(n, prefix);
    if (!n) return;
    const uint64_t rel = (uint64_t)offsets[blockIdx.x] + prefix;  // this call's vertex index
    const uint64_t v0 = (uint64_t)*mesh_size_in + rel;            // appended after the caller's count
    if (!FULL) {
        const int mask = edge_mask(cb.cs);
        uint64_t v = v0;
        for (int e = 0; e < 12; ++e) {
            if (!(mask & (1 << e))) continue;
            float o[3];
            edge_point(g, cb, e, o);
            if (v < mesh_cap) {
                mesh[3 * v] = o[0];
                mesh[3 * v + 1] = o[1];
                mesh[3 * v + 2] = o[2];
            }
            ++v;
        }
        return;
    }
    const unsigned long long t = kMcTri[cb.cs];
    const int nt = (int)(t & 15);
    uint64_t v = v0;
    uint64_t f = (uint64_t)*faces_size_in + rel / 3;
    for (int q = 0; q < nt; ++q) {
        if (!(keep & (1u << q))) continue;
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            const float* p = ep[(t >> (4 + 4 * (3 * q + s))) & 15];
            if (v + s < mesh_cap) {
                mesh[3 * (v + s)] = p[0];
                mesh[3 * (v + s) + 1] = p[1];
                mesh[3 * (v + s) + 2] = p[2];
            }
        }
        if (f < faces_cap) {
            faces[3 * f] = (int32_t)v;
            faces[3 * f + 1] = (int32_t)(v + 1);
            faces[3 * f + 2] = (int32_t)(v + 2);
        }
        v += 3;
        f += 1;
    }
}

__global__ void k_mc_finish(const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ sums, uint32_t nb,
                            uint32_t* mesh_size, uint32_t* faces_size) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t total = offsets[nb - 1] + sums[nb - 1];
    *mesh_size = *mesh_size + total;
    if (faces_size) *faces_size = *faces_size + total / 3;
}

int run_mc(vf_workspace* ws, bool full, float* mesh, uint32_t* mesh_size, uint64_t mesh_cap, int32_t* faces,
           uint32_t* faces_size, uint64_t faces_cap, const float* tsdf, const float* wt, vf_int3 dims, vf_float3 vs,
           vf_float3 size, float iso, hipStream_t stream) {
    uint64_t n;
    int st = check_volume(dims, vs, &n);
    if (st) return st;
    if (!mesh || !mesh_size || !tsdf || !wt || (full && (!faces || !faces_size)))
        return fail(VF_ERR_INVALID_ARGUMENT, "marching cubes: NULL argument");
    if (dims.x < 3 || dims.y < 3 || dims.z < 3) return VF_OK;  // no cube is visited
    McGeom g;
    g.X = dims.x;
    g.Y = dims.y;
    g.Z = dims.z;
    g.cx = dims.x - 2;
    g.cy = dims.y - 2;
    g.vs[0] = vs.x;
    g.vs[1] = vs.y;
    g.vs[2] = vs.z;
    g.halfs[0] = size.x * 0.5f;
    g.halfs[1] = size.y * 0.5f;
    g.halfs[2] = size.z * 0.5f;
    g.iso = iso;
    g.ncubes = (uint64_t)g.cx * (uint64_t)g.cy * (uint64_t)(dims.z - 2);
    const uint64_t nb64 = (g.ncubes + kMcBlock - 1) / kMcBlock;
    if (nb64 > 0x7fffffffull) return fail(VF_ERR_CAPACITY, "marching cubes: %llu blocks", (unsigned long long)nb64);
    const uint32_t nb = (uint32_t)nb64;
    size_t scan_bytes = 0;
    VF_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)nb,
                                            stream));
    const uint64_t sums_bytes = (uint64_t)nb * 4;
    const uint64_t scan_off = (2 * sums_bytes + 255) & ~255ull;
    void* scratch = nullptr;
    st = workspace_scratch(ws, scan_off + scan_bytes, &scratch, stream);
    if (st) return st;
    uint32_t* sums = (uint32_t*)scratch;
    uint32_t* offsets = sums + nb;
    void* scan_tmp = (char*)scratch + scan_off;
    if (full) {
        hipLaunchKernelGGL(k_mc_count<true>, dim3(nb), dim3(kMcBlock), 0, stream, g, tsdf, wt, sums);
    } else {
        hipLaunchKernelGGL(k_mc_count<false>, dim3(nb), dim3(kMcBlock), 0, stream, g, tsdf, wt, sums);
    }
    VF_LAUNCH_CHECK("k_mc_count");
    VF_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, sums, offsets, (int)nb, stream));
    if (full) {
        hipLaunchKernelGGL(k_mc_emit<true>, dim3(nb), dim3(kMcBlock), 0, stream, g, tsdf, wt, offsets, mesh_size,
                           faces_size, mesh, mesh_cap, faces, faces_cap);
    } else {
        hipLaunchKernelGGL(k_mc_emit<false>, dim3(nb), dim3(kMcBlock), 0, stream, g, tsdf, wt, offsets, mesh_size,
                           (const uint32_t*)nullptr, mesh, mesh_cap, (int32_t*)nullptr, (uint64_t)0);
    }
    VF_LAUNCH_CHECK("k_mc_emit");
    hipLaunchKernelGGL(k_mc_finish, dim3(1), dim3(64), 0, stream, offsets, sums, nb, mesh_size,
                       full ? faces_size : (uint32_t*)nullptr);
    VF_LAUNCH_CHECK("k_mc_finish");
    return VF_OK;
}

}  // namespace
}  // namespace vf

using namespace vf;

extern "C" {

int vf_marching_cubes_mesh(vf_workspace* ws, vf_float3* mesh, uint32_t* mesh_size, uint64_t mesh_capacity,
                           const float* tsdf, const float* weights, vf_int3 dims, vf_float3 voxel_size,
                           vf_float3 volume_size, float iso, vf_stream stream) {
    return run_mc(ws, false, (float*)mesh, mesh_size, mesh_capacity, nullptr, nullptr, 0, tsdf, weights, dims,
                  voxel_size, volume_size, iso, (hipStream_t)stream);
}

int vf_marching_cubes_full(vf_workspace* ws, vf_float3* mesh, uint32_t* mesh_size, uint64_t mesh_capacity,
                           vf_int3* faces, uint32_t* faces_size, uint64_t faces_capacity, const float* tsdf,
                           const float* weights, vf_int3 dims, vf_float3 voxel_size, vf_float3 volume_size, float iso,
                           vf_stream stream) {
    return run_mc(ws, true, (float*)mesh, mesh_size, mesh_capacity, (int32_t*)faces, faces_size, faces_capacity, tsdf,
                  weights, dims, voxel_size, volume_size, iso, (hipStream_t)stream);
}

}  // extern "C"
