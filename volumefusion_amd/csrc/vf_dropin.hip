// vf_dropin.hip -- per-entry-point kernels behind the reference's af::run*
// functions.  Each one materialises exactly the buffer its reference kernel
// writes, on the reference's layouts, so the reference's frame loop can call
// them one by one.  Launch shape: one thread per voxel (or element), waves of
// 64 consecutive voxels along x so every access to the z-major arrays is
// coalesced.
#include "vf_host.hpp"

namespace vf {
namespace {

constexpr int kBlock1D = 256;

inline dim3 grid_1d(uint64_t n) { return dim3((unsigned)((n + kBlock1D - 1) / kBlock1D)); }

// Block of 64 x 4 voxels; grid (ceil(X/64), ceil(Y/4), Z).
inline void grid_3d(const Geom& g, dim3* grid, dim3* block) {
    *block = dim3(64, 4, 1);
    *grid = dim3((unsigned)((g.dim[0] + 63) / 64), (unsigned)((g.dim[1] + 3) / 4), (unsigned)g.dim[2]);
}

__global__ void k_translate(float* __restrict__ mesh, uint32_t n, float tx, float ty, float tz) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float* p = mesh + 3ull * i;
    p[0] = p[0] + tx;
    p[1] = p[1] + ty;
    p[2] = p[2] + tz;
}

__global__ void k_eval_weight(const float* __restrict__ radius, const float* __restrict__ dist,
                              float* __restrict__ out, uint64_t n) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float r = radius[i];
    const float rr = r * r;
    out[i] = warp_weight(2.0f * rr, dist[i]);
}

__global__ void k_fill_f32(float* __restrict__ dst, float v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = v;
}

// generateTsdfVerticesKernel, src/TSDFWarp.cu:131-145.
__global__ void k_generate_vertices(float* __restrict__ verts, Geom g) {
    const int32_t x = blockIdx.x * 64 + threadIdx.x;
    const int32_t y = blockIdx.y * 4 + threadIdx.y;
    const int32_t z = blockIdx.z;
    if (x >= g.dim[0] || y >= g.dim[1]) return;
    const uint64_t i = ((uint64_t)z * g.dim[1] + y) * g.dim[0] + x;
    float p[3];
    voxel_to_world(g, x, y, z, p);
    verts[3 * i + 0] = p[0];
    verts[3 * i + 1] = p[1];
    verts[3 * i + 2] = p[2];
}

// computeTsdfWarpWeightsKernel, src/TSDFWarp.cu:88-116.
template <int K>
__global__ void k_weights(float* __restrict__ wout, const float* __restrict__ radii, const uint32_t* __restrict__ idx,
                          const float* __restrict__ dist, const int32_t* __restrict__ cnt, uint64_t n) {
    const uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (v >= n) return;
    int c = cnt[v];
    if (!c) return;
    if (c > K) c = K;
    float w[K];
    float sum = 0.0f;
#pragma unroll
    for (int s = 0; s < K; ++s) {
        w[s] = 0.0f;
        if (s < c) {
            const float r = radii[idx[(uint64_t)K * v + s]];
            const float r2 = r * r;
            w[s] = warp_weight(2.0f * r2, dist[(uint64_t)K * v + s]);
            sum = sum + w[s];
        }
    }
    if (sum != 0.0f) {
#pragma unroll
        for (int s = 0; s < K; ++s) w[s] = w[s] / sum;
    }
#pragma unroll
    for (int s = 0; s < K; ++s) wout[(uint64_t)K * v + s] = w[s];
}

// warpTsdfKernel, src/TSDFWarp.cu:158-190.
template <int K>
__global__ void k_warp(float* __restrict__ verts, const float* __restrict__ T, const uint32_t* __restrict__ idx,
                       const float* __restrict__ w, const int32_t* __restrict__ cnt, uint64_t n) {
    const uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (v >= n) return;
    int c = cnt[v];
    if (!c) return;
    if (c > K) c = K;
    float wv[K];
    bool zero = true;
#pragma unroll
    for (int s = 0; s < K; ++s) {
        wv[s] = w[(uint64_t)K * v + s];
        zero = zero && (__builtin_fabsf(wv[s]) <= 1e-5f);  // Eigen isZero(), dummy precision 1e-5
    }
    if (zero) return;
    float p[3] = {verts[3 * v + 0], verts[3 * v + 1], verts[3 * v + 2]};
    float acc[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < K; ++s) {
        if (s < c) {
            const float* M = T + 16ull * idx[(uint64_t)K * v + s];
            const float4 r0 = make_float4(M[0], M[4], M[8], M[12]);
            const float4 r1 = make_float4(M[1], M[5], M[9], M[13]);
            const float4 r2 = make_float4(M[2], M[6], M[10], M[14]);
            float t[3];
            affine_apply(r0, r1, r2, p, t);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float wt = wv[s] * t[a];
                acc[a] = acc[a] + wt;
            }
        }
    }
    verts[3 * v + 0] = acc[0];
    verts[3 * v + 1] = acc[1];
    verts[3 * v + 2] = acc[2];
}

struct Rot {
    float m[9];
};

// countCollisionsPerVoxelKernel, src/TSDFIntegration.cu:252-283.
__global__ void k_count_collisions(int32_t* __restrict__ col, const float* __restrict__ verts,
                                   const int32_t* __restrict__ cnt, Geom g, Rot rinv, float tx, float ty, float tz,
                                   uint64_t n) {
    const uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (v >= n) return;
    if (!cnt[v]) return;
    const float pw[3] = {verts[3 * v + 0], verts[3 * v + 1], verts[3 * v + 2]};
    const float t[3] = {tx, ty, tz};
    int32_t u[3];
    if (!warped_target(g, rinv.m, t, pw, u)) return;
    const uint64_t ui = ((uint64_t)u[2] * g.dim[1] + u[1]) * g.dim[0] + u[0];
    atomicAdd(col + ui, 1);
}

// integrateWarpedTSDFFromFrameKernel, src/TSDFIntegration.cu:173-226.
__global__ void k_integrate_warped(float* __restrict__ tsdf, float* __restrict__ wt, const float* __restrict__ verts,
                                   const int32_t* __restrict__ col, const int32_t* __restrict__ cnt,
                                   const float* __restrict__ depth, Camera cam, Geom g, Rot rinv, float tx, float ty,
                                   float tz, uint64_t n) {
    const uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (v >= n) return;
    if (!cnt[v]) return;
    const float pw[3] = {verts[3 * v + 0], verts[3 * v + 1], verts[3 * v + 2]};
    const float t[3] = {tx, ty, tz};
    int32_t u[3];
    if (!warped_target(g, rinv.m, t, pw, u)) return;
    const uint64_t ui = ((uint64_t)u[2] * g.dim[1] + u[1]) * g.dim[0] + u[0];
    if (col[ui] > 1) return;
    float tn;
    if (project_sample(cam, depth, pw, &tn) < 0) return;
    float to, wo;
    blend_sample(tsdf[v], wt[v], tn, cam.maxw, &to, &wo);
    tsdf[v] = to;
    wt[v] = wo;
}

// integrateTSDFFromFrameKernel, src/TSDFIntegration.cu:69-96, and with
// CAMERA the camera-frame variant, src/TSDFIntegration.cu:118-148.
template <bool CAMERA>
__global__ void k_integrate_plain(float* __restrict__ tsdf, float* __restrict__ wt, const float* __restrict__ depth,
                                  Camera cam, Geom g, Rot R, float tx, float ty, float tz, Rot Rw2c, float wx,
                                  float wy, float wz) {
    const int32_t x = blockIdx.x * 64 + threadIdx.x;
    const int32_t y = blockIdx.y * 4 + threadIdx.y;
    const int32_t z = blockIdx.z;
    if (x >= g.dim[0] || y >= g.dim[1]) return;
    const uint64_t i = ((uint64_t)z * g.dim[1] + y) * g.dim[0] + x;
    float p[3], q[3];
    voxel_to_world(g, x, y, z, p);
    mat3_vec(R.m, p, q);
    q[0] = q[0] + tx;
    q[1] = q[1] + ty;
    q[2] = q[2] + tz;
    if (CAMERA) {
        float r[3];
        mat3_vec(Rw2c.m, q, r);
        q[0] = r[0] + wx;
        q[1] = r[1] + wy;
        q[2] = r[2] + wz;
    }
    float tn;
    if (project_sample(cam, depth, q, &tn) < 0) return;
    float to, wo;
    blend_sample(tsdf[i], wt[i], tn, cam.maxw, &to, &wo);
    tsdf[i] = to;
    wt[i] = wo;
}

}  // namespace
}  // namespace vf

using namespace vf;

extern "C" {

int vf_translate(vf_float3* mesh, uint32_t mesh_size, vf_float3 t, vf_stream stream) {
    if (mesh_size == 0) return VF_OK;
    if (!mesh) return fail(VF_ERR_INVALID_ARGUMENT, "vf_translate: mesh is NULL");
    hipLaunchKernelGGL(k_translate, grid_1d(mesh_size), dim3(kBlock1D), 0, (hipStream_t)stream, (float*)mesh,
                       mesh_size, t.x, t.y, t.z);
    VF_LAUNCH_CHECK("k_translate");
    return VF_OK;
}

int vf_eval_warp_weight(const float* radius, const float* dist, float* out, size_t n, vf_stream stream) {
    if (!n) return VF_OK;
    if (!radius || !dist || !out) return fail(VF_ERR_INVALID_ARGUMENT, "vf_eval_warp_weight: NULL buffer");
    hipLaunchKernelGGL(k_eval_weight, grid_1d(n), dim3(kBlock1D), 0, (hipStream_t)stream, radius, dist, out,
                       (uint64_t)n);
    VF_LAUNCH_CHECK("k_eval_weight");
    return VF_OK;
}

int vf_fill_f32(float* dst, float value, uint64_t count, vf_stream stream) {
    if (!count) return VF_OK;
    uint64_t blocks = (count + kBlock1D - 1) / kBlock1D;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)blocks), dim3(kBlock1D), 0, (hipStream_t)stream, dst, value,
                       count);
    VF_LAUNCH_CHECK("k_fill_f32");
    return VF_OK;
}

int vf_generate_tsdf_vertices(vf_float3* verts, vf_float3 vs, vf_float3 size, vf_int3 dim, vf_stream stream) {
    uint64_t n;
    int st = check_volume(dim, vs, &n);
    if (st) return st;
    if (!verts) return fail(VF_ERR_INVALID_ARGUMENT, "vf_generate_tsdf_vertices: vertices is NULL");
    const Geom g = make_geom(dim, vs, size);
    dim3 grid, block;
    grid_3d(g, &grid, &block);
    hipLaunchKernelGGL(k_generate_vertices, grid, block, 0, (hipStream_t)stream, (float*)verts, g);
    VF_LAUNCH_CHECK("k_generate_vertices");
    return VF_OK;
}

#define VF_K_SWITCH(k, MACRO) \
    switch (k) {              \
        case 1: MACRO(1); break; \
        case 2: MACRO(2); break; \
        case 3: MACRO(3); break; \
        case 4: MACRO(4); break; \
        case 5: MACRO(5); break; \
        case 6: MACRO(6); break; \
        case 7: MACRO(7); break; \
        case 8: MACRO(8); break; \
    }

int vf_compute_tsdf_warp_weights(int k, float* wout, const float* radii, const uint32_t* idx, const float* dist,
                                 const int32_t* cnt, size_t n, vf_stream stream) {
    int st = check_k(k);
    if (st) return st;
    if (!n) return VF_OK;
    if (!wout || !radii || !idx || !dist || !cnt)
        return fail(VF_ERR_INVALID_ARGUMENT, "vf_compute_tsdf_warp_weights: NULL buffer");
#define VF_LAUNCH_W(KK)                                                                                     \
    hipLaunchKernelGGL(k_weights<KK>, grid_1d(n), dim3(kBlock1D), 0, (hipStream_t)stream, wout, radii, idx, \
                       dist, cnt, (uint64_t)n)
    VF_K_SWITCH(k, VF_LAUNCH_W)
#undef VF_LAUNCH_W
    VF_LAUNCH_CHECK("k_weights");
    return VF_OK;
}

int vf_warp_tsdf(int k, vf_float3* verts, const vf_mat4* T, const uint32_t* idx, const float* w, const int32_t* cnt,
                 vf_float3 vs, vf_float3 size, vf_int3 dim, vf_stream stream) {
    (void)size;
    int st = check_k(k);
    if (st) return st;
    uint64_t n;
    st = check_volume(dim, vs, &n);
    if (st) return st;
    if (!verts || !T || !idx || !w || !cnt) return fail(VF_ERR_INVALID_ARGUMENT, "vf_warp_tsdf: NULL buffer");
#define VF_LAUNCH_WARP(KK)                                                                                        \
    hipLaunchKernelGGL(k_warp<KK>, grid_1d(n), dim3(kBlock1D), 0, (hipStream_t)stream, (float*)verts,           \
                       (const float*)T, idx, w, cnt, n)
    VF_K_SWITCH(k, VF_LAUNCH_WARP)
#undef VF_LAUNCH_WARP
    VF_LAUNCH_CHECK("k_warp");
    return VF_OK;
}

int vf_count_collisions_per_voxel(int32_t* col, const vf_float3* verts, const int32_t* cnt, vf_float3 vs,
                                  vf_float3 size, vf_int3 dim, vf_mat3 R, vf_float3 t, vf_stream stream) {
    uint64_t n;
    int st = check_volume(dim, vs, &n);
    if (st) return st;
    if (!col || !verts || !cnt) return fail(VF_ERR_INVALID_ARGUMENT, "vf_count_collisions_per_voxel: NULL buffer");
    const Geom g = make_geom(dim, vs, size);
    Rot rinv;
    mat3_inverse_eigen(R.m, rinv.m);
    hipLaunchKernelGGL(k_count_collisions, grid_1d(n), dim3(kBlock1D), 0, (hipStream_t)stream, col,
                       (const float*)verts, cnt, g, rinv, t.x, t.y, t.z, n);
    VF_LAUNCH_CHECK("k_count_collisions");
    return VF_OK;
}

int vf_integrate_warped_tsdf(float* tsdf, float* wt, const vf_float3* verts, const int32_t* col, const int32_t* cnt,
                             const float* depth, vf_int2 fd, float delta, vf_float3 vs, vf_float3 size, vf_int3 dim,
                             vf_mat3 R, vf_float3 t, vf_mat3 Kin, float maxw, vf_stream stream) {
    uint64_t n;
    int st = check_volume(dim, vs, &n);
    if (st) return st;
    if (!tsdf || !wt || !verts || !col || !cnt || !depth)
        return fail(VF_ERR_INVALID_ARGUMENT, "vf_integrate_warped_tsdf: NULL buffer");
    if (fd.x < 0 || fd.y < 0) return fail(VF_ERR_INVALID_ARGUMENT, "frame dims must be non-negative");
    const Geom g = make_geom(dim, vs, size);
    Rot rinv;
    mat3_inverse_eigen(R.m, rinv.m);
    const Camera cam = make_camera(Kin, fd.x, fd.y, delta, maxw);
    hipLaunchKernelGGL(k_integrate_warped, grid_1d(n), dim3(kBlock1D), 0, (hipStream_t)stream, tsdf, wt,
                       (const float*)verts, col, cnt, depth, cam, g, rinv, t.x, t.y, t.z, n);
    VF_LAUNCH_CHECK("k_integrate_warped");
    return VF_OK;
}

static int plain_common(bool camera, float* tsdf, float* wt, const float* depth, vf_int2 fd, float delta,
                        vf_int3 dim, vf_float3 vs, vf_float3 size, const vf_mat3* Rw2c, vf_float3 tw2c,
                        vf_mat3 R, vf_float3 t, vf_mat3 Kin, float maxw, vf_stream stream) {
    uint64_t n;
    int st = check_volume(dim, vs, &n);
    if (st) return st;
    if (!tsdf || !wt || !depth) return fail(VF_ERR_INVALID_ARGUMENT, "plain integrate: NULL buffer");
    if (fd.x < 0 || fd.y < 0) return fail(VF_ERR_INVALID_ARGUMENT, "frame dims must be non-negative");
    const Geom g = make_geom(dim, vs, size);
    const Camera cam = make_camera(Kin, fd.x, fd.y, delta, maxw);
    Rot r, w2c;
    for (int i = 0; i < 9; ++i) {
        r.m[i] = R.m[i];
        w2c.m[i] = Rw2c ? Rw2c->m[i] : 0.0f;
    }
    dim3 grid, block;
    grid_3d(g, &grid, &block);
    if (camera)
        hipLaunchKernelGGL(k_integrate_plain<true>, grid, block, 0, (hipStream_t)stream, tsdf, wt, depth, cam, g, r,
                           t.x, t.y, t.z, w2c, tw2c.x, tw2c.y, tw2c.z);
    else
        hipLaunchKernelGGL(k_integrate_plain<false>, grid, block, 0, (hipStream_t)stream, tsdf, wt, depth, cam, g, r,
                           t.x, t.y, t.z, w2c, tw2c.x, tw2c.y, tw2c.z);
    VF_LAUNCH_CHECK("k_integrate_plain");
    return VF_OK;
}

int vf_integrate_tsdf(float* tsdf, float* wt, const float* depth, vf_int2 fd, float delta, vf_int3 dim,
                      vf_float3 vs, vf_float3 size, vf_mat3 R, vf_float3 t, vf_mat3 Kin, float maxw,
                      vf_stream stream) {
    const vf_float3 zero = {0.0f, 0.0f, 0.0f};
    return plain_common(false, tsdf, wt, depth, fd, delta, dim, vs, size, nullptr, zero, R, t, Kin, maxw, stream);
}

int vf_integrate_tsdf_camera_frame(float* tsdf, float* wt, const float* depth, vf_int2 fd, float delta, vf_int3 dim,
                                   vf_float3 vs, vf_float3 size, vf_mat3 Rw2c, vf_float3 tw2c, vf_mat3 R,
                                   vf_float3 t, vf_mat3 Kin, float maxw, vf_stream stream) {
    return plain_common(true, tsdf, wt, depth, fd, delta, dim, vs, size, &Rw2c, tw2c, R, t, Kin, maxw, stream);
}

}  // extern "C"
