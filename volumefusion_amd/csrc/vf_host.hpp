// vf_host.hpp -- host-side helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>

#include "vf_device.hpp"
#include "volumefusion.h"

namespace vf {

// Records a formatted message for vf_last_error() and returns `status`.
int fail(int status, const char* fmt, ...);

#define VF_HIP(call)                                                                                       \
    do {                                                                                                   \
        hipError_t vf_e_ = (call);                                                                         \
        if (vf_e_ != hipSuccess)                                                                           \
            return ::vf::fail(VF_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(vf_e_), __FILE__, \
                              __LINE__);                                                                   \
    } while (0)

#define VF_LAUNCH_CHECK(what)                                                                               \
    do {                                                                                                    \
        hipError_t vf_e_ = hipGetLastError();                                                               \
        if (vf_e_ != hipSuccess)                                                                            \
            return ::vf::fail(VF_ERR_HIP, "launch of %s failed: %s", what, hipGetErrorString(vf_e_));        \
    } while (0)

// Geometry of a volume as the reference derives it (src/TSDFWarp.cu:8-15).
inline Geom make_geom(vf_int3 dim, vf_float3 vs, vf_float3 size) {
    Geom g;
    g.dim[0] = dim.x;
    g.dim[1] = dim.y;
    g.dim[2] = dim.z;
    g.vs[0] = vs.x;
    g.vs[1] = vs.y;
    g.vs[2] = vs.z;
    g.inv[0] = (float)(1.0 / (double)vs.x);
    g.inv[1] = (float)(1.0 / (double)vs.y);
    g.inv[2] = (float)(1.0 / (double)vs.z);
    g.halfs[0] = size.x * 0.5f;
    g.halfs[1] = size.y * 0.5f;
    g.halfs[2] = size.z * 0.5f;
    return g;
}

// Eigen's 3x3 inverse by cofactors (Eigen/src/LU/InverseImpl.h:126-170), the
// host call the reference makes before its collision and integration kernels
// (src/TSDFIntegration.cu:247, 296).  Column-major in and out.
void mat3_inverse_eigen(const float* m, float* out);

// Checks shared by every voxel-sweep entry point.  Returns VF_OK or an error.
int check_volume(vf_int3 dim, vf_float3 vs, uint64_t* n_out);

inline int check_k(int k) {
    if (k < 1 || k > 8) return fail(VF_ERR_UNSUPPORTED_K, "k=%d: supported knn slot counts are 1..8", k);
    return VF_OK;
}

// Device scratch owned by a workspace (defined with the workspace).
int workspace_scratch(vf_workspace* ws, uint64_t bytes, void** ptr, hipStream_t stream);

// What one volume adds, in bytes, to each buffer group of a workspace with a
// record budget of records_per_voxel when a batch holds it: node records,
// per-brick state, collision tiles, pending records, node-list entries,
// descriptors (with the workspace's fixed part) and depth-tile maxima.  A
// batch needs at most the sum over its volumes in each group, and such a
// workspace grows to exactly what its batches need, so no group of it
// exceeds the largest such sum (defined with the workspace).
constexpr int kWsGroups = 7;
void volume_ws_bytes(vf_int3 dim, uint64_t nodes, vf_int3 cube, int width, int height, int brick_shape,
                     float records_per_voxel, uint64_t out[kWsGroups]);

inline Camera make_camera(const vf_mat3& K, int W, int H, float delta, float maxw) {
    Camera c;
    c.fx = K.m[0];
    c.fy = K.m[4];
    c.cx = K.m[6];
    c.cy = K.m[7];
    c.W = W;
    c.H = H;
    c.delta = delta;
    c.maxw = maxw;
    c.rdelta = 1.0 / (double)delta;
    return c;
}

}  // namespace vf
