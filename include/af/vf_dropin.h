// af/vf_dropin.h -- the reference's af:: TSDF-integration entry points on the
// MI355X kernels of volumefusion.h.
//
// Replaces include/af/TSDFWarp.h and include/af/TSDFIntegration.cuh, and
// af::runTranslateKernel, af::runComputeWarpWeightsKernel
// (include/af/VertexManipulation.cuh:16, 19-23) and af::runWarpKernel
// (include/af/ComputeCorrespondence.cuh:11-16), the depth-ingest kernels
// (include/af/VertexManipulation.cuh:11-32, include/af/BilateralFilter.cuh)
// and the marching-cubes kernels (include/af/MarchingCubes.cuh:10-28), for
// callers such as af::runVolumeFusion (src/VolumeFusion.cu:334, 708-741) and
// af::DeviceTsdf::integrate (include/af/DeviceTsdf.cuh:28-32).  Same names,
// argument order and meaning, same Eigen types (spelled out, so this header
// coexists with the reference's eigen_extension.h aliases Vecui<K>/Vecf<K>),
// same behaviour: each call returns after the device work is done (the
// reference ends every wrapper with cudaDeviceSynchronize, e.g.
// src/TSDFWarp.cu:85) and a failure throws std::runtime_error (the
// reference's CUDA_CHECK, src/Helper.cu:7-20).
//
// K is deduced from the knn buffer type (the reference fixes it to
// Constants::motionGraphKNN = 6, include/af/Constants.h:13); 1 <= K <= 8.
#ifndef AF_VF_DROPIN_H
#define AF_VF_DROPIN_H

#include <Eigen/Core>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <utility>
#include <stdexcept>
#include <string>

#include "volumefusion.h"

namespace af {
namespace vf_detail {

static_assert(sizeof(Eigen::Vector3f) == sizeof(vf_float3), "Vec3f must be 3 packed floats");
static_assert(sizeof(Eigen::Vector3i) == sizeof(vf_int3), "Vec3i must be 3 packed ints");
static_assert(sizeof(Eigen::Vector2i) == sizeof(vf_int2), "Vec2i must be 2 packed ints");
static_assert(sizeof(Eigen::Matrix3f) == sizeof(vf_mat3), "Mat3f must be 9 floats");
static_assert(sizeof(Eigen::Matrix4f) == sizeof(vf_mat4), "Mat4f must be 16 floats");

inline void check(int status, const char* what) {
    if (status != VF_OK) throw std::runtime_error(std::string(what) + ": " + vf_last_error());
}

inline void done(int status, const char* what) {
    check(status, what);
    check(vf_device_synchronize(), what);
}

inline vf_float3 f3(const Eigen::Vector3f& v) { return vf_float3{v[0], v[1], v[2]}; }
inline vf_int3 i3(const Eigen::Vector3i& v) { return vf_int3{v[0], v[1], v[2]}; }
inline vf_int2 i2(const Eigen::Vector2i& v) { return vf_int2{v[0], v[1]}; }
inline vf_mat3 m3(const Eigen::Matrix3f& m) {
    vf_mat3 r;
    for (int i = 0; i < 9; ++i) r.m[i] = m.data()[i];  // column-major, as Eigen stores it
    return r;
}

template <int K>
void check_k_layout() {
    static_assert(K >= 1 && K <= 8, "knn slots per voxel must be 1..8");
    static_assert(sizeof(Eigen::Matrix<unsigned int, K, 1>) == 4 * K, "Vecui<K> must be K packed uint32");
    static_assert(sizeof(Eigen::Matrix<float, K, 1>) == 4 * K, "Vecf<K> must be K packed floats");
}

}  // namespace vf_detail

// src/TSDFWarp.cu:70-86
template <int K>
void runTsdfKNNGraph(Eigen::Matrix<unsigned int, K, 1>* knnIdxs, Eigen::Matrix<float, K, 1>* knnDists, int* knnCounts,
                     const Eigen::Vector3f* cGraph, const float* cGraphRadiuses, const std::size_t cGraphSize,
                     const Eigen::Vector3f cVoxelSize, const Eigen::Vector3f cVolumeSize,
                     const Eigen::Vector3i cVolumeDim, const Eigen::Vector3i cVoxelsPerNode) {
    vf_detail::check_k_layout<K>();
    using namespace vf_detail;
    done(vf_tsdf_knn_graph(K, reinterpret_cast<uint32_t*>(knnIdxs), reinterpret_cast<float*>(knnDists), knnCounts,
                           reinterpret_cast<const vf_float3*>(cGraph), cGraphRadiuses, cGraphSize, f3(cVoxelSize),
                           f3(cVolumeSize), i3(cVolumeDim), i3(cVoxelsPerNode), nullptr),
         "runTsdfKNNGraph");
}

// src/TSDFWarp.cu:118-129
template <int K>
void runComputeTsdfWarpWeightsKernel(Eigen::Matrix<float, K, 1>* cKnnWeights, const float* cMotionGraphRadiuses,
                                     const Eigen::Matrix<unsigned int, K, 1>* cKnnIdxs,
                                     const Eigen::Matrix<float, K, 1>* cKnnDists, const int* cKnnCounts,
                                     const std::size_t cGridSize) {
    vf_detail::check_k_layout<K>();
    vf_detail::done(vf_compute_tsdf_warp_weights(K, reinterpret_cast<float*>(cKnnWeights), cMotionGraphRadiuses,
                                                 reinterpret_cast<const uint32_t*>(cKnnIdxs),
                                                 reinterpret_cast<const float*>(cKnnDists), cKnnCounts, cGridSize,
                                                 nullptr),
                    "runComputeTsdfWarpWeightsKernel");
}

// src/TSDFWarp.cu:147-156
inline void runGenerateTsdfVerticesKernel(Eigen::Vector3f* tsdfVertices, const Eigen::Vector3f cVoxelSize,
                                          const Eigen::Vector3f cVolumeSize, const Eigen::Vector3i cVolumeDim) {
    using namespace vf_detail;
    done(vf_generate_tsdf_vertices(reinterpret_cast<vf_float3*>(tsdfVertices), f3(cVoxelSize), f3(cVolumeSize),
                                   i3(cVolumeDim), nullptr),
         "runGenerateTsdfVerticesKernel");
}

// src/TSDFWarp.cu:192-206
template <int K>
void runWarpTsdfKernel(Eigen::Vector3f* tsdfVertices, const Eigen::Matrix4f* cMotionGraph,
                       const Eigen::Matrix<unsigned int, K, 1>* cKnnIdxs,
                       const Eigen::Matrix<float, K, 1>* cKnnWeights, const int* cKnnCounts,
                       const Eigen::Vector3f cVoxelSize, const Eigen::Vector3f cVolumeSize,
                       const Eigen::Vector3i cVolumeDim) {
    vf_detail::check_k_layout<K>();
    using namespace vf_detail;
    done(vf_warp_tsdf(K, reinterpret_cast<vf_float3*>(tsdfVertices), reinterpret_cast<const vf_mat4*>(cMotionGraph),
                      reinterpret_cast<const uint32_t*>(cKnnIdxs), reinterpret_cast<const float*>(cKnnWeights),
                      cKnnCounts, f3(cVoxelSize), f3(cVolumeSize), i3(cVolumeDim), nullptr),
         "runWarpTsdfKernel");
}

// src/TSDFIntegration.cu:98-116
inline void runIntegrateTSDFFromFrameKernel(float* tsdf, float* tsdfWeights, const float* cDepthFrame,
                                            const Eigen::Vector2i cFrameDims, const float cDelta,
                                            const Eigen::Vector3i cDims, const Eigen::Vector3f cVoxelSize,
                                            const Eigen::Vector3f cVolumeSize, const Eigen::Matrix3f cVolumeRotate,
                                            const Eigen::Vector3f cVolumeTranslate,
                                            const Eigen::Matrix3f cDepthIntrinsics, const float cMaxWeight = 15.0f) {
    using namespace vf_detail;
    done(vf_integrate_tsdf(tsdf, tsdfWeights, cDepthFrame, i2(cFrameDims), cDelta, i3(cDims), f3(cVoxelSize),
                           f3(cVolumeSize), m3(cVolumeRotate), f3(cVolumeTranslate), m3(cDepthIntrinsics), cMaxWeight,
                           nullptr),
         "runIntegrateTSDFFromFrameKernel");
}

// src/TSDFIntegration.cu:150-171
inline void runIntegrateTSDFFromCameraFrameKernel(float* tsdf, float* tsdfWeights, const float* cDepthFrame,
                                                  const Eigen::Vector2i cFrameDims, const float cDelta,
                                                  const Eigen::Vector3i cDims, const Eigen::Vector3f cVoxelSize,
                                                  const Eigen::Vector3f cVolumeSize,
                                                  const Eigen::Matrix3f cWorld2CameraRotate,
                                                  const Eigen::Vector3f cWorld2CameraTranslate,
                                                  const Eigen::Matrix3f cVolumeRotate,
                                                  const Eigen::Vector3f cVolumeTranslate,
                                                  const Eigen::Matrix3f cDepthIntrinsics,
                                                  const float cMaxWeight = 15.0f) {
    using namespace vf_detail;
    done(vf_integrate_tsdf_camera_frame(tsdf, tsdfWeights, cDepthFrame, i2(cFrameDims), cDelta, i3(cDims),
                                        f3(cVoxelSize), f3(cVolumeSize), m3(cWorld2CameraRotate),
                                        f3(cWorld2CameraTranslate), m3(cVolumeRotate), f3(cVolumeTranslate),
                                        m3(cDepthIntrinsics), cMaxWeight, nullptr),
         "runIntegrateTSDFFromCameraFrameKernel");
}

// src/TSDFIntegration.cu:228-250
inline void runIntegrateWarpedTSDFFromFrameKernel(float* tsdf, float* tsdfWeights, const Eigen::Vector3f* cTsdfVertices,
                                                  const int* collisionCounts, const int* cKnnCounts,
                                                  const float* cDepthFrame, const Eigen::Vector2i cFrameDims,
                                                  const float cDelta, const Eigen::Vector3f cVoxelSize,
                                                  const Eigen::Vector3f cVolumeSize, const Eigen::Vector3i cVolumeDim,
                                                  const Eigen::Matrix3f cVolumeRotate,
                                                  const Eigen::Vector3f cVolumeTranslate,
                                                  const Eigen::Matrix3f cDepthIntrinsics,
                                                  const float cMaxWeight = 15.0f) {
    using namespace vf_detail;
    done(vf_integrate_warped_tsdf(tsdf, tsdfWeights, reinterpret_cast<const vf_float3*>(cTsdfVertices),
                                  collisionCounts, cKnnCounts, cDepthFrame, i2(cFrameDims), cDelta, f3(cVoxelSize),
                                  f3(cVolumeSize), i3(cVolumeDim), m3(cVolumeRotate), f3(cVolumeTranslate),
                                  m3(cDepthIntrinsics), cMaxWeight, nullptr),
         "runIntegrateWarpedTSDFFromFrameKernel");
}

// src/TSDFIntegration.cu:285-299
inline void runCountCollisionsPerVoxelKernel(int* collisionCounts, const Eigen::Vector3f* tsdfVertices,
                                             const int* cKnnCounts, const Eigen::Vector3f cVoxelSize,
                                             const Eigen::Vector3f cVolumeSize, const Eigen::Vector3i cVolumeDim,
                                             const Eigen::Matrix3f cVolumeRotate,
                                             const Eigen::Vector3f cVolumeTranslate) {
    using namespace vf_detail;
    done(vf_count_collisions_per_voxel(collisionCounts, reinterpret_cast<const vf_float3*>(tsdfVertices), cKnnCounts,
                                       f3(cVoxelSize), f3(cVolumeSize), i3(cVolumeDim), m3(cVolumeRotate),
                                       f3(cVolumeTranslate), nullptr),
         "runCountCollisionsPerVoxelKernel");
}

// src/VertexManipulation.cu:138-145 (the mesh weights have the d >= r cutoff)
template <int K>
void runComputeWarpWeightsKernel(Eigen::Matrix<float, K, 1>* knnWeights, const float* cMotionGraphRadiuses,
                                 const Eigen::Matrix<unsigned int, K, 1>* cKnnIdxs,
                                 const Eigen::Matrix<float, K, 1>* cKnnDists, const std::size_t cMeshSize) {
    vf_detail::check_k_layout<K>();
    vf_detail::done(vf_compute_warp_weights(K, reinterpret_cast<float*>(knnWeights), cMotionGraphRadiuses,
                                            reinterpret_cast<const uint32_t*>(cKnnIdxs),
                                            reinterpret_cast<const float*>(cKnnDists), cMeshSize, 0, nullptr),
                    "runComputeWarpWeightsKernel");
}

// src/ComputeCorrespondence.cu:32-43
template <int K>
void runWarpKernel(Eigen::Vector3f* warpedMesh, const Eigen::Vector3f* cMesh, const Eigen::Matrix4f* cMotionGraph,
                   const Eigen::Matrix<unsigned int, K, 1>* cKnnIdxs, const Eigen::Matrix<float, K, 1>* cKnnWeights,
                   const std::size_t cMeshSize) {
    vf_detail::check_k_layout<K>();
    vf_detail::done(vf_warp_mesh(K, reinterpret_cast<vf_float3*>(warpedMesh), reinterpret_cast<const vf_float3*>(cMesh),
                                 reinterpret_cast<const vf_mat4*>(cMotionGraph),
                                 reinterpret_cast<const uint32_t*>(cKnnIdxs), reinterpret_cast<const float*>(cKnnWeights),
                                 cMeshSize, 0, nullptr),
                    "runWarpKernel");
}

// src/MarchingCubes.cu:510-522 (canonical vertex order; see volumefusion.h)
inline void runMarchingCubesMeshKernel(Eigen::Vector3f* mesh, unsigned int* meshSize, const float* cTsdf,
                                       const float* cTsdfWeights, const Eigen::Vector3i cDims,
                                       const Eigen::Vector3f& cVoxelSize, const Eigen::Vector3f& cVolSize,
                                       const float cIsoValue) {
    using namespace vf_detail;
    done(vf_marching_cubes_mesh(nullptr, reinterpret_cast<vf_float3*>(mesh), meshSize, UINT64_MAX, cTsdf, cTsdfWeights,
                                i3(cDims), f3(cVoxelSize), f3(cVolSize), cIsoValue, nullptr),
         "runMarchingCubesMeshKernel");
}

// src/MarchingCubes.cu:684-702
inline void runMarchingCubesFullKernel(Eigen::Vector3f* mesh, unsigned int* meshSize, Eigen::Vector3i* faces,
                                       unsigned int* facesSize, const float* cTsdf, const float* cTsdfWeights,
                                       const Eigen::Vector3i cDims, const Eigen::Vector3f& cVoxelSize,
                                       const Eigen::Vector3f& cVolSize, const float cIsoValue) {
    using namespace vf_detail;
    done(vf_marching_cubes_full(nullptr, reinterpret_cast<vf_float3*>(mesh), meshSize, UINT64_MAX,
                                reinterpret_cast<vf_int3*>(faces), facesSize, UINT64_MAX, cTsdf, cTsdfWeights,
                                i3(cDims), f3(cVoxelSize), f3(cVolSize), cIsoValue, nullptr),
         "runMarchingCubesFullKernel");
}

// src/VertexManipulation.cu:17-23
inline void runFilterDepthKernel(float* depth, unsigned int depthSize, float threshold,
                                 float fillValue = std::numeric_limits<float>::infinity()) {
    vf_detail::done(vf_filter_depth(depth, depthSize, threshold, fillValue, nullptr), "runFilterDepthKernel");
}

// src/VertexManipulation.cu:34-40
inline void runMaskDepthKernel(float* depth, const bool* cDepthMask, unsigned int depthSize) {
    static_assert(sizeof(bool) == 1, "bool masks are one byte");
    vf_detail::done(vf_mask_depth(depth, reinterpret_cast<const uint8_t*>(cDepthMask), depthSize, nullptr),
                    "runMaskDepthKernel");
}

// src/VertexManipulation.cu:65-71
inline void runDepthToMeshKernel(Eigen::Vector3f* mesh, const float* cDepth, unsigned int depthWidth,
                                 unsigned int depthHeight, const Eigen::Matrix3f& K) {
    using namespace vf_detail;
    done(vf_depth_to_mesh(reinterpret_cast<vf_float3*>(mesh), cDepth, depthWidth, depthHeight, m3(K), nullptr),
         "runDepthToMeshKernel");
}

// src/VertexManipulation.cu:205-217
inline void runDepthVerticesNormalsKernel(Eigen::Vector3f* normals, Eigen::Vector3f* vertices, const float* depth,
                                          int h, int w, std::pair<float, float> normExclAngleRange) {
    vf_detail::done(vf_depth_normals(reinterpret_cast<vf_float3*>(normals), reinterpret_cast<vf_float3*>(vertices),
                                     depth, h, w, normExclAngleRange.first, normExclAngleRange.second, nullptr),
                    "runDepthVerticesNormalsKernel");
}

// src/VertexManipulation.cu:234-241
inline void runRemoveDepthWithoutNormalsKernel(Eigen::Vector3f* normals, Eigen::Vector3f* vertices, float* depth,
                                               int h, int w) {
    vf_detail::done(vf_remove_depth_without_normals(reinterpret_cast<vf_float3*>(normals),
                                                    reinterpret_cast<vf_float3*>(vertices), depth, h, w, nullptr),
                    "runRemoveDepthWithoutNormalsKernel");
}

// src/BilateralFilter.cu:185-193
inline void bilateralFilterTextureOpmShared(float* output, const float* input, int height, int width, int r,
                                            double sI, double sS, float thr) {
    vf_detail::done(vf_bilateral_filter(output, input, height, width, r, sI, sS, thr, nullptr),
                    "bilateralFilterTextureOpmShared");
}

// src/VertexManipulation.cu:81-87
inline void runTranslateKernel(Eigen::Vector3f* mesh, unsigned int meshSize, Eigen::Vector3f translation) {
    using namespace vf_detail;
    done(vf_translate(reinterpret_cast<vf_float3*>(mesh), meshSize, f3(translation), nullptr), "runTranslateKernel");
}

// The whole TSDF-integration block of the frame loop, src/VolumeFusion.cu:702-742
// (volume rotation I, translation = frameCentroid), in five kernels and
// without the per-voxel knn/vertex/collision buffers.  `motionGraph` node
// positions receive the reference's round trip fl(fl(p - c) + c).
template <int K>
void runWarpedTsdfIntegration(vf_workspace* workspace, float* tsdf, float* tsdfWeights, const Eigen::Vector3i volDim,
                              const Eigen::Vector3f volSize, const Eigen::Vector3f frameCentroid,
                              Eigen::Vector3f* motionGraph, const Eigen::Matrix4f* motionGraphTransforms,
                              const float* motionGraphRadiuses, std::size_t graphSize, const float* depthFrame,
                              const Eigen::Vector2i frameDims, const Eigen::Matrix3f depthIntrinsics, float delta,
                              float motionGraphRadius, float maxWeight = 15.0f, int blend = VF_BLEND_LBS) {
    using namespace vf_detail;
    const Eigen::Vector3f voxelSize = volSize.cwiseQuotient(volDim.cast<float>());
    // src/VolumeFusion.cu:702-705
    Eigen::Vector3i perNode;
    for (int a = 0; a < 3; ++a) perNode[a] = 2 * (static_cast<int>(motionGraphRadius / voxelSize[a]) + 1) + 1;
    vf_volume vol{tsdf, tsdfWeights, i3(volDim), f3(voxelSize), f3(volSize), f3(frameCentroid)};
    vf_graph graph{reinterpret_cast<const vf_float3*>(motionGraph), reinterpret_cast<vf_float3*>(motionGraph),
                   reinterpret_cast<const vf_mat4*>(motionGraphTransforms), motionGraphRadiuses, graphSize};
    vf_frame frame{depthFrame, frameDims[0], frameDims[1], m3(depthIntrinsics)};
    vf_warp_params params{K, delta, maxWeight, i3(perNode), blend};
    done(vf_warped_integrate(workspace, &vol, &graph, &frame, &params, nullptr), "runWarpedTsdfIntegration");
}

}  // namespace af

#endif
