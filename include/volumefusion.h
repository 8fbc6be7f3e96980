/*
 * volumefusion.h -- C ABI of the MI355X (gfx950) warped TSDF integration path.
 *
 * Drop-in for VolumeFusion's per-frame TSDF integration: the af:: entry points
 * of include/af/TSDFWarp.h, include/af/TSDFIntegration.cuh and
 * af::runTranslateKernel (include/af/VertexManipulation.cuh:16), on the same
 * data layouts:
 *   - TSDF: two float[N] arrays (tsdf, weight), z-major,
 *     idx = z*X*Y + y*X + x (src/TSDFIntegration.cu:88);
 *   - motion graph: Vec3f[G] positions (12 B packed), Mat4f[G] transforms
 *     (64 B, column-major), float[G] radii (include/af/DeviceMotionGraph.cuh);
 *   - knn buffers: Vecui<K>[N], Vecf<K>[N] (K contiguous values per voxel),
 *     int[N] counts (include/af/DataTypes.cuh);
 *   - depth: float[W*H] row-major in metres; intrinsics Mat3f column-major.
 *
 * Conventions: the caller owns every input/output buffer (device pointers);
 * calls are asynchronous on the given HIP stream (NULL = default stream) and
 * return a status code (VF_OK = 0).  The af:: C++ wrapper in
 * include/af/vf_dropin.h restores the reference's blocking, exception-throwing
 * behaviour.  K (knn slots per voxel) is a run-time argument, 1 <= K <= 8.
 *
 * Canonical knn: where the reference claims knn slots with atomicAdd
 * (src/TSDFWarp.cu:61-67, a race), these kernels keep the K candidates with the
 * smallest (distance, node index), in that order.
 */
#ifndef VOLUMEFUSION_AMD_H
#define VOLUMEFUSION_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vf_float3 { float x, y, z; } vf_float3; /* Eigen::Vector3f, 12 B */
typedef struct vf_int3 { int32_t x, y, z; } vf_int3;   /* Eigen::Vector3i       */
typedef struct vf_int2 { int32_t x, y; } vf_int2;      /* Eigen::Vector2i       */
typedef struct vf_mat3 { float m[9]; } vf_mat3;        /* Eigen::Matrix3f, column-major */
typedef struct vf_mat4 { float m[16]; } vf_mat4;       /* Eigen::Matrix4f, column-major */
typedef void* vf_stream;                               /* hipStream_t */
typedef struct vf_workspace vf_workspace;              /* device scratch, see below */

enum {
    VF_OK = 0,
    VF_ERR_INVALID_ARGUMENT = 1,
    VF_ERR_UNSUPPORTED_K = 2,
    VF_ERR_HIP = 3,
    VF_ERR_OUT_OF_MEMORY = 4,
    VF_ERR_WORKSPACE = 5,
    VF_ERR_CAPACITY = 6,
    VF_ERR_SINGULAR = 7
};

/* Text of a status code, and the message of the last failure on this thread. */
const char* vf_status_string(int status);
const char* vf_last_error(void);
/* Library version string and the device architecture it was built for. */
const char* vf_version(void);

/* ------------------------------------------------------------------------ */
/* Reference entry points                                                   */
/* ------------------------------------------------------------------------ */

/* af::runTsdfKNNGraph, include/af/TSDFWarp.h:9-18 (src/TSDFWarp.cu:23-86).
 * For every node (positions already translated by the caller) and every voxel
 * of its cube (voxels_per_node, centred on worldToVoxel(node)) with
 * |voxelToWorld(v) - node| <= 2 r: candidate.  Writes, for each voxel with at
 * least one candidate, count = min(#candidates, K) and the K best candidates by
 * (distance, node index).  Counts must be zero on entry, as the reference's
 * caller makes them (src/VolumeFusion.cu:358); voxels without candidates are
 * not written. */
int vf_tsdf_knn_graph(int k, uint32_t* knn_idxs, float* knn_dists, int32_t* knn_counts, const vf_float3* graph,
                      const float* graph_radii, size_t graph_size, vf_float3 voxel_size, vf_float3 volume_size,
                      vf_int3 volume_dim, vf_int3 voxels_per_node, vf_stream stream);

/* af::runComputeTsdfWarpWeightsKernel, include/af/TSDFWarp.h:20-25
 * (src/TSDFWarp.cu:88-129): w_j = exp(-d_j^2 / (2 r^2)), normalised to sum 1. */
int vf_compute_tsdf_warp_weights(int k, float* knn_weights, const float* graph_radii, const uint32_t* knn_idxs,
                                 const float* knn_dists, const int32_t* knn_counts, size_t grid_size,
                                 vf_stream stream);

/* af::runGenerateTsdfVerticesKernel, include/af/TSDFWarp.h:27 (src/TSDFWarp.cu:131-156). */
int vf_generate_tsdf_vertices(vf_float3* tsdf_vertices, vf_float3 voxel_size, vf_float3 volume_size,
                              vf_int3 volume_dim, vf_stream stream);

/* af::runWarpTsdfKernel, include/af/TSDFWarp.h:29-36 (src/TSDFWarp.cu:158-206):
 * p' = sum_j w_j (T_j [p; 1]).xyz in place. */
int vf_warp_tsdf(int k, vf_float3* tsdf_vertices, const vf_mat4* motion_graph, const uint32_t* knn_idxs,
                 const float* knn_weights, const int32_t* knn_counts, vf_float3 voxel_size, vf_float3 volume_size,
                 vf_int3 volume_dim, vf_stream stream);

/* af::runCountCollisionsPerVoxelKernel, include/af/TSDFIntegration.cuh:49-56
 * (src/TSDFIntegration.cu:252-299).  collision_counts must be zeroed by the caller. */
int vf_count_collisions_per_voxel(int32_t* collision_counts, const vf_float3* tsdf_vertices,
                                  const int32_t* knn_counts, vf_float3 voxel_size, vf_float3 volume_size,
                                  vf_int3 volume_dim, vf_mat3 volume_rotate, vf_float3 volume_translate,
                                  vf_stream stream);

/* af::runIntegrateWarpedTSDFFromFrameKernel, include/af/TSDFIntegration.cuh:34-47
 * (src/TSDFIntegration.cu:173-250).  max_weight replaces the reference's fixed
 * cap of 15 (src/TSDFIntegration.cu:64); pass 15 for the reference value. */
int vf_integrate_warped_tsdf(float* tsdf, float* tsdf_weights, const vf_float3* tsdf_vertices,
                             const int32_t* collision_counts, const int32_t* knn_counts, const float* depth_frame,
                             vf_int2 frame_dims, float delta, vf_float3 voxel_size, vf_float3 volume_size,
                             vf_int3 volume_dim, vf_mat3 volume_rotate, vf_float3 volume_translate,
                             vf_mat3 depth_intrinsics, float max_weight, vf_stream stream);

/* af::runIntegrateTSDFFromFrameKernel, include/af/TSDFIntegration.cuh:8-18
 * (src/TSDFIntegration.cu:69-116): plain projective integration. */
int vf_integrate_tsdf(float* tsdf, float* tsdf_weights, const float* depth_frame, vf_int2 frame_dims, float delta,
                      vf_int3 dims, vf_float3 voxel_size, vf_float3 volume_size, vf_mat3 volume_rotate,
                      vf_float3 volume_translate, vf_mat3 depth_intrinsics, float max_weight, vf_stream stream);

/* af::runIntegrateTSDFFromCameraFrameKernel, include/af/TSDFIntegration.cuh:20-32
 * (src/TSDFIntegration.cu:118-171). */
int vf_integrate_tsdf_camera_frame(float* tsdf, float* tsdf_weights, const float* depth_frame, vf_int2 frame_dims,
                                   float delta, vf_int3 dims, vf_float3 voxel_size, vf_float3 volume_size,
                                   vf_mat3 world2camera_rotate, vf_float3 world2camera_translate,
                                   vf_mat3 volume_rotate, vf_float3 volume_translate, vf_mat3 depth_intrinsics,
                                   float max_weight, vf_stream stream);

/* af::runTranslateKernel, include/af/VertexManipulation.cuh:16
 * (src/VertexManipulation.cu:73-87): mesh[i] += translation. */
int vf_translate(vf_float3* mesh, uint32_t mesh_size, vf_float3 translation, vf_stream stream);

/* ---- mesh points ---------------------------------------------------------- */
/* af::runComputeWarpWeightsKernel, include/af/VertexManipulation.cuh:19-23
 * (src/VertexManipulation.cu:105-145): for each point, all K slots,
 * w_j = d_j >= r ? 0 : exp(-d_j^2 / (2 r^2)), normalised when the sum is not 0.
 * A slot whose distance is FLT_MAX (a missing neighbour) weighs 0 without
 * reading its radius.  graph_size > 0 also treats indices >= graph_size as
 * missing; 0 leaves them unchecked, as the reference does. */
int vf_compute_warp_weights(int k, float* knn_weights, const float* graph_radii, const uint32_t* knn_idxs,
                            const float* knn_dists, size_t mesh_size, size_t graph_size, vf_stream stream);

/* af::runWarpKernel, include/af/ComputeCorrespondence.cuh:11-16
 * (src/ComputeCorrespondence.cu:5-43): p' = sum_j w_j (T_j [p; 1]).xyz over
 * all K slots, p itself when every |w_j| <= 1e-5.  warped_mesh may alias mesh.
 * Slots of weight exactly 0 are skipped (their index is not read). */
int vf_warp_mesh(int k, vf_float3* warped_mesh, const vf_float3* mesh, const vf_mat4* motion_graph,
                 const uint32_t* knn_idxs, const float* knn_weights, size_t mesh_size, size_t graph_size,
                 vf_stream stream);

/* af::getKnnData<K> (include/af/GraphBuilder.h:20-32; MotionGraph::knnSearch,
 * include/af/MotionGraph.h:50-70) on the device: the K nearest nodes of each
 * point by float L2 distance, slots ordered by (distance, node index);
 * exclude_closest drops the nearest (node-to-node knn).  Missing slots get
 * index UINT32_MAX and distance FLT_MAX.  Uses the workspace's scratch (NULL:
 * the device's default workspace). */
int vf_mesh_knn(vf_workspace* ws, int k, uint32_t* knn_idxs, float* knn_dists, const vf_float3* points, size_t n,
                const vf_float3* nodes, size_t graph_size, int exclude_closest, vf_stream stream);

/* The three above in one kernel (knn, weights, blend) for n points;
 * blend VF_BLEND_LBS is the reference's warp, VF_BLEND_DQB blends the nodes'
 * rigid parts as dual quaternions.  warped may alias points. */
int vf_warp_points(vf_workspace* ws, int k, int blend, vf_float3* warped, const vf_float3* points, size_t n,
                   const vf_float3* nodes, const float* radii, const vf_mat4* transforms, size_t graph_size,
                   vf_stream stream);

/* vf_warp_points without a workspace, allocating nothing: no counterpart in
 * the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Its scratch (the node grid and dual quaternions) is a static 4 MiB device
 * array of the library, one per GPU, held until the process ends; graphs of
 * at most 65536 nodes (VF_ERR_CAPACITY beyond).  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (on different
 * streams or host threads) race and must be serialised by the caller.
 * Same results as vf_warp_points. */
int vf_warp_points_static(int k, int blend, vf_float3* warped, const vf_float3* points, size_t n,
                          const vf_float3* nodes, const float* radii, const vf_mat4* transforms, size_t graph_size,
                          vf_stream stream);

/* ---- depth-frame ingest ---------------------------------------------------- */
/* af::runFilterDepthKernel, include/af/VertexManipulation.cuh:11
 * (src/VertexManipulation.cu:6-23): values above threshold, below 1e-4 or NaN
 * become fill_value (the reference's default: +infinity). */
int vf_filter_depth(float* depth, uint32_t depth_size, float threshold, float fill_value, vf_stream stream);
/* af::runMaskDepthKernel (:12; src/VertexManipulation.cu:25-40): mask[i] == 0
 * sets depth[i] to +infinity. */
int vf_mask_depth(float* depth, const uint8_t* mask, uint32_t depth_size, vf_stream stream);
/* af::bilateralFilterTextureOpmShared, include/af/BilateralFilter.cuh
 * (src/BilateralFilter.cu:52-127, 185-193): clamp-to-edge neighbourhood of
 * radius r <= 9, spatial and quantised intensity weights as the reference
 * tabulates them; infinite centres stay infinite, infinite neighbours and
 * neighbours differing by more than threshold are skipped.  output must not
 * alias input. */
int vf_bilateral_filter(float* output, const float* input, int height, int width, int radius, double sigma_i,
                        double sigma_s, float threshold, vf_stream stream);
/* af::runDepthToMeshKernel (:14; src/VertexManipulation.cu:42-71). */
int vf_depth_to_mesh(vf_float3* mesh, const float* depth, uint32_t width, uint32_t height, vf_mat3 intrinsics,
                     vf_stream stream);
/* af::runDepthVerticesNormalsKernel (:25-30; src/VertexManipulation.cu:147-217):
 * interior pixels; a normal whose angle to +z, acos(n.z), lies in
 * (angle_from, angle_to) is zeroed.  The test is n.z < cos(angle_from) and
 * n.z > cos(angle_to) in binary64, the reference's float acos being
 * unavailable. */
int vf_depth_normals(vf_float3* normals, vf_float3* vertices, const float* depth, int height, int width,
                     float angle_from, float angle_to, vf_stream stream);
/* af::runRemoveDepthWithoutNormalsKernel (:32; src/VertexManipulation.cu:219-241). */
int vf_remove_depth_without_normals(vf_float3* normals, vf_float3* vertices, float* depth, int height, int width,
                                    vf_stream stream);
/* The frame centroid of src/VolumeFusion.cu:216-219: sum of all points over
 * the number with z > 1e-4; the sum in binary64 in a fixed order (256-point
 * binary trees, then the blocks in order), rounded to float.  Blocks on
 * `stream` and returns the centroid and count on the host.  NULL workspace:
 * the device default. */
int vf_depth_centroid(vf_workspace* ws, vf_float3* centroid, uint64_t* count, const vf_float3* mesh, size_t n,
                      vf_stream stream);
/* The whole ingest of src/VolumeFusion.cu:190-219 in three launches: filter
 * (threshold), bilateral filter (bilateral_radius 0 skips it), back-projection,
 * normals, removal of points without normals, centroid.  depth is updated in
 * place as the reference leaves it; vertices and normals are W*H arrays
 * (border normals are not written).  Blocks on `stream` for the centroid. */
int vf_ingest_depth(vf_workspace* ws, float* depth, int width, int height, vf_mat3 intrinsics, float threshold,
                    int bilateral_radius, double sigma_i, double sigma_s, float bilateral_threshold,
                    float angle_from, float angle_to, vf_float3* vertices, vf_float3* normals, vf_float3* centroid,
                    uint64_t* count, vf_stream stream);

/* ---- TSDF volumes on disk ------------------------------------------------- */
/* VF_TSDF_FORMAT_NATIVE: a 64-byte header ("VFTSDF", version 1, header size,
 * dims, size, centre, delta, max weight, flags) then tsdf[N] and weight[N],
 * little-endian float32, z-major.  VF_TSDF_FORMAT_LEGACY: the reference's
 * TSDFVolume::save layout (src/TSDFVolume.cpp:339-358) with its 24-byte size
 * field written as the float size plus 12 zero bytes (the reference writes a
 * Vec3f with sizeof(double) * 3, :344) and zero colour arrays; loading reads
 * only the first 12 bytes of that field and skips the colours. */
enum { VF_TSDF_FORMAT_NATIVE = 1, VF_TSDF_FORMAT_LEGACY = 2 };

typedef struct vf_tsdf_info {
    vf_int3 dim;
    vf_float3 size;
    vf_float3 center;     /* native only (0 from legacy files)              */
    float delta;
    float max_weight;     /* native only                                    */
    int32_t format;       /* set by vf_tsdf_load                            */
} vf_tsdf_info;

/* Device arrays tsdf/weights of info->dim voxels to `path` (streamed through
 * a bounded host buffer; blocks). */
int vf_tsdf_save(const char* path, int format, const vf_tsdf_info* info, const float* tsdf, const float* weights);
/* Reads the header into *info and, when tsdf and weights are not NULL, the
 * volume into device arrays of at least `capacity` voxels. */
int vf_tsdf_load(const char* path, vf_tsdf_info* info, float* tsdf, float* weights, uint64_t capacity);
/* The same with host arrays (no device needed). */
int vf_tsdf_save_host(const char* path, int format, const vf_tsdf_info* info, const float* tsdf,
                      const float* weights);
int vf_tsdf_load_host(const char* path, vf_tsdf_info* info, float* tsdf, float* weights, uint64_t capacity);

/* ---- marching cubes ------------------------------------------------------- */
/* af::runMarchingCubesMeshKernel, include/af/MarchingCubes.cuh:10-17
 * (src/MarchingCubes.cu:422-522): one vertex per crossed edge of every cube
 * (i < X-2, j < Y-2, k < Z-2; a cube with a weight-0 corner is skipped).
 * Vertices are appended after *mesh_size (a device counter the caller zeroes,
 * as the reference does) in canonical order (cubes by linear index, edges by
 * number) where the reference's atomics give a racy order; *mesh_size
 * advances by the full count, and entries at or beyond mesh_capacity are not
 * written.  Uses the workspace's scratch (NULL: the device's default
 * workspace, for callers without one; such calls must not overlap).
 * VF_ERR_INVALID_ARGUMENT if a z slice (X x Y) holds 2^31 voxels or more. */
int vf_marching_cubes_mesh(vf_workspace* ws, vf_float3* mesh, uint32_t* mesh_size, uint64_t mesh_capacity,
                           const float* tsdf, const float* weights, vf_int3 dims, vf_float3 voxel_size,
                           vf_float3 volume_size, float iso, vf_stream stream);

/* af::runMarchingCubesFullKernel, include/af/MarchingCubes.cuh:19-28
 * (src/MarchingCubes.cu:526-702): three vertices and one face per
 * non-degenerate triangle of every cube, triangles in table order; face
 * entries are absolute vertex indices.  Same counters and capacities. */
int vf_marching_cubes_full(vf_workspace* ws, vf_float3* mesh, uint32_t* mesh_size, uint64_t mesh_capacity,
                           vf_int3* faces, uint32_t* faces_size, uint64_t faces_capacity, const float* tsdf,
                           const float* weights, vf_int3 dims, vf_float3 voxel_size, vf_float3 volume_size, float iso,
                           vf_stream stream);

/* Gaussian warp weight of src/TSDFWarp.cu:17 element by element,
 * out[i] = exp(-(dist[i]^2) / (2 radius[i]^2)), unnormalised: the function
 * every weight above is built from, exposed for diagnostics. */
int vf_eval_warp_weight(const float* radius, const float* dist, float* out, size_t n, vf_stream stream);

/* ------------------------------------------------------------------------ */
/* Fused per-frame warped integration                                       */
/* ------------------------------------------------------------------------ */
/* One call performs the whole TSDF-integration block of the reference's frame
 * loop, src/VolumeFusion.cu:702-742 (translate nodes by -c, knn, weights,
 * vertices, translate back, warp, count collisions, integrate, with volume
 * rotation I and translation c), in five kernels and with no per-voxel
 * scratch arrays.  The tsdf/weight arrays are updated in place; when
 * graph.points_out is not NULL it receives the node positions as the
 * reference's two translations leave them, fl(fl(p - c) + c). */

typedef struct vf_volume {
    float* tsdf;          /* float[N], z-major (DeviceTsdf::tsdfPtr)      */
    float* weights;       /* float[N]                                     */
    vf_int3 dim;
    vf_float3 voxel_size; /* size / float(dim) as the caller computes it  */
    vf_float3 size;
    vf_float3 center;     /* c: the frame centroid (volume translation)   */
} vf_volume;

typedef struct vf_graph {
    const vf_float3* points;  /* Vec3f[G] node positions (camera frame)    */
    vf_float3* points_out;    /* optional, may alias points                */
    const vf_mat4* transforms;/* Mat4f[G]                                  */
    const float* radii;       /* float[G]                                  */
    size_t count;             /* G                                         */
} vf_graph;

typedef struct vf_frame {
    const float* depth;       /* float[W*H], metres                        */
    int32_t width, height;
    vf_mat3 intrinsics;
} vf_frame;

/* Blend of the selected nodes' transforms.  VF_BLEND_LBS is the reference's
 * linear blend of Mat4f (src/TSDFWarp.cu:183-189).  VF_BLEND_DQB blends each
 * node's rigid part as a unit dual quaternion (Kavan et al., ACM TOG 27(4),
 * 2008); it has no counterpart in the reference. */
enum { VF_BLEND_LBS = 0, VF_BLEND_DQB = 1 };

typedef struct vf_warp_params {
    int k;                    /* knn slots, 1..8                           */
    float delta;              /* truncation distance                       */
    float max_weight;         /* weight cap (reference: 15)                */
    vf_int3 voxels_per_node;  /* node cube, 2*(int(R/vs)+1)+1 per axis     */
    int32_t blend;            /* VF_BLEND_LBS (default) or VF_BLEND_DQB    */
} vf_warp_params;

typedef struct vf_frame_stats {
    uint64_t supported;       /* voxels with at least one knn candidate    */
    uint64_t in_bounds;       /* supported voxels with target in volume    */
    uint64_t pending;         /* voxels that passed the projective tests   */
    uint64_t collided;        /* pending voxels dropped, target count > 1  */
    uint64_t updated;         /* voxels whose (tsdf, weight) changed       */
    uint64_t active_bricks;   /* bricks pass A swept                       */
    uint64_t overflow_bricks; /* bricks whose node list spilled            */
    uint64_t screened_bricks; /* bricks with a node list the screen proved
                                 unable to change the frame (not swept)    */
    uint64_t redo_bricks;     /* bricks whose pending updates exceeded the
                                 record budget (recomputed by the redo pass) */
} vf_frame_stats;


/* Workspace for up to `max_volumes` volumes per call, each with at most
 * max_voxels voxels and max_nodes nodes (total over the batch).  Allocates
 * device memory on the current device.  Calls on one workspace must be
 * ordered (one stream, or streams synchronised between calls): its buffers,
 * including the last few frames' descriptor sets kept on the device so that
 * a repeated call uploads nothing, are reused from call to call. */
int vf_workspace_create(vf_workspace** ws, int max_volumes, uint64_t max_total_voxels, uint64_t max_total_nodes,
                        vf_stream stream);
int vf_workspace_destroy(vf_workspace* ws);
/* Device bytes held by a workspace. */
uint64_t vf_workspace_bytes(const vf_workspace* ws);

/* One volume, one frame (src/VolumeFusion.cu:702-742). */
int vf_warped_integrate(vf_workspace* ws, const vf_volume* volume, const vf_graph* graph, const vf_frame* frame,
                        const vf_warp_params* params, vf_stream stream);

/* Many independent volumes, one frame each, in the same launches.
 * params[i] belongs to volume i; all volumes must share params[i].k. */
int vf_warped_integrate_batch(vf_workspace* ws, int n, const vf_volume* volumes, const vf_graph* graphs,
                              const vf_frame* frames, const vf_warp_params* params, vf_stream stream);

/* Statistics of the last call on this workspace (blocks on `stream`). */
int vf_workspace_last_stats(vf_workspace* ws, vf_frame_stats* stats, vf_stream stream);

/* The same counters summed over every call since the last reset (blocks on
 * `stream`); reset != 0 zeroes them afterwards. */
int vf_workspace_totals(vf_workspace* ws, vf_frame_stats* totals, int reset, vf_stream stream);

/* Per-kernel timing: when enabled, each call's kernel launches carry HIP
 * events (recorded at the kernels' start and end on the call's stream).
 * enable 1 times k_prep_bin with k_compact, k_pass_a and k_pass_b; enable 2
 * times k_pass_a alone (fewer events between the kernels); 0 turns timing
 * off.  vf_workspace_kernel_ms waits for the last event, returns the summed
 * milliseconds of the three (0 for those not timed) over the timed calls
 * since enabling (or the previous query) and starts over. */
int vf_workspace_set_timing(vf_workspace* ws, int enable);
int vf_workspace_kernel_ms(vf_workspace* ws, float* ms3, int* frames);

/* Launch tuning.  brick_shape selects the voxel brick one pass-A workgroup
 * sweeps per step (one voxel per thread); wg_per_cu caps the resident pass-A
 * workgroups per CU (0 = as many as fit).  Every setting gives identical
 * tsdf, weights and node positions; only the speed differs.  ms_per_frame is
 * what the autotuner measured (0 when set by hand).
 *
 * The brick screen (on unless flags has VF_TUNE_NO_SCREEN) drops, before the
 * voxel sweep, every brick that provably holds no voxel passing the
 * projective tests and no voxel whose warped target can be the target of
 * one: per-node displacement bounds of the blend, depth-tile maxima of the
 * frame.  It is used for LBS volumes.  With it, vf_frame_stats.supported and
 * in_bounds count the swept bricks' voxels only (pending, collided and
 * updated are unchanged). */
enum {
    VF_BRICK_16x8x4 = 0, /* 512 threads (default) */
    VF_BRICK_16x4x4 = 1, /* 256 threads */
    VF_BRICK_8x8x4 = 2,  /* 256 threads */
    VF_BRICK_8x8x8 = 3,  /* 512 threads */
    VF_BRICK_32x4x4 = 4, /* 512 threads */
    VF_BRICK_SHAPES = 5
};
enum { VF_TUNE_NO_SCREEN = 1 };
typedef struct vf_tuning {
    int32_t brick_shape;
    int32_t wg_per_cu;
    float ms_per_frame;
    int32_t flags;            /* VF_TUNE_NO_SCREEN: sweep every brick with a node list */
} vf_tuning;

/* Tuning used by later vf_warped_integrate* calls on this workspace. */
int vf_workspace_set_tuning(vf_workspace* ws, const vf_tuning* tuning);

/* Pending-record budget of later frames on this workspace (no counterpart in
 * the reference, whose per-voxel buffers hold every voxel's result): room for
 * ceil(records_per_voxel x the batch's voxels) pending updates, at most the
 * worst case min(N, G x cube volume) per volume; 0, the default, sizes for the
 * worst case.  The bricks whose updates do not fit are recomputed by one more
 * pass after pass A, which writes them in place; results and statistics are
 * the same.  With a budget the workspace's buffers grow to exactly what a
 * frame needs, without headroom, so that the owner can bound the workspace
 * (the serving loop does).  Range 0..1. */
int vf_workspace_set_record_budget(vf_workspace* ws, float records_per_voxel);

/* ------------------------------------------------------------------------ */
/* Frame plans: a fixed batch, replayed every frame                         */
/* ------------------------------------------------------------------------ */
/* A plan fixes a batch of volumes and the device buffers of their graphs and
 * frames; their contents (depth, transforms, node positions, tsdf) may change
 * between launches, the pointers may not.  The plan owns its workspace, and
 * with VF_PLAN_GRAPH it records its kernels into a HIP graph on the
 * first launch and replays that graph afterwards (the kernels read the frame
 * parity from device memory, so the recorded launches stay valid).  Creating a
 * plan uploads its descriptors on `stream` and waits for them.  A plan's
 * launches must be ordered (one stream, or events); plans do not share
 * buffers. */
typedef struct vf_plan vf_plan;
enum { VF_PLAN_GRAPH = 1 };

/* tuning NULL: the tuning the autotuner last chose for a batch of this size
 * class in this process, or the default. */
int vf_plan_create(vf_plan** plan, int n, const vf_volume* volumes, const vf_graph* graphs, const vf_frame* frames,
                   const vf_warp_params* params, const vf_tuning* tuning, int flags, vf_stream stream);
int vf_plan_destroy(vf_plan* plan);
/* One frame for every volume of the plan, asynchronous on `stream`. */
int vf_plan_launch(vf_plan* plan, vf_stream stream);
/* Times every (brick shape, workgroups per CU) pair for `frames` frames on
 * scratch copies of the plan's volumes (the volumes and node positions are
 * left as they are), keeps the fastest for this plan and for later plans of
 * the same size class, and returns it.  Blocks on `stream`. */
int vf_plan_autotune(vf_plan* plan, int frames, vf_tuning* chosen, vf_stream stream);
int vf_plan_get_tuning(const vf_plan* plan, vf_tuning* tuning);
/* Statistics of the last launch and the totals since the last reset
 * (either pointer may be NULL); blocks on `stream`. */
int vf_plan_stats(vf_plan* plan, vf_frame_stats* last, vf_frame_stats* totals, int reset, vf_stream stream);
uint64_t vf_plan_bytes(const vf_plan* plan);

/* ------------------------------------------------------------------------ */
/* Gauss-Newton motion optimisation                                         */
/* ------------------------------------------------------------------------ */
/* One ICP iteration of the reference (src/VolumeFusion.cu:434-692): residuals
 * of point-to-plane correspondences and of the regularisation graph, the
 * normal equations JTJ x = JTr, their LU solution and the first-order twist
 * update T <- T - [w]x|t of every node.  Vectors are vf_float3, transforms
 * column-major vf_mat4, updates float[6] per node (w0 w1 w2 t0 t1 t2).
 * The reference's kernels one by one:                                      */

/* af::runCorrespondencePairsKernel (include/af/ComputeCorrespondence.cuh:18-26,
 * src/ComputeCorrespondence.cu:46-82): mesh point i pairs with pixel
 * (int(x/z fx + cx), int(y/z fy + cy)) of a depth_dim.x-row, depth_dim.y-column
 * frame when the frame vertex there is within thresh and its depth above
 * 1e-4.  Pairs (i, pixel) are appended after *corr_size (device counter),
 * which advances by their count; they come in ascending i (the reference
 * appends them in the order of an atomic race).  Uses the workspace's
 * scratch (NULL: the device's default workspace). */
int vf_correspondence_pairs(vf_workspace* ws, vf_int2* corr, uint32_t* corr_size, const vf_float3* warped,
                            size_t mesh_size, vf_mat3 intrinsics, const float* depth, vf_int2 depth_dim,
                            const vf_float3* frame_mesh, float thresh, vf_stream stream);
/* af::runFill_r_DepthKernel (include/af/MotionOptimization.cuh:34-40):
 * r[c] = weight * n.dot(warped[corr.x] - depth_mesh[corr.y]). */
int vf_fill_r_depth(float* r, float weight, const vf_int2* corr, uint32_t corr_size, const vf_float3* warped,
                    const vf_float3* depth_mesh, const vf_float3* normals, vf_stream stream);
/* af::runFill_r_MRegKernel (:42-47): r[3 (g reg_k + i) + a] =
 * ((T_g p_g) - (T_nb p_g))[a] * weight for nb = graph_knn[g reg_k + i] < G
 * (rows of absent neighbours untouched). */
int vf_fill_r_mreg(float* r, float weight, const vf_float3* graph, const vf_mat4* transforms, uint32_t graph_size,
                   const uint32_t* graph_knn, int reg_k, vf_stream stream);
/* af::runFill_J_DepthKernel and af::runFill_J_MRegKernel (:11-32): the
 * reference's dense Jacobian, j_height rows of j_width float[6] (zeroed by
 * the caller); k <= 8 mesh knn slots, reg_k <= 8 neighbours. */
int vf_fill_j_depth(float* J, float weight, uint32_t j_height, uint32_t j_width, const vf_int2* corr,
                    uint32_t corr_size, const vf_float3* canon, int k, const uint32_t* knn_idxs,
                    const float* knn_weights, const vf_float3* depth_mesh, const vf_float3* normals,
                    vf_stream stream);
int vf_fill_j_mreg(float* J, float weight, uint32_t j_height, uint32_t j_width, const vf_float3* graph,
                   const vf_mat4* transforms, uint32_t graph_size, const uint32_t* graph_knn, int reg_k,
                   vf_stream stream);
/* af::runUpdateTransformsKernel (:88): T <- T - U, U = [[0 -w2 w1 t0];
 * [w2 0 -w0 t1]; [-w1 w0 0 t2]; [0 0 0 0]]. */
int vf_update_transforms(vf_mat4* transforms, const float* updates, uint32_t graph_size, vf_stream stream);
/* af::linearSolverLUStable (include/af/LinearSolver.cuh:20, src/LinearSolver.cu:179-225):
 * x = A^-1 b by LU with partial pivoting (rocSOLVER getrf, getrs) on a copy
 * of the column-major n x n A; VF_ERR_SINGULAR on a zero pivot.  Blocks. */
int vf_lu_solve(int n, const float* A, int lda, const float* b, float* x, vf_stream stream);

/* The sparse description of one ICP iteration's least-squares problem. */
typedef struct vf_gn_problem {
    const vf_int2* corr;        /* (mesh index, frame pixel) pairs         */
    uint32_t corr_size;
    const vf_float3* canon;     /* canonical mesh                          */
    int k;                      /* mesh knn slots, 1..8 (reference: 6)     */
    const uint32_t* knn_idxs;   /* [mesh][k] nearest nodes                  */
    const float* knn_weights;   /* [mesh][k] normalised weights             */
    const vf_float3* normals;   /* frame normals, by pixel                  */
    const vf_float3* graph;     /* node positions                           */
    uint32_t graph_size;
    const uint32_t* graph_knn;  /* [G][reg_k] regularisation neighbours, >= G: none */
    int reg_k;                  /* 1..8 (reference: 4)                      */
    float weight_depth;         /* settings.energyWeightDepth               */
    float weight_reg;           /* settings.energyWeightMReg                */
} vf_gn_problem;

/* The normal equations the reference assembles from its contributing pairs
 * (src/MotionOptimization.cu:227-479 with the sort and reduce_by_key of
 * src/VolumeFusion.cu:510-613), built here from the sparse rows without a
 * dense J: JTJ (6G x 6G) is the identity except for the 6x6 blocks of node
 * pairs that share a row, each summed over their common rows in ascending
 * row order (depth rows first, then the 3 reg_k rows of each node in order);
 * JTr (G x 6) likewise from r (depth rows, then regularisation rows).  Either
 * output may be NULL.  Blocks. */
int vf_gn_normal_equations(vf_workspace* ws, float* JTJ, float* JTr, const float* r, const vf_gn_problem* problem,
                           vf_stream stream);

/* A solver for one ICP iteration: builds the rows and JTJ and factorises
 * JTJ once (its value does not change during the Gauss-Newton iterations):
 * by Cholesky, JTJ being symmetric positive definite, or, where that finds a
 * non-positive pivot, by the LU with partial pivoting the reference uses
 * (VF_ERR_SINGULAR on a zero pivot).  The problem's buffers must stay alive
 * and unchanged while it exists. */
typedef struct vf_gn_solver vf_gn_solver;
typedef struct vf_gn_report {
    int iterations;       /* updates applied                               */
    int stopped;          /* 1: the energy decrease fell below min_step    */
    double final_energy;  /* energy of the last accepted transforms        */
    double energy[32];    /* energy at the start of each iteration         */
    int cholesky;         /* 1: JTJ was factorised by Cholesky, 0: by LU   */
} vf_gn_report;
int vf_gn_solver_create(vf_gn_solver** solver, const vf_gn_problem* problem, vf_stream stream);
/* No counterpart in the reference (whose solver always runs its LU): the
 * same with the factorisation chosen.  VF_GN_FACTOR_AUTO is
 * vf_gn_solver_create's (Cholesky, the LU where a pivot is not positive);
 * VF_GN_FACTOR_LU always takes the reference's LU with partial pivoting
 * (twice the flops), for callers who want its pivoting. */
enum { VF_GN_FACTOR_AUTO = 0, VF_GN_FACTOR_LU = 1 };
int vf_gn_solver_create_ex(vf_gn_solver** solver, const vf_gn_problem* problem, int factorisation,
                           vf_stream stream);
int vf_gn_solver_destroy(vf_gn_solver* solver);
/* The Gauss-Newton loop of src/VolumeFusion.cu:541-691: pending starts as a
 * copy of transforms; each iteration computes the residuals (depth rows from
 * warped, the mesh warped by pending, and depth_mesh; regularisation rows
 * from pending) and their energy (sum of squares in binary64, fixed order;
 * the reference sums in float in thrust's order); stops when the energy drops
 * by less than min_step; else commits pending to transforms (after the first
 * iteration), solves JTJ x = JTr, updates pending and re-warps the mesh.
 * Blocks; reads back one double per iteration and one int at the end.
 * VF_ERR_HIP if a triangular solve's wait for a solved block gave up (a
 * bounded poll; the solve's waves end regardless). */
int vf_gn_solver_iterate(vf_gn_solver* solver, vf_mat4* transforms, vf_mat4* pending, vf_float3* warped,
                         const vf_float3* canon, size_t mesh_size, const vf_float3* depth_mesh, int max_iters,
                         float min_step, vf_gn_report* report, vf_stream stream);

/* ------------------------------------------------------------------------ */
/* Serving: resident reconstruction streams, continuous batching            */
/* ------------------------------------------------------------------------ */
/* A server keeps each open stream's TSDF volume (tsdf -1, weight 0 at open),
 * node positions, radii, transforms and depth buffer on the device.  Frames
 * (a host depth image and the node transforms) are submitted as they arrive;
 * a dispatcher thread forms a batch from the streams that have a waiting
 * frame and no frame in flight (oldest first, at most max_batch), uploads the
 * inputs and launches one batched frame on one of `lanes` HIP streams, each
 * with its own workspace, so that batches overlap on the device while every
 * stream's frames stay in order.  A completion thread stamps each frame when
 * its batch has finished.  Streams are admitted while the device memory they
 * need fits the budget.  Per-volume results do not depend on what else is in
 * the batch. */
typedef struct vf_server vf_server;

typedef struct vf_server_config {
    uint64_t hbm_budget; /* bytes of device memory for streams and workspaces */
    int32_t k;           /* knn slots, shared by every stream (1..8)          */
    int32_t max_batch;   /* frames per launch, 1..256                         */
    int32_t lanes;       /* batches in flight at once, 1..8                   */
    int32_t device;      /* HIP device                                        */
} vf_server_config;

typedef struct vf_stream_desc {
    vf_int3 dim;
    vf_float3 size;
    vf_float3 center;
    int32_t width, height;    /* depth frame size                            */
    vf_mat3 intrinsics;
    float delta;
    float max_weight;
    vf_int3 voxels_per_node;
    uint32_t nodes;           /* G                                           */
    const vf_float3* points;  /* host Vec3f[G], copied at open               */
    const float* radii;       /* host float[G], copied at open               */
    int32_t blend;            /* VF_BLEND_LBS or VF_BLEND_DQB                */
} vf_stream_desc;

typedef struct vf_frame_result {
    uint64_t ticket;
    int32_t stream_id;
    int32_t status;           /* VF_OK, or the status of its batch's launch  */
    double submit_s;          /* steady-clock seconds at submission          */
    double issue_s;           /* its batch was handed to the device          */
    double done_s;            /* its batch was seen complete                 */
    uint64_t batch;           /* batch sequence number                       */
    uint32_t batch_frames;    /* frames in that batch                        */
    uint32_t lane;
} vf_frame_result;

typedef struct vf_server_stats {
    uint64_t frames_submitted;
    uint64_t frames_done;
    uint64_t batches;
    uint64_t bytes_in_use;    /* admitted streams' device memory estimate    */
    uint64_t hbm_budget;
    int32_t streams_open;
    int32_t max_batch_seen;
    vf_frame_stats totals;    /* kernel counters summed over every batch     */
} vf_server_stats;

/* The most a stream with this description adds to a server's admission
 * charge: its buffers plus `lanes` times what it adds to a lane workspace
 * whose batch holds it.  Lane workspaces carry a pending-record budget of
 * 1/24 record per voxel (vf_workspace_set_record_budget) and grow to exactly
 * what their batches need, so a server charges its open streams' buffers
 * plus `lanes` times, for each of the workspace's buffer groups (node
 * records, brick state, collision tiles, pending records, node lists,
 * descriptors, depth tiles), the sum of its max_batch largest shares; a
 * stream is admitted when that charge, with it, fits the HBM budget. */
uint64_t vf_server_stream_bytes(const vf_stream_desc* desc, int lanes);
int vf_server_create(vf_server** server, const vf_server_config* config);
/* Finishes every submitted frame, then frees everything. */
int vf_server_destroy(vf_server* server);
/* Admits a stream; when its bytes do not fit the budget, waits up to wait_ms
 * for streams to close, then fails with VF_ERR_CAPACITY. */
int vf_server_open(vf_server* server, const vf_stream_desc* desc, int wait_ms, int32_t* stream_id);
/* Waits for the stream's frames, then releases its memory. */
int vf_server_close(vf_server* server, int32_t stream_id);
/* Queues one frame: depth float[W*H] and transforms Mat4f[G] (host memory,
 * copied before the call returns). */
int vf_server_submit(vf_server* server, int32_t stream_id, const float* depth, const vf_mat4* transforms,
                     uint64_t* ticket);
/* The same without the copy: depth and transforms are read where they are
 * (pinned host or device memory) when the frame's batch is issued, so the
 * caller leaves them untouched until the frame is done. */
int vf_server_submit_borrowed(vf_server* server, int32_t stream_id, const float* depth, const vf_mat4* transforms,
                              uint64_t* ticket);
/* Blocks until the frame is done. */
int vf_server_wait(vf_server* server, uint64_t ticket, vf_frame_result* result);
/* Completed frames not returned by an earlier poll, oldest first. */
int vf_server_poll(vf_server* server, vf_frame_result* results, int capacity, int* count);
/* Blocks until every submitted frame is done. */
int vf_server_drain(vf_server* server);
/* Copies a stream's volume and node positions to the host once its frames
 * are done (any pointer may be NULL). */
int vf_server_read(vf_server* server, int32_t stream_id, float* tsdf, float* weights, vf_float3* points);
int vf_server_get_stats(vf_server* server, vf_server_stats* stats);
/* Checkpoint a stream's volume to a TSDF file once its frames are done, and
 * restore it from one (same dims) into an open stream: resume after a
 * restart.  Node positions and transforms are not part of the file. */
int vf_server_save(vf_server* server, int32_t stream_id, const char* path, int format);
int vf_server_restore(vf_server* server, int32_t stream_id, const char* path);

/* ------------------------------------------------------------------------ */
/* Device plumbing for hosts without a GPU framework (Python via ctypes)    */
/* ------------------------------------------------------------------------ */
int vf_device_count(int* count);
int vf_set_device(int device);
int vf_malloc(void** ptr, uint64_t bytes);
int vf_free(void* ptr);
int vf_memcpy_h2d(void* dst, const void* src, uint64_t bytes, vf_stream stream);
int vf_memcpy_d2h(void* dst, const void* src, uint64_t bytes, vf_stream stream);
int vf_memcpy_d2d(void* dst, const void* src, uint64_t bytes, vf_stream stream);
int vf_memset(void* dst, int value, uint64_t bytes, vf_stream stream);
int vf_fill_f32(float* dst, float value, uint64_t count, vf_stream stream);
int vf_stream_create(vf_stream* stream);
int vf_stream_destroy(vf_stream stream);
int vf_stream_synchronize(vf_stream stream);
int vf_device_synchronize(void);
int vf_host_alloc(void** ptr, uint64_t bytes);
int vf_host_free(void* ptr);
/* Event timing on a stream: record start/stop, read milliseconds. */
int vf_event_create(void** event);
int vf_event_destroy(void* event);
int vf_event_record(void* event, vf_stream stream);
int vf_event_elapsed_ms(float* ms, void* start, void* stop);
int vf_event_synchronize(void* event);

#ifdef __cplusplus
}
#endif
#endif
