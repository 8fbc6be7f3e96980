/*
 * vf_oracle.h -- CPU restatement of VolumeFusion's warped TSDF integration path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP kernels
 * are compared against.  Only the test suite, the smoke check and the CPU
 * baseline leg of bench.py load it; the shipped library never links, loads or
 * calls it.
 *
 * Every function restates one reference function (file:line in the comment
 * beside it) with the same floating-point operation order as the reference's
 * device code (Eigen's scalar, coefficient-based evaluation), every operation
 * rounded on its own (no fused multiply-add contraction).  Vectors are packed
 * float triples (Eigen::Vector3f, 12 bytes); 4x4 and 3x3 matrices are
 * column-major float arrays (Eigen default storage).
 *
 * What pins each function (tests/test_oracle_pins.py unless noted):
 *   vfo_expw              glibc exp rounded to binary32 on the weight domain
 *                         and a wide range (0 mismatches), special values.
 *                         Against the reference's device exp (the CUDA math
 *                         library, absent here): parity unpinned.
 *   vfo_weight_tsdf       closed forms: d = 0 -> 1, d = r -> e^-1/2, d = 2r -> e^-2
 *   vfo_weight_mesh_host, vfo_mesh_warp
 *                         MotionGraphTest.Warp golden vectors
 *                         (test/src/MotionGraphTest.cpp:98-133), bit for bit
 *   vfo_div_markstein     IEEE division, exhaustively over mantissas
 *   vfo_mat3_inverse      M * inverse = I; cofactor order of Eigen's inverse
 *   vfo_translate, vfo_generate_vertices
 *                         closed forms; the translate round trip is not the
 *                         identity (a reference quirk the tests keep)
 *   vfo_knn_graph         brute force over every (voxel, node) pair with the
 *                         float decisions replayed, k = 1, 3, 4, 8; cube
 *                         truncation, boundary distances and ties
 *   vfo_warp_weights      closed forms (sum 1, one node -> 1)
 *   vfo_warp_tsdf         closed forms (k = 1 -> T p, identity -> p); Eigen's
 *                         summation order
 *   vfo_count_collisions, vfo_integrate_warped
 *                         identity warp with k = 1 equals plain integration
 *                         except on the voxels a numpy float32 round trip
 *                         predicts to collide or leave the volume
 *   vfo_integrate_plain   fronto-parallel plane closed form, pixel rounding,
 *                         running average and weight cap
 *   vfo_integrate_camera_frame
 *                         identity pose equals vfo_integrate_plain
 *   vfo_tsdfvolume_integrate
 *                         equals the device semantics on one frame within
 *                         4 ulps (the reference's own GPU-vs-CPU test,
 *                         test/src/TSDFIntegrationTest.cu:56-59)
 *   vfo_warped_frame      equals the independently structured voxel-major
 *                         vfo_warped_frame_gather over several frames; every
 *                         stage above
 *   vfo_warped_frame_f64  reporting only (max/RMS error of the float path);
 *                         vfo_warped_frame_gather's binary64 companion
 *                         (tsdf64/wt64 not NULL) equals it bit for bit
 *   vfo_mat4_to_dq, vfo_dqb_apply (DQB, no reference counterpart)
 *                         closed forms: identity -> p bit for bit; one rigid
 *                         node -> R p + t; identity and a rotation by theta,
 *                         equal weights -> rotation by theta/2; +179 and -179
 *                         degrees -> 180 (hemisphere alignment); pure
 *                         translations blend linearly; against the binary64
 *                         versions within a stated bound
 * The knn selection when a voxel has more than K candidates is a race in the
 * reference (src/TSDFWarp.cu:61-67); the canonical (distance, index) rule
 * used here is pinned by brute force, not by the reference.
 */
#ifndef VF_ORACLE_H
#define VF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Volume geometry as the reference passes it: dims, voxel size, volume size. */
typedef struct {
    int32_t dim[3];
    float vs[3];   /* voxel size = size / float(dim), src/VolumeFusion.cu:229 */
    float size[3];
} vfo_vol;

typedef struct {
    int64_t supported;      /* voxels with knn count > 0                       */
    int64_t in_bounds;      /* supported voxels whose warped target is inside  */
    int64_t collided;       /* supported voxels skipped because col[u] > 1     */
    int64_t updated;        /* voxels whose (tsdf, weight) changed             */
    int64_t candidates;     /* (voxel, node) pairs passing the 2r test         */
    int64_t overflow;       /* voxels with more than K candidates              */
} vfo_stats;

/* ---- scalar helpers ---------------------------------------------------- */
float vfo_expw(float x);
float vfo_exp_libm(float x);
float vfo_weight_tsdf(float radius, float dist);
float vfo_weight_mesh_host(float radius, float dist);
float vfo_div_markstein(float a, float b, float recip_b);
int64_t vfo_markstein_mismatches(float b, uint32_t a_lo_bits, uint32_t a_hi_bits);
void vfo_mat3_inverse(const float* m, float* out);
int32_t vfo_cvt_i32_rz(float f);

/* ---- reference kernels, one by one --------------------------------------- */
void vfo_translate(float* pts, size_t n, const float* t);
void vfo_generate_vertices(float* verts, const vfo_vol* vol);
int64_t vfo_knn_graph(int K, uint32_t* idx, float* dist, int32_t* cnt, const float* nodes, const float* radii,
                      size_t G, const vfo_vol* vol, const int32_t* cube);
void vfo_warp_weights(int K, float* w, const float* radii, const uint32_t* idx, const float* dist, const int32_t* cnt,
                      size_t N);
void vfo_warp_tsdf(int K, float* verts, const float* T, const uint32_t* idx, const float* w, const int32_t* cnt,
                   size_t N);
void vfo_count_collisions(int32_t* col, const float* verts, const int32_t* cnt, const vfo_vol* vol, const float* R,
                          const float* t);
int64_t vfo_integrate_warped(float* tsdf, float* wt, const float* verts, const int32_t* col, const int32_t* cnt,
                             const float* depth, int W, int H, float delta, const vfo_vol* vol, const float* R,
                             const float* t, const float* Kin, float maxw);
int64_t vfo_integrate_plain(float* tsdf, float* wt, const float* depth, int W, int H, float delta,
                            const vfo_vol* vol, const float* R, const float* t, const float* Kin, float maxw);
int64_t vfo_integrate_camera_frame(float* tsdf, float* wt, const float* depth, int W, int H, float delta,
                                   const vfo_vol* vol, const float* Rw2c, const float* tw2c, const float* R,
                                   const float* t, const float* Kin, float maxw);
int64_t vfo_tsdfvolume_integrate(float* tsdf, float* wt, const float* depth, int W, int H, float delta,
                                 const vfo_vol* vol, const float* pose, const float* Kin);
void vfo_mesh_warp(float* out, const float* pts, size_t n, const float* nodes, const float* radii, const float* T,
                   size_t G, int K, int host_order, int squared_dist);

/* ---- mesh points: knn, weights, warp -------------------------------------- */
void vfo_mesh_knn(int K, uint32_t* idx, float* dist, const float* pts, size_t n, const float* nodes, size_t G,
                  int exclude_closest);
void vfo_compute_warp_weights(int K, float* w, const float* radii, const uint32_t* idx, const float* dist, size_t n,
                              size_t G);
void vfo_warp_mesh(int K, float* out, const float* pts, const float* T, const uint32_t* idx, const float* w, size_t n,
                   size_t G);
void vfo_warp_points(int K, float* out, const float* pts, size_t n, const float* nodes, const float* radii,
                     const float* T, size_t G, int blend);

/* ---- depth ingest ----------------------------------------------------------- */
void vfo_filter_depth(float* depth, size_t n, float threshold, float fill);
void vfo_mask_depth(float* depth, const uint8_t* mask, size_t n);
void vfo_depth_to_mesh(float* mesh, const float* depth, int W, int H, const float* K);
void vfo_depth_normals(float* normals, const float* v, int H, int W, float a0, float a1);
void vfo_remove_depth_without_normals(const float* normals, float* v, float* depth, int H, int W);
void vfo_depth_centroid(const float* mesh, size_t n, float* out, uint64_t* count);
void vfo_bilateral_tables(int r, double sd, float* g2d, int32_t* clr);
void vfo_bilateral(float* out, const float* in, int H, int W, int r, double sI, double sS, float thr);

/* ---- marching cubes (src/MarchingCubes.cu), canonical output order -------- */
void vfo_marching_cubes(int full, const float* tsdf, const float* wt, const vfo_vol* vol, float iso, float* verts,
                        uint64_t cap_verts, int32_t* faces, uint64_t cap_faces, uint64_t* n_verts,
                        uint64_t* n_faces);

/* ---- dual-quaternion blending (second blend mode; not in the reference) --- */
/* dq: 8 floats, real (x, y, z, w) then dual (x, y, z, w). */
void vfo_mat4_to_dq(const float* M, float* dq);
void vfo_dqb_apply(const float* const* dq, const float* w, int n, const float* p, float* out);
void vfo_mat4_to_dq_f64(const float* M, double* dq);
void vfo_dqb_apply_f64(const double* const* dq, const double* w, int n, const double* p, double* out);
void vfo_warp_tsdf_dqb(int K, float* verts, const float* DQ, const uint32_t* idx, const float* w, const int32_t* cnt,
                       size_t N);

/* ---- the per-frame call sequence (src/VolumeFusion.cu:699-742) ----------- */
/* blend: 0 = the reference's linear blend of Mat4f, 1 = DQB of each node's
 * rigid part. */
/* Optional per-voxel outputs of one frame; any pointer may be NULL. */
typedef struct {
    int32_t* cnt;     /* [N]   knn counts                        */
    uint32_t* idx;    /* [N*K] knn node indices (sorted slots)   */
    float* dist;      /* [N*K] knn distances                     */
    float* w;         /* [N*K] normalised blend weights          */
    float* verts;     /* [N*3] warped vertices (camera frame)    */
    int32_t* col;     /* [N]   collision counts                  */
    int32_t* pix;     /* [N]   depth pixel used, -1 if no update */
} vfo_frame_out;

int vfo_warped_frame(int K, float* tsdf, float* wt, float* nodes, const float* radii, const float* T, size_t G,
                     const vfo_vol* vol, const int32_t* cube, const float* c, const float* depth, int W, int H,
                     float delta, const float* Kin, float maxw, vfo_stats* st, const vfo_frame_out* out,
                     int blend);

int vfo_warped_frame_gather(int K, float* tsdf, float* wt, float* nodes, const float* radii, const float* T,
                            size_t G, const vfo_vol* vol, const int32_t* cube, const float* c, const float* depth,
                            int W, int H, float delta, const float* Kin, float maxw, int nthreads, int z0, int z1,
                            vfo_stats* st, int blend, double* tsdf64, double* wt64);

int vfo_warped_frame_f64(int K, double* tsdf64, double* wt64, const float* nodes, const float* radii,
                         const float* T, size_t G, const vfo_vol* vol, const float* c, const float* depth, int W,
                         float delta, float maxw, const int32_t* cnt, const uint32_t* idx, const int32_t* pix,
                         int blend);

/* ---- Gauss-Newton motion optimisation (vf_gn.c) ---------------------------- */
int64_t vfo_corr_pairs(int32_t* corr, const float* warped, size_t M, const float* Kin, const float* depth, int rows,
                       int cols, const float* frame_mesh, float thresh);
void vfo_gn_r_depth(float* r, float wD, const int32_t* corr, size_t C, const float* warped, const float* dmesh,
                    const float* dnormals);
void vfo_gn_r_mreg(float* r, float wR, const float* graph, const float* T, size_t G, const uint32_t* gknn, int R);
void vfo_gn_j_depth(float* J, float wD, size_t C, size_t G, const int32_t* corr, const float* canon,
                    const uint32_t* knn, const float* kw, int K, const float* dnormals);
void vfo_gn_j_mreg(float* J, float wR, const float* graph, size_t G, const uint32_t* gknn, int R);
int vfo_gn_normal_eq(float* JTJ, float* JTr, const float* J, const float* r, size_t C, size_t G, const int32_t* corr,
                     const uint32_t* knn, const float* kw, int K, const uint32_t* gknn, int R);
void vfo_gn_update(float* T, const float* u, size_t G);
int vfo_lu_solve(int n, const float* A, const float* b, double* x);

#ifdef __cplusplus
}
#endif
#endif
