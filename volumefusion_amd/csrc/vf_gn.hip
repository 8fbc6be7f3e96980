// vf_gn.hip -- Gauss-Newton motion optimisation: the producer of the node
// transforms the warped integration consumes.
//
// The reference runs, per ICP iteration (src/VolumeFusion.cu:434-692):
// projective correspondences, a dense Jacobian J (rows x nodes of Vec6f),
// an atomic enumeration of the node pairs that share a row, thrust sorts and
// reduce_by_key into JTJ / JTr, a dense cuSOLVER LU factorisation and solve,
// and the twist update T <- T - [w]x|t -- with JTJ rebuilt and refactorised on
// every Gauss-Newton iteration although its value never changes inside one
// ICP iteration (J depends on the canonical mesh, the weights, the normals and
// the node positions only).
//
// Here:
//   - the row structure (which nodes each row couples, and their Jacobian
//     entries) is built once per ICP iteration from the sparse inputs; no
//     dense J exists;
//   - each node's rows are sorted by row index, and one wave per node sums
//     its 6x6 blocks and its JTr entry over those rows in ascending order
//     (deterministic; the reference's order is an atomic race);
//   - JTJ is factorised once (a blocked Cholesky factorisation: JTJ is
//     symmetric positive definite, so its Cholesky factor solves the system
//     the reference's LU with partial pivoting solves, in half the work;
//     rocSOLVER getrf where a pivot is not positive), and each Gauss-Newton
//     iteration only forms JTr from the new residuals and back-substitutes;
//   - residuals, energy (binary64, fixed order), update and the mesh warp
//     stay on the device; one read-back per iteration carries the energy the
//     reference's stopping rule needs.
// The reference's per-kernel entry points (residuals, dense J fills, the
// update, the LU solve, correspondences) are kept for drop-in callers.
#include <algorithm>
#include <cstring>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "vf_host.hpp"

namespace vf {
namespace {

constexpr int kGnMaxK = 8;        // mesh knn slots (the reference's motionGraphKNN is 6)
constexpr int kGnMaxR = 8;        // regularisation neighbours (the reference's energyMRegKNN is 4)
constexpr int kGnMaxNbr = 128;    // distinct nodes sharing rows with one node
constexpr int kGnMaxRows = 8192;  // rows of one node (sorted in LDS)
constexpr uint32_t kGnNone = 0xffffffffu;

inline unsigned blocks_for(uint64_t n, int b) { return (unsigned)((n + (uint64_t)b - 1) / (uint64_t)b); }

__device__ __forceinline__ float sum3(float a0, float a1, float a2) { return a0 + (a1 + a2); }

// Row a of gradTv(p), src/MotionOptimization.cu:10-27, as selects (no
// branches: with a run-time row index, the if/else-if form was lowered for
// gfx950 with g[1] = 0 on row 2, seen in the kernel's assembly).
__device__ __forceinline__ void grad_row(const float p[3], int a, float g[6]) {
    const bool r0 = a == 0, r1 = a == 1, r2 = a == 2;
    g[0] = r1 ? -p[2] : (r2 ? p[1] : 0.0f);
    g[1] = r0 ? p[2] : (r2 ? -p[0] : 0.0f);
    g[2] = r0 ? -p[1] : (r1 ? p[0] : 0.0f);
    g[3] = r0 ? 1.0f : 0.0f;
    g[4] = r1 ? 1.0f : 0.0f;
    g[5] = r2 ? 1.0f : 0.0f;
}

// (cCurrDepthNormal * weight).transpose() * gradTv * energyWeight
// (src/MotionOptimization.cu:57): a three-term sum per column, then the weight.
__device__ __forceinline__ void j_depth_value(const float n[3], float w, const float p[3], float wD, float out[6]) {
    const float nw[3] = {n[0] * w, n[1] * w, n[2] * w};
    float g0[6], g1[6], g2[6];
    grad_row(p, 0, g0);
    grad_row(p, 1, g1);
    grad_row(p, 2, g2);
#pragma unroll
    for (int k = 0; k < 6; ++k) out[k] = sum3(nw[0] * g0[k], nw[1] * g1[k], nw[2] * g2[k]) * wD;
}

// (M * p4).head(3) with a materialised Vec4f: (a0 + a1) + (a2 + a3).
__device__ __forceinline__ void mat4_apply4(const float* M, const float p[3], float o[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float a0 = M[0 * 4 + i] * p[0];
        const float a1 = M[1 * 4 + i] * p[1];
        const float a2 = M[2 * 4 + i] * p[2];
        const float a3 = M[3 * 4 + i] * 1.0f;
        o[i] = (a0 + a1) + (a2 + a3);
    }
}

__device__ __forceinline__ void load3(const float* p, float o[3]) {
    o[0] = p[0];
    o[1] = p[1];
    o[2] = p[2];
}

// ---- correspondences (src/ComputeCorrespondence.cu:46-82) --------------------

struct CorrIn {
    const float* warped;
    uint64_t n;
    float fx, fy, cx, cy;
    const float* depth;
    int32_t rows, cols;
    const float* frame_mesh;
    float thresh;
};

// The pixel of mesh point i when it forms a pair, else -1.
__device__ __forceinline__ int64_t corr_pixel(const CorrIn& in, uint64_t i) {
    float v[3];
    load3(in.warped + 3 * i, v);
    const float y0 = v[1] / v[2];
    const float x0 = v[0] / v[2];
    const int32_t y = cvt_i32_rz(y0 * in.fy + in.cy);
    const int32_t x = cvt_i32_rz(x0 * in.fx + in.cx);
    if (y < 0 || x < 0 || y >= in.rows || x >= in.cols) return -1;
    const int64_t d = (int64_t)y * in.cols + x;
    const float* fm = in.frame_mesh + 3 * d;
    const float dx = fm[0] - v[0], dy = fm[1] - v[1], dz = fm[2] - v[2];
    const float nrm = __builtin_sqrtf(sum3(dx * dx, dy * dy, dz * dz));
    if (nrm > in.thresh) return -1;
    return in.depth[d] > 1.e-4f ? d : -1;
}

constexpr int kCorrBlock = 256;

__global__ __launch_bounds__(kCorrBlock) void k_corr_count(CorrIn in, uint32_t* __restrict__ block_count) {
    const uint64_t i = (uint64_t)blockIdx.x * kCorrBlock + threadIdx.x;
    const bool ok = i < in.n && corr_pixel(in, i) >= 0;
    const unsigned long long m = __ballot(ok);
    __shared__ uint32_t s_w[kCorrBlock / 64];
    if (__lane_id() == 0) s_w[threadIdx.x >> 6] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kCorrBlock / 64; ++w) t += s_w[w];
        block_count[blockIdx.x] = t;
    }
}

// Exclusive scan of the block counts after *size; advances *size.
__global__ __launch_bounds__(1024) void k_scan_counts(uint32_t* __restrict__ counts, uint32_t nb,
                                                      uint32_t* __restrict__ size) {
    __shared__ uint32_t s_part[1024];
    const uint32_t per = (nb + 1023) / 1024;
    const uint32_t b = threadIdx.x * per, e = b + per < nb ? b + per : nb;
    uint32_t t = 0;
    for (uint32_t i = b; i < e; ++i) t += counts[i];
    s_part[threadIdx.x] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = size ? *size : 0u;
        for (int i = 0; i < 1024; ++i) {
            const uint32_t v = s_part[i];
            s_part[i] = run;
            run += v;
        }
        if (size) *size = run;
    }
    __syncthreads();
    uint32_t run = s_part[threadIdx.x];
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t v = counts[i];
        counts[i] = run;
        run += v;
    }
}

__global__ __launch_bounds__(kCorrBlock) void k_corr_write(CorrIn in, const uint32_t* __restrict__ block_off,
                                                           int32_t* __restrict__ corr) {
    const uint64_t i = (uint64_t)blockIdx.x * kCorrBlock + threadIdx.x;
    const int64_t d = i < in.n ? corr_pixel(in, i) : -1;
    const bool ok = d >= 0;
    const unsigned long long m = __ballot(ok);
    __shared__ uint32_t s_w[kCorrBlock / 64];
    const int wave = threadIdx.x >> 6;
    if (__lane_id() == 0) s_w[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (ok) {
        uint32_t pos = block_off[blockIdx.x];
        for (int w = 0; w < wave; ++w) pos += s_w[w];
        pos += (uint32_t)__popcll(m & ((1ull << __lane_id()) - 1ull));
        corr[2ull * pos] = (int32_t)i;
        corr[2ull * pos + 1] = (int32_t)d;
    }
}

// ---- residuals, dense Jacobians, update ----------------------------------------

// fill_r_DepthKernel (src/MotionOptimization.cu:165-181).
__global__ void k_r_depth(float* __restrict__ r, float wD, const int32_t* __restrict__ corr, uint32_t C,
                          const float* __restrict__ warped, const float* __restrict__ dmesh,
                          const float* __restrict__ dnormals) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    float v[3], d[3], nn[3];
    load3(warped + 3ull * (uint32_t)corr[2 * c], v);
    load3(dmesh + 3ull * (uint32_t)corr[2 * c + 1], d);
    load3(dnormals + 3ull * (uint32_t)corr[2 * c + 1], nn);
    const float df0 = v[0] - d[0], df1 = v[1] - d[1], df2 = v[2] - d[2];
    r[c] = wD * sum3(nn[0] * df0, nn[1] * df1, nn[2] * df2);
}

// fill_r_MRegKernel (src/MotionOptimization.cu:201-225), one thread per
// (node, neighbour slot).
__global__ void k_r_mreg(float* __restrict__ r, float wR, const float* __restrict__ graph,
                         const float* __restrict__ T, uint32_t G, const uint32_t* __restrict__ gknn, int R) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= G * (uint32_t)R) return;
    const uint32_t g = t / (uint32_t)R;
    const uint32_t nb = gknn[t];
    if (nb >= G) return;
    float p[3], tc[3], tn[3];
    load3(graph + 3ull * g, p);
    mat4_apply4(T + 16ull * g, p, tc);
    mat4_apply4(T + 16ull * nb, p, tn);
#pragma unroll
    for (int a = 0; a < 3; ++a) r[3ull * t + a] = (tc[a] - tn[a]) * wR;
}

// fill_J_DepthKernel (src/MotionOptimization.cu:29-60) into the dense J.
__global__ void k_j_depth(float* __restrict__ J, float wD, uint32_t jW, const int32_t* __restrict__ corr, uint32_t C,
                          const float* __restrict__ canon, int K, const uint32_t* __restrict__ knn,
                          const float* __restrict__ kw, const float* __restrict__ dnormals) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const uint32_t ci = (uint32_t)corr[2 * c], di = (uint32_t)corr[2 * c + 1];
    float p[3], nn[3];
    load3(canon + 3ull * ci, p);
    load3(dnormals + 3ull * di, nn);
    for (int s = 0; s < K; ++s) {
        const float w = kw[(uint64_t)K * ci + s];
        if (w < 1.e-4f) continue;
        const uint32_t j = knn[(uint64_t)K * ci + s];
        float v[6];
        j_depth_value(nn, w, p, wD, v);
        float* o = J + 6ull * ((uint64_t)c * jW + j);
#pragma unroll
        for (int k = 0; k < 6; ++k) o[k] = v[k];
    }
}

// fill_J_MRegKernel (src/MotionOptimization.cu:80-145) into the dense J.
__global__ void k_j_mreg(float* __restrict__ J, float wR, uint32_t jW, const float* __restrict__ graph, uint32_t G,
                         const uint32_t* __restrict__ gknn, int R) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= G * (uint32_t)R) return;
    const uint32_t g = t / (uint32_t)R;
    const uint32_t nb = gknn[t];
    if (nb >= G) return;
    float p[3];
    load3(graph + 3ull * g, p);
    for (int a = 0; a < 3; ++a) {
        float gr[6];
        grad_row(p, a, gr);
        float* jc = J + 6ull * ((3ull * t + a) * jW + g);
        float* jn = J + 6ull * ((3ull * t + a) * jW + nb);
        for (int k = 0; k < 6; ++k) jc[k] = gr[k] * wR;
        for (int k = 0; k < 6; ++k) jn[k] = -(gr[k] * wR);
    }
}

// updateTransformsKernel (src/MotionOptimization.cu:481-498): T <- T - U.
__global__ void k_update(float* __restrict__ T, const float* __restrict__ u, uint32_t G) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G) return;
    const float* x = u + 6ull * g;
    float* M = T + 16ull * g;
    // column-major M(i, j) = M[4 j + i]
    M[4 * 1 + 0] = M[4 * 1 + 0] - (-x[2]);
    M[4 * 2 + 0] = M[4 * 2 + 0] - x[1];
    M[4 * 3 + 0] = M[4 * 3 + 0] - x[3];
    M[4 * 0 + 1] = M[4 * 0 + 1] - x[2];
    M[4 * 2 + 1] = M[4 * 2 + 1] - (-x[0]);
    M[4 * 3 + 1] = M[4 * 3 + 1] - x[4];
    M[4 * 0 + 2] = M[4 * 0 + 2] - (-x[1]);
    M[4 * 1 + 2] = M[4 * 1 + 2] - x[0];
    M[4 * 3 + 2] = M[4 * 3 + 2] - x[5];
    // the zero entries of U: T - 0 is T except for -0 (-0 - 0 = -0): unchanged
}

// ---- normal equations from the sparse rows --------------------------------------
//
// Row r < C is a depth row: its members are the knn slots of mesh point
// corr[r].x with a non-zero weight (the pairs of
// computeJTJAndJTrContribJDepthKernel, src/MotionOptimization.cu:227-259),
// each with the Jacobian entry fill_J_DepthKernel writes (zero below weight
// 1e-4).  Row r >= C is regularisation row 3 (g R + i) + a: members g and its
// neighbour nb (computeJTJAndJTrContribJMRegKernel, :284-314); when nb == g the
// three pairs coincide (the diagonal block takes the product three times, JTr
// the entry twice, with the entry of the last write, -grad).

struct GnRows {
    const int32_t* corr;
    uint32_t C;
    const float* canon;
    int K;
    const uint32_t* knn;
    const float* kw;
    const float* dnormals;
    const float* graph;
    uint32_t G;
    const uint32_t* gknn;
    int R;
    float wD, wR;
};

struct Member {
    uint32_t node;
    int diag;  // multiplicity on the diagonal block
    int jtr;   // multiplicity in JTr
    float j[6];
};

// Members of row `row` (up to kGnMaxK); returns their count.
__device__ int row_members(const GnRows& g, uint32_t row, Member* m) {
    if (row < g.C) {
        const uint32_t ci = (uint32_t)g.corr[2 * row], di = (uint32_t)g.corr[2 * row + 1];
        float p[3], nn[3];
        load3(g.canon + 3ull * ci, p);
        load3(g.dnormals + 3ull * di, nn);
        int n = 0;
        for (int s = 0; s < g.K; ++s) {
            const float w = g.kw[(uint64_t)g.K * ci + s];
            if (w == 0.0f) continue;
            Member& e = m[n++];
            e.node = g.knn[(uint64_t)g.K * ci + s];
            e.diag = e.jtr = 1;
            if (w < 1.e-4f) {
#pragma unroll
                for (int k = 0; k < 6; ++k) e.j[k] = 0.0f;
            } else {
                j_depth_value(nn, w, p, g.wD, e.j);
            }
        }
        return n;
    }
    const uint32_t mr = (row - g.C) / 3u, a = (row - g.C) % 3u;
    const uint32_t node = mr / (uint32_t)g.R;
    const uint32_t nb = g.gknn[mr];
    if (nb >= g.G) return 0;
    float p[3], gr[6];
    load3(g.graph + 3ull * node, p);
    grad_row(p, (int)a, gr);
    if (nb == node) {
        m[0].node = node;
        m[0].diag = 3;
        m[0].jtr = 2;
#pragma unroll
        for (int k = 0; k < 6; ++k) m[0].j[k] = -(gr[k] * g.wR);
        return 1;
    }
    m[0].node = node;
    m[1].node = nb;
    m[0].diag = m[0].jtr = m[1].diag = m[1].jtr = 1;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        m[0].j[k] = gr[k] * g.wR;
        m[1].j[k] = -(gr[k] * g.wR);
    }
    return 2;
}

// Counts each node's rows; flags a member node outside the graph.
__global__ void k_rows_count(GnRows g, uint32_t* __restrict__ cnt, uint32_t* __restrict__ bad) {
    const uint32_t rows = g.C + 3u * g.G * (uint32_t)g.R;
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= rows) return;
    if (row < g.C) {
        const uint32_t ci = (uint32_t)g.corr[2 * row];
        for (int s = 0; s < g.K; ++s) {
            if (g.kw[(uint64_t)g.K * ci + s] == 0.0f) continue;
            const uint32_t j = g.knn[(uint64_t)g.K * ci + s];
            if (j >= g.G) {
                *bad = 1u;
                continue;
            }
            atomicAdd(cnt + j, 1u);
        }
        return;
    }
    const uint32_t mr = (row - g.C) / 3u;
    const uint32_t node = mr / (uint32_t)g.R, nb = g.gknn[mr];
    if (nb >= g.G) return;
    atomicAdd(cnt + node, 1u);
    if (nb != node) atomicAdd(cnt + nb, 1u);
}

__global__ void k_rows_fill(GnRows g, const uint32_t* __restrict__ off, uint32_t* __restrict__ fill,
                            uint32_t* __restrict__ rows_of) {
    const uint32_t rows = g.C + 3u * g.G * (uint32_t)g.R;
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= rows) return;
    if (row < g.C) {
        const uint32_t ci = (uint32_t)g.corr[2 * row];
        for (int s = 0; s < g.K; ++s) {
            if (g.kw[(uint64_t)g.K * ci + s] == 0.0f) continue;
            const uint32_t j = g.knn[(uint64_t)g.K * ci + s];
            if (j >= g.G) continue;
            rows_of[off[j] + atomicAdd(fill + j, 1u)] = row;
        }
        return;
    }
    const uint32_t mr = (row - g.C) / 3u;
    const uint32_t node = mr / (uint32_t)g.R, nb = g.gknn[mr];
    if (nb >= g.G) return;
    rows_of[off[node] + atomicAdd(fill + node, 1u)] = row;
    if (nb != node) rows_of[off[nb] + atomicAdd(fill + nb, 1u)] = row;
}

// Sorts each node's row list (bitonic sort in LDS, one workgroup per node).
__global__ __launch_bounds__(1024) void k_rows_sort(const uint32_t* __restrict__ off, uint32_t* __restrict__ rows_of,
                                                    uint32_t* __restrict__ bad) {
    __shared__ uint32_t s[kGnMaxRows];
    const uint32_t node = blockIdx.x;
    const uint32_t b = off[node], n = off[node + 1] - b;
    if (n > (uint32_t)kGnMaxRows) {
        if (threadIdx.x == 0) *bad = 2u;
        return;
    }
    uint32_t P = 1;
    while (P < n) P <<= 1;
    for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) s[i] = i < n ? rows_of[b + i] : kGnNone;
    __syncthreads();
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const uint32_t x = s[i], y = s[l];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) {
                        s[i] = y;
                        s[l] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) rows_of[b + i] = s[i];
}

// Zero, then the identity diagonal (runInitIdentityKernel before the fill).
__global__ void k_identity(float* __restrict__ A, uint64_t n) {
    const uint64_t total = n * n;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x)
        A[i] = (i / n == i % n) ? 1.0f : 0.0f;
}

// The members of every row, computed once per ICP iteration (one thread per
// row, row_members): slot q < 8 of row r holds node | diag << 24 | jtr << 28
// in node8[8 r + q] (kGnNone past the row's members) and its Jacobian
// entries in j48[48 r + 6 q + k].  A member outside the graph (a knn index
// >= G with a non-zero weight) is dropped and flagged.
__global__ void k_row_data(GnRows g, uint32_t rows, uint32_t* __restrict__ node8, float* __restrict__ j48,
                           uint32_t* __restrict__ bad) {
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= rows) return;
    Member m[kGnMaxK];
    const int n0 = row_members(g, row, m);
    int n = 0;
    for (int q = 0; q < n0; ++q) {
        if (m[q].node >= g.G) {
            *bad = 1u;
            continue;
        }
        m[n++] = m[q];
    }
#pragma unroll
    for (int q = 0; q < kGnMaxK; ++q) {
        uint32_t v = kGnNone;
        float j[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        if (q < n) {
            v = m[q].node | ((uint32_t)m[q].diag << 24) | ((uint32_t)m[q].jtr << 28);
#pragma unroll
            for (int k = 0; k < 6; ++k) j[k] = m[q].j[k];
        }
        node8[8ull * row + q] = v;
#pragma unroll
        for (int k = 0; k < 6; ++k) j48[48ull * row + 6 * q + k] = j[k];
    }
}

// One wave per node i: its 6x6 blocks JTJ(i, j) for every node j sharing a
// row with it, summed over i's rows in ascending order (each accumulator
// starts at -0, the additive identity, so the first term is taken as is),
// and, when r is given, its JTr entry.  Lanes 0..35 own the 36 elements of a
// block; lanes 0..5 the JTr entry.  Each row's members come from the row
// data (node8, j48), the next row's loaded while the current one is summed;
// the row indices come 64 at a time, one per lane.
__global__ __launch_bounds__(64) void k_blocks(uint32_t G, const uint32_t* __restrict__ off,
                                               const uint32_t* __restrict__ rows_of,
                                               const uint32_t* __restrict__ node8, const float* __restrict__ j48,
                                               const float* __restrict__ r, float* __restrict__ JTJ,
                                               float* __restrict__ JTr, uint32_t* __restrict__ bad) {
    __shared__ uint32_t s_ids[kGnMaxNbr];
    __shared__ float s_acc[kGnMaxNbr][36];
    __shared__ float s_j[2][48];
    const uint32_t i = blockIdx.x;
    const int lane = threadIdx.x;
    const uint32_t b = off[i], n = off[i + 1] - b;
    // row e of i's list: rows_of[b + e], held by lane e % 64 of `rowreg`
    // while chunk e / 64 is loaded
    uint32_t rowreg = 0u, chunk = 0xffffffffu;
    const auto row_at = [&](uint32_t e) -> uint32_t {
        if ((e >> 6) != chunk) {
            chunk = e >> 6;
            const uint32_t k = (chunk << 6) + (uint32_t)lane;
            rowreg = k < n ? rows_of[b + k] : 0u;
        }
        return (uint32_t)__shfl((int)rowreg, (int)(e & 63u));
    };
    int nids = 0;
    if (JTJ) {
        // distinct nodes of i's rows, in first-seen order
        uint32_t pn = kGnNone;
        if (n > 0 && lane < 8) pn = node8[8ull * row_at(0) + lane];
        for (uint32_t e = 0; e < n; ++e) {
            const uint32_t cn = pn;
            if (e + 1 < n) {
                const uint32_t r1 = row_at(e + 1);
                pn = lane < 8 ? node8[8ull * r1 + lane] : kGnNone;
            }
            for (int q = 0; q < 8; ++q) {
                const uint32_t mq = (uint32_t)__shfl((int)cn, q);
                if (mq == kGnNone) break;
                const uint32_t j = mq & 0xffffffu;
                const bool seen = __ballot((lane < nids && s_ids[lane] == j) ||
                                           (lane + 64 < nids && s_ids[lane + 64] == j)) != 0;
                if (!seen) {
                    if (nids >= kGnMaxNbr) {
                        if (lane == 0) *bad = 3u;
                        return;
                    }
                    if (lane == 0) s_ids[nids] = j;
                    ++nids;
                }
            }
        }
        for (int t = 0; t < nids; ++t)
            if (lane < 36) s_acc[t][lane] = -0.0f;
    }
    float jtr = -0.0f;
    uint32_t pn = kGnNone;
    float pj = 0.0f, pr = 0.0f;
    if (n > 0) {
        const uint32_t r0 = row_at(0);
        if (lane < 8) pn = node8[8ull * r0 + lane];
        if (lane < 48) pj = j48[48ull * r0 + lane];
        if (r) pr = r[r0];
    }
    for (uint32_t e = 0; e < n; ++e) {
        const uint32_t cn = pn;
        const float cr = pr;
        float* sj = s_j[e & 1u];
        if (lane < 48) sj[lane] = pj;
        if (e + 1 < n) {
            const uint32_t r1 = row_at(e + 1);
            if (lane < 8) pn = node8[8ull * r1 + lane];
            if (lane < 48) pj = j48[48ull * r1 + lane];
            if (r) pr = r[r1];
        }
        // this node's member: the last slot naming i
        const unsigned long long mine = __ballot(lane < 8 && cn != kGnNone && (cn & 0xffffffu) == i);
        if (!mine) {  // cannot happen: the row is in i's list because i is a member
            if (lane == 0) *bad = 4u;
            return;
        }
        const int qi = 63 - __clzll((long long)mine);
        const uint32_t mi = (uint32_t)__shfl((int)cn, qi);
        if (JTr && lane < 6) {
            const int mult = (int)(mi >> 28);
            const float v = sj[6 * qi + lane] * cr;
            for (int t = 0; t < mult; ++t) jtr = jtr + v;
        }
        if (JTJ) {
            for (int q = 0; q < 8; ++q) {
                const uint32_t mq = (uint32_t)__shfl((int)cn, q);
                if (mq == kGnNone) break;
                const uint32_t node = mq & 0xffffffu;
                // the slot of node: lanes test two list entries each
                const unsigned long long h0 = __ballot(lane < nids && s_ids[lane] == node);
                const unsigned long long h1 = __ballot(lane + 64 < nids && s_ids[lane + 64] == node);
                const int slot = h0 ? __ffsll((long long)h0) - 1 : 64 + __ffsll((long long)h1) - 1;
                const int mult = node == i ? (int)((mi >> 24) & 15u) : 1;
                if (lane < 36) {
                    const float v = sj[6 * qi + lane / 6] * sj[6 * q + lane % 6];
                    for (int t = 0; t < mult; ++t) s_acc[slot][lane] = s_acc[slot][lane] + v;
                }
            }
        }
    }
    if (JTr && lane < 6) JTr[6ull * i + lane] = n ? jtr : 0.0f;
    if (JTJ && lane < 36) {
        const uint64_t W = 6ull * G;
        const int a = lane / 6, c = lane % 6;
        for (int t = 0; t < nids; ++t)
            JTJ[(6ull * i + a) * W + 6ull * s_ids[t] + c] = s_acc[t][lane];
    }
}

// Sum of squares in binary64, fixed order: each workgroup sums a contiguous
// range, then one thread adds the workgroups' sums.
constexpr int kEnergyBlocks = 256;
__global__ __launch_bounds__(256) void k_energy_part(const float* __restrict__ r, uint64_t n,
                                                     double* __restrict__ part) {
    __shared__ double s[256];
    const uint64_t per = (n + kEnergyBlocks - 1) / kEnergyBlocks;
    const uint64_t b = blockIdx.x * per, e = b + per < n ? b + per : n;
    double t = 0.0;
    for (uint64_t i = b + threadIdx.x; i < e; i += 256) t += (double)r[i] * (double)r[i];
    s[threadIdx.x] = t;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void k_energy_sum(const double* __restrict__ part, double* __restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < kEnergyBlocks; ++i) t += part[i];
        *out = t;
    }
}

// ---- the Cholesky factorisation of JTJ -------------------------------------------
//
// A right-looking blocked L L^T (lower, column-major) in blocks of
// kCholOuter columns.  Step k factorises the diagonal block k (two halves of
// at most kCholBlock columns by k_chol_block, with a trsm and a syrk
// between them), solves the
// rows below it against it (rocBLAS trsm), updates the next block column
// (syrk on its diagonal block, gemm below it), and then, while the rest of
// the trailing matrix is updated (syrk) on the main stream, factorises the
// next diagonal block on a second stream (look-ahead).  rocBLAS keeps the
// O(n^3) updates on the whole chip; the diagonal blocks, whose factorisation
// is a chain of small dependent steps that rocSOLVER's potrf runs as many
// tiny launches, take one workgroup each (k_chol_block).

constexpr int kCholOuter = 512;                   // outer block (rank of the trailing updates)
constexpr int kCholBlock = 256;                   // largest block of k_chol_block
constexpr int kCholPanel = 32;                    // panel of k_chol_block
constexpr int kCholThreads = 512;                 // k_chol_block workgroup
constexpr int kCholRows = kCholBlock - kCholPanel;  // panel rows after the first panel
constexpr int kCholLd = kCholPanel + 4;           // row stride of the panel factor in LDS

__device__ __forceinline__ float readlane_f(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Column J of the Cholesky factorisation of a pw x pw block (pw <= kCholPanel)
// held by one wave, lane i holding row i in a[] (one instantiation per
// column, so every index into a[] is a constant).  Returns the first column
// whose pivot is not positive (or NaN), else -1.  Lanes at or beyond the
// block hold zeros; entries above the diagonal are left as garbage.
template <int J>
__device__ __forceinline__ int chol_col(float (&a)[kCholPanel], int pw, int lane) {
    if constexpr (J == kCholPanel) {
        return -1;
    } else {
        if (J >= pw) return -1;
        const float s = readlane_f(a[J], J);
        if (!(s > 0.0f)) return J;
        const float d = sqrtf(s);
        const float l = a[J] / d;
        a[J] = lane == J ? d : l;
#pragma unroll
        for (int k = J + 1; k < kCholPanel; ++k) a[k] = __builtin_fmaf(-l, readlane_f(l, k), a[k]);
        return chol_col<J + 1>(a, pw, lane);
    }
}

// L L^T of the kb x kb diagonal block at A (kb <= kCholBlock, leading
// dimension lda), in panels of kCholPanel columns, left-looking: each panel
// subtracts the products of the columns already factorised (staged in LDS),
// one wave factorises its diagonal part in registers, and the rows below are
// solved against it by substitution, one thread per row.  A pivot that is
// not positive stores col0 + its column + 1 in *info (if still 0, as potrf's
// info) and ends the kernel; a kernel that finds *info set does nothing.
__global__ __launch_bounds__(kCholThreads) void k_chol_block(float* __restrict__ A, int lda, int kb, int col0,
                                                             int* __restrict__ info) {
    __shared__ float s_w[32 * kCholPanel];  // L rows of the panel's columns, t-major
    __shared__ float s_r[32 * kCholRows];   // L rows of the panel's rows, t-major
    __shared__ float s_l[kCholPanel * kCholLd];
    __shared__ int s_fail;
    const int tid = threadIdx.x;
    if (*(volatile int*)info != 0) return;  // the same value for every thread
    if (tid == 0) s_fail = -1;
    for (int c0 = 0; c0 < kb; c0 += kCholPanel) {
        const int pw = min(kCholPanel, kb - c0);
        const int rows = kb - c0;  // <= kCholRows after the first panel
        float* P = A + c0 + (size_t)c0 * lda;
        if (c0 > 0) {
            // P[r][c] -= sum_{t < c0} L[c0 + r][t] L[c0 + c][t], in chunks of 32 t
            constexpr int kH = kCholPanel * 256 / kCholThreads;  // columns per thread
            constexpr int kTS = kCholThreads / 256, kTN = 32 / kTS;
            const int rr = tid & 255, cg = tid >> 8;
            float acc[kH];
#pragma unroll
            for (int q = 0; q < kH; ++q) acc[q] = 0.0f;
            for (int t0 = 0; t0 < c0; t0 += 32) {
                for (int e = tid; e < 32 * kCholPanel; e += kCholThreads) {
                    const int c = e % kCholPanel, t = e / kCholPanel;
                    s_w[e] = c < pw ? A[(size_t)(c0 + c) + (size_t)(t0 + t) * lda] : 0.0f;
                }
                float v[kTN];
#pragma unroll
                for (int j = 0; j < kTN; ++j)
                    v[j] = rr < rows ? A[(size_t)(c0 + rr) + (size_t)(t0 + cg + kTS * j) * lda] : 0.0f;
                if (rr < kCholRows) {
#pragma unroll
                    for (int j = 0; j < kTN; ++j) s_r[(cg + kTS * j) * kCholRows + rr] = v[j];
                }
                __syncthreads();
                if (rr < rows) {
#pragma unroll 4
                    for (int t = 0; t < 32; ++t) {
                        const float l0 = s_r[t * kCholRows + rr];
                        const float4* w4 = (const float4*)(s_w + t * kCholPanel + cg * kH);
#pragma unroll
                        for (int q = 0; q < kH / 4; ++q) {
                            const float4 w = w4[q];
                            acc[4 * q + 0] = __builtin_fmaf(l0, w.x, acc[4 * q + 0]);
                            acc[4 * q + 1] = __builtin_fmaf(l0, w.y, acc[4 * q + 1]);
                            acc[4 * q + 2] = __builtin_fmaf(l0, w.z, acc[4 * q + 2]);
                            acc[4 * q + 3] = __builtin_fmaf(l0, w.w, acc[4 * q + 3]);
                        }
                    }
                }
                __syncthreads();
            }
#pragma unroll
            for (int q = 0; q < kH; ++q) {
                const int c = cg * kH + q;
                if (c < pw && rr < rows && rr >= c) P[rr + (size_t)c * lda] -= acc[q];
            }
            __syncthreads();
        }
        if (tid < 64) {
            const int i = tid;
            float a[kCholPanel];
#pragma unroll
            for (int k = 0; k < kCholPanel; ++k) a[k] = (i < pw && k <= i) ? P[i + (size_t)k * lda] : 0.0f;
            const int f = chol_col<0>(a, pw, i);
            if (f >= 0) {
                if (i == 0) s_fail = f;
            } else if (i < kCholPanel) {
#pragma unroll
                for (int k = 0; k < kCholPanel; ++k) {
                    const bool low = i < pw && k <= i;
                    if (low) P[i + (size_t)k * lda] = a[k];
                    s_l[i * kCholLd + k] = low ? a[k] : 0.0f;
                }
            }
        }
        __syncthreads();
        if (s_fail >= 0) {
            if (tid == 0) atomicCAS(info, 0, col0 + c0 + s_fail + 1);
            return;
        }
        // rows below the panel (pw == kCholPanel whenever there are any):
        // x L11^T = b by forward substitution
        const int nbelow = rows - pw;
        if (tid < nbelow) {
            float* B = P + pw + tid;
            float x[kCholPanel];
#pragma unroll
            for (int c = 0; c < kCholPanel; ++c) {
                float v = B[(size_t)c * lda];
#pragma unroll
                for (int t = 0; t < c; ++t) v = __builtin_fmaf(-x[t], s_l[c * kCholLd + t], v);
                x[c] = v / s_l[c * kCholLd + c];
            }
#pragma unroll
            for (int c = 0; c < kCholPanel; ++c) B[(size_t)c * lda] = x[c];
        }
        __syncthreads();
    }
}

// ---- the solve with the factorised JTJ ---------------------------------------
//
// getrs (src/LinearSolver.cu:179-225 calls it through cuSOLVER) as blocked
// substitutions for the solver loop, which solves with the same factors on
// every Gauss-Newton iteration: the row interchanges as one gather (the
// permutation composed once from ipiv), then L y = P b and U x = y in blocks
// of 64 rows.  One launch per triangle: one wave per row block, the blocks
// handed out in the order the waves start (a ticket), so a wave only ever
// waits for blocks taken by waves already running.  Block b's wave loads
// each block column of the factor it needs before that block's solution is
// published, then subtracts the factor block times the solved block (in
// block order), and finally solves its 64 x 64 diagonal block by
// substitution in registers and publishes its solution (each value tagged
// with the solve's epoch in one 64-bit atomic store).  The operations and
// their order are those of a step-by-step blocked substitution (one launch
// per block step, which the step-by-step version took); only the waiting is
// per block instead of per launch.
// Rounding differs from rocSOLVER's getrs (another summation order), not the
// method.
constexpr int kTB = 64;
constexpr uint32_t kChainSpin = 1u << 22;  // polls of a block's flag before the solve gives up

// Lane j's value of x, broadcast (j a constant of the unrolled loops).
__device__ __forceinline__ float lane_value(float x, int j) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), j));
}

__global__ __launch_bounds__(256) void k_gather_perm(const float* __restrict__ b, const int* __restrict__ perm,
                                                     float* __restrict__ y, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = b[perm[i]];
}

// sync[0]: the ticket counter; sync[1]: a wait gave up (sticky until the
// caller clears it).  tagged[i]: (epoch << 32) | bits of the solved value i,
// written by one 64-bit atomic, so a reader that sees this solve's epoch sees
// the value with it; no fence (an agent-scope fence writes back or
// invalidates the whole L2 of the XCD) and no flags to clear between solves.

// This wave's block: the ticket, in start order (one wave per workgroup).
__device__ __forceinline__ int chain_ticket(int* sync) {
    int t = 0;
    if (threadIdx.x == 0) t = atomicAdd(&sync[0], 1);
    return __builtin_amdgcn_readfirstlane(t);
}

// Solved value i of this solve (lanes with need; others get 0): polls the
// tagged word until it carries the epoch, bounded; on giving up, ok turns
// false and the value is what memory holds, so that every wave still ends.
__device__ __forceinline__ float chain_value(const unsigned long long* __restrict__ tagged, int i, bool need,
                                             uint32_t epoch, bool& ok) {
    unsigned long long v = 0;
    if (need) {
        uint32_t spin = 0;
        for (;;) {
            v = __hip_atomic_load(tagged + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((uint32_t)(v >> 32) == epoch) break;
            if (++spin > kChainSpin) {
                ok = false;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    return __uint_as_float((uint32_t)v);
}

// Stores this block's solution: the plain value for the kernels after this
// launch, the tagged one for the waves of this launch.
__device__ __forceinline__ void chain_publish(float* v, unsigned long long* __restrict__ tagged, int i, bool row,
                                              float r, uint32_t epoch, int* sync, bool ok) {
    if (row) {
        v[i] = r;
        __hip_atomic_store(tagged + i, ((unsigned long long)epoch << 32) | __float_as_uint(r), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    if (__ballot(!ok) != 0ull && threadIdx.x == 0) atomicOr(&sync[1], 1);
}

// L y = y in place.  LU is column-major with leading dimension n; L has a unit
// diagonal (UNIT, the LU factors) or its own (the Cholesky factor, divisions
// by it).  Grid: one 64-thread workgroup per row block.
template <bool UNIT>
__global__ __launch_bounds__(64) void k_trsv_lower_chain(const float* __restrict__ LU, int n, float* __restrict__ y,
                                                         unsigned long long* __restrict__ tagged, uint32_t epoch,
                                                         int* __restrict__ sync) {
    const int lane = threadIdx.x;
    const int b = chain_ticket(sync);
    const int i = b * kTB + lane;
    const bool row = i < n;
    float r = row ? y[i] : 0.0f;
    // the diagonal block first: its loads are then off the chain of waits
    const int cb = b * kTB;
    const int mb = n - cb < kTB ? n - cb : kTB;  // rows of this block
    float d[kTB];
#pragma unroll
    for (int j = 0; j < kTB; ++j) d[j] = (row && j < mb) ? LU[(uint64_t)(cb + j) * (uint64_t)n + (uint64_t)i] : 0.0f;
    bool ok = true;
    for (int k = 0; k < b; ++k) {
        const int c0 = k * kTB;  // block k is full: k < b <= last block
        float a[kTB];
#pragma unroll
        for (int j = 0; j < kTB; ++j) a[j] = row ? LU[(uint64_t)(c0 + j) * (uint64_t)n + (uint64_t)i] : 0.0f;
        const float yk = chain_value(tagged, c0 + lane, ok, epoch, ok);
        float acc = 0.0f;
#pragma unroll
        for (int j = 0; j < kTB; ++j) acc = acc + a[j] * lane_value(yk, j);
        r = r - acc;
    }
#pragma unroll
    for (int j = 0; j < kTB; ++j) {
        // final: lane j has subtracted columns 0..j-1 (and, with its own
        // diagonal, divides by it: d[j] of lane j is L(j, j))
        if (!UNIT && lane == j && j < mb) r = r / d[j];
        const float yj = lane_value(r, j);
        if (lane > j) r = r - d[j] * yj;
    }
    chain_publish(y, tagged, i, row, r, epoch, sync, ok);
}

// U x = x in place, blocks from the last: U is the upper triangle of LU, or
// (TRANS) the transpose of the Cholesky factor L in LU's lower triangle:
// U(i, c) = L(c, i), which lane i reads from its column of L.
template <bool TRANS>
__global__ __launch_bounds__(64) void k_trsv_upper_chain(const float* __restrict__ LU, int n, float* __restrict__ x,
                                                         unsigned long long* __restrict__ tagged, uint32_t epoch,
                                                         int* __restrict__ sync) {
    const auto u_at = [&](int i, int c) -> float {
        return TRANS ? LU[(uint64_t)i * (uint64_t)n + (uint64_t)c] : LU[(uint64_t)c * (uint64_t)n + (uint64_t)i];
    };
    const int lane = threadIdx.x;
    const int nb = (n + kTB - 1) / kTB;
    const int b = nb - 1 - chain_ticket(sync);
    const int i = b * kTB + lane;
    const bool row = i < n;
    float r = row ? x[i] : 0.0f;
    // the diagonal block first: its loads are then off the chain of waits
    const int cb = b * kTB;
    const int mb = n - cb < kTB ? n - cb : kTB;
    float d[kTB];
#pragma unroll
    for (int j = 0; j < kTB; ++j) d[j] = (row && j < mb) ? u_at(i, cb + j) : 0.0f;
    bool ok = true;
    // (TRANS) a lane's entries of a block are 64 consecutive floats of its
    // column of L: 16-byte loads where the column starts 16-byte aligned
    const bool vec = TRANS && (n & 3) == 0;
    for (int k = nb - 1; k > b; --k) {
        const int c0 = k * kTB;
        const int m = n - c0 < kTB ? n - c0 : kTB;
        float a[kTB];
        if (vec && m == kTB) {
            const float4* col = (const float4*)(LU + (uint64_t)(row ? i : 0) * (uint64_t)n + (uint64_t)c0);
#pragma unroll
            for (int j = 0; j < kTB / 4; ++j) {
                const float4 v = row ? col[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                a[4 * j] = v.x;
                a[4 * j + 1] = v.y;
                a[4 * j + 2] = v.z;
                a[4 * j + 3] = v.w;
            }
        } else {
#pragma unroll
            for (int j = 0; j < kTB; ++j) a[j] = (row && j < m) ? u_at(i, c0 + j) : 0.0f;
        }
        const float xk = chain_value(tagged, c0 + lane, ok && lane < m, epoch, ok);
        float acc = 0.0f;
#pragma unroll
        for (int j = 0; j < kTB; ++j) acc = acc + a[j] * lane_value(xk, j);
        r = r - acc;
    }
#pragma unroll
    for (int j = kTB - 1; j >= 0; --j) {
        if (j < mb) {
            if (lane == j) r = r / d[j];  // d[j] of lane j is the diagonal U(j, j)
            const float xj = lane_value(r, j);
            if (lane < j) r = r - d[j] * xj;
        }
    }
    chain_publish(x, tagged, i, row, r, epoch, sync, ok);
}

}  // namespace
}  // namespace vf

// ---- host side -------------------------------------------------------------------

struct vf_gn_solver {
    int device = 0;
    vf::GnRows rows;
    uint32_t nrows = 0;
    uint64_t n = 0;  // 6 G
    // device buffers
    uint32_t* cnt = nullptr;   // G + 1: counts, then offsets
    uint32_t* fill = nullptr;  // G
    uint32_t* rows_of = nullptr;
    uint32_t* node8 = nullptr;  // row data: members of each row
    float* j48 = nullptr;       // row data: their Jacobian entries
    uint32_t* bad = nullptr;
    float* JTJ = nullptr;  // the factors after the factorisation: Cholesky (lower) or LU
    bool chol = false;     // JTJ holds the Cholesky factor
    int* ipiv = nullptr;
    int* info = nullptr;
    float* r = nullptr;    // residuals, nrows
    float* x = nullptr;    // JTr
    float* y = nullptr;    // the permuted JTr, solved in place into the update
    int* perm = nullptr;   // row interchanges of the factorisation as one permutation
    int* chain = nullptr;  // the substitutions' ticket and failure flag
    unsigned long long* tagged = nullptr;  // the substitutions' solved values with their solve's epoch
    uint32_t epoch = 0;                    // solves launched (the tag of the next one is epoch + 1)
    double* part = nullptr;
    double* energy = nullptr;
    rocblas_handle blas = nullptr;
    rocblas_handle blas_aux = nullptr;  // rocBLAS on the look-ahead stream
    hipStream_t aux = nullptr;          // the Cholesky look-ahead's stream
    hipEvent_t ev_col = nullptr;        // next block column updated (main stream)
    hipEvent_t ev_diag = nullptr;       // next diagonal block factorised (aux)
};

namespace vf {
namespace {

void gn_free(vf_gn_solver* s) {
    void* ptrs[] = {s->cnt, s->fill, s->rows_of, s->node8, s->j48,  s->bad,  s->JTJ,   s->ipiv,
                    s->info, s->r,   s->x,       s->y,     s->perm, s->part, s->energy, s->chain, s->tagged};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (s->blas) (void)rocblas_destroy_handle(s->blas);
    if (s->blas_aux) (void)rocblas_destroy_handle(s->blas_aux);
    if (s->ev_col) (void)hipEventDestroy(s->ev_col);
    if (s->ev_diag) (void)hipEventDestroy(s->ev_diag);
    if (s->aux) (void)hipStreamDestroy(s->aux);
}

template <typename T>
int gn_alloc(T** p, uint64_t count) {
    const hipError_t e = hipMalloc((void**)p, std::max<uint64_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(e == hipErrorOutOfMemory ? VF_ERR_OUT_OF_MEMORY : VF_ERR_HIP, "gn allocation of %llu B: %s",
                    (unsigned long long)(count * sizeof(T)), hipGetErrorString(e));
    }
    return VF_OK;
}

int check_problem(const vf_gn_problem* p) {
    if (!p) return fail(VF_ERR_INVALID_ARGUMENT, "vf_gn: NULL problem");
    if (p->k < 1 || p->k > kGnMaxK) return fail(VF_ERR_UNSUPPORTED_K, "gn: k=%d, supported 1..%d", p->k, kGnMaxK);
    if (p->reg_k < 1 || p->reg_k > kGnMaxR)
        return fail(VF_ERR_INVALID_ARGUMENT, "gn: reg_k=%d, supported 1..%d", p->reg_k, kGnMaxR);
    if (p->graph_size < 1 || p->graph_size > (1u << 20))
        return fail(VF_ERR_INVALID_ARGUMENT, "gn: graph of %u nodes, supported 1..2^20", p->graph_size);
    if (!p->graph || !p->graph_knn || (p->corr_size && (!p->corr || !p->canon || !p->knn_idxs || !p->knn_weights ||
                                                         !p->normals)))
        return fail(VF_ERR_INVALID_ARGUMENT, "gn: NULL problem buffer");
    const uint64_t rows = (uint64_t)p->corr_size + 3ull * p->graph_size * (uint64_t)p->reg_k;
    if (rows > 0x7fffffffull) return fail(VF_ERR_CAPACITY, "gn: %llu rows exceed 2^31", (unsigned long long)rows);
    const uint64_t n = 6ull * p->graph_size;
    if (n * n > (1ull << 34))
        return fail(VF_ERR_CAPACITY, "gn: the dense normal matrix of %u nodes exceeds 64 GiB", p->graph_size);
    return VF_OK;
}

GnRows make_rows(const vf_gn_problem* p) {
    GnRows g;
    g.corr = (const int32_t*)p->corr;
    g.C = p->corr_size;
    g.canon = (const float*)p->canon;
    g.K = p->k;
    g.knn = p->knn_idxs;
    g.kw = p->knn_weights;
    g.dnormals = (const float*)p->normals;
    g.graph = (const float*)p->graph;
    g.G = p->graph_size;
    g.gknn = p->graph_knn;
    g.R = p->reg_k;
    g.wD = p->weight_depth;
    g.wR = p->weight_reg;
    return g;
}

// Builds each node's sorted row list; returns the host-read status.
int build_rows(vf_gn_solver* s, hipStream_t st) {
    const GnRows& g = s->rows;
    VF_HIP(hipMemsetAsync(s->cnt, 0, sizeof(uint32_t) * (g.G + 1ull), st));
    VF_HIP(hipMemsetAsync(s->fill, 0, sizeof(uint32_t) * g.G, st));
    VF_HIP(hipMemsetAsync(s->bad, 0, sizeof(uint32_t), st));
    hipLaunchKernelGGL(k_rows_count, dim3(blocks_for(s->nrows, 256)), dim3(256), 0, st, g, s->cnt, s->bad);
    VF_LAUNCH_CHECK("k_rows_count");
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, s->cnt, g.G + 1u, (uint32_t*)nullptr);
    VF_LAUNCH_CHECK("k_scan_counts");
    uint32_t total = 0;
    VF_HIP(hipMemcpyAsync(&total, s->cnt + g.G, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    VF_HIP(hipStreamSynchronize(st));
    VF_HIP(hipFree(s->rows_of));
    s->rows_of = nullptr;
    int e = gn_alloc(&s->rows_of, total);
    if (e) return e;
    hipLaunchKernelGGL(k_rows_fill, dim3(blocks_for(s->nrows, 256)), dim3(256), 0, st, g, (const uint32_t*)s->cnt,
                       s->fill, s->rows_of);
    VF_LAUNCH_CHECK("k_rows_fill");
    hipLaunchKernelGGL(k_rows_sort, dim3(g.G), dim3(1024), 0, st, (const uint32_t*)s->cnt, s->rows_of, s->bad);
    VF_LAUNCH_CHECK("k_rows_sort");
    if (!s->node8) {
        e = gn_alloc(&s->node8, 8ull * s->nrows);
        if (!e) e = gn_alloc(&s->j48, 48ull * s->nrows);
        if (e) return e;
    }
    hipLaunchKernelGGL(k_row_data, dim3(blocks_for(s->nrows, 256)), dim3(256), 0, st, g, s->nrows, s->node8, s->j48,
                       s->bad);
    VF_LAUNCH_CHECK("k_row_data");
    return VF_OK;
}

int read_bad(vf_gn_solver* s, hipStream_t st) {
    uint32_t bad = 0;
    VF_HIP(hipMemcpyAsync(&bad, s->bad, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    VF_HIP(hipStreamSynchronize(st));
    switch (bad) {
        case 0: return VF_OK;
        case 1: return fail(VF_ERR_INVALID_ARGUMENT, "gn: a mesh point's knn names a node outside the graph");
        case 2: return fail(VF_ERR_CAPACITY, "gn: a node belongs to more than %d rows", kGnMaxRows);
        case 3: return fail(VF_ERR_CAPACITY, "gn: a node shares rows with more than %d nodes", kGnMaxNbr);
        default: return fail(VF_ERR_HIP, "gn: inconsistent row lists");
    }
}

// The blocked Cholesky factorisation of the n x n matrix A (lower triangle,
// column-major), as described at k_chol_block; *info (device) as potrf's:
// 0, or the first column (from 1) whose pivot is not positive.  The
// rocBLAS calls run on the handle's stream hs, the diagonal blocks after the
// first on the solver's second stream.
int cholesky_lower(vf_gn_solver* s, hipStream_t hs, float* A, int n, int* info) {
    if (hipMemsetAsync(info, 0, sizeof(int), hs) != hipSuccess) return fail(VF_ERR_HIP, "gn: info clear");
    const float one = 1.0f, mone = -1.0f;
    const auto blas_ok = [](rocblas_status b) { return b == rocblas_status_success; };
    // the kb x kb diagonal block at column k (kb <= kCholOuter) on stream st,
    // whose rocBLAS handle is h
    const auto diag = [&](rocblas_handle h, hipStream_t st, int k, int kb) -> int {
        // right-looking over blocks of at most kCholBlock columns
        for (int c = 0; c < kb; c += kCholBlock) {
            float* D = A + (k + c) + (size_t)(k + c) * n;
            const int h1 = std::min(kCholBlock, kb - c), h2 = kb - c - h1;
            hipLaunchKernelGGL(k_chol_block, dim3(1), dim3(kCholThreads), 0, st, D, n, h1, k + c, info);
            if (hipGetLastError() != hipSuccess) return fail(VF_ERR_HIP, "k_chol_block launch");
            if (h2 == 0) break;
            float* D21 = D + h1;
            float* D22 = D21 + (size_t)h1 * n;
            if (!blas_ok(rocblas_strsm(h, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                                       rocblas_diagonal_non_unit, h2, h1, &one, D, n, D21, n)) ||
                !blas_ok(rocblas_ssyrk(h, rocblas_fill_lower, rocblas_operation_none, h2, h1, &mone, D21, n, &one,
                                       D22, n)))
                return fail(VF_ERR_HIP, "gn: rocBLAS call in the Cholesky factorisation");
        }
        return VF_OK;
    };
    int st = diag(s->blas, hs, 0, std::min(kCholOuter, n));
    for (int k = 0; !st && k < n; k += kCholOuter) {
        const int kb = std::min(kCholOuter, n - k);
        const int m = n - k - kb;  // rows below the diagonal block
        if (k > 0 && hipStreamWaitEvent(hs, s->ev_diag, 0) != hipSuccess) st = fail(VF_ERR_HIP, "gn: stream wait");
        if (st || m <= 0) break;
        float* A11 = A + k + (size_t)k * n;
        float* A21 = A11 + kb;
        float* D1 = A + (k + kb) + (size_t)(k + kb) * n;  // the next diagonal block
        const int nb1 = std::min(kCholOuter, m);
        bool ok = blas_ok(rocblas_strsm(s->blas, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                                        rocblas_diagonal_non_unit, m, kb, &one, A11, n, A21, n));
        ok = ok && blas_ok(rocblas_ssyrk(s->blas, rocblas_fill_lower, rocblas_operation_none, nb1, kb, &mone, A21, n,
                                         &one, D1, n));
        if (ok && m > nb1)
            ok = blas_ok(rocblas_sgemm(s->blas, rocblas_operation_none, rocblas_operation_transpose, m - nb1, nb1, kb,
                                       &mone, A21 + nb1, n, A21, n, &one, D1 + nb1, n));
        if (!ok) {
            st = fail(VF_ERR_HIP, "gn: rocBLAS call in the Cholesky factorisation");
            break;
        }
        if (hipEventRecord(s->ev_col, hs) != hipSuccess || hipStreamWaitEvent(s->aux, s->ev_col, 0) != hipSuccess)
            st = fail(VF_ERR_HIP, "gn: look-ahead event");
        if (!st) st = diag(s->blas_aux, s->aux, k + kb, nb1);
        if (!st && hipEventRecord(s->ev_diag, s->aux) != hipSuccess) st = fail(VF_ERR_HIP, "gn: look-ahead event");
        // the rest of the trailing matrix, beside the next diagonal block's factorisation
        if (!st && m > nb1 &&
            !blas_ok(rocblas_ssyrk(s->blas, rocblas_fill_lower, rocblas_operation_none, m - nb1, kb, &mone,
                                   A21 + nb1, n, &one, D1 + nb1 + (size_t)nb1 * n, n)))
            st = fail(VF_ERR_HIP, "gn: rocBLAS call in the Cholesky factorisation");
    }
    return st;
}

// y = JTJ^-1 x with the factors of vf_gn_solver_create: the gather (P, the
// identity for Cholesky), then the forward and backward block steps (L and
// U of the LU factors, or L and L^T of the Cholesky factor).
int lu_solve_factored(vf_gn_solver* s, hipStream_t st) {
    const int n = (int)s->n;
    const int nb = (n + kTB - 1) / kTB;
    hipLaunchKernelGGL(k_gather_perm, dim3(blocks_for((uint64_t)n, 256)), dim3(256), 0, st, (const float*)s->x,
                       (const int*)s->perm, s->y, n);
    VF_LAUNCH_CHECK("k_gather_perm");
    // the ticket cleared before each triangle (not the failure flag); a new tag
    // per triangle (0 is never one: the tags start cleared)
    VF_HIP(hipMemsetAsync(s->chain, 0, sizeof(int), st));
    if (++s->epoch == 0) ++s->epoch;
    if (s->chol)
        hipLaunchKernelGGL(k_trsv_lower_chain<false>, dim3(nb), dim3(kTB), 0, st, (const float*)s->JTJ, n, s->y,
                           s->tagged, s->epoch, s->chain);
    else
        hipLaunchKernelGGL(k_trsv_lower_chain<true>, dim3(nb), dim3(kTB), 0, st, (const float*)s->JTJ, n, s->y,
                           s->tagged, s->epoch, s->chain);
    VF_LAUNCH_CHECK("k_trsv_lower_chain");
    VF_HIP(hipMemsetAsync(s->chain, 0, sizeof(int), st));
    if (++s->epoch == 0) ++s->epoch;
    if (s->chol)
        hipLaunchKernelGGL(k_trsv_upper_chain<true>, dim3(nb), dim3(kTB), 0, st, (const float*)s->JTJ, n, s->y,
                           s->tagged, s->epoch, s->chain);
    else
        hipLaunchKernelGGL(k_trsv_upper_chain<false>, dim3(nb), dim3(kTB), 0, st, (const float*)s->JTJ, n, s->y,
                           s->tagged, s->epoch, s->chain);
    VF_LAUNCH_CHECK("k_trsv_upper_chain");
    return VF_OK;
}

int launch_residuals(vf_gn_solver* s, const float* warped, const vf_float3* dmesh, const vf_mat4* T, hipStream_t st) {
    const GnRows& g = s->rows;
    VF_HIP(hipMemsetAsync(s->r, 0, sizeof(float) * s->nrows, st));
    if (g.C) {
        hipLaunchKernelGGL(k_r_depth, dim3(blocks_for(g.C, 256)), dim3(256), 0, st, s->r, g.wD, g.corr, g.C, warped,
                           (const float*)dmesh, g.dnormals);
        VF_LAUNCH_CHECK("k_r_depth");
    }
    hipLaunchKernelGGL(k_r_mreg, dim3(blocks_for((uint64_t)g.G * g.R, 256)), dim3(256), 0, st, s->r + g.C, g.wR,
                       g.graph, (const float*)T, g.G, g.gknn, g.R);
    VF_LAUNCH_CHECK("k_r_mreg");
    hipLaunchKernelGGL(k_energy_part, dim3(kEnergyBlocks), dim3(256), 0, st, (const float*)s->r, (uint64_t)s->nrows,
                       s->part);
    VF_LAUNCH_CHECK("k_energy_part");
    hipLaunchKernelGGL(k_energy_sum, dim3(1), dim3(64), 0, st, (const double*)s->part, s->energy);
    VF_LAUNCH_CHECK("k_energy_sum");
    return VF_OK;
}

}  // namespace
}  // namespace vf

using namespace vf;

extern "C" {

int vf_correspondence_pairs(vf_workspace* ws, vf_int2* corr, uint32_t* corr_size, const vf_float3* warped,
                            size_t mesh_size, vf_mat3 K, const float* depth, vf_int2 depth_dim,
                            const vf_float3* frame_mesh, float thresh, vf_stream stream) {
    if (!mesh_size) return VF_OK;
    if (!corr || !corr_size || !warped || !depth || !frame_mesh)
        return fail(VF_ERR_INVALID_ARGUMENT, "vf_correspondence_pairs: NULL argument");
    if (depth_dim.x < 0 || depth_dim.y < 0 || mesh_size > 0x7fffffffull)
        return fail(VF_ERR_INVALID_ARGUMENT, "vf_correspondence_pairs: bad sizes");
    CorrIn in;
    in.warped = (const float*)warped;
    in.n = mesh_size;
    in.fx = K.m[0];
    in.fy = K.m[4];
    in.cx = K.m[6];
    in.cy = K.m[7];
    in.depth = depth;
    in.rows = depth_dim.x;
    in.cols = depth_dim.y;
    in.frame_mesh = (const float*)frame_mesh;
    in.thresh = thresh;
    const unsigned nb = blocks_for(mesh_size, kCorrBlock);
    void* mem = nullptr;
    int st = workspace_scratch(ws, sizeof(uint32_t) * (nb + 1ull), &mem, (hipStream_t)stream);
    if (st) return st;
    uint32_t* counts = (uint32_t*)mem;
    hipLaunchKernelGGL(k_corr_count, dim3(nb), dim3(kCorrBlock), 0, (hipStream_t)stream, in, counts);
    VF_LAUNCH_CHECK("k_corr_count");
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, (hipStream_t)stream, counts, nb, corr_size);
    VF_LAUNCH_CHECK("k_scan_counts");
    hipLaunchKernelGGL(k_corr_write, dim3(nb), dim3(kCorrBlock), 0, (hipStream_t)stream, in, (const uint32_t*)counts,
                       (int32_t*)corr);
    VF_LAUNCH_CHECK("k_corr_write");
    return VF_OK;
}

int vf_fill_r_depth(float* r, float weight, const vf_int2* corr, uint32_t corr_size, const vf_float3* warped,
                    const vf_float3* depth_mesh, const vf_float3* normals, vf_stream stream) {
    if (!corr_size) return VF_OK;
    if (!r || !corr || !warped || !depth_mesh || !normals) return fail(VF_ERR_INVALID_ARGUMENT, "vf_fill_r_depth: NULL");
    hipLaunchKernelGGL(k_r_depth, dim3(blocks_for(corr_size, 256)), dim3(256), 0, (hipStream_t)stream, r, weight,
                       (const int32_t*)corr, corr_size, (const float*)warped, (const float*)depth_mesh,
                       (const float*)normals);
    VF_LAUNCH_CHECK("k_r_depth");
    return VF_OK;
}

int vf_fill_r_mreg(float* r, float weight, const vf_float3* graph, const vf_mat4* transforms, uint32_t graph_size,
                   const uint32_t* graph_knn, int reg_k, vf_stream stream) {
    if (!graph_size) return VF_OK;
    if (reg_k < 1 || reg_k > kGnMaxR) return fail(VF_ERR_INVALID_ARGUMENT, "vf_fill_r_mreg: reg_k=%d", reg_k);
    if (!r || !graph || !transforms || !graph_knn) return fail(VF_ERR_INVALID_ARGUMENT, "vf_fill_r_mreg: NULL");
    hipLaunchKernelGGL(k_r_mreg, dim3(blocks_for((uint64_t)graph_size * reg_k, 256)), dim3(256), 0,
                       (hipStream_t)stream, r, weight, (const float*)graph, (const float*)transforms, graph_size,
                       graph_knn, reg_k);
    VF_LAUNCH_CHECK("k_r_mreg");
    return VF_OK;
}

int vf_fill_j_depth(float* J, float weight, uint32_t j_height, uint32_t j_width, const vf_int2* corr,
                    uint32_t corr_size, const vf_float3* canon, int k, const uint32_t* knn_idxs,
                    const float* knn_weights, const vf_float3* depth_mesh, const vf_float3* normals,
                    vf_stream stream) {
    (void)depth_mesh;
    if (!corr_size) return VF_OK;
    if (k < 1 || k > kGnMaxK) return fail(VF_ERR_UNSUPPORTED_K, "vf_fill_j_depth: k=%d", k);
    if (corr_size > j_height) return fail(VF_ERR_INVALID_ARGUMENT, "vf_fill_j_depth: more pairs than J rows");
    if (!J || !corr || !canon || !knn_idxs || !knn_weights || !normals)
        return fail(VF_ERR_INVALID_ARGUMENT, "vf_fill_j_depth: NULL");
    hipLaunchKernelGGL(k_j_depth, dim3(blocks_for(corr_size, 256)), dim3(256), 0, (hipStream_t)stream, J, weight,
                       j_width, (const int32_t*)corr, corr_size, (const float*)canon, k, knn_idxs, knn_weights,
                       (const float*)normals);
    VF_LAUNCH_CHECK("k_j_depth");
    return VF_OK;
}

int vf_fill_j_mreg(float* J, float weight, uint32_t j_height, uint32_t j_width, const vf_float3* graph,
                   const vf_mat4* transforms, uint32_t graph_size, const uint32_t* graph_knn, int reg_k,
                   vf_stream stream) {
    (void)transforms;
    if (!graph_size) return VF_OK;
    if (reg_k < 1 || reg_k > kGnMaxR) return fail(VF_ERR_INVALID_ARGUMENT, "vf_fill_j_mreg: reg_k=%d", reg_k);
    if ((uint64_t)graph_size * reg_k * 3 > j_height || graph_size > j_width)
        return fail(VF_ERR_INVALID_ARGUMENT, "vf_fill_j_mreg: J too small");
    if (!J || !graph || !graph_knn) return fail(VF_ERR_INVALID_ARGUMENT, "vf_fill_j_mreg: NULL");
    hipLaunchKernelGGL(k_j_mreg, dim3(blocks_for((uint64_t)graph_size * reg_k, 256)), dim3(256), 0,
                       (hipStream_t)stream, J, weight, j_width, (const float*)graph, graph_size, graph_knn, reg_k);
    VF_LAUNCH_CHECK("k_j_mreg");
    return VF_OK;
}

int vf_update_transforms(vf_mat4* transforms, const float* updates, uint32_t graph_size, vf_stream stream) {
    if (!graph_size) return VF_OK;
    if (!transforms || !updates) return fail(VF_ERR_INVALID_ARGUMENT, "vf_update_transforms: NULL");
    hipLaunchKernelGGL(k_update, dim3(blocks_for(graph_size, 256)), dim3(256), 0, (hipStream_t)stream,
                       (float*)transforms, updates, graph_size);
    VF_LAUNCH_CHECK("k_update");
    return VF_OK;
}

int vf_lu_solve(int n, const float* A, int lda, const float* b, float* x, vf_stream stream) {
    if (n < 0 || lda < std::max(n, 1)) return fail(VF_ERR_INVALID_ARGUMENT, "vf_lu_solve: n=%d lda=%d", n, lda);
    if (!n) return VF_OK;
    if (!A || !b || !x) return fail(VF_ERR_INVALID_ARGUMENT, "vf_lu_solve: NULL");
    hipStream_t st = (hipStream_t)stream;
    float* Ac = nullptr;
    int* ipiv = nullptr;
    int* info = nullptr;
    rocblas_handle h = nullptr;
    int rc = VF_OK;
    auto cleanup = [&]() {
        if (Ac) (void)hipFree(Ac);
        if (ipiv) (void)hipFree(ipiv);
        if (info) (void)hipFree(info);
        if (h) (void)rocblas_destroy_handle(h);
    };
    if (!rc) rc = gn_alloc(&Ac, (uint64_t)lda * n);
    if (!rc) rc = gn_alloc(&ipiv, (uint64_t)n);
    if (!rc) rc = gn_alloc(&info, 1);
    if (!rc && rocblas_create_handle(&h) != rocblas_status_success) rc = fail(VF_ERR_HIP, "rocblas_create_handle");
    if (!rc && rocblas_set_stream(h, st) != rocblas_status_success) rc = fail(VF_ERR_HIP, "rocblas_set_stream");
    // getrf overwrites A with its factors: factor a copy, solve into x
    if (!rc && hipMemcpyAsync(Ac, A, sizeof(float) * (uint64_t)lda * n, hipMemcpyDeviceToDevice, st) != hipSuccess)
        rc = fail(VF_ERR_HIP, "vf_lu_solve: copy of A");
    if (!rc && hipMemcpyAsync(x, b, sizeof(float) * (uint64_t)n, hipMemcpyDeviceToDevice, st) != hipSuccess)
        rc = fail(VF_ERR_HIP, "vf_lu_solve: copy of b");
    if (!rc && rocsolver_sgetrf(h, n, n, Ac, lda, ipiv, info) != rocblas_status_success)
        rc = fail(VF_ERR_HIP, "rocsolver_sgetrf");
    int hinfo = 0;
    if (!rc && (hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess))
        rc = fail(VF_ERR_HIP, "vf_lu_solve: info read-back");
    if (!rc && hinfo != 0) rc = fail(VF_ERR_SINGULAR, "vf_lu_solve: zero pivot at column %d", hinfo);
    if (!rc && rocsolver_sgetrs(h, rocblas_operation_none, n, 1, Ac, lda, ipiv, x, n) != rocblas_status_success)
        rc = fail(VF_ERR_HIP, "rocsolver_sgetrs");
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = fail(VF_ERR_HIP, "vf_lu_solve: synchronize");
    cleanup();
    return rc;
}

int vf_gn_normal_equations(vf_workspace* ws, float* JTJ, float* JTr, const float* r, const vf_gn_problem* problem,
                           vf_stream stream) {
    (void)ws;
    int st = check_problem(problem);
    if (st) return st;
    if (!JTJ && !JTr) return VF_OK;
    if (JTr && !r) return fail(VF_ERR_INVALID_ARGUMENT, "vf_gn_normal_equations: JTr needs r");
    vf_gn_solver s;
    VF_HIP(hipGetDevice(&s.device));
    s.rows = make_rows(problem);
    s.nrows = s.rows.C + 3u * s.rows.G * (uint32_t)s.rows.R;
    s.n = 6ull * s.rows.G;
    hipStream_t hs = (hipStream_t)stream;
    if (!st) st = gn_alloc(&s.cnt, s.rows.G + 1ull);
    if (!st) st = gn_alloc(&s.fill, s.rows.G);
    if (!st) st = gn_alloc(&s.bad, 1);
    if (!st) st = build_rows(&s, hs);
    if (!st && JTJ) {
        hipLaunchKernelGGL(k_identity, dim3(2048), dim3(256), 0, hs, JTJ, s.n);
        if (hipGetLastError() != hipSuccess) st = fail(VF_ERR_HIP, "k_identity launch");
    }
    if (!st) {
        hipLaunchKernelGGL(k_blocks, dim3(s.rows.G), dim3(64), 0, hs, s.rows.G, (const uint32_t*)s.cnt,
                           (const uint32_t*)s.rows_of, (const uint32_t*)s.node8, (const float*)s.j48, r, JTJ, JTr,
                           s.bad);
        if (hipGetLastError() != hipSuccess) st = fail(VF_ERR_HIP, "k_blocks launch");
    }
    if (!st) st = read_bad(&s, hs);
    gn_free(&s);
    return st;
}

int vf_gn_solver_create(vf_gn_solver** out, const vf_gn_problem* problem, vf_stream stream) {
    return vf_gn_solver_create_ex(out, problem, VF_GN_FACTOR_AUTO, stream);
}

int vf_gn_solver_create_ex(vf_gn_solver** out, const vf_gn_problem* problem, int factorisation, vf_stream stream) {
    if (!out) return fail(VF_ERR_INVALID_ARGUMENT, "vf_gn_solver_create: out is NULL");
    if (factorisation != VF_GN_FACTOR_AUTO && factorisation != VF_GN_FACTOR_LU)
        return fail(VF_ERR_INVALID_ARGUMENT, "vf_gn_solver_create: factorisation %d", factorisation);
    *out = nullptr;
    int st = check_problem(problem);
    if (st) return st;
    vf_gn_solver* s = new vf_gn_solver();
    hipStream_t hs = (hipStream_t)stream;
    (void)hipGetDevice(&s->device);
    s->rows = make_rows(problem);
    s->nrows = s->rows.C + 3u * s->rows.G * (uint32_t)s->rows.R;
    s->n = 6ull * s->rows.G;
    if (!st) st = gn_alloc(&s->cnt, s->rows.G + 1ull);
    if (!st) st = gn_alloc(&s->fill, s->rows.G);
    if (!st) st = gn_alloc(&s->bad, 1);
    if (!st) st = gn_alloc(&s->JTJ, s->n * s->n);
    if (!st) st = gn_alloc(&s->ipiv, s->n);
    if (!st) st = gn_alloc(&s->info, 1);
    if (!st) st = gn_alloc(&s->r, s->nrows);
    if (!st) st = gn_alloc(&s->x, s->n);
    if (!st) st = gn_alloc(&s->y, s->n);
    if (!st) st = gn_alloc(&s->perm, s->n);
    if (!st) st = gn_alloc(&s->chain, 2);
    if (!st) st = gn_alloc(&s->tagged, s->n);
    if (!st && (hipMemsetAsync(s->chain, 0, 2 * sizeof(int), hs) != hipSuccess ||
                hipMemsetAsync(s->tagged, 0, sizeof(unsigned long long) * s->n, hs) != hipSuccess))
        st = fail(VF_ERR_HIP, "gn: substitution state clear");
    if (!st) st = gn_alloc(&s->part, kEnergyBlocks);
    if (!st) st = gn_alloc(&s->energy, 1);
    if (!st && rocblas_create_handle(&s->blas) != rocblas_status_success)
        st = fail(VF_ERR_HIP, "rocblas_create_handle");
    if (!st && rocblas_set_stream(s->blas, hs) != rocblas_status_success) st = fail(VF_ERR_HIP, "rocblas_set_stream");
    if (!st && (hipStreamCreateWithFlags(&s->aux, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&s->ev_col, hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&s->ev_diag, hipEventDisableTiming) != hipSuccess))
        st = fail(VF_ERR_HIP, "gn: look-ahead stream and events");
    if (!st && (rocblas_create_handle(&s->blas_aux) != rocblas_status_success ||
                rocblas_set_stream(s->blas_aux, s->aux) != rocblas_status_success))
        st = fail(VF_ERR_HIP, "gn: look-ahead rocBLAS handle");
    if (!st) st = build_rows(s, hs);
    const auto normal_matrix = [&]() -> int {
        hipLaunchKernelGGL(k_identity, dim3(2048), dim3(256), 0, hs, s->JTJ, s->n);
        hipLaunchKernelGGL(k_blocks, dim3(s->rows.G), dim3(64), 0, hs, s->rows.G, (const uint32_t*)s->cnt,
                           (const uint32_t*)s->rows_of, (const uint32_t*)s->node8, (const float*)s->j48,
                           (const float*)nullptr, s->JTJ, (float*)nullptr, s->bad);
        if (hipGetLastError() != hipSuccess) return fail(VF_ERR_HIP, "normal-equation kernels launch");
        return read_bad(s, hs);
    };
    const auto read_info = [&](int* hinfo) -> int {
        if (hipMemcpyAsync(hinfo, s->info, sizeof(int), hipMemcpyDeviceToHost, hs) != hipSuccess ||
            hipStreamSynchronize(hs) != hipSuccess)
            return fail(VF_ERR_HIP, "gn: info read-back");
        return VF_OK;
    };
    if (!st) st = normal_matrix();
    // JTJ does not change within the ICP iteration: factorise it once.  JTJ
    // (J^T J, with the identity where a node has no rows) is symmetric
    // positive definite unless the problem is degenerate, so the Cholesky
    // factorisation L L^T (half the work of the LU and no pivot search;
    // blocked, cholesky_lower) solves the same system; the reference calls an
    // LU with partial pivoting (linearSolverLUStable,
    // src/LinearSolver.cu:179-225).  Where the Cholesky factorisation finds a
    // non-positive pivot, JTJ is rebuilt and factorised as the reference
    // does (getrf), which then reports a zero pivot as the reference's solver
    // would.
    if (!st && factorisation == VF_GN_FACTOR_AUTO) {
        st = cholesky_lower(s, hs, s->JTJ, (int)s->n, s->info);
        int hinfo = 0;
        if (!st) st = read_info(&hinfo);
        s->chol = !st && hinfo == 0;
    }
    if (!st && !s->chol) {
        // (after a Cholesky attempt JTJ holds its partial factor)
        if (factorisation == VF_GN_FACTOR_AUTO) st = normal_matrix();
        // Partial pivoting does interchange rows on these matrices (the
        // rotation columns' couplings exceed their diagonals), so the LU
        // without pivot searches is not an option.
        if (!st && rocsolver_sgetrf(s->blas, (int)s->n, (int)s->n, s->JTJ, (int)s->n, s->ipiv, s->info) !=
                       rocblas_status_success)
            st = fail(VF_ERR_HIP, "rocsolver_sgetrf");
        if (!st) {
            int hinfo = 0;
            st = read_info(&hinfo);
            if (!st && hinfo != 0)
                st = fail(VF_ERR_SINGULAR, "gn: the normal matrix has a zero pivot at column %d", hinfo);
        }
    }
    if (!st && s->chol) {
        std::vector<int> perm(s->n);
        for (uint64_t i = 0; i < s->n; ++i) perm[i] = (int)i;
        if (hipMemcpyAsync(s->perm, perm.data(), sizeof(int) * s->n, hipMemcpyHostToDevice, hs) != hipSuccess ||
            hipStreamSynchronize(hs) != hipSuccess)
            st = fail(VF_ERR_HIP, "gn: permutation upload");
    } else if (!st) {
        // the interchanges (row i with row ipiv[i] - 1, in order) as one gather
        std::vector<int> piv(s->n), perm(s->n);
        if (hipMemcpyAsync(piv.data(), s->ipiv, sizeof(int) * s->n, hipMemcpyDeviceToHost, hs) != hipSuccess ||
            hipStreamSynchronize(hs) != hipSuccess) {
            st = fail(VF_ERR_HIP, "gn: pivot read-back");
        } else {
            for (uint64_t i = 0; i < s->n; ++i) perm[i] = (int)i;
            for (uint64_t i = 0; i < s->n && !st; ++i) {
                const int t = piv[i] - 1;
                if (t < 0 || (uint64_t)t >= s->n) st = fail(VF_ERR_HIP, "gn: pivot %d out of range", piv[i]);
                else std::swap(perm[i], perm[(size_t)t]);
            }
            if (!st && (hipMemcpyAsync(s->perm, perm.data(), sizeof(int) * s->n, hipMemcpyHostToDevice, hs) !=
                            hipSuccess ||
                        hipStreamSynchronize(hs) != hipSuccess))
                st = fail(VF_ERR_HIP, "gn: permutation upload");
        }
    }
    if (st) {
        gn_free(s);
        delete s;
        return st;
    }
    *out = s;
    return VF_OK;
}

int vf_gn_solver_destroy(vf_gn_solver* s) {
    if (!s) return VF_OK;
    gn_free(s);
    delete s;
    return VF_OK;
}

int vf_gn_solver_iterate(vf_gn_solver* s, vf_mat4* transforms, vf_mat4* pending, vf_float3* warped,
                         const vf_float3* canon, size_t mesh_size, const vf_float3* depth_mesh, int max_iters,
                         float min_step, vf_gn_report* report, vf_stream stream) {
    if (!s || !transforms || !pending || !warped || !canon || !report)
        return fail(VF_ERR_INVALID_ARGUMENT, "vf_gn_solver_iterate: NULL argument");
    if (s->rows.C && !depth_mesh) return fail(VF_ERR_INVALID_ARGUMENT, "vf_gn_solver_iterate: NULL depth mesh");
    if (max_iters < 0 || max_iters > 1000)
        return fail(VF_ERR_INVALID_ARGUMENT, "vf_gn_solver_iterate: max_iters %d", max_iters);
    hipStream_t hs = (hipStream_t)stream;
    if (rocblas_set_stream(s->blas, hs) != rocblas_status_success) return fail(VF_ERR_HIP, "rocblas_set_stream");
    memset(report, 0, sizeof(*report));
    report->cholesky = s->chol ? 1 : 0;
    const uint32_t G = s->rows.G;
    // the loop of src/VolumeFusion.cu:541-691: pending starts as the transforms
    VF_HIP(hipMemcpyAsync(pending, transforms, sizeof(vf_mat4) * G, hipMemcpyDeviceToDevice, hs));
    VF_HIP(hipMemsetAsync(s->chain + 1, 0, sizeof(int), hs));  // the substitutions' failure flag
    double energy = 1.0e300;
    for (int it = 0; it < max_iters; ++it) {
        int st = launch_residuals(s, (const float*)warped, depth_mesh, pending, hs);
        if (st) return st;
        double e = 0.0;
        VF_HIP(hipMemcpyAsync(&e, s->energy, sizeof(double), hipMemcpyDeviceToHost, hs));
        VF_HIP(hipStreamSynchronize(hs));
        if (it < (int)(sizeof(report->energy) / sizeof(report->energy[0]))) report->energy[it] = e;
        report->iterations = it;
        if (energy - e < (double)min_step) {
            report->stopped = 1;
            break;
        }
        energy = e;
        report->final_energy = e;
        if (it != 0) VF_HIP(hipMemcpyAsync(transforms, pending, sizeof(vf_mat4) * G, hipMemcpyDeviceToDevice, hs));
        // JTr from the new residuals, then the back substitution
        hipLaunchKernelGGL(k_blocks, dim3(G), dim3(64), 0, hs, G, (const uint32_t*)s->cnt,
                           (const uint32_t*)s->rows_of, (const uint32_t*)s->node8, (const float*)s->j48,
                           (const float*)s->r, (float*)nullptr, s->x, s->bad);
        VF_LAUNCH_CHECK("k_blocks");
        st = lu_solve_factored(s, hs);
        if (st) return st;
        hipLaunchKernelGGL(k_update, dim3(blocks_for(G, 256)), dim3(256), 0, hs, (float*)pending, (const float*)s->y,
                           G);
        VF_LAUNCH_CHECK("k_update");
        st = vf_warp_mesh(s->rows.K, warped, canon, pending, s->rows.knn, s->rows.kw, mesh_size, G, stream);
        if (st) return st;
        report->iterations = it + 1;
    }
    int gave_up = 0;
    VF_HIP(hipMemcpyAsync(&gave_up, s->chain + 1, sizeof(int), hipMemcpyDeviceToHost, hs));
    VF_HIP(hipStreamSynchronize(hs));
    if (gave_up) return fail(VF_ERR_HIP, "gn: a triangular solve's wait for a block gave up");
    return VF_OK;
}

}  // extern "C"
