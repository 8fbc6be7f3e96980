/* vf_gn.c -- CPU restatement of the reference's Gauss-Newton motion
 * optimisation for the checker (TEST INFRASTRUCTURE ONLY: loaded by tests/,
 * never by the volumefusion_amd package).
 *
 * Covers one ICP iteration of src/VolumeFusion.cu:434-692:
 *   vfo_corr_pairs        correspondencePairsKernel,
 *                         src/ComputeCorrespondence.cu:46-82 (pairs in
 *                         ascending mesh index; the reference appends them
 *                         with an atomic counter, so its order is a race)
 *   vfo_gn_r_depth        fill_r_DepthKernel, src/MotionOptimization.cu:165-181
 *   vfo_gn_r_mreg         fill_r_MRegKernel, :201-225
 *   vfo_gn_j_depth        fill_J_DepthKernel, :29-60 (dense J of Vec6f)
 *   vfo_gn_j_mreg         fill_J_MRegKernel, :80-145
 *   vfo_gn_normal_eq      JTJ and JTr from the contributing pairs
 *                         (:227-479 with the sort and reduce_by_key of
 *                         src/VolumeFusion.cu:510-613): identity diagonal, a
 *                         6x6 block for every node pair that shares a row,
 *                         summed over the rows in ascending row order (the
 *                         reference's order is an atomic race), first term
 *                         first
 *   vfo_gn_update         updateTransformsKernel, :481-498: T <- T - [w]x|t
 *   vfo_lu_solve          linearSolverLUStable, src/LinearSolver.cu:179-225:
 *                         LU with partial pivoting (getrf + getrs), binary64
 *
 * Pins (tests/test_oracle_gn.py): the Jacobians against finite differences of
 * the residuals under the update rule; one Gauss-Newton step solving an
 * exactly linear problem; the update's closed form; the LU solve against
 * LAPACK; correspondences of a plane at known pixels.  The float expression
 * orders follow the reference's Eigen expressions (checked against them in
 * tests/cpp/eigen_order_check.cpp). */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "vf_oracle.h"

#define M3(m, i, j) ((m)[(j) * 3 + (i)])

/* (M * p4).head(3) with a materialised Vec4f (Eigen's four-term reduction
 * (a0 + a1) + (a2 + a3), the mesh-warp order). */
static void mat4_apply4(const float* M, const float* p, float* o) {
    for (int i = 0; i < 3; ++i) {
        const float a0 = M[0 * 4 + i] * p[0];
        const float a1 = M[1 * 4 + i] * p[1];
        const float a2 = M[2 * 4 + i] * p[2];
        const float a3 = M[3 * 4 + i] * 1.0f;
        o[i] = (a0 + a1) + (a2 + a3);
    }
}

/* three-term Eigen reduction a0 + (a1 + a2) */
static inline float sum3(float a0, float a1, float a2) { return a0 + (a1 + a2); }

int64_t vfo_corr_pairs(int32_t* corr, const float* warped, size_t M, const float* Kin, const float* depth, int rows,
                       int cols, const float* frame_mesh, float thresh) {
    const float fx = M3(Kin, 0, 0), fy = M3(Kin, 1, 1), cx = M3(Kin, 0, 2), cy = M3(Kin, 1, 2);
    int64_t n = 0;
    for (size_t i = 0; i < M; ++i) {
        const float* v = warped + 3 * i;
        const float y0 = v[1] / v[2];
        const float x0 = v[0] / v[2];
        const int32_t y = vfo_cvt_i32_rz(y0 * fy + cy);
        const int32_t x = vfo_cvt_i32_rz(x0 * fx + cx);
        if (y < 0 || x < 0 || y >= rows || x >= cols)
            continue;
        const size_t d = (size_t)y * (size_t)cols + (size_t)x;
        const float* fm = frame_mesh + 3 * d;
        const float dx = fm[0] - v[0], dy = fm[1] - v[1], dz = fm[2] - v[2];
        const float nrm = sqrtf(sum3(dx * dx, dy * dy, dz * dz));
        if (nrm > thresh)
            continue;
        if (depth[d] > 1.e-4f) {
            corr[2 * n] = (int32_t)i;
            corr[2 * n + 1] = (int32_t)d;
            ++n;
        }
    }
    return n;
}

void vfo_gn_r_depth(float* r, float wD, const int32_t* corr, size_t C, const float* warped, const float* dmesh,
                    const float* dnormals) {
    for (size_t c = 0; c < C; ++c) {
        const float* v = warped + 3 * (size_t)corr[2 * c];
        const float* d = dmesh + 3 * (size_t)corr[2 * c + 1];
        const float* nn = dnormals + 3 * (size_t)corr[2 * c + 1];
        const float df[3] = {v[0] - d[0], v[1] - d[1], v[2] - d[2]};
        r[c] = wD * sum3(nn[0] * df[0], nn[1] * df[1], nn[2] * df[2]);
    }
}

void vfo_gn_r_mreg(float* r, float wR, const float* graph, const float* T, size_t G, const uint32_t* gknn, int R) {
    for (size_t g = 0; g < G; ++g) {
        float tc[3];
        mat4_apply4(T + 16 * g, graph + 3 * g, tc);
        for (int i = 0; i < R; ++i) {
            const uint32_t nb = gknn[(size_t)R * g + i];
            if (nb >= G)
                continue;
            float tn[3];
            mat4_apply4(T + 16 * (size_t)nb, graph + 3 * g, tn);
            const size_t row = (g * (size_t)R + (size_t)i) * 3;
            for (int a = 0; a < 3; ++a)
                r[row + a] = (tc[a] - tn[a]) * wR;
        }
    }
}

/* row a of gradTv(p) (src/MotionOptimization.cu:10-27) */
static void grad_row(const float* p, int a, float g[6]) {
    memset(g, 0, 6 * sizeof(float));
    if (a == 0) {
        g[1] = p[2]; g[2] = -p[1]; g[3] = 1.0f;
    } else if (a == 1) {
        g[0] = -p[2]; g[2] = p[0]; g[4] = 1.0f;
    } else {
        g[0] = p[1]; g[1] = -p[0]; g[5] = 1.0f;
    }
}

/* ((n w)^T gradTv(p)) wD: each entry a three-term reduction over the rows of
 * gradTv (zeros included), then the weight. */
static void j_depth_entry(const float* n, float w, const float* p, float wD, float out[6]) {
    const float nw[3] = {n[0] * w, n[1] * w, n[2] * w};
    float G[3][6];
    for (int a = 0; a < 3; ++a)
        grad_row(p, a, G[a]);
    for (int k = 0; k < 6; ++k)
        out[k] = sum3(nw[0] * G[0][k], nw[1] * G[1][k], nw[2] * G[2][k]) * wD;
}

void vfo_gn_j_depth(float* J, float wD, size_t C, size_t G, const int32_t* corr, const float* canon,
                    const uint32_t* knn, const float* kw, int K, const float* dnormals) {
    for (size_t c = 0; c < C; ++c) {
        const size_t ci = (size_t)corr[2 * c], di = (size_t)corr[2 * c + 1];
        for (int s = 0; s < K; ++s) {
            const float w = kw[(size_t)K * ci + s];
            if (w < 1.e-4f)
                continue;
            const uint32_t j = knn[(size_t)K * ci + s];
            j_depth_entry(dnormals + 3 * di, w, canon + 3 * ci, wD, J + 6 * (c * G + j));
        }
    }
}

void vfo_gn_j_mreg(float* J, float wR, const float* graph, size_t G, const uint32_t* gknn, int R) {
    for (size_t g = 0; g < G; ++g) {
        float gr[3][6];
        for (int a = 0; a < 3; ++a) {
            grad_row(graph + 3 * g, a, gr[a]);
            for (int k = 0; k < 6; ++k)
                gr[a][k] = gr[a][k] * wR;
        }
        for (int i = 0; i < R; ++i) {
            const uint32_t nb = gknn[(size_t)R * g + i];
            if (nb >= G)
                continue;
            const size_t row = (g * (size_t)R + (size_t)i) * 3;
            for (int a = 0; a < 3; ++a) {
                float* jc = J + 6 * ((row + a) * G + g);
                float* jn = J + 6 * ((row + a) * G + nb);
                for (int k = 0; k < 6; ++k) {
                    jc[k] = gr[a][k];
                    jn[k] = -gr[a][k];
                }
            }
        }
    }
}

/* A row's members: the nodes the reference's contribution enumeration
 * (src/MotionOptimization.cu:227-259, 284-314) pairs in that row, each with a
 * multiplicity for the diagonal block and for JTr. */
typedef struct {
    uint32_t node[64];
    int diag_mult[64];
    int jtr_mult[64];
    int n;
} members_t;

static void row_members(size_t row, size_t C, size_t G, const int32_t* corr, const uint32_t* knn, const float* kw,
                        int K, const uint32_t* gknn, int R, members_t* m) {
    m->n = 0;
    if (row < C) {
        const size_t ci = (size_t)corr[2 * row];
        for (int s = 0; s < K; ++s) {
            if (kw[(size_t)K * ci + s] == 0.0f)
                continue;
            m->node[m->n] = knn[(size_t)K * ci + s];
            m->diag_mult[m->n] = 1;
            m->jtr_mult[m->n] = 1;
            m->n++;
        }
        return;
    }
    const size_t mr = (row - C) / 3;
    const size_t g = mr / (size_t)R;
    const uint32_t nb = gknn[mr];
    if (g >= G || nb >= G)
        return;
    if (nb == g) {
        /* the pairs (g, nb), (g, g), (nb, nb) coincide; JTr takes (row, g) twice */
        m->node[0] = (uint32_t)g;
        m->diag_mult[0] = 3;
        m->jtr_mult[0] = 2;
        m->n = 1;
        return;
    }
    m->node[0] = (uint32_t)g;
    m->node[1] = nb;
    m->diag_mult[0] = m->diag_mult[1] = 1;
    m->jtr_mult[0] = m->jtr_mult[1] = 1;
    m->n = 2;
}

/* JTJ (6G x 6G, row-major == column-major) and JTr (G x 6) from the dense J
 * (jH rows x G columns of Vec6f) and r.  Rows [0, C) are depth rows, then
 * 3 R rows per node. */
int vfo_gn_normal_eq(float* JTJ, float* JTr, const float* J, const float* r, size_t C, size_t G, const int32_t* corr,
                     const uint32_t* knn, const float* kw, int K, const uint32_t* gknn, int R) {
    const size_t jH = C + G * (size_t)R * 3;
    const size_t W = 6 * G;
    uint8_t* started = (uint8_t*)calloc(G * G, 1);
    uint8_t* jstarted = (uint8_t*)calloc(G, 1);
    members_t* m = (members_t*)malloc(sizeof(members_t));
    if (!started || !jstarted || !m) {
        free(started); free(jstarted); free(m);
        return -1;
    }
    memset(JTJ, 0, sizeof(float) * W * W);
    for (size_t i = 0; i < W; ++i)
        JTJ[i * W + i] = 1.0f;
    memset(JTr, 0, sizeof(float) * 6 * G);
    for (size_t row = 0; row < jH; ++row) {
        row_members(row, C, G, corr, knn, kw, K, gknn, R, m);
        for (int p = 0; p < m->n; ++p) {
            const uint32_t i = m->node[p];
            const float* ji = J + 6 * (row * G + i);
            for (int t = 0; t < m->jtr_mult[p]; ++t) {
                for (int k = 0; k < 6; ++k) {
                    const float v = ji[k] * r[row];
                    JTr[6 * i + k] = jstarted[i] ? JTr[6 * i + k] + v : v;
                }
                jstarted[i] = 1;
            }
            for (int q = 0; q < m->n; ++q) {
                const uint32_t j = m->node[q];
                const float* jj = J + 6 * (row * G + j);
                const int mult = (i == j) ? m->diag_mult[p] : 1;
                for (int t = 0; t < mult; ++t) {
                    uint8_t* st = started + (size_t)i * G + j;
                    for (int a = 0; a < 6; ++a)
                        for (int b = 0; b < 6; ++b) {
                            float* e = JTJ + (6 * (size_t)i + a) * W + 6 * (size_t)j + b;
                            const float v = ji[a] * jj[b];
                            *e = *st ? *e + v : v;
                        }
                    *st = 1;
                }
            }
        }
    }
    free(started);
    free(jstarted);
    free(m);
    return 0;
}

void vfo_gn_update(float* T, const float* u, size_t G) {
    for (size_t g = 0; g < G; ++g) {
        const float* x = u + 6 * g;
        float U[16] = {0};
        /* column-major: U(i, j) = U[4 j + i] */
        U[4 * 1 + 0] = -x[2]; U[4 * 2 + 0] = x[1]; U[4 * 3 + 0] = x[3];
        U[4 * 0 + 1] = x[2]; U[4 * 2 + 1] = -x[0]; U[4 * 3 + 1] = x[4];
        U[4 * 0 + 2] = -x[1]; U[4 * 1 + 2] = x[0]; U[4 * 3 + 2] = x[5];
        for (int e = 0; e < 16; ++e)
            T[16 * g + e] = T[16 * g + e] - U[e];
    }
}

/* A x = b by LU with partial pivoting (the getrf / getrs pair), binary64.
 * Returns 0, or the 1-based index of the first zero pivot. */
int vfo_lu_solve(int n, const float* A, const float* b, double* x) {
    double* M = (double*)malloc(sizeof(double) * (size_t)n * (size_t)n);
    int* piv = (int*)malloc(sizeof(int) * (size_t)(n ? n : 1));
    if (!M || !piv) {
        free(M); free(piv);
        return -1;
    }
    for (size_t i = 0; i < (size_t)n * (size_t)n; ++i)
        M[i] = (double)A[i];
    for (int i = 0; i < n; ++i)
        x[i] = (double)b[i];
    int info = 0;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = fabs(M[(size_t)k * n + k]);
        for (int i = k + 1; i < n; ++i) {
            const double v = fabs(M[(size_t)i * n + k]);
            if (v > best) {
                best = v;
                p = i;
            }
        }
        piv[k] = p;
        if (best == 0.0) {
            if (!info) info = k + 1;
            continue;
        }
        if (p != k) {
            for (int j = 0; j < n; ++j) {
                const double t = M[(size_t)k * n + j];
                M[(size_t)k * n + j] = M[(size_t)p * n + j];
                M[(size_t)p * n + j] = t;
            }
            const double t = x[k];
            x[k] = x[p];
            x[p] = t;
        }
        const double d = M[(size_t)k * n + k];
        for (int i = k + 1; i < n; ++i) {
            const double l = M[(size_t)i * n + k] / d;
            M[(size_t)i * n + k] = l;
            if (l == 0.0) continue;
            for (int j = k + 1; j < n; ++j)
                M[(size_t)i * n + j] -= l * M[(size_t)k * n + j];
            x[i] -= l * x[k];
        }
    }
    if (!info) {
        for (int i = n - 1; i >= 0; --i) {
            double s = x[i];
            for (int j = i + 1; j < n; ++j)
                s -= M[(size_t)i * n + j] * x[j];
            x[i] = s / M[(size_t)i * n + i];
        }
    }
    free(M);
    free(piv);
    return info;
}
