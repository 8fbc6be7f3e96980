/*
 * vf_oracle.c -- CPU restatement of VolumeFusion's warped TSDF integration path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels.  Loaded by the
 * test suite, the smoke check and the CPU-baseline leg of bench.py; never by
 * the shipped library.
 *
 * Build with -ffp-contract=off: every float operation below is rounded on its
 * own, in the order the reference's device code evaluates it (Eigen's scalar
 * coefficient-based path, Eigen/src/Core/Redux.h:99-113 for sums of three
 * terms, c0 + (c1 + c2)).  fma() appears only where the restated algorithm
 * itself is defined with a fused multiply-add (the exponential below).
 *
 * Reference semantics that are reproduced on purpose (they are the reference):
 *  - voxel "centre" is the voxel's min corner (src/TSDFWarp.cu:8-10);
 *  - world->voxel truncates toward zero with inv = float(1.0/double(vs))
 *    (src/TSDFWarp.cu:11-15);
 *  - pixel = int(u + 0.5f), truncation (src/TSDFIntegration.cu:16-22);
 *  - sdf < -delta updates with -1 (src/TSDFIntegration.cu:52-58);
 *  - running average uses the uncapped w_old + 1 (src/TSDFIntegration.cu:63-64);
 *  - voxels with knn count 0 are never integrated on the warped path
 *    (src/TSDFIntegration.cu:197-200);
 *  - the tsdf Gaussian has no radius cutoff (src/TSDFWarp.cu:17).
 *
 * Canonical choices where the reference is a GPU race or is not pinned:
 *  - knn selection: of all candidates (node cube and dist <= 2r,
 *    src/TSDFWarp.cu:40-56) keep the K smallest by (dist, node index),
 *    slots in that order.  The reference claims slots with atomicAdd
 *    (src/TSDFWarp.cu:61-67), so its slot order is a race; its chosen set equals
 *    this one whenever a voxel has at most K candidates.
 *  - exp: the reference calls the CUDA device exp (src/TSDFWarp.cu:17).  Here
 *    e^x is evaluated by vfo_expw below (Tang's table-driven method in
 *    binary64, one final rounding to binary32); the HIP kernels evaluate the
 *    same algorithm so both agree bit for bit.  Its last bit against the
 *    reference's device exp is not pinned (that library is not available).
 *  - pow(radius, 2) is radius*radius in float (the device overload of pow for
 *    (float, int)).
 */
#include "vf_oracle.h"
#include "vf_mc_tables.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* scalar helpers                                                           */
/* ------------------------------------------------------------------------ */

/* float -> int32 truncation toward zero, saturating, NaN -> 0 (the gfx9
 * v_cvt_i32_f32 behaviour; in-range values match a C cast exactly). */
int32_t vfo_cvt_i32_rz(float f) {
    if (f != f)
        return 0;
    if (f >= 2147483648.0f)
        return INT32_MAX;
    if (f < -2147483648.0f)
        return INT32_MIN;
    return (int32_t)f;
}

/* 2^(j/32), j = 0..31, correctly rounded to binary64. */
static const double kExp2Table[32] = {
    0x1.0000000000000p+0, 0x1.059b0d3158574p+0, 0x1.0b5586cf9890fp+0, 0x1.11301d0125b51p+0,
    0x1.172b83c7d517bp+0, 0x1.1d4873168b9aap+0, 0x1.2387a6e756238p+0, 0x1.29e9df51fdee1p+0,
    0x1.306fe0a31b715p+0, 0x1.371a7373aa9cbp+0, 0x1.3dea64c123422p+0, 0x1.44e086061892dp+0,
    0x1.4bfdad5362a27p+0, 0x1.5342b569d4f82p+0, 0x1.5ab07dd485429p+0, 0x1.6247eb03a5585p+0,
    0x1.6a09e667f3bcdp+0, 0x1.71f75e8ec5f74p+0, 0x1.7a11473eb0187p+0, 0x1.82589994cce13p+0,
    0x1.8ace5422aa0dbp+0, 0x1.93737b0cdc5e5p+0, 0x1.9c49182a3f090p+0, 0x1.a5503b23e255dp+0,
    0x1.ae89f995ad3adp+0, 0x1.b7f76f2fb5e47p+0, 0x1.c199bdd85529cp+0, 0x1.cb720dcef9069p+0,
    0x1.d5818dcfba487p+0, 0x1.dfc97337b9b5fp+0, 0x1.ea4afa2a490dap+0, 0x1.f50765b6e4540p+0};

/* e^x evaluated in binary64 and rounded once to binary32, Tang's
 * table-driven method (ACM TOMS 15(2), 1989): k = rint(32 x / ln2),
 * r = x - k ln2/32 by a two-constant Cody-Waite reduction (fdlibm's ln2 split
 * divided by 32), |r| <= ln2/64; e^r by its degree-5 Taylor polynomial in
 * Horner form with fused multiply-adds (truncation < 2^-48 relative);
 * e^x = 2^(k>>5) * 2^((k&31)/32) * e^r.  Inputs below -87 return 0. */
float vfo_expw(float xf) {
    if (xf != xf)
        return xf;
    if (xf < -87.0f)
        return 0.0f;
    if (xf > 88.5f)
        return INFINITY;
    const double x = (double)xf;
    const double kd = rint(x * 0x1.71547652b82fep+5);
    const int k = (int)kd;
    double r = fma(-kd, 0x1.62e42fee00000p-6, x);
    r = fma(-kd, 0x1.a39ef35793c76p-38, r);
    double p = fma(0x1.1111111111111p-7, r, 0x1.5555555555555p-5);
    p = fma(p, r, 0x1.5555555555555p-3);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    const double t = p * kExp2Table[k & 31];
    return (float)ldexp(t, k >> 5);
}

/* glibc's binary64 exp rounded to binary32: the host-compiled meaning of the
 * reference's exp; used only to bound vfo_expw's error in the tests. */
float vfo_exp_libm(float x) { return (float)exp((double)x); }

/* detail::weight, src/TSDFWarp.cu:17: exp(-(d*d) / (2*pow(r,2))), no cutoff. */
float vfo_weight_tsdf(float radius, float dist) {
    const float r2 = radius * radius;
    const float den = 2.0f * r2;
    const float num = -(dist * dist);
    return vfo_expw(num / den);
}

/* MotionGraph::weight, src/MotionGraph.cpp:39-41, as compiled for the host:
 * pow(float,int) promotes to double, so the quotient and exp are binary64 and
 * the result is rounded to float on return.  Has the d >= r cutoff. */
float vfo_weight_mesh_host(float radius, float dist) {
    if (dist >= radius)
        return 0.0f;
    const double r = (double)radius;
    const double num = -(double)(dist * dist);
    return (float)exp(num / (2.0 * pow(r, 2.0)));
}

/* a/b from a precomputed correctly rounded reciprocal (Markstein 1990):
 * q0 = RN(a*y), r = a - q0*b (exact by fma), q = RN(q0 + r*y).  Used by the
 * tests to check that this equals IEEE division on the weight domain. */
float vfo_div_markstein(float a, float b, float y) {
    const float q0 = a * y;
    const float r = fmaf(-q0, b, a);
    return fmaf(r, y, q0);
}

/* Number of floats a in [a_lo, a_hi] (bit patterns, positive) for which
 * vfo_div_markstein(a, b, RN(1/b)) differs from the IEEE quotient a/b. */
int64_t vfo_markstein_mismatches(float b, uint32_t a_lo_bits, uint32_t a_hi_bits) {
    const float y = 1.0f / b;
    int64_t bad = 0;
    for (uint32_t u = a_lo_bits; u <= a_hi_bits; ++u) {
        float a;
        memcpy(&a, &u, 4);
        const float q = a / b;
        const float m = vfo_div_markstein(a, b, y);
        uint32_t qb, mb;
        memcpy(&qb, &q, 4);
        memcpy(&mb, &m, 4);
        bad += qb != mb;
        if (u == 0xffffffffu) break;
    }
    return bad;
}

/* Eigen's 3x3 inverse (Eigen/src/LU/InverseImpl.h:126-170), as the reference
 * calls it on the host (src/TSDFIntegration.cu:247, 296).  Column-major. */
#define M3(m, i, j) ((m)[(j)*3 + (i)])
static float cof3(const float* m, int i, int j) {
    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    const float a = M3(m, i1, j1) * M3(m, i2, j2);
    const float b = M3(m, i1, j2) * M3(m, i2, j1);
    return a - b;
}
void vfo_mat3_inverse(const float* m, float* out) {
    const float c0 = cof3(m, 0, 0), c1 = cof3(m, 1, 0), c2 = cof3(m, 2, 0);
    const float det = c0 * M3(m, 0, 0) + (c1 * M3(m, 1, 0) + c2 * M3(m, 2, 0));
    const float invdet = 1.0f / det;
    M3(out, 0, 0) = c0 * invdet;
    M3(out, 0, 1) = c1 * invdet;
    M3(out, 0, 2) = c2 * invdet;
    M3(out, 1, 0) = cof3(m, 0, 1) * invdet;
    M3(out, 1, 1) = cof3(m, 1, 1) * invdet;
    M3(out, 1, 2) = cof3(m, 2, 1) * invdet;
    M3(out, 2, 0) = cof3(m, 0, 2) * invdet;
    M3(out, 2, 1) = cof3(m, 1, 2) * invdet;
    M3(out, 2, 2) = cof3(m, 2, 2) * invdet;
}

/* Mat3f * Vec3f, Eigen device order: row i = a0 + (a1 + a2). */
static void mat3_vec(const float* m, const float* v, float* o) {
    for (int i = 0; i < 3; ++i) {
        const float a0 = M3(m, i, 0) * v[0];
        const float a1 = M3(m, i, 1) * v[1];
        const float a2 = M3(m, i, 2) * v[2];
        o[i] = a0 + (a1 + a2);
    }
}

/* (Mat4f * p.homogeneous()).head(3): Eigen evaluates the homogeneous product
 * as leftCols<3>() * p plus the last column
 * (Eigen/src/Geometry/Homogeneous.h:268-276), each row a0 + (a1 + a2) + m3. */
static void mat4_apply(const float* T, const float* p, float* o) {
    for (int i = 0; i < 3; ++i) {
        const float a0 = T[0 * 4 + i] * p[0];
        const float a1 = T[1 * 4 + i] * p[1];
        const float a2 = T[2 * 4 + i] * p[2];
        const float s = a0 + (a1 + a2);
        o[i] = s + T[3 * 4 + i];
    }
}

typedef struct {
    int32_t dim[3];
    float vs[3];
    float inv[3];   /* float(1.0 / double(vs)), src/TSDFWarp.cu:12 */
    float halfs[3]; /* size * 0.5f == 0.5f * size (exact)          */
    size_t n;
} geom_t;

static void geom_init(geom_t* g, const vfo_vol* v) {
    for (int a = 0; a < 3; ++a) {
        g->dim[a] = v->dim[a];
        g->vs[a] = v->vs[a];
        g->inv[a] = (float)(1.0 / (double)v->vs[a]);
        g->halfs[a] = v->size[a] * 0.5f;
    }
    g->n = (size_t)v->dim[0] * (size_t)v->dim[1] * (size_t)v->dim[2];
}

/* voxelToWorld, src/TSDFWarp.cu:8-10 (== src/TSDFIntegration.cu:6-8). */
static inline void voxel_to_world(const geom_t* g, int64_t x, int64_t y, int64_t z, float* p) {
    const float vx = (float)x * g->vs[0];
    const float vy = (float)y * g->vs[1];
    const float vz = (float)z * g->vs[2];
    p[0] = vx - g->halfs[0];
    p[1] = vy - g->halfs[1];
    p[2] = vz - g->halfs[2];
}

/* worldToVoxel, src/TSDFWarp.cu:11-15 (== src/TSDFIntegration.cu:10-14). */
static inline void world_to_voxel(const geom_t* g, const float* p, int32_t* v) {
    for (int a = 0; a < 3; ++a) {
        const float s = p[a] + g->halfs[a];
        v[a] = vfo_cvt_i32_rz(s * g->inv[a]);
    }
}

static inline int in_volume(const geom_t* g, int64_t x, int64_t y, int64_t z) {
    return x >= 0 && y >= 0 && z >= 0 && x < g->dim[0] && y < g->dim[1] && z < g->dim[2];
}

static inline size_t lin(const geom_t* g, int64_t x, int64_t y, int64_t z) {
    return ((size_t)z * (size_t)g->dim[1] + (size_t)y) * (size_t)g->dim[0] + (size_t)x;
}

/* strict (dist, index) order of the canonical knn selection */
static inline int knn_less(float da, uint32_t ia, float db, uint32_t ib) {
    return da < db || (da == db && ia < ib);
}

/* insert (d, j) into an ascending list of length *n (capacity K) */
static inline void knn_insert(int K, float* dl, uint32_t* il, int* n, float d, uint32_t j) {
    int pos = *n;
    if (pos == K) {
        if (!knn_less(d, j, dl[K - 1], il[K - 1]))
            return;
        pos = K - 1;
    } else {
        (*n)++;
    }
    while (pos > 0 && knn_less(d, j, dl[pos - 1], il[pos - 1])) {
        dl[pos] = dl[pos - 1];
        il[pos] = il[pos - 1];
        --pos;
    }
    dl[pos] = d;
    il[pos] = j;
}

/* ------------------------------------------------------------------------ */
/* reference kernels                                                        */
/* ------------------------------------------------------------------------ */

/* translateKernel, src/VertexManipulation.cu:73-79: mesh[i] += t. */
void vfo_translate(float* pts, size_t n, const float* t) {
    for (size_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a)
            pts[3 * i + a] = pts[3 * i + a] + t[a];
}

/* generateTsdfVerticesKernel, src/TSDFWarp.cu:131-145. */
void vfo_generate_vertices(float* verts, const vfo_vol* vol) {
    geom_t g;
    geom_init(&g, vol);
    for (int64_t z = 0; z < g.dim[2]; ++z)
        for (int64_t y = 0; y < g.dim[1]; ++y)
            for (int64_t x = 0; x < g.dim[0]; ++x)
                voxel_to_world(&g, x, y, z, verts + 3 * lin(&g, x, y, z));
}

/* Candidate test of tsdfKNNGraph for one (node, voxel) pair,
 * src/TSDFWarp.cu:52-56: d = |voxelToWorld(v) - node|, reject d > 2r. */
static inline int knn_candidate(const geom_t* g, const float* m, float r, int64_t x, int64_t y, int64_t z, float* d) {
    float p[3];
    voxel_to_world(g, x, y, z, p);
    const float dx = p[0] - m[0], dy = p[1] - m[1], dz = p[2] - m[2];
    const float s = dx * dx + (dy * dy + dz * dz);
    *d = sqrtf(s);
    return !(*d > r * 2.0f);
}

/* Lower corner of node j's cube, src/TSDFWarp.cu:46-48:
 * worldToVoxel(node) - (cube - 1) / 2 (integer division). */
static inline void node_cube_lo(const geom_t* g, const float* m, const int32_t* cube, int64_t* lo) {
    int32_t nv[3];
    world_to_voxel(g, m, nv);
    for (int a = 0; a < 3; ++a)
        lo[a] = (int64_t)nv[a] - (int64_t)((cube[a] - 1) / 2);
}

/* tsdfKNNGraph, src/TSDFWarp.cu:23-68, node-major like the reference launch
 * (one (node, cube offset) pair per thread), canonical selection.  Counts must
 * be zero on entry, as every reference caller makes them
 * (src/VolumeFusion.cu:358).  Voxels without candidates are left untouched.
 * Returns the number of voxels that had more than K candidates. */
int64_t vfo_knn_graph(int K, uint32_t* idx, float* dist, int32_t* cnt, const float* nodes, const float* radii,
                      size_t G, const vfo_vol* vol, const int32_t* cube) {
    geom_t g;
    geom_init(&g, vol);
    int64_t overflow = 0;
    if (cube[0] < 1 || cube[1] < 1 || cube[2] < 1)
        return 0;
    int32_t* seen = (int32_t*)calloc(g.n, sizeof(int32_t)); /* candidates seen per voxel */
    for (size_t j = 0; j < G; ++j) {
        const float* m = nodes + 3 * j;
        int64_t lo[3];
        node_cube_lo(&g, m, cube, lo);
        for (int64_t oz = 0; oz < cube[2]; ++oz)
            for (int64_t oy = 0; oy < cube[1]; ++oy)
                for (int64_t ox = 0; ox < cube[0]; ++ox) {
                    const int64_t x = lo[0] + ox, y = lo[1] + oy, z = lo[2] + oz;
                    if (!in_volume(&g, x, y, z))
                        continue;
                    float d;
                    if (!knn_candidate(&g, m, radii[j], x, y, z, &d))
                        continue;
                    const size_t v = lin(&g, x, y, z);
                    if (seen[v] == 0)
                        cnt[v] = 0;
                    seen[v]++;
                    int n = cnt[v];
                    knn_insert(K, dist + (size_t)K * v, idx + (size_t)K * v, &n, d, (uint32_t)j);
                    cnt[v] = n;
                }
    }
    for (size_t v = 0; v < g.n; ++v)
        overflow += seen[v] > K;
    free(seen);
    return overflow;
}

/* computeTsdfWarpWeightsKernel, src/TSDFWarp.cu:88-116. */
void vfo_warp_weights(int K, float* w, const float* radii, const uint32_t* idx, const float* dist, const int32_t* cnt,
                      size_t N) {
    for (size_t v = 0; v < N; ++v) {
        int c = cnt[v];
        if (!c)
            continue;
        if (c > K)
            c = K;
        float* wv = w + (size_t)K * v;
        float sum = 0.0f;
        for (int s = 0; s < K; ++s)
            wv[s] = 0.0f;
        for (int s = 0; s < c; ++s) {
            wv[s] = vfo_weight_tsdf(radii[idx[(size_t)K * v + s]], dist[(size_t)K * v + s]);
            sum = sum + wv[s];
        }
        if (sum != 0.0f)
            for (int s = 0; s < K; ++s)
                wv[s] = wv[s] / sum;
    }
}

/* Eigen isZero() with the float dummy precision 1e-5 (src/TSDFWarp.cu:180). */
static inline int weights_zero(const float* w, int K) {
    for (int s = 0; s < K; ++s)
        if (!(fabsf(w[s]) <= 1e-5f))
            return 0;
    return 1;
}

/* warpTsdfKernel, src/TSDFWarp.cu:158-190: p' = sum_j w_j (T_j [p;1]).head(3). */
void vfo_warp_tsdf(int K, float* verts, const float* T, const uint32_t* idx, const float* w, const int32_t* cnt,
                   size_t N) {
    for (size_t v = 0; v < N; ++v) {
        int c = cnt[v];
        if (!c)
            continue;
        if (c > K)
            c = K;
        const float* wv = w + (size_t)K * v;
        if (weights_zero(wv, K))
            continue;
        float acc[3] = {0.0f, 0.0f, 0.0f};
        const float* p = verts + 3 * v;
        for (int s = 0; s < c; ++s) {
            float t[3];
            mat4_apply(T + 16 * (size_t)idx[(size_t)K * v + s], p, t);
            for (int a = 0; a < 3; ++a) {
                const float wt = wv[s] * t[a];
                acc[a] = acc[a] + wt;
            }
        }
        verts[3 * v + 0] = acc[0];
        verts[3 * v + 1] = acc[1];
        verts[3 * v + 2] = acc[2];
    }
}

/* ------------------------------------------------------------------------ */
/* dual-quaternion blending (second blend mode; not in the reference)        */
/* ------------------------------------------------------------------------ */
/* The reference blends Mat4f node transforms linearly (LBS,
 * src/TSDFWarp.cu:183-189).  DQB (Kavan, Collins, Zara, O'Sullivan, ACM TOG
 * 27(4), 2008) blends each node's rigid part as a unit dual quaternion.  The
 * float32 operation order below is the one the device evaluates (no fused
 * multiply-add); the binary64 versions are the accuracy reference.  Dual
 * quaternions are 8 floats: real (x, y, z, w), dual (x, y, z, w). */

/* Rotation part of a column-major Mat4f (Shepperd's method, largest
 * diagonal branch), normalised, sign chosen with w >= 0; dual part
 * 0.5 (0, t) q. */
void vfo_mat4_to_dq(const float* M, float* dq) {
    const float r00 = M[0], r10 = M[1], r20 = M[2];
    const float r01 = M[4], r11 = M[5], r21 = M[6];
    const float r02 = M[8], r12 = M[9], r22 = M[10];
    const float tr = (r00 + r11) + r22;
    float x, y, z, w;
    if (tr > 0.0f) {
        const float s = sqrtf(tr + 1.0f) * 2.0f;
        w = 0.25f * s;
        x = (r21 - r12) / s;
        y = (r02 - r20) / s;
        z = (r10 - r01) / s;
    } else if (r00 > r11 && r00 > r22) {
        const float s = sqrtf(((1.0f + r00) - r11) - r22) * 2.0f;
        w = (r21 - r12) / s;
        x = 0.25f * s;
        y = (r01 + r10) / s;
        z = (r02 + r20) / s;
    } else if (r11 > r22) {
        const float s = sqrtf(((1.0f + r11) - r00) - r22) * 2.0f;
        w = (r02 - r20) / s;
        x = (r01 + r10) / s;
        y = 0.25f * s;
        z = (r12 + r21) / s;
    } else {
        const float s = sqrtf(((1.0f + r22) - r00) - r11) * 2.0f;
        w = (r10 - r01) / s;
        x = (r02 + r20) / s;
        y = (r12 + r21) / s;
        z = 0.25f * s;
    }
    const float nn = sqrtf(((x * x + y * y) + z * z) + w * w);
    x = x / nn;
    y = y / nn;
    z = z / nn;
    w = w / nn;
    if (w < 0.0f) {
        x = -x;
        y = -y;
        z = -z;
        w = -w;
    }
    const float tx = M[12], ty = M[13], tz = M[14];
    dq[0] = x;
    dq[1] = y;
    dq[2] = z;
    dq[3] = w;
    /* (0, t) q = (w t + t x q_v, -t . q_v) */
    dq[4] = 0.5f * (w * tx + (ty * z - tz * y));
    dq[5] = 0.5f * (w * ty + (tz * x - tx * z));
    dq[6] = 0.5f * (w * tz + (tx * y - ty * x));
    dq[7] = 0.5f * -((tx * x + ty * y) + tz * z);
}

/* Blend n dual quaternions (dq[s]: 8 floats each) with weights w[s] (the
 * normalised knn weights), each aligned to the hemisphere of the first, and
 * transform p:  b = sum w_s sign_s dq_s;  (r, d) = b / |b_real|;
 * p' = p + 2 r_v x (r_v x p + r_w p) + 2 (r_w d_v - d_w r_v + r_v x d_v). */
void vfo_dqb_apply(const float* const* dq, const float* w, int n, const float* p, float* out) {
    float br[4] = {0.0f, 0.0f, 0.0f, 0.0f}, bd[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    const float* f = dq[0];
    for (int s = 0; s < n; ++s) {
        const float* q = dq[s];
        const float dot = ((q[0] * f[0] + q[1] * f[1]) + q[2] * f[2]) + q[3] * f[3];
        const float ws = dot < 0.0f ? -w[s] : w[s];
        for (int a = 0; a < 4; ++a) {
            br[a] = br[a] + ws * q[a];
            bd[a] = bd[a] + ws * q[4 + a];
        }
    }
    const float len = sqrtf(((br[0] * br[0] + br[1] * br[1]) + br[2] * br[2]) + br[3] * br[3]);
    if (!(len > 0.0f)) {
        out[0] = p[0];
        out[1] = p[1];
        out[2] = p[2];
        return;
    }
    const float il = 1.0f / len;
    float r[4], d[4];
    for (int a = 0; a < 4; ++a) {
        r[a] = br[a] * il;
        d[a] = bd[a] * il;
    }
    /* rotation: t1 = r_v x p; t2 = t1 + r_w p; t3 = r_v x t2; p + 2 t3 */
    const float t1x = r[1] * p[2] - r[2] * p[1];
    const float t1y = r[2] * p[0] - r[0] * p[2];
    const float t1z = r[0] * p[1] - r[1] * p[0];
    const float t2x = t1x + r[3] * p[0];
    const float t2y = t1y + r[3] * p[1];
    const float t2z = t1z + r[3] * p[2];
    const float t3x = r[1] * t2z - r[2] * t2y;
    const float t3y = r[2] * t2x - r[0] * t2z;
    const float t3z = r[0] * t2y - r[1] * t2x;
    /* translation: 2 ((r_w d_v - d_w r_v) + r_v x d_v) */
    const float ux = r[1] * d[2] - r[2] * d[1];
    const float uy = r[2] * d[0] - r[0] * d[2];
    const float uz = r[0] * d[1] - r[1] * d[0];
    const float trx = 2.0f * ((r[3] * d[0] - d[3] * r[0]) + ux);
    const float try_ = 2.0f * ((r[3] * d[1] - d[3] * r[1]) + uy);
    const float trz = 2.0f * ((r[3] * d[2] - d[3] * r[2]) + uz);
    out[0] = (p[0] + 2.0f * t3x) + trx;
    out[1] = (p[1] + 2.0f * t3y) + try_;
    out[2] = (p[2] + 2.0f * t3z) + trz;
}

/* The same in binary64 (conversion and blend), the accuracy reference. */
void vfo_mat4_to_dq_f64(const float* Mf, double* dq) {
    double M[16];
    for (int i = 0; i < 16; ++i)
        M[i] = (double)Mf[i];
    const double r00 = M[0], r10 = M[1], r20 = M[2], r01 = M[4], r11 = M[5], r21 = M[6];
    const double r02 = M[8], r12 = M[9], r22 = M[10];
    const double tr = r00 + r11 + r22;
    double x, y, z, w;
    if (tr > 0.0) {
        const double s = sqrt(tr + 1.0) * 2.0;
        w = 0.25 * s; x = (r21 - r12) / s; y = (r02 - r20) / s; z = (r10 - r01) / s;
    } else if (r00 > r11 && r00 > r22) {
        const double s = sqrt(1.0 + r00 - r11 - r22) * 2.0;
        w = (r21 - r12) / s; x = 0.25 * s; y = (r01 + r10) / s; z = (r02 + r20) / s;
    } else if (r11 > r22) {
        const double s = sqrt(1.0 + r11 - r00 - r22) * 2.0;
        w = (r02 - r20) / s; x = (r01 + r10) / s; y = 0.25 * s; z = (r12 + r21) / s;
    } else {
        const double s = sqrt(1.0 + r22 - r00 - r11) * 2.0;
        w = (r10 - r01) / s; x = (r02 + r20) / s; y = (r12 + r21) / s; z = 0.25 * s;
    }
    const double nn = sqrt(x * x + y * y + z * z + w * w);
    x /= nn; y /= nn; z /= nn; w /= nn;
    if (w < 0.0) { x = -x; y = -y; z = -z; w = -w; }
    const double tx = M[12], ty = M[13], tz = M[14];
    dq[0] = x; dq[1] = y; dq[2] = z; dq[3] = w;
    dq[4] = 0.5 * (w * tx + ty * z - tz * y);
    dq[5] = 0.5 * (w * ty + tz * x - tx * z);
    dq[6] = 0.5 * (w * tz + tx * y - ty * x);
    dq[7] = -0.5 * (tx * x + ty * y + tz * z);
}

void vfo_dqb_apply_f64(const double* const* dq, const double* w, int n, const double* p, double* out) {
    double br[4] = {0, 0, 0, 0}, bd[4] = {0, 0, 0, 0};
    const double* f = dq[0];
    for (int s = 0; s < n; ++s) {
        const double* q = dq[s];
        const double dot = q[0] * f[0] + q[1] * f[1] + q[2] * f[2] + q[3] * f[3];
        const double ws = dot < 0.0 ? -w[s] : w[s];
        for (int a = 0; a < 4; ++a) {
            br[a] += ws * q[a];
            bd[a] += ws * q[4 + a];
        }
    }
    const double len = sqrt(br[0] * br[0] + br[1] * br[1] + br[2] * br[2] + br[3] * br[3]);
    if (!(len > 0.0)) {
        out[0] = p[0]; out[1] = p[1]; out[2] = p[2];
        return;
    }
    double r[4], d[4];
    for (int a = 0; a < 4; ++a) {
        r[a] = br[a] / len;
        d[a] = bd[a] / len;
    }
    const double t1[3] = {r[1] * p[2] - r[2] * p[1], r[2] * p[0] - r[0] * p[2], r[0] * p[1] - r[1] * p[0]};
    const double t2[3] = {t1[0] + r[3] * p[0], t1[1] + r[3] * p[1], t1[2] + r[3] * p[2]};
    const double t3[3] = {r[1] * t2[2] - r[2] * t2[1], r[2] * t2[0] - r[0] * t2[2], r[0] * t2[1] - r[1] * t2[0]};
    const double u[3] = {r[1] * d[2] - r[2] * d[1], r[2] * d[0] - r[0] * d[2], r[0] * d[1] - r[1] * d[0]};
    for (int a = 0; a < 3; ++a)
        out[a] = p[a] + 2.0 * t3[a] + 2.0 * (r[3] * d[a] - d[3] * r[a] + u[a]);
}

/* warpTsdfKernel's loop (src/TSDFWarp.cu:158-190) with DQB in place of the
 * linear blend; DQ holds each node's dual quaternion (8 floats). */
void vfo_warp_tsdf_dqb(int K, float* verts, const float* DQ, const uint32_t* idx, const float* w, const int32_t* cnt,
                       size_t N) {
    for (size_t v = 0; v < N; ++v) {
        int c = cnt[v];
        if (!c)
            continue;
        if (c > K)
            c = K;
        const float* wv = w + (size_t)K * v;
        if (weights_zero(wv, K))
            continue;
        const float* q[64];
        for (int s = 0; s < c; ++s)
            q[s] = DQ + 8 * (size_t)idx[(size_t)K * v + s];
        float o[3];
        vfo_dqb_apply(q, wv, c, verts + 3 * v, o);
        verts[3 * v + 0] = o[0];
        verts[3 * v + 1] = o[1];
        verts[3 * v + 2] = o[2];
    }
}

/* Target voxel of a warped vertex, src/TSDFIntegration.cu:272-279 and
 * :210-216: u = worldToVoxel(Rinv * (p' - t)).  Returns 0 when outside. */
static inline int warped_target(const geom_t* g, const float* Rinv, const float* t, const float* pw, size_t* u) {
    float q[3], o[3];
    for (int a = 0; a < 3; ++a)
        q[a] = pw[a] - t[a];
    mat3_vec(Rinv, q, o);
    int32_t v[3];
    world_to_voxel(g, o, v);
    if (v[0] < 0 || v[1] < 0 || v[2] < 0 || v[0] >= g->dim[0] || v[1] >= g->dim[1] || v[2] >= g->dim[2])
        return 0;
    *u = lin(g, v[0], v[1], v[2]);
    return 1;
}

/* countCollisionsPerVoxelKernel, src/TSDFIntegration.cu:252-283 (R is the
 * volume rotation; its inverse is taken on the host, :296). */
void vfo_count_collisions(int32_t* col, const float* verts, const int32_t* cnt, const vfo_vol* vol, const float* R,
                          const float* t) {
    geom_t g;
    geom_init(&g, vol);
    float Rinv[9];
    vfo_mat3_inverse(R, Rinv);
    for (size_t v = 0; v < g.n; ++v) {
        if (!cnt[v])
            continue;
        size_t u;
        if (warped_target(&g, Rinv, t, verts + 3 * v, &u))
            col[u] += 1;
    }
}

/* project3dToFrame + updateTsdf, src/TSDFIntegration.cu:16-67.
 * Kin is the column-major 3x3 intrinsics.  Returns the pixel index used when
 * the voxel is updated, else -1. */
static inline int64_t update_tsdf(float* tsdf, float* wt, const float* p, size_t i, const float* depth, int W, int H,
                                  float delta, const float* Kin, float maxw) {
    if (p[2] < 1.e-4f)
        return -1;
    const float zInv = 1.0f / p[2];
    const float ux = M3(Kin, 0, 0) * p[0];
    const float uf = ux * zInv + M3(Kin, 0, 2);
    const float vy = M3(Kin, 1, 1) * p[1];
    const float vf = vy * zInv + M3(Kin, 1, 2);
    const int32_t px = vfo_cvt_i32_rz(uf + 0.5f);
    const int32_t py = vfo_cvt_i32_rz(vf + 0.5f);
    if (px < 0 || px >= W || py < 0 || py >= H)
        return -1;
    const size_t pi = (size_t)py * (size_t)W + (size_t)px;
    const float d = depth[pi];
    if (isinf(d) || d < 1.e-4f)
        return -1;
    const float sdf = p[2] - d;
    if (sdf > delta)
        return -1;
    const float tn = (sdf < -delta) ? -1.0f : sdf / delta;
    const float wo = wt[i];
    const float to = tsdf[i];
    const float a = to * wo;
    const float b = tn * 1.0f;
    const float wn = wo + 1.0f;
    tsdf[i] = (a + b) / wn;
    wt[i] = (maxw < wn) ? maxw : wn;
    return (int64_t)pi;
}

/* integrateWarpedTSDFFromFrameKernel, src/TSDFIntegration.cu:173-226. */
int64_t vfo_integrate_warped(float* tsdf, float* wt, const float* verts, const int32_t* col, const int32_t* cnt,
                             const float* depth, int W, int H, float delta, const vfo_vol* vol, const float* R,
                             const float* t, const float* Kin, float maxw) {
    geom_t g;
    geom_init(&g, vol);
    float Rinv[9];
    vfo_mat3_inverse(R, Rinv);
    int64_t updated = 0;
    for (size_t v = 0; v < g.n; ++v) {
        if (!cnt[v])
            continue;
        size_t u;
        if (!warped_target(&g, Rinv, t, verts + 3 * v, &u))
            continue;
        if (col[u] > 1)
            continue;
        updated += update_tsdf(tsdf, wt, verts + 3 * v, v, depth, W, H, delta, Kin, maxw) >= 0;
    }
    return updated;
}

/* integrateTSDFFromFrameKernel, src/TSDFIntegration.cu:69-96:
 * p = R * voxelToWorld(v) + t, then updateTsdf. */
int64_t vfo_integrate_plain(float* tsdf, float* wt, const float* depth, int W, int H, float delta,
                            const vfo_vol* vol, const float* R, const float* t, const float* Kin, float maxw) {
    geom_t g;
    geom_init(&g, vol);
    int64_t updated = 0;
    for (int64_t z = 0; z < g.dim[2]; ++z)
        for (int64_t y = 0; y < g.dim[1]; ++y)
            for (int64_t x = 0; x < g.dim[0]; ++x) {
                float p[3], q[3];
                voxel_to_world(&g, x, y, z, p);
                mat3_vec(R, p, q);
                for (int a = 0; a < 3; ++a)
                    q[a] = q[a] + t[a];
                updated += update_tsdf(tsdf, wt, q, lin(&g, x, y, z), depth, W, H, delta, Kin, maxw) >= 0;
            }
    return updated;
}

/* integrateTSDFFromCameraFrameKernel, src/TSDFIntegration.cu:118-148:
 * p = Rw2c * (R * voxelToWorld(v) + t) + tw2c, then updateTsdf. */
int64_t vfo_integrate_camera_frame(float* tsdf, float* wt, const float* depth, int W, int H, float delta,
                                   const vfo_vol* vol, const float* Rw2c, const float* tw2c, const float* R,
                                   const float* t, const float* Kin, float maxw) {
    geom_t g;
    geom_init(&g, vol);
    int64_t updated = 0;
    for (int64_t z = 0; z < g.dim[2]; ++z)
        for (int64_t y = 0; y < g.dim[1]; ++y)
            for (int64_t x = 0; x < g.dim[0]; ++x) {
                float p[3], q[3], r[3];
                voxel_to_world(&g, x, y, z, p);
                mat3_vec(R, p, q);
                for (int a = 0; a < 3; ++a)
                    q[a] = q[a] + t[a];
                mat3_vec(Rw2c, q, r);
                for (int a = 0; a < 3; ++a)
                    r[a] = r[a] + tw2c[a];
                updated += update_tsdf(tsdf, wt, r, lin(&g, x, y, z), depth, W, H, delta, Kin, maxw) >= 0;
            }
    return updated;
}

/* TSDFVolume::integrate, src/TSDFVolume.cpp:55-136 (tsdf and weights only;
 * the colour branch is not restated).  pose is a column-major 4x4.  Differs
 * from the GPU kernel in its validity tests (z <= 0, depth == 0 or NaN), in
 * truncation by clamp * (1/delta) and in having no weight cap. */
int64_t vfo_tsdfvolume_integrate(float* tsdf, float* wt, const float* depth, int W, int H, float delta,
                                 const vfo_vol* vol, const float* pose, const float* Kin) {
    geom_t g;
    geom_init(&g, vol);
    float R[9], t[3];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j)
            M3(R, i, j) = pose[j * 4 + i];
        t[i] = pose[3 * 4 + i];
    }
    const float deltaInv = 1.0f / delta;
    int64_t updated = 0;
    for (int64_t z = 0; z < g.dim[2]; ++z)
        for (int64_t y = 0; y < g.dim[1]; ++y)
            for (int64_t x = 0; x < g.dim[0]; ++x) {
                const size_t off = lin(&g, x, y, z);
                float pt[3], q[3];
                voxel_to_world(&g, x, y, z, pt);
                mat3_vec(R, pt, q);
                for (int a = 0; a < 3; ++a)
                    q[a] = q[a] + t[a];
                if (q[2] <= 0.0f)
                    continue;
                const float zInv = 1.0f / q[2];
                const float ux = M3(Kin, 0, 0) * q[0];
                const float uf = ux * zInv + M3(Kin, 0, 2);
                const float vy = M3(Kin, 1, 1) * q[1];
                const float vf = vy * zInv + M3(Kin, 1, 2);
                const int32_t u = vfo_cvt_i32_rz(uf + 0.5f);
                const int32_t v = vfo_cvt_i32_rz(vf + 0.5f);
                if (!(u >= 0 && u < W && v >= 0 && v < H))
                    continue;
                const float d = depth[(size_t)v * (size_t)W + (size_t)u];
                if (d == 0.0f || isnan(d))
                    continue;
                const float sdf = q[2] - d;
                float tv = sdf;
                if (tv > delta)
                    tv = delta;
                else if (tv < -delta)
                    tv = -delta;
                tv = tv * deltaInv;
                float w = 0.0f;
                if (sdf <= -delta)
                    w = 1.0f;
                else if (sdf <= delta)
                    w = 1.0f;
                if (w > 0.0f) {
                    const float wo = wt[off];
                    const float to = tsdf[off];
                    const float wn = wo + w;
                    const float a = to * wo;
                    const float b = tv * w;
                    tsdf[off] = (a + b) / wn;
                    wt[off] = wn;
                    updated++;
                }
            }
    return updated;
}

/* ------------------------------------------------------------------------ */
/* mesh points: knn, weights, warp (the device kernels of the point path)    */
/* ------------------------------------------------------------------------ */

/* K nearest nodes of each point, MotionGraph::knnSearch<K>
 * (include/af/MotionGraph.h:50-70; nanoflann's float squared distance
 * ((dx^2 + dy^2) + dz^2), then sqrt).  Canonical order (distance, node index)
 * where nanoflann keeps its tree's visit order among equal distances.  With
 * exclude_closest the first result is dropped (node-to-node knn,
 * src/WarpField.cu:31).  Missing slots: index UINT32_MAX, distance FLT_MAX
 * (include/af/MotionGraph.h:66-69).  idx/dist are [n][K]. */
void vfo_mesh_knn(int K, uint32_t* idx, float* dist, const float* pts, size_t n, const float* nodes, size_t G,
                  int exclude_closest) {
    const int KK = K + (exclude_closest ? 1 : 0);
    float dl[65];
    uint32_t il[65];
    for (size_t i = 0; i < n; ++i) {
        int c = 0;
        for (size_t j = 0; j < G; ++j) {
            const float dx = pts[3 * i + 0] - nodes[3 * j + 0];
            const float dy = pts[3 * i + 1] - nodes[3 * j + 1];
            const float dz = pts[3 * i + 2] - nodes[3 * j + 2];
            const float s2 = (dx * dx + dy * dy) + dz * dz;
            knn_insert(KK, dl, il, &c, sqrtf(s2), (uint32_t)j);
        }
        const int off = exclude_closest ? 1 : 0;
        for (int s = 0; s < K; ++s) {
            const int t = s + off;
            idx[(size_t)K * i + s] = t < c ? il[t] : UINT32_MAX;
            dist[(size_t)K * i + s] = t < c ? dl[t] : FLT_MAX;
        }
    }
}

/* computeWarpWeightsKernel, src/VertexManipulation.cu:105-136: every one of
 * the K slots, Gaussian with the d >= r cutoff (device float semantics, the
 * exponential as vfo_expw), normalised when the sum is not 0.  A slot whose
 * distance is FLT_MAX (a missing neighbour) weighs 0 without reading the
 * radius; indices >= G are treated the same way (the reference would read
 * out of bounds). */
void vfo_compute_warp_weights(int K, float* w, const float* radii, const uint32_t* idx, const float* dist, size_t n,
                              size_t G) {
    for (size_t i = 0; i < n; ++i) {
        float sum = 0.0f;
        float* wv = w + (size_t)K * i;
        for (int s = 0; s < K; ++s) {
            const uint32_t j = idx[(size_t)K * i + s];
            const float d = dist[(size_t)K * i + s];
            float x = 0.0f;
            if (d < FLT_MAX && j < G) {
                const float r = radii[j];
                if (!(d >= r)) {
                    const float r2 = r * r;
                    const float den = 2.0f * r2;
                    x = vfo_expw(-(d * d) / den);
                }
            }
            wv[s] = x;
            sum = sum + x;
        }
        if (sum != 0.0f)
            for (int s = 0; s < K; ++s)
                wv[s] = wv[s] / sum;
    }
}

/* warpKernel, src/ComputeCorrespondence.cu:5-30: all K slots; the vertex is
 * materialised as a Vec4f, so each row of T p is Eigen's four-term reduction
 * (a0 + a1) + (a2 + a3) (Eigen/src/Core/Redux.h:99-113); weights all within
 * 1e-5 of 0 leave the point as it is.  A slot of weight exactly 0 adds
 * nothing: for a finite transform w t is a signed zero and the running sum,
 * never -0, is unchanged; skipping it keeps a missing neighbour's index
 * (UINT32_MAX) from being read.  Indices >= G contribute 0. */
void vfo_warp_mesh(int K, float* out, const float* pts, const float* T, const uint32_t* idx, const float* w, size_t n,
                   size_t G) {
    for (size_t i = 0; i < n; ++i) {
        const float* p = pts + 3 * i;
        const float* wv = w + (size_t)K * i;
        if (weights_zero(wv, K)) {
            out[3 * i + 0] = p[0];
            out[3 * i + 1] = p[1];
            out[3 * i + 2] = p[2];
            continue;
        }
        float acc[3] = {0.0f, 0.0f, 0.0f};
        for (int s = 0; s < K; ++s) {
            const uint32_t j = idx[(size_t)K * i + s];
            if (wv[s] == 0.0f)
                continue;
            float t[3] = {0.0f, 0.0f, 0.0f};
            if (j < G) {
                const float* M = T + 16 * (size_t)j;
                for (int a = 0; a < 3; ++a) {
                    const float a0 = M[0 * 4 + a] * p[0];
                    const float a1 = M[1 * 4 + a] * p[1];
                    const float a2 = M[2 * 4 + a] * p[2];
                    const float a3 = M[3 * 4 + a] * 1.0f;
                    t[a] = (a0 + a1) + (a2 + a3);
                }
            }
            for (int a = 0; a < 3; ++a) {
                const float wt = wv[s] * t[a];
                acc[a] = acc[a] + wt;
            }
        }
        out[3 * i + 0] = acc[0];
        out[3 * i + 1] = acc[1];
        out[3 * i + 2] = acc[2];
    }
}

/* The point path in one call: knn, weights, then the linear blend (blend 0)
 * or DQB (blend 1, weights as above, missing slots skipped). */
void vfo_warp_points(int K, float* out, const float* pts, size_t n, const float* nodes, const float* radii,
                     const float* T, size_t G, int blend) {
    uint32_t* idx = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)K * (n ? n : 1));
    float* dist = (float*)malloc(sizeof(float) * (size_t)K * (n ? n : 1));
    float* w = (float*)malloc(sizeof(float) * (size_t)K * (n ? n : 1));
    if (!idx || !dist || !w) {
        free(idx); free(dist); free(w);
        return;
    }
    vfo_mesh_knn(K, idx, dist, pts, n, nodes, G, 0);
    vfo_compute_warp_weights(K, w, radii, idx, dist, n, G);
    if (!blend) {
        vfo_warp_mesh(K, out, pts, T, idx, w, n, G);
    } else {
        for (size_t i = 0; i < n; ++i) {
            const float* wv = w + (size_t)K * i;
            if (weights_zero(wv, K)) {
                memcpy(out + 3 * i, pts + 3 * i, 3 * sizeof(float));
                continue;
            }
            float dqs[64][8], ws[64];
            const float* q[64];
            int m = 0;
            for (int s = 0; s < K; ++s) {
                const uint32_t j = idx[(size_t)K * i + s];
                if (j >= G)
                    continue;
                vfo_mat4_to_dq(T + 16 * (size_t)j, dqs[m]);
                q[m] = dqs[m];
                ws[m] = wv[s];
                ++m;
            }
            vfo_dqb_apply(q, ws, m, pts + 3 * i, out + 3 * i);
        }
    }
    free(idx); free(dist); free(w);
}

/* MotionGraph::warp, src/MotionGraph.cpp:62-90, host semantics: the K nearest
 * nodes (nanoflann's float L2, ((0 + d0^2) + d1^2) + d2^2,
 * third_party/include/nanoflann/nanoflann.hpp:432-440, then sqrt), Gaussian
 * weight with cutoff in binary64, p' = sum_j (w_j / sum) * (T_j p).
 * host_order selects Eigen's SSE product order ((a0 + a1) + a2) + m3 instead
 * of the scalar a0 + (a1 + a2) + m3.  squared_dist feeds nanoflann's squared
 * distance to the weight without the sqrt of src/MotionGraph.cpp:59. */
void vfo_mesh_warp(float* out, const float* pts, size_t n, const float* nodes, const float* radii, const float* T,
                   size_t G, int K, int host_order, int squared_dist) {
    float* dl = (float*)malloc(sizeof(float) * (size_t)(K > 0 ? K : 1));
    uint32_t* il = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(K > 0 ? K : 1));
    float* w = (float*)malloc(sizeof(float) * (size_t)(K > 0 ? K : 1));
    for (size_t i = 0; i < n; ++i) {
        const float* p = pts + 3 * i;
        int cnt = 0;
        for (size_t j = 0; j < G; ++j) {
            float s = 0.0f;
            for (int a = 0; a < 3; ++a) {
                const float diff = p[a] - nodes[3 * j + a];
                s = s + diff * diff;
            }
            knn_insert(K, dl, il, &cnt, s, (uint32_t)j);
        }
        float sum = 0.0f;
        for (int s = 0; s < cnt; ++s) {
            w[s] = vfo_weight_mesh_host(radii[il[s]], squared_dist ? dl[s] : sqrtf(dl[s]));
            sum = sum + w[s];
        }
        if (sum == 0.0f) {
            out[3 * i + 0] = p[0];
            out[3 * i + 1] = p[1];
            out[3 * i + 2] = p[2];
            continue;
        }
        float acc[3] = {0.0f, 0.0f, 0.0f};
        for (int s = 0; s < cnt; ++s) {
            const float* M = T + 16 * (size_t)il[s];
            float t[3];
            if (host_order) {
                for (int a = 0; a < 3; ++a) {
                    float r = M[0 * 4 + a] * p[0];
                    r = r + M[1 * 4 + a] * p[1];
                    r = r + M[2 * 4 + a] * p[2];
                    t[a] = r + M[3 * 4 + a];
                }
            } else {
                mat4_apply(M, p, t);
            }
            const float ws = w[s] / sum;
            for (int a = 0; a < 3; ++a)
                acc[a] = acc[a] + ws * t[a];
        }
        out[3 * i + 0] = acc[0];
        out[3 * i + 1] = acc[1];
        out[3 * i + 2] = acc[2];
    }
    free(dl);
    free(il);
    free(w);
}

/* ------------------------------------------------------------------------ */
/* the per-frame sequence                                                   */
/* ------------------------------------------------------------------------ */

/* src/VolumeFusion.cu:699-742 with R = I and t = c, as runVolumeFusion calls
 * it:  nodes -= c; knn; weights; vertices; nodes += c; vertices += c; warp;
 * count collisions; integrate.  `nodes` is updated in place exactly as the
 * reference's two translations leave it.  Returns 0, or -1 on allocation
 * failure. */
int vfo_warped_frame(int K, float* tsdf, float* wt, float* nodes, const float* radii, const float* T, size_t G,
                     const vfo_vol* vol, const int32_t* cube, const float* c, const float* depth, int W, int H,
                     float delta, const float* Kin, float maxw, vfo_stats* st, const vfo_frame_out* out,
                     int blend) {
    geom_t g;
    geom_init(&g, vol);
    const size_t N = g.n;
    int32_t* cnt = (int32_t*)calloc(N, sizeof(int32_t));
    uint32_t* idx = (uint32_t*)malloc(sizeof(uint32_t) * N * (size_t)K);
    float* dist = (float*)malloc(sizeof(float) * N * (size_t)K);
    float* w = (float*)malloc(sizeof(float) * N * (size_t)K);
    float* verts = (float*)malloc(sizeof(float) * N * 3);
    int32_t* col = (int32_t*)calloc(N, sizeof(int32_t));
    float* DQ = blend ? (float*)malloc(sizeof(float) * 8 * (G ? G : 1)) : NULL;
    if (!cnt || !idx || !dist || !w || !verts || !col || (blend && !DQ)) {
        free(cnt); free(idx); free(dist); free(w); free(verts); free(col); free(DQ);
        return -1;
    }
    for (size_t j = 0; blend && j < G; ++j)
        vfo_mat4_to_dq(T + 16 * j, DQ + 8 * j);
    const float negc[3] = {-c[0], -c[1], -c[2]};
    static const float I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};

    vfo_translate(nodes, G, negc);                                  /* :708 */
    const int64_t overflow = vfo_knn_graph(K, idx, dist, cnt, nodes, radii, G, vol, cube); /* :709 */
    vfo_warp_weights(K, w, radii, idx, dist, cnt, N);               /* :714 */
    vfo_generate_vertices(verts, vol);                              /* :718 */
    vfo_translate(nodes, G, c);                                     /* :721 */
    vfo_translate(verts, N, c);                                     /* :723 */
    if (blend)
        vfo_warp_tsdf_dqb(K, verts, DQ, idx, w, cnt, N);
    else
        vfo_warp_tsdf(K, verts, T, idx, w, cnt, N);                 /* :726 */
    vfo_count_collisions(col, verts, cnt, vol, I3, c);              /* :730 */

    /* :735-741, counting the statistics on the way */
    float Rinv[9];
    vfo_mat3_inverse(I3, Rinv);
    vfo_stats s;
    memset(&s, 0, sizeof(s));
    s.overflow = overflow;
    for (size_t v = 0; v < N; ++v) {
        if (out && out->pix)
            out->pix[v] = -1;
        if (!cnt[v])
            continue;
        s.supported++;
        s.candidates += cnt[v];
        size_t u;
        if (!warped_target(&g, Rinv, c, verts + 3 * v, &u))
            continue;
        s.in_bounds++;
        if (col[u] > 1) {
            s.collided++;
            continue;
        }
        const int64_t pi = update_tsdf(tsdf, wt, verts + 3 * v, v, depth, W, H, delta, Kin, maxw);
        if (pi >= 0) {
            s.updated++;
            if (out && out->pix)
                out->pix[v] = (int32_t)pi;
        }
    }
    if (st)
        *st = s;
    if (out) {
        if (out->cnt) memcpy(out->cnt, cnt, sizeof(int32_t) * N);
        if (out->idx) memcpy(out->idx, idx, sizeof(uint32_t) * N * (size_t)K);
        if (out->dist) memcpy(out->dist, dist, sizeof(float) * N * (size_t)K);
        if (out->w) memcpy(out->w, w, sizeof(float) * N * (size_t)K);
        if (out->verts) memcpy(out->verts, verts, sizeof(float) * N * 3);
        if (out->col) memcpy(out->col, col, sizeof(int32_t) * N);
    }
    free(cnt); free(idx); free(dist); free(w); free(verts); free(col); free(DQ);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* voxel-major (gather) restatement of the same frame, multi-threaded       */
/* ------------------------------------------------------------------------ */
/* Nodes are binned into 8x8x8 voxel bricks by their cubes; each voxel then
 * visits its brick's nodes in ascending index order and keeps the canonical
 * (dist, index) top-K.  Same operations, same order, so it must equal
 * vfo_warped_frame bit for bit; it is the CPU baseline that bench.py times.
 * Sources are restricted to z in [z0, z1) (z1 <= 0 means the whole volume);
 * collisions are counted among those sources only. */

static void f64_warped_point(const geom_t* g, const vfo_vol* vol, int64_t x, int64_t y, int64_t z, int n,
                             const uint32_t* il, const float* nodes, const float* radii, const float* T,
                             const float* c, int blend, double acc[3]);
static void f64_update(double* tsdf64, double* wt64, size_t v, double z, float d, float delta, float maxw);

typedef struct {
    float m[3];
    float r2;       /* 2r, the candidate bound          */
    int64_t lo[3];  /* cube lower corner                */
    int64_t hi[3];  /* cube upper corner (inclusive)    */
} gnode_t;

typedef struct {
    int K;
    const geom_t* g;
    const gnode_t* gn;
    const float* radii;
    const float* T;
    const float* DQ;      /* per-node dual quaternions when blending with DQB */
    const float* c;
    const uint32_t* boff; /* brick list offsets */
    const uint32_t* bl;   /* brick lists        */
    int32_t bdim[3];
    int z0, z1;
    float* pw;            /* [N*3] warped positions of supported voxels */
    uint8_t* sup;         /* [N] 1 = supported and target in bounds      */
    uint32_t* tgt;        /* [N] target voxel                            */
    int32_t* col;         /* [N] collision counts (atomic)               */
    float* tsdf;
    float* wt;
    const float* depth;
    int W, H;
    float delta, maxw;
    const float* Kin;
    int nthreads, tid;
    int64_t supported, in_bounds, collided, updated, candidates, overflow;
    /* binary64 companion (optional): the frame's positions before the
     * translations, the warped z of each in-bounds voxel, the state */
    const vfo_vol* vol;
    const float* nodes0;
    double* z64;
    double* tsdf64;
    double* wt64;
} gwork_t;

static void* gather_phase1(void* arg) {
    gwork_t* wk = (gwork_t*)arg;
    const geom_t* g = wk->g;
    const int K = wk->K;
    float dl[64];
    uint32_t il[64];
    static const float I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int64_t z = wk->z0 + wk->tid; z < wk->z1; z += wk->nthreads) {
        for (int64_t y = 0; y < g->dim[1]; ++y) {
            for (int64_t x = 0; x < g->dim[0]; ++x) {
                const size_t b = ((size_t)(z >> 3) * (size_t)wk->bdim[1] + (size_t)(y >> 3)) * (size_t)wk->bdim[0] +
                                 (size_t)(x >> 3);
                const size_t v = lin(g, x, y, z);
                int n = 0, seen = 0;
                for (uint32_t e = wk->boff[b]; e < wk->boff[b + 1]; ++e) {
                    const uint32_t j = wk->bl[e];
                    const gnode_t* nd = wk->gn + j;
                    if (x < nd->lo[0] || x > nd->hi[0] || y < nd->lo[1] || y > nd->hi[1] || z < nd->lo[2] ||
                        z > nd->hi[2])
                        continue;
                    float d;
                    float p[3];
                    voxel_to_world(g, x, y, z, p);
                    const float dx = p[0] - nd->m[0], dy = p[1] - nd->m[1], dz = p[2] - nd->m[2];
                    const float s = dx * dx + (dy * dy + dz * dz);
                    d = sqrtf(s);
                    if (d > nd->r2)
                        continue;
                    seen++;
                    knn_insert(K, dl, il, &n, d, j);
                }
                if (!n)
                    continue;
                wk->supported++;
                wk->candidates += n;
                wk->overflow += seen > K;
                /* weights, src/TSDFWarp.cu:101-113 */
                float w[64], sum = 0.0f;
                for (int s = 0; s < n; ++s) {
                    w[s] = vfo_weight_tsdf(wk->radii[il[s]], dl[s]);
                    sum = sum + w[s];
                }
                if (sum != 0.0f)
                    for (int s = 0; s < n; ++s)
                        w[s] = w[s] / sum;
                /* vertex + c, warp, src/TSDFWarp.cu:183-189 */
                float p[3];
                voxel_to_world(g, x, y, z, p);
                for (int a = 0; a < 3; ++a)
                    p[a] = p[a] + wk->c[a];
                float acc[3] = {0.0f, 0.0f, 0.0f};
                if (!weights_zero(w, n) && wk->DQ) {
                    const float* q[64];
                    for (int s = 0; s < n; ++s)
                        q[s] = wk->DQ + 8 * (size_t)il[s];
                    vfo_dqb_apply(q, w, n, p, acc);
                } else if (!weights_zero(w, n)) {
                    for (int s = 0; s < n; ++s) {
                        float t[3];
                        mat4_apply(wk->T + 16 * (size_t)il[s], p, t);
                        for (int a = 0; a < 3; ++a)
                            acc[a] = acc[a] + w[s] * t[a];
                    }
                } else {
                    acc[0] = p[0]; acc[1] = p[1]; acc[2] = p[2];
                }
                size_t u;
                if (!warped_target(g, I3, wk->c, acc, &u))
                    continue;
                wk->in_bounds++;
                wk->sup[v] = 1;
                wk->tgt[v] = (uint32_t)u;
                if (wk->z64) {
                    double acc64[3];
                    f64_warped_point(g, wk->vol, x, y, z, n, il, wk->nodes0, wk->radii, wk->T, wk->c, wk->DQ != NULL,
                                     acc64);
                    wk->z64[v] = acc64[2];
                }
                wk->pw[3 * v + 0] = acc[0];
                wk->pw[3 * v + 1] = acc[1];
                wk->pw[3 * v + 2] = acc[2];
                __atomic_fetch_add(wk->col + u, 1, __ATOMIC_RELAXED);
            }
        }
    }
    return NULL;
}

static void* gather_phase2(void* arg) {
    gwork_t* wk = (gwork_t*)arg;
    const geom_t* g = wk->g;
    for (int64_t z = wk->z0 + wk->tid; z < wk->z1; z += wk->nthreads)
        for (int64_t y = 0; y < g->dim[1]; ++y)
            for (int64_t x = 0; x < g->dim[0]; ++x) {
                const size_t v = lin(g, x, y, z);
                if (!wk->sup[v])
                    continue;
                if (wk->col[wk->tgt[v]] > 1) {
                    wk->collided++;
                    continue;
                }
                const int64_t pi = update_tsdf(wk->tsdf, wk->wt, wk->pw + 3 * v, v, wk->depth, wk->W, wk->H,
                                               wk->delta, wk->Kin, wk->maxw);
                if (pi >= 0) {
                    wk->updated++;
                    if (wk->tsdf64)
                        f64_update(wk->tsdf64, wk->wt64, v, wk->z64[v], wk->depth[pi], wk->delta, wk->maxw);
                }
            }
    return NULL;
}

int vfo_warped_frame_gather(int K, float* tsdf, float* wt, float* nodes, const float* radii, const float* T,
                            size_t G, const vfo_vol* vol, const int32_t* cube, const float* c, const float* depth,
                            int W, int H, float delta, const float* Kin, float maxw, int nthreads, int z0, int z1,
                            vfo_stats* st, int blend, double* tsdf64, double* wt64) {
    if (K < 1 || K > 64)
        return -2;
    geom_t g;
    geom_init(&g, vol);
    const size_t N = g.n;
    if (z1 <= 0 || z1 > g.dim[2])
        z1 = g.dim[2];
    if (z0 < 0)
        z0 = 0;
    if (nthreads < 1)
        nthreads = 1;
    const int32_t bdim[3] = {(g.dim[0] + 7) / 8, (g.dim[1] + 7) / 8, (g.dim[2] + 7) / 8};
    const size_t NB = (size_t)bdim[0] * (size_t)bdim[1] * (size_t)bdim[2];
    gnode_t* gn = (gnode_t*)malloc(sizeof(gnode_t) * (G ? G : 1));
    uint32_t* boff = (uint32_t*)calloc(NB + 1, sizeof(uint32_t));
    float* pw = (float*)malloc(sizeof(float) * 3 * N);
    uint8_t* sup = (uint8_t*)calloc(N, 1);
    uint32_t* tgt = (uint32_t*)malloc(sizeof(uint32_t) * N);
    int32_t* col = (int32_t*)calloc(N, sizeof(int32_t));
    float* DQ = blend ? (float*)malloc(sizeof(float) * 8 * (G ? G : 1)) : NULL;
    double* z64 = tsdf64 ? (double*)malloc(sizeof(double) * (N ? N : 1)) : NULL;
    if (!gn || !boff || !pw || !sup || !tgt || !col || (blend && !DQ) || (tsdf64 && !z64)) {
        free(gn); free(boff); free(pw); free(sup); free(tgt); free(col); free(DQ); free(z64);
        return -1;
    }
    for (size_t j = 0; blend && j < G; ++j)
        vfo_mat4_to_dq(T + 16 * j, DQ + 8 * j);
    const int valid_cube = cube[0] >= 1 && cube[1] >= 1 && cube[2] >= 1;
    /* node prep: nodes - c (src/VolumeFusion.cu:708), cube, 2r */
    for (size_t j = 0; j < G; ++j) {
        float m[3];
        for (int a = 0; a < 3; ++a)
            m[a] = nodes[3 * j + a] + (-c[a]);
        int64_t lo[3];
        node_cube_lo(&g, m, cube, lo);
        for (int a = 0; a < 3; ++a) {
            gn[j].m[a] = m[a];
            gn[j].lo[a] = lo[a];
            gn[j].hi[a] = lo[a] + cube[a] - 1;
        }
        gn[j].r2 = radii[j] * 2.0f;
    }
    /* bin nodes into bricks (counting pass, then fill in ascending j) */
    for (int pass = 0; pass < 2; ++pass) {
        uint32_t* fill = NULL;
        if (pass == 1) {
            for (size_t b = 0; b < NB; ++b)
                boff[b + 1] += boff[b];
            fill = (uint32_t*)calloc(NB, sizeof(uint32_t));
        }
        uint32_t* bl = pass == 1 ? (uint32_t*)malloc(sizeof(uint32_t) * (boff[NB] ? boff[NB] : 1)) : NULL;
        for (size_t j = 0; valid_cube && j < G; ++j) {
            int64_t blo[3], bhi[3];
            int empty = 0;
            for (int a = 0; a < 3; ++a) {
                int64_t lo = gn[j].lo[a] < 0 ? 0 : gn[j].lo[a];
                int64_t hi = gn[j].hi[a] >= g.dim[a] ? g.dim[a] - 1 : gn[j].hi[a];
                if (lo > hi)
                    empty = 1;
                blo[a] = lo >> 3;
                bhi[a] = hi >> 3;
            }
            if (empty)
                continue;
            for (int64_t bz = blo[2]; bz <= bhi[2]; ++bz)
                for (int64_t by = blo[1]; by <= bhi[1]; ++by)
                    for (int64_t bx = blo[0]; bx <= bhi[0]; ++bx) {
                        const size_t b = ((size_t)bz * (size_t)bdim[1] + (size_t)by) * (size_t)bdim[0] + (size_t)bx;
                        if (pass == 0)
                            boff[b + 1]++;
                        else
                            bl[boff[b] + fill[b]++] = (uint32_t)j;
                    }
        }
        if (pass == 1) {
            free(fill);
            /* phase 1 and 2 over threads */
            pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
            gwork_t* wk = (gwork_t*)calloc((size_t)nthreads, sizeof(gwork_t));
            for (int t = 0; t < nthreads; ++t) {
                gwork_t* w = wk + t;
                w->K = K; w->g = &g; w->gn = gn; w->radii = radii; w->T = T; w->DQ = DQ; w->c = c; w->boff = boff; w->bl = bl;
                w->bdim[0] = bdim[0]; w->bdim[1] = bdim[1]; w->bdim[2] = bdim[2];
                w->z0 = z0; w->z1 = z1; w->pw = pw; w->sup = sup; w->tgt = tgt; w->col = col;
                w->tsdf = tsdf; w->wt = wt; w->depth = depth; w->W = W; w->H = H; w->delta = delta; w->maxw = maxw;
                w->Kin = Kin; w->nthreads = nthreads; w->tid = t;
                w->vol = vol; w->nodes0 = nodes; w->z64 = z64; w->tsdf64 = tsdf64; w->wt64 = wt64;
            }
            for (int t = 0; t < nthreads; ++t)
                pthread_create(th + t, NULL, gather_phase1, wk + t);
            for (int t = 0; t < nthreads; ++t)
                pthread_join(th[t], NULL);
            for (int t = 0; t < nthreads; ++t)
                pthread_create(th + t, NULL, gather_phase2, wk + t);
            for (int t = 0; t < nthreads; ++t)
                pthread_join(th[t], NULL);
            vfo_stats s;
            memset(&s, 0, sizeof(s));
            for (int t = 0; t < nthreads; ++t) {
                s.supported += wk[t].supported;
                s.in_bounds += wk[t].in_bounds;
                s.collided += wk[t].collided;
                s.updated += wk[t].updated;
                s.candidates += wk[t].candidates;
                s.overflow += wk[t].overflow;
            }
            if (st)
                *st = s;
            free(th);
            free(wk);
            free(bl);
        }
    }
    /* nodes after the two translations, src/VolumeFusion.cu:708, 721 */
    for (size_t j = 0; j < G; ++j)
        for (int a = 0; a < 3; ++a)
            nodes[3 * j + a] = gn[j].m[a] + c[a];
    free(gn); free(boff); free(pw); free(sup); free(tgt); free(col); free(DQ); free(z64);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* binary64 companion for error reporting                                   */
/* ------------------------------------------------------------------------ */
/* Same discrete decisions as the float frame (knn set and order, which voxels
 * update, which depth pixel), continuous values in binary64: node offsets,
 * distances, Gaussian weights, the blended position and the running average.
 * tsdf64/wt64 carry the binary64 state across frames.  nodes are the
 * positions before the frame's translations. */
/* The binary64 warped position of voxel (x, y, z) with its n selected nodes
 * il[0..n): node offsets, squared distances, Gaussian weights, normalised
 * blend (LBS over Mat4f or DQB) in binary64. */
static void f64_warped_point(const geom_t* g, const vfo_vol* vol, int64_t x, int64_t y, int64_t z, int n,
                             const uint32_t* il, const float* nodes, const float* radii, const float* T,
                             const float* c, int blend, double acc[3]) {
    double p[3];
    p[0] = (double)x * (double)g->vs[0] - 0.5 * (double)vol->size[0];
    p[1] = (double)y * (double)g->vs[1] - 0.5 * (double)vol->size[1];
    p[2] = (double)z * (double)g->vs[2] - 0.5 * (double)vol->size[2];
    double w[64], sum = 0.0;
    for (int s = 0; s < n; ++s) {
        const uint32_t j = il[s];
        double d2 = 0.0;
        for (int a = 0; a < 3; ++a) {
            const double dd = p[a] - ((double)nodes[3 * j + a] - (double)c[a]);
            d2 += dd * dd;
        }
        const double r = (double)radii[j];
        w[s] = exp(-d2 / (2.0 * r * r));
        sum += w[s];
    }
    double pc[3];
    for (int a = 0; a < 3; ++a) {
        pc[a] = p[a] + (double)c[a];
        acc[a] = 0.0;
    }
    if (blend) {
        double dqs[64][8], wn[64];
        const double* q[64];
        for (int s = 0; s < n; ++s) {
            vfo_mat4_to_dq_f64(T + 16 * (size_t)il[s], dqs[s]);
            q[s] = dqs[s];
            wn[s] = w[s] / sum;
        }
        vfo_dqb_apply_f64(q, wn, n, pc, acc);
        return;
    }
    for (int s = 0; s < n; ++s) {
        const float* M = T + 16 * (size_t)il[s];
        for (int a = 0; a < 3; ++a) {
            const double t = (double)M[0 * 4 + a] * pc[0] + (double)M[1 * 4 + a] * pc[1] +
                             (double)M[2 * 4 + a] * pc[2] + (double)M[3 * 4 + a];
            acc[a] += (w[s] / sum) * t;
        }
    }
}

/* The binary64 running average of src/TSDFIntegration.cu:46-64 for a voxel
 * the float frame updated from depth pixel value d. */
static void f64_update(double* tsdf64, double* wt64, size_t v, double z, float d, float delta, float maxw) {
    const double sdf = z - (double)d;
    const double dl = (double)delta;
    const double tn = sdf < -dl ? -1.0 : sdf / dl;
    const double wo = wt64[v];
    tsdf64[v] = (tsdf64[v] * wo + tn) / (wo + 1.0);
    wt64[v] = (wo + 1.0) < (double)maxw ? (wo + 1.0) : (double)maxw;
}

int vfo_warped_frame_f64(int K, double* tsdf64, double* wt64, const float* nodes, const float* radii,
                         const float* T, size_t G, const vfo_vol* vol, const float* c, const float* depth, int W,
                         float delta, float maxw, const int32_t* cnt, const uint32_t* idx, const int32_t* pix,
                         int blend) {
    (void)G;
    (void)W;
    geom_t g;
    geom_init(&g, vol);
    for (int64_t z = 0; z < g.dim[2]; ++z)
        for (int64_t y = 0; y < g.dim[1]; ++y)
            for (int64_t x = 0; x < g.dim[0]; ++x) {
                const size_t v = lin(&g, x, y, z);
                if (pix[v] < 0 || !cnt[v])
                    continue;
                const int n = cnt[v] > K ? K : cnt[v];
                double acc[3];
                f64_warped_point(&g, vol, x, y, z, n, idx + (size_t)K * v, nodes, radii, T, c, blend, acc);
                f64_update(tsdf64, wt64, v, acc[2], depth[pix[v]], delta, maxw);
            }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* marching cubes over the TSDF                                             */
/* ------------------------------------------------------------------------ */

static const uint64_t kMcTri[256] = VF_MC_TRI_TABLE;
static const int kMcEdge[12][2] = VF_MC_EDGE_CORNERS;
static const int kMcCorner[8][3] = VF_MC_CORNER_OFFSETS;

/* interpolate, src/MarchingCubes.cu:294-312: the ratio in float, widened to
 * binary64, clamped, and the lerp in binary64 rounded to float. */
static void mc_interp(float t0, float t1, const float* p0, const float* p1, float iso, float* out) {
    if (fabs((double)(iso - t0)) < 1e-7) {
        memcpy(out, p0, 3 * sizeof(float));
        return;
    }
    if (fabs((double)(iso - t1)) < 1e-7) {
        memcpy(out, p1, 3 * sizeof(float));
        return;
    }
    if (fabs((double)(t0 - t1)) < 1e-7) {
        memcpy(out, p0, 3 * sizeof(float));
        return;
    }
    double mu = (double)((iso - t0) / (t1 - t0));
    if (mu > 1.0)
        mu = 1.0;
    else if (mu < 0)
        mu = 0.0;
    for (int a = 0; a < 3; ++a)
        out[a] = (float)((double)p0[a] + mu * (double)(p1[a] - p0[a]));
}

/* Case of the cube at (i, j, k), computeLutIndex (src/MarchingCubes.cu:336-377):
 * 0 when any corner has weight 0. */
static int mc_case(const geom_t* g, const float* tsdf, const float* wt, int64_t i, int64_t j, int64_t k, float iso) {
    size_t off[8];
    for (int c = 0; c < 8; ++c)
        off[c] = lin(g, i + kMcCorner[c][0], j + kMcCorner[c][1], k + kMcCorner[c][2]);
    for (int c = 0; c < 8; ++c)
        if (wt[off[c]] == 0.0f)
            return 0;
    int idx = 0;
    for (int c = 0; c < 8; ++c)
        if (tsdf[off[c]] > iso)
            idx |= 1 << c;
    return idx;
}

static void mc_edge_point(const geom_t* g, const float* tsdf, int64_t i, int64_t j, int64_t k, int e, float iso,
                          float* out) {
    const int a = kMcEdge[e][0], b = kMcEdge[e][1];
    const int64_t ia = i + kMcCorner[a][0], ja = j + kMcCorner[a][1], ka = k + kMcCorner[a][2];
    const int64_t ib = i + kMcCorner[b][0], jb = j + kMcCorner[b][1], kb = k + kMcCorner[b][2];
    float pa[3], pb[3];
    voxel_to_world(g, ia, ja, ka, pa);
    voxel_to_world(g, ib, jb, kb, pb);
    mc_interp(tsdf[lin(g, ia, ja, ka)], tsdf[lin(g, ib, jb, kb)], pa, pb, iso, out);
}

static int mc_edge_mask(int c) {
    int m = 0;
    for (int e = 0; e < 12; ++e)
        if (((c >> kMcEdge[e][0]) ^ (c >> kMcEdge[e][1])) & 1)
            m |= 1 << e;
    return m;
}

/* marchingCubesMeshKernel (full == 0, src/MarchingCubes.cu:422-508: one
 * vertex per crossed edge) and marchingCubesFullKernel (full == 1,
 * :526-682 with computeTriangles :395-420: three vertices and one face per
 * non-degenerate triangle).  Cubes at i < X-2, j < Y-2, k < Z-2, as the
 * reference launches them.  The reference appends with atomics, so its order
 * is a race; here the order is canonical: cubes by (k, j, i), edges by
 * index, triangles in table order.  Writes at most cap_verts vertices and
 * cap_faces faces (face indices count from 0) and returns the full counts. */
void vfo_marching_cubes(int full, const float* tsdf, const float* wt, const vfo_vol* vol, float iso, float* verts,
                        uint64_t cap_verts, int32_t* faces, uint64_t cap_faces, uint64_t* n_verts,
                        uint64_t* n_faces) {
    geom_t g;
    geom_init(&g, vol);
    uint64_t nv = 0, nf = 0;
    for (int64_t k = 0; k + 2 < g.dim[2]; ++k)
        for (int64_t j = 0; j + 2 < g.dim[1]; ++j)
            for (int64_t i = 0; i + 2 < g.dim[0]; ++i) {
                const int c = mc_case(&g, tsdf, wt, i, j, k, iso);
                if (c == 0 || c == 255)
                    continue;
                const int mask = mc_edge_mask(c);
                float ep[12][3];
                for (int e = 0; e < 12; ++e)
                    if (mask & (1 << e))
                        mc_edge_point(&g, tsdf, i, j, k, e, iso, ep[e]);
                if (!full) {
                    for (int e = 0; e < 12; ++e)
                        if (mask & (1 << e)) {
                            if (nv < cap_verts)
                                memcpy(verts + 3 * nv, ep[e], 3 * sizeof(float));
                            ++nv;
                        }
                    continue;
                }
                const uint64_t t = kMcTri[c];
                const int nt = (int)(t & 15);
                for (int q = 0; q < nt; ++q) {
                    const float* p[3];
                    for (int v = 0; v < 3; ++v)
                        p[v] = ep[(t >> (4 + 4 * (3 * q + v))) & 15];
                    const int same01 = p[0][0] == p[1][0] && p[0][1] == p[1][1] && p[0][2] == p[1][2];
                    const int same02 = p[0][0] == p[2][0] && p[0][1] == p[2][1] && p[0][2] == p[2][2];
                    const int same12 = p[1][0] == p[2][0] && p[1][1] == p[2][1] && p[1][2] == p[2][2];
                    if (same01 || same02 || same12)
                        continue;
                    for (int v = 0; v < 3; ++v) {
                        if (nv + v < cap_verts)
                            memcpy(verts + 3 * (nv + v), p[v], 3 * sizeof(float));
                    }
                    if (nf < cap_faces) {
                        faces[3 * nf + 0] = (int32_t)nv;
                        faces[3 * nf + 1] = (int32_t)(nv + 1);
                        faces[3 * nf + 2] = (int32_t)(nv + 2);
                    }
                    nv += 3;
                    nf += 1;
                }
            }
    *n_verts = nv;
    *n_faces = nf;
}

/* ------------------------------------------------------------------------ */
/* depth ingest (src/VertexManipulation.cu:6-71, 147-217; BilateralFilter.cu) */
/* ------------------------------------------------------------------------ */

/* filterDepthKernel, src/VertexManipulation.cu:6-15. */
void vfo_filter_depth(float* depth, size_t n, float threshold, float fill) {
    for (size_t i = 0; i < n; ++i) {
        const float v = depth[i];
        if (v > threshold || v < 1.e-4f || v != v)
            depth[i] = fill;
    }
}

/* maskDepthKernel, src/VertexManipulation.cu:25-32. */
void vfo_mask_depth(float* depth, const uint8_t* mask, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!mask[i])
            depth[i] = INFINITY;
}

/* depthToMeshKernel, src/VertexManipulation.cu:42-63: back-projection with
 * the reciprocals 1/fx, 1/fy; invalid depth gives (0, 0, 0).  K column-major. */
void vfo_depth_to_mesh(float* mesh, const float* depth, int W, int H, const float* K) {
    const float rfx = 1 / K[0], rfy = 1 / K[4];
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t i = (size_t)y * W + x;
            const float d = depth[i];
            if (isinf(d) || d < 1.e-4f) {
                mesh[3 * i] = mesh[3 * i + 1] = mesh[3 * i + 2] = 0.0f;
                continue;
            }
            const float x0 = ((float)x - K[6]) * rfx;
            const float y0 = ((float)y - K[7]) * rfy;
            mesh[3 * i] = x0 * d;
            mesh[3 * i + 1] = y0 * d;
            mesh[3 * i + 2] = d;
        }
}

/* depthVerticesNormalsKernel, src/VertexManipulation.cu:147-203, interior
 * pixels only.  The angle test acos(n.z) in (a0, a1) is evaluated as
 * n.z < cos(a0) and n.z > cos(a1) in binary64 (acos is decreasing); the
 * reference's float acos is not available, so its boundary ulp is not
 * pinned. */
void vfo_depth_normals(float* normals, const float* v, int H, int W, float a0, float a1) {
    const double c0 = cos((double)a0), c1 = cos((double)a1);
    for (int y = 1; y < H - 1; ++y)
        for (int x = 1; x < W - 1; ++x) {
            const size_t i = (size_t)y * W + x;
            float* nrm = normals + 3 * i;
            if (v[3 * i + 2] < 1.e-4f) {
                nrm[0] = nrm[1] = nrm[2] = 0.0f;
                continue;
            }
            const float* xr = v + 3 * (i + 1);
            const float* xl = v + 3 * (i - 1);
            const float* yb = v + 3 * (i + W);
            const float* yt = v + 3 * (i - W);
            if (xr[2] < 1.e-4f || xl[2] < 1.e-4f || yb[2] < 1.e-4f || yt[2] < 1.e-4f) {
                nrm[0] = nrm[1] = nrm[2] = 0.0f;
                continue;
            }
            float dx[3], dy[3], gx[3], gy[3], c[3];
            for (int a = 0; a < 3; ++a) {
                dx[a] = xr[a] - xl[a];
                dy[a] = yb[a] - yt[a];
            }
            const float nx = sqrtf(dx[0] * dx[0] + dx[1] * dx[1]);
            const float ny = sqrtf(dy[0] * dy[0] + dy[1] * dy[1]);
            for (int a = 0; a < 3; ++a) {
                gx[a] = dx[a] / nx;
                gy[a] = dy[a] / ny;
            }
            /* gY.cross(gX) */
            c[0] = gy[1] * gx[2] - gy[2] * gx[1];
            c[1] = gy[2] * gx[0] - gy[0] * gx[2];
            c[2] = gy[0] * gx[1] - gy[1] * gx[0];
            const float sq = c[0] * c[0] + (c[1] * c[1] + c[2] * c[2]);
            if (sq > 0.0f) {
                const float len = sqrtf(sq);
                for (int a = 0; a < 3; ++a)
                    c[a] = c[a] / len;
            }
            /* (0, 0, 1) . n in Eigen's order: 0 n.x + (0 n.y + 1 n.z) */
            const float dot = 0.0f * c[0] + (0.0f * c[1] + 1.0f * c[2]);
            if ((double)dot < c0 && (double)dot > c1) {
                nrm[0] = nrm[1] = nrm[2] = 0.0f;
                continue;
            }
            nrm[0] = c[0];
            nrm[1] = c[1];
            nrm[2] = c[2];
        }
}

/* removeDepthWithoutNormalsKernel, src/VertexManipulation.cu:219-232. */
void vfo_remove_depth_without_normals(const float* normals, float* v, float* depth, int H, int W) {
    for (int y = 1; y < H - 1; ++y)
        for (int x = 1; x < W - 1; ++x) {
            const size_t i = (size_t)y * W + x;
            const float* nrm = normals + 3 * i;
            if (fabsf(nrm[0]) <= 1.e-4f && fabsf(nrm[1]) <= 1.e-4f && fabsf(nrm[2]) <= 1.e-4f) {
                v[3 * i] = v[3 * i + 1] = v[3 * i + 2] = 0.0f;
                depth[i] = INFINITY;
            }
        }
}

/* Centroid of the back-projected points (src/VolumeFusion.cu:216-219): the
 * sum of all points divided by the number with z > 1e-4.  The reference sums
 * in float in thrust's order; here in binary64 in a fixed order (256-point
 * blocks reduced as a binary tree, a[i] += a[i + s] for s = 128 .. 1, then the
 * block sums in order), rounded once to float. */
void vfo_depth_centroid(const float* mesh, size_t n, float* out, uint64_t* count) {
    /* block totals: a binary tree over each 256-point block */
    const size_t nb = (n + 255) / 256;
    double* bt = (double*)calloc(3 * (nb ? nb : 1), sizeof(double));
    uint64_t cnt = 0;
    for (size_t b = 0; b < nb; ++b) {
        double blk[256][3];
        for (int t = 0; t < 256; ++t)
            for (int a = 0; a < 3; ++a)
                blk[t][a] = (256 * b + t < n) ? (double)mesh[3 * (256 * b + t) + a] : 0.0;
        for (int s = 128; s >= 1; s >>= 1)
            for (int t = 0; t < s; ++t)
                for (int a = 0; a < 3; ++a)
                    blk[t][a] = blk[t][a] + blk[t + s][a];
        for (int a = 0; a < 3; ++a)
            bt[3 * b + a] = blk[0][a];
        for (int t = 0; t < 256 && 256 * b + t < n; ++t)
            cnt += mesh[3 * (256 * b + t) + 2] > 1.e-4f;
    }
    /* the totals: 1024 chunks of ceil(nb / 1024) consecutive blocks, each
     * summed in order from 0, then a binary tree over the chunks */
    double part[1024][3];
    const size_t c = (nb + 1023) / 1024;
    for (size_t t = 0; t < 1024; ++t) {
        double q[3] = {0.0, 0.0, 0.0};
        for (size_t b = t * c; b < (t + 1) * c && b < nb; ++b)
            for (int a = 0; a < 3; ++a)
                q[a] = q[a] + bt[3 * b + a];
        for (int a = 0; a < 3; ++a)
            part[t][a] = q[a];
    }
    for (int s = 512; s >= 1; s >>= 1)
        for (int t = 0; t < s; ++t)
            for (int a = 0; a < 3; ++a)
                part[t][a] = part[t][a] + part[t + s][a];
    free(bt);
    for (int a = 0; a < 3; ++a)
        out[a] = (float)part[0][a] / (float)cnt;
    *count = cnt;
}

/* updateGaussian16 (src/BilateralFilter.cu:12-47): spatial weights
 * expf(-(x^2) / (2 sd^2)) per offset, their 2D products; intensity weights
 * as integers (int)(e^(-(i/10)^2 / 2) / e^(-4.5)), i = 0..30. */
void vfo_bilateral_tables(int r, double sd, float* g2d, int32_t* clr) {
    float g[64];
    for (int i = 0; i < 2 * r + 1; ++i) {
        const float x = (float)(i - r);
        g[i] = expf((float)(-(x * x) / (2 * sd * sd)));
    }
    int idx = 0;
    for (int dy = -r; dy <= r; ++dy)
        for (int dx = -r; dx <= r; ++dx)
            g2d[idx++] = g[dy + r] * g[dx + r];
    float fc[31];
    for (int i = 0; i <= 30; ++i)
        fc[i] = (float)exp(-(pow(i / 10.0, 2)) / 2);
    for (int i = 0; i <= 30; ++i)
        clr[i] = (int32_t)(fc[i] / fc[30]);
}

/* gpuFilterTextureOpmShared (src/BilateralFilter.cu:52-127): clamp-to-edge
 * neighbourhood of radius r; an infinite centre stays infinite; neighbours
 * that are infinite or differ by more than thr are skipped; the rest weigh
 * spatial(b, a) * intensity((int)|diff| 30/(3 sI)), summed with a (x offset)
 * outer and b (y offset) inner. */
void vfo_bilateral(float* out, const float* in, int H, int W, int r, double sI, double sS, float thr) {
    float g2d[400];
    int32_t clr[31];
    vfo_bilateral_tables(r, sS, g2d, clr);
    const float norm = (float)(30 / (3 * sI));
    const int kd = 2 * r + 1;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const float c = in[(size_t)y * W + x];
            if (isinf(c)) {
                out[(size_t)y * W + x] = INFINITY;
                continue;
            }
            float acc = 0.0f, wsum = 0.0f;
            int ig = 0;
            for (int a = 0; a < kd; ++a)
                for (int b = 0; b < kd; ++b, ++ig) {
                    int xx = x + a - r, yy = y + b - r;
                    xx = xx < 0 ? 0 : (xx >= W ? W - 1 : xx);
                    yy = yy < 0 ? 0 : (yy >= H ? H - 1 : yy);
                    const float p = in[(size_t)yy * W + xx];
                    if (isinf(p))
                        continue;
                    const float diff = fabsf(c - p);
                    if (diff > thr)
                        continue;
                    int x0 = vfo_cvt_i32_rz(fabsf(diff * norm));
                    if (x0 > 30 || x0 < 0)
                        x0 = 30;
                    const float w = g2d[ig] * (float)clr[x0];
                    acc = acc + p * w;
                    wsum = wsum + w;
                }
            out[(size_t)y * W + x] = acc / wsum;
        }
}
