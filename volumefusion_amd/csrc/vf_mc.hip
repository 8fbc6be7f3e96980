// vf_mc.hip -- marching cubes over the TSDF (af::runMarchingCubesMeshKernel
// and af::runMarchingCubesFullKernel, src/MarchingCubes.cu:422-702).
//
// The reference appends every vertex (and face) with an atomicAdd on one
// counter, so its output order is a race and the counter a hot spot.  Here
// the output is compacted in a canonical order (cubes by linear index, edges
// by number, triangles in table order) with two sweeps:
//   k_mc_count   per cube: case and output count; per run of 60 cubes of a
//                row: sum.  With faces, a cube whose triangles may be
//                degenerate flags its run for the next kernel.
//   k_mc_exact   (faces only) the flagged runs' exact sums
//   (scan)       exclusive scan of the run sums (rocPRIM)
//   k_mc_emit    per cube again, only in runs with output: wave-local
//                prefix, then the writes
//   k_mc_finish  advances the caller's counters
// The count sweep reads every voxel's tsdf and weight about once (waves march
// along z and share x neighbours across lanes); the emit sweep reads only the
// runs on the surface.  HBM-bound: about 8 B per voxel.
#include <hipcub/hipcub.hpp>

#include "vf_host.hpp"
#include "vf_mc_tables.h"

namespace vf {
namespace {

constexpr int kMcBlock = 256;

__constant__ unsigned long long kMcTri[256] = VF_MC_TRI_TABLE;
__constant__ int kMcEdge[12][2] = VF_MC_EDGE_CORNERS;
__constant__ int kMcCorner[8][3] = VF_MC_CORNER_OFFSETS;
constexpr int kCornerOff[8][3] = VF_MC_CORNER_OFFSETS;  // compile-time copies for register selection
constexpr int kMcEdgeC[12][2] = VF_MC_EDGE_CORNERS;
constexpr unsigned long long kMcTriC[256] = VF_MC_TRI_TABLE;

// Case lookup of the count sweep.  Index: bits 0-3 the above-iso flags of
// this lane's corners (row j and j + 1 of slice k, then of slice k + 1),
// bits 4-7 the same flags of the next lane's (x + 1) corners.  Entry: the
// case (bits 0-7), its crossed-edge count (8-15), its triangle count (16-23).
struct McCaseLut {
    uint32_t v[256];
    constexpr McCaseLut() : v() {
        for (int idx = 0; idx < 256; ++idx) {
            int cs = 0;
            for (int c = 0; c < 8; ++c) {
                const int dx = kCornerOff[c][0], dy = kCornerOff[c][1], dz = kCornerOff[c][2];
                const int bit = (idx >> (4 * dx + dy + 2 * dz)) & 1;
                cs |= bit << c;
            }
            int nedge = 0;
            for (int e = 0; e < 12; ++e) nedge += ((cs >> kMcEdgeC[e][0]) ^ (cs >> kMcEdgeC[e][1])) & 1;
            const int ntri = (cs == 0 || cs == 255) ? 0 : (int)(kMcTriC[cs] & 15);
            if (cs == 0 || cs == 255) nedge = 0;
            v[idx] = (uint32_t)cs | ((uint32_t)nedge << 8) | ((uint32_t)ntri << 16);
        }
    }
};
__device__ const McCaseLut kMcCaseLut = McCaseLut();

struct McGeom {
    int32_t X, Y, Z;
    int32_t cx, cy;  // cubes per row (X - 2), per column (Y - 2)
    float vs[3];
    float halfs[3];
    float iso;
    uint64_t ncubes;
    int32_t cz;  // cube layers (Z - 2)
    int32_t nr;  // runs of kRun cubes per row
    // the count sweep's snap screen over the whole volume: lmin <= every
    // cube's edge length L and pscale >= 2^-19 times every cube's largest
    // corner coordinate P, with margins of 2^-30 (see screen_ok)
    double lmin, pscale;
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

// The cube's corner positions differ in every axis (always, unless the
// voxel grid is too fine for binary32 where the cube sits).
__device__ __forceinline__ bool corners_distinct(const McGeom& g, const Cube& cb) {
    float p0[3], p1[3];
    mc_world(g, cb.i, cb.j, cb.k, p0);
    mc_world(g, cb.i + 1, cb.j + 1, cb.k + 1, p1);
    return p0[0] != p1[0] && p0[1] != p1[1] && p0[2] != p1[2];
}

// For each crossed edge e, four bits at 4 e: 1 + the corner its point equals
// exactly, or 0.  An edge point keeps the edge line's two constant
// coordinates exactly and its third between the corners', so two points of
// different edges are equal only where the two edge lines meet: a corner
// both edges end at, when the corner positions differ in every axis.  Two
// points are then equal exactly when their codes are equal and non-zero.
__device__ __forceinline__ unsigned long long edge_codes(const McGeom& g, const Cube& cb, int mask) {
    unsigned long long c = 0;
#pragma unroll
    for (int e = 0; e < 12; ++e) {
        if (mask & (1 << e)) {
            const int a = kMcEdgeC[e][0], b = kMcEdgeC[e][1];
            float pa[3], pb[3], o[3];
            mc_world(g, cb.i + kCornerOff[a][0], cb.j + kCornerOff[a][1], cb.k + kCornerOff[a][2], pa);
            mc_world(g, cb.i + kCornerOff[b][0], cb.j + kCornerOff[b][1], cb.k + kCornerOff[b][2], pb);
            mc_interp(cb.t[a], cb.t[b], pa, pb, g.iso, o);
            const unsigned long long code = same3(o, pa) ? (unsigned long long)(a + 1)
                                          : same3(o, pb) ? (unsigned long long)(b + 1)
                                                         : 0ull;
            c |= code << (4 * e);
        }
    }
    return c;
}

// No edge point of the cube can equal a corner (a screen before the exact
// test of edge_codes).  mc_interp returns a corner when a corner value is
// within 1e-7 of iso, or when the lerp's displacement mu * L along the edge
// is at most half an ulp of the corner coordinate (<= 2^-24 P, P the largest
// |coordinate| of the cube's corners, L the edge length as the float
// difference of the corners).  With m and M the smallest and largest
// |t - iso| over the corners, every crossed edge has mu and 1 - mu at least
// (1 - 2^-22) m / (2 M) after float rounding; m >= 2e-7, m >= 2^-16 M and
// m L >= 2^-19 P M then put the displacement above 2^-21 P on both ends.
// NaN or infinite corner values fail the screen.
__device__ __forceinline__ bool no_snap(const McGeom& g, const Cube& cb) {
    float m = 3.0e38f, M = 0.0f;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        const float d = fabsf(cb.t[v] - g.iso);
        if (!(d <= 3.0e38f)) return false;
        m = fminf(m, d);
        M = fmaxf(M, d);
    }
    if (!(m >= 2.0e-7f)) return false;
    float p0[3], p1[3];
    mc_world(g, cb.i, cb.j, cb.k, p0);
    mc_world(g, cb.i + 1, cb.j + 1, cb.k + 1, p1);
    double L = 3.0e38, P = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        L = fmin(L, (double)(p1[a] - p0[a]));
        P = fmax(P, fmax(fabs((double)p0[a]), fabs((double)p1[a])));
    }
    if (!(L > 0.0)) return false;
    return (double)m >= 0x1p-16 * (double)M && (double)m * L >= 0x1p-19 * P * (double)M;
}

// no_snap over the whole volume, for the count sweep: with m and M the
// smallest and largest |t - iso| over a cube's corners, m >= 2e-7,
// m >= 2^-16 M and m lmin >= pscale M imply no_snap's m L >= 2^-19 P M for
// every cube of the volume (lmin and pscale carry 2^-30 margins, far above
// the binary64 rounding of the two products).  Only m and M of the cube are
// needed, not its coordinates.
__device__ __forceinline__ bool screen_ok(const McGeom& g, float m, float M) {
    return m >= 2.0e-7f && (double)m >= 0x1p-16 * (double)M && (double)m * g.lmin >= g.pscale * (double)M;
}

// Output vertex count of a cube: 3 per triangle of its case (FULL, dropping
// triangles with two equal points, src/MarchingCubes.cu:400-410) or 1 per
// crossed edge.  With FULL, bit q of keep marks the kept triangles.
template <bool FULL>
__device__ __forceinline__ int cube_outputs(const McGeom& g, const Cube& cb, unsigned& keep) {
    keep = 0;
    if (cb.cs == 0 || cb.cs == 255) return 0;
    const int mask = edge_mask(cb.cs);
    if (!FULL) return __popc(mask);
    const unsigned long long t = kMcTri[cb.cs];
    const int nt = (int)(t & 15);
    int n = 0;
    if (no_snap(g, cb)) {
        // every edge point lies strictly inside its edge: no two coincide
        keep = (1u << nt) - 1u;
        n = 3 * nt;
    } else if (corners_distinct(g, cb)) {
        const unsigned long long codes = edge_codes(g, cb, mask);
        for (int q = 0; q < nt; ++q) {
            const int c0 = (int)(codes >> (4 * ((t >> (4 + 12 * q)) & 15))) & 15;
            const int c1 = (int)(codes >> (4 * ((t >> (8 + 12 * q)) & 15))) & 15;
            const int c2 = (int)(codes >> (4 * ((t >> (12 + 12 * q)) & 15))) & 15;
            const bool degenerate = (c0 && (c0 == c1 || c0 == c2)) || (c1 && c1 == c2);
            if (!degenerate) {
                keep |= 1u << q;
                n += 3;
            }
        }
    } else {
        // corners that share a coordinate: compare the points themselves
        for (int q = 0; q < nt; ++q) {
            float p0[3], p1[3], p2[3];
            edge_point(g, cb, (int)((t >> (4 + 12 * q)) & 15), p0);
            edge_point(g, cb, (int)((t >> (8 + 12 * q)) & 15), p1);
            edge_point(g, cb, (int)((t >> (12 + 12 * q)) & 15), p2);
            if (!same3(p0, p1) && !same3(p0, p2) && !same3(p1, p2)) {
                keep |= 1u << q;
                n += 3;
            }
        }
    }
    return n;
}

// The sweeps march along z.  In the exact and emit sweeps a wave owns a run
// of kRun cubes of one row (j, k): lane l holds voxel column x = r * kRun + l,
// and the cube at x takes its x + 1 corners from lane l + 1 (lane kRun only
// supplies corners, the lanes after it idle).  Each step down z loads one new
// slice: 4 coalesced loads per lane (tsdf and weight of rows j and j + 1).
// The count sweep takes four runs per wave, four cubes per lane (run_pos4).  Runs are numbered q = (k * cy + j) * nr + r,
// which is the cubes' linear order, so per-run sums scanned in q order give
// every cube's output offset.
constexpr int kRun = 60;          // 15 lanes x 4 cubes in the count sweep (one 16-lane row per run)
constexpr int kRunLanes = kRun / 4;
static_assert(kRunLanes == 15, "a run fills one 16-lane row but its last lane");
constexpr int kMarch = 16;  // z steps per wave
static_assert(kMarch <= 64, "one lane per step holds the run sums");

struct Slice {
    float t0, t1;  // tsdf at rows j, j + 1 of this lane's column
    float w0, w1;  // weights
};

// Columns beyond the row's end read the last column instead: only lanes
// that own no cube see them, through their own values or their left
// neighbour's, which owns none either.
__device__ __forceinline__ Slice load_slice(const McGeom& g, const float* __restrict__ tsdf,
                                            const float* __restrict__ wt, int32_t x, int32_t j, int32_t k) {
    const int32_t xc = x < g.X ? x : g.X - 1;
    // the slice's base (scalar where k is uniform, as in the sweeps) and the
    // lane's 32-bit place in it (run_mc checks that a slice has < 2^31 voxels)
    const uint64_t sk = (uint64_t)k * (uint64_t)g.X * (uint64_t)g.Y;
    const float* __restrict__ tk = tsdf + sk;
    const float* __restrict__ wk = wt + sk;
    const uint32_t o0 = (uint32_t)j * (uint32_t)g.X + (uint32_t)xc;
    const uint32_t o1 = o0 + (uint32_t)g.X;
    Slice s;
    s.t0 = tk[o0];
    s.t1 = tk[o1];
    s.w0 = wk[o0];
    s.w1 = wk[o1];
    return s;
}

// The value of lane l + 1 (DPP wave_shl:1, one VALU instruction); lane 63
// gets 0.
__device__ __forceinline__ float from_next_lane(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, false));
}

// The cube at (x, j, k) from this lane's slices k (lo) and k + 1 (hi) and the
// next lane's.  kCornerOff gives each corner's (dx, dy, dz).
__device__ __forceinline__ void make_cube(const McGeom& g, int32_t x, int32_t j, int32_t k, const Slice& lo,
                                          const Slice& hi, Cube* cb) {
    const float nt0 = from_next_lane(lo.t0), nt1 = from_next_lane(lo.t1);
    const float nw0 = from_next_lane(lo.w0), nw1 = from_next_lane(lo.w1);
    const float mt0 = from_next_lane(hi.t0), mt1 = from_next_lane(hi.t1);
    const float mw0 = from_next_lane(hi.w0), mw1 = from_next_lane(hi.w1);
    // [dz][dy][dx]
    const float T[2][2][2] = {{{lo.t0, nt0}, {lo.t1, nt1}}, {{hi.t0, mt0}, {hi.t1, mt1}}};
    const float W[2][2][2] = {{{lo.w0, nw0}, {lo.w1, nw1}}, {{hi.w0, mw0}, {hi.w1, mw1}}};
    cb->i = x;
    cb->j = j;
    cb->k = k;
    bool zero = false;
    int cs = 0;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        const int dx = kCornerOff[v][0], dy = kCornerOff[v][1], dz = kCornerOff[v][2];
        zero = zero || (W[dz][dy][dx] == 0.0f);
        cb->t[v] = T[dz][dy][dx];
        if (cb->t[v] > g.iso) cs |= 1 << v;
    }
    cb->cs = zero ? 0 : cs;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readlane((int)wave_inclusive_sum(v), 63);
}

__device__ __forceinline__ uint32_t wave_exclusive_sum(uint32_t v, int) { return wave_inclusive_sum(v) - v; }

struct RunPos {
    int32_t r, j, k0;
    bool live;
};

// A workgroup's four waves take four consecutive rows j of the same run and
// z segment, so the voxel row two waves share (j + 1 of one, j of the next)
// is fetched from HBM once and read again from the CU's caches.  Workgroups
// are dispatched to the eight XCDs round-robin, and each XCD has its own L2:
// tile numbers b go to the XCDs in chunks of kMcChunk row groups (every run
// of each), so neighbouring runs of a row (whose 252 B spans share a 128 B
// line) and neighbouring row groups (which share a row) mostly meet in one
// L2, while the chunks' round-robin keeps the surface work spread over the
// XCDs.  Tiles beyond the last whole round of chunks keep their number.
constexpr uint32_t kMcChunk = 8;
__device__ __forceinline__ RunPos run_pos(const McGeom& g) {
    const uint32_t groups = ((uint32_t)g.cy + (kMcBlock / 64) - 1) / (kMcBlock / 64);
    const uint32_t chunk = kMcChunk * (uint32_t)g.nr;
    const uint32_t whole = gridDim.x / (8u * chunk) * (8u * chunk);
    uint32_t b = blockIdx.x;
    if (b < whole) {
        const uint32_t i = b >> 3;  // this workgroup's place among its XCD's
        b = ((i / chunk) * 8u + (blockIdx.x & 7u)) * chunk + i % chunk;
    }
    RunPos p;
    p.r = (int32_t)(b % (uint32_t)g.nr);
    p.j = (int32_t)(((b / (uint32_t)g.nr) % groups) * (kMcBlock / 64) + (threadIdx.x >> 6));
    p.k0 = (int32_t)(b / ((uint32_t)g.nr * groups)) * kMarch;
    p.live = p.k0 < g.cz && p.j < g.cy;
    return p;
}

// The count sweep's mapping: a workgroup's four waves take four consecutive
// rows j of the same four runs and z segment (rows two waves share are read
// from HBM once, as in run_pos); in a wave, 16-lane row q takes run r + q,
// its lane s the four columns x0 = (r + q) * kRun + 4 s (lane 15 only
// supplies the corners at x0 of its row).  Tile numbers go to the XCDs in
// chunks as in run_pos.
__device__ __forceinline__ RunPos run_pos4(const McGeom& g) {
    const uint32_t groups = ((uint32_t)g.cy + (kMcBlock / 64) - 1) / (kMcBlock / 64);
    const uint32_t rg = ((uint32_t)g.nr + 3u) / 4u;  // groups of four runs per row
    const uint32_t chunk = kMcChunk * rg;
    const uint32_t whole = gridDim.x / (8u * chunk) * (8u * chunk);
    uint32_t b = blockIdx.x;
    if (b < whole) {
        const uint32_t i = b >> 3;
        b = ((i / chunk) * 8u + (blockIdx.x & 7u)) * chunk + i % chunk;
    }
    RunPos p;
    p.r = (int32_t)(b % rg) * 4;
    p.j = (int32_t)(((b / rg) % groups) * (kMcBlock / 64) + (threadIdx.x >> 6));
    p.k0 = (int32_t)(b / (rg * groups)) * kMarch;
    p.live = p.k0 < g.cz && p.j < g.cy;
    return p;
}

// One slice of four columns: tsdf and weight of rows j and j + 1.
struct Slice4 {
    float t0[4], t1[4], w0[4], w1[4];
};

// VEC: 16-byte loads (rows 16-byte aligned); a lane whose four columns start
// past the row's end reads the row's last four instead.  Else one load per
// column, columns beyond the row reading the last one (as load_slice).
// Either way only cubes no lane owns see columns past the row's end (their
// x + 1 corner of a lane's last cube lies past cx otherwise), and the loads
// are the same on every pass (no branch around them), so the waits before a
// step cover only the slices it reads.
template <bool VEC>
__device__ __forceinline__ Slice4 load_slice4(const McGeom& g, const float* __restrict__ tsdf,
                                              const float* __restrict__ wt, int32_t x0, int32_t j, int32_t k) {
    const uint64_t sk = (uint64_t)k * (uint64_t)g.X * (uint64_t)g.Y;
    const float* __restrict__ tk = tsdf + sk;
    const float* __restrict__ wk = wt + sk;
    const uint32_t o0 = (uint32_t)j * (uint32_t)g.X, o1 = o0 + (uint32_t)g.X;
    Slice4 s;
    if (VEC) {
        const uint32_t xs = (uint32_t)(x0 < g.X - 4 ? x0 : g.X - 4);
        const float4 a = *(const float4*)(tk + o0 + xs), b = *(const float4*)(tk + o1 + xs);
        const float4 c = *(const float4*)(wk + o0 + xs), d = *(const float4*)(wk + o1 + xs);
        s.t0[0] = a.x, s.t0[1] = a.y, s.t0[2] = a.z, s.t0[3] = a.w;
        s.t1[0] = b.x, s.t1[1] = b.y, s.t1[2] = b.z, s.t1[3] = b.w;
        s.w0[0] = c.x, s.w0[1] = c.y, s.w0[2] = c.z, s.w0[3] = c.w;
        s.w1[0] = d.x, s.w1[1] = d.y, s.w1[2] = d.z, s.w1[3] = d.w;
    } else {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const uint32_t xc = (uint32_t)(x0 + c < g.X ? x0 + c : g.X - 1);
            s.t0[c] = tk[o0 + xc];
            s.t1[c] = tk[o1 + xc];
            s.w0[c] = wk[o0 + xc];
            s.w1[c] = wk[o1 + xc];
        }
    }
    return s;
}

// Inclusive sum within each 16-lane row (DPP row shifts; VALU only).
__device__ __forceinline__ uint32_t row_inclusive_sum(uint32_t x) {
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);  // row_shr:8
    return x;
}

// With FULL the sweep counts 3 per triangle where screen_ok rules out
// degenerate triangles and leaves the other runs to k_mc_exact: the exact
// test's registers stay out of the sweep.  Four cubes per lane: a wave's
// step reads four 1 KB row segments (tsdf and weight of rows j, j + 1) and
// the per-step work (addresses, sums, loop) is shared by 240 cubes.
template <bool FULL, bool VEC>
__global__ __launch_bounds__(kMcBlock) void k_mc_count(McGeom g, const float* __restrict__ tsdf,
                                                        const float* __restrict__ wt, uint32_t* __restrict__ sums,
                                                        uint32_t* __restrict__ exact_list) {
    __shared__ uint32_t s_lut[256];
    static_assert(kMcBlock == 256, "one lookup entry per thread");
    s_lut[threadIdx.x] = kMcCaseLut.v[threadIdx.x];
    __syncthreads();
    const RunPos rp = run_pos4(g);
    if (!rp.live) return;
    const int lane = __lane_id();
    const int q = lane >> 4, sub = lane & 15;
    const int32_t r = rp.r + q;  // this 16-lane row's run
    const bool run_ok = r < g.nr;
    const int32_t x0 = r * kRun + 4 * sub;
    uint32_t ownm = 0;  // bit c: the cube at x0 + c is this lane's
#pragma unroll
    for (int c = 0; c < 4; ++c) ownm |= (uint32_t)(run_ok && sub < kRunLanes && x0 + c < g.cx) << c;
    const int32_t k1 = rp.k0 + kMarch < g.cz ? rp.k0 + kMarch : g.cz;
    // a ring of four slices as in the one-cube sweep: step k reads slices k
    // and k + 1, then reloads slice k's registers with slice k + 4 (past the
    // segment's end the last slice used, k1, again; a ring of three, at five
    // waves per SIMD instead of three, was no faster)
    Slice4 s0 = load_slice4<VEC>(g, tsdf, wt, x0, rp.j, rp.k0);
    Slice4 s1 = load_slice4<VEC>(g, tsdf, wt, x0, rp.j, rp.k0 + 1);
    Slice4 s2 = load_slice4<VEC>(g, tsdf, wt, x0, rp.j, min(rp.k0 + 2, k1));
    Slice4 s3 = load_slice4<VEC>(g, tsdf, wt, x0, rp.j, min(rp.k0 + 3, k1));
    uint32_t pending = 0;          // the previous step's count, 16 bits, reduced with this one
    uint32_t my_sum = 0;           // lane 16 q + t: the sum of run r + q at step k0 + t
    unsigned long long flagged = 0;  // bit 16 q + t: run r + q at step k0 + t has a cube for k_mc_exact
    const auto step = [&](int32_t k, const Slice4& lo, const Slice4& hi) {
        // per column: above-iso flags (rows j, j + 1 of slices k, k + 1) and a
        // zero-weight bit; column 4 is the next lane's column 0
        uint32_t code[5];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            code[c] = (lo.t0[c] > g.iso ? 1u : 0u) | (lo.t1[c] > g.iso ? 2u : 0u) | (hi.t0[c] > g.iso ? 4u : 0u) |
                      (hi.t1[c] > g.iso ? 8u : 0u) |
                      ((lo.w0[c] == 0.0f || lo.w1[c] == 0.0f || hi.w0[c] == 0.0f || hi.w1[c] == 0.0f) ? 16u : 0u);
        }
        code[4] = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)code[0], 0x130, 0xf, 0xf, false);
        uint32_t e[4], cs[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const bool zero = ((code[c] | code[c + 1]) & 16u) != 0u;
            e[c] = s_lut[(code[c] & 15u) | ((code[c + 1] & 15u) << 4)];
            cs[c] = zero ? 0u : (e[c] & 255u);
        }
        const int t = k - rp.k0;
        uint32_t n = 0;
        if (!FULL) {
#pragma unroll
            for (int c = 0; c < 4; ++c)
                n += ((ownm >> c) & 1u) && ((code[c] | code[c + 1]) & 16u) == 0u ? (e[c] >> 8) & 255u : 0u;
        } else {
            uint32_t surf = 0;
#pragma unroll
            for (int c = 0; c < 4; ++c) surf |= (uint32_t)(((ownm >> c) & 1u) && cs[c] != 0u && cs[c] != 255u) << c;
            // most steps of most waves hold no surface cube: the screen is
            // skipped for the whole wave
            if (__ballot(surf != 0u) != 0ull) {
                // 3 per triangle where the screen rules out snapped edge
                // points (then no triangle is degenerate); else the run is
                // flagged and k_mc_exact counts it
                float dmn[5], dmx[5];
                uint32_t bad[5];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const float d0 = fabsf(lo.t0[c] - g.iso), d1 = fabsf(lo.t1[c] - g.iso);
                    const float d2 = fabsf(hi.t0[c] - g.iso), d3 = fabsf(hi.t1[c] - g.iso);
                    bad[c] = (!(d0 <= 3.0e38f) || !(d1 <= 3.0e38f) || !(d2 <= 3.0e38f) || !(d3 <= 3.0e38f)) ? 1u : 0u;
                    dmn[c] = fminf(fminf(d0, d1), fminf(d2, d3));
                    dmx[c] = fmaxf(fmaxf(d0, d1), fmaxf(d2, d3));
                }
                dmn[4] = from_next_lane(dmn[0]);
                dmx[4] = from_next_lane(dmx[0]);
                bad[4] = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)bad[0], 0x130, 0xf, 0xf, false);
                bool exact_any = false;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const float m = fminf(dmn[c], dmn[c + 1]), M = fmaxf(dmx[c], dmx[c + 1]);
                    const bool surface = ((surf >> c) & 1u) != 0u;
                    const bool exact = surface && ((bad[c] | bad[c + 1]) != 0u || !screen_ok(g, m, M));
                    n += surface && !exact ? 3u * ((e[c] >> 16) & 255u) : 0u;
                    exact_any = exact_any || exact;
                }
                const unsigned long long em = __ballot(exact_any);
#pragma unroll
                for (int rq = 0; rq < 4; ++rq)
                    if ((em >> (16 * rq)) & 0xffffull) flagged |= 1ull << (16 * rq + t);
            }
        }
        // two steps per row reduction: a run's count is below 2^16 (60 cubes
        // x at most 15 vertices)
        const bool odd = (t & 1) != 0;
        if (!odd && k + 1 < k1) {
            pending = n;
        } else {
            const uint32_t tot = row_inclusive_sum(odd ? (pending | (n << 16)) : n);
            uint32_t both = (uint32_t)__builtin_amdgcn_readlane((int)tot, 15);
            both = q == 1 ? (uint32_t)__builtin_amdgcn_readlane((int)tot, 31) : both;
            both = q == 2 ? (uint32_t)__builtin_amdgcn_readlane((int)tot, 47) : both;
            both = q == 3 ? (uint32_t)__builtin_amdgcn_readlane((int)tot, 63) : both;
            if (odd) {
                my_sum = sub == t - 1 ? (both & 0xffffu) : my_sum;
                my_sum = sub == t ? (both >> 16) : my_sum;
            } else {
                my_sum = sub == t ? both : my_sum;
            }
        }
    };
    for (int32_t k = rp.k0; k < k1; k += 4) {
        step(k, s0, s1);
        s0 = load_slice4<VEC>(g, tsdf, wt, x0, rp.j, min(k + 4, k1));
        if (k + 1 < k1) step(k + 1, s1, s2);
        s1 = load_slice4<VEC>(g, tsdf, wt, x0, rp.j, min(k + 5, k1));
        if (k + 2 < k1) step(k + 2, s2, s3);
        s2 = load_slice4<VEC>(g, tsdf, wt, x0, rp.j, min(k + 6, k1));
        if (k + 3 < k1) step(k + 3, s3, s0);
        s3 = load_slice4<VEC>(g, tsdf, wt, x0, rp.j, min(k + 7, k1));
    }
    // run q of step k0 + t: q0 + t * cy * nr
    const uint64_t kstride = (uint64_t)g.cy * (uint64_t)g.nr;
    const uint64_t q0 = ((uint64_t)rp.k0 * (uint64_t)g.cy + (uint64_t)rp.j) * (uint64_t)g.nr + (uint64_t)r;
    if (run_ok && sub < k1 - rp.k0) sums[q0 + (uint64_t)sub * kstride] = my_sum;
    if (FULL && flagged) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(exact_list, (uint32_t)__popcll(flagged));
        base = (uint32_t)__shfl((int)base, 0);
        if ((flagged >> lane) & 1ull)
            exact_list[1u + base + (uint32_t)__popcll(flagged & ((1ull << lane) - 1ull))] =
                (uint32_t)(q0 + (uint64_t)sub * kstride);
    }
}

// The exact sums of the runs k_mc_count flagged (list: count, then run
// numbers q), a wave per run.
__global__ __launch_bounds__(kMcBlock) void k_mc_exact(McGeom g, const float* __restrict__ tsdf,
                                                        const float* __restrict__ wt, uint32_t* __restrict__ sums,
                                                        const uint32_t* __restrict__ exact_list) {
    const uint32_t n = exact_list[0];
    const int lane = __lane_id();
    const uint32_t waves = gridDim.x * (kMcBlock / 64);
    for (uint32_t w = blockIdx.x * (kMcBlock / 64) + (threadIdx.x >> 6); w < n; w += waves) {
        const uint32_t q = exact_list[1u + w];
        const int32_t r = (int32_t)(q % (uint32_t)g.nr);
        const uint32_t jk = q / (uint32_t)g.nr;
        const int32_t j = (int32_t)(jk % (uint32_t)g.cy);
        const int32_t k = (int32_t)(jk / (uint32_t)g.cy);
        const int32_t x = r * kRun + lane;
        const bool own = lane < kRun && x < g.cx;
        const Slice lo = load_slice(g, tsdf, wt, x, j, k);
        const Slice hi = load_slice(g, tsdf, wt, x, j, k + 1);
        Cube cb;
        make_cube(g, x, j, k, lo, hi, &cb);
        unsigned keep;
        const uint32_t c = own ? (uint32_t)cube_outputs<true>(g, cb, keep) : 0u;
        const uint32_t total = wave_sum(c);
        if (lane == 0) sums[q] = total;
    }
}

template <bool FULL>
__global__ __launch_bounds__(kMcBlock) void k_mc_emit(McGeom g, const float* __restrict__ tsdf,
                                                       const float* __restrict__ wt,
                                                       const uint32_t* __restrict__ sums,
                                                       const uint32_t* __restrict__ offsets,
                                                       const uint32_t* __restrict__ mesh_size_in,
                                                       const uint32_t* __restrict__ faces_size_in,
                                                       float* __restrict__ mesh, uint64_t mesh_cap,
                                                       int32_t* __restrict__ faces, uint64_t faces_cap) {
    // FULL: this thread's edge points, [edge][axis][thread]
    __shared__ float s_ep[FULL ? 12 * 3 * kMcBlock : 1];
    const RunPos rp = run_pos(g);
    if (!rp.live) return;
    const int lane = __lane_id();
    const int32_t x = rp.r * kRun + lane;
    const bool own = lane < kRun && x < g.cx;
    const int32_t k1 = rp.k0 + kMarch < g.cz ? rp.k0 + kMarch : g.cz;
    // the column's run sums and offsets, lane l holding step k0 + l; most
    // runs hold no surface and are skipped without reading the volume
    const uint64_t kstride = (uint64_t)g.cy * (uint64_t)g.nr;
    const uint64_t q0 = ((uint64_t)rp.k0 * (uint64_t)g.cy + (uint64_t)rp.j) * (uint64_t)g.nr + (uint64_t)rp.r;
    uint32_t my_sum = 0, my_off = 0;
    if (lane < k1 - rp.k0) {
        my_sum = sums[q0 + (uint64_t)lane * kstride];
        my_off = offsets[q0 + (uint64_t)lane * kstride];
    }
    unsigned long long steps = __ballot(my_sum != 0u);
    // the next step with output: its slices are loaded while this one
    // computes and writes (a step right after the last reuses its upper slice)
    Slice nlo{}, nhi{};
    if (steps) {
        const int32_t k = rp.k0 + __builtin_ctzll(steps);
        nlo = load_slice(g, tsdf, wt, x, rp.j, k);
        nhi = load_slice(g, tsdf, wt, x, rp.j, k + 1);
    }
    while (steps) {
        const int kk = __builtin_ctzll(steps);
        steps &= steps - 1ull;
        const int32_t k = rp.k0 + kk;
        const uint32_t run_off = __shfl(my_off, kk);
        const Slice lo = nlo, hi = nhi;
        if (steps) {
            const int kn = __builtin_ctzll(steps);
            nlo = kn == kk + 1 ? hi : load_slice(g, tsdf, wt, x, rp.j, rp.k0 + kn);
            nhi = load_slice(g, tsdf, wt, x, rp.j, rp.k0 + kn + 1);
        }
        Cube cb;
        make_cube(g, x, rp.j, k, lo, hi, &cb);
        unsigned keep = 0;
        uint32_t n = 0;
        if (own) n = (uint32_t)cube_outputs<FULL>(g, cb, keep);
        const uint32_t prefix = wave_exclusive_sum(n, lane);
        if (!n) continue;
        const uint64_t rel = (uint64_t)run_off + prefix;  // this call's vertex index
        const uint64_t v0 = (uint64_t)*mesh_size_in + rel;  // appended after the caller's count
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
            continue;
        }
        const unsigned long long t = kMcTri[cb.cs];
        const int nt = (int)(t & 15);
        const int mask = edge_mask(cb.cs);
        float* ep = s_ep + threadIdx.x;
#pragma unroll
        for (int e = 0; e < 12; ++e) {
            if (mask & (1 << e)) {
                float o[3];
                edge_point(g, cb, e, o);
                ep[(3 * e + 0) * kMcBlock] = o[0];
                ep[(3 * e + 1) * kMcBlock] = o[1];
                ep[(3 * e + 2) * kMcBlock] = o[2];
            }
        }
        uint64_t v = v0;
        uint64_t f = (uint64_t)*faces_size_in + rel / 3;
        for (int qq = 0; qq < nt; ++qq) {
            if (!(keep & (1u << qq))) continue;
            for (int s = 0; s < 3; ++s) {
                const int e = (int)((t >> (4 + 4 * (3 * qq + s))) & 15);
                if (v + s < mesh_cap) {
                    mesh[3 * (v + s)] = ep[(3 * e + 0) * kMcBlock];
                    mesh[3 * (v + s) + 1] = ep[(3 * e + 1) * kMcBlock];
                    mesh[3 * (v + s) + 2] = ep[(3 * e + 2) * kMcBlock];
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
}

__global__ void k_mc_finish(const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ sums, uint64_t nq,
                            uint32_t* mesh_size, uint32_t* faces_size) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t total = offsets[nq - 1] + sums[nq - 1];
    *mesh_size = *mesh_size + total;
    if (faces_size) *faces_size = *faces_size + total / 3;
}

int run_mc(vf_workspace* ws, bool full, float* mesh, uint32_t* mesh_size, uint64_t mesh_cap, int32_t* faces,
           uint32_t* faces_size, uint64_t faces_cap, const float* tsdf, const float* wt, vf_int3 dims, vf_float3 vs,
           vf_float3 size, float iso, hipStream_t stream) {
    uint64_t n;
    int st = check_volume(dims, vs, &n);
    if (st) return st;
    if ((uint64_t)dims.x * (uint64_t)dims.y >= (1ull << 31))
        return fail(VF_ERR_INVALID_ARGUMENT, "marching cubes: a z slice of %d x %d voxels (at most 2^31 - 1)", dims.x,
                    dims.y);
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
    g.cz = dims.z - 2;
    g.nr = (g.cx + kRun - 1) / kRun;
    {
        // the smallest edge length and the largest |coordinate| over the
        // cubes' corners 0 .. c per axis, in the kernels' float arithmetic
        double lmin = 3.0e38, pmax = 0.0;
        const int32_t cmax[3] = {g.cx, g.cy, g.cz};
        for (int a = 0; a < 3; ++a) {
            float prev = (float)0 * g.vs[a] - g.halfs[a];
            pmax = std::max(pmax, std::fabs((double)prev));
            for (int32_t i = 1; i <= cmax[a]; ++i) {
                const float cur = (float)i * g.vs[a] - g.halfs[a];
                lmin = std::min(lmin, (double)(float)(cur - prev));
                pmax = std::max(pmax, std::fabs((double)cur));
                prev = cur;
            }
        }
        g.lmin = lmin > 0.0 ? lmin * (1.0 - 0x1p-30) : 0.0;
        g.pscale = 0x1p-19 * pmax * (1.0 + 0x1p-30);
    }
    const uint64_t nq = (uint64_t)g.cz * (uint64_t)g.cy * (uint64_t)g.nr;  // runs
    const uint64_t groups = ((uint64_t)g.cy + kMcBlock / 64 - 1) / (kMcBlock / 64);
    const uint64_t nb64 = (uint64_t)g.nr * groups * (uint64_t)((g.cz + kMarch - 1) / kMarch);
    if (nb64 > 0x7fffffffull || nq > 0x7fffffffull)
        return fail(VF_ERR_CAPACITY, "marching cubes: %llu runs", (unsigned long long)nq);
    const uint32_t nb = (uint32_t)nb64;
    // the count sweep: a workgroup per four runs (16-lane rows) of four rows
    const uint32_t nb4 =
        (uint32_t)(((uint64_t)g.nr + 3) / 4 * groups * (uint64_t)((g.cz + kMarch - 1) / kMarch));
    // 16-byte loads where every row starts 16-byte aligned
    const bool vec = (g.X & 3) == 0 && (((uintptr_t)tsdf | (uintptr_t)wt) & 15u) == 0;
    size_t scan_bytes = 0;
    VF_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)nq,
                                            stream));
    // scratch: run sums, their offsets, (faces) the flagged-run list (a
    // count, then at most one entry per run), the scan's temporary storage
    const uint64_t sums_bytes = nq * 4;
    const uint64_t list_off = 2 * sums_bytes;
    const uint64_t scan_off = (list_off + (full ? (nq + 1) * 4 : 0) + 255) & ~255ull;
    void* scratch = nullptr;
    st = workspace_scratch(ws, scan_off + scan_bytes, &scratch, stream);
    if (st) return st;
    uint32_t* sums = (uint32_t*)scratch;
    uint32_t* offsets = sums + nq;
    uint32_t* exact_list = (uint32_t*)((char*)scratch + list_off);
    void* scan_tmp = (char*)scratch + scan_off;
    if (full) {
        VF_HIP(hipMemsetAsync(exact_list, 0, sizeof(uint32_t), stream));
        if (vec)
            hipLaunchKernelGGL((k_mc_count<true, true>), dim3(nb4), dim3(kMcBlock), 0, stream, g, tsdf, wt, sums,
                               exact_list);
        else
            hipLaunchKernelGGL((k_mc_count<true, false>), dim3(nb4), dim3(kMcBlock), 0, stream, g, tsdf, wt, sums,
                               exact_list);
        VF_LAUNCH_CHECK("k_mc_count");
        const uint32_t nb_exact = (uint32_t)std::min<uint64_t>(nb, 512);
        hipLaunchKernelGGL(k_mc_exact, dim3(nb_exact), dim3(kMcBlock), 0, stream, g, tsdf, wt, sums,
                           (const uint32_t*)exact_list);
        VF_LAUNCH_CHECK("k_mc_exact");
    } else {
        if (vec)
            hipLaunchKernelGGL((k_mc_count<false, true>), dim3(nb4), dim3(kMcBlock), 0, stream, g, tsdf, wt, sums,
                               (uint32_t*)nullptr);
        else
            hipLaunchKernelGGL((k_mc_count<false, false>), dim3(nb4), dim3(kMcBlock), 0, stream, g, tsdf, wt, sums,
                               (uint32_t*)nullptr);
        VF_LAUNCH_CHECK("k_mc_count");
    }
    VF_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, sums, offsets, (int)nq, stream));
    if (full) {
        hipLaunchKernelGGL(k_mc_emit<true>, dim3(nb), dim3(kMcBlock), 0, stream, g, tsdf, wt, sums, offsets,
                           mesh_size, faces_size, mesh, mesh_cap, faces, faces_cap);
    } else {
        hipLaunchKernelGGL(k_mc_emit<false>, dim3(nb), dim3(kMcBlock), 0, stream, g, tsdf, wt, sums, offsets,
                           mesh_size, (const uint32_t*)nullptr, mesh, mesh_cap, (int32_t*)nullptr, (uint64_t)0);
    }
    VF_LAUNCH_CHECK("k_mc_emit");
    hipLaunchKernelGGL(k_mc_finish, dim3(1), dim3(64), 0, stream, offsets, sums, nq, mesh_size,
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
