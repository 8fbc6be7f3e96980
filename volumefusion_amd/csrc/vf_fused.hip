// vf_fused.hip -- the warped TSDF integration of one frame (or a batch of
// independent volumes, one frame each) in five kernels (four with the brick
// screen off).
//
// The reference runs nine dense launches with ~100 B of per-voxel scratch
// (src/VolumeFusion.cu:699-742, buffers at :252-260).  Here:
//
//   k_prep_bin  one thread per (node, brick overlapped by the node's cube):
//               node record (translated position, cube corner, 2r, 2r^2 and
//               its reciprocal, transform rows); a slot in the brick's node
//               list; the brick's displacement bound.  Also clears the
//               collision tiles the previous frame flagged, and (extra
//               workgroups) takes the depth-tile maxima of the frame.
//   k_compact   two bricks per thread: the bricks with a non-empty node list
//               in brick order within each workgroup's run, each with its
//               place in the node-list pool and the brick screen's verdict;
//               bricks that may hold a pending voxel mark their target tiles.
//   k_select    the active list: the candidates the screen keeps (see the
//               brick screen below); fills the node lists (k_fill_lists
//               does that without the screen).
//   k_pass_a    one workgroup per active brick at a time, bricks drawn from
//               per-XCD queues (one voxel per thread, each wave a 4x4x4
//               block): stage the brick's node list in LDS (each cube as
//               brick-local byte ranges), per voxel select the canonical
//               top-K (ties to the smaller node index), weights, blend,
//               target voxel.
//               Collisions are counted in an LDS window of 8^3 tiles around
//               the brick (2 bits per voxel: seen once / seen twice) and
//               merged into the global bitmap once per brick, a word at a
//               time.  Voxels that pass the projective tests append a 16 B
//               pending record (voxel, target, new tsdf, new weight) to the
//               workgroup's own record arena (a brick's records together:
//               there, else in a shared overflow arena, else nowhere and
//               the brick joins the redo list).
//   (redo)      only with a record budget: pass A again over the redo list,
//               writing each pending voxel whose target was reached once.
//   k_pass_b    one thread per pending record: drop it if its target was
//               reached twice, else write tsdf and weight.
//
// Unsupported voxels (no node cube covers their brick) are never touched, so
// HBM traffic scales with the supported shell, not with the dense volume;
// the screen also skips supported bricks that cannot change the frame.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include <hip/hip_ext.h>

#include "vf_host.hpp"

namespace vf {
namespace {

constexpr int kCap = 128;       // node slots per brick list (and chunk size when a list spills)
constexpr int kSlotBits = 7;    // list slot bits of a knn id (node << kSlotBits | slot)
constexpr uint32_t kSlotMask = (1u << kSlotBits) - 1u;
static_assert(kCap == 1 << kSlotBits, "a list slot fits the id's low bits");
constexpr int kMargin = 8;      // collision window reaches 8 voxels beyond the brick
constexpr int kStatSlots = 64;  // statistics counters are spread over 64 addresses
constexpr uint32_t kNone = 0xffffffffu;

// Tiles of 8 voxels that cover [x0 - kMargin, x0 + b - 1 + kMargin] for every
// brick origin x0 (a multiple of b); origins repeat modulo lcm(b, 8) <= 8 b.
constexpr int window_tiles(int b) {
    int m = 0;
    for (int x0 = 0; x0 < 8 * b; x0 += b) {
        const int lo = (x0 - kMargin) >= 0 ? (x0 - kMargin) / 8 : -((kMargin - x0 + 7) / 8);
        const int t = (x0 + b - 1 + kMargin) / 8 - lo + 1;
        m = t > m ? t : m;
    }
    return m;
}

// Brick shape: one voxel per thread, x fastest, one brick per workgroup step.
template <int BX_, int BY_, int BZ_>
struct Shape {
    static constexpr int BX = BX_, BY = BY_, BZ = BZ_;
    static constexpr int kThreads = BX * BY * BZ;
    static constexpr int WX = window_tiles(BX), WY = window_tiles(BY), WZ = window_tiles(BZ);
    static constexpr int kWinTiles = WX * WY * WZ;
    static_assert(kThreads % 64 == 0 && kThreads >= kCap && kThreads <= 1024, "brick must be whole waves");
};

// The brick shapes the kernels are built for (VF_BRICK_* in volumefusion.h).
using Shape0 = Shape<16, 8, 4>;  // 512 threads, 64 B rows
using Shape1 = Shape<16, 4, 4>;  // 256 threads
using Shape2 = Shape<8, 8, 4>;   // 256 threads
using Shape3 = Shape<8, 8, 8>;   // 512 threads
using Shape4 = Shape<32, 4, 4>;  // 512 threads, 128 B rows
constexpr int kNumShapes = 5;
constexpr int kShapeDims[kNumShapes][3] = {{16, 8, 4}, {16, 4, 4}, {8, 8, 4}, {8, 8, 8}, {32, 4, 4}};

enum {
    kStSupported,
    kStInBounds,
    kStPending,
    kStCollided,
    kStUpdated,
    kStActive,
    kStSpilled,
    kStScreened,
    kStRedo,
    kStCount = 9
};

// Depth-tile maxima for the brick screen: tiles of 16 and of 128 pixels.
constexpr int kTile0 = 16;
constexpr int kTile1 = 128;

struct VolDesc {
    Geom g;
    Camera cam;
    float c[3];     // node / vertex translation (the frame centroid)
    float rinv[9];  // inverse volume rotation, Eigen order (identity in the reference's call)
    const float* depth;
    float* tsdf;
    float* wt;
    const float* nodes;  // Vec3f[G]
    float* nodes_out;    // optional round-trip write-back
    const float* radii;  // float[G]
    const float* T;      // Mat4f[G], column-major
    int32_t cube[3];
    int32_t bdim[3];
    // floor(b / bdim[0]) and floor(b / (bdim[0] bdim[1])) for b < 2^24 as
    // (b * bmag[i]) >> bsh[i] (Granlund and Montgomery, PLDI 1994:
    // bsh = 24 + ceil(log2 d), bmag = floor(2^bsh / d) + 1)
    uint32_t bmag[2];
    uint32_t bsh[2];
    int32_t tdim[3];
    int32_t maxb[3];
    uint32_t G;
    uint32_t node_base;
    uint32_t brick_base;
    uint32_t tile_base;
    uint32_t voxel_base;  // first voxel of this volume in the batch-wide voxel numbering
    uint32_t k1_base;
    uint32_t k1_count;
    int32_t blend;  // VF_BLEND_LBS or VF_BLEND_DQB
    // Brick screen: pass A sweeps only the bricks that can hold a pending
    // voxel or whose voxels can reach the target of one (screen != 0).
    int32_t screen;
    uint32_t pyr_blk;    // first depth-tile workgroup of this volume in k_prep_bin's grid
    uint32_t pyr_base0;  // first 16-pixel tile maximum (pw0 x ph0, row-major)
    uint32_t pyr_base1;  // first 128-pixel tile maximum (pw1 x ph1)
    int32_t pw0, ph0, pw1, ph1;
};

struct NodeRec {
    float4 m;   // translated position; .w = 2r (the candidate bound)
    int4 lo;    // cube lower corner; .w unused
    float4 t0;  // transform rows (T(i,0), T(i,1), T(i,2), T(i,3))
    float4 t1;
    float4 t2;
    float4 w;   // .x = 2*(r*r) (weight denominator), .y = RN(1 / .x)
};

constexpr uint32_t kQueues = 8;       // pass-A brick queues, one per XCD
constexpr uint32_t kQueueChunk = 16;  // list entries per queue chunk
constexpr uint32_t kQueueStride = 64; // counters 256 bytes apart

// A list's length (low half) and, when k_compact appends to it, the node-list
// pool entries handed out with it (high half): one 64-bit atomic reserves both.
__device__ __forceinline__ uint32_t run_length(unsigned long long run) { return (uint32_t)run; }

struct Ctrl {
    unsigned long long active_run[2];  // active list, by frame parity
    unsigned long long cand_run[2];    // bricks with a node list before the screen, by frame parity
    uint32_t queue[2][kQueues * kQueueStride];  // pass-A queue counters by frame parity
    uint32_t overflow_records;  // reserved in the shared overflow arena (may exceed its capacity)
    uint32_t redo_count;        // bricks whose records fit no arena, recomputed by the redo pass
    uint32_t frame;  // frames completed (k_pass_b counts); its parity serves replayed launches
};

// Frame parity: from the host, or (parity_arg < 0) from the device frame
// counter, so that a recorded launch sequence can be replayed unchanged.
__device__ __forceinline__ uint32_t frame_parity(const Ctrl* ctrl, int parity_arg) {
    return parity_arg >= 0 ? (uint32_t)parity_arg : (ctrl->frame & 1u);
}

struct KnnOut {
    uint32_t* idx;
    float* dist;
    int32_t* cnt;
};

struct Arena {
    uint4* records;
    uint32_t* counts;      // per pass-A workgroup
    uint32_t cap;          // records per workgroup arena
    uint32_t n_arenas;
    uint4* overflow;       // shared overflow arena (atomic append)
    uint32_t overflow_cap;
};

__device__ __forceinline__ void stat_add(unsigned long long* stat, unsigned long long* totals, int counter,
                                         unsigned long long v) {
    if (!v) return;
    const int slot = counter * kStatSlots + (int)(blockIdx.x & (kStatSlots - 1));
    atomicAdd(stat + slot, v);
    atomicAdd(totals + slot, v);
}

constexpr int kPrepThreads = 1024;
constexpr int kMaxBatch = 256;  // volumes per call

// ---- brick screen ----------------------------------------------------------
//
// A supported voxel matters to the frame only if it is pending (passes the
// projective tests of src/TSDFIntegration.cu:32-58) or if its warped target is
// the target of a pending voxel (the count > 1 rule, :252-299).  The screen
// proves both false for whole bricks and drops them before pass A:
//
//  * Every candidate node's warp moves a voxel of its cube by at most D_j
//    (per axis): T_j p - p is affine in p, so |T_j p - p| is largest at a
//    corner of the cube (clipped to the brick).  The LBS blend is a convex
//    combination of the candidates' T_j p (weights exp(.) in [e^-2, 1],
//    normalised), so the warped point lies within D_b = max_j D_j of the
//    voxel, up to rounding (covered by a relative slack of 2^-16).
//  * The brick's warped points lie in a box; no point of the box passes the
//    projective tests when the box is behind the largest valid depth over its
//    pixel footprint by more than delta (depth tile maxima), or has no pixel
//    in the frame, or lies in front of z = 1e-4.
//  * The targets lie in the box's voxel range (+-1).  Pending-capable bricks
//    mark the 8^3 tiles of their target range; any other brick is dropped
//    when its target range meets no marked tile: none of its voxels can share
//    a target with a pending voxel.
//
// Dropped bricks change no tsdf, weight or collision decision.  Statistics of
// supported and in-bounds voxels then count the swept bricks only.
// The screen is off for DQB volumes (the blend is not a convex combination)
// and for nodes whose weights could be NaN.

constexpr uint32_t kInfBits = 0x7f800000u;  // +inf as float bits: unbounded displacement

// D_j over the voxels [vlo, vhi] of node j's cube, in binary64: per axis a,
// f_a(p) = (T p)_a - p_a is affine in p, so its largest magnitude over the
// box is |f_a(centre)| + sum_k |T_ak - [a == k]| halfwidth_k; plus a
// rounding slack of 2^-16 times a bound on the terms of the float product
// (sum_k |T_ak| P_k + |T_a3| + P_a, P_k the largest |p_k| on the box).
// +inf when the node's weights could leave [e^-2, 1] (radius not positive or
// 2r^2 outside the Markstein-safe range), its position or transform is not
// finite, or the bound does not fit a float.
__device__ uint32_t disp_bound_bits(const VolDesc& vd, const float* M, const float m[3], float r,
                                    const int32_t vlo[3], const int32_t vhi[3]) {
    if (!(r > 0.0f) || !den_safe(2.0f * (r * r))) return kInfBits;
    if (!__builtin_isfinite(m[0]) || !__builtin_isfinite(m[1]) || !__builtin_isfinite(m[2])) return kInfBits;
    float e[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) e[i] = M[i];  // column-major: M(a, j) = e[4 j + a]
    bool fin = true;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int j = 0; j < 4; ++j) fin = fin && __builtin_isfinite(e[4 * j + a]);
    if (!fin) return kInfBits;
    double c[3], h[3], P[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double off = (double)vd.c[a] - (double)vd.g.halfs[a], vs = (double)vd.g.vs[a];
        const double lo = (double)vlo[a] * vs + off, hi = (double)vhi[a] * vs + off;
        c[a] = 0.5 * (lo + hi);
        h[a] = 0.5 * (hi - lo);
        P[a] = fmax(fabs(lo), fabs(hi));
    }
    double dmax = 0.0, mag = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double f = (double)e[12 + a] - c[a], w = 0.0, s = fabs((double)e[12 + a]) + P[a];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double t = (double)e[4 * k + a];
            f += t * c[k];
            w += fabs(t - (a == k ? 1.0 : 0.0)) * h[k];
            s += fabs(t) * P[k];
        }
        dmax = fmax(dmax, fabs(f) + w);
        mag = fmax(mag, s);
    }
    const double D = dmax * (1.0 + 0x1p-40) + mag * 0x1p-16 + 0x1p-40;
    if (!(D <= 1.0e38)) return kInfBits;
    float fD = (float)D;
    if ((double)fD < D) fD = __int_as_float(__float_as_int(fD) + 1);  // round up
    return __float_as_uint(fD);
}

// A depth pixel as the screen sees it: the projective test rejects infinite
// and too small depths (-inf here: never the maximum); NaN passes the test
// (its sdf compares false both ways), so it counts as +inf.
__device__ __forceinline__ float depth_key(float d) {
    if (d != d) return kInf;
    if (__builtin_isinf(d) || d < 1.e-4f) return -kInf;
    return d;
}

// Tile maxima of one 128 x 128 pixel block of a volume's frame: the 64
// 16-pixel tiles and the one 128-pixel tile.  One workgroup of kPrepThreads.
__device__ void depth_tiles(const VolDesc* __restrict__ vols, int nvol, uint32_t blk, float* __restrict__ dmax) {
    __shared__ uint32_t s_pb[kMaxBatch];
    __shared__ float s_seg[kTile1][kTile1 / kTile0];
    static_assert(kPrepThreads == kTile1 * (kTile1 / kTile0), "one thread per 16-pixel row segment");
    const int tid = threadIdx.x;
    for (int v = tid; v < nvol; v += blockDim.x) s_pb[v] = vols[v].pyr_blk;
    __syncthreads();
    int v = 0, hi = nvol - 1;  // last volume whose first block is <= blk
    while (v < hi) {
        const int mid = (v + hi + 1) >> 1;
        if (s_pb[mid] <= blk)
            v = mid;
        else
            hi = mid - 1;
    }
    const VolDesc& vd = vols[v];
    const uint32_t local = blk - vd.pyr_blk;
    const int32_t bx = (int32_t)(local % (uint32_t)vd.pw1), by = (int32_t)(local / (uint32_t)vd.pw1);
    const int32_t W = vd.cam.W, H = vd.cam.H;
    const int row = tid / (kTile1 / kTile0), seg = tid % (kTile1 / kTile0);
    const int32_t y = by * kTile1 + row;
    const int32_t x0 = bx * kTile1 + seg * kTile0;
    float m = -kInf;
    if (y < H) {
        const float* rowp = vd.depth + (int64_t)y * W;
#pragma unroll
        for (int i = 0; i < kTile0; ++i)
            if (x0 + i < W) m = fmaxf(m, depth_key(gload(rowp + x0 + i)));
    }
    s_seg[row][seg] = m;
    __syncthreads();
    if (tid < 64) {
        const int ty = tid / (kTile1 / kTile0), tx = tid % (kTile1 / kTile0);
        float t = -kInf;
#pragma unroll
        for (int r = 0; r < kTile0; ++r) t = fmaxf(t, s_seg[ty * kTile0 + r][tx]);
        const int32_t gx = bx * (kTile1 / kTile0) + tx, gy = by * (kTile1 / kTile0) + ty;
        if (gx < vd.pw0 && gy < vd.ph0) dmax[vd.pyr_base0 + (uint32_t)(gy * vd.pw0 + gx)] = t;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) t = fmaxf(t, __shfl_xor(t, o));
        if (tid == 0) dmax[vd.pyr_base1 + (uint32_t)(by * vd.pw1 + bx)] = t;
    }
}

// The prep kernel: workgroups [0, pair_blocks) take the (node, brick) items;
// workgroups [pair_blocks, gridDim.x) compute the depth-tile maxima of the
// screened volumes' frames instead (one 128 x 128 pixel block each).
template <class S>
__global__ __launch_bounds__(kPrepThreads) void k_prep_bin(
    const VolDesc* __restrict__ vols, int nvol, uint32_t total, uint32_t pair_blocks, NodeRec* __restrict__ nrec,
    uint32_t* __restrict__ brick_count, uint2* __restrict__ items, uint32_t* __restrict__ brick_disp,
    float* __restrict__ dmax, Ctrl* __restrict__ ctrl, uint32_t* __restrict__ bitmap,
    uint32_t* __restrict__ tile_flag, uint32_t clear_tiles, unsigned long long* __restrict__ stat) {
    if (blockIdx.x >= pair_blocks) {
        depth_tiles(vols, nvol, blockIdx.x - pair_blocks, dmax);
        return;
    }
    __shared__ uint32_t s_k1[kMaxBatch];  // first (node, brick) item of each volume
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nthreads = pair_blocks * blockDim.x;
    for (int v = threadIdx.x; v < nvol; v += blockDim.x) s_k1[v] = vols[v].k1_base;

    // Clear the collision tiles the previous frame on this workspace flagged
    // (tiles [0, clear_tiles) hold every flag it can have set).
    for (uint32_t t = gid; t < clear_tiles; t += nthreads) {
        if (tile_flag[t]) {
            uint4* w = (uint4*)(bitmap + (uint64_t)t * 32u);
#pragma unroll
            for (int k = 0; k < 8; ++k) w[k] = make_uint4(0u, 0u, 0u, 0u);
            tile_flag[t] = 0u;
        }
    }
    if (gid < kStCount * kStatSlots) stat[gid] = 0ull;
    if (gid == 0) {
        ctrl->overflow_records = 0u;
        ctrl->redo_count = 0u;
    }
    __syncthreads();

    // This thread's (node, brick) pair: the node takes a slot in the brick's
    // list, recorded as (brick, node << 7 | slot) for the list fill once
    // k_compact has placed every list in the pool ({kNone, 0}: no entry).
    uint2 entry = make_uint2(kNone, 0u);
    do {
        if (gid >= total) break;
        int v = 0, hi = nvol - 1;  // last volume whose first item is <= gid
        while (v < hi) {
            const int mid = (v + hi + 1) >> 1;
            if (s_k1[mid] <= gid)
                v = mid;
            else
                hi = mid - 1;
        }
        const VolDesc& vd = vols[v];
        const uint32_t local = gid - vd.k1_base;
        if (local >= vd.k1_count) break;
        const uint32_t nb = (uint32_t)(vd.maxb[0] * vd.maxb[1] * vd.maxb[2]);
        const uint32_t j = local / nb;
        const uint32_t o = local - j * nb;

        // translateKernel(nodes, -c), src/VolumeFusion.cu:708
        const float* np = vd.nodes + 3ull * j;
        float m[3];
        m[0] = np[0] + (-vd.c[0]);
        m[1] = np[1] + (-vd.c[1]);
        m[2] = np[2] + (-vd.c[2]);
        // cube corner, src/TSDFWarp.cu:46-48
        int32_t nv[3];
        world_to_voxel(vd.g, m, nv);
        int32_t lo[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            long long l = (long long)nv[a] - (long long)((vd.cube[a] - 1) / 2);
            if (l < -(1ll << 30)) l = -(1ll << 30);
            if (l > (1ll << 30)) l = (1ll << 30);
            lo[a] = (int32_t)l;
        }
        if (o == 0) {
            const float r = vd.radii[j];
            const float rr = r * r;
            const float den = 2.0f * rr;
            NodeRec rec;
            rec.m = make_float4(m[0], m[1], m[2], r * 2.0f);
            rec.lo = make_int4(lo[0], lo[1], lo[2], 0);
            if (vd.T && vd.blend == VF_BLEND_DQB) {
                // t0, t1: the node's dual quaternion (real, dual)
                mat4_to_dq(vd.T + 16ull * j, &rec.t0, &rec.t1);
                rec.t2 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            } else if (vd.T) {  // the knn-only entry point has no transforms
                const float* M = vd.T + 16ull * j;
                rec.t0 = make_float4(M[0], M[4], M[8], M[12]);
                rec.t1 = make_float4(M[1], M[5], M[9], M[13]);
                rec.t2 = make_float4(M[2], M[6], M[10], M[14]);
            } else {
                rec.t0 = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
                rec.t1 = make_float4(0.0f, 1.0f, 0.0f, 0.0f);
                rec.t2 = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
            }
            rec.w = make_float4(den, 1.0f / den, 0.0f, 0.0f);
            nrec[vd.node_base + j] = rec;
        }
        // bricks overlapped by the cube clipped to the volume
        int32_t blo[3], bhi[3];
        const int32_t bs[3] = {S::BX, S::BY, S::BZ};
        bool empty = false;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const long long l = lo[a] < 0 ? 0 : lo[a];
            long long h = (long long)lo[a] + vd.cube[a] - 1;
            if (h > vd.g.dim[a] - 1) h = vd.g.dim[a] - 1;
            empty = empty || l > h;
            blo[a] = (int32_t)(l / bs[a]);
            bhi[a] = (int32_t)(h / bs[a]);
        }
        if (empty) break;
        const int32_t ox = (int32_t)(o % (uint32_t)vd.maxb[0]);
        const int32_t oy = (int32_t)((o / (uint32_t)vd.maxb[0]) % (uint32_t)vd.maxb[1]);
        const int32_t oz = (int32_t)(o / (uint32_t)(vd.maxb[0] * vd.maxb[1]));
        const int32_t bx = blo[0] + ox, by = blo[1] + oy, bz = blo[2] + oz;
        if (bx > bhi[0] || by > bhi[1] || bz > bhi[2]) break;
        const uint32_t b = ((uint32_t)bz * (uint32_t)vd.bdim[1] + (uint32_t)by) * (uint32_t)vd.bdim[0] + (uint32_t)bx;
        const uint32_t gb = vd.brick_base + b;
        const uint32_t slot = atomicAdd(brick_count + gb, 1u);
        if (slot < (uint32_t)kCap) entry = make_uint2(gb, (j << kSlotBits) | slot);
        if (vd.screen) {
            // displacement bound of this node over its cube within this brick
            const int32_t bb[3] = {bx, by, bz};
            int32_t vlo[3], vhi[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int32_t o = bb[a] * bs[a];
                vlo[a] = lo[a] > o ? lo[a] : o;
                long long h = (long long)lo[a] + vd.cube[a] - 1;
                if (h > o + bs[a] - 1) h = o + bs[a] - 1;
                if (h > vd.g.dim[a] - 1) h = vd.g.dim[a] - 1;
                vhi[a] = (int32_t)h;
            }
            atomicMax(brick_disp + gb, disp_bound_bits(vd, vd.T + 16ull * j, m, vd.radii[j], vlo, vhi));
        }
    } while (false);
    if (gid < total) items[gid] = entry;
}

// The frame's active bricks (those with a non-empty node list), in brick
// order within each workgroup's run of kCompactPer * kPrepThreads bricks; one
// list append per workgroup (appends to one counter serialise, so a
// workgroup takes many bricks).
constexpr int kCompactPer = 2;
constexpr int kCompactWaves = kPrepThreads / 64;

// Order-preserving compaction of one workgroup's kCompactPer x kPrepThreads
// slots (bit r of `flags`: slot r * kPrepThreads + thread is kept): one
// atomic on *counter reserves the run, s_cnt receives the first position of
// each (round, wave) within it.  Returns the run's first list position.
__device__ __forceinline__ uint32_t compact_run(uint32_t flags, uint32_t* counter, uint32_t* s_cnt,
                                                uint32_t* s_base, uint32_t* s_total = nullptr) {
    const int lane = __lane_id();
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < kCompactPer; ++r) {
        const unsigned long long m = __ballot((flags >> r) & 1u);
        if (lane == 0) s_cnt[r * kCompactWaves + wave] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (wave == 0) {
        // exclusive prefix over (round, wave), round-major: two entries a lane
        static_assert(kCompactPer * kCompactWaves <= 128, "two table entries per lane of one wave");
        const int e0 = 2 * lane;
        const uint32_t c0 = e0 < kCompactPer * kCompactWaves ? s_cnt[e0] : 0u;
        const uint32_t c1 = e0 + 1 < kCompactPer * kCompactWaves ? s_cnt[e0 + 1] : 0u;
        const uint32_t incl = wave_inclusive_sum(c0 + c1);
        const uint32_t excl = incl - (c0 + c1);
        if (e0 < kCompactPer * kCompactWaves) s_cnt[e0] = excl;
        if (e0 + 1 < kCompactPer * kCompactWaves) s_cnt[e0 + 1] = excl + c0;
        const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
        if (lane == 0) {
            *s_base = total ? atomicAdd(counter, total) : 0u;
            if (s_total) *s_total = total;
        }
    }
    __syncthreads();
    return *s_base;
}

// compact_run for k_compact, which also places the run's node lists in the
// pool: v is this thread's pool entries, *voff receives their first entry
// (an exclusive prefix over the workgroup plus the run's pool base).  The run
// and its pool entries are reserved together by one 64-bit atomic on *run
// (low half: list positions, high half: pool entries; neither half can carry
// into the other, both stay below 2^32), so the workgroup waits for one
// contended atomic instead of two in a row.
__device__ __forceinline__ uint32_t compact_alloc(uint32_t flags, uint32_t v, unsigned long long* run,
                                                  uint32_t* s_cnt, uint32_t* s_w, uint2* s_base,
                                                  uint32_t* s_total, uint32_t* voff) {
    const int lane = __lane_id();
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < kCompactPer; ++r) {
        const unsigned long long m = __ballot((flags >> r) & 1u);
        if (lane == 0) s_cnt[r * kCompactWaves + wave] = (uint32_t)__popcll(m);
    }
    const uint32_t incl = wave_inclusive_sum(v);
    if (lane == 63) s_w[wave] = incl;
    __syncthreads();
    if (wave == 0) {
        static_assert(kCompactPer * kCompactWaves <= 128, "two table entries per lane of one wave");
        static_assert(kCompactWaves <= 64, "one wave total a lane");
        const int e0 = 2 * lane;
        const uint32_t c0 = e0 < kCompactPer * kCompactWaves ? s_cnt[e0] : 0u;
        const uint32_t c1 = e0 + 1 < kCompactPer * kCompactWaves ? s_cnt[e0 + 1] : 0u;
        const uint32_t ci = wave_inclusive_sum(c0 + c1);
        const uint32_t ce = ci - (c0 + c1);
        if (e0 < kCompactPer * kCompactWaves) s_cnt[e0] = ce;
        if (e0 + 1 < kCompactPer * kCompactWaves) s_cnt[e0 + 1] = ce + c0;
        const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)ci, 63);
        const uint32_t x = lane < kCompactWaves ? s_w[lane] : 0u;
        const uint32_t xi = wave_inclusive_sum(x);
        if (lane < kCompactWaves) s_w[lane] = xi - x;
        const uint32_t ptot = (uint32_t)__builtin_amdgcn_readlane((int)xi, 63);
        if (lane == 0) {
            // a thread with pool entries has a listed brick, so ptot > 0 implies total > 0
            const unsigned long long old =
                total ? atomicAdd(run, ((unsigned long long)ptot << 32) | total) : 0ull;
            *s_base = make_uint2((uint32_t)old, (uint32_t)(old >> 32));
            *s_total = total;
        }
    }
    __syncthreads();
    const uint2 b = *s_base;
    *voff = b.y + s_w[wave] + (incl - v);
    return b.x;
}

// List position of this thread's slot r after compact_run (whole waves call it).
__device__ __forceinline__ uint32_t compact_pos(uint32_t base, const uint32_t* s_cnt, int r, bool keep) {
    const unsigned long long m = __ballot(keep);
    return base + s_cnt[r * kCompactWaves + (threadIdx.x >> 6)] +
           (uint32_t)__popcll(m & ((1ull << __lane_id()) - 1ull));
}

constexpr uint32_t kScPend = 1u;   // may hold a pending voxel: kept, marks its target tiles
constexpr uint32_t kScAll = 2u;    // kept, and every brick of its volume is kept
constexpr uint32_t kScEmpty = 4u;  // no target inside the volume: dropped
constexpr uint32_t kScKeep = 8u;   // screen off for the volume: kept
constexpr int kMaxScreenTiles = 64;

// Lanes that select one candidate together (k_select) and that screen one
// brick together (k_compact).  Measured at 512^3 (prep, compaction and
// selection together): 39.9 us with one lane per screened brick, 41.1, 44.5
// and 51.6 us with 2, 4 and 8; the per-brick work is short, and fewer lanes
// per brick mean fewer rounds in the workgroups whose run is dense.
constexpr int kGrp = 8;
constexpr int kGrpC = 1;

// Largest of x over the aligned group of GL lanes (every lane of the group
// active).
template <int GL>
__device__ __forceinline__ float group_max(float x) {
#pragma unroll
    for (int o = 1; o < GL; o <<= 1) x = fmaxf(x, __shfl_xor(x, o, GL));
    return x;
}

// Any of the group's lanes (every lane of the group active).
template <int GL>
__device__ __forceinline__ bool group_any(bool x) {
    const unsigned long long m = __ballot(x);
    return ((m >> (__lane_id() & (64 - GL))) & ((1ull << GL) - 1ull)) != 0ull;
}

// The screen's verdict on brick b of vd with displacement bound dbits: its
// target range in 8^3 tiles (lo | hi << 16 per axis in .x, .y, .z) and flags
// in .w.  Called by the GL lanes of a group together (q: the lane's rank in
// the group); each computes the same box, the depth-tile loads are split
// over the group.
template <class S, int GL>
__device__ uint4 screen_brick(const VolDesc& vd, uint32_t b, uint32_t dbits, const float* __restrict__ dmax,
                              int q) {
    if (!vd.screen) return make_uint4(0u, 0u, 0u, kScKeep);
    if (dbits >= kInfBits) return make_uint4(0u, 0u, 0u, kScPend | kScAll);
    const uint32_t bd0 = (uint32_t)vd.bdim[0], bd1 = (uint32_t)vd.bdim[1];
    const int32_t bi[3] = {(int32_t)(b % bd0), (int32_t)((b / bd0) % bd1), (int32_t)(b / (bd0 * bd1))};
    const int32_t bs[3] = {S::BX, S::BY, S::BZ};
    const double D = (double)__uint_as_float(dbits);
    double lo[3], hi[3];  // box of the brick's warped voxel positions (camera frame)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int32_t v0 = bi[a] * bs[a];
        const int32_t v1 = v0 + bs[a] - 1 < vd.g.dim[a] - 1 ? v0 + bs[a] - 1 : vd.g.dim[a] - 1;
        const double h = (double)vd.g.halfs[a], c = (double)vd.c[a], vs = (double)vd.g.vs[a];
        lo[a] = ((double)v0 * vs - h + c) - D;
        hi[a] = ((double)v1 * vs - h + c) + D;
    }
    // targets worldToVoxel(Rinv (p - c)), src/TSDFIntegration.cu:272-279,
    // over the box by interval arithmetic, +-1 voxel for rounding
    uint32_t tr[3];
    bool empty = false;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double olo = 0.0, ohi = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double r = (double)vd.rinv[i + 3 * j];
            const double ql = r * (lo[j] - (double)vd.c[j]), qh = r * (hi[j] - (double)vd.c[j]);
            olo += fmin(ql, qh);
            ohi += fmax(ql, qh);
        }
        const double fl = (olo + (double)vd.g.halfs[i]) * (double)vd.g.inv[i];
        const double fh = (ohi + (double)vd.g.halfs[i]) * (double)vd.g.inv[i];
        const double vl = fmax(floor(fl - fabs(fl) * 0x1p-20) - 1.0, 0.0);
        const double vh = fmin(ceil(fh + fabs(fh) * 0x1p-20) + 1.0, (double)(vd.g.dim[i] - 1));
        if (!(vl <= vh)) {
            empty = true;
            tr[i] = 0u;
        } else {
            tr[i] = ((uint32_t)vl >> 3) | (((uint32_t)vh >> 3) << 16);
        }
    }
    if (empty) return make_uint4(0u, 0u, 0u, kScEmpty);
    // projective tests, src/TSDFIntegration.cu:16-58
    const Camera& cam = vd.cam;
    bool pend;
    if (hi[2] < 0.99e-4) {
        pend = false;  // every point in front of z = 1e-4
    } else if (!(lo[2] > 1.0e-3)) {
        pend = true;  // footprint unbounded near z = 0
    } else {
        const double fx = cam.fx, fy = cam.fy;
        double u0 = fx * lo[0] / lo[2], u1 = u0, v0 = fy * lo[1] / lo[2], v1 = v0;
        const double us[3] = {fx * lo[0] / hi[2], fx * hi[0] / lo[2], fx * hi[0] / hi[2]};
        const double vs_[3] = {fy * lo[1] / hi[2], fy * hi[1] / lo[2], fy * hi[1] / hi[2]};
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            u0 = fmin(u0, us[q]);
            u1 = fmax(u1, us[q]);
            v0 = fmin(v0, vs_[q]);
            v1 = fmax(v1, vs_[q]);
        }
        // pixel = int(u + 0.5f) (truncation), +-1 for rounding
        const double px0 = fmax(floor(u0 + (double)cam.cx + 0.5) - 1.0, 0.0);
        const double px1 = fmin(ceil(u1 + (double)cam.cx + 0.5) + 1.0, (double)cam.W - 1.0);
        const double py0 = fmax(floor(v0 + (double)cam.cy + 0.5) - 1.0, 0.0);
        const double py1 = fmin(ceil(v1 + (double)cam.cy + 0.5) + 1.0, (double)cam.H - 1.0);
        if (!__builtin_isfinite(px0 + px1 + py0 + py1)) {
            pend = true;
        } else if (!(px0 <= px1 && py0 <= py1)) {
            pend = false;  // no pixel of the frame
        } else {
            const int32_t x0 = (int32_t)px0, x1 = (int32_t)px1, y0 = (int32_t)py0, y1 = (int32_t)py1;
            // the largest valid depth over the footprint: 16-pixel tiles when
            // at most 16 of them cover it, else 128-pixel tiles (at most 40),
            // else unbounded; lane q of the group loads tiles q, q + GL, ...,
            // eight at a time
            float md = -kInf;
            int32_t tx0 = x0 / kTile0, ty0 = y0 / kTile0, nx = x1 / kTile0 - tx0 + 1, ny = y1 / kTile0 - ty0 + 1;
            uint32_t base = vd.pyr_base0, pw = (uint32_t)vd.pw0;
            if (nx * ny > 16) {
                tx0 = x0 / kTile1, ty0 = y0 / kTile1, nx = x1 / kTile1 - tx0 + 1, ny = y1 / kTile1 - ty0 + 1;
                base = vd.pyr_base1, pw = (uint32_t)vd.pw1;
            }
            if (nx * ny > 40) {
                md = kInf;
            } else {
                for (int32_t t0 = q; t0 < nx * ny; t0 += 8 * GL) {
                    float v[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int32_t t = t0 + u * GL;
                        v[u] = t < nx * ny ? dmax[base + (uint32_t)((ty0 + t / nx) * (int32_t)pw + tx0 + t % nx)]
                                           : -kInf;
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) md = fmaxf(md, v[u]);
                }
                md = group_max<GL>(md);
            }
            // every sdf = z - d > delta: z >= lo[2] and d <= md, and a margin
            // above one ulp of delta keeps the float subtraction above delta
            const double thr = (double)cam.delta + fabs((double)cam.delta) * 0x1p-20 + 0x1p-40;
            pend = !(lo[2] - (double)md > thr);
        }
    }
    return make_uint4(tr[0], tr[1], tr[2], pend ? kScPend : 0u);
}

__device__ __forceinline__ uint32_t tile_span(uint32_t r) { return (r >> 16) - (r & 0xffffu) + 1u; }

// The frame's active bricks (those with a non-empty node list), in brick
// order within each workgroup's run of kCompactPer * kPrepThreads bricks; one
// list append per workgroup (appends to one counter serialise, so a
// workgroup takes many bricks).  Each brick whose list holds at most kCap
// nodes gets its place in the node-list pool (one pool allocation per
// workgroup); the lists are filled after this kernel.  With SCREEN they go to the candidate list
// with the screen's verdict, and pending-capable bricks mark their target
// tiles (tag: this frame's number + 1) for k_select.  The run's candidates
// are first packed in LDS, then screened by groups of kGrpC lanes, so that
// they occupy the first waves of the workgroup rather than scattered lanes
// of every wave.
template <class S, bool SCREEN>
__global__ __launch_bounds__(kPrepThreads) void k_compact(
    const VolDesc* __restrict__ vols, int nvol, uint32_t nbricks, const uint32_t* __restrict__ brick_count,
    uint32_t* __restrict__ brick_off, const uint32_t* __restrict__ brick_disp, const float* __restrict__ dmax,
    uint32_t* __restrict__ active2, uint2* __restrict__ active_oc2,
    uint32_t* __restrict__ cand, uint4* __restrict__ cand_info, uint32_t* __restrict__ tile_mark,
    uint32_t* __restrict__ vol_mark, Ctrl* __restrict__ ctrl, uint64_t brick_cap, int parity_arg) {
    __shared__ uint32_t s_cnt[kCompactPer * kCompactWaves];  // per (round, wave): count, then first position
    __shared__ uint32_t s_total;
    __shared__ uint2 s_base;              // the run's first list position and first pool entry
    __shared__ uint32_t s_lw[kCompactWaves];
    __shared__ uint32_t s_bb[kMaxBatch];  // first brick of each volume
    __shared__ uint32_t s_item[SCREEN ? kCompactPer * kPrepThreads : 1];  // the run's candidates, in list order
    for (int v = threadIdx.x; v < nvol; v += blockDim.x) s_bb[v] = vols[v].brick_base;
    const uint32_t parity = frame_parity(ctrl, parity_arg);
    const uint64_t gb0 = (uint64_t)blockIdx.x * (kPrepThreads * kCompactPer) + threadIdx.x;
    uint32_t flags = 0;  // bit r: brick gb0 + r * kPrepThreads is active
    uint32_t lw[kCompactPer], lsum = 0;  // pool entries of each brick (0 for a spilled list)
    uint2 oc[kCompactPer];               // each brick's (first pool entry, list count)
#pragma unroll
    for (int r = 0; r < kCompactPer; ++r) {
        const uint64_t gb = gb0 + (uint64_t)(r * kPrepThreads);
        const uint32_t c = gb < nbricks ? brick_count[gb] : 0u;
        flags |= (uint32_t)(c != 0u) << r;
        lw[r] = c <= (uint32_t)kCap ? c : 0u;
        lsum += lw[r];
        oc[r] = make_uint2(kNone, c);
    }
    uint32_t off;
    const uint32_t base = compact_alloc(flags, lsum, SCREEN ? &ctrl->cand_run[parity] : &ctrl->active_run[parity],
                                        s_cnt, s_lw, &s_base, &s_total, &off);
#pragma unroll
    for (int r = 0; r < kCompactPer; ++r) {
        // a spilled list has no place: its entries are dropped
        if (lw[r]) oc[r].x = off;
        if ((flags >> r) & 1u) brick_off[gb0 + (uint64_t)(r * kPrepThreads)] = oc[r].x;
        off += lw[r];
    }
#pragma unroll
    for (int r = 0; r < kCompactPer; ++r) {
        const bool act = (flags >> r) & 1u;
        const uint32_t pos = compact_pos(base, s_cnt, r, act);
        if (act) {
            const uint32_t gb = (uint32_t)(gb0 + (uint64_t)(r * kPrepThreads));
            int lo = 0, hi = nvol - 1;  // last volume whose first brick is <= gb
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_bb[mid] <= gb)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            const uint32_t item = ((uint32_t)lo << 24) | (gb - s_bb[lo]);
            if (!SCREEN) {
                active2[(uint64_t)parity * brick_cap + pos] = item;
                active_oc2[(uint64_t)parity * brick_cap + pos] = oc[r];
            } else {
                s_item[pos - base] = item;
            }
        }
    }
    if (!SCREEN) return;
    __syncthreads();
    const uint32_t tag = ctrl->frame + 1u;
    const int q = (int)(threadIdx.x % kGrpC);
    for (uint32_t c = threadIdx.x / kGrpC; c < s_total; c += kPrepThreads / kGrpC) {
        const uint32_t item = s_item[c];
        const uint32_t v = item >> 24, local = item & 0xffffffu;
        const VolDesc& vd = vols[v];
        const uint32_t gb = s_bb[v] + local;
        const uint4 info = screen_brick<S, kGrpC>(vd, local, brick_disp[gb], dmax, q);
        if (q == 0) {
            cand[base + c] = item;
            cand_info[base + c] = info;
        }
        if (info.w & kScPend) {
            const uint32_t nx = tile_span(info.x), nxy = nx * tile_span(info.y);
            const uint32_t nt = nxy * tile_span(info.z);
            if ((info.w & kScAll) || nt > (uint32_t)kMaxScreenTiles) {
                if (q == 0) vol_mark[v] = tag;
            } else {
                const uint32_t x0 = info.x & 0xffffu, y0 = info.y & 0xffffu, z0 = info.z & 0xffffu;
                for (uint32_t t = (uint32_t)q; t < nt; t += kGrpC) {
                    const uint32_t tz = z0 + t / nxy, ty = y0 + (t % nxy) / nx, tx = x0 + t % nx;
                    tile_mark[vd.tile_base + (tz * (uint32_t)vd.tdim[1] + ty) * (uint32_t)vd.tdim[0] + tx] = tag;
                }
            }
        }
    }
}

// The node lists: every (node, brick) entry of k_prep_bin into the brick's
// place in the pool, at its slot (a brick whose list spilled has none).
// Four entries a thread at a time: their loads, then their offsets' loads,
// then the stores (two dependent loads in flight for all four).
__device__ __forceinline__ void fill_lists(const uint2* __restrict__ items, uint32_t total,
                                           const uint32_t* __restrict__ brick_off, uint32_t* __restrict__ brick_nodes,
                                           uint32_t gid, uint32_t nthreads) {
    constexpr int kU = 4;
    for (uint64_t i0 = gid; i0 < total; i0 += (uint64_t)kU * nthreads) {
        uint2 e[kU];
        uint32_t off[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const uint64_t i = i0 + (uint64_t)u * nthreads;
            e[u] = i < total ? items[i] : make_uint2(kNone, 0u);
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) off[u] = e[u].x != kNone ? brick_off[e[u].x] : kNone;
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (off[u] != kNone) brick_nodes[off[u] + (e[u].y & kSlotMask)] = e[u].y >> kSlotBits;
    }
}

// The list fill alone, for frames without the brick screen (k_select does it
// otherwise).
__global__ __launch_bounds__(kPrepThreads) void k_fill_lists(const uint2* __restrict__ items, uint32_t total,
                                                             const uint32_t* __restrict__ brick_off,
                                                             uint32_t* __restrict__ brick_nodes) {
    fill_lists(items, total, brick_off, brick_nodes, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// The active list from the candidates: a brick stays when the screen kept it
// or when its target tiles meet a tile a pending-capable brick marked (or its
// volume is marked whole).  Dropped bricks get their node-list count cleared
// here (pass A clears the others); every candidate's displacement bound is
// cleared for the next frame.  A group of kGrp lanes takes one candidate
// (lane q tests tiles q, q + 8, ...); a workgroup takes kSelPerBlock
// candidates at a time, in list order, so the active list keeps the
// candidates' order.  The workgroups also fill the node lists.
constexpr uint32_t kSelPerRound = kPrepThreads / kGrp;
constexpr uint32_t kSelPerBlock = kCompactPer * kSelPerRound;

__global__ __launch_bounds__(kPrepThreads) void k_select(
    const VolDesc* __restrict__ vols, const uint32_t* __restrict__ cand, const uint4* __restrict__ cand_info,
    uint32_t* __restrict__ brick_count, uint32_t* __restrict__ brick_disp, const uint32_t* __restrict__ tile_mark,
    const uint32_t* __restrict__ vol_mark, uint32_t* __restrict__ active2, uint2* __restrict__ active_oc2,
    Ctrl* __restrict__ ctrl, uint64_t brick_cap, int parity_arg, const uint2* __restrict__ items, uint32_t total_items,
    const uint32_t* __restrict__ brick_off, uint32_t* __restrict__ brick_nodes, unsigned long long* __restrict__ stat,
    unsigned long long* __restrict__ totals) {
    __shared__ uint32_t s_cnt[kCompactPer * kCompactWaves];
    __shared__ uint32_t s_base;
    __shared__ unsigned long long s_drop;
    if (threadIdx.x == 0) s_drop = 0ull;
    fill_lists(items, total_items, brick_off, brick_nodes, blockIdx.x * blockDim.x + threadIdx.x,
               gridDim.x * blockDim.x);
    const uint32_t parity = frame_parity(ctrl, parity_arg);
    const uint32_t n = run_length(ctrl->cand_run[parity]);
    const uint32_t tag = ctrl->frame + 1u;
    const int q = (int)(threadIdx.x % kGrp);
    uint32_t dropped = 0;
    // the loop bound is the same for every thread of the workgroup
    for (uint64_t c0 = (uint64_t)blockIdx.x * kSelPerBlock; c0 < n; c0 += (uint64_t)gridDim.x * kSelPerBlock) {
        uint32_t flags = 0, items[kCompactPer];
        uint2 ocs[kCompactPer];
#pragma unroll
        for (int r = 0; r < kCompactPer; ++r) {
            const uint64_t i = c0 + (uint64_t)r * kSelPerRound + threadIdx.x / kGrp;
            items[r] = 0u;
            ocs[r] = make_uint2(0u, 0u);
            if (i >= n) continue;
            const uint32_t item = cand[i];
            const uint4 info = cand_info[i];
            const uint32_t v = item >> 24;
            const VolDesc& vd = vols[v];
            const uint32_t gb = vd.brick_base + (item & 0xffffffu);
            bool keep;
            if (info.w & (kScPend | kScAll | kScKeep)) {
                keep = true;
            } else if (info.w & kScEmpty) {
                keep = false;
            } else if (vol_mark[v] == tag) {
                keep = true;
            } else if (tile_span(info.x) * tile_span(info.y) * tile_span(info.z) > (uint32_t)kMaxScreenTiles) {
                keep = true;
            } else {
                const uint32_t x0 = info.x & 0xffffu, y0 = info.y & 0xffffu, z0 = info.z & 0xffffu;
                const uint32_t nx = tile_span(info.x), nxy = nx * tile_span(info.y);
                const uint32_t nt = nxy * tile_span(info.z);
                static_assert(kMaxScreenTiles <= 8 * kGrp, "eight tiles per lane cover the range");
                uint32_t m[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const uint32_t t = (uint32_t)q + (uint32_t)(u * kGrp);
                    m[u] = 0u;
                    if (t < nt) {
                        const uint32_t tz = z0 + t / nxy, ty = y0 + (t % nxy) / nx, tx = x0 + t % nx;
                        m[u] = tile_mark[vd.tile_base + (tz * (uint32_t)vd.tdim[1] + ty) * (uint32_t)vd.tdim[0] + tx];
                    }
                }
                bool hit = false;
#pragma unroll
                for (int u = 0; u < 8; ++u) hit = hit || m[u] == tag;
                keep = group_any<kGrp>(hit);
            }
            if (q == 0) {
                brick_disp[gb] = 0u;
                if (!keep) {
                    brick_count[gb] = 0u;
                    ++dropped;
                } else {
                    ocs[r] = make_uint2(brick_off[gb], brick_count[gb]);
                }
                items[r] = item;
                flags |= (uint32_t)keep << r;
            }
        }
        // the low half of the 64-bit run counter (little-endian)
        const uint32_t base = compact_run(flags, (uint32_t*)&ctrl->active_run[parity], s_cnt, &s_base);
#pragma unroll
        for (int r = 0; r < kCompactPer; ++r) {
            const bool keep = (flags >> r) & 1u;
            const uint32_t pos = compact_pos(base, s_cnt, r, keep);
            if (keep) {
                active2[(uint64_t)parity * brick_cap + pos] = items[r];
                active_oc2[(uint64_t)parity * brick_cap + pos] = ocs[r];
            }
        }
        __syncthreads();  // s_cnt and s_base are rewritten by the next round
    }
    if (dropped) atomicAdd(&s_drop, (unsigned long long)dropped);
    __syncthreads();
    if (threadIdx.x == 0) stat_add(stat, totals, kStScreened, s_drop);
}

// A node cube as the ranges [A, B] of brick-local coordinates it covers per
// axis, for a brick at origin o: A = lo - o clipped to [0, 33], B = lo + c -
// 1 - o, also at most dim - 1 - o (the volume's last voxel), clipped to
// [-1, 32]; one byte per axis, B biased by 128 (.y).  For local coordinates
// l in 0..31 packed the same way (lpk), bit 7 of every byte of
// (lpk + 0x808080) - A and of B - lpk is set exactly when A <= l <= B on
// that axis: no byte borrows from the next.
__device__ __forceinline__ uint2 cube_bytes(const int32_t lo[3], const int32_t o[3], const int32_t c[3],
                                            const int32_t dim[3]) {
    uint32_t apk = 0, bpk = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        long long la = (long long)lo[a] - o[a];
        long long lb = (long long)lo[a] + c[a] - 1 - o[a];
        if (lb > (long long)dim[a] - 1 - o[a]) lb = (long long)dim[a] - 1 - o[a];
        la = la < 0 ? 0 : (la > 33 ? 33 : la);
        lb = lb < -1 ? -1 : (lb > 32 ? 32 : lb);
        apk |= (uint32_t)la << (8 * a);
        bpk |= (uint32_t)(lb + 128) << (8 * a);
    }
    return make_uint2(apk, bpk);
}

__device__ __forceinline__ bool cube_covers(uint32_t lpk, uint32_t apk, uint32_t bpk) {
    return (((lpk + 0x808080u) - apk) & (bpk - lpk) & 0x808080u) == 0x808080u;
}

// Five waves per SIMD (at most 96 VGPRs): with k = 8 the compiler keeps one
// value of the sample path in scratch memory, which costs less than the
// occupancy of four waves (measured 267 us against 303 us per 512^3 frame).
//
// REDO: the redo pass.  Bricks whose pending records fit neither the
// workgroup's arena nor the shared overflow arena are appended to the redo
// list (their records are not stored); once pass A has finished, and with it
// the collision bitmap, this instantiation recomputes them and writes each
// pending voxel whose target was reached once straight into the volume, as
// k_pass_b would have.  A voxel's new value depends only on its own tsdf and
// weight, which no other kernel of the frame writes, so the result is the
// same whichever way a brick's records go.
template <int K, bool KNN_ONLY, class S, bool REDO = false>
__global__ __launch_bounds__(S::kThreads) __attribute__((amdgpu_waves_per_eu(5))) void k_pass_a(
    const VolDesc* __restrict__ vols, const NodeRec* __restrict__ nrec, uint32_t* __restrict__ brick_count,
    const uint32_t* __restrict__ brick_nodes, const uint32_t* __restrict__ brick_off,
    const uint32_t* __restrict__ active2, const uint2* __restrict__ active_oc2, uint32_t* __restrict__ redo,
    Ctrl* __restrict__ ctrl,
    uint32_t* __restrict__ bitmap, uint32_t* __restrict__ tile_flag, uint64_t brick_cap, int parity_arg,
    Arena arena, unsigned long long* __restrict__ stat,
    unsigned long long* __restrict__ totals, KnnOut kout) {
    static_assert(!(KNN_ONLY && REDO), "the knn-only frame has no records");
    constexpr int kPassAThreads = S::kThreads;
    constexpr int kBX = S::BX, kBY = S::BY, kBZ = S::BZ;
    constexpr int kWX = S::WX, kWY = S::WY, kWZ = S::WZ;
    constexpr int kWinTiles = S::kWinTiles;
    __shared__ float4 s_m[kCap + 1];
    __shared__ int4 s_lo[kCap + 1];
    __shared__ float4 s_t0[kCap];
    __shared__ float4 s_t1[kCap];
    __shared__ float4 s_t2[kCap];
    __shared__ float2 s_w[kCap];
    __shared__ double s_exp2[32];
    __shared__ unsigned long long s_keys[K * kPassAThreads];
    __shared__ uint32_t s_bm[KNN_ONLY || REDO ? 1 : kWinTiles * 32];
    __shared__ uint32_t s_wave_emit[kPassAThreads / 64];
    __shared__ unsigned long long s_stat[5];
    __shared__ uint32_t s_next, s_first, s_ovf;

    const int tid = threadIdx.x;
    const int lane = __lane_id();
    const int wave = tid >> 6;
    const uint32_t parity = frame_parity(ctrl, parity_arg);
    const uint32_t* __restrict__ active = REDO ? redo : active2 + (uint64_t)parity * brick_cap;
    // each active brick's (first node-list entry, list count), read with the
    // entry (the redo pass reads them from the brick arrays)
    const uint2* __restrict__ active_oc = active_oc2 + (uint64_t)parity * brick_cap;
    const auto brick_oc = [&](uint32_t a_i, uint32_t gbi) -> uint2 {
        return REDO ? make_uint2(brick_off[gbi], brick_count[gbi]) : active_oc[a_i];
    };
    if (tid < 5) s_stat[tid] = 0ull;
    if (tid < 32) s_exp2[tid] = kExp2Table[tid];
    if (!KNN_ONLY && !REDO) {
        for (int w = tid; w < kWinTiles * 32; w += kPassAThreads) s_bm[w] = 0u;
    }
    if (!REDO && blockIdx.x == 0 && tid == 0) {
        // the other parity's list and counters belong to the previous frame,
        // whose kernels have finished
        ctrl->active_run[parity ^ 1u] = 0ull;
        ctrl->cand_run[parity ^ 1u] = 0ull;
        for (uint32_t q = 0; q < kQueues; ++q) ctrl->queue[parity ^ 1u][q * kQueueStride] = 0u;
    }
    const uint32_t nactive = REDO ? ctrl->redo_count : run_length(ctrl->active_run[parity]);

    // Each wave takes a 4x4x4 block of the brick: neighbour counts vary
    // less inside a compact block than along a 64-voxel slab, and a wave's
    // per-neighbour work is set by its largest count.
    static_assert(kBX % 4 == 0 && kBY % 4 == 0 && kBZ % 4 == 0, "bricks are tiled by 4x4x4 wave blocks");
    constexpr int kWavesX = kBX / 4, kWavesY = kBY / 4;
    const int lane_in = tid & 63, wave_in = tid >> 6;
    const int lx = (wave_in % kWavesX) * 4 + (lane_in & 3);
    const int ly = ((wave_in / kWavesX) % kWavesY) * 4 + ((lane_in >> 2) & 3);
    const int lz = (wave_in / (kWavesX * kWavesY)) * 4 + (lane_in >> 4);
    static_assert(kBX <= 32 && kBY <= 32 && kBZ <= 32, "brick-local coordinates fit the byte-parallel cube test");
    const uint32_t lpk = (uint32_t)lx | ((uint32_t)ly << 8) | ((uint32_t)lz << 16);
    uint32_t fill = 0;  // records in this workgroup's arena (uniform)

    // Dynamic schedule over kQueues queues: queue x owns the list chunks
    // x, x + kQueues, ... of kQueueChunk entries and hands out its entries in
    // order from a counter of its own.  Workgroup w draws from queue w % 8,
    // the queue of its XCD (workgroups are dispatched to the 8 XCDs
    // round-robin), so an XCD works through runs of neighbouring list entries
    // and no workgroup waits on a fixed share; one counter for all workgroups
    // measured as a serialising same-address hot spot.  Each grab is issued
    // one brick ahead and read a brick later.  The brick after the current
    // one is fetched while the current one computes: its entry, its list
    // count and, for this thread's list slot, the node id and the node record.
    // The redo pass takes its few bricks round-robin.
    const uint32_t qx = blockIdx.x % kQueues;
    uint32_t* const qctr = &ctrl->queue[parity][qx * kQueueStride];
    auto queue_pos = [qx](uint32_t j) -> uint32_t {
        return ((j / kQueueChunk) * kQueues + qx) * kQueueChunk + j % kQueueChunk;
    };
    if (tid == 0) {
        if (REDO) {
            s_first = blockIdx.x;
            s_next = blockIdx.x + gridDim.x;
        } else {
            s_first = queue_pos(atomicAdd(qctr, 1u));
            s_next = queue_pos(atomicAdd(qctr, 1u));
        }
    }
    __syncthreads();
    uint32_t a = s_first;
    uint32_t item = a < nactive ? active[a] : kNone;
    uint32_t cnt = 0, pre_id = 0;
    const float4 z4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float4 pm = z4, plo = z4, pt0 = z4, pt1 = z4, pt2 = z4, pw = z4;  // this thread's prefetched node record
    if (item != kNone) {
        const VolDesc& v0 = vols[item >> 24];
        const uint32_t gb0 = v0.brick_base + (item & 0xffffffu);
        const uint2 oc0 = brick_oc(a, gb0);
        cnt = oc0.y;
        if (cnt <= (uint32_t)kCap && (uint32_t)tid < cnt) {
            pre_id = brick_nodes[oc0.x + tid];
            const float4* r = (const float4*)(nrec + v0.node_base + pre_id);
            pm = r[0];
            plo = r[1];
            pt0 = r[2];
            pt1 = r[3];
            pt2 = r[4];
            pw = r[5];
        }
    }
    __syncthreads();
    uint32_t a_n = s_next;

    while (item != kNone) {
        item = __builtin_amdgcn_readfirstlane(item);
        cnt = __builtin_amdgcn_readfirstlane(cnt);
        const uint32_t vol = item >> 24;
        const uint32_t b = item & 0xffffffu;
        const VolDesc& vd = vols[vol];
        const int32_t dimx = vd.g.dim[0], dimy = vd.g.dim[1], dimz = vd.g.dim[2];
        const uint32_t gb = vd.brick_base + b;
        const bool spill = cnt > (uint32_t)kCap;

        const uint32_t bq = (uint32_t)(((uint64_t)b * vd.bmag[0]) >> vd.bsh[0]);
        const int32_t bzi = (int32_t)(((uint64_t)b * vd.bmag[1]) >> vd.bsh[1]);
        const int32_t bxi = (int32_t)(b - bq * (uint32_t)vd.bdim[0]);
        const int32_t byi = (int32_t)(bq - (uint32_t)bzi * (uint32_t)vd.bdim[1]);
        const int32_t ox = bxi * kBX, oy = byi * kBY, oz = bzi * kBZ;  // brick origin
        const int32_t x = ox + lx, y = oy + ly, z = oz + lz;
        const int32_t o3[3] = {ox, oy, oz};
        const int32_t c3[3] = {vd.cube[0], vd.cube[1], vd.cube[2]};
        const int32_t d3[3] = {dimx, dimy, dimz};
        // voxels beyond the volume's last (x, y or z) are covered by no cube
        const uint32_t vidx = ((uint32_t)z * (uint32_t)dimy + (uint32_t)y) * (uint32_t)dimx + (uint32_t)x;

        // next brick: entry, count and this thread's list id; the grab of the
        // one after it
        const uint32_t item_n = a_n < nactive ? active[a_n] : kNone;
        uint32_t a_nn = 0;
        if (tid == 0 && item_n != kNone) a_nn = REDO ? a_n + gridDim.x : queue_pos(atomicAdd(qctr, 1u));
        uint32_t cnt_n = 0, id_n = 0, nbase_n = 0;
        if (item_n != kNone) {
            const VolDesc& vn = vols[item_n >> 24];
            const uint32_t gbn = vn.brick_base + (item_n & 0xffffffu);
            nbase_n = vn.node_base;
            const uint2 ocn = brick_oc(a_n, gbn);
            cnt_n = ocn.y;
            if (cnt_n <= (uint32_t)kCap && (uint32_t)tid < cnt_n) id_n = brick_nodes[ocn.x + tid];
        }

        float p[3];
        voxel_to_world(vd.g, x, y, z, p);
        KnnSlots<K, kPassAThreads> ks;
        ks.init(s_keys, tid);
        bool slow = true;  // the spilled path always takes IEEE divisions

        if (!spill) {
            // This brick's list with its node records, in list order (the
            // order the binning appended them).  Each entry carries the knn
            // id (node << 7) | slot: ids order as node indices do, so ties
            // in distance fall to the smaller node index, and the low bits
            // locate the entry's records here.
            if ((uint32_t)tid < cnt) {
                const int32_t lo3[3] = {__float_as_int(plo.x), __float_as_int(plo.y), __float_as_int(plo.z)};
                const uint2 ab = cube_bytes(lo3, o3, c3, d3);
                s_m[tid] = pm;
                s_lo[tid] = make_int4((int)ab.x, (int)ab.y, 0, (int)((pre_id << kSlotBits) | (uint32_t)tid));
                s_t0[tid] = pt0;
                s_t1[tid] = pt1;
                s_t2[tid] = pt2;
                s_w[tid] = make_float2(pw.x, pw.y);
            }
            // a node whose 2r^2 is outside the Markstein-safe range sends the
            // brick's weights through IEEE divisions
            slow = __syncthreads_or((uint32_t)tid < cnt && !den_safe(pw.x)) != 0;
            // Branch-free test of every list entry, the next entry's LDS
            // reads issued before this one's arithmetic (the arrays hold one
            // slot beyond kCap, so the read after the last entry stays inside).
            int4 l = s_lo[0];
            float4 nm = s_m[0];
#pragma unroll 2
            for (uint32_t e = 0; e < cnt; ++e) {
                const int4 lc = l;
                const float4 mc = nm;
                l = s_lo[e + 1];
                nm = s_m[e + 1];
                const bool inside = cube_covers(lpk, (uint32_t)lc.x, (uint32_t)lc.y);
                const float dx = p[0] - mc.x, dy = p[1] - mc.y, dz = p[2] - mc.z;
                const float s = dx * dx + (dy * dy + dz * dz);
                const float d = sqrt_rn(s);
                if (inside & !(d > mc.w)) ks.offer(d, (uint32_t)lc.w);
            }
        } else {
            // the list spilled: walk every node of the volume in chunks
            const uint32_t nchunks = (vd.G + kCap - 1) / kCap;
            for (uint32_t ch = 0; ch < nchunks; ++ch) {
                const uint32_t base = ch * (uint32_t)kCap;
                const uint32_t m = vd.G - base < (uint32_t)kCap ? vd.G - base : (uint32_t)kCap;
                __syncthreads();
                if ((uint32_t)tid < m) {
                    const NodeRec rec = nrec[vd.node_base + base + tid];
                    const int32_t lo3[3] = {rec.lo.x, rec.lo.y, rec.lo.z};
                    const uint2 ab = cube_bytes(lo3, o3, c3, d3);
                    s_m[tid] = rec.m;
                    s_lo[tid] = make_int4((int)ab.x, (int)ab.y, 0, 0);
                }
                __syncthreads();
                for (uint32_t e = 0; e < m; ++e) {
                    const int4 l = s_lo[e];
                    if (cube_covers(lpk, (uint32_t)l.x, (uint32_t)l.y)) {
                        const float4 nm = s_m[e];
                        const float dx = p[0] - nm.x, dy = p[1] - nm.y, dz = p[2] - nm.z;
                        const float s = dx * dx + (dy * dy + dz * dz);
                        const float d = sqrt_rn(s);
                        if (!(d > nm.w)) ks.offer(d, base + e);
                    }
                }
            }
        }
        // the list count is cleared for the next frame; pass A clears it
        // after the record step, as the brick may go to the redo pass
        if (tid == 0) {
            if (KNN_ONLY || REDO) brick_count[gb] = 0u;
            if (spill && !REDO) s_stat[3] += 1ull;
        }

        const int n = ks.n;
        struct {
            float d[K];
            uint32_t i[K];
        } kl;
        ks.finish(kl.d, kl.i);

        // the next brick's node record for this thread's list slot
        float4 pm_n = z4, plo_n = z4, pt0_n = z4, pt1_n = z4, pt2_n = z4, pw_n = z4;
#define VF_FETCH_NEXT_RECORD()                                                         \
    do {                                                                               \
        if (item_n != kNone && cnt_n <= (uint32_t)kCap && (uint32_t)tid < cnt_n) {     \
            const float4* r = (const float4*)(nrec + nbase_n + id_n);                  \
            pm_n = r[0];                                                               \
            plo_n = r[1];                                                              \
            pt0_n = r[2];                                                              \
            pt1_n = r[3];                                                              \
            pt2_n = r[4];                                                              \
            pw_n = r[5];                                                               \
        }                                                                              \
    } while (0)

        if (KNN_ONLY) {
            VF_FETCH_NEXT_RECORD();
            if (n > 0) {
                kout.cnt[vidx] = n;
#pragma unroll
                for (int s = 0; s < K; ++s) {
                    if (s < n) {
                        kout.idx[(uint64_t)K * vidx + s] = spill ? kl.i[s] : kl.i[s] >> kSlotBits;
                        kout.dist[(uint64_t)K * vidx + s] = kl.d[s];
                    }
                }
            }
        } else {
        // collision window of this brick: kWX x kWY x kWZ tiles from (wx0, wy0, wz0)
        const int32_t wx0 = (bxi * kBX - kMargin) >> 3;
        const int32_t wy0 = (byi * kBY - kMargin) >> 3;
        const int32_t wz0 = (bzi * kBZ - kMargin) >> 3;

        bool inb = false, emit = false;
        uint32_t ucode = 0;
        float tnew = 0.0f, wnew = 0.0f;
        if (n > 0) {
            // computeTsdfWarpWeightsKernel body, src/TSDFWarp.cu:101-113
            float w[K];
            float sum = 0.0f;
#pragma unroll
            for (int s = 0; s < K; ++s) {
                w[s] = 0.0f;
                if (s < n) {
                    float2 dw;
                    if (spill) {
                        const float4 nw = nrec[vd.node_base + kl.i[s]].w;
                        dw = make_float2(nw.x, nw.y);
                    } else {
                        dw = s_w[kl.i[s] & kSlotMask];
                    }
                    w[s] = slow ? warp_weight(dw.x, kl.d[s]) : warp_weight_rcp(dw.x, dw.y, kl.d[s], s_exp2);
                    sum = sum + w[s];
                }
            }
            // isZero() of the normalised weights (src/TSDFWarp.cu:179-181):
            // every weight at most 1e-5 in magnitude.  Each weight is NaN or
            // in [-0, 1] (exp of a non-positive argument), so a zero sum means
            // every weight is zero, a NaN sum makes every normalised weight
            // NaN (not zero), and a positive sum of n <= 8 terms leaves the
            // largest normalised weight at least 1/8: the test is sum == 0.
            // The division by the sum (src/TSDFWarp.cu:112-113) is applied
            // to each selected weight where it is used below.
            const bool zero = sum == 0.0f;
            // vertex + c (src/VolumeFusion.cu:723), warp (src/TSDFWarp.cu:183-189)
            float pc[3];
            pc[0] = p[0] + vd.c[0];
            pc[1] = p[1] + vd.c[1];
            pc[2] = p[2] + vd.c[2];
            float acc[3] = {0.0f, 0.0f, 0.0f};
            if (zero) {
                acc[0] = pc[0];
                acc[1] = pc[1];
                acc[2] = pc[2];
            } else {
            float rs = rcp_rn(sum), sm = sum;
            if (slow) {
                // IEEE quotients (src/TSDFWarp.cu:112-113); then
                // div_rcp(w, 1, 1) below returns w unchanged
#pragma unroll
                for (int s = 0; s < K; ++s)
                    if (s < n) w[s] = w[s] / sum;
                sm = 1.0f;
                rs = 1.0f;
            }
            if (vd.blend == VF_BLEND_DQB) {
                // dual-quaternion blend of the selected nodes' rigid parts
                float4 br = make_float4(0.0f, 0.0f, 0.0f, 0.0f), bd = br;
                const float4 f = spill ? nrec[vd.node_base + kl.i[0]].t0 : s_t0[kl.i[0] & kSlotMask];
#pragma unroll
                for (int s = 0; s < K; ++s) {
                    if (s < n) {
                        float4 qr, qd;
                        if (spill) {
                            const NodeRec& rec = nrec[vd.node_base + kl.i[s]];
                            qr = rec.t0;
                            qd = rec.t1;
                        } else {
                            qr = s_t0[kl.i[s] & kSlotMask];
                            qd = s_t1[kl.i[s] & kSlotMask];
                        }
                        dqb_accumulate(qr, qd, f, div_rcp(w[s], sm, rs), &br, &bd);
                    }
                }
                dqb_transform(br, bd, pc, acc);
            } else {
#pragma unroll
                for (int s = 0; s < K; ++s) {
                    if (s < n) {
                        float4 r0, r1, r2;
                        if (spill) {
                            const NodeRec& rec = nrec[vd.node_base + kl.i[s]];
                            r0 = rec.t0;
                            r1 = rec.t1;
                            r2 = rec.t2;
                        } else {
                            r0 = s_t0[kl.i[s] & kSlotMask];
                            r1 = s_t1[kl.i[s] & kSlotMask];
                            r2 = s_t2[kl.i[s] & kSlotMask];
                        }
                        const float ws = div_rcp(w[s], sm, rs);
                        float t[3];
                        affine_apply(r0, r1, r2, pc, t);
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            const float wt = ws * t[a];
                            acc[a] = acc[a] + wt;
                        }
                    }
                }
            }
            }
            // collision target (src/TSDFIntegration.cu:272-282)
            int32_t u[3];
            inb = warped_target(vd.g, vd.rinv, vd.c, acc, u);
            if (inb) {
                const int32_t tx = u[0] >> 3, ty = u[1] >> 3, tz = u[2] >> 3;
                const uint32_t bit = ((uint32_t)(u[2] & 7) << 6) | ((uint32_t)(u[1] & 7) << 3) | (uint32_t)(u[0] & 7);
                const uint32_t tloc = ((uint32_t)tz * (uint32_t)vd.tdim[1] + (uint32_t)ty) * (uint32_t)vd.tdim[0] +
                                      (uint32_t)tx;
                ucode = ((vd.tile_base + tloc) << 9) | bit;
                const uint32_t mask = 1u << (bit & 31u);
                const uint32_t lx_t = (uint32_t)(tx - wx0), ly_t = (uint32_t)(ty - wy0), lz_t = (uint32_t)(tz - wz0);
                if (REDO) {
                    // pass A has counted this brick's targets
                } else if (lx_t < (uint32_t)kWX && ly_t < (uint32_t)kWY && lz_t < (uint32_t)kWZ) {
                    uint32_t* blk = s_bm + ((lz_t * kWY + ly_t) * kWX + lx_t) * 32u;
                    const uint32_t old = atomicOr(blk + (bit >> 5), mask);
                    if (old & mask) atomicOr(blk + 16 + (bit >> 5), mask);
                } else {
                    // target beyond the window: straight to the global bitmap
                    const uint32_t tile = vd.tile_base + tloc;
                    uint32_t* blk = bitmap + (uint64_t)tile * 32u;
                    const uint32_t old = atomicOr(blk + (bit >> 5), mask);
                    if (old & mask) atomicOr(blk + 16 + (bit >> 5), mask);
                    tile_flag[tile] = 1u;
                }
                // projective tests and the new sample (src/TSDFIntegration.cu:32-64)
                float tn;
                if (project_sample(vd.cam, vd.depth, acc, &tn) >= 0) {
                    blend_sample(gload(vd.tsdf + vidx), gload(vd.wt + vidx), tn, vd.cam.maxw, &tnew, &wnew);
                    emit = true;
                }
            }
        }
        // next brick's node record for this thread's list slot, loaded once
        // this brick's knn state is no longer live
        VF_FETCH_NEXT_RECORD();
        if (REDO) {
            // k_pass_b's step for this voxel: write it unless its target
            // was reached twice
            bool coll = false, upd = false;
            if (emit) {
                const uint32_t word = bitmap[(uint64_t)(ucode >> 9) * 32u + 16u + ((ucode & 511u) >> 5)];
                coll = ((word >> (ucode & 31u)) & 1u) != 0u;
                upd = !coll;
                if (upd) {
                    vd.tsdf[vidx] = tnew;
                    vd.wt[vidx] = wnew;
                }
            }
            const unsigned long long mc = __ballot(coll);
            const unsigned long long mu = __ballot(upd);
            if (lane == 0) {
                if (mc) atomicAdd(&s_stat[0], (unsigned long long)__popcll(mc));
                if (mu) atomicAdd(&s_stat[1], (unsigned long long)__popcll(mu));
            }
        } else {
        const unsigned long long msup = __ballot(n > 0);
        const unsigned long long minb = __ballot(inb);
        const unsigned long long memit = __ballot(emit);
        if (lane == 0) {
            s_wave_emit[wave] = (uint32_t)__popcll(memit);
            if (msup) atomicAdd(&s_stat[0], (unsigned long long)__popcll(msup));
            if (minb) atomicAdd(&s_stat[1], (unsigned long long)__popcll(minb));
            if (memit) atomicAdd(&s_stat[2], (unsigned long long)__popcll(memit));
        }
        __syncthreads();
        // The brick's pending records (in lane order), all in one place: the
        // workgroup's arena if they fit, else the shared overflow arena if
        // they fit there, else none, and the brick goes to the redo list
        // (its reserved overflow slots get null records).
        {
            uint32_t before = 0, total = 0;
#pragma unroll
            for (int wv = 0; wv < kPassAThreads / 64; ++wv) {
                const uint32_t c = s_wave_emit[wv];
                before += wv < wave ? c : 0u;
                total += c;
            }
            bool to_redo = false;
            if (total) {  // uniform, as are fill and the branches below
                uint4* dst;
                uint64_t room;  // slots of dst that may be written
                if ((uint64_t)fill + total <= arena.cap) {
                    dst = arena.records + (uint64_t)blockIdx.x * arena.cap + fill;
                    room = total;
                    fill += total;
                } else {
                    if (tid == 0) s_ovf = atomicAdd(&ctrl->overflow_records, total);
                    __syncthreads();
                    const uint32_t o = s_ovf;
                    dst = arena.overflow + o;
                    room = o < arena.overflow_cap ? arena.overflow_cap - o : 0u;
                    to_redo = room < total;
                    if (to_redo && tid == 0) {
                        redo[atomicAdd(&ctrl->redo_count, 1u)] = item;
                        s_stat[4] += 1ull;
                    }
                }
                if (emit) {
                    const uint32_t rank = before + (uint32_t)__popcll(memit & ((1ull << lane) - 1ull));
                    const uint4 rec = to_redo ? make_uint4(kNone, 0u, 0u, 0u)
                                              : make_uint4(vd.voxel_base + (uint32_t)vidx, ucode,
                                                           __float_as_uint(tnew), __float_as_uint(wnew));
                    if (rank < room) dst[rank] = rec;
                }
            }
            if (tid == 0 && !to_redo) brick_count[gb] = 0u;
        }
        // merge the window into the global bitmap, one word at a time
        for (int w = tid; w < kWinTiles * 16; w += kPassAThreads) {
            const int lt = w >> 4, wi = w & 15;
            uint32_t* blk = s_bm + lt * 32;
            const uint32_t l1 = blk[wi];
            if (l1) {
                const uint32_t l2 = blk[16 + wi];
                const int32_t tx = wx0 + (lt % kWX), ty = wy0 + ((lt / kWX) % kWY), tz = wz0 + lt / (kWX * kWY);
                // hits only land in tiles inside the volume
                const uint32_t tile =
                    vd.tile_base + ((uint32_t)tz * (uint32_t)vd.tdim[1] + (uint32_t)ty) * (uint32_t)vd.tdim[0] +
                    (uint32_t)tx;
                uint32_t* g = bitmap + (uint64_t)tile * 32u;
                const uint32_t old = atomicOr(g + wi, l1);
                const uint32_t two = l2 | (old & l1);
                if (two) atomicOr(g + 16 + wi, two);
                tile_flag[tile] = 1u;
                blk[wi] = 0u;
                blk[16 + wi] = 0u;
            }
        }
        }  // !REDO
        }  // !KNN_ONLY
#undef VF_FETCH_NEXT_RECORD
        // every thread read s_next before this brick's first barrier
        if (tid == 0) s_next = a_nn;
        __syncthreads();
        a = a_n;
        a_n = s_next;
        item = item_n;
        cnt = cnt_n;
        pre_id = id_n;
        pm = pm_n;
        plo = plo_n;
        pt0 = pt0_n;
        pt1 = pt1_n;
        pt2 = pt2_n;
        pw = pw_n;
    }
    __syncthreads();
    if (REDO && tid == 0) {
        stat_add(stat, totals, kStCollided, s_stat[0]);
        stat_add(stat, totals, kStUpdated, s_stat[1]);
    } else if (!KNN_ONLY && tid == 0) {
        arena.counts[blockIdx.x] = fill;
        stat_add(stat, totals, kStSupported, s_stat[0]);
        stat_add(stat, totals, kStInBounds, s_stat[1]);
        stat_add(stat, totals, kStPending, s_stat[2]);
        stat_add(stat, totals, kStSpilled, s_stat[3]);
        stat_add(stat, totals, kStRedo, s_stat[4]);
        if (blockIdx.x == 0) {
            stat[kStActive * kStatSlots] = nactive;
            atomicAdd(totals + kStActive * kStatSlots, (unsigned long long)nactive);
        }
    }
}

struct VolPtr {
    float* tsdf;
    float* wt;
    uint32_t voxel_base;
    uint32_t node_base;
};

// The collision word of a pending record's target: bit (target & 31) is set
// when the target was reached twice.
__device__ __forceinline__ uint32_t record_word(const uint4 rec, const uint32_t* __restrict__ bitmap) {
    const uint32_t tile = rec.y >> 9;
    const uint32_t bit = rec.y & 511u;
    return bitmap[(uint64_t)tile * 32u + 16u + (bit >> 5)];
}

// One pending record: write it unless its target was reached twice.
__device__ __forceinline__ bool apply_record(const uint4 rec, uint32_t word, const VolPtr* s_vp, int nvol) {
    if ((word >> (rec.y & 31u)) & 1u) return false;
    // volume of the voxel: last volume whose first voxel is <= rec.x
    int lo = 0, hi = nvol - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_vp[mid].voxel_base <= rec.x)
            lo = mid;
        else
            hi = mid - 1;
    }
    const VolPtr vp = s_vp[lo];
    const uint32_t v = rec.x - vp.voxel_base;
    vp.tsdf[v] = __uint_as_float(rec.z);
    vp.wt[v] = __uint_as_float(rec.w);
    return true;
}

__global__ __launch_bounds__(256) void k_pass_b(const VolPtr* __restrict__ vptr, int nvol, uint32_t total_nodes,
                                                Arena arena, Ctrl* __restrict__ ctrl,
                                                const uint32_t* __restrict__ bitmap,
                                                const VolDesc* __restrict__ vols, const NodeRec* __restrict__ nrec,
                                                unsigned long long* __restrict__ stat,
                                                unsigned long long* __restrict__ totals) {
    __shared__ VolPtr s_vp[kMaxBatch];
    __shared__ unsigned long long s_cnt[2];
    const int tid = threadIdx.x;
    for (int v = tid; v < nvol; v += blockDim.x) s_vp[v] = vptr[v];
    if (tid < 2) s_cnt[tid] = 0ull;
    __syncthreads();
    unsigned long long collided = 0, updated = 0;
    // this workgroup's share of the per-workgroup arenas, kUnrollB records per
    // thread in flight (their loads issued before any is applied)
    constexpr uint32_t kUnrollB = 8;
    for (uint32_t a = blockIdx.x; a < arena.n_arenas; a += gridDim.x) {
        const uint32_t n = arena.counts[a];
        const uint4* recs = arena.records + (uint64_t)a * arena.cap;
        for (uint32_t r0 = 0; r0 < n; r0 += kUnrollB * blockDim.x) {
            uint4 rec[kUnrollB];
            uint32_t word[kUnrollB];
#pragma unroll
            for (uint32_t u = 0; u < kUnrollB; ++u) {
                const uint32_t r = r0 + u * blockDim.x + tid;
                rec[u] = r < n ? recs[r] : make_uint4(0u, 0u, 0u, 0u);
            }
#pragma unroll
            for (uint32_t u = 0; u < kUnrollB; ++u)
                word[u] = r0 + u * blockDim.x + tid < n ? record_word(rec[u], bitmap) : 0u;
#pragma unroll
            for (uint32_t u = 0; u < kUnrollB; ++u) {
                if (r0 + u * blockDim.x + tid < n) {
                    if (apply_record(rec[u], word[u], s_vp, nvol))
                        ++updated;
                    else
                        ++collided;
                }
            }
        }
    }
    // the shared overflow arena (null records: slots of bricks that went to
    // the redo pass)
    uint32_t no = ctrl->overflow_records;
    if (no > arena.overflow_cap) no = arena.overflow_cap;
    for (uint32_t r = blockIdx.x * blockDim.x + tid; r < no; r += gridDim.x * blockDim.x) {
        const uint4 rec = arena.overflow[r];
        if (rec.x == kNone) continue;
        if (apply_record(rec, record_word(rec, bitmap), s_vp, nvol))
            ++updated;
        else
            ++collided;
    }
    // node positions as the reference's two translations leave them,
    // fl(fl(p - c) + c) (src/VolumeFusion.cu:708, 721): one pass over the
    // batch's nodes, each thread finding its node's volume in LDS (a loop
    // over the volumes would chain one descriptor load per volume)
    const uint32_t gid = blockIdx.x * blockDim.x + tid;
    const uint32_t nthreads = gridDim.x * blockDim.x;
    for (uint32_t g = gid; g < total_nodes; g += nthreads) {
        int lo = 0, hi = nvol - 1;  // last volume whose first node is <= g
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_vp[mid].node_base <= g)
                lo = mid;
            else
                hi = mid - 1;
        }
        const VolDesc& vd = vols[lo];
        const uint32_t j = g - vd.node_base;
        if (!vd.nodes_out || j >= vd.G) continue;
        const float4 m = nrec[g].m;
        vd.nodes_out[3ull * j + 0] = m.x + vd.c[0];
        vd.nodes_out[3ull * j + 1] = m.y + vd.c[1];
        vd.nodes_out[3ull * j + 2] = m.z + vd.c[2];
    }
    if (collided) atomicAdd(&s_cnt[0], collided);
    if (updated) atomicAdd(&s_cnt[1], updated);
    __syncthreads();
    if (tid == 0) {
        stat_add(stat, totals, kStCollided, s_cnt[0]);
        stat_add(stat, totals, kStUpdated, s_cnt[1]);
        // no kernel of this frame reads the counter any more
        if (blockIdx.x == 0) ctrl->frame = ctrl->frame + 1u;
    }
}

}  // namespace

// Device buffers of the brick machinery, grown on demand.
constexpr int kDescSlots = 4;  // resident descriptor sets per workspace

struct Buffers {
    Ctrl* ctrl = nullptr;
    NodeRec* nrec = nullptr;
    uint32_t* brick_count = nullptr;
    uint32_t* brick_off = nullptr;    // per brick: first entry of its node list in brick_nodes
    uint32_t* brick_nodes = nullptr;  // node-list pool: item_cap entries
    uint2* items = nullptr;           // per (node, brick) item: its list entry (brick, node << 7 | slot)
    uint32_t* active = nullptr;  // 2 x brick_cap (frame parity); cand doubles as the redo list
    uint2* active_oc = nullptr;  // 2 x brick_cap: each active brick's (first node-list entry, list count)
    uint32_t* bitmap = nullptr;
    uint32_t* tile_flag = nullptr;
    uint4* records = nullptr;     // arena_slots x arena_cap, then the overflow arena
    uint32_t* arena_counts = nullptr;
    unsigned long long* stat = nullptr;  // kStCount x kStatSlots, this frame
    unsigned long long* totals = nullptr;
    VolDesc* vols = nullptr;
    VolPtr* vptr = nullptr;
    // brick screen
    uint32_t* brick_disp = nullptr;  // per brick: displacement bound (float bits), max over its (node, brick) pairs
    uint32_t* cand = nullptr;        // bricks with a node list, before the screen
    uint4* cand_info = nullptr;      // the screen's verdict per candidate
    uint32_t* tile_mark = nullptr;   // per 8^3 tile: tag of the last frame that may have a pending target there
    uint32_t* vol_mark = nullptr;    // per volume: tag of the last frame that keeps every brick
    float* dmax = nullptr;           // depth-tile maxima
    uint64_t node_cap = 0, brick_cap = 0, tile_cap = 0, record_cap = 0, vol_cap = 0, arena_slots = 0, pyr_cap = 0,
             item_cap = 0;
    uint64_t bytes = 0;
};

}  // namespace vf

struct vf_workspace {
    int device = 0;
    vf::Buffers buf;
    uint64_t frame = 0;  // frame parity selects the active-brick list
    uint64_t dirty_tiles = 0;  // collision tiles [0, dirty_tiles) may hold flags from the last frame
    std::vector<vf::VolDesc> vols_h;
    std::vector<vf::VolPtr> vptr_h;
    // Device copies of recent descriptor sets (buf.vols / buf.vptr hold
    // kDescSlots sets of vol_cap entries): a frame whose descriptors equal a
    // resident set launches on it without an upload.  desc_n[s] = volumes in
    // set s (0: empty), desc_use[s] = last frame that used it.
    std::vector<vf::VolDesc> desc_vols[vf::kDescSlots];
    std::vector<vf::VolPtr> desc_vptr[vf::kDescSlots];
    int desc_n[vf::kDescSlots] = {};
    uint64_t desc_use[vf::kDescSlots] = {};
    uint64_t desc_clock = 0;
    // optional per-kernel timing: 6 events per frame, carried by the kernel
    // launches themselves (start of k_prep_bin, end of k_compact, start and
    // end of pass A, start and end of pass B), so no marker is queued
    // between kernels; timing 1: all three, 2: pass A only
    int timing = 0;
    std::vector<hipEvent_t> events;
    int timed_frames = 0;
    vf_tuning tuning = {0, 0, 0.0f, 0};
    // pending-record budget in records per voxel (0: the worst case); with a
    // budget the buffers grow to exactly what a frame needs
    float record_ratio = 0.0f;
    void* scratch = nullptr;  // general scratch for other entry points (point path)
    uint64_t scratch_bytes = 0;
};

namespace vf {
namespace {

constexpr int kMaxTimed = 4096;
constexpr int kTimedEvents = 6;  // per timed frame

void free_buffers(Buffers& b) {
    void* ptrs[] = {b.ctrl,      b.nrec,         b.brick_count,  b.brick_nodes, b.active,    b.bitmap,
                    b.tile_flag, b.records,      b.arena_counts, b.stat,        b.totals,    b.vols,
                    b.vptr,      b.brick_disp,   b.cand,         b.cand_info,   b.tile_mark, b.vol_mark,
                    b.dmax,      b.brick_off,    b.items,        b.active_oc};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    b = Buffers();
}

template <typename T>
int dev_alloc(T** p, uint64_t count, uint64_t* bytes) {
    const uint64_t nb = count * sizeof(T);
    if (nb == 0) {
        *p = nullptr;
        return VF_OK;
    }
    hipError_t e = hipMalloc((void**)p, nb);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(e == hipErrorOutOfMemory ? VF_ERR_OUT_OF_MEMORY : VF_ERR_HIP, "workspace allocation of %llu B: %s",
                    (unsigned long long)nb, hipGetErrorString(e));
    }
    *bytes += nb;
    return VF_OK;
}

struct Need {
    uint64_t nodes = 0, bricks = 0, tiles = 0, records = 0, nvol = 0, arena_slots = 0, pyr = 0, items = 0;
};

// (Re)allocates to at least the requested capacities; new buffers are zeroed.
int ensure(vf_workspace* ws, const Need& need, hipStream_t stream) {
    Buffers& b = ws->buf;
    if (b.ctrl && need.nodes <= b.node_cap && need.bricks <= b.brick_cap && need.tiles <= b.tile_cap &&
        need.records <= b.record_cap && need.nvol <= b.vol_cap && need.arena_slots <= b.arena_slots &&
        need.pyr <= b.pyr_cap && need.items <= b.item_cap)
        return VF_OK;
    if (b.ctrl) VF_HIP(hipStreamSynchronize(stream));
    // a capacity that has to grow gets half again as much headroom, so that
    // batches of varying size settle after a few frames; a workspace with a
    // record budget grows to exactly what is asked (its owner bounds it)
    const bool exact = ws->record_ratio > 0.0f;
    const auto grow = [exact](uint64_t want, uint64_t have) {
        if (exact) return std::max<uint64_t>(std::max(want, have), 1);
        return std::max<uint64_t>(want <= have ? have : std::max(want, have + have / 2), 1);
    };
    Buffers nbuf;
    nbuf.node_cap = grow(need.nodes, b.node_cap);
    nbuf.brick_cap = grow(need.bricks, b.brick_cap);
    nbuf.tile_cap = grow(need.tiles, b.tile_cap);
    nbuf.record_cap = grow(need.records, b.record_cap);
    nbuf.vol_cap = grow(need.nvol, b.vol_cap);
    nbuf.arena_slots = grow(need.arena_slots, b.arena_slots);
    nbuf.pyr_cap = grow(need.pyr, b.pyr_cap);
    nbuf.item_cap = grow(need.items, b.item_cap);
    // the running totals survive the reallocation
    unsigned long long* totals = b.totals;
    b.totals = nullptr;
    free_buffers(b);
    int st = VF_OK;
    uint64_t bytes = totals ? (uint64_t)kStCount * kStatSlots * 8 : 0;
    nbuf.totals = totals;
    if (!st) st = dev_alloc(&nbuf.ctrl, 1, &bytes);
    if (!st) st = dev_alloc(&nbuf.nrec, nbuf.node_cap, &bytes);
    if (!st) st = dev_alloc(&nbuf.brick_count, nbuf.brick_cap, &bytes);
    if (!st) st = dev_alloc(&nbuf.brick_off, nbuf.brick_cap, &bytes);
    if (!st) st = dev_alloc(&nbuf.brick_nodes, nbuf.item_cap, &bytes);
    if (!st) st = dev_alloc(&nbuf.items, nbuf.item_cap, &bytes);
    if (!st) st = dev_alloc(&nbuf.active, nbuf.brick_cap * 2, &bytes);
    if (!st) st = dev_alloc(&nbuf.active_oc, nbuf.brick_cap * 2, &bytes);
    if (!st) st = dev_alloc(&nbuf.bitmap, nbuf.tile_cap * 32, &bytes);
    if (!st) st = dev_alloc(&nbuf.tile_flag, nbuf.tile_cap, &bytes);
    if (!st) st = dev_alloc(&nbuf.records, nbuf.record_cap, &bytes);
    if (!st) st = dev_alloc(&nbuf.arena_counts, nbuf.arena_slots, &bytes);
    if (!st) st = dev_alloc(&nbuf.stat, (uint64_t)kStCount * kStatSlots, &bytes);
    if (!st && !nbuf.totals) {
        st = dev_alloc(&nbuf.totals, (uint64_t)kStCount * kStatSlots, &bytes);
        if (!st) VF_HIP(hipMemsetAsync(nbuf.totals, 0, (uint64_t)kStCount * kStatSlots * 8, stream));
    }
    if (!st) st = dev_alloc(&nbuf.vols, nbuf.vol_cap * kDescSlots, &bytes);
    if (!st) st = dev_alloc(&nbuf.vptr, nbuf.vol_cap * kDescSlots, &bytes);
    if (!st) st = dev_alloc(&nbuf.brick_disp, nbuf.brick_cap, &bytes);
    if (!st) st = dev_alloc(&nbuf.cand, nbuf.brick_cap, &bytes);
    if (!st) st = dev_alloc(&nbuf.cand_info, nbuf.brick_cap, &bytes);
    if (!st) st = dev_alloc(&nbuf.tile_mark, nbuf.tile_cap, &bytes);
    if (!st) st = dev_alloc(&nbuf.vol_mark, nbuf.vol_cap, &bytes);
    if (!st) st = dev_alloc(&nbuf.dmax, nbuf.pyr_cap, &bytes);
    if (st) {
        free_buffers(nbuf);
        return st;
    }
    nbuf.bytes = bytes;
    // Zero the state the kernels rely on: control block, brick counts, tiles,
    // statistics.
    VF_HIP(hipMemsetAsync(nbuf.ctrl, 0, sizeof(Ctrl), stream));
    VF_HIP(hipMemsetAsync(nbuf.brick_count, 0, nbuf.brick_cap * sizeof(uint32_t), stream));
    VF_HIP(hipMemsetAsync(nbuf.bitmap, 0, nbuf.tile_cap * 32 * sizeof(uint32_t), stream));
    VF_HIP(hipMemsetAsync(nbuf.tile_flag, 0, nbuf.tile_cap * sizeof(uint32_t), stream));
    VF_HIP(hipMemsetAsync(nbuf.arena_counts, 0, nbuf.arena_slots * sizeof(uint32_t), stream));
    VF_HIP(hipMemsetAsync(nbuf.stat, 0, (uint64_t)kStCount * kStatSlots * 8, stream));
    VF_HIP(hipMemsetAsync(nbuf.brick_disp, 0, nbuf.brick_cap * sizeof(uint32_t), stream));
    VF_HIP(hipMemsetAsync(nbuf.tile_mark, 0, nbuf.tile_cap * sizeof(uint32_t), stream));
    VF_HIP(hipMemsetAsync(nbuf.vol_mark, 0, nbuf.vol_cap * sizeof(uint32_t), stream));
    VF_HIP(hipStreamSynchronize(stream));
    b = nbuf;
    ws->frame = 0;
    ws->dirty_tiles = 0;
    for (int s = 0; s < kDescSlots; ++s) ws->desc_n[s] = 0;
    return VF_OK;
}

inline int32_t ceil_div(int32_t a, int32_t b) { return (a + b - 1) / b; }

// Fills `vd` for one volume, with bricks of shape `shape`.
int make_desc(const vf_volume& vol, const vf_graph& gr, const vf_frame& fr, const vf_warp_params& p, bool knn_only,
              int shape, bool screen, VolDesc* vd, uint64_t* n_out) {
    uint64_t n;
    int st = check_volume(vol.dim, vol.voxel_size, &n);
    if (st) return st;
    if (!knn_only && (!vol.tsdf || !vol.weights))
        return fail(VF_ERR_INVALID_ARGUMENT, "volume tsdf/weights must not be NULL");
    if (gr.count > 0 && (!gr.points || !gr.radii || (!knn_only && !gr.transforms)))
        return fail(VF_ERR_INVALID_ARGUMENT, "graph buffers must not be NULL");
    if (gr.count > 0xffffffull)
        return fail(VF_ERR_INVALID_ARGUMENT, "graph of %llu nodes exceeds 2^24-1", (unsigned long long)gr.count);
    if (!knn_only) {
        if (fr.width < 0 || fr.height < 0) return fail(VF_ERR_INVALID_ARGUMENT, "frame dims must be non-negative");
        if ((uint64_t)fr.width * (uint64_t)fr.height > 0 && !fr.depth)
            return fail(VF_ERR_INVALID_ARGUMENT, "depth frame must not be NULL");
    }
    memset(vd, 0, sizeof(*vd));
    vd->g = make_geom(vol.dim, vol.voxel_size, vol.size);
    vd->cam = make_camera(fr.intrinsics, fr.width, fr.height, p.delta, p.max_weight);
    vd->c[0] = vol.center.x;
    vd->c[1] = vol.center.y;
    vd->c[2] = vol.center.z;
    const float I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    mat3_inverse_eigen(I3, vd->rinv);
    vd->depth = fr.depth;
    vd->tsdf = vol.tsdf;
    vd->wt = vol.weights;
    vd->nodes = (const float*)gr.points;
    vd->nodes_out = (float*)gr.points_out;
    vd->radii = gr.radii;
    vd->T = (const float*)gr.transforms;
    vd->cube[0] = p.voxels_per_node.x;
    vd->cube[1] = p.voxels_per_node.y;
    vd->cube[2] = p.voxels_per_node.z;
    if (p.blend != VF_BLEND_LBS && p.blend != VF_BLEND_DQB)
        return fail(VF_ERR_INVALID_ARGUMENT, "blend mode %d: supported VF_BLEND_LBS (0), VF_BLEND_DQB (1)", p.blend);
    vd->blend = knn_only ? VF_BLEND_LBS : p.blend;
    const int32_t* bs = kShapeDims[shape];
    const int32_t dims[3] = {vol.dim.x, vol.dim.y, vol.dim.z};
    for (int a = 0; a < 3; ++a) {
        vd->bdim[a] = ceil_div(dims[a], bs[a]);
        vd->tdim[a] = ceil_div(dims[a], 8);
        const int32_t c = vd->cube[a] < 1 ? 1 : vd->cube[a];
        vd->maxb[a] = ceil_div(c - 1, bs[a]) + 1;
        if (vd->maxb[a] > vd->bdim[a]) vd->maxb[a] = vd->bdim[a];
    }
    {
        // magic divisors of brick indices (b < 2^24: the volume holds at most 2^24 bricks)
        const uint64_t ds[2] = {(uint64_t)std::max(vd->bdim[0], 1),
                                (uint64_t)std::max(vd->bdim[0], 1) * (uint64_t)std::max(vd->bdim[1], 1)};
        for (int i = 0; i < 2; ++i) {
            const uint64_t d = std::min<uint64_t>(ds[i], 1ull << 24);
            uint32_t l = 0;
            while ((1ull << l) < d) ++l;
            vd->bsh[i] = 24 + l;
            vd->bmag[i] = (uint32_t)((1ull << (24 + l)) / d + 1);
        }
    }
    // the brick screen: LBS volumes (the blend's displacement bound needs a
    // convex combination), tile coordinates that fit 16 bits
    vd->screen = (screen && !knn_only && vd->blend == VF_BLEND_LBS && vd->tdim[0] <= 0xffff &&
                  vd->tdim[1] <= 0xffff && vd->tdim[2] <= 0xffff)
                     ? 1
                     : 0;
    if (vd->screen) {
        vd->pw0 = ceil_div(fr.width, kTile0);
        vd->ph0 = ceil_div(fr.height, kTile0);
        vd->pw1 = ceil_div(fr.width, kTile1);
        vd->ph1 = ceil_div(fr.height, kTile1);
    }
    const bool no_cube = vd->cube[0] < 1 || vd->cube[1] < 1 || vd->cube[2] < 1;
    vd->G = (uint32_t)gr.count;
    const uint64_t k1 = no_cube ? 0 : (uint64_t)vd->G * (uint64_t)(vd->maxb[0] * vd->maxb[1] * vd->maxb[2]);
    if (k1 > 0xffffffffull) return fail(VF_ERR_CAPACITY, "node x brick work of %llu exceeds 2^32", (unsigned long long)k1);
    vd->k1_count = (uint32_t)k1;
    *n_out = n;
    return VF_OK;
}

// Everything a frame's launches need besides the device buffers.
struct FrameSetup {
    int nvol = 0;
    int k = 0;
    int shape = 0;
    bool knn_only = false;
    bool screen = false;        // some volume is screened: k_compact feeds k_select
    uint64_t blocks_pair = 0;   // k_prep_bin workgroups over (node, brick) pairs
    uint64_t blocks_pyr = 0;    // k_prep_bin workgroups over depth blocks
    uint32_t k1 = 0;
    uint64_t blocks_prep = 0, blocks_a = 0, blocks_b = 0, blocks_redo = 0, arena_cap = 0, overflow_cap = 0, bricks = 0;
    bool redo = false;         // launch the redo pass (a record budget: some bricks may not fit)
    uint64_t tiles = 0;        // collision tiles of this batch
    uint64_t nodes = 0;        // motion-graph nodes of this batch
    uint64_t clear_tiles = 0;  // tiles whose flags prep clears: every tile a previous frame or a replay may flag
    const VolDesc* vols = nullptr;  // the frame's device descriptor set
    const VolPtr* vptr = nullptr;
    int timing = 0;                 // events on all three kernel groups (1) or on pass A only (2)
};

// A launch that carries timing events (recorded at the kernel's start and
// end, no separate marker in the queue) when either is given.
template <typename F, typename... Args>
void launch(F kernel, dim3 grid, dim3 block, hipStream_t stream, hipEvent_t start, hipEvent_t stop, Args... args) {
    if (start || stop)
        hipExtLaunchKernelGGL(kernel, grid, block, 0, stream, start, stop, 0, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, 0, stream, args...);
}

template <class S>
const void* pass_a_kernel(int k, bool knn_only) {
#define VF_PASS_A_FN(KK) \
    return knn_only ? (const void*)k_pass_a<KK, true, S> : (const void*)k_pass_a<KK, false, S>
    switch (k) {
        case 1: VF_PASS_A_FN(1);
        case 2: VF_PASS_A_FN(2);
        case 3: VF_PASS_A_FN(3);
        case 4: VF_PASS_A_FN(4);
        case 5: VF_PASS_A_FN(5);
        case 6: VF_PASS_A_FN(6);
        case 7: VF_PASS_A_FN(7);
        default: VF_PASS_A_FN(8);
    }
#undef VF_PASS_A_FN
}

// Resident pass-A workgroups per CU for (k, shape).
int pass_a_occupancy(int k, bool knn_only, int shape, int* per_cu) {
    const void* fn = nullptr;
    int threads = 0;
    switch (shape) {
        case 0: fn = pass_a_kernel<Shape0>(k, knn_only); threads = Shape0::kThreads; break;
        case 1: fn = pass_a_kernel<Shape1>(k, knn_only); threads = Shape1::kThreads; break;
        case 2: fn = pass_a_kernel<Shape2>(k, knn_only); threads = Shape2::kThreads; break;
        case 3: fn = pass_a_kernel<Shape3>(k, knn_only); threads = Shape3::kThreads; break;
        default: fn = pass_a_kernel<Shape4>(k, knn_only); threads = Shape4::kThreads; break;
    }
    VF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, fn, threads, 0));
    if (*per_cu < 1) *per_cu = 1;
    return VF_OK;
}

int check_tuning(const vf_tuning& t) {
    if (t.brick_shape < 0 || t.brick_shape >= kNumShapes)
        return fail(VF_ERR_INVALID_ARGUMENT, "brick shape %d: supported 0..%d", t.brick_shape, kNumShapes - 1);
    if (t.wg_per_cu < 0 || t.wg_per_cu > 64)
        return fail(VF_ERR_INVALID_ARGUMENT, "workgroups per CU %d: supported 0 (occupancy limit) .. 64", t.wg_per_cu);
    if (t.flags & ~VF_TUNE_NO_SCREEN) return fail(VF_ERR_INVALID_ARGUMENT, "tuning flags 0x%x: unknown bits", t.flags);
    return VF_OK;
}

}  // namespace

// The groups follow setup_frame and ensure: Need.nodes (NodeRec), Need.bricks
// (brick_count, brick_off, active x 2, active_oc x 2, brick_disp, cand,
// cand_info) with
// Need.arena_slots (at most one pass-A workgroup per brick), Need.tiles
// (bitmap, tile_flag, tile_mark), Need.records (the budget, at least one),
// Need.items (items, brick_nodes), Need.nvol (descriptor sets, vol_mark) and
// Need.pyr (the screen's depth tiles, whether or not the volume is screened).
void volume_ws_bytes(vf_int3 dim, uint64_t nodes, vf_int3 cube, int width, int height, int brick_shape,
                     float records_per_voxel, uint64_t out[kWsGroups]) {
    const int32_t* bs = kShapeDims[brick_shape >= 0 && brick_shape < kNumShapes ? brick_shape : 0];
    const int32_t d[3] = {std::max(dim.x, 0), std::max(dim.y, 0), std::max(dim.z, 0)};
    const int32_t c[3] = {cube.x, cube.y, cube.z};
    uint64_t n = 1, bricks = 1, tiles = 1, maxb = 1;
    for (int a = 0; a < 3; ++a) {
        const int32_t bd = ceil_div(d[a], bs[a]);
        n *= (uint64_t)d[a];
        bricks *= (uint64_t)bd;
        tiles *= (uint64_t)ceil_div(d[a], 8);
        maxb *= (uint64_t)std::min(ceil_div(std::max(c[a], 1) - 1, bs[a]) + 1, bd);
    }
    const bool no_cube = c[0] < 1 || c[1] < 1 || c[2] < 1;
    const uint64_t k1 = no_cube ? 0 : nodes * maxb;
    const uint64_t w = (uint64_t)std::max(width, 0), h = (uint64_t)std::max(height, 0);
    const uint64_t pyr = (w + kTile0 - 1) / kTile0 * ((h + kTile0 - 1) / kTile0) +
                         (w + kTile1 - 1) / kTile1 * ((h + kTile1 - 1) / kTile1);
    const uint64_t fixed = sizeof(Ctrl) + 2ull * kStCount * kStatSlots * 8 + sizeof(uint4) + sizeof(uint2) + 4;
    out[0] = nodes * sizeof(NodeRec);
    out[1] = bricks * (4 + 4 + 8 + 16 + 4 + 4 + 16 + 4);
    out[2] = tiles * (32 * 4 + 4 + 4);
    out[3] = (uint64_t)std::ceil((double)n * (double)records_per_voxel) * sizeof(uint4);
    out[4] = k1 * (sizeof(uint2) + 4);
    out[5] = (sizeof(VolDesc) + sizeof(VolPtr)) * kDescSlots + 4 + fixed;
    out[6] = pyr * sizeof(float);
}

namespace {

// Builds the descriptors of a batch, sizes the launches, grows the buffers and
// uploads the descriptors on `stream`.
int setup_frame(vf_workspace* ws, int nvol, const vf_volume* vols, const vf_graph* graphs, const vf_frame* frames,
                const vf_warp_params* p, bool knn_only, const vf_tuning& tuning, FrameSetup* fs,
                hipStream_t stream) {
    int st = check_k(p->k);
    if (st) return st;
    st = check_tuning(tuning);
    if (st) return st;
    if (nvol < 1 || nvol > kMaxBatch)
        return fail(VF_ERR_INVALID_ARGUMENT, "batch of %d volumes: supported 1..%d", nvol, kMaxBatch);
    for (int v = 1; v < nvol; ++v)
        if (p[v].k != p[0].k) return fail(VF_ERR_INVALID_ARGUMENT, "batched volumes must share k");
    const int shape = knn_only ? 0 : tuning.brick_shape;
    ws->vols_h.resize(nvol);
    ws->vptr_h.resize(nvol);
    Need need;
    uint64_t voxels = 0, k1 = 0, rec_bound = 0, pyr_blk = 0;
    bool any_screen = false;
    for (int v = 0; v < nvol; ++v) {
        const vf_frame empty_frame = {nullptr, 0, 0, {{0}}};
        VolDesc& vd = ws->vols_h[v];
        memset(&vd, 0, sizeof(vd));  // padding too: descriptor sets are compared bytewise
        uint64_t n;
        st = make_desc(vols[v], graphs[v], frames ? frames[v] : empty_frame, p[v], knn_only, shape,
                       !(tuning.flags & VF_TUNE_NO_SCREEN), &vd, &n);
        if (st) return st;
        if (vd.screen) {
            any_screen = true;
            vd.pyr_blk = (uint32_t)pyr_blk;
            vd.pyr_base0 = (uint32_t)need.pyr;
            need.pyr += (uint64_t)vd.pw0 * vd.ph0;
            vd.pyr_base1 = (uint32_t)need.pyr;
            need.pyr += (uint64_t)vd.pw1 * vd.ph1;
            pyr_blk += (uint64_t)vd.pw1 * vd.ph1;
        } else {
            vd.pyr_blk = (uint32_t)pyr_blk;  // no blocks: keeps the block bases sorted
        }
        if (pyr_blk > 0x7fffffffull || need.pyr > 0xffffffffull)
            return fail(VF_ERR_CAPACITY, "depth frames of the batch too large for the brick screen");
        vd.node_base = (uint32_t)need.nodes;
        vd.brick_base = (uint32_t)need.bricks;
        vd.tile_base = (uint32_t)need.tiles;
        vd.voxel_base = (uint32_t)voxels;
        vd.k1_base = (uint32_t)k1;
        need.nodes += vd.G;
        const uint64_t nb = (uint64_t)vd.bdim[0] * vd.bdim[1] * vd.bdim[2];
        need.bricks += nb;
        need.tiles += (uint64_t)vd.tdim[0] * vd.tdim[1] * vd.tdim[2];
        voxels += n;
        k1 += vd.k1_count;
        // pending records <= supported voxels <= min(N, G * cube volume)
        const uint64_t cubevol = (uint64_t)std::max(vd.cube[0], 0) * std::max(vd.cube[1], 0) * std::max(vd.cube[2], 0);
        rec_bound += std::min<uint64_t>(n, (uint64_t)vd.G * cubevol);
        ws->vptr_h[v].tsdf = vd.tsdf;
        ws->vptr_h[v].wt = vd.wt;
        ws->vptr_h[v].voxel_base = vd.voxel_base;
        ws->vptr_h[v].node_base = vd.node_base;
        if (nb > 0xffffffull) return fail(VF_ERR_CAPACITY, "volume has more than 2^24 bricks");
        if (need.bricks > 0xffffffffull || need.tiles >= (1ull << 23) || k1 > 0xffffffffull ||
            voxels > 0xffffffffull)
            return fail(VF_ERR_CAPACITY,
                        "batch too large: %llu voxels, %llu tiles (limits 2^32 voxels and 2^23 8^3 tiles per call)",
                        (unsigned long long)voxels, (unsigned long long)need.tiles);
    }
    need.nvol = (uint64_t)nvol;

    int dev = 0, cus = 0;
    VF_HIP(hipGetDevice(&dev));
    VF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    int per_cu = 1;
    st = pass_a_occupancy(p->k, knn_only, shape, &per_cu);
    if (st) return st;
    const int wg = tuning.wg_per_cu > 0 ? std::min(tuning.wg_per_cu, per_cu) : per_cu;
    const uint64_t blocks_a = std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)cus * wg, need.bricks));
    const int threads_a = kShapeDims[shape][0] * kShapeDims[shape][1] * kShapeDims[shape][2];
    // records: one arena per pass-A workgroup, plus a shared overflow arena.
    // Without a budget the arenas take an even share of the bound and the
    // overflow arena the whole bound, so every brick's records fit.  With a
    // budget of R records the arenas hold R / 2 between them and the overflow
    // arena the rest; the bricks that fit neither go to the redo pass.
    uint64_t arena_cap = 0, overflow_cap = 0;
    bool redo = false;
    if (!knn_only) {
        if (ws->record_ratio > 0.0f) {
            const uint64_t budget =
                std::max<uint64_t>(1, (uint64_t)std::ceil((double)voxels * (double)ws->record_ratio));
            const uint64_t r = std::min(rec_bound, budget);
            arena_cap = r / (2 * blocks_a);
            overflow_cap = std::max<uint64_t>(r - blocks_a * arena_cap, 1);
            redo = true;
        } else {
            arena_cap = (rec_bound + blocks_a - 1) / blocks_a + (uint64_t)threads_a;
            overflow_cap = rec_bound;
        }
        if (arena_cap > 0xffffffffull || overflow_cap > 0xffffffffull)
            return fail(VF_ERR_CAPACITY, "record bound %llu exceeds 2^32", (unsigned long long)rec_bound);
        need.records = blocks_a * arena_cap + overflow_cap;
        need.arena_slots = blocks_a;
    } else {
        need.records = 1;
        need.arena_slots = 1;
    }
    need.items = std::max<uint64_t>(k1, 1);
    st = ensure(ws, need, stream);
    if (st) return st;
    Buffers& B = ws->buf;
    // the descriptor set: a resident copy when one matches, else upload into
    // the least recently used slot (stream-ordered after the frames that
    // read it)
    int slot = -1;
    for (int s = 0; s < kDescSlots && slot < 0; ++s)
        if (ws->desc_n[s] == nvol && !memcmp(ws->desc_vols[s].data(), ws->vols_h.data(), sizeof(VolDesc) * nvol) &&
            !memcmp(ws->desc_vptr[s].data(), ws->vptr_h.data(), sizeof(VolPtr) * nvol))
            slot = s;
    if (slot < 0) {
        slot = 0;
        for (int s = 1; s < kDescSlots; ++s)
            if (ws->desc_use[s] < ws->desc_use[slot]) slot = s;
        ws->desc_n[slot] = 0;
        VF_HIP(hipMemcpyAsync(B.vols + (uint64_t)slot * B.vol_cap, ws->vols_h.data(), sizeof(VolDesc) * nvol,
                              hipMemcpyHostToDevice, stream));
        VF_HIP(hipMemcpyAsync(B.vptr + (uint64_t)slot * B.vol_cap, ws->vptr_h.data(), sizeof(VolPtr) * nvol,
                              hipMemcpyHostToDevice, stream));
        ws->desc_vols[slot].assign(ws->vols_h.begin(), ws->vols_h.begin() + nvol);
        ws->desc_vptr[slot].assign(ws->vptr_h.begin(), ws->vptr_h.begin() + nvol);
        ws->desc_n[slot] = nvol;
    }
    ws->desc_use[slot] = ++ws->desc_clock;
    fs->vols = B.vols + (uint64_t)slot * B.vol_cap;
    fs->vptr = B.vptr + (uint64_t)slot * B.vol_cap;

    fs->nvol = nvol;
    fs->k = p->k;
    fs->shape = shape;
    fs->knn_only = knn_only;
    fs->k1 = (uint32_t)k1;
    fs->blocks_pair = std::max<uint64_t>((k1 + kPrepThreads - 1) / kPrepThreads, 256);
    fs->blocks_pyr = pyr_blk;
    fs->blocks_prep = fs->blocks_pair + fs->blocks_pyr;
    fs->screen = any_screen;
    fs->blocks_a = blocks_a;
    fs->bricks = need.bricks;
    fs->tiles = need.tiles;
    fs->nodes = need.nodes;
    fs->clear_tiles = std::max<uint64_t>(ws->dirty_tiles, need.tiles);
    fs->blocks_b = std::min<uint64_t>(blocks_a, (uint64_t)cus * 8);
    // the redo list is short or empty on most frames: one workgroup per CU
    // starts and exits quickly, and takes the list round-robin when it is long
    fs->blocks_redo = std::min<uint64_t>(blocks_a, (uint64_t)std::max(cus, 1));
    fs->arena_cap = arena_cap;
    fs->overflow_cap = overflow_cap;
    fs->redo = redo;
    return VF_OK;
}

template <class S>
int launch_kernels_shape(Buffers& B, const FrameSetup& fs, int parity_arg, KnnOut kout, hipEvent_t* ev,
                         hipStream_t stream) {
    hipEvent_t e[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    if (ev)
        for (int i = 0; i < 6; ++i) e[i] = (fs.timing == 2 && (i < 2 || i > 3)) ? nullptr : ev[i];
    const VolDesc* const vols = fs.vols;
    launch(k_prep_bin<S>, dim3((unsigned)fs.blocks_prep), dim3(kPrepThreads), stream, e[0], nullptr,
           vols, fs.nvol, fs.k1, (uint32_t)fs.blocks_pair, B.nrec, B.brick_count, B.items, B.brick_disp,
           B.dmax, B.ctrl, B.bitmap, B.tile_flag, (uint32_t)fs.clear_tiles, B.stat);
    VF_LAUNCH_CHECK("k_prep_bin");
    const dim3 grid_c((unsigned)((fs.bricks + kPrepThreads * kCompactPer - 1) / (kPrepThreads * kCompactPer)));
    if (fs.screen) {
        launch(k_compact<S, true>, grid_c, dim3(kPrepThreads), stream, nullptr, nullptr, vols, fs.nvol,
               (uint32_t)fs.bricks, (const uint32_t*)B.brick_count, B.brick_off, (const uint32_t*)B.brick_disp,
               (const float*)B.dmax, B.active, B.active_oc, B.cand, B.cand_info, B.tile_mark, B.vol_mark, B.ctrl,
               B.brick_cap, parity_arg);
        VF_LAUNCH_CHECK("k_compact");
        launch(k_select, grid_c, dim3(kPrepThreads), stream, nullptr, e[1], vols, (const uint32_t*)B.cand,
               (const uint4*)B.cand_info, B.brick_count, B.brick_disp, (const uint32_t*)B.tile_mark,
               (const uint32_t*)B.vol_mark, B.active, B.active_oc, B.ctrl, B.brick_cap, parity_arg,
               (const uint2*)B.items,
               fs.k1, (const uint32_t*)B.brick_off, B.brick_nodes, B.stat, B.totals);
        VF_LAUNCH_CHECK("k_select");
    } else {
        launch(k_compact<S, false>, grid_c, dim3(kPrepThreads), stream, nullptr, nullptr, vols, fs.nvol,
               (uint32_t)fs.bricks, (const uint32_t*)B.brick_count, B.brick_off, (const uint32_t*)B.brick_disp,
               (const float*)B.dmax, B.active, B.active_oc, B.cand, B.cand_info, B.tile_mark, B.vol_mark, B.ctrl,
               B.brick_cap, parity_arg);
        VF_LAUNCH_CHECK("k_compact");
        launch(k_fill_lists, dim3((unsigned)fs.blocks_pair), dim3(kPrepThreads), stream, nullptr, e[1],
               (const uint2*)B.items, fs.k1, (const uint32_t*)B.brick_off, B.brick_nodes);
        VF_LAUNCH_CHECK("k_fill_lists");
    }
    Arena arena;
    arena.records = B.records;
    arena.counts = B.arena_counts;
    arena.cap = (uint32_t)fs.arena_cap;
    arena.n_arenas = (uint32_t)fs.blocks_a;
    arena.overflow = B.records + fs.blocks_a * fs.arena_cap;
    arena.overflow_cap = (uint32_t)fs.overflow_cap;
#define VF_PASS_A_ARGS                                                                                    \
    vols, (const NodeRec*)B.nrec, B.brick_count, (const uint32_t*)B.brick_nodes, (const uint32_t*)B.brick_off, \
        (const uint32_t*)B.active, (const uint2*)B.active_oc, B.cand, B.ctrl, B.bitmap, B.tile_flag, B.brick_cap, \
        parity_arg, arena, B.stat,                                                                                  \
        B.totals, kout
#define VF_PASS_A(KK)                                                                                       \
    if (fs.knn_only) {                                                                                      \
        launch(k_pass_a<KK, true, S>, dim3((unsigned)fs.blocks_a), dim3(S::kThreads), stream, e[2], e[3],   \
               VF_PASS_A_ARGS);                                                                             \
    } else {                                                                                                \
        launch(k_pass_a<KK, false, S>, dim3((unsigned)fs.blocks_a), dim3(S::kThreads), stream, e[2], e[3],  \
               VF_PASS_A_ARGS);                                                                             \
        if (fs.redo) {                                                                                      \
            VF_LAUNCH_CHECK("k_pass_a");                                                                    \
            launch(k_pass_a<KK, false, S, true>, dim3((unsigned)fs.blocks_redo), dim3(S::kThreads), stream, \
                   nullptr, nullptr, VF_PASS_A_ARGS);                                                       \
        }                                                                                                   \
    }
    switch (fs.k) {
        case 1: VF_PASS_A(1); break;
        case 2: VF_PASS_A(2); break;
        case 3: VF_PASS_A(3); break;
        case 4: VF_PASS_A(4); break;
        case 5: VF_PASS_A(5); break;
        case 6: VF_PASS_A(6); break;
        case 7: VF_PASS_A(7); break;
        case 8: VF_PASS_A(8); break;
    }
#undef VF_PASS_A
#undef VF_PASS_A_ARGS
    VF_LAUNCH_CHECK(fs.redo ? "k_pass_a (redo)" : "k_pass_a");
    if (!fs.knn_only) {
        launch(k_pass_b, dim3((unsigned)fs.blocks_b), dim3(256), stream, e[4], e[5], fs.vptr, fs.nvol,
               (uint32_t)fs.nodes, arena, B.ctrl,
               (const uint32_t*)B.bitmap, vols, (const NodeRec*)B.nrec, B.stat, B.totals);
        VF_LAUNCH_CHECK("k_pass_b");
    }
    return VF_OK;
}

// The frame's kernels on `stream`: k_prep_bin, k_compact, k_pass_a,
// k_pass_b.  With parity_arg < 0 the kernels take the frame parity from the
// device counter (what a recorded graph needs); ev (6 events, or NULL) times
// the prep-and-list kernels, pass A and pass B.
int launch_kernels(vf_workspace* ws, const FrameSetup& fs, int parity_arg, KnnOut kout, hipEvent_t* ev,
                   hipStream_t stream) {
    Buffers& B = ws->buf;
    switch (fs.shape) {
        case 0: return launch_kernels_shape<Shape0>(B, fs, parity_arg, kout, ev, stream);
        case 1: return launch_kernels_shape<Shape1>(B, fs, parity_arg, kout, ev, stream);
        case 2: return launch_kernels_shape<Shape2>(B, fs, parity_arg, kout, ev, stream);
        case 3: return launch_kernels_shape<Shape3>(B, fs, parity_arg, kout, ev, stream);
        default: return launch_kernels_shape<Shape4>(B, fs, parity_arg, kout, ev, stream);
    }
}

// At least `frames` timed frames' events exist (created in doubling steps).
int grow_events(vf_workspace* ws, int frames) {
    const size_t want = (size_t)kTimedEvents * (size_t)frames;
    if (ws->events.size() >= want) return VF_OK;
    const size_t have = ws->events.size();
    const size_t n = std::min<size_t>(std::max(want, 2 * have), (size_t)kTimedEvents * kMaxTimed);
    ws->events.resize(n, nullptr);
    for (size_t i = have; i < n; ++i) VF_HIP(hipEventCreate(&ws->events[i]));
    return VF_OK;
}

int launch_frame(vf_workspace* ws, int nvol, const vf_volume* vols, const vf_graph* graphs, const vf_frame* frames,
                 const vf_warp_params* p, bool knn_only, KnnOut kout, hipStream_t stream) {
    FrameSetup fs;
    int st = setup_frame(ws, nvol, vols, graphs, frames, p, knn_only, ws->tuning, &fs, stream);
    if (st) return st;
    hipEvent_t* ev = nullptr;
    if (ws->timing && !knn_only) {
        if (ws->timed_frames >= kMaxTimed) return fail(VF_ERR_CAPACITY, "more than %d timed frames", kMaxTimed);
        st = grow_events(ws, ws->timed_frames + 1);
        if (st) return st;
        ev = &ws->events[kTimedEvents * ws->timed_frames];
        ws->timed_frames++;
    }
    fs.timing = ws->timing;
    st = launch_kernels(ws, fs, (int)(ws->frame & 1u), kout, ev, stream);
    if (st) return st;
    ws->frame++;
    ws->dirty_tiles = knn_only ? 0 : fs.tiles;
    return VF_OK;
}

int read_stats(vf_workspace* ws, bool totals, vf_frame_stats* out, hipStream_t stream) {
    memset(out, 0, sizeof(*out));
    if (!ws->buf.ctrl) return VF_OK;
    unsigned long long h[kStCount * kStatSlots];
    VF_HIP(hipMemcpyAsync(h, totals ? ws->buf.totals : ws->buf.stat, sizeof(h), hipMemcpyDeviceToHost, stream));
    VF_HIP(hipStreamSynchronize(stream));
    unsigned long long s[kStCount] = {0};
    for (int c = 0; c < kStCount; ++c)
        for (int k = 0; k < kStatSlots; ++k) s[c] += h[c * kStatSlots + k];
    out->supported = s[kStSupported];
    out->in_bounds = s[kStInBounds];
    out->pending = s[kStPending];
    out->collided = s[kStCollided];
    out->updated = s[kStUpdated];
    out->active_bricks = s[kStActive];
    out->overflow_bricks = s[kStSpilled];
    out->screened_bricks = s[kStScreened];
    out->redo_bricks = s[kStRedo];
    return VF_OK;
}

// ---- frame plans -----------------------------------------------------------

// Size class of a batch for the tuning cache: k, volume count, and the
// log2 buckets of its voxel and node totals.
struct TuneKey {
    int k, n, vox_log2, node_log2;
    bool operator<(const TuneKey& o) const {
        if (k != o.k) return k < o.k;
        if (n != o.n) return n < o.n;
        if (vox_log2 != o.vox_log2) return vox_log2 < o.vox_log2;
        return node_log2 < o.node_log2;
    }
};

int ilog2_u64(uint64_t v) {
    int l = 0;
    while (v > 1) {
        v >>= 1;
        ++l;
    }
    return l;
}

TuneKey tune_key(int n, const vf_volume* vols, const vf_graph* graphs, const vf_warp_params* p) {
    uint64_t vox = 0, nodes = 0;
    for (int v = 0; v < n; ++v) {
        vox += (uint64_t)std::max(vols[v].dim.x, 0) * std::max(vols[v].dim.y, 0) * std::max(vols[v].dim.z, 0);
        nodes += graphs[v].count;
    }
    return TuneKey{p[0].k, n, ilog2_u64(vox), ilog2_u64(nodes)};
}

std::mutex g_tune_mutex;
std::map<TuneKey, vf_tuning> g_tune_cache;

// Device memory and events released on scope exit (autotuner scratch).
struct ScratchGuard {
    std::vector<void*> ptrs;
    vf_workspace* ws = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~ScratchGuard() {
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
        if (ws) free_buffers(ws->buf);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

}  // namespace
}  // namespace vf

struct vf_plan {
    vf_workspace ws;
    int n = 0;
    int flags = 0;
    std::vector<vf_volume> vols;
    std::vector<vf_graph> graphs;
    std::vector<vf_frame> frames;
    std::vector<vf_warp_params> params;
    vf_tuning tuning = {0, 0, 0.0f, 0};
    vf::FrameSetup fs;
    hipStream_t capture_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace vf {
namespace {

void plan_drop_graph(vf_plan* pl) {
    if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
    if (pl->graph) (void)hipGraphDestroy(pl->graph);
    pl->exec = nullptr;
    pl->graph = nullptr;
}

// (Re)computes the plan's launch setup for its tuning and uploads its
// descriptors; waits for the upload.
int plan_setup(vf_plan* pl, hipStream_t stream) {
    plan_drop_graph(pl);
    int st = setup_frame(&pl->ws, pl->n, pl->vols.data(), pl->graphs.data(), pl->frames.data(), pl->params.data(),
                         false, pl->tuning, &pl->fs, stream);
    if (st) return st;
    VF_HIP(hipStreamSynchronize(stream));
    return VF_OK;
}

int plan_record_graph(vf_plan* pl) {
    if (!pl->capture_stream) VF_HIP(hipStreamCreateWithFlags(&pl->capture_stream, hipStreamNonBlocking));
    VF_HIP(hipStreamBeginCapture(pl->capture_stream, hipStreamCaptureModeThreadLocal));
    const int st = launch_kernels(&pl->ws, pl->fs, -1, KnnOut{nullptr, nullptr, nullptr}, nullptr, pl->capture_stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(pl->capture_stream, &g);
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return fail(VF_ERR_HIP, "graph capture of the frame failed: %s", hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(&pl->exec, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(g);
        pl->exec = nullptr;
        return fail(VF_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(ei));
    }
    pl->graph = g;
    return VF_OK;
}

int plan_autotune(vf_plan* pl, int frames, vf_tuning* chosen, hipStream_t stream) {
    frames = std::max(1, std::min(frames, 64));
    const int n = pl->n;
    ScratchGuard guard;
    vf_workspace tmp;
    VF_HIP(hipGetDevice(&tmp.device));
    guard.ws = &tmp;
    // scratch copies of the volumes; node positions are not written back
    std::vector<vf_volume> vols = pl->vols;
    std::vector<vf_graph> graphs = pl->graphs;
    for (int v = 0; v < n; ++v) {
        const uint64_t nv = (uint64_t)vols[v].dim.x * vols[v].dim.y * vols[v].dim.z;
        float* t = nullptr;
        float* w = nullptr;
        VF_HIP(hipMalloc((void**)&t, nv * sizeof(float)));
        guard.ptrs.push_back(t);
        VF_HIP(hipMalloc((void**)&w, nv * sizeof(float)));
        guard.ptrs.push_back(w);
        VF_HIP(hipMemcpyAsync(t, pl->vols[v].tsdf, nv * sizeof(float), hipMemcpyDeviceToDevice, stream));
        VF_HIP(hipMemcpyAsync(w, pl->vols[v].weights, nv * sizeof(float), hipMemcpyDeviceToDevice, stream));
        vols[v].tsdf = t;
        vols[v].weights = w;
        graphs[v].points_out = nullptr;
    }
    VF_HIP(hipEventCreate(&guard.ev[0]));
    VF_HIP(hipEventCreate(&guard.ev[1]));
    vf_tuning best = {-1, 0, 0.0f, pl->tuning.flags};
    for (int shape = 0; shape < kNumShapes; ++shape) {
        int per_cu = 1;
        int st = pass_a_occupancy(pl->params[0].k, false, shape, &per_cu);
        if (st) return st;
        for (int wg = 1; wg <= per_cu; ++wg) {
            const vf_tuning t = {shape, wg, 0.0f, pl->tuning.flags};
            FrameSetup fs;
            st = setup_frame(&tmp, n, vols.data(), graphs.data(), pl->frames.data(), pl->params.data(), false, t, &fs,
                             stream);
            if (st) return st;
            st = launch_kernels(&tmp, fs, -1, KnnOut{nullptr, nullptr, nullptr}, nullptr, stream);
            if (st) return st;
            VF_HIP(hipEventRecord(guard.ev[0], stream));
            for (int f = 0; f < frames; ++f) {
                st = launch_kernels(&tmp, fs, -1, KnnOut{nullptr, nullptr, nullptr}, nullptr, stream);
                if (st) return st;
            }
            VF_HIP(hipEventRecord(guard.ev[1], stream));
            VF_HIP(hipEventSynchronize(guard.ev[1]));
            float ms = 0.0f;
            VF_HIP(hipEventElapsedTime(&ms, guard.ev[0], guard.ev[1]));
            ms /= (float)frames;
            if (best.brick_shape < 0 || ms < best.ms_per_frame) best = vf_tuning{shape, wg, ms, t.flags};
        }
    }
    VF_HIP(hipStreamSynchronize(stream));
    pl->tuning = best;
    {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        g_tune_cache[tune_key(n, pl->vols.data(), pl->graphs.data(), pl->params.data())] = best;
    }
    const int st = plan_setup(pl, stream);
    if (st) return st;
    if (chosen) *chosen = best;
    return VF_OK;
}

// Workspace behind vf_tsdf_knn_graph, which has the reference's signature and
// so no workspace argument: one per device, grown on demand.
std::mutex g_knn_ws_mutex;
std::vector<vf_workspace*> g_knn_ws;

}  // namespace
}  // namespace vf

using namespace vf;

namespace vf {

namespace {
std::mutex g_default_ws_mutex;
std::vector<vf_workspace*> g_default_ws;
}  // namespace

// At least `bytes` of device scratch owned by the workspace; grows after the
// stream's earlier work is done.  Contents are not preserved.  A NULL
// workspace selects the current device's default one (for entry points with
// the reference's signatures; calls that share it must not overlap).
int workspace_scratch(vf_workspace* ws, uint64_t bytes, void** ptr, hipStream_t stream) {
    if (!ws) {
        int dev = 0;
        VF_HIP(hipGetDevice(&dev));
        std::lock_guard<std::mutex> lock(g_default_ws_mutex);
        if ((int)g_default_ws.size() <= dev) g_default_ws.resize(dev + 1, nullptr);
        if (!g_default_ws[dev]) {
            g_default_ws[dev] = new vf_workspace();
            g_default_ws[dev]->device = dev;
        }
        ws = g_default_ws[dev];
    }
    if (bytes > ws->scratch_bytes) {
        VF_HIP(hipStreamSynchronize(stream));
        if (ws->scratch) VF_HIP(hipFree(ws->scratch));
        ws->scratch = nullptr;
        ws->scratch_bytes = 0;
        const uint64_t nb = std::max<uint64_t>(bytes, 1u << 16);
        const hipError_t e = hipMalloc(&ws->scratch, nb);
        if (e != hipSuccess)
            return fail(e == hipErrorOutOfMemory ? VF_ERR_OUT_OF_MEMORY : VF_ERR_HIP, "workspace scratch of %llu B: %s",
                        (unsigned long long)nb, hipGetErrorString(e));
        ws->scratch_bytes = nb;
    }
    *ptr = ws->scratch;
    return VF_OK;
}

}  // namespace vf

extern "C" {

int vf_workspace_create(vf_workspace** out, int max_volumes, uint64_t max_total_voxels, uint64_t max_total_nodes,
                        vf_stream stream) {
    if (!out) return fail(VF_ERR_INVALID_ARGUMENT, "vf_workspace_create: out is NULL");
    *out = nullptr;
    vf_workspace* ws = new vf_workspace();
    VF_HIP(hipGetDevice(&ws->device));
    (void)max_volumes;
    (void)max_total_voxels;
    (void)max_total_nodes;
    (void)stream;
    *out = ws;
    return VF_OK;
}

int vf_workspace_destroy(vf_workspace* ws) {
    if (!ws) return VF_OK;
    if (ws->scratch) (void)hipFree(ws->scratch);
    for (auto& e : ws->events) (void)hipEventDestroy(e);
    free_buffers(ws->buf);
    delete ws;
    return VF_OK;
}

uint64_t vf_workspace_bytes(const vf_workspace* ws) { return ws ? ws->buf.bytes + ws->scratch_bytes : 0; }

int vf_warped_integrate(vf_workspace* ws, const vf_volume* volume, const vf_graph* graph, const vf_frame* frame,
                        const vf_warp_params* params, vf_stream stream) {
    if (!ws || !volume || !graph || !frame || !params)
        return fail(VF_ERR_INVALID_ARGUMENT, "vf_warped_integrate: NULL argument");
    return launch_frame(ws, 1, volume, graph, frame, params, false, KnnOut{nullptr, nullptr, nullptr},
                        (hipStream_t)stream);
}

int vf_warped_integrate_batch(vf_workspace* ws, int n, const vf_volume* volumes, const vf_graph* graphs,
                              const vf_frame* frames, const vf_warp_params* params, vf_stream stream) {
    if (!ws || !volumes || !graphs || !frames || !params)
        return fail(VF_ERR_INVALID_ARGUMENT, "vf_warped_integrate_batch: NULL argument");
    return launch_frame(ws, n, volumes, graphs, frames, params, false, KnnOut{nullptr, nullptr, nullptr},
                        (hipStream_t)stream);
}

int vf_workspace_last_stats(vf_workspace* ws, vf_frame_stats* stats, vf_stream stream) {
    if (!ws || !stats) return fail(VF_ERR_INVALID_ARGUMENT, "vf_workspace_last_stats: NULL argument");
    return read_stats(ws, false, stats, (hipStream_t)stream);
}

int vf_workspace_totals(vf_workspace* ws, vf_frame_stats* totals, int reset, vf_stream stream) {
    if (!ws || !totals) return fail(VF_ERR_INVALID_ARGUMENT, "vf_workspace_totals: NULL argument");
    int st = read_stats(ws, true, totals, (hipStream_t)stream);
    if (st) return st;
    if (reset && ws->buf.totals) {
        VF_HIP(hipMemsetAsync(ws->buf.totals, 0, (uint64_t)kStCount * kStatSlots * 8, (hipStream_t)stream));
        VF_HIP(hipStreamSynchronize((hipStream_t)stream));
    }
    return VF_OK;
}

int vf_workspace_set_timing(vf_workspace* ws, int enable) {
    if (!ws) return fail(VF_ERR_INVALID_ARGUMENT, "vf_workspace_set_timing: NULL workspace");
    if (enable < 0 || enable > 2) return fail(VF_ERR_INVALID_ARGUMENT, "timing mode %d: supported 0, 1, 2", enable);
    ws->timing = enable;
    ws->timed_frames = 0;
    // events for the first frames now, so that creating them does not fall
    // between the timed frames' launches
    return ws->timing ? grow_events(ws, 256) : VF_OK;
}

int vf_workspace_kernel_ms(vf_workspace* ws, float* ms3, int* frames) {
    if (!ws || !ms3 || !frames) return fail(VF_ERR_INVALID_ARGUMENT, "vf_workspace_kernel_ms: NULL argument");
    ms3[0] = ms3[1] = ms3[2] = 0.0f;
    *frames = ws->timed_frames;
    if (ws->timed_frames) VF_HIP(hipEventSynchronize(ws->events[kTimedEvents * ws->timed_frames - 1]));
    for (int f = 0; f < ws->timed_frames; ++f) {
        for (int k = 0; k < 3; ++k) {
            if (ws->timing == 2 && k != 1) continue;
            float t = 0.0f;
            const hipEvent_t* e = &ws->events[kTimedEvents * f + 2 * k];
            VF_HIP(hipEventElapsedTime(&t, e[0], e[1]));
            ms3[k] += t;
        }
    }
    ws->timed_frames = 0;
    return VF_OK;
}

int vf_workspace_set_tuning(vf_workspace* ws, const vf_tuning* tuning) {
    if (!ws || !tuning) return fail(VF_ERR_INVALID_ARGUMENT, "vf_workspace_set_tuning: NULL argument");
    int st = check_tuning(*tuning);
    if (st) return st;
    ws->tuning = *tuning;
    return VF_OK;
}

int vf_workspace_set_record_budget(vf_workspace* ws, float records_per_voxel) {
    if (!ws) return fail(VF_ERR_INVALID_ARGUMENT, "vf_workspace_set_record_budget: NULL workspace");
    if (!(records_per_voxel >= 0.0f && records_per_voxel <= 1.0f))
        return fail(VF_ERR_INVALID_ARGUMENT, "record budget %g per voxel: supported 0 (the worst case) .. 1",
                    (double)records_per_voxel);
    ws->record_ratio = records_per_voxel;
    return VF_OK;
}

int vf_plan_create(vf_plan** out, int n, const vf_volume* volumes, const vf_graph* graphs, const vf_frame* frames,
                   const vf_warp_params* params, const vf_tuning* tuning, int flags, vf_stream stream) {
    if (!out) return fail(VF_ERR_INVALID_ARGUMENT, "vf_plan_create: out is NULL");
    *out = nullptr;
    if (!volumes || !graphs || !frames || !params) return fail(VF_ERR_INVALID_ARGUMENT, "vf_plan_create: NULL argument");
    if (n < 1 || n > 256) return fail(VF_ERR_INVALID_ARGUMENT, "plan of %d volumes: supported 1..256", n);
    if (flags & ~VF_PLAN_GRAPH) return fail(VF_ERR_INVALID_ARGUMENT, "vf_plan_create: unknown flags 0x%x", flags);
    vf_plan* pl = new vf_plan();
    hipError_t e = hipGetDevice(&pl->ws.device);
    if (e != hipSuccess) {
        delete pl;
        return fail(VF_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e));
    }
    pl->n = n;
    pl->flags = flags;
    pl->vols.assign(volumes, volumes + n);
    pl->graphs.assign(graphs, graphs + n);
    pl->frames.assign(frames, frames + n);
    pl->params.assign(params, params + n);
    if (tuning) {
        pl->tuning = *tuning;
    } else {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        auto it = g_tune_cache.find(tune_key(n, volumes, graphs, params));
        if (it != g_tune_cache.end()) pl->tuning = it->second;
    }
    int st = plan_setup(pl, (hipStream_t)stream);
    if (st) {
        free_buffers(pl->ws.buf);
        delete pl;
        return st;
    }
    *out = pl;
    return VF_OK;
}

int vf_plan_destroy(vf_plan* pl) {
    if (!pl) return VF_OK;
    plan_drop_graph(pl);
    if (pl->capture_stream) (void)hipStreamDestroy(pl->capture_stream);
    for (auto& e : pl->ws.events) (void)hipEventDestroy(e);
    free_buffers(pl->ws.buf);
    delete pl;
    return VF_OK;
}

int vf_plan_launch(vf_plan* pl, vf_stream stream) {
    if (!pl) return fail(VF_ERR_INVALID_ARGUMENT, "vf_plan_launch: NULL plan");
    if (pl->flags & VF_PLAN_GRAPH) {
        if (!pl->exec) {
            int st = plan_record_graph(pl);
            if (st) return st;
        }
        VF_HIP(hipGraphLaunch(pl->exec, (hipStream_t)stream));
        return VF_OK;
    }
    return launch_kernels(&pl->ws, pl->fs, -1, KnnOut{nullptr, nullptr, nullptr}, nullptr, (hipStream_t)stream);
}

int vf_plan_autotune(vf_plan* pl, int frames, vf_tuning* chosen, vf_stream stream) {
    if (!pl) return fail(VF_ERR_INVALID_ARGUMENT, "vf_plan_autotune: NULL plan");
    return plan_autotune(pl, frames, chosen, (hipStream_t)stream);
}

int vf_plan_get_tuning(const vf_plan* pl, vf_tuning* tuning) {
    if (!pl || !tuning) return fail(VF_ERR_INVALID_ARGUMENT, "vf_plan_get_tuning: NULL argument");
    *tuning = pl->tuning;
    return VF_OK;
}

int vf_plan_stats(vf_plan* pl, vf_frame_stats* last, vf_frame_stats* totals, int reset, vf_stream stream) {
    if (!pl) return fail(VF_ERR_INVALID_ARGUMENT, "vf_plan_stats: NULL plan");
    int st = VF_OK;
    if (last) st = read_stats(&pl->ws, false, last, (hipStream_t)stream);
    if (!st && totals) st = read_stats(&pl->ws, true, totals, (hipStream_t)stream);
    if (!st && reset && pl->ws.buf.totals) {
        VF_HIP(hipMemsetAsync(pl->ws.buf.totals, 0, (uint64_t)kStCount * kStatSlots * 8, (hipStream_t)stream));
        VF_HIP(hipStreamSynchronize((hipStream_t)stream));
    }
    return st;
}

uint64_t vf_plan_bytes(const vf_plan* pl) { return pl ? pl->ws.buf.bytes : 0; }

int vf_tsdf_knn_graph(int k, uint32_t* knn_idxs, float* knn_dists, int32_t* knn_counts, const vf_float3* graph,
                      const float* graph_radii, size_t graph_size, vf_float3 voxel_size, vf_float3 volume_size,
                      vf_int3 volume_dim, vf_int3 voxels_per_node, vf_stream stream) {
    int st = check_k(k);
    if (st) return st;
    uint64_t n;
    st = check_volume(volume_dim, voxel_size, &n);
    if (st) return st;
    if (graph_size == 0) return VF_OK;
    if (!knn_idxs || !knn_dists || !knn_counts) return fail(VF_ERR_INVALID_ARGUMENT, "vf_tsdf_knn_graph: NULL buffer");
    int dev = 0;
    VF_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_knn_ws_mutex);
    if ((int)g_knn_ws.size() <= dev) g_knn_ws.resize(dev + 1, nullptr);
    if (!g_knn_ws[dev]) {
        g_knn_ws[dev] = new vf_workspace();
        g_knn_ws[dev]->device = dev;
    }
    vf_volume vol;
    memset(&vol, 0, sizeof(vol));
    vol.dim = volume_dim;
    vol.voxel_size = voxel_size;
    vol.size = volume_size;
    vf_graph gr;
    memset(&gr, 0, sizeof(gr));
    gr.points = graph;
    gr.radii = graph_radii;
    gr.count = graph_size;
    vf_warp_params p;
    memset(&p, 0, sizeof(p));
    p.k = k;
    p.voxels_per_node = voxels_per_node;
    return launch_frame(g_knn_ws[dev], 1, &vol, &gr, nullptr, &p, true, KnnOut{knn_idxs, knn_dists, knn_counts},
                        (hipStream_t)stream);
}

}  // extern "C"
