// vf_fused.hip -- the warped TSDF integration of one frame (or a batch of
// independent volumes, one frame each) in three kernels.
//
// The reference runs nine dense launches with ~100 B of per-voxel scratch
// (src/VolumeFusion.cu:699-742, buffers at :252-260).  Here:
//
//   k_prep_bin  one thread per (node, brick overlapped by the node's cube):
//               node record (translated position, cube corner, 2r, 2r^2,
//               transform rows), append the node to the brick's list.  Also
//               clears the collision tiles the previous frame touched.
//   k_pass_a    one 512-thread workgroup per active brick (bricks of
//               BX x BY x BZ voxels, x fastest): stage the brick's node list in
//               LDS (sorted by node index), per voxel select the canonical
//               top-K, weights, blend, target voxel; count collisions in a
//               2-bit-per-voxel saturating bitmap (8^3 tiles); voxels that
//               pass the projective tests append a 16 B pending record
//               (voxel, target, new tsdf, new weight).
//   k_pass_b    one thread per pending record: drop it if its target was
//               reached twice, else write tsdf/weight.
//
// Unsupported voxels (no node cube covers their brick) are never touched, so
// HBM traffic scales with the supported shell, not with the dense volume.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "vf_host.hpp"

namespace vf {
namespace {

constexpr int kCap = 128;           // node slots per brick list (and chunk size when a list spills)
constexpr int kPassAThreads = 512;  // one voxel per thread, one brick per workgroup
constexpr uint32_t kNone = 0xffffffffu;
constexpr int kBX = 16, kBY = 8, kBZ = 4;  // brick shape; 64 B rows of tsdf per brick row

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
    int32_t tdim[3];
    int32_t maxb[3];
    uint32_t G;
    uint32_t node_base;
    uint32_t brick_base;
    uint32_t tile_base;
    uint32_t k1_base;
    uint32_t k1_count;
};

struct NodeRec {
    float4 m;   // translated position; .w = 2r (the candidate bound)
    int4 lo;    // cube lower corner; .w = bits of 2*(r*r) (the weight denominator)
    float4 t0;  // transform rows (T(i,0), T(i,1), T(i,2), T(i,3))
    float4 t1;
    float4 t2;
};

struct Ctrl {
    uint32_t active_count;
    uint32_t work_next;
    uint32_t record_count;
    uint32_t touched_count[2];
    uint32_t overflow_bricks;
    uint32_t pad[2];
    unsigned long long stats[8];   // this frame: supported, in_bounds, pending, collided, updated, active bricks
    unsigned long long totals[8];  // since the last reset: the same counters summed over frames
};

struct KnnOut {
    uint32_t* idx;
    float* dist;
    int32_t* cnt;
};

__device__ __forceinline__ uint32_t tile_code(const VolDesc& vd, const int32_t u[3]) {
    const uint32_t t = ((uint32_t)(u[2] >> 3) * (uint32_t)vd.tdim[1] + (uint32_t)(u[1] >> 3)) * (uint32_t)vd.tdim[0] +
                       (uint32_t)(u[0] >> 3);
    const uint32_t bit = ((uint32_t)(u[2] & 7) << 6) | ((uint32_t)(u[1] & 7) << 3) | (uint32_t)(u[0] & 7);
    return (t << 9) | bit;
}

__global__ __launch_bounds__(256) void k_prep_bin(const VolDesc* __restrict__ vols, int nvol, uint32_t total,
                                                  NodeRec* __restrict__ nrec, uint32_t* __restrict__ brick_count,
                                                  uint32_t* __restrict__ brick_nodes, uint32_t* __restrict__ active,
                                                  Ctrl* __restrict__ ctrl, uint32_t* __restrict__ bitmap,
                                                  uint32_t* __restrict__ tile_flag,
                                                  const uint32_t* __restrict__ touched_prev, uint32_t prev_parity) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nthreads = gridDim.x * blockDim.x;

    // Clear the collision tiles the previous frame on this workspace touched.
    const uint32_t nt = ctrl->touched_count[prev_parity];
    for (uint32_t w = gid; w < nt * 33u; w += nthreads) {
        const uint32_t tile = touched_prev[w / 33u];
        const uint32_t k = w % 33u;
        if (k < 32u)
            bitmap[(uint64_t)tile * 32u + k] = 0u;
        else
            tile_flag[tile] = 0u;
    }
    if (gid == 0) {
        ctrl->record_count = 0u;
        ctrl->overflow_bricks = 0u;
        for (int s = 0; s < 8; ++s) ctrl->stats[s] = 0ull;
    }
    if (gid >= total) return;

    int v = 0;
    while (v + 1 < nvol && vols[v + 1].k1_base <= gid) ++v;
    const VolDesc& vd = vols[v];
    const uint32_t local = gid - vd.k1_base;
    if (local >= vd.k1_count) return;
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
        rec.lo = make_int4(lo[0], lo[1], lo[2], __float_as_int(den));
        if (vd.T) {  // the knn-only entry point has no transforms
            const float* M = vd.T + 16ull * j;
            rec.t0 = make_float4(M[0], M[4], M[8], M[12]);
            rec.t1 = make_float4(M[1], M[5], M[9], M[13]);
            rec.t2 = make_float4(M[2], M[6], M[10], M[14]);
        } else {
            rec.t0 = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
            rec.t1 = make_float4(0.0f, 1.0f, 0.0f, 0.0f);
            rec.t2 = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
        }
        nrec[vd.node_base + j] = rec;
    }
    // bricks overlapped by the cube clipped to the volume
    int32_t blo[3], bhi[3];
    const int32_t bs[3] = {kBX, kBY, kBZ};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const long long l = lo[a] < 0 ? 0 : lo[a];
        long long h = (long long)lo[a] + vd.cube[a] - 1;
        if (h > vd.g.dim[a] - 1) h = vd.g.dim[a] - 1;
        if (l > h) return;
        blo[a] = (int32_t)(l / bs[a]);
        bhi[a] = (int32_t)(h / bs[a]);
    }
    const int32_t ox = (int32_t)(o % (uint32_t)vd.maxb[0]);
    const int32_t oy = (int32_t)((o / (uint32_t)vd.maxb[0]) % (uint32_t)vd.maxb[1]);
    const int32_t oz = (int32_t)(o / (uint32_t)(vd.maxb[0] * vd.maxb[1]));
    const int32_t bx = blo[0] + ox, by = blo[1] + oy, bz = blo[2] + oz;
    if (bx > bhi[0] || by > bhi[1] || bz > bhi[2]) return;
    const uint32_t b = ((uint32_t)bz * (uint32_t)vd.bdim[1] + (uint32_t)by) * (uint32_t)vd.bdim[0] + (uint32_t)bx;
    const uint32_t gb = vd.brick_base + b;
    const uint32_t slot = atomicAdd(brick_count + gb, 1u);
    if (slot < (uint32_t)kCap)
        brick_nodes[(uint64_t)gb * kCap + slot] = j;
    else if (slot == (uint32_t)kCap)
        atomicAdd(&ctrl->overflow_bricks, 1u);
    if (slot == 0) {
        const uint32_t a = atomicAdd(&ctrl->active_count, 1u);
        active[a] = ((uint32_t)v << 24) | b;
    }
}

template <int K, bool KNN_ONLY>
__global__ __launch_bounds__(kPassAThreads) void k_pass_a(
    const VolDesc* __restrict__ vols, const NodeRec* __restrict__ nrec, uint32_t* __restrict__ brick_count,
    const uint32_t* __restrict__ brick_nodes, const uint32_t* __restrict__ active, Ctrl* __restrict__ ctrl,
    uint32_t* __restrict__ bitmap, uint32_t* __restrict__ tile_flag, uint32_t* __restrict__ touched_cur,
    uint32_t parity, uint4* __restrict__ records, uint32_t vbits, KnnOut kout) {
    __shared__ uint32_t s_raw[kCap];
    __shared__ uint32_t s_ids[kCap];
    __shared__ float4 s_m[kCap];
    __shared__ int4 s_lo[kCap];
    __shared__ float4 s_t0[kCap];
    __shared__ float4 s_t1[kCap];
    __shared__ float4 s_t2[kCap];
    __shared__ uint32_t s_item;
    __shared__ unsigned long long s_stat[3];

    const int tid = threadIdx.x;
    const int lane = __lane_id();
    if (tid < 3) s_stat[tid] = 0ull;
    if (!KNN_ONLY && blockIdx.x == 0 && tid == 0) ctrl->touched_count[parity ^ 1u] = 0u;
    const uint32_t nactive = ctrl->active_count;

    const int lx = tid % kBX;
    const int ly = (tid / kBX) % kBY;
    const int lz = tid / (kBX * kBY);

    for (;;) {
        if (tid == 0) {
            const uint32_t i = atomicAdd(&ctrl->work_next, 1u);
            s_item = i < nactive ? active[i] : kNone;
        }
        __syncthreads();
        const uint32_t item = __builtin_amdgcn_readfirstlane(s_item);
        if (item == kNone) break;
        const uint32_t vol = item >> 24;
        const uint32_t b = item & 0xffffffu;
        const VolDesc& vd = vols[vol];
        const uint32_t gb = vd.brick_base + b;
        const uint32_t cnt = brick_count[gb];
        const bool spill = cnt > (uint32_t)kCap;
        const uint32_t nchunks = spill ? (vd.G + kCap - 1) / kCap : 1u;

        const int32_t bxi = (int32_t)(b % (uint32_t)vd.bdim[0]);
        const int32_t byi = (int32_t)((b / (uint32_t)vd.bdim[0]) % (uint32_t)vd.bdim[1]);
        const int32_t bzi = (int32_t)(b / (uint32_t)(vd.bdim[0] * vd.bdim[1]));
        const int32_t x = bxi * kBX + lx, y = byi * kBY + ly, z = bzi * kBZ + lz;
        const bool valid = x < vd.g.dim[0] && y < vd.g.dim[1] && z < vd.g.dim[2];

        float p[3];
        voxel_to_world(vd.g, x, y, z, p);
        KnnList<K> kl;
        kl.init();

        for (uint32_t ch = 0; ch < nchunks; ++ch) {
            uint32_t m;
            if (!spill) {
                m = cnt;
                if ((uint32_t)tid < m) s_raw[tid] = brick_nodes[(uint64_t)gb * kCap + tid];
                __syncthreads();
                if ((uint32_t)tid < m) {
                    const uint32_t id = s_raw[tid];
                    uint32_t rank = 0;
                    for (uint32_t e = 0; e < m; ++e) rank += s_raw[e] < id;
                    s_ids[rank] = id;
                }
                __syncthreads();
            } else {
                const uint32_t base = ch * (uint32_t)kCap;
                m = vd.G - base < (uint32_t)kCap ? vd.G - base : (uint32_t)kCap;
                if ((uint32_t)tid < m) s_ids[tid] = base + tid;
                __syncthreads();
            }
            if ((uint32_t)tid < m) {
                const NodeRec rec = nrec[vd.node_base + s_ids[tid]];
                s_m[tid] = rec.m;
                s_lo[tid] = rec.lo;
                s_t0[tid] = rec.t0;
                s_t1[tid] = rec.t1;
                s_t2[tid] = rec.t2;
            }
            __syncthreads();
            if (valid) {
                for (uint32_t e = 0; e < m; ++e) {
                    const int4 l = s_lo[e];
                    if ((uint32_t)(x - l.x) < (uint32_t)vd.cube[0] && (uint32_t)(y - l.y) < (uint32_t)vd.cube[1] &&
                        (uint32_t)(z - l.z) < (uint32_t)vd.cube[2]) {
                        const float4 nm = s_m[e];
                        const float dx = p[0] - nm.x, dy = p[1] - nm.y, dz = p[2] - nm.z;
                        const float s = dx * dx + (dy * dy + dz * dz);
                        const float d = __builtin_sqrtf(s);
                        if (!(d > nm.w)) kl.offer(d, spill ? s_ids[e] : e);
                    }
                }
            }
            if (spill && ch + 1 < nchunks) __syncthreads();
        }
        if (tid == 0) brick_count[gb] = 0u;

        const int n = valid ? kl.n : 0;
        const uint64_t vidx = ((uint64_t)z * vd.g.dim[1] + y) * vd.g.dim[0] + x;

        if (KNN_ONLY) {
            if (n > 0) {
                kout.cnt[vidx] = n;
#pragma unroll
                for (int s = 0; s < K; ++s) {
                    if (s < n) {
                        kout.idx[(uint64_t)K * vidx + s] = spill ? kl.i[s] : s_ids[kl.i[s]];
                        kout.dist[(uint64_t)K * vidx + s] = kl.d[s];
                    }
                }
            }
        } else {
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
                        const int32_t denb = spill ? nrec[vd.node_base + kl.i[s]].lo.w : s_lo[kl.i[s]].w;
                        w[s] = warp_weight(__int_as_float(denb), kl.d[s]);
                        sum = sum + w[s];
                    }
                }
                bool zero = true;
                if (sum != 0.0f) {
#pragma unroll
                    for (int s = 0; s < K; ++s)
                        if (s < n) w[s] = w[s] / sum;
                }
#pragma unroll
                for (int s = 0; s < K; ++s)
                    if (s < n) zero = zero && (__builtin_fabsf(w[s]) <= 1e-5f);
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
                                r0 = s_t0[kl.i[s]];
                                r1 = s_t1[kl.i[s]];
                                r2 = s_t2[kl.i[s]];
                            }
                            float t[3];
                            affine_apply(r0, r1, r2, pc, t);
#pragma unroll
                            for (int a = 0; a < 3; ++a) {
                                const float wt = w[s] * t[a];
                                acc[a] = acc[a] + wt;
                            }
                        }
                    }
                }
                // collision target (src/TSDFIntegration.cu:272-282)
                int32_t u[3];
                inb = warped_target(vd.g, vd.rinv, vd.c, acc, u);
                if (inb) {
                    ucode = tile_code(vd, u);
                    const uint32_t tile = vd.tile_base + (ucode >> 9);
                    const uint32_t bit = ucode & 511u;
                    uint32_t* blk = bitmap + (uint64_t)tile * 32u;
                    const uint32_t mask = 1u << (bit & 31u);
                    const uint32_t old = atomicOr(blk + (bit >> 5), mask);
                    if (old & mask) atomicOr(blk + 16 + (bit >> 5), mask);
                    if (__atomic_load_n(tile_flag + tile, __ATOMIC_RELAXED) == 0u) {
                        if (atomicCAS(tile_flag + tile, 0u, 1u) == 0u) {
                            const uint32_t k = atomicAdd(&ctrl->touched_count[parity], 1u);
                            touched_cur[k] = tile;
                        }
                    }
                    // projective tests and the new sample (src/TSDFIntegration.cu:32-64)
                    float tn;
                    if (project_sample(vd.cam, vd.depth, acc, &tn) >= 0) {
                        blend_sample(vd.tsdf[vidx], vd.wt[vidx], tn, vd.cam.maxw, &tnew, &wnew);
                        emit = true;
                    }
                }
            }
            const unsigned long long msup = __ballot(n > 0);
            const unsigned long long minb = __ballot(inb);
            const unsigned long long memit = __ballot(emit);
            if (lane == 0) {
                atomicAdd(&s_stat[0], (unsigned long long)__popcll(msup));
                atomicAdd(&s_stat[1], (unsigned long long)__popcll(minb));
                atomicAdd(&s_stat[2], (unsigned long long)__popcll(memit));
            }
            if (memit) {
                const int leader = __ffsll((long long)memit) - 1;
                uint32_t base = 0;
                if (lane == leader) base = atomicAdd(&ctrl->record_count, (uint32_t)__popcll(memit));
                base = __shfl(base, leader);
                if (emit) {
                    const uint32_t rank = (uint32_t)__popcll(memit & ((1ull << lane) - 1ull));
                    const uint32_t vcode = vbits >= 32u ? (uint32_t)vidx : (((uint32_t)vol << vbits) | (uint32_t)vidx);
                    records[base + rank] = make_uint4(vcode, ucode, __float_as_uint(tnew), __float_as_uint(wnew));
                }
            }
        }
        __syncthreads();
    }
    __syncthreads();
    if (!KNN_ONLY && tid == 0) {
        atomicAdd(&ctrl->stats[0], s_stat[0]);
        atomicAdd(&ctrl->stats[1], s_stat[1]);
        atomicAdd(&ctrl->stats[2], s_stat[2]);
        atomicAdd(&ctrl->totals[0], s_stat[0]);
        atomicAdd(&ctrl->totals[1], s_stat[1]);
        atomicAdd(&ctrl->totals[2], s_stat[2]);
    }
}

struct VolPtr {
    float* tsdf;
    float* wt;
    uint32_t tile_base;
    uint32_t pad;
};

__global__ __launch_bounds__(256) void k_pass_b(const VolPtr* __restrict__ vptr, const uint4* __restrict__ records,
                                                Ctrl* __restrict__ ctrl, const uint32_t* __restrict__ bitmap,
                                                uint32_t vbits, const VolDesc* __restrict__ vols, int nvol,
                                                const NodeRec* __restrict__ nrec) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nthreads = gridDim.x * blockDim.x;
    if (gid == 0) {
        ctrl->stats[5] = ctrl->active_count;
        ctrl->totals[5] += ctrl->active_count;
        ctrl->active_count = 0u;
        ctrl->work_next = 0u;
    }
    const uint32_t n = ctrl->record_count;
    const uint32_t vmask = vbits >= 32u ? 0xffffffffu : ((1u << vbits) - 1u);
    unsigned long long collided = 0, updated = 0;
    for (uint32_t r = gid; r < n; r += nthreads) {
        const uint4 rec = records[r];
        const uint32_t vol = vbits >= 32u ? 0u : rec.x >> vbits;
        const uint32_t v = rec.x & vmask;
        const VolPtr vp = vptr[vol];
        const uint32_t tile = vp.tile_base + (rec.y >> 9);
        const uint32_t bit = rec.y & 511u;
        const uint32_t word = bitmap[(uint64_t)tile * 32u + 16u + (bit >> 5)];
        if ((word >> (bit & 31u)) & 1u) {
            ++collided;
        } else {
            vp.tsdf[v] = __uint_as_float(rec.z);
            vp.wt[v] = __uint_as_float(rec.w);
            ++updated;
        }
    }
    // node positions as the reference's two translations leave them,
    // fl(fl(p - c) + c) (src/VolumeFusion.cu:708, 721)
    for (int v = 0; v < nvol; ++v) {
        const VolDesc& vd = vols[v];
        if (!vd.nodes_out) continue;
        for (uint32_t j = gid; j < vd.G; j += nthreads) {
            const float4 m = nrec[vd.node_base + j].m;
            vd.nodes_out[3ull * j + 0] = m.x + vd.c[0];
            vd.nodes_out[3ull * j + 1] = m.y + vd.c[1];
            vd.nodes_out[3ull * j + 2] = m.z + vd.c[2];
        }
    }
    // wave sums, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
        collided += __shfl_down(collided, off);
        updated += __shfl_down(updated, off);
    }
    if (__lane_id() == 0) {
        if (collided) {
            atomicAdd(&ctrl->stats[3], collided);
            atomicAdd(&ctrl->totals[3], collided);
        }
        if (updated) {
            atomicAdd(&ctrl->stats[4], updated);
            atomicAdd(&ctrl->totals[4], updated);
        }
    }
}

__global__ void k_reset_after_knn(Ctrl* ctrl) {
    ctrl->active_count = 0u;
    ctrl->work_next = 0u;
}

}  // namespace

// Device buffers of the brick machinery, grown on demand.
struct Buffers {
    Ctrl* ctrl = nullptr;
    NodeRec* nrec = nullptr;
    uint32_t* brick_count = nullptr;
    uint32_t* brick_nodes = nullptr;
    uint32_t* active = nullptr;
    uint32_t* bitmap = nullptr;
    uint32_t* tile_flag = nullptr;
    uint32_t* touched = nullptr;  // 2 x tile_cap
    uint4* records = nullptr;
    VolDesc* vols = nullptr;
    VolPtr* vptr = nullptr;
    uint64_t node_cap = 0, brick_cap = 0, tile_cap = 0, record_cap = 0, vol_cap = 0;
    uint64_t bytes = 0;
};

}  // namespace vf

struct vf_workspace {
    int device = 0;
    vf::Buffers buf;
    uint64_t frame = 0;  // selects the touched-tile list (frame parity)
    std::vector<vf::VolDesc> vols_h;
    std::vector<vf::VolPtr> vptr_h;
    // optional per-kernel timing: 4 events per frame (before k_prep_bin, after
    // each of the three kernels), recycled every kMaxTimed frames
    bool timing = false;
    std::vector<hipEvent_t> events;
    int timed_frames = 0;
};

namespace vf {
namespace {

void free_buffers(Buffers& b) {
    void* ptrs[] = {b.ctrl, b.nrec, b.brick_count, b.brick_nodes, b.active, b.bitmap,
                    b.tile_flag, b.touched, b.records, b.vols, b.vptr};
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

// (Re)allocates to at least the requested capacities; new buffers are zeroed.
int ensure(vf_workspace* ws, uint64_t nodes, uint64_t bricks, uint64_t tiles, uint64_t records, uint64_t nvol,
           hipStream_t stream) {
    Buffers& b = ws->buf;
    if (b.ctrl && nodes <= b.node_cap && bricks <= b.brick_cap && tiles <= b.tile_cap && records <= b.record_cap &&
        nvol <= b.vol_cap)
        return VF_OK;
    if (b.ctrl) VF_HIP(hipStreamSynchronize(stream));
    Buffers nbuf;
    nbuf.node_cap = std::max<uint64_t>(std::max(nodes, b.node_cap), 1);
    nbuf.brick_cap = std::max<uint64_t>(std::max(bricks, b.brick_cap), 1);
    nbuf.tile_cap = std::max<uint64_t>(std::max(tiles, b.tile_cap), 1);
    nbuf.record_cap = std::max<uint64_t>(std::max(records, b.record_cap), 1);
    nbuf.vol_cap = std::max<uint64_t>(std::max(nvol, b.vol_cap), 1);
    free_buffers(b);
    int st = VF_OK;
    uint64_t bytes = 0;
    if (!st) st = dev_alloc(&nbuf.ctrl, 1, &bytes);
    if (!st) st = dev_alloc(&nbuf.nrec, nbuf.node_cap, &bytes);
    if (!st) st = dev_alloc(&nbuf.brick_count, nbuf.brick_cap, &bytes);
    if (!st) st = dev_alloc(&nbuf.brick_nodes, nbuf.brick_cap * kCap, &bytes);
    if (!st) st = dev_alloc(&nbuf.active, nbuf.brick_cap, &bytes);
    if (!st) st = dev_alloc(&nbuf.bitmap, nbuf.tile_cap * 32, &bytes);
    if (!st) st = dev_alloc(&nbuf.tile_flag, nbuf.tile_cap, &bytes);
    if (!st) st = dev_alloc(&nbuf.touched, nbuf.tile_cap * 2, &bytes);
    if (!st) st = dev_alloc(&nbuf.records, nbuf.record_cap, &bytes);
    if (!st) st = dev_alloc(&nbuf.vols, nbuf.vol_cap, &bytes);
    if (!st) st = dev_alloc(&nbuf.vptr, nbuf.vol_cap, &bytes);
    if (st) {
        free_buffers(nbuf);
        return st;
    }
    nbuf.bytes = bytes;
    // Zero state the kernels rely on: control block, brick counts, tiles.
    VF_HIP(hipMemsetAsync(nbuf.ctrl, 0, sizeof(Ctrl), stream));
    VF_HIP(hipMemsetAsync(nbuf.brick_count, 0, nbuf.brick_cap * sizeof(uint32_t), stream));
    VF_HIP(hipMemsetAsync(nbuf.bitmap, 0, nbuf.tile_cap * 32 * sizeof(uint32_t), stream));
    VF_HIP(hipMemsetAsync(nbuf.tile_flag, 0, nbuf.tile_cap * sizeof(uint32_t), stream));
    VF_HIP(hipStreamSynchronize(stream));
    b = nbuf;
    ws->frame = 0;
    return VF_OK;
}

inline int32_t ceil_div(int32_t a, int32_t b) { return (a + b - 1) / b; }

// Fills `vd` for one volume; accumulates the capacities the call needs.
int make_desc(const vf_volume& vol, const vf_graph& gr, const vf_frame& fr, const vf_warp_params& p, bool knn_only,
              VolDesc* vd, uint64_t* n_out) {
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
    const int32_t bs[3] = {kBX, kBY, kBZ};
    const int32_t dims[3] = {vol.dim.x, vol.dim.y, vol.dim.z};
    for (int a = 0; a < 3; ++a) {
        vd->bdim[a] = ceil_div(dims[a], bs[a]);
        vd->tdim[a] = ceil_div(dims[a], 8);
        const int32_t c = vd->cube[a] < 1 ? 1 : vd->cube[a];
        vd->maxb[a] = ceil_div(c - 1, bs[a]) + 1;
        if (vd->maxb[a] > vd->bdim[a]) vd->maxb[a] = vd->bdim[a];
    }
    const bool no_cube = vd->cube[0] < 1 || vd->cube[1] < 1 || vd->cube[2] < 1;
    vd->G = (uint32_t)gr.count;
    const uint64_t k1 = no_cube ? 0 : (uint64_t)vd->G * (uint64_t)(vd->maxb[0] * vd->maxb[1] * vd->maxb[2]);
    if (k1 > 0xffffffffull) return fail(VF_ERR_CAPACITY, "node x brick work of %llu exceeds 2^32", (unsigned long long)k1);
    vd->k1_count = (uint32_t)k1;
    *n_out = n;
    return VF_OK;
}

int launch_frame(vf_workspace* ws, int nvol, const vf_volume* vols, const vf_graph* graphs, const vf_frame* frames,
                 const vf_warp_params* p, bool knn_only, KnnOut kout, hipStream_t stream) {
    int st = check_k(p->k);
    if (st) return st;
    if (nvol < 1 || nvol > 256) return fail(VF_ERR_INVALID_ARGUMENT, "batch of %d volumes: supported 1..256", nvol);
    for (int v = 1; v < nvol; ++v)
        if (p[v].k != p[0].k) return fail(VF_ERR_INVALID_ARGUMENT, "batched volumes must share k");
    ws->vols_h.resize(nvol);
    ws->vptr_h.resize(nvol);
    uint64_t nodes = 0, bricks = 0, tiles = 0, voxels = 0, maxn = 0, k1 = 0;
    for (int v = 0; v < nvol; ++v) {
        const vf_frame empty_frame = {nullptr, 0, 0, {{0}}};
        VolDesc& vd = ws->vols_h[v];
        uint64_t n;
        st = make_desc(vols[v], graphs[v], frames ? frames[v] : empty_frame, p[v], knn_only, &vd, &n);
        if (st) return st;
        vd.node_base = (uint32_t)nodes;
        vd.brick_base = (uint32_t)bricks;
        vd.tile_base = (uint32_t)tiles;
        vd.k1_base = (uint32_t)k1;
        nodes += vd.G;
        bricks += (uint64_t)vd.bdim[0] * vd.bdim[1] * vd.bdim[2];
        tiles += (uint64_t)vd.tdim[0] * vd.tdim[1] * vd.tdim[2];
        voxels += n;
        k1 += vd.k1_count;
        maxn = std::max(maxn, n);
        ws->vptr_h[v].tsdf = vd.tsdf;
        ws->vptr_h[v].wt = vd.wt;
        ws->vptr_h[v].tile_base = vd.tile_base;
        ws->vptr_h[v].pad = 0;
        if (bricks > 0xffffffffull || tiles > 0x7fffffull || k1 > 0xffffffffull)
            return fail(VF_ERR_CAPACITY, "batch too large for 32-bit brick/tile/work indices");
        if ((uint64_t)vd.bdim[0] * vd.bdim[1] * vd.bdim[2] > 0xffffffull)
            return fail(VF_ERR_CAPACITY, "volume has more than 2^24 bricks");
    }
    uint32_t vbits = 32;
    if (nvol > 1) {
        vbits = 1;
        while (vbits < 32 && (1ull << vbits) < maxn) ++vbits;
        if (vbits + (uint32_t)std::ceil(std::log2((double)nvol)) > 32)
            return fail(VF_ERR_CAPACITY, "batch of %d volumes with %llu voxels each does not fit 32-bit records", nvol,
                        (unsigned long long)maxn);
    }
    st = ensure(ws, nodes, bricks, tiles, knn_only ? 1 : voxels, (uint64_t)nvol, stream);
    if (st) return st;
    Buffers& B = ws->buf;
    VF_HIP(hipMemcpyAsync(B.vols, ws->vols_h.data(), sizeof(VolDesc) * nvol, hipMemcpyHostToDevice, stream));
    VF_HIP(hipMemcpyAsync(B.vptr, ws->vptr_h.data(), sizeof(VolPtr) * nvol, hipMemcpyHostToDevice, stream));

    const uint32_t parity = (uint32_t)(ws->frame & 1u);
    const uint32_t prev = parity ^ 1u;
    hipEvent_t* ev = nullptr;
    if (ws->timing && !knn_only) {
        constexpr int kMaxTimed = 4096;
        if (ws->events.empty()) {
            ws->events.resize(4 * kMaxTimed);
            for (auto& e : ws->events) VF_HIP(hipEventCreate(&e));
        }
        if (ws->timed_frames >= kMaxTimed) return fail(VF_ERR_CAPACITY, "more than %d timed frames", kMaxTimed);
        ev = &ws->events[4 * ws->timed_frames];
        ws->timed_frames++;
        VF_HIP(hipEventRecord(ev[0], stream));
    }
    {
        uint64_t blocks = std::max<uint64_t>((k1 + 255) / 256, 1024);
        hipLaunchKernelGGL(k_prep_bin, dim3((unsigned)blocks), dim3(256), 0, stream, B.vols, nvol, (uint32_t)k1, B.nrec,
                           B.brick_count, B.brick_nodes, B.active, B.ctrl, B.bitmap, B.tile_flag,
                           B.touched + (uint64_t)prev * B.tile_cap, prev);
        VF_LAUNCH_CHECK("k_prep_bin");
        if (ev) VF_HIP(hipEventRecord(ev[1], stream));
    }
    int dev = 0, cus = 0;
    VF_HIP(hipGetDevice(&dev));
    VF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    {
        int per_cu = 0;
#define VF_PASS_A_OCC(KK)                                                                                        \
    VF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(                                                         \
        &per_cu, knn_only ? (const void*)k_pass_a<KK, true> : (const void*)k_pass_a<KK, false>, kPassAThreads, 0))
        switch (p->k) {
            case 1: VF_PASS_A_OCC(1); break;
            case 2: VF_PASS_A_OCC(2); break;
            case 3: VF_PASS_A_OCC(3); break;
            case 4: VF_PASS_A_OCC(4); break;
            case 5: VF_PASS_A_OCC(5); break;
            case 6: VF_PASS_A_OCC(6); break;
            case 7: VF_PASS_A_OCC(7); break;
            case 8: VF_PASS_A_OCC(8); break;
        }
#undef VF_PASS_A_OCC
        if (per_cu < 1) per_cu = 1;
        uint64_t blocks = (uint64_t)cus * (uint64_t)per_cu;
        blocks = std::min<uint64_t>(blocks, std::max<uint64_t>(bricks, 1));
        uint32_t* touched_cur = B.touched + (uint64_t)parity * B.tile_cap;
#define VF_PASS_A(KK)                                                                                                \
    if (knn_only)                                                                                                    \
        hipLaunchKernelGGL((k_pass_a<KK, true>), dim3((unsigned)blocks), dim3(kPassAThreads), 0, stream, B.vols,     \
                           B.nrec, B.brick_count, B.brick_nodes, B.active, B.ctrl, B.bitmap, B.tile_flag,            \
                           touched_cur, parity, B.records, vbits, kout);                                             \
    else                                                                                                             \
        hipLaunchKernelGGL((k_pass_a<KK, false>), dim3((unsigned)blocks), dim3(kPassAThreads), 0, stream, B.vols,    \
                           B.nrec, B.brick_count, B.brick_nodes, B.active, B.ctrl, B.bitmap, B.tile_flag,            \
                           touched_cur, parity, B.records, vbits, kout)
        switch (p->k) {
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
        VF_LAUNCH_CHECK("k_pass_a");
        if (ev) VF_HIP(hipEventRecord(ev[2], stream));
    }
    if (knn_only) {
        hipLaunchKernelGGL(k_reset_after_knn, dim3(1), dim3(1), 0, stream, B.ctrl);
        VF_LAUNCH_CHECK("k_reset_after_knn");
    } else {
        const uint64_t blocks = (uint64_t)cus * 8;
        hipLaunchKernelGGL(k_pass_b, dim3((unsigned)blocks), dim3(256), 0, stream, B.vptr, B.records, B.ctrl, B.bitmap,
                           vbits, B.vols, nvol, B.nrec);
        VF_LAUNCH_CHECK("k_pass_b");
        if (ev) VF_HIP(hipEventRecord(ev[3], stream));
        ws->frame++;
    }
    return VF_OK;
}

// Workspace behind vf_tsdf_knn_graph, which has the reference's signature and
// so no workspace argument: one per device, grown on demand.
std::mutex g_knn_ws_mutex;
std::vector<vf_workspace*> g_knn_ws;

}  // namespace
}  // namespace vf

using namespace vf;

extern "C" {

int vf_workspace_create(vf_workspace** out, int max_volumes, uint64_t max_total_voxels, uint64_t max_total_nodes,
                        vf_stream stream) {
    if (!out) return fail(VF_ERR_INVALID_ARGUMENT, "vf_workspace_create: out is NULL");
    *out = nullptr;
    vf_workspace* ws = new vf_workspace();
    VF_HIP(hipGetDevice(&ws->device));
    if (max_total_voxels || max_total_nodes || max_volumes > 0) {
        const uint64_t bricks = max_total_voxels / (kBX * kBY * kBZ) + 64ull * (uint64_t)std::max(max_volumes, 1) * 64;
        const uint64_t tiles = max_total_voxels / 512 + 64ull * (uint64_t)std::max(max_volumes, 1) * 64;
        int st = ensure(ws, max_total_nodes, bricks, tiles, max_total_voxels, (uint64_t)std::max(max_volumes, 1),
                        (hipStream_t)stream);
        if (st) {
            delete ws;
            return st;
        }
    }
    *out = ws;
    return VF_OK;
}

int vf_workspace_destroy(vf_workspace* ws) {
    if (!ws) return VF_OK;
    for (auto& e : ws->events) (void)hipEventDestroy(e);
    free_buffers(ws->buf);
    delete ws;
    return VF_OK;
}

uint64_t vf_workspace_bytes(const vf_workspace* ws) { return ws ? ws->buf.bytes : 0; }

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
    memset(stats, 0, sizeof(*stats));
    if (!ws->buf.ctrl) return VF_OK;
    Ctrl c;
    VF_HIP(hipMemcpyAsync(&c, ws->buf.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, (hipStream_t)stream));
    VF_HIP(hipStreamSynchronize((hipStream_t)stream));
    stats->supported = c.stats[0];
    stats->in_bounds = c.stats[1];
    stats->pending = c.stats[2];
    stats->collided = c.stats[3];
    stats->updated = c.stats[4];
    stats->active_bricks = c.stats[5];
    stats->overflow_bricks = c.overflow_bricks;
    return VF_OK;
}

int vf_workspace_set_timing(vf_workspace* ws, int enable) {
    if (!ws) return fail(VF_ERR_INVALID_ARGUMENT, "vf_workspace_set_timing: NULL workspace");
    ws->timing = enable != 0;
    ws->timed_frames = 0;
    return VF_OK;
}

int vf_workspace_kernel_ms(vf_workspace* ws, float* ms3, int* frames) {
    if (!ws || !ms3 || !frames) return fail(VF_ERR_INVALID_ARGUMENT, "vf_workspace_kernel_ms: NULL argument");
    ms3[0] = ms3[1] = ms3[2] = 0.0f;
    *frames = ws->timed_frames;
    if (ws->timed_frames) VF_HIP(hipEventSynchronize(ws->events[4 * ws->timed_frames - 1]));
    for (int f = 0; f < ws->timed_frames; ++f) {
        for (int k = 0; k < 3; ++k) {
            float t = 0.0f;
            VF_HIP(hipEventElapsedTime(&t, ws->events[4 * f + k], ws->events[4 * f + k + 1]));
            ms3[k] += t;
        }
    }
    ws->timed_frames = 0;
    return VF_OK;
}

int vf_workspace_totals(vf_workspace* ws, vf_frame_stats* totals, int reset, vf_stream stream) {
    if (!ws || !totals) return fail(VF_ERR_INVALID_ARGUMENT, "vf_workspace_totals: NULL argument");
    memset(totals, 0, sizeof(*totals));
    if (!ws->buf.ctrl) return VF_OK;
    Ctrl c;
    VF_HIP(hipMemcpyAsync(&c, ws->buf.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, (hipStream_t)stream));
    VF_HIP(hipStreamSynchronize((hipStream_t)stream));
    totals->supported = c.totals[0];
    totals->in_bounds = c.totals[1];
    totals->pending = c.totals[2];
    totals->collided = c.totals[3];
    totals->updated = c.totals[4];
    totals->active_bricks = c.totals[5];
    if (reset) {
        VF_HIP(hipMemsetAsync(&ws->buf.ctrl->totals, 0, sizeof(c.totals), (hipStream_t)stream));
        VF_HIP(hipStreamSynchronize((hipStream_t)stream));
    }
    return VF_OK;
}

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
