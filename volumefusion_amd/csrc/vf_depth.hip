// vf_depth.hip -- depth-frame ingest: the steps before integration in the
// reference's frame loop (src/VolumeFusion.cu:185-219).
//
//   filterDepth      src/VertexManipulation.cu:6-23
//   maskDepth        src/VertexManipulation.cu:25-40
//   bilateral filter src/BilateralFilter.cu:52-127, 185-193
//   depthToMesh      src/VertexManipulation.cu:42-71
//   normals          src/VertexManipulation.cu:147-217
//   remove points without normals  src/VertexManipulation.cu:219-241
//   centroid         src/VolumeFusion.cu:216-219
// as drop-in entry points, and fused into three launches
// (vf_ingest_depth): filter+bilateral, then back-projection + normals +
// removal + per-block centroid sums, then the final sum.  Every step is
// per-pixel work on a 1-4 MB frame: memory and launch bound.
#include <cfloat>
#include <cmath>
#include <cstring>

#include "vf_host.hpp"

namespace vf {
namespace {

constexpr int kTileX = 16, kTileY = 16;
constexpr int kMaxRadius = 9;  // the reference's 400-entry spatial table

struct BilateralTables {
    float g2d[400];
    int32_t clr[31];
};

// updateGaussian16, src/BilateralFilter.cu:12-47 (host).
void bilateral_tables(int r, double sd, BilateralTables* t) {
    float g[2 * kMaxRadius + 1];
    for (int i = 0; i < 2 * r + 1; ++i) {
        const float x = (float)(i - r);
        g[i] = expf((float)(-(x * x) / (2 * sd * sd)));
    }
    int idx = 0;
    for (int dy = -r; dy <= r; ++dy)
        for (int dx = -r; dx <= r; ++dx) t->g2d[idx++] = g[dy + r] * g[dx + r];
    float fc[31];
    for (int i = 0; i <= 30; ++i) fc[i] = (float)exp(-(pow(i / 10.0, 2)) / 2);
    for (int i = 0; i <= 30; ++i) t->clr[i] = (int32_t)(fc[i] / fc[30]);
}

__device__ __forceinline__ float filtered(float v, float thr, float fill) {
    return (v > thr || v < 1.e-4f || v != v) ? fill : v;
}

__global__ __launch_bounds__(256) void k_filter_depth(float* __restrict__ d, uint64_t n, float thr, float fill) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = filtered(d[i], thr, fill);
}

__global__ __launch_bounds__(256) void k_mask_depth(float* __restrict__ d, const uint8_t* __restrict__ mask,
                                                    uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !mask[i]) d[i] = __builtin_inff();
}

// Bilateral filter with clamp-to-edge neighbourhoods.  With pre_filter the
// input passes filterDepth on load (the fused ingest).
template <bool PRE>
__global__ __launch_bounds__(kTileX* kTileY) void k_bilateral(float* __restrict__ out, const float* __restrict__ in,
                                                              int H, int W, int r, float norm, float thr,
                                                              float pre_thr, float pre_fill, BilateralTables tab) {
    __shared__ float tile[(kTileX + 2 * kMaxRadius) * (kTileY + 2 * kMaxRadius)];
    const int sw = kTileX + 2 * r, sh = kTileY + 2 * r;
    const int x0 = blockIdx.x * kTileX - r, y0 = blockIdx.y * kTileY - r;
    for (int e = threadIdx.y * kTileX + threadIdx.x; e < sw * sh; e += kTileX * kTileY) {
        int gx = x0 + e % sw, gy = y0 + e / sw;
        gx = gx < 0 ? 0 : (gx >= W ? W - 1 : gx);
        gy = gy < 0 ? 0 : (gy >= H ? H - 1 : gy);
        const float v = in[(uint64_t)gy * W + gx];
        tile[e] = PRE ? filtered(v, pre_thr, pre_fill) : v;
    }
    __syncthreads();
    const int x = blockIdx.x * kTileX + threadIdx.x, y = blockIdx.y * kTileY + threadIdx.y;
    if (x >= W || y >= H) return;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const float c = tile[(ty + r) * sw + tx + r];
    if (__builtin_isinf(c)) {
        out[(uint64_t)y * W + x] = __builtin_inff();
        return;
    }
    const int kd = 2 * r + 1;
    float acc = 0.0f, wsum = 0.0f;
    int ig = 0;
    for (int a = 0; a < kd; ++a) {
        for (int b = 0; b < kd; ++b, ++ig) {
            const float p = tile[(ty + b) * sw + tx + a];
            if (__builtin_isinf(p)) continue;
            const float diff = __builtin_fabsf(c - p);
            if (diff > thr) continue;
            int q = cvt_i32_rz(__builtin_fabsf(diff * norm));
            if (q > 30 || q < 0) q = 30;
            const float w = tab.g2d[ig] * (float)tab.clr[q];
            acc = acc + p * w;
            wsum = wsum + w;
        }
    }
    out[(uint64_t)y * W + x] = acc / wsum;
}

struct Intr {
    float rfx, rfy, cx, cy;
};

// depthToMesh for one pixel.
__device__ __forceinline__ void back_project(const Intr& k, int x, int y, float d, float v[3]) {
    if (__builtin_isinf(d) || d < 1.e-4f) {
        v[0] = v[1] = v[2] = 0.0f;
        return;
    }
    const float x0 = ((float)x - k.cx) * k.rfx;
    const float y0 = ((float)y - k.cy) * k.rfy;
    v[0] = x0 * d;
    v[1] = y0 * d;
    v[2] = d;
}

__global__ __launch_bounds__(256) void k_depth_to_mesh(float* __restrict__ mesh, const float* __restrict__ depth,
                                                       int W, int H, Intr k) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)W * H) return;
    const int x = (int)(i % W), y = (int)(i / W);
    float v[3];
    back_project(k, x, y, depth[i], v);
    mesh[3 * i] = v[0];
    mesh[3 * i + 1] = v[1];
    mesh[3 * i + 2] = v[2];
}

// depthVerticesNormalsKernel body; returns false (normal 0) for invalid
// pixels.  The angle test acos(n.z) in (a0, a1) is n.z < cos a0 and
// n.z > cos a1 in binary64.
__device__ __forceinline__ void pixel_normal(const float* c, const float* xr, const float* xl, const float* yb,
                                             const float* yt, double c0, double c1, float n[3]) {
    n[0] = n[1] = n[2] = 0.0f;
    if (c[2] < 1.e-4f) return;
    if (xr[2] < 1.e-4f || xl[2] < 1.e-4f || yb[2] < 1.e-4f || yt[2] < 1.e-4f) return;
    float dx[3], dy[3], gx[3], gy[3], q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        dx[a] = xr[a] - xl[a];
        dy[a] = yb[a] - yt[a];
    }
    const float nx = __builtin_sqrtf(dx[0] * dx[0] + dx[1] * dx[1]);
    const float ny = __builtin_sqrtf(dy[0] * dy[0] + dy[1] * dy[1]);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        gx[a] = dx[a] / nx;
        gy[a] = dy[a] / ny;
    }
    q[0] = gy[1] * gx[2] - gy[2] * gx[1];
    q[1] = gy[2] * gx[0] - gy[0] * gx[2];
    q[2] = gy[0] * gx[1] - gy[1] * gx[0];
    const float sq = q[0] * q[0] + (q[1] * q[1] + q[2] * q[2]);
    if (sq > 0.0f) {
        const float len = __builtin_sqrtf(sq);
#pragma unroll
        for (int a = 0; a < 3; ++a) q[a] = q[a] / len;
    }
    const float dot = 0.0f * q[0] + (0.0f * q[1] + 1.0f * q[2]);
    if ((double)dot < c0 && (double)dot > c1) return;
    n[0] = q[0];
    n[1] = q[1];
    n[2] = q[2];
}

__global__ __launch_bounds__(256) void k_depth_normals(float* __restrict__ normals, const float* __restrict__ v, int H,
                                                       int W, double c0, double c1) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t iw = (uint64_t)(W - 2), ih = (uint64_t)(H - 2);
    if (t >= iw * ih) return;
    const int x = (int)(t % iw) + 1, y = (int)(t / iw) + 1;
    const uint64_t i = (uint64_t)y * W + x;
    float n[3];
    pixel_normal(v + 3 * i, v + 3 * (i + 1), v + 3 * (i - 1), v + 3 * (i + W), v + 3 * (i - W), c0, c1, n);
    normals[3 * i] = n[0];
    normals[3 * i + 1] = n[1];
    normals[3 * i + 2] = n[2];
}

__device__ __forceinline__ bool normal_is_zero(const float n[3]) {
    return __builtin_fabsf(n[0]) <= 1.e-4f && __builtin_fabsf(n[1]) <= 1.e-4f && __builtin_fabsf(n[2]) <= 1.e-4f;
}

__global__ __launch_bounds__(256) void k_remove_without_normals(const float* __restrict__ normals,
                                                                float* __restrict__ v, float* __restrict__ depth,
                                                                int H, int W) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t iw = (uint64_t)(W - 2), ih = (uint64_t)(H - 2);
    if (t >= iw * ih) return;
    const int x = (int)(t % iw) + 1, y = (int)(t / iw) + 1;
    const uint64_t i = (uint64_t)y * W + x;
    const float n[3] = {normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]};
    if (normal_is_zero(n)) {
        v[3 * i] = v[3 * i + 1] = v[3 * i + 2] = 0.0f;
        depth[i] = __builtin_inff();
    }
}

// One 256-point block of the centroid sum: binary tree in binary64,
// a[t] += a[t + s] for s = 128 .. 1; plus the count of points with z > 1e-4.
__device__ __forceinline__ void centroid_block(const float p[3], bool in, double (*s_sum)[3], uint32_t* s_cnt,
                                               double* block_sums, uint32_t* block_counts) {
    const int t = threadIdx.x;
    s_sum[t][0] = in ? (double)p[0] : 0.0;
    s_sum[t][1] = in ? (double)p[1] : 0.0;
    s_sum[t][2] = in ? (double)p[2] : 0.0;
    if (t == 0) *s_cnt = 0;
    __syncthreads();
    if (in && p[2] > 1.e-4f) atomicAdd(s_cnt, 1u);
    for (int s = 128; s >= 1; s >>= 1) {
        if (t < s) {
            s_sum[t][0] = s_sum[t][0] + s_sum[t + s][0];
            s_sum[t][1] = s_sum[t][1] + s_sum[t + s][1];
            s_sum[t][2] = s_sum[t][2] + s_sum[t + s][2];
        }
        __syncthreads();
    }
    if (t == 0) {
        block_sums[3 * blockIdx.x] = s_sum[0][0];
        block_sums[3 * blockIdx.x + 1] = s_sum[0][1];
        block_sums[3 * blockIdx.x + 2] = s_sum[0][2];
        block_counts[blockIdx.x] = *s_cnt;
    }
}

__global__ __launch_bounds__(256) void k_centroid_blocks(const float* __restrict__ mesh, uint64_t n,
                                                         double* __restrict__ block_sums,
                                                         uint32_t* __restrict__ block_counts) {
    __shared__ double s_sum[256][3];
    __shared__ uint32_t s_cnt;
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    float p[3] = {0.0f, 0.0f, 0.0f};
    if (i < n) {
        p[0] = mesh[3 * i];
        p[1] = mesh[3 * i + 1];
        p[2] = mesh[3 * i + 2];
    }
    centroid_block(p, i < n, s_sum, &s_cnt, block_sums, block_counts);
}

// The frame total from the block sums in a fixed order: 1024 chunks of
// ceil(nb / 1024) consecutive blocks, each summed in order from 0, then a
// binary tree over the chunks; centroid = float(total) / float(count).
constexpr int kFinalThreads = 1024;

__global__ __launch_bounds__(kFinalThreads) void k_centroid_final(const double* __restrict__ block_sums,
                                                                  const uint32_t* __restrict__ block_counts,
                                                                  uint32_t nb, float* __restrict__ out) {
    __shared__ double s_part[3][kFinalThreads];
    __shared__ unsigned long long s_cnt;
    const uint32_t t = threadIdx.x;
    if (t == 0) s_cnt = 0ull;
    const uint32_t c = (nb + kFinalThreads - 1) / kFinalThreads;
    double q0 = 0.0, q1 = 0.0, q2 = 0.0;
    uint64_t cnt = 0;
    for (uint32_t b = t * c; b < (t + 1) * c && b < nb; ++b) {
        q0 = q0 + block_sums[3 * b];
        q1 = q1 + block_sums[3 * b + 1];
        q2 = q2 + block_sums[3 * b + 2];
        cnt += block_counts[b];
    }
    s_part[0][t] = q0;
    s_part[1][t] = q1;
    s_part[2][t] = q2;
    __syncthreads();
    if (cnt) atomicAdd(&s_cnt, (unsigned long long)cnt);
    for (uint32_t s = kFinalThreads / 2; s >= 1; s >>= 1) {
        if (t < s) {
#pragma unroll
            for (int a = 0; a < 3; ++a) s_part[a][t] = s_part[a][t] + s_part[a][t + s];
        }
        __syncthreads();
    }
    if (t == 0) {
        const uint64_t total = s_cnt;
        out[0] = (float)s_part[0][0] / (float)total;
        out[1] = (float)s_part[1][0] / (float)total;
        out[2] = (float)s_part[2][0] / (float)total;
        out[3] = __uint_as_float((uint32_t)total);  // count, bit pattern (frames are < 2^32 pixels)
    }
}

// Fused back-projection, normal, removal and centroid block sums: each pixel
// back-projects its four neighbours itself.
__global__ __launch_bounds__(256) void k_ingest_points(const float* __restrict__ depth_f, int W, int H, Intr k,
                                                       double c0, double c1, float* __restrict__ verts,
                                                       float* __restrict__ normals, float* __restrict__ depth_out,
                                                       double* __restrict__ block_sums,
                                                       uint32_t* __restrict__ block_counts) {
    __shared__ double s_sum[256][3];
    __shared__ uint32_t s_cnt;
    const uint64_t n = (uint64_t)W * H;
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    float v[3] = {0.0f, 0.0f, 0.0f};
    if (i < n) {
        const int x = (int)(i % W), y = (int)(i / W);
        float d = depth_f[i];
        back_project(k, x, y, d, v);
        if (x >= 1 && x < W - 1 && y >= 1 && y < H - 1) {
            float xr[3], xl[3], yb[3], yt[3], nrm[3];
            back_project(k, x + 1, y, depth_f[i + 1], xr);
            back_project(k, x - 1, y, depth_f[i - 1], xl);
            back_project(k, x, y + 1, depth_f[i + W], yb);
            back_project(k, x, y - 1, depth_f[i - W], yt);
            pixel_normal(v, xr, xl, yb, yt, c0, c1, nrm);
            normals[3 * i] = nrm[0];
            normals[3 * i + 1] = nrm[1];
            normals[3 * i + 2] = nrm[2];
            if (normal_is_zero(nrm)) {
                v[0] = v[1] = v[2] = 0.0f;
                d = __builtin_inff();
            }
        }
        verts[3 * i] = v[0];
        verts[3 * i + 1] = v[1];
        verts[3 * i + 2] = v[2];
        depth_out[i] = d;
    }
    centroid_block(v, i < n, s_sum, &s_cnt, block_sums, block_counts);
}

inline unsigned g1(uint64_t n) { return (unsigned)((n + 255) / 256); }

int check_frame(int W, int H) {
    if (W < 0 || H < 0 || (uint64_t)W * (uint64_t)H > 0xffffffffull)
        return fail(VF_ERR_INVALID_ARGUMENT, "frame %dx%d: supported up to 2^32 pixels", W, H);
    return VF_OK;
}

Intr make_intr(const vf_mat3& K) {
    Intr k;
    k.rfx = 1 / K.m[0];
    k.rfy = 1 / K.m[4];
    k.cx = K.m[6];
    k.cy = K.m[7];
    return k;
}

}  // namespace
}  // namespace vf

using namespace vf;

extern "C" {

int vf_filter_depth(float* depth, uint32_t depth_size, float threshold, float fill_value, vf_stream stream) {
    if (!depth_size) return VF_OK;
    if (!depth) return fail(VF_ERR_INVALID_ARGUMENT, "vf_filter_depth: NULL depth");
    hipLaunchKernelGGL(k_filter_depth, dim3(g1(depth_size)), dim3(256), 0, (hipStream_t)stream, depth,
                       (uint64_t)depth_size, threshold, fill_value);
    VF_LAUNCH_CHECK("k_filter_depth");
    return VF_OK;
}

int vf_mask_depth(float* depth, const uint8_t* mask, uint32_t depth_size, vf_stream stream) {
    if (!depth_size) return VF_OK;
    if (!depth || !mask) return fail(VF_ERR_INVALID_ARGUMENT, "vf_mask_depth: NULL buffer");
    hipLaunchKernelGGL(k_mask_depth, dim3(g1(depth_size)), dim3(256), 0, (hipStream_t)stream, depth, mask,
                       (uint64_t)depth_size);
    VF_LAUNCH_CHECK("k_mask_depth");
    return VF_OK;
}

int vf_bilateral_filter(float* output, const float* input, int height, int width, int radius, double sigma_i,
                        double sigma_s, float threshold, vf_stream stream) {
    int st = check_frame(width, height);
    if (st) return st;
    if (radius < 0 || radius > kMaxRadius)
        return fail(VF_ERR_INVALID_ARGUMENT, "bilateral radius %d: supported 0..%d", radius, kMaxRadius);
    if (!width || !height) return VF_OK;
    if (!output || !input || output == input)
        return fail(VF_ERR_INVALID_ARGUMENT, "vf_bilateral_filter: NULL or aliased buffers");
    BilateralTables tab;
    memset(&tab, 0, sizeof(tab));
    bilateral_tables(radius, sigma_s, &tab);
    const float norm = (float)(30 / (3 * sigma_i));
    dim3 grid((width + kTileX - 1) / kTileX, (height + kTileY - 1) / kTileY);
    hipLaunchKernelGGL(k_bilateral<false>, grid, dim3(kTileX, kTileY), 0, (hipStream_t)stream, output, input, height,
                       width, radius, norm, threshold, 0.0f, 0.0f, tab);
    VF_LAUNCH_CHECK("k_bilateral");
    return VF_OK;
}

int vf_depth_to_mesh(vf_float3* mesh, const float* depth, uint32_t width, uint32_t height, vf_mat3 intrinsics,
                     vf_stream stream) {
    int st = check_frame((int)std::min<uint32_t>(width, 0x7fffffff), (int)std::min<uint32_t>(height, 0x7fffffff));
    if (st) return st;
    const uint64_t n = (uint64_t)width * height;
    if (!n) return VF_OK;
    if (!mesh || !depth) return fail(VF_ERR_INVALID_ARGUMENT, "vf_depth_to_mesh: NULL buffer");
    hipLaunchKernelGGL(k_depth_to_mesh, dim3(g1(n)), dim3(256), 0, (hipStream_t)stream, (float*)mesh, depth,
                       (int)width, (int)height, make_intr(intrinsics));
    VF_LAUNCH_CHECK("k_depth_to_mesh");
    return VF_OK;
}

int vf_depth_normals(vf_float3* normals, vf_float3* vertices, const float* depth, int height, int width,
                     float angle_from, float angle_to, vf_stream stream) {
    (void)depth;  // unused by the reference kernel as well
    int st = check_frame(width, height);
    if (st) return st;
    if (width < 3 || height < 3) return VF_OK;
    if (!normals || !vertices) return fail(VF_ERR_INVALID_ARGUMENT, "vf_depth_normals: NULL buffer");
    const uint64_t n = (uint64_t)(width - 2) * (height - 2);
    hipLaunchKernelGGL(k_depth_normals, dim3(g1(n)), dim3(256), 0, (hipStream_t)stream, (float*)normals,
                       (const float*)vertices, height, width, cos((double)angle_from), cos((double)angle_to));
    VF_LAUNCH_CHECK("k_depth_normals");
    return VF_OK;
}

int vf_remove_depth_without_normals(vf_float3* normals, vf_float3* vertices, float* depth, int height, int width,
                                    vf_stream stream) {
    int st = check_frame(width, height);
    if (st) return st;
    if (width < 3 || height < 3) return VF_OK;
    if (!normals || !vertices || !depth) return fail(VF_ERR_INVALID_ARGUMENT, "vf_remove_depth_without_normals: NULL");
    const uint64_t n = (uint64_t)(width - 2) * (height - 2);
    hipLaunchKernelGGL(k_remove_without_normals, dim3(g1(n)), dim3(256), 0, (hipStream_t)stream,
                       (const float*)normals, (float*)vertices, depth, height, width);
    VF_LAUNCH_CHECK("k_remove_without_normals");
    return VF_OK;
}

int vf_depth_centroid(vf_workspace* ws, vf_float3* centroid, uint64_t* count, const vf_float3* mesh, size_t n,
                      vf_stream stream) {
    if (!centroid || !count || (n && !mesh)) return fail(VF_ERR_INVALID_ARGUMENT, "vf_depth_centroid: NULL argument");
    if (n > 0xffffffffull) return fail(VF_ERR_INVALID_ARGUMENT, "vf_depth_centroid: more than 2^32 points");
    const uint32_t nb = (uint32_t)std::max<uint64_t>(1, (n + 255) / 256);
    void* scratch = nullptr;
    int st = workspace_scratch(ws, (uint64_t)nb * 28 + 64, &scratch, (hipStream_t)stream);
    if (st) return st;
    double* sums = (double*)scratch;
    uint32_t* counts = (uint32_t*)(sums + 3ull * nb);
    float* out = (float*)(counts + nb + (nb & 1));
    hipLaunchKernelGGL(k_centroid_blocks, dim3(nb), dim3(256), 0, (hipStream_t)stream, (const float*)mesh,
                       (uint64_t)n, sums, counts);
    VF_LAUNCH_CHECK("k_centroid_blocks");
    hipLaunchKernelGGL(k_centroid_final, dim3(1), dim3(kFinalThreads), 0, (hipStream_t)stream, sums, counts, nb,
                       out);
    VF_LAUNCH_CHECK("k_centroid_final");
    float h[4];
    VF_HIP(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, (hipStream_t)stream));
    VF_HIP(hipStreamSynchronize((hipStream_t)stream));
    *centroid = vf_float3{h[0], h[1], h[2]};
    *count = __builtin_bit_cast(uint32_t, h[3]);
    return VF_OK;
}

int vf_ingest_depth(vf_workspace* ws, float* depth, int width, int height, vf_mat3 intrinsics, float threshold,
                    int bilateral_radius, double sigma_i, double sigma_s, float bilateral_threshold,
                    float angle_from, float angle_to, vf_float3* vertices, vf_float3* normals, vf_float3* centroid,
                    uint64_t* count, vf_stream stream) {
    int st = check_frame(width, height);
    if (st) return st;
    if (bilateral_radius < 0 || bilateral_radius > kMaxRadius)
        return fail(VF_ERR_INVALID_ARGUMENT, "bilateral radius %d: supported 0..%d", bilateral_radius, kMaxRadius);
    if (!depth || !vertices || !normals || !centroid || !count)
        return fail(VF_ERR_INVALID_ARGUMENT, "vf_ingest_depth: NULL argument");
    const uint64_t n = (uint64_t)width * height;
    if (!n) return VF_OK;
    hipStream_t s = (hipStream_t)stream;
    const uint32_t nb = (uint32_t)((n + 255) / 256);
    const uint64_t depth_bytes = (n * 4 + 255) & ~255ull;
    void* scratch = nullptr;
    st = workspace_scratch(ws, depth_bytes + (uint64_t)nb * 28 + 64, &scratch, s);
    if (st) return st;
    float* depth_f = (float*)scratch;
    double* sums = (double*)((char*)scratch + depth_bytes);
    uint32_t* counts = (uint32_t*)(sums + 3ull * nb);
    float* out = (float*)(counts + nb + (nb & 1));
    // 1. filter (and the bilateral filter of the filtered frame)
    if (bilateral_radius > 0) {
        BilateralTables tab;
        memset(&tab, 0, sizeof(tab));
        bilateral_tables(bilateral_radius, sigma_s, &tab);
        const float norm = (float)(30 / (3 * sigma_i));
        dim3 grid((width + kTileX - 1) / kTileX, (height + kTileY - 1) / kTileY);
        hipLaunchKernelGGL(k_bilateral<true>, grid, dim3(kTileX, kTileY), 0, s, depth_f, (const float*)depth, height,
                           width, bilateral_radius, norm, bilateral_threshold, threshold, __builtin_inff(), tab);
        VF_LAUNCH_CHECK("k_bilateral");
    } else {
        VF_HIP(hipMemcpyAsync(depth_f, depth, n * 4, hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_filter_depth, dim3(g1(n)), dim3(256), 0, s, depth_f, n, threshold, __builtin_inff());
        VF_LAUNCH_CHECK("k_filter_depth");
    }
    // 2. back-projection, normals, removal, centroid block sums
    hipLaunchKernelGGL(k_ingest_points, dim3(nb), dim3(256), 0, s, (const float*)depth_f, width, height,
                       make_intr(intrinsics), cos((double)angle_from), cos((double)angle_to), (float*)vertices,
                       (float*)normals, depth, sums, counts);
    VF_LAUNCH_CHECK("k_ingest_points");
    // 3. the centroid
    hipLaunchKernelGGL(k_centroid_final, dim3(1), dim3(kFinalThreads), 0, s, sums, counts, nb, out);
    VF_LAUNCH_CHECK("k_centroid_final");
    float h[4];
    VF_HIP(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, s));
    VF_HIP(hipStreamSynchronize(s));
    *centroid = vf_float3{h[0], h[1], h[2]};
    *count = __builtin_bit_cast(uint32_t, h[3]);
    return VF_OK;
}

}  // extern "C"
