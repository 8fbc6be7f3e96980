// Checks of the af:: drop-in header (include/af/vf_dropin.h) from C++.
//
//   dropin_check host   argument errors surface as std::runtime_error (no GPU)
//   dropin_check gpu    two frames of the reference's per-frame sequence
//                       (src/VolumeFusion.cu:356-359, 702-742) called through
//                       the af:: functions one by one, against the fused
//                       af::runWarpedTsdfIntegration on a copy of the same
//                       state; tsdf, weights and node positions must agree
//                       bit for bit.
// Exit status 0 on success; the failing check is printed otherwise.
#include <af/vf_dropin.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

namespace {

using Vec3f = Eigen::Vector3f;
using Vec3i = Eigen::Vector3i;
using Vec2i = Eigen::Vector2i;
using Mat3f = Eigen::Matrix3f;
using Mat4f = Eigen::Matrix4f;
constexpr int K = 6;  // Constants::motionGraphKNN of the reference
using Vecui = Eigen::Matrix<unsigned int, K, 1>;
using Vecf = Eigen::Matrix<float, K, 1>;

int failures = 0;

void expect(bool ok, const char* what) {
    if (!ok) {
        std::printf("FAIL: %s\n", what);
        ++failures;
    }
}

template <class F>
void expect_throw(F&& f, const char* needle, const char* what) {
    try {
        f();
    } catch (const std::runtime_error& e) {
        if (std::strstr(e.what(), needle)) return;
        std::printf("FAIL: %s: wrong message '%s'\n", what, e.what());
        ++failures;
        return;
    }
    std::printf("FAIL: %s: no exception\n", what);
    ++failures;
}

int host_checks() {
    // every function of the header, instantiated at the reference's K
    void (*knn)(Vecui*, Vecf*, int*, const Vec3f*, const float*, std::size_t, Vec3f, Vec3f, Vec3i, Vec3i) =
        &af::runTsdfKNNGraph<K>;
    void (*wts)(Vecf*, const float*, const Vecui*, const Vecf*, const int*, std::size_t) =
        &af::runComputeTsdfWarpWeightsKernel<K>;
    void (*warp)(Vec3f*, const Mat4f*, const Vecui*, const Vecf*, const int*, Vec3f, Vec3f, Vec3i) =
        &af::runWarpTsdfKernel<K>;
    void (*mesh_w)(Vecf*, const float*, const Vecui*, const Vecf*, std::size_t) = &af::runComputeWarpWeightsKernel<K>;
    void (*mesh_warp)(Vec3f*, const Vec3f*, const Mat4f*, const Vecui*, const Vecf*, std::size_t) =
        &af::runWarpKernel<K>;
    (void)wts;
    (void)warp;
    (void)mesh_w;
    (void)mesh_warp;
    const Vec3f vs(0.01f, 0.01f, 0.01f), size(0.1f, 0.1f, 0.1f);
    const Vec3i dim(10, 10, 10);
    expect_throw([&] { af::runTranslateKernel(nullptr, 5, Vec3f(1, 2, 3)); }, "mesh is NULL",
                 "runTranslateKernel(nullptr)");
    expect_throw([&] { knn(nullptr, nullptr, nullptr, nullptr, nullptr, 4, vs, size, dim, Vec3i(3, 3, 3)); },
                 "runTsdfKNNGraph", "runTsdfKNNGraph(nullptr)");
    expect_throw(
        [&] {
            af::runIntegrateTSDFFromFrameKernel(nullptr, nullptr, nullptr, Vec2i(4, 4), 0.02f, dim, vs, size,
                                                Mat3f::Identity(), Vec3f::Zero(), Mat3f::Identity());
        },
        "runIntegrateTSDFFromFrameKernel", "runIntegrateTSDFFromFrameKernel(nullptr)");
    expect_throw(
        [&] {
            af::runWarpedTsdfIntegration<K>(nullptr, nullptr, nullptr, dim, size, Vec3f::Zero(), nullptr, nullptr,
                                            nullptr, 0, nullptr, Vec2i(4, 4), Mat3f::Identity(), 0.02f, 0.02f);
        },
        "runWarpedTsdfIntegration", "runWarpedTsdfIntegration(no workspace)");
    return failures;
}

// ---- gpu: a small deterministic scene --------------------------------------

struct Lcg {
    uint64_t s;
    float uniform() {  // [0, 1)
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return static_cast<float>((s >> 40) & 0xffffff) / 16777216.0f;
    }
};

template <class T>
T* dev_alloc(std::size_t n) {
    void* p = nullptr;
    af::vf_detail::check(vf_malloc(&p, n * sizeof(T)), "vf_malloc");
    return static_cast<T*>(p);
}

template <class T>
T* dev_upload(const std::vector<T>& h) {
    T* d = dev_alloc<T>(h.size());
    af::vf_detail::check(vf_memcpy_h2d(d, h.data(), h.size() * sizeof(T), nullptr), "vf_memcpy_h2d");
    return d;
}

template <class T>
std::vector<T> dev_download(const T* d, std::size_t n) {
    std::vector<T> h(n);
    af::vf_detail::check(vf_memcpy_d2h(h.data(), d, n * sizeof(T), nullptr), "vf_memcpy_d2h");
    af::vf_detail::check(vf_device_synchronize(), "sync");
    return h;
}

bool same_bits(const void* a, const void* b, std::size_t bytes) { return std::memcmp(a, b, bytes) == 0; }

int gpu_checks() {
    int n_dev = 0;
    af::vf_detail::check(vf_device_count(&n_dev), "vf_device_count");
    if (n_dev < 1) {
        std::printf("FAIL: no device\n");
        return 1;
    }
    const Vec3i dim(48, 40, 36);
    const Vec3f size(0.24f, 0.20f, 0.18f);
    const Vec3f vs = size.cwiseQuotient(dim.cast<float>());
    const Vec3f centroid(0.01f, -0.02f, 0.60f);
    const std::size_t N = static_cast<std::size_t>(dim.prod());
    const int W = 64, H = 48;
    Mat3f intr;
    intr << 90.f, 0.f, 31.5f, 0.f, 90.f, 23.5f, 0.f, 0.f, 1.f;
    const float delta = 0.02f, graphRadius = 0.025f;

    // nodes scattered in the volume, radii around the graph radius
    const std::size_t G = 48;
    Lcg rng{12345};
    std::vector<Vec3f> nodes(G);
    std::vector<float> radii(G);
    for (std::size_t j = 0; j < G; ++j) {
        for (int a = 0; a < 3; ++a) nodes[j][a] = centroid[a] + (rng.uniform() - 0.5f) * size[a] * 0.9f;
        radii[j] = graphRadius * (0.8f + 0.4f * rng.uniform());
    }
    // a wavy surface in front of the camera
    std::vector<float> depth(static_cast<std::size_t>(W) * H);
    for (int v = 0; v < H; ++v)
        for (int u = 0; u < W; ++u)
            depth[static_cast<std::size_t>(v) * W + u] = 0.58f + 0.015f * std::sin(0.21f * u) * std::cos(0.17f * v);
    auto transforms_for = [&](int frame) {
        std::vector<Mat4f> T(G);
        for (std::size_t j = 0; j < G; ++j) {
            const float ang = 0.02f * (frame + 1) * (rng.uniform() - 0.5f);
            Mat4f m = Mat4f::Identity();
            m(0, 0) = std::cos(ang);
            m(0, 1) = -std::sin(ang);
            m(1, 0) = std::sin(ang);
            m(1, 1) = std::cos(ang);
            for (int a = 0; a < 3; ++a) m(a, 3) = 0.004f * (rng.uniform() - 0.5f);
            T[j] = m;
        }
        return T;
    };

    std::vector<float> tsdf0(N, -1.0f), w0(N, 0.0f);
    float* tsdfA = dev_upload(tsdf0);
    float* wA = dev_upload(w0);
    float* tsdfB = dev_upload(tsdf0);
    float* wB = dev_upload(w0);
    Vec3f* nodesA = dev_upload(nodes);
    Vec3f* nodesB = dev_upload(nodes);
    float* radii_d = dev_upload(radii);
    float* depth_d = dev_upload(depth);
    Vecui* knnIdxs = dev_alloc<Vecui>(N);
    Vecf* knnDists = dev_alloc<Vecf>(N);
    Vecf* knnWeights = dev_alloc<Vecf>(N);
    int* knnCounts = dev_alloc<int>(N);
    int* collisions = dev_alloc<int>(N);
    Vec3f* vertices = dev_alloc<Vec3f>(N);
    vf_workspace* ws = nullptr;
    af::vf_detail::check(vf_workspace_create(&ws, 1, N, G, nullptr), "vf_workspace_create");

    Vec3f grapMinRadius3d;
    grapMinRadius3d.fill(graphRadius);
    const Vec3i perNode = 2 * (grapMinRadius3d.cwiseQuotient(vs).cast<int>() + Vec3i::Ones()) + Vec3i::Ones();

    uint64_t updated = 0;
    for (int frame = 0; frame < 2; ++frame) {
        std::vector<Mat4f> T = transforms_for(frame);
        Mat4f* T_d = dev_upload(T);
        // the reference's sequence, call by call
        af::vf_detail::check(vf_memset(knnCounts, 0, N * sizeof(int), nullptr), "zero counts");
        af::vf_detail::check(vf_memset(collisions, 0, N * sizeof(int), nullptr), "zero collisions");
        af::runTranslateKernel(nodesA, G, -centroid);
        af::runTsdfKNNGraph<K>(knnIdxs, knnDists, knnCounts, nodesA, radii_d, G, vs, size, dim, perNode);
        af::runComputeTsdfWarpWeightsKernel<K>(knnWeights, radii_d, knnIdxs, knnDists, knnCounts, N);
        af::runGenerateTsdfVerticesKernel(vertices, vs, size, dim);
        af::runTranslateKernel(nodesA, G, centroid);
        af::runTranslateKernel(vertices, static_cast<unsigned>(N), centroid);
        af::runWarpTsdfKernel<K>(vertices, T_d, knnIdxs, knnWeights, knnCounts, vs, size, dim);
        af::runCountCollisionsPerVoxelKernel(collisions, vertices, knnCounts, vs, size, dim, Mat3f::Identity(),
                                             centroid);
        af::runIntegrateWarpedTSDFFromFrameKernel(tsdfA, wA, vertices, collisions, knnCounts, depth_d, Vec2i(W, H),
                                                  delta, vs, size, dim, Mat3f::Identity(), centroid, intr);
        // the fused call
        af::runWarpedTsdfIntegration<K>(ws, tsdfB, wB, dim, size, centroid, nodesB, T_d, radii_d, G, depth_d,
                                        Vec2i(W, H), intr, delta, graphRadius);
        vf_frame_stats st{};
        af::vf_detail::check(vf_workspace_last_stats(ws, &st, nullptr), "last_stats");
        updated += st.updated;
        af::vf_detail::check(vf_free(T_d), "vf_free");
    }

    const auto ta = dev_download(tsdfA, N), tb = dev_download(tsdfB, N);
    const auto wa = dev_download(wA, N), wb = dev_download(wB, N);
    const auto na = dev_download(nodesA, G), nb = dev_download(nodesB, G);
    expect(same_bits(ta.data(), tb.data(), N * 4), "tsdf: af:: sequence == fused");
    expect(same_bits(wa.data(), wb.data(), N * 4), "weights: af:: sequence == fused");
    expect(same_bits(na.data(), nb.data(), G * 12), "node positions: af:: sequence == fused");
    std::size_t changed = 0;
    for (std::size_t i = 0; i < N; ++i) changed += wa[i] != 0.0f;
    expect(changed > 1000, "scene integrates a meaningful number of voxels");
    std::printf("voxels integrated: %zu, fused updates over 2 frames: %llu\n", changed,
                static_cast<unsigned long long>(updated));

    for (void* p : std::initializer_list<void*>{tsdfA, wA, tsdfB, wB, nodesA, nodesB, radii_d, depth_d, knnIdxs,
                                                knnDists, knnWeights, knnCounts, collisions, vertices})
        vf_free(p);
    vf_workspace_destroy(ws);
    return failures;
}

}  // namespace

int main(int argc, char** argv) {
    const char* mode = argc > 1 ? argv[1] : "host";
    int rc = 0;
    try {
        rc = std::strcmp(mode, "gpu") == 0 ? gpu_checks() : host_checks();
    } catch (const std::exception& e) {
        std::printf("FAIL: exception %s\n", e.what());
        rc = 1;
    }
    std::printf("%s: %s\n", mode, rc ? "FAILED" : "ok");
    return rc ? 1 : 0;
}
