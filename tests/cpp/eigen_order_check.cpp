// eigen_order_check -- the oracle's evaluation orders against the reference's
// own Eigen expressions, compiled here on the host with the vendored Eigen.
//
// The reference's kernels are Eigen expressions evaluated in device code,
// where Eigen takes its scalar (unvectorised) path.  Built with
// -DEIGEN_DONT_VECTORIZE and -ffp-contract=off, the same expressions on the
// host follow that path with every operation rounded on its own, which is the
// order the oracle (oracle/vf_oracle.c) and the kernels restate.  Each check
// below evaluates a reference expression with Eigen types on inputs chosen so
// that the plausible orders differ, compares the oracle's function bit for
// bit, and shows that a wrong variant (another summation order, a transposed
// rotation) disagrees with the oracle somewhere, so the check can fail.
//
//   warp       (M * p.homogeneous()).head(3), weighted sum over the slots
//              (src/TSDFWarp.cu:185-187)            vs vfo_warp_tsdf
//   mesh       (M * p4).head(3) with a materialised Vec4f
//              (src/VertexManipulation.cu:138-141)  vs vfo_warp_mesh
//   distance   (tsdfPoint - node).norm() (src/TSDFWarp.cu:53-54)
//                                                   vs vfo_knn_graph
//   weights    running sum, weights /= sum (src/TSDFWarp.cu:101-113)
//                                                   vs vfo_warp_weights
//   iszero     weights.isZero() (src/TSDFWarp.cu:179-181)
//                                                   vs vfo_warp_tsdf's skip
//   inverse    Mat3f::inverse() of rotations and general matrices
//              (src/TSDFIntegration.cu:247, 296)    vs vfo_mat3_inverse
//   plain      integrateTSDFFromFrameKernel + updateTsdf with a rotated
//              volume (src/TSDFIntegration.cu:16-96) vs vfo_integrate_plain
//   collide    Rinv * (p - t) and worldToVoxel with a rotated volume
//              (src/TSDFIntegration.cu:252-282)     vs vfo_count_collisions
//   gn_*       the Gauss-Newton rows: (n w)^T gradTv wD, wD n.dot(diff),
//              ((T_g p4) - (T_nb p4)).head(3) * wR
//              (src/MotionOptimization.cu:29-225)   vs vfo_gn_j_depth,
//                                                     vfo_gn_r_depth, vfo_gn_r_mreg
//
// Exit status 0 when every check passes; each prints one line.
#include <Eigen/Dense>
#include <Eigen/Geometry>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "vf_oracle.h"

namespace {

using Vec3f = Eigen::Vector3f;
using Vec4f = Eigen::Vector4f;
using Vec3i = Eigen::Vector3i;
using Mat3f = Eigen::Matrix3f;
using Mat4f = Eigen::Matrix4f;
constexpr int K = 8;
using VecKf = Eigen::Matrix<float, K, 1>;

int failures = 0;

bool same(float a, float b) {
    uint32_t x, y;
    std::memcpy(&x, &a, 4);
    std::memcpy(&y, &b, 4);
    return x == y;
}

void report(const char* name, long samples, long mismatches, long wrong_differs) {
    const bool ok = mismatches == 0 && wrong_differs > 0;
    std::printf("%s %s: %ld samples, %ld mismatches against the oracle, wrong variant differs on %ld\n",
                ok ? "ok  " : "FAIL", name, samples, mismatches, wrong_differs);
    if (!ok) ++failures;
}

// Floats with random exponents and mantissas, so that rounding depends on
// the order in which terms meet.
struct Rng {
    std::mt19937_64 g{20250101};
    float uni(float lo, float hi) { return std::uniform_real_distribution<float>(lo, hi)(g); }
    float wide() {
        const float m = uni(1.0f, 2.0f) * (uni(0.0f, 1.0f) < 0.5f ? -1.0f : 1.0f);
        return std::ldexp(m, (int)std::uniform_int_distribution<int>(-6, 6)(g));
    }
    Mat3f rotation() {
        Eigen::Quaternionf q(uni(-1, 1), uni(-1, 1), uni(-1, 1), uni(-1, 1));
        q.normalize();
        return q.toRotationMatrix();
    }
};

// The voxel warp of src/TSDFWarp.cu:183-189 for one vertex.
Vec3f ref_warp(const std::vector<Mat4f>& M, const VecKf& w, int n, const Vec3f& p) {
    Vec3f acc = Vec3f::Zero();
    for (int j = 0; j < n; ++j) acc += w[j] * (M[j] * p.homogeneous()).head(3);
    return acc;
}

void check_warp(Rng& r) {
    long samples = 0, bad = 0, differs = 0;
    for (int t = 0; t < 20000; ++t) {
        const int n = 1 + t % K;
        std::vector<Mat4f> M(n);
        std::vector<float> T(16 * n);
        for (int j = 0; j < n; ++j) {
            for (int e = 0; e < 16; ++e) M[j].data()[e] = r.wide();
            M[j].row(3) << 0, 0, 0, 1;
            std::memcpy(&T[16 * j], M[j].data(), 64);
        }
        VecKf w = VecKf::Zero();
        for (int j = 0; j < n; ++j) w[j] = r.uni(0.05f, 1.0f);
        const Vec3f p(r.wide(), r.wide(), r.wide());
        const Vec3f e = ref_warp(M, w, n, p);
        float v[3] = {p[0], p[1], p[2]};
        std::vector<uint32_t> idx(K);
        for (int j = 0; j < K; ++j) idx[j] = j < n ? j : 0;
        const int32_t cnt = n;
        vfo_warp_tsdf(K, v, T.data(), idx.data(), w.data(), &cnt, 1);
        // wrong variant: each row summed left to right, ((a0 + a1) + a2) + m3
        float wr[3] = {0, 0, 0};
        for (int j = 0; j < n; ++j)
            for (int a = 0; a < 3; ++a) {
                const float s = ((M[j](a, 0) * p[0] + M[j](a, 1) * p[1]) + M[j](a, 2) * p[2]) + M[j](a, 3);
                wr[a] = wr[a] + w[j] * s;
            }
        for (int a = 0; a < 3; ++a) {
            ++samples;
            bad += !same(e[a], v[a]);
            differs += !same(wr[a], v[a]);
        }
    }
    report("warp", samples, bad, differs);
}

void check_mesh(Rng& r) {
    long samples = 0, bad = 0, differs = 0;
    const int G = 16;
    std::vector<Mat4f> M(G);
    std::vector<float> T(16 * G);
    for (int j = 0; j < G; ++j) {
        for (int e = 0; e < 16; ++e) M[j].data()[e] = r.wide();
        std::memcpy(&T[16 * j], M[j].data(), 64);
    }
    for (int t = 0; t < 20000; ++t) {
        std::vector<uint32_t> idx(K);
        VecKf w;
        for (int j = 0; j < K; ++j) {
            idx[j] = (uint32_t)((t * 7 + j * 3) % G);
            w[j] = r.uni(0.05f, 1.0f);
        }
        const Vec3f p(r.wide(), r.wide(), r.wide());
        // src/VertexManipulation.cu:131-141: the homogeneous point is a Vec4f
        Vec3f e = Vec3f::Zero();
        const Vec4f ph = p.homogeneous();
        for (int j = 0; j < K; ++j) e += w[j] * (M[idx[j]] * ph).head(3);
        float out[3];
        vfo_warp_mesh(K, out, p.data(), T.data(), idx.data(), w.data(), 1, G);
        // wrong variant: the voxel kernel's order a0 + (a1 + a2) + a3
        float wr[3] = {0, 0, 0};
        for (int j = 0; j < K; ++j)
            for (int a = 0; a < 3; ++a) {
                const Mat4f& m = M[idx[j]];
                const float s = (m(a, 0) * p[0] + (m(a, 1) * p[1] + m(a, 2) * p[2])) + m(a, 3);
                wr[a] = wr[a] + w[j] * s;
            }
        for (int a = 0; a < 3; ++a) {
            ++samples;
            bad += !same(e[a], out[a]);
            differs += !same(wr[a], out[a]);
        }
    }
    report("mesh", samples, bad, differs);
}

// voxelToWorld, src/TSDFWarp.cu:8-10
Vec3f voxel_to_world(const Vec3i& v, const Vec3f& vs, const Vec3f& size) {
    return v.cast<float>().cwiseProduct(vs) - size * 0.5f;
}

void check_distance(Rng& r) {
    long samples = 0, bad = 0, differs = 0;
    for (int t = 0; t < 12; ++t) {
        const Vec3i dim(20 + t % 3, 18, 16 + t % 4);
        const float s = r.uni(0.0031f, 0.0047f);
        const Vec3f vs(s, s * r.uni(0.9f, 1.1f), s);
        const Vec3f size = dim.cast<float>().cwiseProduct(vs);
        vfo_vol vol;
        for (int a = 0; a < 3; ++a) vol.dim[a] = dim[a], vol.vs[a] = vs[a], vol.size[a] = size[a];
        const Vec3f node(r.uni(-0.02f, 0.02f), r.uni(-0.02f, 0.02f), r.uni(-0.02f, 0.02f));
        const float radius = 1.0f;  // every voxel of the cube is a candidate
        const int32_t cube[3] = {15, 15, 15};
        const size_t N = (size_t)dim.prod();
        std::vector<uint32_t> idx(N);
        std::vector<float> dist(N);
        std::vector<int32_t> cnt(N, 0);
        vfo_knn_graph(1, idx.data(), dist.data(), cnt.data(), node.data(), &radius, 1, &vol, cube);
        for (int z = 0; z < dim[2]; ++z)
            for (int y = 0; y < dim[1]; ++y)
                for (int x = 0; x < dim[0]; ++x) {
                    const size_t v = ((size_t)z * dim[1] + y) * dim[0] + x;
                    if (!cnt[v]) continue;
                    const Vec3f tsdfPoint = voxel_to_world(Vec3i(x, y, z), vs, size);
                    const float d = (tsdfPoint - node).norm();
                    const Vec3f q = tsdfPoint - node;
                    const float dw = std::sqrt((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]);
                    ++samples;
                    bad += !same(d, dist[v]);
                    differs += !same(dw, dist[v]);
                }
    }
    report("distance", samples, bad, differs);
}

void check_weights(Rng& r) {
    long samples = 0, bad = 0, differs = 0;
    const int G = 32;
    std::vector<float> radii(G);
    for (auto& x : radii) x = r.uni(0.01f, 0.03f);
    for (int t = 0; t < 20000; ++t) {
        const int n = 1 + t % K;
        std::vector<uint32_t> idx(K, 0);
        VecKf dist = VecKf::Zero();
        for (int j = 0; j < n; ++j) {
            idx[j] = (uint32_t)((t + 5 * j) % G);
            dist[j] = r.uni(0.0f, 2.0f * radii[idx[j]]);
        }
        // src/TSDFWarp.cu:101-113 (the Gaussian itself is the oracle's)
        VecKf weights;
        weights.fill(0.f);
        float weightsSum = 0;
        for (int j = 0; j < n; j++) {
            weights[j] = vfo_weight_tsdf(radii[idx[j]], dist[j]);
            weightsSum += weights[j];
        }
        VecKf wrong = weights * (1.0f / weightsSum);
        if (weightsSum != 0.f) weights /= weightsSum;
        float w[K];
        const int32_t cnt = n;
        vfo_warp_weights(K, w, radii.data(), idx.data(), dist.data(), &cnt, 1);
        for (int j = 0; j < K; ++j) {
            ++samples;
            bad += !same(weights[j], w[j]);
            differs += !same(wrong[j], w[j]);
        }
    }
    report("weights", samples, bad, differs);
}

void check_iszero() {
    const float vals[] = {0.0f, -0.0f, 1e-5f, -1e-5f, std::nextafter(1e-5f, 1.0f), -std::nextafter(1e-5f, 1.0f),
                          9.99e-6f, 1e-30f, NAN, 0.5f};
    const int nv = (int)(sizeof(vals) / sizeof(vals[0]));
    long samples = 0, bad = 0, differs = 0;
    Mat4f I = Mat4f::Identity();
    I(0, 3) = 1.0f;  // a translation, so a warped vertex moves
    std::vector<float> T(16 * K);
    for (int j = 0; j < K; ++j) std::memcpy(&T[16 * j], I.data(), 64);
    std::vector<uint32_t> idx(K, 0);
    for (int a = 0; a < nv; ++a)
        for (int b = 0; b < nv; ++b)
            for (int c = 0; c < nv; ++c) {
                VecKf w = VecKf::Zero();
                w[0] = vals[a], w[1] = vals[b], w[2] = vals[c];
                const bool zero = w.isZero();
                float v[3] = {0.25f, 0.5f, 0.75f};
                const int32_t cnt = 3;
                vfo_warp_tsdf(K, v, T.data(), idx.data(), w.data(), &cnt, 1);
                const bool skipped = same(v[0], 0.25f) && same(v[1], 0.5f) && same(v[2], 0.75f);
                // wrong variant: a strict bound, |w| < 1e-5
                bool wrong = true;
                for (int j = 0; j < 3; ++j) wrong = wrong && std::fabs(w[j]) < 1e-5f;
                ++samples;
                bad += zero != skipped;
                differs += wrong != skipped;
            }
    report("iszero", samples, bad, differs);
}

void check_inverse(Rng& r) {
    long samples = 0, bad = 0, differs = 0;
    for (int t = 0; t < 20000; ++t) {
        Mat3f m;
        if (t % 2) {
            m = r.rotation();
        } else {
            for (int e = 0; e < 9; ++e) m.data()[e] = r.wide();
        }
        const Mat3f inv = m.inverse();
        float out[9];
        vfo_mat3_inverse(m.data(), out);
        // wrong variant: the transpose of the adjugate's layout, i.e. the
        // inverse of the transposed matrix (equal only for symmetric input)
        const Mat3f wrong = Mat3f(m.transpose()).inverse();
        for (int e = 0; e < 9; ++e) {
            ++samples;
            bad += !same(inv.data()[e], out[e]);
            differs += !same(wrong.data()[e], out[e]);
        }
    }
    report("inverse", samples, bad, differs);
}

// updateTsdf, src/TSDFIntegration.cu:16-67 (weight cap as a parameter; the
// reference's is 15)
void ref_update(std::vector<float>& tsdf, std::vector<float>& wts, const Vec3f& p, size_t i,
                const std::vector<float>& depth, const Eigen::Vector2i& fd, float delta, const Mat3f& Kin) {
    if (p[2] < 1.e-4f) return;
    const float zInv = 1.0f / p[2];
    const float uf = (Kin(0, 0) * p[0] * zInv) + Kin(0, 2);
    const float vf = (Kin(1, 1) * p[1] * zInv) + Kin(1, 2);
    const Eigen::Vector2i px(static_cast<int>(uf + 0.5f), static_cast<int>(vf + 0.5f));
    if (px[0] < 0 || px[0] >= fd[0] || px[1] < 0 || px[1] >= fd[1]) return;
    const float d = depth[(size_t)px[1] * fd[0] + px[0]];
    if (std::isinf(d) || d < 1.e-4f) return;
    const float sdf = p[2] - d;
    if (sdf > delta) return;
    const float tn = sdf < -delta ? -1.f : sdf / delta;
    const float wn = 1.0f;
    tsdf[i] = (tsdf[i] * wts[i] + tn * wn) / (wts[i] + wn);
    wts[i] = std::min(wts[i] + wn, 15.0f);
}

void check_plain(Rng& r) {
    long samples = 0, bad = 0, differs = 0;
    const int W = 160, H = 120;
    Mat3f Kin;
    Kin << 126.0f, 0, 80.0f, 0, 126.0f, 60.0f, 0, 0, 1;
    const Eigen::Vector2i fd(W, H);
    std::vector<float> depth((size_t)W * H);
    for (int v = 0; v < H; ++v)
        for (int u = 0; u < W; ++u)  // a tilted plane, a few invalid pixels
            depth[(size_t)v * W + u] = (u * 7 + v * 3) % 97 == 0 ? 0.0f : 0.55f + 0.0011f * (u - 80) - 0.0007f * (v - 60);
    for (int t = 0; t < 6; ++t) {
        const Vec3i dim(24, 20, 16);
        const Vec3f vs(0.005f, 0.005f, 0.005f);
        const Vec3f size = dim.cast<float>().cwiseProduct(vs);
        vfo_vol vol;
        for (int a = 0; a < 3; ++a) vol.dim[a] = dim[a], vol.vs[a] = vs[a], vol.size[a] = size[a];
        const Mat3f R = r.rotation();
        const Vec3f tr(r.uni(-0.01f, 0.01f), r.uni(-0.01f, 0.01f), 0.55f + r.uni(-0.01f, 0.01f));
        const size_t N = (size_t)dim.prod();
        std::vector<float> te(N, -1.0f), we(N, 0.0f), tw(N, -1.0f), ww(N, 0.0f), to(N, -1.0f), wo(N, 0.0f);
        for (int z = 0; z < dim[2]; ++z)
            for (int y = 0; y < dim[1]; ++y)
                for (int x = 0; x < dim[0]; ++x) {
                    const size_t i = ((size_t)z * dim[1] + y) * dim[0] + x;
                    const Vec3f p = voxel_to_world(Vec3i(x, y, z), vs, size);
                    ref_update(te, we, R * p + tr, i, depth, fd, 0.012f, Kin);
                    ref_update(tw, ww, R.transpose() * p + tr, i, depth, fd, 0.012f, Kin);
                }
        vfo_integrate_plain(to.data(), wo.data(), depth.data(), W, H, 0.012f, &vol, R.data(), tr.data(), Kin.data(),
                            15.0f);
        for (size_t i = 0; i < N; ++i) {
            ++samples;
            bad += !same(te[i], to[i]) || !same(we[i], wo[i]);
            differs += !same(tw[i], to[i]) || !same(ww[i], wo[i]);
        }
    }
    report("plain", samples, bad, differs);
}

void check_collide(Rng& r) {
    long samples = 0, bad = 0, differs = 0;
    for (int t = 0; t < 6; ++t) {
        const Vec3i dim(24, 20, 16);
        const Vec3f vs(0.005f, 0.0045f, 0.005f);
        const Vec3f size = dim.cast<float>().cwiseProduct(vs);
        vfo_vol vol;
        for (int a = 0; a < 3; ++a) vol.dim[a] = dim[a], vol.vs[a] = vs[a], vol.size[a] = size[a];
        const Mat3f R = r.rotation();
        const Vec3f tr(r.uni(-0.01f, 0.01f), r.uni(-0.01f, 0.01f), 0.5f);
        const size_t N = (size_t)dim.prod();
        std::vector<float> verts(3 * N);
        std::vector<int32_t> cnt(N, 1), col(N, 0), ce(N, 0), cw(N, 0);
        for (size_t i = 0; i < N; ++i) {
            // points of the rotated volume, moved by up to two voxels
            const Vec3i v((int)(i % dim[0]), (int)((i / dim[0]) % dim[1]), (int)(i / ((size_t)dim[0] * dim[1])));
            const Vec3f p = R * voxel_to_world(v, vs, size) + tr +
                            Vec3f(r.uni(-0.01f, 0.01f), r.uni(-0.01f, 0.01f), r.uni(-0.01f, 0.01f));
            std::memcpy(&verts[3 * i], p.data(), 12);
        }
        vfo_count_collisions(col.data(), verts.data(), cnt.data(), &vol, R.data(), tr.data());
        // src/TSDFIntegration.cu:270-282, Rinv = cVolumeRotate.inverse()
        const Mat3f Rinv = R.inverse(), Rwrong = Mat3f(R.transpose()).inverse();
        const Vec3f vsInv(1.0 / vs[0], 1.0 / vs[1], 1.0 / vs[2]);
        for (size_t i = 0; i < N; ++i) {
            const Vec3f wv(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
            for (int variant = 0; variant < 2; ++variant) {
                const Vec3f o = (variant ? Rwrong : Rinv) * (wv - tr);
                const Vec3f f = (o + 0.5f * size).cwiseProduct(vsInv);
                const Vec3i u = f.cast<int>();
                if ((u.array() < 0).any() || (u.array() >= dim.array()).any()) continue;
                (variant ? cw : ce)[((size_t)u[2] * dim[1] + u[1]) * dim[0] + u[0]] += 1;
            }
        }
        for (size_t i = 0; i < N; ++i) {
            ++samples;
            bad += ce[i] != col[i];
            differs += cw[i] != col[i];
        }
    }
    report("collide", samples, bad, differs);
}

// Gauss-Newton rows (src/MotionOptimization.cu:29-225): the depth Jacobian
// (n w)^T gradTv(p) wD, the depth residual wD n.dot(diff) and the
// regularisation residual ((T_g p4) - (T_nb p4)).head(3) * wR.
void check_gn(Rng& r) {
    using Mat3_6f = Eigen::Matrix<float, 3, 6>;
    long samples = 0, bad = 0, differs = 0;
    for (int t = 0; t < 20000; ++t) {
        const Vec3f p(r.wide(), r.wide(), r.wide()), n(r.wide(), r.wide(), r.wide());
        const float w = r.uni(0.01f, 1.0f), wD = r.uni(0.5f, 2.0f);
        Mat3_6f gradTv;
        gradTv << 0, p[2], -p[1], 1, 0, 0, -p[2], 0, p[0], 0, 1, 0, p[1], -p[0], 0, 0, 0, 1;
        const Eigen::Matrix<float, 1, 6> e = (n * w).transpose() * gradTv * wD;
        const Eigen::Matrix<float, 1, 6> wrong = (n * (w * wD)).transpose() * gradTv;
        float J[6] = {0};
        const int32_t corr[2] = {0, 0};
        const uint32_t knn = 0;
        vfo_gn_j_depth(J, wD, 1, 1, corr, p.data(), &knn, &w, 1, n.data());
        for (int k = 0; k < 6; ++k) {
            ++samples;
            bad += !same(e[k], J[k]);
            differs += !same(wrong[k], J[k]);
        }
    }
    report("gn_j_depth", samples, bad, differs);
    samples = bad = differs = 0;
    for (int t = 0; t < 20000; ++t) {
        const Vec3f v(r.wide(), r.wide(), r.wide()), d(r.wide(), r.wide(), r.wide()), n(r.wide(), r.wide(), r.wide());
        const float wD = r.uni(0.5f, 2.0f);
        const Vec3f diff = v - d;
        const float e = wD * n.dot(diff);
        const float wrong = wD * ((n[0] * diff[0] + n[1] * diff[1]) + n[2] * diff[2]);
        float out = 0.0f;
        const int32_t corr[2] = {0, 0};
        vfo_gn_r_depth(&out, wD, corr, 1, v.data(), d.data(), n.data());
        ++samples;
        bad += !same(e, out);
        differs += !same(wrong, out);
    }
    report("gn_r_depth", samples, bad, differs);
    samples = bad = differs = 0;
    for (int t = 0; t < 20000; ++t) {
        Mat4f Tg, Tn;
        for (int e = 0; e < 16; ++e) Tg.data()[e] = r.wide(), Tn.data()[e] = r.wide();
        std::vector<float> T(32);
        std::memcpy(&T[0], Tg.data(), 64);
        std::memcpy(&T[16], Tn.data(), 64);
        const Vec3f g0(r.wide(), r.wide(), r.wide()), g1(r.wide(), r.wide(), r.wide());
        const float graph[6] = {g0[0], g0[1], g0[2], g1[0], g1[1], g1[2]};
        const float wR = r.uni(0.5f, 2.0f);
        const Vec4f ph = g0.homogeneous();
        const Vec3f diff = (Tg * ph).head(3) - (Tn * ph).head(3);
        const Vec3f e = diff * wR;
        Vec3f wrong;
        for (int a = 0; a < 3; ++a) {
            const float x = (Tg(a, 0) * g0[0] + (Tg(a, 1) * g0[1] + Tg(a, 2) * g0[2])) + Tg(a, 3);
            const float y = (Tn(a, 0) * g0[0] + (Tn(a, 1) * g0[1] + Tn(a, 2) * g0[2])) + Tn(a, 3);
            wrong[a] = (x - y) * wR;
        }
        const uint32_t gknn[2] = {1, 0xffffffffu};
        float out[6] = {0};
        vfo_gn_r_mreg(out, wR, graph, T.data(), 2, gknn, 1);
        for (int a = 0; a < 3; ++a) {
            ++samples;
            bad += !same(e[a], out[a]);
            differs += !same(wrong[a], out[a]);
        }
    }
    report("gn_r_mreg", samples, bad, differs);
}

}  // namespace

int main() {
    Rng r;
    check_warp(r);
    check_mesh(r);
    check_distance(r);
    check_weights(r);
    check_iszero();
    check_inverse(r);
    check_plain(r);
    check_collide(r);
    check_gn(r);
    std::printf("eigen_order_check: %s\n", failures ? "FAILED" : "ok");
    return failures ? 1 : 0;
}
