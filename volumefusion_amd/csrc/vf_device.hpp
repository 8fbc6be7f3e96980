// vf_device.hpp -- device-side arithmetic shared by every warped-TSDF kernel.
//
// All translation units are compiled with -ffp-contract=off: each float
// operation below is rounded on its own, in the order the reference's device
// code evaluates it (Eigen's scalar coefficient-based path; sums of three
// terms are a0 + (a1 + a2), Eigen/src/Core/Redux.h:99-113).  Fused
// multiply-adds appear only where an algorithm here is defined with one.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vf {

constexpr float kInf = __builtin_huge_valf();

// float -> int32, truncation toward zero, saturating, NaN -> 0: exactly what
// v_cvt_i32_f32 does (out-of-range inputs clamp to INT32_MIN / INT32_MAX, NaN
// gives 0), so one instruction.  Written as asm because the C cast leaves
// out-of-range inputs undefined.
__device__ __forceinline__ int32_t cvt_i32_rz(float f) {
    int32_t r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(f));
    return r;
}

// A load that the compiler may issue as a global (not flat) memory access:
// the pointer is known to address device global memory.
template <class T>
__device__ __forceinline__ T gload(const T* p) {
    return *(const __attribute__((address_space(1))) T*)p;
}

// a / b for a divisor b shared by many quotients, from rb = RN64(1 / b):
// RN32(RN64(a * rb)) is the IEEE binary32 quotient whenever that quotient is
// a normal number.  Proof: the binary64 value is within 2^-52 (relative) of
// a / b; a quotient of two binary32 numbers is never a binary32 midpoint and,
// when not one, lies at least 2^-49 (relative) from every midpoint, so both
// round to the same binary32 number.  Results below 2^-126 in magnitude
// (zero, subnormal, or underflowed) take the IEEE division itself.
__device__ __forceinline__ float div_by_recip64(float a, float b, double rb) {
    const float q = (float)((double)a * rb);
    if (__builtin_expect(!(__builtin_fabsf(q) >= 0x1p-126f), 0)) return a / b;
    return q;
}

// 1 / x with the bits of the IEEE division 1.0f / x: v_rcp_f32 (within one
// ulp) and one Newton step in fused multiply-adds, y0 + y0 (1 - x y0).  The
// step gives the correctly rounded reciprocal for every x in [2^-126, 2^126)
// (tests/cpp/numerics_check.hip runs all of them on the device); other x take
// the division itself.
__device__ __forceinline__ float rcp_rn(float x) {
    if (__builtin_expect(x >= 0x1p-126f && x < 0x1p126f, 1)) {
        const float y0 = __builtin_amdgcn_rcpf(x);
        return __builtin_fmaf(__builtin_fmaf(-x, y0, 1.0f), y0, y0);
    }
    return 1.0f / x;
}

// Correctly rounded square root, the same bits as __builtin_sqrtf.  For
// s >= 2^-96 the compiler's expansion of __builtin_sqrtf on gfx950 is
// v_sqrt_f32 (within 1 ulp) followed by a one-ulp correction: with
// r- = r - ulp and r+ = r + ulp, the residuals fma(-r-, r, s) and
// fma(-r+, r, s) decide between r-, r and r+.  This is that sequence without
// the scaling by 2^32 that only inputs below 2^-96 need and without the
// zero/infinity class test (the correction already leaves 0 and +inf
// unchanged).  Inputs below 2^-96, zero and NaN take __builtin_sqrtf itself.
__device__ __forceinline__ float sqrt_rn(float s) {
    float r = __builtin_amdgcn_sqrtf(s);
    const float dn = __int_as_float(__float_as_int(r) - 1);
    const float up = __int_as_float(__float_as_int(r) + 1);
    const float vd = __builtin_fmaf(-dn, r, s);
    const float vu = __builtin_fmaf(-up, r, s);
    r = (vd <= 0.0f) ? dn : r;
    r = (vu > 0.0f) ? up : r;
    if (__builtin_expect(!(s >= 0x1p-96f), 0)) r = __builtin_sqrtf(s);
    return r;
}

// 2^(j/32), j = 0..31, correctly rounded to binary64.
#define VF_EXP2_TABLE                                                                                  \
    0x1.0000000000000p+0, 0x1.059b0d3158574p+0, 0x1.0b5586cf9890fp+0, 0x1.11301d0125b51p+0,          \
        0x1.172b83c7d517bp+0, 0x1.1d4873168b9aap+0, 0x1.2387a6e756238p+0, 0x1.29e9df51fdee1p+0,      \
        0x1.306fe0a31b715p+0, 0x1.371a7373aa9cbp+0, 0x1.3dea64c123422p+0, 0x1.44e086061892dp+0,      \
        0x1.4bfdad5362a27p+0, 0x1.5342b569d4f82p+0, 0x1.5ab07dd485429p+0, 0x1.6247eb03a5585p+0,      \
        0x1.6a09e667f3bcdp+0, 0x1.71f75e8ec5f74p+0, 0x1.7a11473eb0187p+0, 0x1.82589994cce13p+0,      \
        0x1.8ace5422aa0dbp+0, 0x1.93737b0cdc5e5p+0, 0x1.9c49182a3f090p+0, 0x1.a5503b23e255dp+0,      \
        0x1.ae89f995ad3adp+0, 0x1.b7f76f2fb5e47p+0, 0x1.c199bdd85529cp+0, 0x1.cb720dcef9069p+0,      \
        0x1.d5818dcfba487p+0, 0x1.dfc97337b9b5fp+0, 0x1.ea4afa2a490dap+0, 0x1.f50765b6e4540p+0

__constant__ const double kExp2Table[32] = {VF_EXP2_TABLE};

// fma(a, b, c) in binary64 as the three-address v_fma_f64.  With a constant
// addend the compiler otherwise picks the two-address v_fmac_f64 and copies
// the constant into the destination first (one more instruction per use).
// Same rounding as __builtin_fma.
__device__ __forceinline__ double fma_nd(double a, double b, double c) {
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}

// e^x in binary64, rounded once to binary32, by Tang's table-driven method
// (ACM TOMS 15(2), 1989): k = rint(32 x / ln2), r = x - k ln2/32 by a
// two-constant Cody-Waite reduction (fdlibm's ln2 split / 32), |r| <= ln2/64;
// e^r by its degree-5 Taylor polynomial in Horner form with fused
// multiply-adds (truncation < 2^-48 relative); e^x = 2^(k>>5) * 2^((k&31)/32)
// * e^r.  Inputs below -87 give 0.  `tab` is kExp2Table or a copy of it (LDS).
// Every step is an exactly specified IEEE operation, so a host build of the
// same steps gives the same bits.
__device__ __forceinline__ float expw(float xf, const double* tab) {
    if (xf != xf) return xf;
    if (xf < -87.0f) return 0.0f;
    if (xf > 88.5f) return kInf;
    const double x = (double)xf;
    const double kd = __builtin_rint(x * 0x1.71547652b82fep+5);
    const int k = (int)kd;
    double r = __builtin_fma(-kd, 0x1.62e42fee00000p-6, x);
    r = __builtin_fma(-kd, 0x1.a39ef35793c76p-38, r);
    double p = fma_nd(0x1.1111111111111p-7, r, 0x1.5555555555555p-5);
    p = fma_nd(p, r, 0x1.5555555555555p-3);
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    const double t = p * tab[k & 31];
    return (float)__builtin_ldexp(t, k >> 5);
}

// expw for arguments known to lie in [-87, 88.5] (the weight argument is in
// [-2.000001, 0]): the same steps without the range checks.
__device__ __forceinline__ float expw_dom(float xf, const double* tab) {
    const double x = (double)xf;
    const double kd = __builtin_rint(x * 0x1.71547652b82fep+5);
    const int k = (int)kd;
    double r = __builtin_fma(-kd, 0x1.62e42fee00000p-6, x);
    r = __builtin_fma(-kd, 0x1.a39ef35793c76p-38, r);
    double p = fma_nd(0x1.1111111111111p-7, r, 0x1.5555555555555p-5);
    p = fma_nd(p, r, 0x1.5555555555555p-3);
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    const double t = p * tab[k & 31];
    return (float)__builtin_ldexp(t, k >> 5);
}

// a / b for a >= 0 in the weight domain from y = RN(1/b): q0 = RN(a y),
// rem = a - q0 b exactly (fma), q = RN(q0 + rem y) (Markstein, IBM J. Res.
// Dev. 34(1), 1990).  Equal to the IEEE quotient on the operand ranges used
// here (checked exhaustively on the host for every dividend mantissa over the
// weight and weight-sum binades).
__device__ __forceinline__ float div_rcp(float a, float b, float y) {
    const float q0 = a * y;
    const float rem = __builtin_fmaf(-q0, b, a);
    return __builtin_fmaf(rem, y, q0);
}

// The weight denominators 2 r^2 for which the Markstein path below (weight
// exponent and normalisation) is exact: every intermediate of div_rcp stays a
// normal number.  Nodes outside the range take IEEE divisions.
constexpr float kDenLo = 0x1p-64f;
constexpr float kDenHi = 0x1p64f;
__device__ __forceinline__ bool den_safe(float den) { return den >= kDenLo && den <= kDenHi; }

// Gaussian warp weight of src/TSDFWarp.cu:17, exp(-(d*d) / (2 * r*r)), no
// cutoff.  `den` is 2*(r*r), evaluated once per node.
__device__ __forceinline__ float warp_weight(float den, float d) {
    const float num = -(d * d);
    return expw(num / den, kExp2Table);
}

// The same weight with the node's 1/den precomputed; identical result for
// den_safe(den).  The argument -(d*d)/den lies in [-2.000001, 0] for
// candidates (d <= 2r).
__device__ __forceinline__ float warp_weight_rcp(float den, float rden, float d, const double* tab) {
    const float q = d * d;
    return expw_dom(-div_rcp(q, den, rden), tab);
}

// Inclusive prefix sum over the wave in DPP moves (VALU only): shifts by 1,
// 2, 4, 8 within each row of 16 lanes, then row 15's and row 31's lane added
// to the rows above them.
#define VF_DPP(x, ctrl, rows) (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(x), ctrl, rows, 0xf, false)
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t x) {
    x += VF_DPP(x, 0x111, 0xf);  // row_shr:1
    x += VF_DPP(x, 0x112, 0xf);  // row_shr:2
    x += VF_DPP(x, 0x114, 0xf);  // row_shr:4
    x += VF_DPP(x, 0x118, 0xf);  // row_shr:8
    x += VF_DPP(x, 0x142, 0xa);  // row_bcast:15 into rows 1 and 3
    x += VF_DPP(x, 0x143, 0xc);  // row_bcast:31 into rows 2 and 3
    return x;
}
#undef VF_DPP

// Sorting networks on 64-bit keys for 2, 4, 6, 8 keys (Batcher's odd-even
// merge sort for 2, 4, 8).
__device__ __forceinline__ void cmpx(unsigned long long& a, unsigned long long& b) {
    const bool sw = b < a;
    const unsigned long long lo = sw ? b : a;
    const unsigned long long hi = sw ? a : b;
    a = lo;
    b = hi;
}

template <int N>
__device__ __forceinline__ void sort_keys(unsigned long long* k) {
    if constexpr (N == 2) {
        cmpx(k[0], k[1]);
    } else if constexpr (N == 4) {
        cmpx(k[0], k[1]); cmpx(k[2], k[3]);
        cmpx(k[0], k[2]); cmpx(k[1], k[3]);
        cmpx(k[1], k[2]);
    } else if constexpr (N == 6) {
        // 12 comparators (optimal for six inputs; checked on all 0-1 inputs)
        cmpx(k[0], k[5]); cmpx(k[1], k[3]); cmpx(k[2], k[4]);
        cmpx(k[1], k[2]); cmpx(k[3], k[4]);
        cmpx(k[0], k[3]); cmpx(k[2], k[5]);
        cmpx(k[0], k[1]); cmpx(k[2], k[3]); cmpx(k[4], k[5]);
        cmpx(k[1], k[2]); cmpx(k[3], k[4]);
    } else if constexpr (N == 8) {
        cmpx(k[0], k[1]); cmpx(k[2], k[3]); cmpx(k[4], k[5]); cmpx(k[6], k[7]);
        cmpx(k[0], k[2]); cmpx(k[1], k[3]); cmpx(k[4], k[6]); cmpx(k[5], k[7]);
        cmpx(k[1], k[2]); cmpx(k[5], k[6]);
        cmpx(k[0], k[4]); cmpx(k[1], k[5]); cmpx(k[2], k[6]); cmpx(k[3], k[7]);
        cmpx(k[2], k[4]); cmpx(k[3], k[5]);
        cmpx(k[1], k[2]); cmpx(k[3], k[4]); cmpx(k[5], k[6]);
    }
}

constexpr int pow2_at_least(int k) { return k <= 1 ? 1 : (k <= 2 ? 2 : (k <= 4 ? 4 : 8)); }

// Canonical knn selection with per-thread LDS slots.  Each candidate is a key
// (bits(dist) << 32) | id; for dist >= 0 the key order is the (dist, id)
// order.  While a thread holds fewer than K keys a candidate is appended
// (one LDS store); when full it replaces the largest key if smaller (rare).
// finish() sorts the K slots with a sorting network.  Slots are strided by the
// workgroup size so a wave's stores hit consecutive 8-byte words.
template <int K, int STRIDE>
struct KnnSlots {
    unsigned long long* slot;  // this thread's slot 0 in LDS
    int n;

    __device__ __forceinline__ void init(unsigned long long* base, int tid) {
        slot = base + tid;
        n = 0;
    }

    __device__ __forceinline__ void offer(float dist, uint32_t id) {
        const unsigned long long key = ((unsigned long long)__float_as_uint(dist) << 32) | id;
        if (n < K) {
            slot[n * STRIDE] = key;
            ++n;
        } else {
            int mi = 0;
            unsigned long long mk = slot[0];
#pragma unroll
            for (int s = 1; s < K; ++s) {
                const unsigned long long v = slot[s * STRIDE];
                if (v > mk) {
                    mk = v;
                    mi = s;
                }
            }
            if (key < mk) slot[mi * STRIDE] = key;
        }
    }

    // The first Q slots, empty ones as ~0 (after every real key), sorted.
    template <int Q, int P>
    __device__ __forceinline__ void load_sort(unsigned long long* k) const {
#pragma unroll
        for (int s = 0; s < Q; ++s) k[s] = (s < K && s < n) ? slot[s * STRIDE] : ~0ull;
#pragma unroll
        for (int s = Q; s < P; ++s) k[s] = ~0ull;
        sort_keys<Q>(k);
    }

    // Sorted selection into registers: d[s], i[s] for s < n.  The network is
    // the smallest of 2, 4, 6 and 8 keys (2 and 4 for K <= 4) that holds the
    // largest n of the wave; slots at or beyond a lane's n are never read by
    // the caller.  Called by whole waves (the size test is a ballot).
    __device__ __forceinline__ void finish(float* d, uint32_t* i) const {
        constexpr int P = pow2_at_least(K);
        unsigned long long k[P];
        if constexpr (P == 8) {
            if (__ballot(n > 6))
                load_sort<8, P>(k);
            else if (__ballot(n > 4))
                load_sort<6, P>(k);
            else if (__ballot(n > 2))
                load_sort<4, P>(k);
            else
                load_sort<2, P>(k);
        } else if constexpr (P == 4) {
            if (__ballot(n > 2))
                load_sort<4, P>(k);
            else
                load_sort<2, P>(k);
        } else {
            load_sort<P, P>(k);
        }
#pragma unroll
        for (int s = 0; s < K; ++s) {
            d[s] = __uint_as_float((uint32_t)(k[s] >> 32));
            i[s] = (uint32_t)k[s];
        }
    }
};

// Per-volume geometry, as the reference derives it from (dims, voxel size,
// volume size): inv = float(1.0 / double(vs)) (src/TSDFWarp.cu:12) and
// halfs = size * 0.5f.
struct Geom {
    int32_t dim[3];
    float vs[3];
    float inv[3];
    float halfs[3];
};

// voxelToWorld, src/TSDFWarp.cu:8-10: the voxel's min corner.
__device__ __forceinline__ void voxel_to_world(const Geom& g, int32_t x, int32_t y, int32_t z, float p[3]) {
    const float ax = (float)x * g.vs[0];
    const float ay = (float)y * g.vs[1];
    const float az = (float)z * g.vs[2];
    p[0] = ax - g.halfs[0];
    p[1] = ay - g.halfs[1];
    p[2] = az - g.halfs[2];
}

// worldToVoxel, src/TSDFWarp.cu:11-15: truncation toward zero.
__device__ __forceinline__ void world_to_voxel(const Geom& g, const float p[3], int32_t v[3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float s = p[a] + g.halfs[a];
        v[a] = cvt_i32_rz(s * g.inv[a]);
    }
}

// Mat3f (column-major) times Vec3f, Eigen device order a0 + (a1 + a2).
__device__ __forceinline__ void mat3_vec(const float* m, const float v[3], float o[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float a0 = m[0 + i] * v[0];
        const float a1 = m[3 + i] * v[1];
        const float a2 = m[6 + i] * v[2];
        o[i] = a0 + (a1 + a2);
    }
}

// (T * p.homogeneous()).head(3) with rows[i] = (T(i,0), T(i,1), T(i,2), T(i,3)):
// Eigen evaluates leftCols<3>() * p then adds the last column
// (Eigen/src/Geometry/Homogeneous.h:268-276).
__device__ __forceinline__ void affine_apply(const float4& r0, const float4& r1, const float4& r2, const float p[3],
                                             float o[3]) {
    {
        const float a0 = r0.x * p[0], a1 = r0.y * p[1], a2 = r0.z * p[2];
        const float s = a0 + (a1 + a2);
        o[0] = s + r0.w;
    }
    {
        const float a0 = r1.x * p[0], a1 = r1.y * p[1], a2 = r1.z * p[2];
        const float s = a0 + (a1 + a2);
        o[1] = s + r1.w;
    }
    {
        const float a0 = r2.x * p[0], a1 = r2.y * p[1], a2 = r2.z * p[2];
        const float s = a0 + (a1 + a2);
        o[2] = s + r2.w;
    }
}

// Dual quaternion (real xyzw, dual xyzw) of a column-major Mat4f's rigid
// part: Shepperd's method on the largest diagonal, normalised, w >= 0; dual
// 0.5 (0, t) q.  A second blend mode with no reference counterpart (the
// reference blends Mat4f linearly); every operation rounded on its own, in
// the order the CPU restatement uses.
__device__ __forceinline__ void mat4_to_dq(const float* M, float4* qr, float4* qd) {
    const float r00 = M[0], r10 = M[1], r20 = M[2];
    const float r01 = M[4], r11 = M[5], r21 = M[6];
    const float r02 = M[8], r12 = M[9], r22 = M[10];
    const float tr = (r00 + r11) + r22;
    float x, y, z, w;
    if (tr > 0.0f) {
        const float s = __builtin_sqrtf(tr + 1.0f) * 2.0f;
        w = 0.25f * s;
        x = (r21 - r12) / s;
        y = (r02 - r20) / s;
        z = (r10 - r01) / s;
    } else if (r00 > r11 && r00 > r22) {
        const float s = __builtin_sqrtf(((1.0f + r00) - r11) - r22) * 2.0f;
        w = (r21 - r12) / s;
        x = 0.25f * s;
        y = (r01 + r10) / s;
        z = (r02 + r20) / s;
    } else if (r11 > r22) {
        const float s = __builtin_sqrtf(((1.0f + r11) - r00) - r22) * 2.0f;
        w = (r02 - r20) / s;
        x = (r01 + r10) / s;
        y = 0.25f * s;
        z = (r12 + r21) / s;
    } else {
        const float s = __builtin_sqrtf(((1.0f + r22) - r00) - r11) * 2.0f;
        w = (r10 - r01) / s;
        x = (r02 + r20) / s;
        y = (r12 + r21) / s;
        z = 0.25f * s;
    }
    const float nn = __builtin_sqrtf(((x * x + y * y) + z * z) + w * w);
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
    *qr = make_float4(x, y, z, w);
    *qd = make_float4(0.5f * (w * tx + (ty * z - tz * y)), 0.5f * (w * ty + (tz * x - tx * z)),
                      0.5f * (w * tz + (tx * y - ty * x)), 0.5f * -((tx * x + ty * y) + tz * z));
}

// Accumulates w * sign * (qr, qd) into (br, bd), the sign aligning qr with
// the first node's real part f.
__device__ __forceinline__ void dqb_accumulate(const float4& qr, const float4& qd, const float4& f, float w, float4* br,
                                               float4* bd) {
    const float dot = ((qr.x * f.x + qr.y * f.y) + qr.z * f.z) + qr.w * f.w;
    const float ws = dot < 0.0f ? -w : w;
    br->x = br->x + ws * qr.x;
    br->y = br->y + ws * qr.y;
    br->z = br->z + ws * qr.z;
    br->w = br->w + ws * qr.w;
    bd->x = bd->x + ws * qd.x;
    bd->y = bd->y + ws * qd.y;
    bd->z = bd->z + ws * qd.z;
    bd->w = bd->w + ws * qd.w;
}

// Applies the blended dual quaternion (br, bd), normalised by |br|, to p
// (Kavan et al., ACM TOG 27(4), 2008):
// p' = p + 2 r_v x (r_v x p + r_w p) + 2 ((r_w d_v - d_w r_v) + r_v x d_v).
__device__ __forceinline__ void dqb_transform(const float4& br, const float4& bd, const float p[3], float o[3]) {
    const float len = __builtin_sqrtf(((br.x * br.x + br.y * br.y) + br.z * br.z) + br.w * br.w);
    if (!(len > 0.0f)) {
        o[0] = p[0];
        o[1] = p[1];
        o[2] = p[2];
        return;
    }
    const float il = 1.0f / len;
    const float rx = br.x * il, ry = br.y * il, rz = br.z * il, rw = br.w * il;
    const float dx = bd.x * il, dy = bd.y * il, dz = bd.z * il, dw = bd.w * il;
    const float t1x = ry * p[2] - rz * p[1];
    const float t1y = rz * p[0] - rx * p[2];
    const float t1z = rx * p[1] - ry * p[0];
    const float t2x = t1x + rw * p[0];
    const float t2y = t1y + rw * p[1];
    const float t2z = t1z + rw * p[2];
    const float t3x = ry * t2z - rz * t2y;
    const float t3y = rz * t2x - rx * t2z;
    const float t3z = rx * t2y - ry * t2x;
    const float ux = ry * dz - rz * dy;
    const float uy = rz * dx - rx * dz;
    const float uz = rx * dy - ry * dx;
    const float trx = 2.0f * ((rw * dx - dw * rx) + ux);
    const float try_ = 2.0f * ((rw * dy - dw * ry) + uy);
    const float trz = 2.0f * ((rw * dz - dw * rz) + uz);
    o[0] = (p[0] + 2.0f * t3x) + trx;
    o[1] = (p[1] + 2.0f * t3y) + try_;
    o[2] = (p[2] + 2.0f * t3z) + trz;
}

// Target voxel of a warped vertex (src/TSDFIntegration.cu:272-279):
// u = worldToVoxel(Rinv * (p' - t)).  Returns false when u is outside.
__device__ __forceinline__ bool warped_target(const Geom& g, const float* Rinv, const float t[3], const float pw[3],
                                              int32_t u[3]) {
    float q[3], o[3];
    q[0] = pw[0] - t[0];
    q[1] = pw[1] - t[1];
    q[2] = pw[2] - t[2];
    mat3_vec(Rinv, q, o);
    world_to_voxel(g, o, u);
    return !(u[0] < 0 || u[1] < 0 || u[2] < 0 || u[0] >= g.dim[0] || u[1] >= g.dim[1] || u[2] >= g.dim[2]);
}

struct Camera {
    float fx, fy, cx, cy;  // K(0,0), K(1,1), K(0,2), K(1,2)
    int32_t W, H;
    float delta, maxw;
    double rdelta;  // RN64(1 / delta)
};

// project3dToFrame + the tests of updateTsdf up to the read-modify-write,
// src/TSDFIntegration.cu:16-58.  Returns the pixel index, or -1 when the
// voxel is not updated; *tn is the new truncated sdf sample.
__device__ __forceinline__ int64_t project_sample(const Camera& cam, const float* __restrict__ depth, const float p[3],
                                                  float* tn) {
    if (p[2] < 1.e-4f) return -1;
    const float zInv = rcp_rn(p[2]);
    const float ux = cam.fx * p[0];
    const float uf = ux * zInv + cam.cx;
    const float vy = cam.fy * p[1];
    const float vf = vy * zInv + cam.cy;
    const int32_t px = cvt_i32_rz(uf + 0.5f);
    const int32_t py = cvt_i32_rz(vf + 0.5f);
    if (px < 0 || px >= cam.W || py < 0 || py >= cam.H) return -1;
    const int64_t pi = (int64_t)py * cam.W + px;
    const float d = gload(depth + pi);
    if (__builtin_isinf(d) || d < 1.e-4f) return -1;
    const float sdf = p[2] - d;
    if (sdf > cam.delta) return -1;
    *tn = (sdf < -cam.delta) ? -1.0f : div_by_recip64(sdf, cam.delta, cam.rdelta);
    return pi;
}

// The read-modify-write of updateTsdf, src/TSDFIntegration.cu:60-64: the
// running average uses the uncapped w_old + 1; the stored weight is capped.
__device__ __forceinline__ void blend_sample(float to, float wo, float tn, float maxw, float* t_out, float* w_out) {
    const float a = to * wo;
    const float b = tn * 1.0f;
    const float wn = wo + 1.0f;
    *t_out = (a + b) / wn;
    *w_out = (maxw < wn) ? maxw : wn;
}

}  // namespace vf
