// kernels.hpp -- gfx950 device code of liboctofmm: FP64 pair arithmetic, the
// AoS->record staging kernel, the level interaction kernel and the root kernel.
//
// Pair arithmetic follows Octo-Tiger's templates
// (octotiger/multipole_interactions/kernel/compute_kernel_templates.hpp:67-391,
//  octotiger/monopole_interactions/kernel/monopole_kernel_templates.hpp:16-239):
// the same derivative tensor D of -1/r, the same order-3 contraction with
// multiplicity factors, and the same angular-momentum correction from the
// order-4 tensor.  Only the summation order over pairs differs.
#pragma once

#include <hip/hip_runtime.h>

#include "octofmm.h"
#include "octofmm_internal.hpp"

namespace octofmm {

// ---------------------------------------------------------------------------
// Pair arithmetic
//
// Partner records carry the multiplicity factors of the contraction folded
// into the multipole: m4, m7, m9 are stored halved (factor_half) and the
// octupole m10..m19 scaled by factor_sixth = {1,3,3,3,6,3,1,3,3,1}/6, so every
// term of the contraction is one fused multiply-add.  The target's own
// octupole used by the angular-momentum correction carries the same scaling,
// which makes the scaled octupole difference directly the correction weight.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr double record_scale(int c) {
    return (c == 4 || c == 7 || c == 9) ? 0.5
        : (c == 10 || c == 16 || c == 19) ? 1.0 / 6.0
        : (c == 14) ? 1.0
        : (c >= 11 && c <= 18) ? 0.5
        : 1.0;
}

// Sum over the 8 lanes of an aligned group, by three DPP lane moves (VALU ops,
// no LDS round trip): the pair of xor 1, the pair of xor 2 within each quad,
// then the mirrored lane of the other quad (7 - i within each half row, once
// every lane holds its quad's sum).  Every lane ends with the same value:
// ((v0 + v1) + (v2 + v3)) + ((v4 + v5) + (v6 + v7)), in the same order on
// every run.  The 8 lanes of a group must be active together.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double sum8(double v) {
    v += dpp_move<0xB1>(v);    // quad_perm [1, 0, 3, 2]
    v += dpp_move<0x4E>(v);    // quad_perm [2, 3, 0, 1]
    v += dpp_move<0x141>(v);   // row_half_mirror
    return v;
}

struct DTensor {
    double D[20];
    double d2, d3;
    // d3 times monomials of the separation, shared with the correction
    double d3x00, d3x11, d3x22, d3x01, d3x02, d3x12;
};

// 1/sqrt(x): the hardware estimate (relative error <= 5.2e-8, measured on
// MI355X over x in [1e-12, 1e6]) refined by one Newton step, which squares the
// error: <= 4.1e-15 measured (a second step would give 1.4e-16).  The
// derivative tensor scales it by at most 7 (d3 ~ r^-7), far inside the 1e-12
// gate, and the step saves three instructions per pair.
__device__ __forceinline__ double rsqrt_nr(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-((0.5 * x) * y), y, 0.5);
    return fma(y, e, y);
}

// Derivatives of -1/r up to order 3 at dX (compute_d_factors), with the
// reference's floor r^2 >= 1e-20.
__device__ __forceinline__ void d_tensor(double x, double y, double z, DTensor& t) {
    const double x00 = x * x;
    double r2 = fma(z, z, fma(y, y, x00));
    r2 = fmax(r2, 1.0e-20);   // one v_max_f64; r2 of finite separations is never NaN
    const double ri = rsqrt_nr(r2);           // 1/r
    const double r2inv = ri * ri;
    const double d1 = ri * r2inv;             // 1/r^3
    const double d2 = (-3.0 * r2inv) * d1;    // -3/r^5
    const double d3 = (-5.0 * r2inv) * d2;    // 15/r^7
    t.d2 = d2;
    t.d3 = d3;
    double* D = t.D;
    D[0] = -ri;
    D[1] = x * d1;
    D[2] = y * d1;
    D[3] = z * d1;
    const double d2x = d2 * x, d2y = d2 * y, d2z = d2 * z;
    // second order from the d2 x_a products the order-3 terms need anyway
    D[4] = fma(d2, x00, d1);
    D[5] = d2x * y;
    D[6] = d2x * z;
    D[7] = fma(d2y, y, d1);
    D[8] = d2y * z;
    D[9] = fma(d2z, z, d1);
    const double d3x = d3 * x, d3y = d3 * y;
    const double d3x00 = d3x * x, d3x11 = d3y * y, d3x22 = (d3 * z) * z;
    const double d3x01 = d3x * y, d3x02 = d3x * z, d3x12 = d3y * z;
    t.d3x00 = d3x00;
    t.d3x11 = d3x11;
    t.d3x22 = d3x22;
    t.d3x01 = d3x01;
    t.d3x02 = d3x02;
    t.d3x12 = d3x12;
    D[10] = fma(d3x00, x, 3.0 * d2x);
    D[11] = fma(d3x00, y, d2y);
    D[12] = fma(d3x00, z, d2z);
    D[13] = fma(d3x11, x, d2x);
    D[14] = d3x12 * x;
    D[15] = fma(d3x22, x, d2x);
    D[16] = fma(d3x11, y, 3.0 * d2y);
    D[17] = fma(d3x11, z, d2z);
    D[18] = fma(d3x22, y, d2y);
    D[19] = fma(d3x22, z, 3.0 * d2z);
}

// Order-2/3 part of the contraction (compute_interaction_multipole_non_rho)
// with factor-scaled partner components m.  NCOMP == 4: potential and
// gradient rows only (p2m targets).
template <int NCOMP>
__device__ __forceinline__ void contract_high(const double* m, const double* D, double* L) {
    double l0 = L[0], l1 = L[1], l2 = L[2], l3 = L[3];
    l0 = fma(m[4], D[4], l0);   l1 = fma(m[4], D[10], l1);  l2 = fma(m[4], D[11], l2);  l3 = fma(m[4], D[12], l3);
    l0 = fma(m[5], D[5], l0);   l1 = fma(m[5], D[11], l1);  l2 = fma(m[5], D[13], l2);  l3 = fma(m[5], D[14], l3);
    l0 = fma(m[6], D[6], l0);   l1 = fma(m[6], D[12], l1);  l2 = fma(m[6], D[14], l2);  l3 = fma(m[6], D[15], l3);
    l0 = fma(m[7], D[7], l0);   l1 = fma(m[7], D[13], l1);  l2 = fma(m[7], D[16], l2);  l3 = fma(m[7], D[17], l3);
    l0 = fma(m[8], D[8], l0);   l1 = fma(m[8], D[14], l1);  l2 = fma(m[8], D[17], l2);  l3 = fma(m[8], D[18], l3);
    l0 = fma(m[9], D[9], l0);   l1 = fma(m[9], D[15], l1);  l2 = fma(m[9], D[18], l2);  l3 = fma(m[9], D[19], l3);
#pragma unroll
    for (int k = 10; k < 20; ++k) l0 = fma(-m[k], D[k], l0);
    L[0] = l0; L[1] = l1; L[2] = l2; L[3] = l3;
    if constexpr (NCOMP == 20) {
        // L_ab += m0 D_ab - m_c D_abc
        L[4] = fma(m[0], D[4], L[4]); L[4] = fma(-m[1], D[10], L[4]); L[4] = fma(-m[2], D[11], L[4]); L[4] = fma(-m[3], D[12], L[4]);
        L[5] = fma(m[0], D[5], L[5]); L[5] = fma(-m[1], D[11], L[5]); L[5] = fma(-m[2], D[13], L[5]); L[5] = fma(-m[3], D[14], L[5]);
        L[6] = fma(m[0], D[6], L[6]); L[6] = fma(-m[1], D[12], L[6]); L[6] = fma(-m[2], D[14], L[6]); L[6] = fma(-m[3], D[15], L[6]);
        L[7] = fma(m[0], D[7], L[7]); L[7] = fma(-m[1], D[13], L[7]); L[7] = fma(-m[2], D[16], L[7]); L[7] = fma(-m[3], D[17], L[7]);
        L[8] = fma(m[0], D[8], L[8]); L[8] = fma(-m[1], D[14], L[8]); L[8] = fma(-m[2], D[17], L[8]); L[8] = fma(-m[3], D[18], L[8]);
        L[9] = fma(m[0], D[9], L[9]); L[9] = fma(-m[1], D[15], L[9]); L[9] = fma(-m[2], D[18], L[9]); L[9] = fma(-m[3], D[19], L[9]);
#pragma unroll
        for (int i = 10; i < 20; ++i) L[i] = fma(m[0], D[i], L[i]);
    }
}

// Dipole rows of the DRHODT kernels (compute_kernel_non_rho, :374-389).  The
// reference subtracts the whole first row of the second-order tensor from
// each gradient component; kept as is.
__device__ __forceinline__ void dipole_rows(const double* m, const double* D, double* L) {
    L[0] = fma(-m[1], D[1], L[0]);
    L[0] = fma(-m[2], D[2], L[0]);
    L[0] = fma(-m[3], D[3], L[0]);
    L[1] = fma(-m[1], D[4] + D[5] + D[6], L[1]);
    L[2] = fma(-m[2], D[5] + D[7] + D[8], L[2]);
    L[3] = fma(-m[3], D[6] + D[8] + D[9], L[3]);
}

// Angular-momentum correction from the order-4 tensor (compute_interaction_multipole_rho,
// :224-334; compute_interaction_p2m_rho, monopole_kernel_templates.hpp:69-187).
// a[k] = factor-scaled octupole difference (M2L) or partner octupole (p2m).
__device__ __forceinline__ void am_correction_t(double d2, double d3x00, double d3x11, double d3x22, double d3x01,
    double d3x02, double d3x12, const double* a, double* C) {
    // 6.0 in a scalar register pair: as a literal it makes the compiler use
    // the two-operand FMA, whose tied destination costs a register copy for
    // each of U0 and U10 (the three-operand form takes no 64-bit literal)
    double six = 6.0;
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+s"(six));
#endif
    const double t2 = 3.0 * d2;
    const double U0 = fma(six, d3x00, t2);
    const double U1 = 3.0 * d3x01;
    const double U2 = 3.0 * d3x02;
    const double U3 = d2 + d3x11 + d3x00;
    const double U4 = d3x12;
    const double U5 = d2 + d3x22 + d3x00;
    const double U7 = d3x02;
    const double U8 = d3x01;
    const double U10 = fma(six, d3x11, t2);
    const double U12 = d2 + d3x22 + d3x11;
    const double U13 = 3.0 * d3x12;
    const double U14 = fma(six, d3x22, t2);
    double c0 = C[0], c1 = C[1], c2 = C[2];
    c0 = fma(-a[0], U0, c0);  c1 = fma(-a[0], U1, c1);  c2 = fma(-a[0], U2, c2);
    c0 = fma(-a[1], U1, c0);  c1 = fma(-a[1], U3, c1);  c2 = fma(-a[1], U4, c2);
    c0 = fma(-a[2], U2, c0);  c1 = fma(-a[2], U4, c1);  c2 = fma(-a[2], U5, c2);
    c0 = fma(-a[3], U3, c0);  c1 = fma(-a[3], U1, c1);  c2 = fma(-a[3], U7, c2);
    c0 = fma(-a[4], U4, c0);  c1 = fma(-a[4], U7, c1);  c2 = fma(-a[4], U8, c2);
    c0 = fma(-a[5], U5, c0);  c1 = fma(-a[5], U8, c1);  c2 = fma(-a[5], U2, c2);
    c0 = fma(-a[6], U1, c0);  c1 = fma(-a[6], U10, c1); c2 = fma(-a[6], U13, c2);
    c0 = fma(-a[7], U7, c0);  c1 = fma(-a[7], U13, c1); c2 = fma(-a[7], U12, c2);
    c0 = fma(-a[8], U8, c0);  c1 = fma(-a[8], U12, c1); c2 = fma(-a[8], U13, c2);
    c0 = fma(-a[9], U2, c0);  c1 = fma(-a[9], U13, c1); c2 = fma(-a[9], U14, c2);
    C[0] = c0; C[1] = c1; C[2] = c2;
}

// The same correction from the separation and d2, d3 (used where no
// derivative tensor with the shared products is at hand).
__device__ __forceinline__ void am_correction(double x, double y, double z, double d2, double d3,
    const double* a, double* C) {
    const double d3x = d3 * x, d3y = d3 * y;
    am_correction_t(d2, d3x * x, d3y * y, (d3 * z) * z, d3x * y, d3x * z, d3y * z, a, C);
}

__device__ __forceinline__ void load_record20(const double* __restrict__ p, double* m) {
    const double2* p2 = reinterpret_cast<const double2*>(p);
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        const double2 v = p2[i];
        m[2 * i] = v.x;
        m[2 * i + 1] = v.y;
    }
}

// A partner record held in registers (24 doubles as 12 double2).
struct Rec {
    double2 v[12];
};

__device__ __forceinline__ void load_rec(const double* __restrict__ p, Rec& r) {
    const double2* p2 = reinterpret_cast<const double2*>(p);
#pragma unroll
    for (int i = 0; i < 12; ++i) r.v[i] = p2[i];
}

// One far M2L pair: partner record (m0, scaled M1..M19, Y, mon) onto the
// target at X with mass reciprocal inv_mc0 and scaled octupole mc[0..9].
template <bool RHO>
__device__ __forceinline__ void m2l_far_rec(const Rec& r, double X0, double X1, double X2, double inv_mc0,
    const double* mc, double* L, double* C) {
    double m[20];
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        m[2 * i] = r.v[i].x;
        m[2 * i + 1] = r.v[i].y;
    }
    const double x = X0 - r.v[10].x, y = X1 - r.v[10].y, z = X2 - r.v[11].x;
    DTensor t;
    d_tensor(x, y, z, t);
#pragma unroll
    for (int i = 0; i < 4; ++i) L[i] = fma(m[0], t.D[i], L[i]);
    if constexpr (!RHO) dipole_rows(m, t.D, L);
    contract_high<20>(m, t.D, L);
    if constexpr (RHO) {
        const double n0c = m[0] * inv_mc0;
        double a[10];
#pragma unroll
        for (int k = 0; k < 10; ++k) a[k] = fma(-mc[k], n0c, m[10 + k]);
        am_correction_t(t.d2, t.d3x00, t.d3x11, t.d3x22, t.d3x01, t.d3x02, t.d3x12, a, C);
    }
}

// FP32 half of a mixed-precision far pair: the derivative tensor of -1/r at
// the separation (x, y, z), already rounded to FP32, and the 20-component
// contraction with the factor-scaled partner multipole m, accumulated into
// the FP32 partial sums Lf.  Returns the FP32 estimate of 1/r, from which the
// FP64 correction starts.
template <bool RHO>
__device__ __forceinline__ float m2l_fp32_core(float x, float y, float z, const float* m, float* Lf) {
    const float x00 = x * x, x11 = y * y, x22 = z * z;
    float r2 = fmaf(z, z, fmaf(y, y, x00));
    r2 = fmaxf(r2, 1.0e-20f);   // one v_max_f32; r2 of finite separations is never NaN
    const float ri = __builtin_amdgcn_rsqf(r2);
    const float r2inv = ri * ri;
    const float d1 = ri * r2inv;
    const float d2 = (-3.0f * r2inv) * d1;
    const float d3 = (-5.0f * r2inv) * d2;
    float D[20];
    D[0] = -ri;
    D[1] = x * d1;
    D[2] = y * d1;
    D[3] = z * d1;
    D[4] = fmaf(d2, x00, d1);
    D[5] = d2 * (x * y);
    D[6] = d2 * (x * z);
    D[7] = fmaf(d2, x11, d1);
    D[8] = d2 * (y * z);
    D[9] = fmaf(d2, x22, d1);
    const float d2x = d2 * x, d2y = d2 * y, d2z = d2 * z;
    const float d3x00 = d3 * x00, d3x11 = d3 * x11, d3x22 = d3 * x22;
    D[10] = fmaf(d3x00, x, 3.0f * d2x);
    D[11] = fmaf(d3x00, y, d2y);
    D[12] = fmaf(d3x00, z, d2z);
    D[13] = fmaf(d3x11, x, d2x);
    D[14] = (d3 * x) * (y * z);
    D[15] = fmaf(d3x22, x, d2x);
    D[16] = fmaf(d3x11, y, 3.0f * d2y);
    D[17] = fmaf(d3x11, z, d2z);
    D[18] = fmaf(d3x22, y, d2y);
    D[19] = fmaf(d3x22, z, 3.0f * d2z);
#pragma unroll
    for (int i = 0; i < 4; ++i) Lf[i] = fmaf(m[0], D[i], Lf[i]);
    if constexpr (!RHO) {
        Lf[0] = fmaf(-m[1], D[1], Lf[0]);
        Lf[0] = fmaf(-m[2], D[2], Lf[0]);
        Lf[0] = fmaf(-m[3], D[3], Lf[0]);
        Lf[1] = fmaf(-m[1], D[4] + D[5] + D[6], Lf[1]);
        Lf[2] = fmaf(-m[2], D[5] + D[7] + D[8], Lf[2]);
        Lf[3] = fmaf(-m[3], D[6] + D[8] + D[9], Lf[3]);
    }
    float l0 = Lf[0], l1 = Lf[1], l2 = Lf[2], l3 = Lf[3];
    l0 = fmaf(m[4], D[4], l0);   l1 = fmaf(m[4], D[10], l1);  l2 = fmaf(m[4], D[11], l2);  l3 = fmaf(m[4], D[12], l3);
    l0 = fmaf(m[5], D[5], l0);   l1 = fmaf(m[5], D[11], l1);  l2 = fmaf(m[5], D[13], l2);  l3 = fmaf(m[5], D[14], l3);
    l0 = fmaf(m[6], D[6], l0);   l1 = fmaf(m[6], D[12], l1);  l2 = fmaf(m[6], D[14], l2);  l3 = fmaf(m[6], D[15], l3);
    l0 = fmaf(m[7], D[7], l0);   l1 = fmaf(m[7], D[13], l1);  l2 = fmaf(m[7], D[16], l2);  l3 = fmaf(m[7], D[17], l3);
    l0 = fmaf(m[8], D[8], l0);   l1 = fmaf(m[8], D[14], l1);  l2 = fmaf(m[8], D[17], l2);  l3 = fmaf(m[8], D[18], l3);
    l0 = fmaf(m[9], D[9], l0);   l1 = fmaf(m[9], D[15], l1);  l2 = fmaf(m[9], D[18], l2);  l3 = fmaf(m[9], D[19], l3);
#pragma unroll
    for (int k = 10; k < 20; ++k) l0 = fmaf(-m[k], D[k], l0);
    Lf[0] = l0; Lf[1] = l1; Lf[2] = l2; Lf[3] = l3;
    static constexpr int kRowB[6][3] = {{10, 11, 12}, {11, 13, 14}, {12, 14, 15}, {13, 16, 17}, {14, 17, 18},
        {15, 18, 19}};
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        float t = fmaf(m[0], D[4 + i], Lf[4 + i]);
        t = fmaf(-m[1], D[kRowB[i][0]], t);
        t = fmaf(-m[2], D[kRowB[i][1]], t);
        Lf[4 + i] = fmaf(-m[3], D[kRowB[i][2]], t);
    }
#pragma unroll
    for (int i = 10; i < 20; ++i) Lf[i] = fmaf(m[0], D[i], Lf[i]);
    return ri;
}

// FP64 half of a mixed-precision far pair (RHO): the angular-momentum
// correction at the FP64 separation, with 1/r refined from the FP32 estimate
// ri by one FP64 Newton step, and the octupole difference formed from the
// partner's FP64 mass m0 and scaled octupole oct[0..9].
__device__ __forceinline__ void mixed_correction(double xd, double yd, double zd, float ri, double m0,
    const double* oct, double inv_mc0, const double* mc, double* C) {
    double r2d = fma(zd, zd, fma(yd, yd, xd * xd));
    r2d = fmax(r2d, 1.0e-20);
    double rid = (double)ri;
    rid = fma(rid, fma(-(0.5 * r2d) * rid, rid, 0.5), rid);
    const double r2i = rid * rid;
    const double e1 = rid * r2i;
    const double e2 = (-3.0 * r2i) * e1;
    const double e3 = (-5.0 * r2i) * e2;
    const double n0c = m0 * inv_mc0;
    double a[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) a[k] = fma(-mc[k], n0c, oct[k]);
    am_correction(xd, yd, zd, e2, e3, a, C);
}

// Mixed precision far pair from an FP64 record (lane mapping 0): the
// separation is formed in FP64 and rounded to FP32, the record's components
// are rounded to FP32 for the derivative tensor and the 20-component
// contraction, which accumulate into the FP32 partial sums Lf of the current
// coarse offset (at most 8 pairs, added to the FP64 expansion by the caller).
// The angular-momentum correction stays FP64 (mixed_correction).
template <bool RHO>
__device__ __forceinline__ void m2l_far_mixed(const Rec& r, double X0, double X1, double X2, double inv_mc0,
    const double* mc, float* Lf, double* C) {
    const double xd = X0 - r.v[10].x, yd = X1 - r.v[10].y, zd = X2 - r.v[11].x;
    float m[20];
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        m[2 * i] = (float)r.v[i].x;
        m[2 * i + 1] = (float)r.v[i].y;
    }
    const float ri = m2l_fp32_core<RHO>((float)xd, (float)yd, (float)zd, m, Lf);
    if constexpr (RHO) {
        double oct[10];
#pragma unroll
        for (int k = 0; k < 10; ++k) oct[k] = (k & 1) ? r.v[5 + (k >> 1)].y : r.v[5 + (k >> 1)].x;
        mixed_correction(xd, yd, zd, ri, r.v[0].x, oct, inv_mc0, mc, C);
    }
}

// Mixed-precision partner record (192 bytes, 12 chunks of 16), written by the
// staging kernel beside the FP64 record for mixed-precision plans: chunks 0-4
// the factor-scaled multipole m0..m19 rounded to FP32, chunk 5 (Y0, Y1),
// chunk 6 (Y2, m0), chunks 7-11 the factor-scaled octupole m10..m19 in FP64.
constexpr int kRecmY01 = 5;
constexpr int kRecmY2M0 = 6;
constexpr int kRecmOct = 7;

// Mixed precision far pair from a mixed-precision record r (in LDS): no
// conversion of the multipole, the separation formed in FP64 and rounded.
template <bool RHO>
__device__ __forceinline__ void m2l_far_m32(const double2* __restrict__ r, double X0, double X1, double X2,
    double inv_mc0, const double* mc, float* Lf, double* C) {
    const double2 y01 = r[kRecmY01], y2m = r[kRecmY2M0];
    const double xd = X0 - y01.x, yd = X1 - y01.y, zd = X2 - y2m.x;
    float m[20];
    const float4* f4 = reinterpret_cast<const float4*>(r);
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        const float4 v = f4[i];
        m[4 * i] = v.x;
        m[4 * i + 1] = v.y;
        m[4 * i + 2] = v.z;
        m[4 * i + 3] = v.w;
    }
    const float ri = m2l_fp32_core<RHO>((float)xd, (float)yd, (float)zd, m, Lf);
    if constexpr (RHO) {
        double oct[10];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const double2 v = r[kRecmOct + k];
            oct[2 * k] = v.x;
            oct[2 * k + 1] = v.y;
        }
        mixed_correction(xd, yd, zd, ri, y2m.y, oct, inv_mc0, mc, C);
    }
}

template <bool RHO>
__device__ __forceinline__ void m2l_far(const double* __restrict__ p, double X0, double X1, double X2,
    double inv_mc0, const double* mc, double* L, double* C) {
    double m[20];
    load_record20(p, m);
    const double2 yv = reinterpret_cast<const double2*>(p)[10];
    const double x = X0 - yv.x, y = X1 - yv.y, z = X2 - p[22];
    DTensor t;
    d_tensor(x, y, z, t);
#pragma unroll
    for (int i = 0; i < 4; ++i) L[i] = fma(m[0], t.D[i], L[i]);
    if constexpr (!RHO) dipole_rows(m, t.D, L);
    contract_high<20>(m, t.D, L);
    if constexpr (RHO) {
        const double n0c = m[0] * inv_mc0;
        double a[10];
#pragma unroll
        for (int k = 0; k < 10; ++k) a[k] = fma(-mc[k], n0c, m[10 + k]);
        am_correction_t(t.d2, t.d3x00, t.d3x11, t.d3x22, t.d3x01, t.d3x02, t.d3x12, a, C);
    }
}

// One near pair with a monopole-only partner of mass mn at Y.
template <bool RHO>
__device__ __forceinline__ void m2l_near(double mn, double Y0, double Y1, double Y2, double X0, double X1,
    double X2, double inv_mc0, const double* mc, double* L, double* C) {
    const double x = X0 - Y0, y = X1 - Y1, z = X2 - Y2;
    DTensor t;
    d_tensor(x, y, z, t);
#pragma unroll
    for (int i = 0; i < 20; ++i) L[i] = fma(mn, t.D[i], L[i]);
    if constexpr (RHO) {
        const double n0c = mn * inv_mc0;
        double a[10];
#pragma unroll
        for (int k = 0; k < 10; ++k) a[k] = -(mc[k] * n0c);
        am_correction_t(t.d2, t.d3x00, t.d3x11, t.d3x22, t.d3x01, t.d3x02, t.d3x12, a, C);
    }
}

// One p2m pair: refined-neighbour multipole record r (held in registers)
// onto a leaf cell at X.
template <bool RHO>
__device__ __forceinline__ void p2m_pair_rec(const Rec& r, double X0, double X1, double X2, double* L, double* C) {
    double m[20];
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        m[2 * i] = r.v[i].x;
        m[2 * i + 1] = r.v[i].y;
    }
    const double x = X0 - r.v[10].x, y = X1 - r.v[10].y, z = X2 - r.v[11].x;
    DTensor t;
    d_tensor(x, y, z, t);
#pragma unroll
    for (int i = 0; i < 4; ++i) L[i] = fma(m[0], t.D[i], L[i]);
    if constexpr (!RHO) dipole_rows(m, t.D, L);
    contract_high<4>(m, t.D, L);
    if constexpr (RHO) am_correction_t(t.d2, t.d3x00, t.d3x11, t.d3x22, t.d3x01, t.d3x02, t.d3x12, m + 10, C);
}

// ---------------------------------------------------------------------------
// Kernel arguments
// ---------------------------------------------------------------------------
struct StageArgs {
    const double* M;        // [nslots][512][20]
    const double* com;      // [nslots][512][3]
    const double* mon;      // [nslots][512]
    const int32_t* kind;    // [nslots]
    const double* fxbase;   // [nslots][3] frame origin for leaf-cell centres
    const int32_t* foff;    // [nslots][3] frame offset of the slot's cell (0,0,0)
    double* rec;            // [nslots][512][24]
    double* recm;           // mixed-precision records [nslots][512][24], or null
    const int32_t* list;    // slots to stage (null: all nslots)
    double dx;
    int32_t nslots;         // slots to stage (length of list when given)
};

struct LevelArgs {
    const double* rec;      // partner records
    const double* com;      // input centres of mass (leaf targets)
    double* L;
    double* Lc;
    const int32_t* nbr;     // [nslots][27]
    const int32_t* kind;    // [nslots]
    const int2* items;      // (slot, flags)
    const Tables* tab;
    const uint64_t* mixbits;   // per coarse offset: byte k, bit j = far pair evaluated in FP32
    // mixed-precision plans, lane mappings 3 and 7: the mixed-precision
    // records, and per entry of the far schedule (Tables::sched) 1 where
    // every far pair of the entry's offset is evaluated in FP32
    const double* recm;
    const uint8_t* mixoff;
    const double* mon;      // [nslots][512] monopoles (p2p window of leaf targets)
    // leaf targets with refined neighbours: the p2m sums (L0..3, L_c, 0) of
    // each cell, [p2m item][512][8], written by p2m_kernel for leaf_kernel
    double* p2m_out;
    // Whole-tree solves (l2l != 0): the level's expansions are written, not
    // accumulated, as interactions + the parent cell's final expansion
    // re-expanded about the cell's centre of mass (and L_c + the parent's
    // L_c); leaf cells also get phi and g.  parent / octant per slot, pL /
    // pLc / pcom of the level above (null on level 0).
    const int32_t* parent;
    const int32_t* octant;
    const double* pL;
    const double* pLc;
    const double* pcom;
    double* g;
    double G;
    int32_t l2l;
    int32_t correct_am;
    int32_t leaf_rows;      // whole-tree solves: 0 = leaf cells get g (and L_c) only, no L rows
    // persistent form of lane mapping 7: 8 work counters, one per XCD
    // (cleared by queue_reset before each launch)
    uint32_t* queue;
    int32_t nitems;
    int32_t item0;          // first item of this launch
    double inv_dx;
    double neg_inv_dx;
};

struct RootArgs {
    const double* rec;      // records of the root slot (512 x 24)
    const Tables* tab;
    double* partial;        // [nchunks][512][24]
    int32_t per_chunk;
    int32_t nchunks;
};

struct RootReduceArgs {
    const double* partial;
    double* L;              // [512][20]
    double* Lc;             // [512][3]
    int32_t nchunks;
    int32_t rho;
    int32_t overwrite;      // whole-tree solve: L is written, not accumulated
};

// ---------------------------------------------------------------------------
// Tree passes: expansion shifts of taylor<4> (octotiger/taylor.hpp:222-311)
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int tmap2(int a, int b) {
    const int i = a < b ? a : b, j = a < b ? b : a;
    return 4 + i * 3 - (i * (i - 1)) / 2 + (j - i);
}

__host__ __device__ constexpr int tmap3(int a, int b, int c) {
    const int t[27] = {10, 11, 12, 11, 13, 14, 12, 14, 15, 11, 13, 14, 13, 16, 17, 14, 17, 18,
        12, 14, 15, 14, 17, 18, 15, 18, 19};
    return t[a * 9 + b * 3 + c];
}

// taylor::operator>>= : B = A moved from its centre by X (X = old - new centre).
__device__ __forceinline__ void multipole_shift(const double* A, const double* X, double* B) {
#pragma unroll
    for (int i = 0; i < 20; ++i) B[i] = A[i];
#pragma unroll
    for (int a = 0; a < 3; ++a) B[1 + a] += A[0] * X[a];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = a; b < 3; ++b) {
            B[tmap2(a, b)] += A[1 + a] * X[b] + X[a] * A[1 + b];
            B[tmap2(a, b)] += A[0] * X[a] * X[b];
        }
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = a; b < 3; ++b)
#pragma unroll
            for (int c = b; c < 3; ++c) {
                double t = B[tmap3(a, b, c)];
                t += A[tmap2(a, b)] * X[c] + A[tmap2(b, c)] * X[a] + A[tmap2(c, a)] * X[b];
                t += A[1 + a] * X[b] * X[c] + A[1 + b] * X[c] * X[a] + A[1 + c] * X[a] * X[b];
                t += A[0] * X[a] * X[b] * X[c];
                B[tmap3(a, b, c)] = t;
            }
}

// multipole_shift for A = (m, 0, ..., 0), a cell that holds only its mass:
// the terms of the general form that remain, each formed in the same order.
__device__ __forceinline__ void monopole_shift(double m, const double* X, double* B) {
    B[0] = m;
#pragma unroll
    for (int a = 0; a < 3; ++a) B[1 + a] = m * X[a];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = a; b < 3; ++b) B[tmap2(a, b)] = m * X[a] * X[b];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = a; b < 3; ++b)
#pragma unroll
            for (int c = b; c < 3; ++c) B[tmap3(a, b, c)] = m * X[a] * X[b] * X[c];
}

// taylor::operator<<= : B = local expansion A re-expanded at +X.
__device__ __forceinline__ void local_shift(const double* A, const double* X, double* B) {
#pragma unroll
    for (int i = 0; i < 20; ++i) B[i] = A[i];
#pragma unroll
    for (int a = 0; a < 3; ++a) B[0] += A[1 + a] * X[a];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) B[0] += A[tmap2(a, b)] * X[a] * X[b] * 0.5;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) B[1 + a] += A[tmap2(a, b)] * X[b];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c) B[0] += A[tmap3(a, b, c)] * X[a] * X[b] * X[c] * (1.0 / 6.0);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c) B[1 + a] += A[tmap3(a, b, c)] * X[b] * X[c] * 0.5;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int c = a; c < 3; ++c) B[tmap2(a, c)] += A[tmap3(a, b, c)] * X[b];
}

// parent cell holding coarse cell (ip, jp, kp) of a node in octant o
// (node_server::compute_fmm, src/node_server.cpp:613-636, 756-768)
__host__ __device__ __forceinline__ int parent_cell(int o, int ip, int jp, int kp) {
    return (ip + 4 * (o & 1)) * 64 + (jp + 4 * ((o >> 1) & 1)) * 8 + (kp + 4 * ((o >> 2) & 1));
}

struct M2MArgs {
    const double* M;        // [n][512][20] this level
    double* com;            // [n][512][3]  this level (leaf centres written for RHO)
    const double* mon;      // [n][512]
    const int32_t* kind;
    const int32_t* parent;
    const int32_t* octant;
    const double* xbase;    // [n][3]
    double* pM;             // parent level [np][512][20]
    double* pcom;           // parent level [np][512][3]
    double dx;
    int32_t nnodes;
    int32_t rho;
};


}   // namespace octofmm
