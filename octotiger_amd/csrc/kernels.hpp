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
// ---------------------------------------------------------------------------
struct DTensor {
    double D[20];
    double d2, d3;
};

// Derivatives of -1/r up to order 3 at dX (compute_d_factors).
__device__ __forceinline__ void d_tensor(double x, double y, double z, DTensor& t) {
    const double x00 = x * x, x11 = y * y, x22 = z * z;
    const double r2 = x00 + x11 + x22;
    const double r2inv = 1.0 / (r2 < 1.0e-20 ? 1.0e-20 : r2);
    const double d0 = -sqrt(r2inv);
    const double d1 = -d0 * r2inv;
    const double d2 = -3.0 * d1 * r2inv;
    const double d3 = -5.0 * d2 * r2inv;
    t.d2 = d2;
    t.d3 = d3;
    double* D = t.D;
    D[0] = d0;
    D[1] = x * d1;
    D[2] = y * d1;
    D[3] = z * d1;
    D[4] = d2 * x00 + d1;
    D[5] = d2 * (x * y);
    D[6] = d2 * (x * z);
    D[7] = d2 * x11 + d1;
    D[8] = d2 * (y * z);
    D[9] = d2 * x22 + d1;
    const double d2x = d2 * x, d2y = d2 * y, d2z = d2 * z;
    D[10] = d3 * x00 * x + 3.0 * d2x;
    D[11] = d3 * x00 * y + d2y;
    D[12] = d3 * x00 * z + d2z;
    D[13] = d3 * x * x11 + d2x;
    D[14] = d3 * x * (y * z);
    D[15] = d3 * x * x22 + d2x;
    D[16] = d3 * x11 * y + 3.0 * d2y;
    D[17] = d3 * x11 * z + d2z;
    D[18] = d3 * y * x22 + d2y;
    D[19] = d3 * x22 * z + 3.0 * d2z;
}

// Order-2/3 part of the contraction of a 20-component partner multipole m with
// D into the local expansion L (compute_interaction_multipole_non_rho).  With
// NCOMP == 4 only the potential and gradient rows are formed (p2m targets).
template <int NCOMP>
__device__ __forceinline__ void contract_high(const double* m, const double* D, double* L) {
    // multiplicity factors / 2 for order 2 and / 6 for order 3
    L[0] += m[4] * (D[4] * 0.5) + m[5] * D[5] + m[6] * D[6] + m[7] * (D[7] * 0.5) + m[8] * D[8] +
        m[9] * (D[9] * 0.5);
    L[1] += m[4] * (D[10] * 0.5) + m[5] * D[11] + m[6] * D[12] + m[7] * (D[13] * 0.5) + m[8] * D[14] +
        m[9] * (D[15] * 0.5);
    L[2] += m[4] * (D[11] * 0.5) + m[5] * D[13] + m[6] * D[14] + m[7] * (D[16] * 0.5) + m[8] * D[17] +
        m[9] * (D[18] * 0.5);
    L[3] += m[4] * (D[12] * 0.5) + m[5] * D[14] + m[6] * D[15] + m[7] * (D[17] * 0.5) + m[8] * D[18] +
        m[9] * (D[19] * 0.5);
    constexpr double s6 = 1.0 / 6.0;
    L[0] -= m[10] * (D[10] * s6) + m[11] * (D[11] * 0.5) + m[12] * (D[12] * 0.5) + m[13] * (D[13] * 0.5) +
        m[14] * D[14] + m[15] * (D[15] * 0.5) + m[16] * (D[16] * s6) + m[17] * (D[17] * 0.5) +
        m[18] * (D[18] * 0.5) + m[19] * (D[19] * s6);
    if constexpr (NCOMP == 20) {
        L[4] += m[0] * D[4] - m[1] * D[10] - m[2] * D[11] - m[3] * D[12];
        L[5] += m[0] * D[5] - m[1] * D[11] - m[2] * D[13] - m[3] * D[14];
        L[6] += m[0] * D[6] - m[1] * D[12] - m[2] * D[14] - m[3] * D[15];
        L[7] += m[0] * D[7] - m[1] * D[13] - m[2] * D[16] - m[3] * D[17];
        L[8] += m[0] * D[8] - m[1] * D[14] - m[2] * D[17] - m[3] * D[18];
        L[9] += m[0] * D[9] - m[1] * D[15] - m[2] * D[18] - m[3] * D[19];
#pragma unroll
        for (int i = 10; i < 20; ++i) L[i] += m[0] * D[i];
    }
}

// Dipole rows of the DRHODT kernels (compute_kernel_non_rho, :374-389).  The
// reference subtracts the whole first row of the second-order tensor from
// each gradient component; kept as is.
__device__ __forceinline__ void dipole_rows(const double* m, const double* D, double* L) {
    L[0] -= m[1] * D[1] + m[2] * D[2] + m[3] * D[3];
    L[1] -= m[1] * (D[4] + D[5] + D[6]);
    L[2] -= m[2] * (D[5] + D[7] + D[8]);
    L[3] -= m[3] * (D[6] + D[8] + D[9]);
}

// Angular-momentum correction from the order-4 tensor (compute_interaction_multipole_rho,
// :224-334; compute_interaction_p2m_rho, monopole_kernel_templates.hpp:69-187).
// n[k] = octupole difference (M2L) or partner octupole (p2m), k = 0..9.
__device__ __forceinline__ void am_correction(double x, double y, double z, double d2, double d3,
    const double* n, double* C) {
    const double x00 = x * x, x11 = y * y, x22 = z * z;
    const double d3x00 = d3 * x00, d3x11 = d3 * x11, d3x22 = d3 * x22;
    const double d3x01 = d3 * x * y, d3x02 = d3 * x * z, d3x12 = d3 * y * z;
    const double U0 = x00 * d3 + 3.0 * d2 + 5.0 * d3x00;
    const double U1 = 3.0 * d3x01;
    const double U2 = 3.0 * d3x02;
    const double U3 = d2 + d3x11 + d3x00;
    const double U4 = d3x12;
    const double U5 = d2 + d3x22 + d3x00;
    const double U7 = d3x02;
    const double U8 = d3x01;
    const double U10 = x11 * d3 + 3.0 * d2 + 5.0 * d3x11;
    const double U12 = d2 + d3x22 + d3x11;
    const double U13 = 3.0 * d3x12;
    const double U14 = x22 * d3 + 3.0 * d2 + 5.0 * d3x22;
    constexpr double s6 = 1.0 / 6.0;
    // factor_sixth = multiplicity / 6 : {1,3,3,3,6,3,1,3,3,1} / 6
    const double a0 = n[0] * s6, a1 = n[1] * 0.5, a2 = n[2] * 0.5, a3 = n[3] * 0.5, a4 = n[4],
                 a5 = n[5] * 0.5, a6 = n[6] * s6, a7 = n[7] * 0.5, a8 = n[8] * 0.5, a9 = n[9] * s6;
    C[0] -= a0 * U0 + a1 * U1 + a2 * U2 + a3 * U3 + a4 * U4 + a5 * U5 + a6 * U1 + a7 * U7 + a8 * U8 +
        a9 * U2;
    C[1] -= a0 * U1 + a1 * U3 + a2 * U4 + a3 * U1 + a4 * U7 + a5 * U8 + a6 * U10 + a7 * U13 + a8 * U12 +
        a9 * U13;
    C[2] -= a0 * U2 + a1 * U4 + a2 * U5 + a3 * U7 + a4 * U8 + a5 * U2 + a6 * U13 + a7 * U12 + a8 * U13 +
        a9 * U14;
}

// One far M2L pair: partner record p (m0, M1..M19, Y, mon) onto target X.
template <bool RHO>
__device__ __forceinline__ void m2l_far(const double* __restrict__ p, double X0, double X1, double X2,
    double mc0, const double* mc, double* L, double* C) {
    double m[20];
    const double2* p2 = reinterpret_cast<const double2*>(p);
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        const double2 v = p2[i];
        m[2 * i] = v.x;
        m[2 * i + 1] = v.y;
    }
    const double2 yv = p2[10];
    const double x = X0 - yv.x, y = X1 - yv.y, z = X2 - p[22];
    DTensor t;
    d_tensor(x, y, z, t);
    L[0] += m[0] * t.D[0];
    L[1] += m[0] * t.D[1];
    L[2] += m[0] * t.D[2];
    L[3] += m[0] * t.D[3];
    if constexpr (!RHO) dipole_rows(m, t.D, L);
    contract_high<20>(m, t.D, L);
    if constexpr (RHO) {
        const double n0c = m[0] / mc0;
        double n[10];
#pragma unroll
        for (int k = 0; k < 10; ++k) n[k] = m[10 + k] - mc[k] * n0c;
        am_correction(x, y, z, t.d2, t.d3, n, C);
    }
}

// One near pair with a monopole-only partner of mass mn at Y.
template <bool RHO>
__device__ __forceinline__ void m2l_near(double mn, double Y0, double Y1, double Y2, double X0, double X1,
    double X2, double mc0, const double* mc, double* L, double* C) {
    const double x = X0 - Y0, y = X1 - Y1, z = X2 - Y2;
    DTensor t;
    d_tensor(x, y, z, t);
#pragma unroll
    for (int i = 0; i < 20; ++i) L[i] += mn * t.D[i];
    if constexpr (RHO) {
        const double n0c = mn / mc0;
        double n[10];
#pragma unroll
        for (int k = 0; k < 10; ++k) n[k] = -(mc[k] * n0c);
        am_correction(x, y, z, t.d2, t.d3, n, C);
    }
}

// One p2m pair: refined-neighbour multipole record p onto a leaf cell at X.
template <bool RHO>
__device__ __forceinline__ void p2m_pair(const double* __restrict__ p, double X0, double X1, double X2,
    double* L, double* C) {
    double m[20];
    const double2* p2 = reinterpret_cast<const double2*>(p);
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        const double2 v = p2[i];
        m[2 * i] = v.x;
        m[2 * i + 1] = v.y;
    }
    const double x = X0 - p[20], y = X1 - p[21], z = X2 - p[22];
    DTensor t;
    d_tensor(x, y, z, t);
    L[0] += m[0] * t.D[0];
    L[1] += m[0] * t.D[1];
    L[2] += m[0] * t.D[2];
    L[3] += m[0] * t.D[3];
    if constexpr (!RHO) dipole_rows(m, t.D, L);
    contract_high<4>(m, t.D, L);
    if constexpr (RHO) am_correction(x, y, z, t.d2, t.d3, m + 10, C);
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
    double dx;
    int32_t nslots;
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
    int32_t nitems;
    int32_t pad;
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
};

}   // namespace octofmm
