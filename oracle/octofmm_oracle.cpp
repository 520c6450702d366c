// octofmm_oracle.cpp -- CPU FP64 oracle for the FMM gravity interaction path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: a scalar,
// FP64 restatement of Octo-Tiger's interaction kernels.  Only tests/, the
// smoke() check and the cpu_baseline leg of bench.py may load it, and only as
// the checker or the timed CPU baseline -- never as a product code path.
//
// What it restates (reference paths relative to the Octo-Tiger tree):
//   * stencils       src/multipole_interactions/util/calculate_stencil.cpp:23-161
//                    src/monopole_interactions/util/calculate_stencil.cpp:19-177
//                    octotiger/common_kernel/helper.hpp:20-31 (reciprocal_distance)
//   * pair math      octotiger/multipole_interactions/kernel/compute_kernel_templates.hpp:67-391
//                    octotiger/monopole_interactions/kernel/monopole_kernel_templates.hpp:16-239
//   * AoS->SoA       octotiger/multipole_interactions/legacy/multipole_interaction_interface.hpp:88-281
//                    octotiger/monopole_interactions/legacy/monopole_interaction_interface.hpp:64-143
//                    octotiger/monopole_interactions/legacy/p2m_interaction_interface.hpp:102-160
//                    octotiger/common_kernel/interactions_iterators.hpp:19-128
//   * kernels        octotiger/multipole_interactions/kernel/kokkos_kernel.hpp:137-321 (non-root RHO),
//                    :474-633 (non-root DRHODT), :324-471 / :637-764 (root),
//                    octotiger/monopole_interactions/kernel/kokkos_kernel.hpp:143-320 (p2p),
//                    :517-942 (p2m); all walked with SIMD width 1, i.e. in the
//                    per-cell order the Kokkos host executor uses.
//   * write-back     octotiger/multipole_interactions/kernel/kokkos_kernel.hpp:944-983,
//                    octotiger/monopole_interactions/kernel/kokkos_kernel.hpp:1279-1353
//
// Built with -ffp-contract=off so every operation rounds as written, in the
// reference's operation order; its pair functions therefore agree bit for bit
// with the reference's own templates compiled the same way.
#include "octofmm.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int INX = 8;
constexpr double THETA_FLOOR = 0.34;                               // CMakeLists.txt:66
constexpr int STENCIL_WIDTH_HELPER = static_cast<int>(2.0 / THETA_FLOOR);
constexpr int STENCIL_WIDTH =
    (STENCIL_WIDTH_HELPER % 2 == 0) ? STENCIL_WIDTH_HELPER + 1 : STENCIL_WIDTH_HELPER;
constexpr int STENCIL_MIN = -STENCIL_WIDTH;
constexpr int STENCIL_MAX = STENCIL_WIDTH;
constexpr int STENCIL_INX = 2 * STENCIL_WIDTH_HELPER + 1;
constexpr int FULL_STENCIL = STENCIL_INX * STENCIL_INX * STENCIL_INX;
constexpr int PAD_DEPTH = INX;                                    // INNER_CELLS_PADDING_DEPTH
constexpr int PAD_OFFSET = PAD_DEPTH - STENCIL_MAX;               // PADDING_OFFSET
constexpr int PSTRIDE = INX + 2 * STENCIL_MAX;                    // PADDED_STRIDE
constexpr int ENTRIES = PSTRIDE * PSTRIDE * PSTRIDE;
constexpr int SOA_PAD = 19;
constexpr int CLEN = ENTRIES + SOA_PAD;                           // component_length
constexpr int NCELL = INX * INX * INX;

static_assert(STENCIL_INX == 11 && PSTRIDE == 18, "geometry");

// multiplicity factors of the symmetric tensors (compute_kernel_templates.hpp:23-35)
const double kFactor[20] = {1, 1, 1, 1, 1, 2, 2, 1, 2, 1, 1, 3, 3, 3, 6, 3, 1, 3, 3, 1};
inline double fhalf(int i) { return kFactor[i] / 2.0; }
inline double fsixth(int i) { return kFactor[i] / 6.0; }

inline int padded_flat(int x, int y, int z) {
    return (x - PAD_OFFSET) * PSTRIDE * PSTRIDE + (y - PAD_OFFSET) * PSTRIDE + (z - PAD_OFFSET);
}
inline int inner_flat(int x, int y, int z) { return x * INX * INX + y * INX + z; }
inline int coarse(int i) { return ((i + INX) >> 1) - INX / 2; }   // multiindex::transform_coarse

// geo::direction raw index -> components (geometry.hpp:114-125)
inline void dir_components(int raw, int d[3]) {
    d[0] = raw % 3 - 1;
    d[1] = (raw / 3) % 3 - 1;
    d[2] = raw / 9 - 1;
}
inline int flat_to_raw(int flat) { return flat < 13 ? flat : flat + 1; }

// ---------------------------------------------------------------------------
// Stencils
// ---------------------------------------------------------------------------
double reciprocal_distance(long long i0, long long i1, long long i2, long long j0, long long j1,
    long long j2) {
    const double t = static_cast<double>((i0 - j0) * (i0 - j0) + (i1 - j1) * (i1 - j1) +
        (i2 - j2) * (i2 - j2));
    if (t > 0.0) return 1.0 / std::sqrt(t);
    return 1.0e+10;
}

struct Elem {
    int x, y, z;
    bool phase;   // true: multipole (far) partner, false: monopole-only (near)
};

// Union over the 8 parity positions of the offsets whose coarse parents pass
// the opening criterion; phase true iff the fine distance is outside 1/theta.
bool build_multipole_stencil(double theta, std::vector<Elem>& out, bool with_phase) {
    std::vector<Elem> per[8];
    for (int i0 = 0; i0 < 2; ++i0)
        for (int i1 = 0; i1 < 2; ++i1)
            for (int i2 = 0; i2 < 2; ++i2) {
                std::vector<Elem>& st = per[i0 * 4 + i1 * 2 + i2];
                for (int j0 = i0 - INX; j0 < i0 + INX; ++j0)
                    for (int j1 = i1 - INX; j1 < i1 + INX; ++j1)
                        for (int j2 = i2 - INX; j2 < i2 + INX; ++j2) {
                            if (i0 == j0 && i1 == j1 && i2 == j2) continue;
                            const int c_i0 = (i0 + INX) / 2 - INX / 2, c_i1 = (i1 + INX) / 2 - INX / 2,
                                      c_i2 = (i2 + INX) / 2 - INX / 2;
                            const int c_j0 = (j0 + INX) / 2 - INX / 2, c_j1 = (j1 + INX) / 2 - INX / 2,
                                      c_j2 = (j2 + INX) / 2 - INX / 2;
                            const double th_f = reciprocal_distance(i0, i1, i2, j0, j1, j2);
                            const double th_c =
                                reciprocal_distance(c_i0, c_i1, c_i2, c_j0, c_j1, c_j2);
                            if (th_c > theta && (!with_phase || th_f <= theta)) {
                                st.push_back({j0 - i0, j1 - i1, j2 - i2, true});
                            } else if (th_c > theta) {
                                st.push_back({j0 - i0, j1 - i1, j2 - i2, false});
                            }
                        }
            }
    out.clear();
    for (int i = 0; i < 8; ++i)
        for (const Elem& e : per[i]) {
            bool found = false;
            for (const Elem& s : out)
                if (s.x == e.x && s.y == e.y && s.z == e.z) {
                    found = true;
                    break;
                }
            if (!found) out.push_back(e);
        }
    int mnx = 0, mxx = 0, mny = 0, mxy = 0, mnz = 0, mxz = 0;
    for (const Elem& e : out) {
        mnx = e.x < mnx ? e.x : mnx; mxx = e.x > mxx ? e.x : mxx;
        mny = e.y < mny ? e.y : mny; mxy = e.y > mxy ? e.y : mxy;
        mnz = e.z < mnz ? e.z : mnz; mxz = e.z > mxz ? e.z : mxz;
    }
    if (mxz != mxy || mxy != mxx || mnz != mny || mny != mnx || mnz != -mxz) return false;
    if (STENCIL_WIDTH < mxx) return false;
    return true;
}

inline int stencil_index(int sx, int sy, int sz) {
    return (sx + STENCIL_MAX) * STENCIL_INX * STENCIL_INX + (sy + STENCIL_MAX) * STENCIL_INX +
        (sz + STENCIL_MAX);
}

struct Stencils {
    uint8_t mmask[FULL_STENCIL];   // multipole masks
    uint8_t mind[FULL_STENCIL];    // multipole phase indicators (default true)
    uint8_t pmask[FULL_STENCIL];   // monopole (p2p) masks
    double four[FULL_STENCIL][4];  // p2p constants
};

bool build_stencils(double theta, Stencils& s) {
    if (!(theta >= THETA_FLOOR)) return false;
    std::vector<Elem> st;
    if (!build_multipole_stencil(theta, st, true)) return false;
    for (int i = 0; i < FULL_STENCIL; ++i) {
        s.mmask[i] = 0;
        s.mind[i] = 1;
    }
    for (const Elem& e : st) {
        const int idx = stencil_index(e.x, e.y, e.z);
        s.mmask[idx] = 1;
        if (!e.phase) s.mind[idx] = 0;
    }
    std::vector<Elem> mono;
    if (!build_multipole_stencil(theta, mono, false)) return false;
    for (int i = 0; i < FULL_STENCIL; ++i) {
        s.pmask[i] = 0;
        for (int c = 0; c < 4; ++c) s.four[i][c] = 0.0;
    }
    for (const Elem& e : mono) {
        const int idx = stencil_index(e.x, e.y, e.z);
        s.pmask[idx] = 1;
        const double x = e.x, y = e.y, z = e.z;
        const double r = std::sqrt(x * x + y * y + z * z);
        const double r3 = r * r * r;
        s.four[idx][0] = -1.0 / r;
        s.four[idx][1] = x / r3;
        s.four[idx][2] = y / r3;
        s.four[idx][3] = z / r3;
    }
    return true;
}

// ---------------------------------------------------------------------------
// Pair arithmetic (compute_kernel_templates.hpp, monopole_kernel_templates.hpp)
// ---------------------------------------------------------------------------
struct DF {
    double d2, d3, X00, X11, X22;
    double D[20];
};

inline void d_factors(const double dX[3], DF& f) {
    f.X00 = dX[0] * dX[0];
    f.X11 = dX[1] * dX[1];
    f.X22 = dX[2] * dX[2];
    const double r2 = f.X00 + f.X11 + f.X22;
    const double r2inv = 1.0 / (r2 < 1.0e-20 ? 1.0e-20 : r2);
    const double d0 = -std::sqrt(r2inv);
    const double d1 = -d0 * r2inv;
    f.d2 = -3.0 * d1 * r2inv;
    f.d3 = -5.0 * f.d2 * r2inv;
    const double d2 = f.d2, d3 = f.d3;
    double* D = f.D;
    D[0] = d0;
    D[1] = dX[0] * d1;
    D[2] = dX[1] * d1;
    D[3] = dX[2] * d1;
    const double X12 = dX[1] * dX[2], X01 = dX[0] * dX[1], X02 = dX[0] * dX[2];
    D[4] = d2 * f.X00 + d1;
    D[5] = d2 * X01;
    D[6] = d2 * X02;
    D[7] = d2 * f.X11 + d1;
    D[8] = d2 * X12;
    D[9] = d2 * f.X22 + d1;
    const double d2x0 = d2 * dX[0], d2x1 = d2 * dX[1], d2x2 = d2 * dX[2];
    D[10] = d3 * f.X00 * dX[0] + 3.0 * d2x0;
    D[11] = d3 * f.X00 * dX[1] + d2 * dX[1];
    D[12] = d3 * f.X00 * dX[2] + d2 * dX[2];
    D[13] = d3 * dX[0] * f.X11 + d2 * dX[0];
    D[14] = d3 * dX[0] * X12;
    D[15] = d3 * dX[0] * f.X22 + d2x0;
    D[16] = d3 * f.X11 * dX[1] + 3.0 * d2x1;
    D[17] = d3 * f.X11 * dX[2] + d2 * dX[2];
    D[18] = d3 * dX[1] * f.X22 + d2 * dX[1];
    D[19] = d3 * f.X22 * dX[2] + 3.0 * d2x2;
}

// D index of the third-order tensor e_g + (second-order pair j), g = 1..3,
// j = 4..9 : used by the gradient rows of the contraction.
const int kGradIdx[3][6] = {
    {10, 11, 12, 13, 14, 15},   // x * {xx, xy, xz, yy, yz, zz}
    {11, 13, 14, 16, 17, 18},   // y * ...
    {12, 14, 15, 17, 18, 19},   // z * ...
};

// compute_interaction_multipole_non_rho (compute_kernel_templates.hpp:141-222);
// `ncomp` = 20 for M2L, 4 for p2m (compute_interaction_p2m_non_rho).
inline void contract_quadrupole_octupole(const double* m, double* L, const double* D, int ncomp) {
    for (int j = 4; j < 10; ++j) {
        L[0] += m[j] * (D[j] * fhalf(j));
        for (int g = 0; g < 3; ++g) L[1 + g] += m[j] * (D[kGradIdx[g][j - 4]] * fhalf(j));
    }
    for (int j = 10; j < 20; ++j) L[0] -= m[j] * (D[j] * fsixth(j));
    if (ncomp == 4) return;
    for (int i = 4; i < 10; ++i) L[i] += m[0] * D[i];
    for (int g = 0; g < 3; ++g)
        for (int i = 4; i < 10; ++i) L[i] -= m[1 + g] * D[kGradIdx[g][i - 4]];
    for (int i = 10; i < 20; ++i) L[i] += m[0] * D[i];
}

// Angular-momentum correction (compute_interaction_multipole_rho, :224-334 and
// compute_interaction_p2m_rho, monopole_kernel_templates.hpp:69-187).  For M2L
// the octupole difference is n0 = m_p - m_c * (m_p0 / m_c0); for p2m the
// partner octupole enters directly (m_cell == nullptr).
inline void correction(const DF& f, const double dX[3], const double* mp, const double* mc,
    double* corr) {
    const double d2 = f.d2, d3 = f.d3;
    const double n0c = mc ? mp[0] / mc[0] : 0.0;
    auto n0 = [&](int k) { return mc ? mp[k] - mc[k] * n0c : mp[k]; };
    double U[15];
    const double d3X00 = d3 * f.X00, d3X11 = d3 * f.X11, d3X22 = d3 * f.X22;
    const double d3X01 = d3 * dX[0] * dX[1], d3X02 = d3 * dX[0] * dX[2], d3X12 = d3 * dX[1] * dX[2];
    U[0] = dX[0] * dX[0] * d3 + 2.0 * d2;
    U[0] += d2;
    U[0] += 5.0 * d3X00;
    U[1] = 3.0 * d3X01;
    U[2] = 3.0 * d3X02;
    U[3] = d2;
    U[3] += d3X11;
    U[3] += d3 * f.X00;
    U[4] = d3X12;
    U[5] = d2;
    U[5] += d3X22;
    U[5] += d3X00;
    U[6] = 3.0 * d3X01;
    U[7] = d3 * dX[0] * dX[2];
    U[8] = d3 * dX[0] * dX[1];
    U[9] = 3.0 * d3X02;
    U[10] = dX[1] * dX[1] * d3 + 2.0 * d2;
    U[10] += d2;
    U[10] += 5.0 * d3X11;
    U[11] = 3.0 * d3X12;
    U[12] = d2;
    U[12] += d3X22;
    U[12] += d3X11;
    U[13] = 3.0 * d3X12;
    U[14] = dX[2] * dX[2] * d3 + 2.0 * d2;
    U[14] += d2;
    U[14] += 5.0 * d3X22;
    // rows of U used by octupole component 10+k for the x, y and z correction
    static const int kRows[10][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}, {3, 6, 7}, {4, 7, 8},
        {5, 8, 9}, {6, 10, 11}, {7, 11, 12}, {8, 12, 13}, {9, 13, 14}};
    for (int k = 0; k < 10; ++k) {
        const double t = n0(10 + k);
        for (int a = 0; a < 3; ++a) corr[a] -= t * (U[kRows[k][a]] * fsixth(10 + k));
    }
}

// compute_kernel_rho / compute_kernel_non_rho (:336-391)
inline void pair_m2l(const double X[3], const double Y[3], const double* mp, const double* mc,
    double* L, double* corr, bool rho) {
    const double dX[3] = {X[0] - Y[0], X[1] - Y[1], X[2] - Y[2]};
    DF f;
    d_factors(dX, f);
    const double* D = f.D;
    for (int i = 0; i < 4; ++i) L[i] += mp[0] * D[i];
    if (!rho) {
        for (int g = 0; g < 3; ++g) {
            L[0] -= mp[1 + g] * D[1 + g];
            // the reference subtracts the whole gradient row of D from the
            // same component (compute_kernel_templates.hpp:378-389)
            static const int kRow[3][3] = {{4, 5, 6}, {5, 7, 8}, {6, 8, 9}};
            for (int a = 0; a < 3; ++a) L[1 + g] -= mp[1 + g] * D[kRow[g][a]];
        }
    }
    contract_quadrupole_octupole(mp, L, D, 20);
    if (rho) correction(f, dX, mp, mc, corr);
}

// compute_kernel_p2m_rho / _non_rho (monopole_kernel_templates.hpp:188-239)
inline void pair_p2m(const double X[3], const double Y[3], const double* mp, double* L,
    double* corr, bool rho) {
    const double dX[3] = {X[0] - Y[0], X[1] - Y[1], X[2] - Y[2]};
    DF f;
    d_factors(dX, f);
    const double* D = f.D;
    for (int i = 0; i < 4; ++i) L[i] += mp[0] * D[i];
    if (!rho) {
        static const int kRow[3][3] = {{4, 5, 6}, {5, 7, 8}, {6, 8, 9}};
        for (int g = 0; g < 3; ++g) {
            L[0] -= mp[1 + g] * D[1 + g];
            for (int a = 0; a < 3; ++a) L[1 + g] -= mp[1 + g] * D[kRow[g][a]];
        }
    }
    contract_quadrupole_octupole(mp, L, D, 4);
    if (rho) correction(f, dX, mp, nullptr, corr);
}

// ---------------------------------------------------------------------------
// AoS -> SoA staging of the 18^3 window
// ---------------------------------------------------------------------------
struct Window {
    std::vector<double> mon, M, com;
    Window() : mon(ENTRIES, 0.0), M(20 * CLEN, 0.0), com(3 * CLEN, 0.0) {}
    void set_M(int flat, const double* v) {
        for (int c = 0; c < 20; ++c) M[c * CLEN + flat] = v ? v[c] : 0.0;
    }
    void set_com(int flat, const double* v) {
        for (int c = 0; c < 3; ++c) com[c * CLEN + flat] = v ? v[c] : 0.0;
    }
};

// iterate_inner_cells_padding (interactions_iterators.hpp:65-128): visits the
// neighbour cells of direction d that fall inside the window; hands over the
// window-frame index (0..23) and the neighbour-local flat index.
template <typename F> void for_padding_cells(const int d[3], F&& f) {
    int s[3], e[3], off[3];
    for (int k = 0; k < 3; ++k) {
        s[k] = d[k] == -1 ? INX - STENCIL_MAX : 0;
        e[k] = d[k] == 1 ? STENCIL_MAX : INX;
        off[k] = d[k] == 1 ? 2 * INX : (d[k] == 0 ? INX : 0);
    }
    for (int i0 = s[0]; i0 < e[0]; ++i0)
        for (int i1 = s[1]; i1 < e[1]; ++i1)
            for (int i2 = s[2]; i2 < e[2]; ++i2)
                f(i0 + off[0], i1 + off[1], i2 + off[2], inner_flat(i0, i1, i2));
}

bool in_window(int x, int y, int z) {
    return x >= PAD_OFFSET && x < PAD_OFFSET + PSTRIDE && y >= PAD_OFFSET &&
        y < PAD_OFFSET + PSTRIDE && z >= PAD_OFFSET && z < PAD_OFFSET + PSTRIDE;
}

// multipole_interactions::update_input
void stage_multipole(const octofmm_subgrid* in, Window& w) {
    const bool root = in->use_root_stencil != 0;
    for (int i0 = 0; i0 < INX; ++i0)
        for (int i1 = 0; i1 < INX; ++i1)
            for (int i2 = 0; i2 < INX; ++i2) {
                const int fl = padded_flat(i0 + PAD_DEPTH, i1 + PAD_DEPTH, i2 + PAD_DEPTH);
                const int u = inner_flat(i0, i1, i2);
                w.set_M(fl, in->M + 20 * u);
                w.set_com(fl, in->com + 3 * u);
                if (!root) w.mon[fl] = 0.0;
            }
    if (root) return;
    const double dx = in->dx;
    for (int flat = 0; flat < 26; ++flat) {
        int d[3];
        dir_components(flat_to_raw(flat), d);
        const octofmm_neighbor& nb = in->neighbors[flat];
        auto zero_all = [&](int x, int y, int z, int) {
            const int fl = padded_flat(x, y, z);
            w.set_M(fl, nullptr);
            w.set_com(fl, nullptr);
            w.mon[fl] = 0.0;
        };
        auto frame_of = [&](int u, int& x, int& y, int& z) {
            x = u / 64 + PAD_DEPTH + d[0] * PAD_DEPTH;
            y = (u / 8) % 8 + PAD_DEPTH + d[1] * PAD_DEPTH;
            z = u % 8 + PAD_DEPTH + d[2] * PAD_DEPTH;
        };
        if (nb.kind == OCTOFMM_NEIGHBOR_MULTIPOLE && nb.M && nb.com) {
            if (nb.count == NCELL && !nb.cells) {
                for_padding_cells(d, [&](int x, int y, int z, int u) {
                    const int fl = padded_flat(x, y, z);
                    w.set_M(fl, nb.M + 20 * u);
                    w.set_com(fl, nb.com + 3 * u);
                    w.mon[fl] = 0.0;
                });
            } else {
                for_padding_cells(d, zero_all);
                for (int c = 0; c < nb.count; ++c) {
                    int x, y, z;
                    frame_of(nb.cells[c], x, y, z);
                    if (!in_window(x, y, z)) continue;
                    const int fl = padded_flat(x, y, z);
                    w.set_M(fl, nb.M + 20 * c);
                    w.set_com(fl, nb.com + 3 * c);
                }
            }
        } else if (nb.kind == OCTOFMM_NEIGHBOR_MONOPOLE && nb.mon) {
            // leaf neighbour: monopoles seen at the geometric cell centres
            auto centre = [&](int x, int y, int z, int fl) {
                const double e[3] = {x * dx + in->xbase[0] - INX * dx, y * dx + in->xbase[1] - INX * dx,
                    z * dx + in->xbase[2] - INX * dx};
                w.set_com(fl, e);
                w.set_M(fl, nullptr);
            };
            if (nb.count == NCELL && !nb.cells) {
                for_padding_cells(d, [&](int x, int y, int z, int u) {
                    const int fl = padded_flat(x, y, z);
                    centre(x, y, z, fl);
                    w.mon[fl] = nb.mon[u];
                });
            } else {
                for_padding_cells(d, [&](int x, int y, int z, int) {
                    const int fl = padded_flat(x, y, z);
                    centre(x, y, z, fl);
                    w.mon[fl] = 0.0;
                });
                for (int c = 0; c < nb.count; ++c) {
                    int x, y, z;
                    frame_of(nb.cells[c], x, y, z);
                    if (!in_window(x, y, z)) continue;
                    w.mon[padded_flat(x, y, z)] = nb.mon[c];
                }
            }
        } else {
            for_padding_cells(d, zero_all);
        }
    }
}

// monopole_interactions::update_input
void stage_monopole(const octofmm_subgrid* in, std::vector<double>& mon) {
    for (int i0 = 0; i0 < INX; ++i0)
        for (int i1 = 0; i1 < INX; ++i1)
            for (int i2 = 0; i2 < INX; ++i2)
                mon[padded_flat(i0 + PAD_DEPTH, i1 + PAD_DEPTH, i2 + PAD_DEPTH)] =
                    in->mon[inner_flat(i0, i1, i2)];
    for (int flat = 0; flat < 26; ++flat) {
        int d[3];
        dir_components(flat_to_raw(flat), d);
        const octofmm_neighbor& nb = in->neighbors[flat];
        if (nb.kind == OCTOFMM_NEIGHBOR_MONOPOLE && nb.mon) {
            if (nb.count == NCELL && !nb.cells) {
                for_padding_cells(d, [&](int x, int y, int z, int u) {
                    mon[padded_flat(x, y, z)] = nb.mon[u];
                });
            } else {
                for_padding_cells(d, [&](int x, int y, int z, int) { mon[padded_flat(x, y, z)] = 0.0; });
                for (int c = 0; c < nb.count; ++c) {
                    const int u = nb.cells[c];
                    const int x = u / 64 + PAD_DEPTH + d[0] * PAD_DEPTH;
                    const int y = (u / 8) % 8 + PAD_DEPTH + d[1] * PAD_DEPTH;
                    const int z = u % 8 + PAD_DEPTH + d[2] * PAD_DEPTH;
                    if (!in_window(x, y, z)) continue;
                    mon[padded_flat(x, y, z)] = nb.mon[c];
                }
            }
        } else {
            for_padding_cells(d, [&](int x, int y, int z, int) { mon[padded_flat(x, y, z)] = 0.0; });
        }
    }
}

// ---------------------------------------------------------------------------
// Kernels (scalar walk of the Kokkos host kernels)
// ---------------------------------------------------------------------------
void kernel_nonroot(const Window& w, const Stencils& s, double theta, bool rho,
    double (*res)[20], double (*cor)[3]) {
    const double th_rec2 = (1.0 / theta) * (1.0 / theta);
    for (int ix = 0; ix < INX; ++ix)
        for (int iy = 0; iy < INX; ++iy)
            for (int iz = 0; iz < INX; ++iz) {
                const int cx = ix + PAD_DEPTH, cy = iy + PAD_DEPTH, cz = iz + PAD_DEPTH;
                const int cf = padded_flat(cx, cy, cz);
                const int ccx = coarse(cx), ccy = coarse(cy), ccz = coarse(cz);
                double mc[20], X[3], L[20] = {0}, C[3] = {0};
                for (int c = 0; c < 20; ++c) mc[c] = w.M[c * CLEN + cf];
                for (int c = 0; c < 3; ++c) X[c] = w.com[c * CLEN + cf];
                for (int x = 0; x < STENCIL_INX; ++x) {
                    const int px = cx + x + STENCIL_MIN;
                    const int dcx = (ccx - coarse(px)) * (ccx - coarse(px));
                    for (int y = 0; y < STENCIL_INX; ++y) {
                        const int py = cy + y + STENCIL_MIN;
                        const int dcy = (ccy - coarse(py)) * (ccy - coarse(py));
                        for (int sz = STENCIL_MIN; sz <= STENCIL_MAX; ++sz) {
                            const int idx = x * STENCIL_INX * STENCIL_INX + y * STENCIL_INX +
                                (sz - STENCIL_MIN);
                            if (!s.mmask[idx]) continue;
                            const int pz = cz + sz;
                            const int dcz = (ccz - coarse(pz)) * (ccz - coarse(pz));
                            if (!(static_cast<double>(dcx + dcy + dcz) < th_rec2)) continue;
                            const int pf = padded_flat(px, py, pz);
                            const double Y[3] = {w.com[pf], w.com[CLEN + pf], w.com[2 * CLEN + pf]};
                            double mp[20];
                            const bool far = s.mind[idx] != 0;
                            mp[0] = w.mon[pf] + (far ? w.M[pf] : 0.0);
                            for (int c = 1; c < 20; ++c) mp[c] = far ? w.M[c * CLEN + pf] : 0.0;
                            pair_m2l(X, Y, mp, mc, L, C, rho);
                        }
                    }
                }
                const int u = inner_flat(ix, iy, iz);
                for (int c = 0; c < 20; ++c) res[u][c] = L[c];
                for (int c = 0; c < 3; ++c) cor[u][c] = C[c];
            }
}

void kernel_root(const Window& w, const Stencils& s, bool rho, double (*res)[20], double (*cor)[3]) {
    for (int ix = 0; ix < INX; ++ix)
        for (int iy = 0; iy < INX; ++iy)
            for (int iz = 0; iz < INX; ++iz) {
                const int cf = padded_flat(ix + PAD_DEPTH, iy + PAD_DEPTH, iz + PAD_DEPTH);
                double mc[20], X[3], L[20] = {0}, C[3] = {0};
                for (int c = 0; c < 20; ++c) mc[c] = w.M[c * CLEN + cf];
                for (int c = 0; c < 3; ++c) X[c] = w.com[c * CLEN + cf];
                for (int x = 0; x < INX; ++x)
                    for (int y = 0; y < INX; ++y)
                        for (int z = 0; z < INX; ++z) {
                            const int sx = x - ix, sy = y - iy, sz = z - iz;
                            // near offsets and the cell itself are excluded
                            // (kokkos_kernel.hpp:385-412)
                            if (sx >= STENCIL_MIN && sx <= STENCIL_MAX && sy >= STENCIL_MIN &&
                                sy <= STENCIL_MAX && sz >= STENCIL_MIN && sz <= STENCIL_MAX) {
                                if (!s.mind[stencil_index(sx, sy, sz)] || (sx == 0 && sy == 0 && sz == 0))
                                    continue;
                            }
                            const int pf = padded_flat(x + INX, y + INX, z + INX);
                            const double Y[3] = {w.com[pf], w.com[CLEN + pf], w.com[2 * CLEN + pf]};
                            double mp[20];
                            for (int c = 0; c < 20; ++c) mp[c] = w.M[c * CLEN + pf];
                            pair_m2l(X, Y, mp, mc, L, C, rho);
                        }
                const int u = inner_flat(ix, iy, iz);
                for (int c = 0; c < 20; ++c) res[u][c] = L[c];
                for (int c = 0; c < 3; ++c) cor[u][c] = C[c];
            }
}

void kernel_p2p(const std::vector<double>& mon, const Stencils& s, double theta, double dx,
    double (*res)[4]) {
    const double th_rec2 = (1.0 / theta) * (1.0 / theta);
    const double dcomp[2] = {1.0 / dx, -1.0 / dx};
    for (int ix = 0; ix < INX; ++ix)
        for (int iy = 0; iy < INX; ++iy)
            for (int iz = 0; iz < INX; ++iz) {
                const int cx = ix + PAD_DEPTH, cy = iy + PAD_DEPTH, cz = iz + PAD_DEPTH;
                const int ccx = coarse(cx), ccy = coarse(cy), ccz = coarse(cz);
                double L[4] = {0, 0, 0, 0};
                for (int x = 0; x < STENCIL_INX; ++x) {
                    const int px = cx + x + STENCIL_MIN;
                    const int dcx = (ccx - coarse(px)) * (ccx - coarse(px));
                    for (int sy = STENCIL_MIN; sy <= STENCIL_MAX; ++sy) {
                        const int py = cy + sy;
                        const int dcy = (ccy - coarse(py)) * (ccy - coarse(py));
                        for (int sz = STENCIL_MIN; sz <= STENCIL_MAX; ++sz) {
                            const int idx = x * STENCIL_INX * STENCIL_INX + (sy - STENCIL_MIN) * STENCIL_INX +
                                (sz - STENCIL_MIN);
                            if (!s.pmask[idx]) continue;
                            const int pz = cz + sz;
                            const int dcz = (ccz - coarse(pz)) * (ccz - coarse(pz));
                            if (!(static_cast<double>(dcx + dcy + dcz) < th_rec2)) continue;
                            const double m = mon[padded_flat(px, py, pz)] * dcomp[0];
                            // monopole_kernel_templates.hpp:16-23 and kokkos_kernel.hpp:274-284
                            L[0] += s.four[idx][0] * m;
                            L[1] += s.four[idx][1] * m * dcomp[1];
                            L[2] += s.four[idx][2] * m * dcomp[1];
                            L[3] += s.four[idx][3] * m * dcomp[1];
                        }
                    }
                }
                const int u = inner_flat(ix, iy, iz);
                for (int c = 0; c < 4; ++c) res[u][c] = L[c];
            }
}

// p2m against one refined neighbour direction (kokkos_kernel.hpp:517-942,
// launched per direction with trimmed target ranges, :1051-1165).
void kernel_p2m_dir(const octofmm_subgrid* in, const octofmm_neighbor& nb, const int d[3],
    const Stencils& s, double theta, bool rho, double (*res)[4], double (*cor)[3]) {
    int st[3], en[3], sz3[3], cs[3], ce[3];
    for (int k = 0; k < 3; ++k) {
        st[k] = d[k] == -1 ? INX - STENCIL_MAX : 0;
        en[k] = d[k] == 1 ? STENCIL_MAX : INX;
        sz3[k] = en[k] - st[k];
        cs[k] = d[k] == 1 ? INX - (STENCIL_MAX + 1) : 0;
        ce[k] = d[k] == -1 ? (STENCIL_MAX + 1) : INX;
    }
    const int nsize = sz3[0] * sz3[1] * sz3[2];
    // update_neighbor_input: compact slab of the neighbour's cells
    std::vector<double> sM(20 * (size_t)nsize, 0.0), sY(3 * (size_t)nsize, 0.0);
    auto slab = [&](int x, int y, int z) {
        return (x - st[0]) * (sz3[1] * sz3[2]) + (y - st[1]) * sz3[2] + (z - st[2]);
    };
    if (nb.count == NCELL && !nb.cells) {
        for (int x = st[0]; x < en[0]; ++x)
            for (int y = st[1]; y < en[1]; ++y)
                for (int z = st[2]; z < en[2]; ++z) {
                    const int f = slab(x, y, z), u = inner_flat(x, y, z);
                    for (int c = 0; c < 20; ++c) sM[c * nsize + f] = nb.M[20 * u + c];
                    for (int c = 0; c < 3; ++c) sY[c * nsize + f] = nb.com[3 * u + c];
                }
    } else {
        for (int c = 0; c < nb.count; ++c) {
            const int u = nb.cells[c];
            const int x = u / 64, y = (u / 8) % 8, z = u % 8;
            if (x < st[0] || x >= en[0] || y < st[1] || y >= en[1] || z < st[2] || z >= en[2]) continue;
            const int f = slab(x, y, z);
            for (int k = 0; k < 20; ++k) sM[k * nsize + f] = nb.M[20 * c + k];
            for (int k = 0; k < 3; ++k) sY[k * nsize + f] = nb.com[3 * c + k];
        }
    }
    const double th_rec2 = (1.0 / theta) * (1.0 / theta);
    for (int ix = cs[0]; ix < ce[0]; ++ix)
        for (int iy = cs[1]; iy < ce[1]; ++iy)
            for (int iz = cs[2]; iz < ce[2]; ++iz) {
                const int cx = ix + PAD_DEPTH, cy = iy + PAD_DEPTH, cz = iz + PAD_DEPTH;
                const int u = inner_flat(ix, iy, iz);
                const double X[3] = {in->com[3 * u], in->com[3 * u + 1], in->com[3 * u + 2]};
                for (int x = st[0]; x < en[0]; ++x)
                    for (int y = st[1]; y < en[1]; ++y)
                        for (int z = st[2]; z < en[2]; ++z) {
                            const int px = PAD_DEPTH + d[0] * PAD_DEPTH + x;
                            const int py = PAD_DEPTH + d[1] * PAD_DEPTH + y;
                            const int pz = PAD_DEPTH + d[2] * PAD_DEPTH + z;
                            const long long sfl = (long long)(px - cx - STENCIL_MIN) * STENCIL_INX * STENCIL_INX +
                                (long long)(py - cy - STENCIL_MIN) * STENCIL_INX + (pz - cz - STENCIL_MIN);
                            const bool m1 = sfl >= 0 && sfl < FULL_STENCIL && s.pmask[sfl];
                            if (!m1) continue;
                            const int dcx = (coarse(cx) - coarse(px)) * (coarse(cx) - coarse(px));
                            const int dcy = (coarse(cy) - coarse(py)) * (coarse(cy) - coarse(py));
                            const int dcz = (coarse(cz) - coarse(pz)) * (coarse(cz) - coarse(pz));
                            if (!(static_cast<double>(dcx + dcy + dcz) < th_rec2)) continue;
                            const int f = slab(x, y, z);
                            const double Y[3] = {sY[f], sY[nsize + f], sY[2 * nsize + f]};
                            double mp[20];
                            for (int c = 0; c < 20; ++c) mp[c] = sM[c * nsize + f];
                            pair_p2m(X, Y, mp, res[u], cor[u], rho);
                        }
            }
}

bool valid_subgrid(const octofmm_subgrid* in, bool leaf) {
    if (!in || !in->com) return false;
    if (!leaf && !in->M) return false;
    if (leaf && !in->mon) return false;
    if (in->type != OCTOFMM_RHO && in->type != OCTOFMM_DRHODT) return false;
    for (int f = 0; f < 26; ++f) {
        const octofmm_neighbor& nb = in->neighbors[f];
        if (nb.kind < 0 || nb.kind > 2) return false;
        if (nb.kind != OCTOFMM_NEIGHBOR_EMPTY) {
            if (nb.count < 0 || nb.count > NCELL) return false;
            if (nb.count != NCELL && !nb.cells && nb.count != 0) return false;
        }
    }
    return true;
}

}   // namespace

extern "C" {

int oracle_build_stencils(double theta, uint8_t* mmask, uint8_t* mind, uint8_t* pmask, double* four) {
    Stencils s;
    if (!build_stencils(theta, s)) return OCTOFMM_ERR_THETA;
    if (mmask) std::memcpy(mmask, s.mmask, FULL_STENCIL);
    if (mind) std::memcpy(mind, s.mind, FULL_STENCIL);
    if (pmask) std::memcpy(pmask, s.pmask, FULL_STENCIL);
    if (four) std::memcpy(four, s.four, sizeof(s.four));
    return OCTOFMM_OK;
}

// The multipole stencil as an ordered element list (calculate_stencil order).
int oracle_multipole_stencil(double theta, int32_t* elems, uint8_t* phase, int32_t* n) {
    if (!(theta >= THETA_FLOOR)) return OCTOFMM_ERR_THETA;
    std::vector<Elem> st;
    if (!build_multipole_stencil(theta, st, true)) return OCTOFMM_ERR_THETA;
    *n = static_cast<int32_t>(st.size());
    for (size_t i = 0; i < st.size(); ++i) {
        if (elems) {
            elems[3 * i] = st[i].x;
            elems[3 * i + 1] = st[i].y;
            elems[3 * i + 2] = st[i].z;
        }
        if (phase) phase[i] = st[i].phase;
    }
    return OCTOFMM_OK;
}

void oracle_d_factors(const double* dX, double* out25) {
    DF f;
    d_factors(dX, f);
    out25[0] = f.d2;
    out25[1] = f.d3;
    out25[2] = f.X00;
    out25[3] = f.X11;
    out25[4] = f.X22;
    for (int i = 0; i < 20; ++i) out25[5 + i] = f.D[i];
}

void oracle_pair_m2l(const double* X, const double* Y, const double* mp, const double* mc, double* L,
    double* corr, int32_t rho) {
    pair_m2l(X, Y, mp, mc, L, corr, rho != 0);
}

void oracle_pair_p2m(const double* X, const double* Y, const double* mp, double* L, double* corr,
    int32_t rho) {
    pair_p2m(X, Y, mp, L, corr, rho != 0);
}

int oracle_multipole_interactions(double theta, const octofmm_subgrid* in, double* L, double* L_c) {
    if (!valid_subgrid(in, false) || !L || (in->type == OCTOFMM_RHO && !L_c))
        return OCTOFMM_ERR_INVALID_ARGUMENT;
    Stencils s;
    if (!build_stencils(theta, s)) return OCTOFMM_ERR_THETA;
    Window w;
    stage_multipole(in, w);
    std::vector<double> res(NCELL * 20), cor(NCELL * 3);
    const bool rho = in->type == OCTOFMM_RHO;
    if (in->use_root_stencil)
        kernel_root(w, s, rho, (double(*)[20])res.data(), (double(*)[3])cor.data());
    else
        kernel_nonroot(w, s, theta, rho, (double(*)[20])res.data(), (double(*)[3])cor.data());
    for (int u = 0; u < NCELL; ++u) {
        for (int c = 0; c < 20; ++c) L[20 * u + c] += res[20 * u + c];
        if (rho)
            for (int c = 0; c < 3; ++c) L_c[3 * u + c] = cor[3 * u + c];
    }
    return OCTOFMM_OK;
}

int oracle_monopole_interactions(double theta, const octofmm_subgrid* in,
    int32_t contains_multipole_neighbor, double* L, double* L_c) {
    if (!valid_subgrid(in, true) || !L) return OCTOFMM_ERR_INVALID_ARGUMENT;
    const bool rho = in->type == OCTOFMM_RHO;
    if (rho && contains_multipole_neighbor && !L_c) return OCTOFMM_ERR_INVALID_ARGUMENT;
    if (contains_multipole_neighbor && !in->com) return OCTOFMM_ERR_INVALID_ARGUMENT;
    Stencils s;
    if (!build_stencils(theta, s)) return OCTOFMM_ERR_THETA;
    std::vector<double> mon(ENTRIES, 0.0);
    stage_monopole(in, mon);
    std::vector<double> res(NCELL * 4, 0.0), cor(NCELL * 3, 0.0);
    kernel_p2p(mon, s, theta, in->dx, (double(*)[4])res.data());
    if (contains_multipole_neighbor) {
        for (int flat = 0; flat < 26; ++flat) {
            const octofmm_neighbor& nb = in->neighbors[flat];
            if (nb.kind != OCTOFMM_NEIGHBOR_MULTIPOLE || !nb.M || !nb.com) continue;
            int d[3];
            dir_components(flat_to_raw(flat), d);
            kernel_p2m_dir(in, nb, d, s, theta, rho, (double(*)[4])res.data(), (double(*)[3])cor.data());
        }
    }
    for (int u = 0; u < NCELL; ++u) {
        for (int c = 0; c < 4; ++c) L[20 * u + c] += res[4 * u + c];
        if (rho && contains_multipole_neighbor)
            for (int c = 0; c < 3; ++c) L_c[3 * u + c] = cor[3 * u + c];
    }
    return OCTOFMM_OK;
}

}   // extern "C"

// ---------------------------------------------------------------------------
// Level driver: every target slot of one level, each through the per-sub-grid
// entry points above, exactly as node_server::compute_fmm calls them once per
// octree node (src/node_server.cpp:711-723).  OpenMP over targets; each
// target's arithmetic is independent of the thread count.
// ---------------------------------------------------------------------------
extern "C" int oracle_level_interactions(double theta, int32_t nslots, const int32_t* neighbors,
    const int32_t* kind, const double* xbase, double dx, int32_t ntargets, const int32_t* targets,
    int32_t type, const double* M, const double* com, const double* mon, double* L, double* L_c,
    int32_t nthreads) {
    if (nslots <= 0 || !neighbors || !kind || !xbase || !targets || !com || !L) return OCTOFMM_ERR_INVALID_ARGUMENT;
    Stencils probe;
    if (!build_stencils(theta, probe)) return OCTOFMM_ERR_THETA;
    int status = OCTOFMM_OK;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
    for (int t = 0; t < ntargets; ++t) {
        const int s = targets[t];
        if (s < 0 || s >= nslots) {
#pragma omp critical
            status = OCTOFMM_ERR_INVALID_ARGUMENT;
            continue;
        }
        octofmm_subgrid in;
        std::memset(&in, 0, sizeof(in));
        in.M = M ? M + (size_t)s * NCELL * 20 : nullptr;
        in.com = com + (size_t)s * NCELL * 3;
        in.mon = mon ? mon + (size_t)s * NCELL : nullptr;
        in.dx = dx;
        for (int k = 0; k < 3; ++k) in.xbase[k] = xbase[3 * s + k];
        in.type = type;
        in.use_root_stencil = 0;
        bool has_multipole = false;
        for (int f = 0; f < 26; ++f) {
            const int n = neighbors[(size_t)s * 27 + flat_to_raw(f)];
            octofmm_neighbor& nb = in.neighbors[f];
            if (n < 0) {
                nb.kind = OCTOFMM_NEIGHBOR_EMPTY;
            } else if (kind[n] == OCTOFMM_SLOT_LEAF) {
                nb.kind = OCTOFMM_NEIGHBOR_MONOPOLE;
                nb.count = NCELL;
                nb.mon = mon + (size_t)n * NCELL;
            } else {
                nb.kind = OCTOFMM_NEIGHBOR_MULTIPOLE;
                nb.count = NCELL;
                nb.M = M + (size_t)n * NCELL * 20;
                nb.com = com + (size_t)n * NCELL * 3;
                has_multipole = true;
            }
        }
        double* Ls = L + (size_t)s * NCELL * 20;
        double* Lcs = L_c ? L_c + (size_t)s * NCELL * 3 : nullptr;
        const int st = kind[s] == OCTOFMM_SLOT_LEAF
            ? oracle_monopole_interactions(theta, &in, has_multipole ? 1 : 0, Ls, Lcs)
            : oracle_multipole_interactions(theta, &in, Ls, Lcs);
        if (st != OCTOFMM_OK) {
#pragma omp critical
            status = st;
        }
    }
    return status;
}

// ---------------------------------------------------------------------------
// Tree passes either side of the interaction kernels: the upward M2M pass and
// the downward L2L pass with the leaf force assembly, one octree level at a
// time.  Nodes of a level are 8^3 sub-grids; node n's parent is a non-leaf
// node of the level above and octant[n] (bit 0 x, bit 1 y, bit 2 z, as
// geo::octant::get_side, octotiger/geometry.hpp:333-335) says which 4^3 block
// of the parent's cells the node covers.
// ---------------------------------------------------------------------------
namespace {

// taylor<4> accessors (octotiger/taylor.hpp, src/taylor.cpp:10-58)
constexpr int kMap2[3][3] = {{4, 5, 6}, {5, 7, 8}, {6, 8, 9}};
constexpr int kMap3[3][3][3] = {{{10, 11, 12}, {11, 13, 14}, {12, 14, 15}},
    {{11, 13, 14}, {13, 16, 17}, {14, 17, 18}}, {{12, 14, 15}, {14, 17, 18}, {15, 18, 19}}};

// taylor::operator>>= (octotiger/taylor.hpp:222-255): moves a multipole
// from its expansion centre by X (X = old centre - new centre).
void multipole_shift(const double A[20], const double X[3], double B[20]) {
    for (int i = 0; i < 20; ++i) B[i] = A[i];
    for (int a = 0; a < 3; ++a) B[1 + a] += A[0] * X[a];
    for (int a = 0; a < 3; ++a)
        for (int b = a; b < 3; ++b) {
            B[kMap2[a][b]] += A[1 + a] * X[b] + X[a] * A[1 + b];
            B[kMap2[a][b]] += A[0] * X[a] * X[b];
        }
    for (int a = 0; a < 3; ++a)
        for (int b = a; b < 3; ++b)
            for (int c = b; c < 3; ++c) {
                double& t = B[kMap3[a][b][c]];
                t += A[kMap2[a][b]] * X[c] + A[kMap2[b][c]] * X[a] + A[kMap2[c][a]] * X[b];
                t += A[1 + a] * X[b] * X[c] + A[1 + b] * X[c] * X[a] + A[1 + c] * X[a] * X[b];
                t += A[0] * X[a] * X[b] * X[c];
            }
}

// taylor::operator<<= (octotiger/taylor.hpp:263-311): re-expands a local
// expansion about a point displaced by X (X = new centre - old centre).
void local_shift(const double A[20], const double X[3], double B[20]) {
    for (int i = 0; i < 20; ++i) B[i] = A[i];
    for (int a = 0; a < 3; ++a) B[0] += A[1 + a] * X[a];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) B[0] += A[kMap2[a][b]] * X[a] * X[b] * 0.5;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) B[1 + a] += A[kMap2[a][b]] * X[b];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            for (int c = 0; c < 3; ++c) B[0] += A[kMap3[a][b][c]] * X[a] * X[b] * X[c] * (1.0 / 6.0);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            for (int c = 0; c < 3; ++c) B[1 + a] += A[kMap3[a][b][c]] * X[b] * X[c] * 0.5;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            for (int c = a; c < 3; ++c) B[kMap2[a][c]] += A[kMap3[a][b][c]] * X[b];
}

// parent cell that holds coarse cell (ip, jp, kp) of a node in octant o
// (node_server::compute_fmm gather, src/node_server.cpp:613-636)
inline int parent_cell(int o, int ip, int jp, int kp) {
    return (ip + 4 * (o & 1)) * 64 + (jp + 4 * ((o >> 1) & 1)) * 8 + (kp + 4 * ((o >> 2) & 1));
}

inline int child_cell(int ip, int jp, int kp, int ci) {
    return (2 * ip + (ci & 1)) * 64 + (2 * jp + ((ci >> 1) & 1)) * 8 + (2 * kp + ((ci >> 2) & 1));
}

}   // namespace

extern "C" {

void oracle_multipole_shift(const double* A, const double* X, double* B) { multipole_shift(A, X, B); }
void oracle_local_shift(const double* A, const double* X, double* B) { local_shift(A, X, B); }

// Upward pass of one level (grid::compute_multipoles, src/grid_fmm.cpp:1639-1803).
// Leaf nodes (RHO): their cell centres com = xbase + i*dx are written first
// (:1662-1675).  Every node with a parent then hands its 4^3 coarse
// multipoles and centres of mass to the parent's cells (:1696-1790); DRHODT
// keeps the centres of mass of the RHO pass (:1700-1723 run for RHO only).
int oracle_m2m_level(int32_t nnodes, const int32_t* kind, const int32_t* parent, const int32_t* octant,
    const double* xbase, double dx, int32_t type, const double* M, double* com, const double* mon,
    double* parent_M, double* parent_com) {
    if (nnodes <= 0 || !kind || !parent || !octant || !xbase || !com) return OCTOFMM_ERR_INVALID_ARGUMENT;
    const bool rho = type == OCTOFMM_RHO;
    for (int n = 0; n < nnodes; ++n) {
        const bool leaf = kind[n] == OCTOFMM_SLOT_LEAF;
        if (leaf && !mon) return OCTOFMM_ERR_INVALID_ARGUMENT;
        if (!leaf && !M) return OCTOFMM_ERR_INVALID_ARGUMENT;
        double* c0 = com + (size_t)n * NCELL * 3;
        if (leaf && rho)
            for (int i = 0; i < INX; ++i)
                for (int j = 0; j < INX; ++j)
                    for (int k = 0; k < INX; ++k) {
                        double* p = c0 + 3 * inner_flat(i, j, k);
                        p[0] = xbase[3 * n + 0] + i * dx;
                        p[1] = xbase[3 * n + 1] + j * dx;
                        p[2] = xbase[3 * n + 2] + k * dx;
                    }
        const int pn = parent[n];
        if (pn < 0) continue;
        if (!parent_M || !parent_com) return OCTOFMM_ERR_INVALID_ARGUMENT;
        for (int ip = 0; ip < 4; ++ip)
            for (int jp = 0; jp < 4; ++jp)
                for (int kp = 0; kp < 4; ++kp) {
                    const int pc = parent_cell(octant[n], ip, jp, kp);
                    double* Y = parent_com + ((size_t)pn * NCELL + pc) * 3;
                    if (rho) {
                        double mc[8], X[3][8];
                        for (int ci = 0; ci < 8; ++ci) {
                            const int u = child_cell(ip, jp, kp, ci);
                            mc[ci] = leaf ? mon[(size_t)n * NCELL + u] : M[((size_t)n * NCELL + u) * 20];
                            for (int d = 0; d < 3; ++d) X[d][ci] = c0[3 * u + d];
                        }
                        double mtot = 0.0;
                        for (int ci = 0; ci < 8; ++ci) mtot += mc[ci];
                        for (int d = 0; d < 3; ++d) {
                            double s = 0.0;
                            for (int ci = 0; ci < 8; ++ci) s += X[d][ci] * mc[ci];
                            Y[d] = s / mtot;
                        }
                    }
                    double MM[20] = {0.0};
                    for (int ci = 0; ci < 8; ++ci) {
                        const int u = child_cell(ip, jp, kp, ci);
                        double A[20] = {0.0}, B[20];
                        if (leaf)
                            A[0] = mon[(size_t)n * NCELL + u];
                        else
                            for (int j = 0; j < 20; ++j) A[j] = M[((size_t)n * NCELL + u) * 20 + j];
                        const double dX[3] = {c0[3 * u] - Y[0], c0[3 * u + 1] - Y[1], c0[3 * u + 2] - Y[2]};
                        multipole_shift(A, dX, B);
                        for (int j = 0; j < 20; ++j) MM[j] += B[j];
                    }
                    double* Mo = parent_M + ((size_t)pn * NCELL + pc) * 20;
                    for (int j = 0; j < 20; ++j) Mo[j] = MM[j];
                }
    }
    return OCTOFMM_OK;
}

// Downward pass of one level (grid::compute_expansions, src/grid_fmm.cpp:1524-1637,
// fed as node_server::compute_fmm does, src/node_server.cpp:738-772): the
// parent's final expansion of the cell above is re-expanded about each child
// cell's centre of mass and added to L, the parent's L_c is added unchanged,
// and leaf nodes assemble phi = G L0 and g = -G (L_1..3 + L_c) into
// g[node][cell][4] (DRHODT: phi slot only).  Nodes without a parent add zeros.
int oracle_l2l_level(int32_t nnodes, const int32_t* kind, const int32_t* parent, const int32_t* octant,
    int32_t type, const double* com, const double* parent_L, const double* parent_Lc, const double* parent_com,
    double* L, double* L_c, double G, int32_t correct_am, double* g) {
    if (nnodes <= 0 || !kind || !parent || !octant || !com || !L) return OCTOFMM_ERR_INVALID_ARGUMENT;
    const bool rho = type == OCTOFMM_RHO;
    if (rho && !L_c) return OCTOFMM_ERR_INVALID_ARGUMENT;
    for (int n = 0; n < nnodes; ++n) {
        const int pn = parent[n];
        if (pn >= 0 && (!parent_L || !parent_com || (rho && !parent_Lc))) return OCTOFMM_ERR_INVALID_ARGUMENT;
        for (int u = 0; u < NCELL; ++u) {
            const int i = u / 64, j = (u / 8) % 8, k = u % 8;
            double l[20] = {0.0}, lc[3] = {0.0, 0.0, 0.0};
            if (pn >= 0) {
                const int pc = parent_cell(octant[n], i >> 1, j >> 1, k >> 1);
                const double* A = parent_L + ((size_t)pn * NCELL + pc) * 20;
                const double* Y = parent_com + ((size_t)pn * NCELL + pc) * 3;
                const double* X = com + ((size_t)n * NCELL + u) * 3;
                const double dX[3] = {X[0] - Y[0], X[1] - Y[1], X[2] - Y[2]};
                local_shift(A, dX, l);
                if (rho)
                    for (int d = 0; d < 3; ++d) lc[d] = parent_Lc[((size_t)pn * NCELL + pc) * 3 + d];
            }
            double* Lo = L + ((size_t)n * NCELL + u) * 20;
            for (int c = 0; c < 20; ++c) Lo[c] += l[c];
            double* Co = rho ? L_c + ((size_t)n * NCELL + u) * 3 : nullptr;
            if (rho)
                for (int d = 0; d < 3; ++d) Co[d] += lc[d];
            if (g && kind[n] == OCTOFMM_SLOT_LEAF) {
                double* go = g + ((size_t)n * NCELL + u) * 4;
                go[0] = G * Lo[0];
                if (rho)
                    for (int d = 0; d < 3; ++d) {
                        go[1 + d] = -G * Lo[1 + d];
                        if (correct_am) go[1 + d] -= G * Co[d];
                    }
            }
        }
    }
    return OCTOFMM_OK;
}

}   // extern "C"
