// ref_driver.cpp -- harness that compiles Octo-Tiger's own interaction
// templates from the reference tree (never copied into this repository) into
// oracle/_ref/libocto_ref.so.
//
// TEST INFRASTRUCTURE ONLY: used to pin the oracle (octofmm_oracle.cpp) and to
// generate the golden fixtures under tests/golden/; optionally timed as the CPU
// baseline.  Built by oracle/Makefile (target `ref`) with
//   -I<reference> -DOCTOTIGER_GRIDDIM=8 -DOCTOTIGER_THETA_MINIMUM=0.34
//   -DOCTOTIGER_KOKKOS_MULTIPOLE_TASKS=1 -DOCTOTIGER_KOKKOS_MONOPOLE_TASKS=1
// i.e. the compile definitions CMakeLists.txt:871-950 attaches to octolib.
//
// What comes from the reference tree (compiled where it lies):
//   octotiger/multipole_interactions/kernel/compute_kernel_templates.hpp
//   octotiger/monopole_interactions/kernel/monopole_kernel_templates.hpp
//   octotiger/common_kernel/{interaction_constants,multiindex,helper}.hpp
//   octotiger/geometry.hpp, octotiger/defs.hpp, octotiger/hydro_defs.hpp, octotiger/real.hpp
// What this file adds: the scalar walk of the Kokkos host kernels and the
// AoS->SoA staging (restated, since kokkos_kernel.hpp needs Kokkos/HPX and
// calculate_stencil.cpp needs HPX options, neither of which exists here).
#include "octotiger/common_kernel/helper.hpp"
#include "octotiger/common_kernel/interaction_constants.hpp"
#include "octotiger/common_kernel/multiindex.hpp"
#include "octotiger/geometry.hpp"
#include "octotiger/monopole_interactions/kernel/monopole_kernel_templates.hpp"
#include "octotiger/multipole_interactions/kernel/compute_kernel_templates.hpp"

#include "octofmm.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

using namespace octotiger::fmm;
namespace mi = octotiger::fmm::multipole_interactions;
namespace mo = octotiger::fmm::monopole_interactions;

namespace {

constexpr size_t CLEN = ENTRIES + SOA_PADDING;
auto dmax = [](const double& a, const double& b) { return std::max(a, b); };

struct RefStencils {
    std::vector<int> mask, ind, pmask;
    std::vector<double> four;
};

// calculate_stencil + calculate_stencil_masks, both variants, with the
// reference's reciprocal_distance and multiindex.
bool ref_stencils(double theta0, RefStencils& s) {
    if (theta0 < THETA_FLOOR) return false;
    for (int with_phase = 1; with_phase >= 0; --with_phase) {
        std::vector<two_phase_stencil> st(8);
        for (int64_t i0 = 0; i0 < 2; ++i0)
            for (int64_t i1 = 0; i1 < 2; ++i1)
                for (int64_t i2 = 0; i2 < 2; ++i2) {
                    two_phase_stencil& t = st[i0 * 4 + i1 * 2 + i2];
                    for (int64_t j0 = i0 - INX; j0 < i0 + INX; ++j0)
                        for (int64_t j1 = i1 - INX; j1 < i1 + INX; ++j1)
                            for (int64_t j2 = i2 - INX; j2 < i2 + INX; ++j2) {
                                if (i0 == j0 && i1 == j1 && i2 == j2) continue;
                                const int64_t a0 = (i0 + INX) / 2 - INX / 2, a1 = (i1 + INX) / 2 - INX / 2,
                                              a2 = (i2 + INX) / 2 - INX / 2;
                                const int64_t b0 = (j0 + INX) / 2 - INX / 2, b1 = (j1 + INX) / 2 - INX / 2,
                                              b2 = (j2 + INX) / 2 - INX / 2;
                                const real tf = detail::reciprocal_distance(i0, i1, i2, j0, j1, j2);
                                const real tc = detail::reciprocal_distance(a0, a1, a2, b0, b1, b2);
                                if (tc > theta0) {
                                    t.stencil_elements.emplace_back(j0 - i0, j1 - i1, j2 - i2);
                                    t.stencil_phase_indicator.emplace_back(!with_phase || tf <= theta0);
                                }
                            }
                }
        two_phase_stencil sup;
        for (int i = 0; i < 8; ++i)
            for (size_t e = 0; e < st[i].stencil_elements.size(); ++e) {
                bool found = false;
                for (auto& s2 : sup.stencil_elements)
                    if (st[i].stencil_elements[e].compare(s2)) {
                        found = true;
                        break;
                    }
                if (!found) {
                    sup.stencil_elements.push_back(st[i].stencil_elements[e]);
                    sup.stencil_phase_indicator.push_back(st[i].stencil_phase_indicator[e]);
                }
            }
        if (with_phase) {
            s.mask.assign(FULL_STENCIL_SIZE, 0);
            s.ind.assign(FULL_STENCIL_SIZE, 1);
            for (size_t e = 0; e < sup.stencil_elements.size(); ++e) {
                const auto& el = sup.stencil_elements[e];
                const size_t idx = (el.x + STENCIL_MAX) * STENCIL_INX * STENCIL_INX +
                    (el.y + STENCIL_MAX) * STENCIL_INX + (el.z + STENCIL_MAX);
                s.mask[idx] = 1;
                if (!sup.stencil_phase_indicator[e]) s.ind[idx] = 0;
            }
        } else {
            s.pmask.assign(FULL_STENCIL_SIZE, 0);
            s.four.assign(4 * FULL_STENCIL_SIZE, 0.0);
            for (const auto& el : sup.stencil_elements) {
                const real x = el.x, y = el.y, z = el.z;
                const real r = std::sqrt(mi::sqr(x) + mi::sqr(y) + mi::sqr(z));
                const real r3 = r * r * r;
                const size_t idx = (el.x + STENCIL_MAX) * STENCIL_INX * STENCIL_INX +
                    (el.y + STENCIL_MAX) * STENCIL_INX + (el.z + STENCIL_MAX);
                s.pmask[idx] = 1;
                s.four[4 * idx + 0] = -1.0 / r;
                s.four[4 * idx + 1] = x / r3;
                s.four[4 * idx + 2] = y / r3;
                s.four[4 * idx + 3] = z / r3;
            }
        }
    }
    return true;
}

template <typename F> void padding_cells(const geo::direction& dir, F&& f) {
    int s[3], e[3], o[3];
    for (int k = 0; k < 3; ++k) {
        s[k] = dir[k] == -1 ? INNER_CELLS_PER_DIRECTION - STENCIL_MAX : 0;
        e[k] = dir[k] == 1 ? STENCIL_MAX : INNER_CELLS_PER_DIRECTION;
        o[k] = dir[k] == 1 ? 2 * INNER_CELLS_PER_DIRECTION : (dir[k] == 0 ? INNER_CELLS_PER_DIRECTION : 0);
    }
    for (int i0 = s[0]; i0 < e[0]; ++i0)
        for (int i1 = s[1]; i1 < e[1]; ++i1)
            for (int i2 = s[2]; i2 < e[2]; ++i2) {
                const multiindex<> m(i0 + o[0], i1 + o[1], i2 + o[2]);
                f(m, to_flat_index_padded(m), to_inner_flat_index_not_padded(multiindex<>(i0, i1, i2)));
            }
}

struct RefWindow {
    std::vector<double> mon, M, com;
    RefWindow() : mon(ENTRIES, 0.0), M(20 * CLEN, 0.0), com(3 * CLEN, 0.0) {}
};

void ref_stage(const octofmm_subgrid* in, RefWindow& w) {
    for (int i0 = 0; i0 < INX; ++i0)
        for (int i1 = 0; i1 < INX; ++i1)
            for (int i2 = 0; i2 < INX; ++i2) {
                const multiindex<> m(i0 + INNER_CELLS_PADDING_DEPTH, i1 + INNER_CELLS_PADDING_DEPTH,
                    i2 + INNER_CELLS_PADDING_DEPTH);
                const size_t f = to_flat_index_padded(m);
                const size_t u = to_inner_flat_index_not_padded(multiindex<>(i0, i1, i2));
                for (int c = 0; c < 20; ++c) w.M[c * CLEN + f] = in->M[20 * u + c];
                for (int c = 0; c < 3; ++c) w.com[c * CLEN + f] = in->com[3 * u + c];
                w.mon[f] = 0.0;
            }
    if (in->use_root_stencil) return;
    const double dx = in->dx;
    for (const geo::direction& dir : geo::direction::full_set()) {
        const octofmm_neighbor& nb = in->neighbors[dir.flat_index()];
        auto zero = [&](const multiindex<>&, size_t f, size_t) {
            for (int c = 0; c < 20; ++c) w.M[c * CLEN + f] = 0.0;
            for (int c = 0; c < 3; ++c) w.com[c * CLEN + f] = 0.0;
            w.mon[f] = 0.0;
        };
        auto frame = [&](int u) {
            const multiindex<> off = flat_index_to_multiindex_not_padded(u);
            return multiindex<>(off.x + INNER_CELLS_PADDING_DEPTH + dir[0] * INNER_CELLS_PADDING_DEPTH,
                off.y + INNER_CELLS_PADDING_DEPTH + dir[1] * INNER_CELLS_PADDING_DEPTH,
                off.z + INNER_CELLS_PADDING_DEPTH + dir[2] * INNER_CELLS_PADDING_DEPTH);
        };
        auto inside = [](const multiindex<>& m) {
            const int lo = PADDING_OFFSET, hi = PADDING_OFFSET + PADDED_STRIDE;
            return m.x >= lo && m.x < hi && m.y >= lo && m.y < hi && m.z >= lo && m.z < hi;
        };
        if (nb.kind == OCTOFMM_NEIGHBOR_MULTIPOLE && nb.M && nb.com) {
            if (nb.count == 512 && !nb.cells) {
                padding_cells(dir, [&](const multiindex<>&, size_t f, size_t u) {
                    for (int c = 0; c < 20; ++c) w.M[c * CLEN + f] = nb.M[20 * u + c];
                    for (int c = 0; c < 3; ++c) w.com[c * CLEN + f] = nb.com[3 * u + c];
                    w.mon[f] = 0.0;
                });
            } else {
                padding_cells(dir, zero);
                for (int k = 0; k < nb.count; ++k) {
                    const multiindex<> m = frame(nb.cells[k]);
                    if (!inside(m)) continue;
                    const size_t f = to_flat_index_padded(m);
                    for (int c = 0; c < 20; ++c) w.M[c * CLEN + f] = nb.M[20 * k + c];
                    for (int c = 0; c < 3; ++c) w.com[c * CLEN + f] = nb.com[3 * k + c];
                }
            }
        } else if (nb.kind == OCTOFMM_NEIGHBOR_MONOPOLE && nb.mon) {
            const bool full = nb.count == 512 && !nb.cells;
            padding_cells(dir, [&](const multiindex<>& i, size_t f, size_t u) {
                const double e[3] = {(i.x) * dx + in->xbase[0] - INNER_CELLS_PER_DIRECTION * dx,
                    (i.y) * dx + in->xbase[1] - INNER_CELLS_PER_DIRECTION * dx,
                    (i.z) * dx + in->xbase[2] - INNER_CELLS_PER_DIRECTION * dx};
                for (int c = 0; c < 3; ++c) w.com[c * CLEN + f] = e[c];
                for (int c = 0; c < 20; ++c) w.M[c * CLEN + f] = 0.0;
                w.mon[f] = full ? nb.mon[u] : 0.0;
            });
            if (!full)
                for (int k = 0; k < nb.count; ++k) {
                    const multiindex<> m = frame(nb.cells[k]);
                    if (inside(m)) w.mon[to_flat_index_padded(m)] = nb.mon[k];
                }
        } else {
            padding_cells(dir, zero);
        }
    }
}

}   // namespace

extern "C" {

void ref_d_factors(const double* dX_in, double* out25) {
    double dX[3] = {dX_in[0], dX_in[1], dX_in[2]};
    double d2, d3, X00, X11, X22, D[20];
    mi::compute_d_factors(d2, d3, X00, X11, X22, D, dX, dmax);
    out25[0] = d2;
    out25[1] = d3;
    out25[2] = X00;
    out25[3] = X11;
    out25[4] = X22;
    for (int i = 0; i < 20; ++i) out25[5 + i] = D[i];
}

void ref_pair_m2l(const double* X_in, const double* Y_in, const double* mp_in, const double* mc_in,
    double* L_io, double* corr_io, int32_t rho) {
    double X[3], Y[3], mp[20], mc[20], L[20], C[3];
    std::copy(X_in, X_in + 3, X);
    std::copy(Y_in, Y_in + 3, Y);
    std::copy(mp_in, mp_in + 20, mp);
    std::copy(mc_in, mc_in + 20, mc);
    std::copy(L_io, L_io + 20, L);
    std::copy(corr_io, corr_io + 3, C);
    if (rho)
        mi::compute_kernel_rho(X, Y, mp, L, C, mc, dmax);
    else
        mi::compute_kernel_non_rho(X, Y, mp, L, dmax);
    std::copy(L, L + 20, L_io);
    std::copy(C, C + 3, corr_io);
}

void ref_pair_p2m(const double* X_in, const double* Y_in, const double* mp_in, double* L_io,
    double* corr_io, int32_t rho) {
    double X[3], Y[3], mp[20], L[4], C[3];
    std::copy(X_in, X_in + 3, X);
    std::copy(Y_in, Y_in + 3, Y);
    std::copy(mp_in, mp_in + 20, mp);
    std::copy(L_io, L_io + 4, L);
    std::copy(corr_io, corr_io + 3, C);
    if (rho)
        mo::compute_kernel_p2m_rho(X, Y, mp, L, C, dmax);
    else
        mo::compute_kernel_p2m_non_rho(X, Y, mp, L, dmax);
    std::copy(L, L + 4, L_io);
    std::copy(C, C + 3, corr_io);
}

int ref_build_stencils(double theta, uint8_t* mmask, uint8_t* mind, uint8_t* pmask, double* four) {
    RefStencils s;
    if (!ref_stencils(theta, s)) return OCTOFMM_ERR_THETA;
    for (int i = 0; i < FULL_STENCIL_SIZE; ++i) {
        if (mmask) mmask[i] = (uint8_t)s.mask[i];
        if (mind) mind[i] = (uint8_t)s.ind[i];
        if (pmask) pmask[i] = (uint8_t)s.pmask[i];
    }
    if (four) std::memcpy(four, s.four.data(), sizeof(double) * 4 * FULL_STENCIL_SIZE);
    return OCTOFMM_OK;
}

// Non-root / root M2L of one sub-grid: the Kokkos host kernels
// (kokkos_kernel.hpp:137-321, 324-471, 474-633, 637-764) walked with SIMD width 1.
int ref_multipole_interactions(double theta, const octofmm_subgrid* in, double* Lout, double* Lcout) {
    RefStencils s;
    if (!ref_stencils(theta, s)) return OCTOFMM_ERR_THETA;
    RefWindow w;
    ref_stage(in, w);
    const bool rho = in->type == OCTOFMM_RHO;
    const double th2 = (1.0 / theta) * (1.0 / theta);
    for (int idx = 0; idx < INX; ++idx)
        for (int idy = 0; idy < INX; ++idy)
            for (int idz = 0; idz < INX; ++idz) {
                const multiindex<> ci(idx + INNER_CELLS_PADDING_DEPTH, idy + INNER_CELLS_PADDING_DEPTH,
                    idz + INNER_CELLS_PADDING_DEPTH);
                const size_t cf = to_flat_index_padded(ci);
                multiindex<> cc = ci;
                cc.transform_coarse();
                double mc[20], X[3], L[20] = {0}, C[3] = {0}, mp[20], Y[3];
                for (int c = 0; c < 20; ++c) mc[c] = w.M[c * CLEN + cf];
                for (int c = 0; c < 3; ++c) X[c] = w.com[c * CLEN + cf];
                if (!in->use_root_stencil) {
                    for (int x = 0; x < STENCIL_INX; ++x)
                        for (int y = 0; y < STENCIL_INX; ++y)
                            for (int sz = STENCIL_MIN; sz <= STENCIL_MAX; ++sz) {
                                const size_t index = x * STENCIL_INX * STENCIL_INX + y * STENCIL_INX + (sz - STENCIL_MIN);
                                if (!s.mask[index]) continue;
                                multiindex<> pi(ci.x + x + STENCIL_MIN, ci.y + y + STENCIL_MIN, ci.z + sz);
                                multiindex<> pc = pi;
                                pc.transform_coarse();
                                const double dist = distance_squared_reciprocal(cc, pc);
                                if (!(dist < th2)) continue;
                                const size_t pf = to_flat_index_padded(pi);
                                for (int c = 0; c < 3; ++c) Y[c] = w.com[c * CLEN + pf];
                                mp[0] = w.mon[pf];
                                const bool far = s.ind[index] != 0;
                                mp[0] += far ? w.M[pf] : 0.0;
                                for (int c = 1; c < 20; ++c) mp[c] = far ? w.M[c * CLEN + pf] : 0.0;
                                if (rho)
                                    mi::compute_kernel_rho(X, Y, mp, L, C, mc, dmax);
                                else
                                    mi::compute_kernel_non_rho(X, Y, mp, L, dmax);
                            }
                } else {
                    for (int x = 0; x < INX; ++x)
                        for (int y = 0; y < INX; ++y)
                            for (int z = 0; z < INX; ++z) {
                                const int sx = x - idx, sy = y - idy, sz = z - idz;
                                if (sx >= STENCIL_MIN && sx <= STENCIL_MAX && sy >= STENCIL_MIN && sy <= STENCIL_MAX &&
                                    sz >= STENCIL_MIN && sz <= STENCIL_MAX) {
                                    const size_t index = (sx - STENCIL_MIN) * STENCIL_INX * STENCIL_INX +
                                        (sy - STENCIL_MIN) * STENCIL_INX + (sz - STENCIL_MIN);
                                    if (!s.ind[index] || (sx == 0 && sy == 0 && sz == 0)) continue;
                                }
                                const size_t pf = to_flat_index_padded(multiindex<>(x + INX, y + INX, z + INX));
                                for (int c = 0; c < 3; ++c) Y[c] = w.com[c * CLEN + pf];
                                for (int c = 0; c < 20; ++c) mp[c] = w.M[c * CLEN + pf];
                                if (rho)
                                    mi::compute_kernel_rho(X, Y, mp, L, C, mc, dmax);
                                else
                                    mi::compute_kernel_non_rho(X, Y, mp, L, dmax);
                            }
                }
                const size_t u = to_inner_flat_index_not_padded(multiindex<>(idx, idy, idz));
                for (int c = 0; c < 20; ++c) Lout[20 * u + c] += L[c];
                if (rho)
                    for (int c = 0; c < 3; ++c) Lcout[3 * u + c] = C[c];
            }
    return OCTOFMM_OK;
}

// p2p (kokkos_kernel.hpp:143-320) + p2m per refined direction (:517-942).
int ref_monopole_interactions(double theta, const octofmm_subgrid* in, int32_t with_p2m, double* Lout,
    double* Lcout) {
    RefStencils s;
    if (!ref_stencils(theta, s)) return OCTOFMM_ERR_THETA;
    std::vector<double> mon(ENTRIES, 0.0);
    for (int u = 0; u < 512; ++u) {
        const multiindex<> o = flat_index_to_multiindex_not_padded(u);
        mon[to_flat_index_padded(multiindex<>(o.x + INX, o.y + INX, o.z + INX))] = in->mon[u];
    }
    for (const geo::direction& dir : geo::direction::full_set()) {
        const octofmm_neighbor& nb = in->neighbors[dir.flat_index()];
        const bool full = nb.kind == OCTOFMM_NEIGHBOR_MONOPOLE && nb.mon && nb.count == 512 && !nb.cells;
        padding_cells(dir, [&](const multiindex<>&, size_t f, size_t u) { mon[f] = full ? nb.mon[u] : 0.0; });
        if (nb.kind == OCTOFMM_NEIGHBOR_MONOPOLE && nb.mon && !full)
            for (int k = 0; k < nb.count; ++k) {
                const multiindex<> o = flat_index_to_multiindex_not_padded(nb.cells[k]);
                const multiindex<> m(o.x + INX + dir[0] * INX, o.y + INX + dir[1] * INX, o.z + INX + dir[2] * INX);
                const int lo = PADDING_OFFSET, hi = PADDING_OFFSET + PADDED_STRIDE;
                if (m.x >= lo && m.x < hi && m.y >= lo && m.y < hi && m.z >= lo && m.z < hi)
                    mon[to_flat_index_padded(m)] = nb.mon[k];
            }
    }
    const bool rho = in->type == OCTOFMM_RHO;
    const double th2 = (1.0 / theta) * (1.0 / theta);
    const double dcomp[2] = {1.0 / in->dx, -1.0 / in->dx};
    std::vector<double> res(512 * 4, 0.0), cor(512 * 3, 0.0);
    for (int idx = 0; idx < INX; ++idx)
        for (int idy = 0; idy < INX; ++idy)
            for (int idz = 0; idz < INX; ++idz) {
                const multiindex<> ci(idx + INX, idy + INX, idz + INX);
                multiindex<> cc = ci;
                cc.transform_coarse();
                double L[4] = {0, 0, 0, 0};
                for (int x = 0; x < STENCIL_INX; ++x)
                    for (int sy = STENCIL_MIN; sy <= STENCIL_MAX; ++sy)
                        for (int sz = STENCIL_MIN; sz <= STENCIL_MAX; ++sz) {
                            const size_t index =
                                x * STENCIL_INX * STENCIL_INX + (sy - STENCIL_MIN) * STENCIL_INX + (sz - STENCIL_MIN);
                            if (!s.pmask[index]) continue;
                            multiindex<> pi(ci.x + x + STENCIL_MIN, ci.y + sy, ci.z + sz);
                            multiindex<> pc = pi;
                            pc.transform_coarse();
                            if (!(static_cast<double>(distance_squared_reciprocal(cc, pc)) < th2)) continue;
                            const double m = mon[to_flat_index_padded(pi)] * dcomp[0];
                            const double four[4] = {s.four[4 * index], s.four[4 * index + 1], s.four[4 * index + 2],
                                s.four[4 * index + 3]};
                            L[0] += four[0] * m;
                            L[1] += four[1] * m * dcomp[1];
                            L[2] += four[2] * m * dcomp[1];
                            L[3] += four[3] * m * dcomp[1];
                        }
                const size_t u = to_inner_flat_index_not_padded(multiindex<>(idx, idy, idz));
                for (int c = 0; c < 4; ++c) res[4 * u + c] = L[c];
            }
    if (with_p2m) {
        for (const geo::direction& dir : geo::direction::full_set()) {
            const octofmm_neighbor& nb = in->neighbors[dir.flat_index()];
            if (nb.kind != OCTOFMM_NEIGHBOR_MULTIPOLE || !nb.M || !nb.com) continue;
            int st[3], en[3], sz3[3], cs[3], ce[3];
            for (int k = 0; k < 3; ++k) {
                st[k] = dir[k] == -1 ? INX - STENCIL_MAX : 0;
                en[k] = dir[k] == 1 ? STENCIL_MAX : INX;
                sz3[k] = en[k] - st[k];
                cs[k] = dir[k] == 1 ? INX - (STENCIL_MAX + 1) : 0;
                ce[k] = dir[k] == -1 ? (STENCIL_MAX + 1) : INX;
            }
            const int ns = sz3[0] * sz3[1] * sz3[2];
            std::vector<double> sM(20 * ns, 0.0), sY(3 * ns, 0.0);
            auto put = [&](int x, int y, int z, const double* Mv, const double* Yv) {
                if (x < st[0] || x >= en[0] || y < st[1] || y >= en[1] || z < st[2] || z >= en[2]) return;
                const int f = (x - st[0]) * (sz3[1] * sz3[2]) + (y - st[1]) * sz3[2] + (z - st[2]);
                for (int c = 0; c < 20; ++c) sM[c * ns + f] = Mv[c];
                for (int c = 0; c < 3; ++c) sY[c * ns + f] = Yv[c];
            };
            if (nb.count == 512 && !nb.cells) {
                for (int u = 0; u < 512; ++u) put(u / 64, (u / 8) % 8, u % 8, nb.M + 20 * u, nb.com + 3 * u);
            } else {
                for (int k = 0; k < nb.count; ++k) {
                    const int u = nb.cells[k];
                    put(u / 64, (u / 8) % 8, u % 8, nb.M + 20 * k, nb.com + 3 * k);
                }
            }
            for (int ix = cs[0]; ix < ce[0]; ++ix)
                for (int iy = cs[1]; iy < ce[1]; ++iy)
                    for (int iz = cs[2]; iz < ce[2]; ++iz) {
                        const multiindex<> ci(ix + INX, iy + INX, iz + INX);
                        multiindex<> cc = ci;
                        cc.transform_coarse();
                        const size_t u = to_inner_flat_index_not_padded(multiindex<>(ix, iy, iz));
                        double X[3] = {in->com[3 * u], in->com[3 * u + 1], in->com[3 * u + 2]};
                        double L[4], C[3], Y[3], mp[20];
                        for (int c = 0; c < 4; ++c) L[c] = res[4 * u + c];
                        for (int c = 0; c < 3; ++c) C[c] = cor[3 * u + c];
                        for (int x = st[0]; x < en[0]; ++x)
                            for (int y = st[1]; y < en[1]; ++y)
                                for (int z = st[2]; z < en[2]; ++z) {
                                    const multiindex<> pi(INX + dir[0] * INX + x, INX + dir[1] * INX + y,
                                        INX + dir[2] * INX + z);
                                    const long long sfl =
                                        (long long)(pi.x - ci.x - STENCIL_MIN) * STENCIL_INX * STENCIL_INX +
                                        (long long)(pi.y - ci.y - STENCIL_MIN) * STENCIL_INX + (pi.z - ci.z - STENCIL_MIN);
                                    if (!(sfl >= 0 && sfl < FULL_STENCIL_SIZE && s.pmask[sfl])) continue;
                                    multiindex<> pc = pi;
                                    pc.transform_coarse();
                                    if (!(static_cast<double>(distance_squared_reciprocal(cc, pc)) < th2)) continue;
                                    const int f = (x - st[0]) * (sz3[1] * sz3[2]) + (y - st[1]) * sz3[2] + (z - st[2]);
                                    for (int c = 0; c < 3; ++c) Y[c] = sY[c * ns + f];
                                    for (int c = 0; c < 20; ++c) mp[c] = sM[c * ns + f];
                                    if (rho)
                                        mo::compute_kernel_p2m_rho(X, Y, mp, L, C, dmax);
                                    else
                                        mo::compute_kernel_p2m_non_rho(X, Y, mp, L, dmax);
                                }
                        for (int c = 0; c < 4; ++c) res[4 * u + c] = L[c];
                        for (int c = 0; c < 3; ++c) cor[3 * u + c] = C[c];
                    }
        }
    }
    for (int u = 0; u < 512; ++u) {
        for (int c = 0; c < 4; ++c) Lout[20 * u + c] += res[4 * u + c];
        if (rho && with_p2m)
            for (int c = 0; c < 3; ++c) Lcout[3 * u + c] = cor[3 * u + c];
    }
    return OCTOFMM_OK;
}

}   // extern "C"

// Level loop over the reference templates: every target slot of a level, each
// through ref_multipole_interactions / ref_monopole_interactions, OpenMP over
// targets (one sub-grid per task, as the reference schedules one task per
// octree node).
extern "C" int ref_level_interactions(double theta, int32_t nslots, const int32_t* neighbors,
    const int32_t* kind, const double* xbase, double dx, int32_t ntargets, const int32_t* targets,
    int32_t type, const double* M, const double* com, const double* mon, double* L, double* L_c,
    int32_t nthreads) {
    int status = OCTOFMM_OK;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
    for (int t = 0; t < ntargets; ++t) {
        const int s = targets[t];
        if (s < 0 || s >= nslots) continue;
        octofmm_subgrid in;
        std::memset(&in, 0, sizeof(in));
        in.M = M + (size_t)s * 512 * 20;
        in.com = com + (size_t)s * 512 * 3;
        in.mon = mon + (size_t)s * 512;
        in.dx = dx;
        for (int k = 0; k < 3; ++k) in.xbase[k] = xbase[3 * s + k];
        in.type = type;
        bool has_multipole = false;
        for (int f = 0; f < 26; ++f) {
            const int raw = f < 13 ? f : f + 1;
            const int n = neighbors[(size_t)s * 27 + raw];
            octofmm_neighbor& nb = in.neighbors[f];
            if (n < 0) {
                nb.kind = OCTOFMM_NEIGHBOR_EMPTY;
            } else if (kind[n] == OCTOFMM_SLOT_LEAF) {
                nb.kind = OCTOFMM_NEIGHBOR_MONOPOLE;
                nb.count = 512;
                nb.mon = mon + (size_t)n * 512;
            } else {
                nb.kind = OCTOFMM_NEIGHBOR_MULTIPOLE;
                nb.count = 512;
                nb.M = M + (size_t)n * 512 * 20;
                nb.com = com + (size_t)n * 512 * 3;
                has_multipole = true;
            }
        }
        double* Ls = L + (size_t)s * 512 * 20;
        double* Lcs = L_c + (size_t)s * 512 * 3;
        const int st = kind[s] == OCTOFMM_SLOT_LEAF ? ref_monopole_interactions(theta, &in, has_multipole, Ls, Lcs)
                                                    : ref_multipole_interactions(theta, &in, Ls, Lcs);
        if (st != OCTOFMM_OK) {
#pragma omp critical
            status = st;
        }
    }
    return status;
}

// ---------------------------------------------------------------------------
// The timed CPU baseline: the Kokkos host kernels' z-row SIMD walk
// (multipole_kernel_rho_impl / multipole_kernel_non_rho_impl,
// octotiger/multipole_interactions/kernel/kokkos_kernel.hpp:137-321, 474-633,
// launched by launch_interface :865-940 with one z-row of INX cells split into
// INX / W vectors of W lanes), with the reference's own compute_kernel_rho /
// compute_kernel_non_rho instantiated on a W-wide vector of doubles in place
// of the Kokkos SIMD type.  One sub-grid per OpenMP task, staged AoS -> SoA
// per sub-grid as update_input does; stencils built once per level call, as
// the reference builds them once at start-up.  The library is built three
// times by oracle/Makefile: the pinning build (-O2 -ffp-contract=off, no
// vector walk: REF_SIMD_WIDTH 1) and the timed builds W = 8 (-O3
// -march=x86-64-v4) and W = 4 (-O3 -march=x86-64-v3), which export
// ref_level_interactions_simd.
// ---------------------------------------------------------------------------
#ifndef REF_SIMD_WIDTH
#define REF_SIMD_WIDTH 1
#endif

extern "C" int ref_simd_width(void) { return REF_SIMD_WIDTH; }

#if REF_SIMD_WIDTH > 1

namespace refsimd {

template <int W>
struct vec {
    typedef double raw __attribute__((vector_size(W * 8)));
    typedef long long mask __attribute__((vector_size(W * 8)));
    raw x;
    vec() = default;
    vec(double a) { x = raw{} + a; }
    static vec from(raw r) {
        vec v;
        v.x = r;
        return v;
    }
    vec& operator+=(const vec& o) { x += o.x; return *this; }
    vec& operator-=(const vec& o) { x -= o.x; return *this; }
    vec& operator*=(const vec& o) { x *= o.x; return *this; }
};
template <int W> vec<W> operator+(const vec<W>& a, const vec<W>& b) { return vec<W>::from(a.x + b.x); }
template <int W> vec<W> operator-(const vec<W>& a, const vec<W>& b) { return vec<W>::from(a.x - b.x); }
template <int W> vec<W> operator*(const vec<W>& a, const vec<W>& b) { return vec<W>::from(a.x * b.x); }
template <int W> vec<W> operator/(const vec<W>& a, const vec<W>& b) { return vec<W>::from(a.x / b.x); }
template <int W> vec<W> operator+(double a, const vec<W>& b) { return vec<W>::from(a + b.x); }
template <int W> vec<W> operator-(double a, const vec<W>& b) { return vec<W>::from(a - b.x); }
template <int W> vec<W> operator*(double a, const vec<W>& b) { return vec<W>::from(a * b.x); }
template <int W> vec<W> operator/(double a, const vec<W>& b) { return vec<W>::from(a / b.x); }
template <int W> vec<W> operator+(const vec<W>& a, double b) { return vec<W>::from(a.x + b); }
template <int W> vec<W> operator-(const vec<W>& a, double b) { return vec<W>::from(a.x - b); }
template <int W> vec<W> operator*(const vec<W>& a, double b) { return vec<W>::from(a.x * b); }
template <int W> vec<W> operator/(const vec<W>& a, double b) { return vec<W>::from(a.x / b); }
template <int W> vec<W> operator-(const vec<W>& a) { return vec<W>::from(-a.x); }
template <int W> vec<W> sqrt(const vec<W>& a) {
    vec<W> r;
    for (int i = 0; i < W; ++i) r.x[i] = __builtin_sqrt(a.x[i]);
    return r;
}
template <int W> vec<W> vmax(const vec<W>& a, const vec<W>& b) { return vec<W>::from(a.x > b.x ? a.x : b.x); }
template <int W> vec<W> load(const double* p) {
    vec<W> r;
    std::memcpy(&r.x, p, sizeof(r.x));
    return r;
}

// One z-row segment of W target cells of one sub-grid (kokkos_kernel.hpp:137-321).
template <int W>
void row_segment(const RefStencils& s, const RefWindow& w, bool rho, double th2, int idx, int idy, int idz,
    double* Lout, double* Lcout) {
    typedef vec<W> V;
    typedef typename V::mask lraw;
    const multiindex<> ci(idx + INNER_CELLS_PADDING_DEPTH, idy + INNER_CELLS_PADDING_DEPTH,
        idz * W + INNER_CELLS_PADDING_DEPTH);
    const size_t cf = to_flat_index_padded(ci);
    const int32_t ccx = ((ci.x + INX) >> 1) - (INX / 2);
    const int32_t ccy = ((ci.y + INX) >> 1) - (INX / 2);
    int32_t ccz[W];
    for (int i = 0; i < W; ++i) ccz[i] = ((ci.z + i + INX) >> 1) - (INX / 2);
    V mc[20], X[3], L[20], C[3], mp[20], Y[3];
    for (int c = 0; c < 20; ++c) mc[c] = load<W>(&w.M[c * CLEN + cf]);
    for (int c = 0; c < 3; ++c) X[c] = load<W>(&w.com[c * CLEN + cf]);
    for (int c = 0; c < 20; ++c) L[c] = V(0.0);
    for (int c = 0; c < 3; ++c) C[c] = V(0.0);
    auto vmax_f = [](const V& a, const V& b) -> V { return vmax(a, b); };
    for (int x = 0; x < STENCIL_INX; ++x) {
        const int px = ci.x + x + STENCIL_MIN;
        const int32_t dxc = ccx - (((px + INX) >> 1) - (INX / 2));
        for (int y = 0; y < STENCIL_INX; ++y) {
            const int py = ci.y + y + STENCIL_MIN;
            const int32_t dyc = ccy - (((py + INX) >> 1) - (INX / 2));
            for (int sz = STENCIL_MIN; sz <= STENCIL_MAX; ++sz) {
                const size_t index = x * STENCIL_INX * STENCIL_INX + y * STENCIL_INX + (sz - STENCIL_MIN);
                if (!s.mask[index]) continue;
                const int pz = ci.z + sz;
                V tc;
                for (int i = 0; i < W; ++i) {
                    const int32_t dzc = ccz[i] - (((pz + i + INX) >> 1) - (INX / 2));
                    tc.x[i] = static_cast<double>(dxc * dxc + dyc * dyc + dzc * dzc);
                }
                const lraw mask = tc.x < th2;
                bool any = false;
                for (int i = 0; i < W; ++i) any = any || mask[i] != 0;
                if (!any) continue;
                const size_t pf = to_flat_index_padded(multiindex<>(px, py, pz));
                for (int c = 0; c < 3; ++c) Y[c] = load<W>(&w.com[c * CLEN + pf]);
                const typename V::raw zero = {};
                mp[0] = V::from(mask ? load<W>(&w.mon[pf]).x : zero);
                const lraw fmask = s.ind[index] ? mask : lraw{};
                mp[0].x += fmask ? load<W>(&w.M[pf]).x : zero;
                for (int c = 1; c < 20; ++c) mp[c] = V::from(fmask ? load<W>(&w.M[c * CLEN + pf]).x : zero);
                if (rho)
                    mi::compute_kernel_rho(X, Y, mp, L, C, mc, vmax_f);
                else
                    mi::compute_kernel_non_rho(X, Y, mp, L, vmax_f);
            }
        }
    }
    for (int i = 0; i < W; ++i) {
        const size_t u = to_inner_flat_index_not_padded(multiindex<>(idx, idy, idz * W + i));
        for (int c = 0; c < 20; ++c) Lout[20 * u + c] += L[c].x[i];
        if (rho)
            for (int c = 0; c < 3; ++c) Lcout[3 * u + c] = C[c].x[i];
    }
}

}   // namespace refsimd

// Non-root M2L over the target slots of a level (non-leaf targets with full
// or empty neighbours; leaf targets and leaf neighbours go through the scalar
// walk above), one sub-grid per OpenMP task.
extern "C" int ref_level_interactions_simd(double theta, int32_t nslots, const int32_t* neighbors,
    const int32_t* kind, const double* xbase, double dx, int32_t ntargets, const int32_t* targets,
    int32_t type, const double* M, const double* com, const double* mon, double* L, double* L_c,
    int32_t nthreads) {
    static_assert(INX % REF_SIMD_WIDTH == 0, "z-rows split into whole vectors");
    RefStencils s;
    if (!ref_stencils(theta, s)) return OCTOFMM_ERR_THETA;
    const double th2 = (1.0 / theta) * (1.0 / theta);
    const bool rho = type == OCTOFMM_RHO;
    int status = OCTOFMM_OK;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
    for (int t = 0; t < ntargets; ++t) {
        const int sl = targets[t];
        if (sl < 0 || sl >= nslots || kind[sl] != OCTOFMM_SLOT_NONLEAF) {
#pragma omp critical
            status = OCTOFMM_ERR_INVALID_ARGUMENT;
            continue;
        }
        octofmm_subgrid in;
        std::memset(&in, 0, sizeof(in));
        in.M = M + (size_t)sl * 512 * 20;
        in.com = com + (size_t)sl * 512 * 3;
        in.mon = mon + (size_t)sl * 512;
        in.dx = dx;
        for (int k = 0; k < 3; ++k) in.xbase[k] = xbase[3 * sl + k];
        in.type = type;
        for (int f = 0; f < 26; ++f) {
            const int raw = f < 13 ? f : f + 1;
            const int n = neighbors[(size_t)sl * 27 + raw];
            octofmm_neighbor& nb = in.neighbors[f];
            if (n < 0) {
                nb.kind = OCTOFMM_NEIGHBOR_EMPTY;
            } else if (kind[n] == OCTOFMM_SLOT_LEAF) {
                nb.kind = OCTOFMM_NEIGHBOR_MONOPOLE;
                nb.count = 512;
                nb.mon = mon + (size_t)n * 512;
            } else {
                nb.kind = OCTOFMM_NEIGHBOR_MULTIPOLE;
                nb.count = 512;
                nb.M = M + (size_t)n * 512 * 20;
                nb.com = com + (size_t)n * 512 * 3;
            }
        }
        RefWindow w;
        ref_stage(&in, w);
        double* Ls = L + (size_t)sl * 512 * 20;
        double* Lcs = L_c + (size_t)sl * 512 * 3;
        for (int idx = 0; idx < INX; ++idx)
            for (int idy = 0; idy < INX; ++idy)
                for (int idz = 0; idz < INX / REF_SIMD_WIDTH; ++idz)
                    refsimd::row_segment<REF_SIMD_WIDTH>(s, w, rho, th2, idx, idy, idz, Ls, Lcs);
    }
    return status;
}
#endif   // REF_SIMD_WIDTH > 1
