// octofmm_interface.hpp -- header-only C++ adapter that keeps Octo-Tiger's
// interaction-kernel signatures and forwards to the C-ABI of octofmm.h.
//
// Drop-in for
//   octotiger::fmm::multipole_interactions::multipole_interaction_interface
//     ::compute_multipole_interactions
//     (octotiger/multipole_interactions/legacy/multipole_interaction_interface.hpp:33-77)
//   octotiger::fmm::monopole_interactions::monopole_interaction_interface
//     ::compute_interactions
//     (octotiger/monopole_interactions/legacy/monopole_interaction_interface.hpp:29-61)
//   and the free functions multipole_kernel_interface / monopole_kernel_interface
//     (src/multipole_interactions/multipole_kernel_interface.cpp:73-77,
//      src/monopole_interactions/monopole_kernel_interface.cpp:97-101).
//
// The adapter is templated on the caller's types so it compiles inside an
// Octo-Tiger build without this repository depending on Octo-Tiger:
//   multipole / expansion : taylor<4, real>   (20 components, operator[])
//   space_vector          : std::array<real, 3>-like (operator[])
//   neighbor_gravity_type : { data.M, data.m, data.x (shared_ptr<vector>), is_monopole }
//   grid                  : get_L(), get_L_c(), get_ilist_n_bnd(dir) (partial halos)
// Results are accumulated into grid->get_L() and written to grid->get_L_c(),
// as the reference does.  Errors throw std::runtime_error with the library's
// message (the reference aborts instead).
#pragma once

#include <array>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "octofmm.h"

namespace octofmm_adapter {

// One GPU and one opening angle; build once at start-up (the reference builds
// its stencils at start-up too, frontend/init_methods.cpp:189-192).
class context {
public:
    explicit context(double theta = 0.5, int device = 0) {
        octofmm_config cfg{};
        cfg.theta = theta;
        cfg.device = device;
        const octofmm_status s = octofmm_create(&cfg, &ctx_);
        if (s != OCTOFMM_OK) throw std::runtime_error(std::string("octofmm_create: ") + octofmm_status_string(s));
    }
    ~context() { octofmm_destroy(ctx_); }
    context(const context&) = delete;
    context& operator=(const context&) = delete;
    octofmm_context* get() const { return ctx_; }
    void check(octofmm_status s) const {
        if (s != OCTOFMM_OK)
            throw std::runtime_error(std::string("octofmm: ") + octofmm_status_string(s) + ": " +
                octofmm_last_error(ctx_));
    }

private:
    octofmm_context* ctx_ = nullptr;
};

namespace detail {

// Pointer to n contiguous records of `width` doubles: the caller's storage
// when its element type is exactly `width` doubles, else a packed copy.
template <typename T>
double component(const T& x, int c) {
    if constexpr (std::is_arithmetic<T>::value) {
        (void)c;
        return static_cast<double>(x);
    } else {
        return static_cast<double>(x[c]);
    }
}

template <typename T>
const double* packed(const std::vector<T>& v, int width, std::vector<double>& scratch) {
    if (std::is_standard_layout<T>::value && sizeof(T) == width * sizeof(double))
        return reinterpret_cast<const double*>(v.data());
    scratch.resize(v.size() * width);
    for (size_t i = 0; i < v.size(); ++i)
        for (int c = 0; c < width; ++c) scratch[i * width + c] = component(v[i], c);
    return scratch.data();
}

struct staging {
    std::vector<double> buf[3 * OCTOFMM_NDIR + 3];
    std::vector<int32_t> cells[OCTOFMM_NDIR];
};

template <typename Grid, typename Dir>
void ilist_cells(Grid& grid, Dir dir, std::vector<int32_t>& out) {
    out.clear();
    for (const auto& e : grid.get_ilist_n_bnd(dir)) out.push_back(static_cast<int32_t>(e.second));
}

// Fills an octofmm_subgrid from Octo-Tiger containers
// (update_input, multipole_interaction_interface.hpp:88-281).
template <typename real, typename multipole, typename space_vector, typename neighbor_gravity_type,
    typename Grid>
void marshal(octofmm_subgrid& sg, staging& st, const std::vector<real>* monopoles,
    const std::vector<multipole>* M, const std::vector<std::shared_ptr<std::vector<space_vector>>>& com_ptr,
    std::vector<neighbor_gravity_type>& neighbors, int type, real dx, const std::array<real, 3>& xbase,
    bool root, Grid* grid) {
    std::memset(&sg, 0, sizeof(sg));
    if (M && !M->empty()) sg.M = packed(*M, 20, st.buf[78]);
    sg.com = packed(*com_ptr[0], 3, st.buf[79]);
    if (monopoles && !monopoles->empty()) sg.mon = packed(*monopoles, 1, st.buf[80]);
    sg.dx = static_cast<double>(dx);
    for (int k = 0; k < 3; ++k) sg.xbase[k] = static_cast<double>(xbase[k]);
    sg.type = type;
    sg.use_root_stencil = root ? 1 : 0;
    if (neighbors.size() != OCTOFMM_NDIR) throw std::runtime_error("octofmm: 26 neighbours expected");
    for (int f = 0; f < OCTOFMM_NDIR; ++f) {
        auto& nb = neighbors[f];
        octofmm_neighbor& d = sg.neighbors[f];
        if (!nb.is_monopole) {
            if (!nb.data.M || !nb.data.x) {
                d.kind = OCTOFMM_NEIGHBOR_EMPTY;
                continue;
            }
            d.kind = OCTOFMM_NEIGHBOR_MULTIPOLE;
            d.count = static_cast<int32_t>(nb.data.M->size());
            d.M = packed(*nb.data.M, 20, st.buf[3 * f]);
            d.com = packed(*nb.data.x, 3, st.buf[3 * f + 1]);
        } else {
            if (!nb.data.m) {
                d.kind = OCTOFMM_NEIGHBOR_EMPTY;
                continue;
            }
            d.kind = OCTOFMM_NEIGHBOR_MONOPOLE;
            d.count = static_cast<int32_t>(nb.data.m->size());
            d.mon = packed(*nb.data.m, 1, st.buf[3 * f + 2]);
        }
        if (d.count != OCTOFMM_CELLS) {
            if (!grid) throw std::runtime_error("octofmm: partial halo needs the grid's interaction list");
            ilist_cells(*grid, nb.direction, st.cells[f]);
            if (st.cells[f].size() != static_cast<size_t>(d.count))
                throw std::runtime_error("octofmm: interaction list and halo sizes differ");
            d.cells = st.cells[f].data();
        }
    }
}

template <typename Expansion, typename SpaceVector>
void write_back(std::vector<Expansion>& L, std::vector<SpaceVector>* L_c, const double* l, const double* lc,
    int ncomp) {
    for (int u = 0; u < OCTOFMM_CELLS; ++u) {
        for (int c = 0; c < ncomp; ++c) L[u][c] += l[20 * u + c];
        if (L_c)
            for (int c = 0; c < 3; ++c) (*L_c)[u][c] = lc[3 * u + c];
    }
}

}   // namespace detail

// multipole_interaction_interface (non-leaf sub-grids, root included)
template <typename Grid>
class multipole_interaction_interface {
public:
    explicit multipole_interaction_interface(context& ctx) : ctx_(ctx) {}
    void set_grid_ptr(std::shared_ptr<Grid> ptr) { grid_ = std::move(ptr); }

    template <typename real, typename multipole, typename space_vector, typename neighbor_gravity_type,
        typename gsolve_type, size_t NDIR>
    void compute_multipole_interactions(std::vector<real>& monopoles, std::vector<multipole>& M_ptr,
        std::vector<std::shared_ptr<std::vector<space_vector>>>& com_ptr,
        std::vector<neighbor_gravity_type>& neighbors, gsolve_type type, real dx,
        std::array<bool, NDIR>& is_direction_empty, std::array<real, 3> xbase, const bool use_root_stencil) {
        (void)is_direction_empty;   // empty directions are the neighbours without data
        octofmm_subgrid sg;
        detail::marshal(sg, st_, &monopoles, &M_ptr, com_ptr, neighbors, static_cast<int>(type), dx, xbase,
            use_root_stencil, grid_.get());
        L_.assign(20 * OCTOFMM_CELLS, 0.0);
        Lc_.assign(3 * OCTOFMM_CELLS, 0.0);
        ctx_.check(octofmm_compute_multipole_interactions(ctx_.get(), &sg, L_.data(), Lc_.data()));
        const bool rho = static_cast<int>(type) == OCTOFMM_RHO;
        detail::write_back(grid_->get_L(), rho ? &grid_->get_L_c() : nullptr, L_.data(), Lc_.data(), 20);
    }

private:
    context& ctx_;
    std::shared_ptr<Grid> grid_;
    detail::staging st_;
    std::vector<double> L_, Lc_;
};

// monopole_interaction_interface (leaf sub-grids: p2p and p2m)
template <typename Grid>
class monopole_interaction_interface {
public:
    explicit monopole_interaction_interface(context& ctx) : ctx_(ctx) {}

    template <typename real, typename space_vector, typename neighbor_gravity_type, typename gsolve_type,
        size_t NDIR>
    void compute_interactions(const std::vector<real>& monopoles,
        std::vector<std::shared_ptr<std::vector<space_vector>>>& com_ptr,
        std::vector<neighbor_gravity_type>& neighbors, gsolve_type type, real dx,
        std::array<bool, NDIR>& is_direction_empty, std::shared_ptr<Grid>& grid_ptr,
        const bool contains_multipole_neighbor, std::array<real, 3> xbase = {0, 0, 0}) {
        (void)is_direction_empty;
        octofmm_subgrid sg;
        detail::marshal<real, std::array<real, 20>>(sg, st_, &monopoles, nullptr, com_ptr, neighbors,
            static_cast<int>(type), dx, xbase, false, grid_ptr.get());
        L_.assign(20 * OCTOFMM_CELLS, 0.0);
        Lc_.assign(3 * OCTOFMM_CELLS, 0.0);
        ctx_.check(octofmm_compute_monopole_interactions(ctx_.get(), &sg, contains_multipole_neighbor ? 1 : 0,
            L_.data(), Lc_.data()));
        const bool write_lc = static_cast<int>(type) == OCTOFMM_RHO && contains_multipole_neighbor;
        detail::write_back(grid_ptr->get_L(), write_lc ? &grid_ptr->get_L_c() : nullptr, L_.data(), Lc_.data(),
            4);
    }

private:
    context& ctx_;
    detail::staging st_;
    std::vector<double> L_, Lc_;
};

}   // namespace octofmm_adapter
