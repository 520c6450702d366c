// Drives include/octofmm_interface.hpp with stand-in types shaped like
// Octo-Tiger's (taylor<4,double>, space_vector, neighbor_gravity_type, grid):
// reads one 27-slot case from a binary file, runs the adapter, writes L, L_c.
//
// input : int32 target(0 multipole, 1 monopole) type root contains_multipole_neighbor,
//         double dx xbase[3], then per raw direction 0..26:
//         int32 kind (-1 empty, 0 non-leaf, 1 leaf), int32 n, int32 cells[n],
//         double M[n*20] com[n*3] mon[n]
//         (n = 512 with cells 0..511: a full sub-grid; n < 512: a partial halo whose
//         cells the stand-in grid lists in get_ilist_n_bnd, in the order of the arrays)
// output: double L[512*20] L_c[512*3]
#include <array>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <memory>
#include <utility>
#include <vector>

#include "octofmm_interface.hpp"

template <int N, class T>
struct taylor {
    std::array<T, 20> data{};
    T& operator[](int i) { return data[i]; }
    const T& operator[](int i) const { return data[i]; }
};
using multipole = taylor<4, double>;
using expansion = taylor<4, double>;
struct space_vector : std::array<double, 3> {};
struct gravity_boundary_type {
    std::shared_ptr<std::vector<multipole>> M;
    std::shared_ptr<std::vector<double>> m;
    std::shared_ptr<std::vector<space_vector>> x;
};
struct neighbor_gravity_type {
    gravity_boundary_type data;
    bool is_monopole = false;
    int direction = 0;
};
enum gsolve_type { RHO, DRHODT };
// grid::get_ilist_n_bnd's element type: `second` is the neighbour-side cell
// index (octotiger/interaction_types.hpp:134-150)
struct boundary_interaction_type {
    std::uint16_t second = 0;
};
struct grid {
    std::vector<expansion> L = std::vector<expansion>(512);
    std::vector<space_vector> L_c = std::vector<space_vector>(512);
    std::vector<boundary_interaction_type> ilist[27];
    std::vector<expansion>& get_L() { return L; }
    std::vector<space_vector>& get_L_c() { return L_c; }
    const std::vector<boundary_interaction_type>& get_ilist_n_bnd(int dir) { return ilist[dir]; }
};

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    std::ifstream in(argv[1], std::ios::binary);
    int32_t hdr[4];
    double dx, xb[3];
    in.read((char*)hdr, sizeof(hdr));
    in.read((char*)&dx, 8);
    in.read((char*)xb, 24);
    std::vector<multipole> M[27];
    std::vector<space_vector> X[27];
    std::vector<double> mon[27];
    int32_t kind[27];
    auto g = std::make_shared<grid>();
    for (int s = 0; s < 27; ++s) {
        int32_t n = 0;
        in.read((char*)&kind[s], 4);
        in.read((char*)&n, 4);
        if (n < 0 || n > 512) return 3;
        std::vector<int32_t> cells(n);
        in.read((char*)cells.data(), 4 * n);
        M[s].resize(n);
        X[s].resize(n);
        mon[s].resize(n);
        in.read((char*)M[s].data(), n * 20 * 8);
        in.read((char*)X[s].data(), n * 3 * 8);
        in.read((char*)mon[s].data(), n * 8);
        for (int32_t c : cells) {
            boundary_interaction_type b;
            b.second = static_cast<std::uint16_t>(c);
            g->ilist[s].push_back(b);
        }
    }
    if (!in) return 3;
    std::vector<neighbor_gravity_type> nbrs(26);
    for (int f = 0; f < 26; ++f) {
        const int raw = f < 13 ? f : f + 1;
        nbrs[f].direction = raw;
        if (kind[raw] == 0) {
            nbrs[f].data.M = std::make_shared<std::vector<multipole>>(M[raw]);
            nbrs[f].data.x = std::make_shared<std::vector<space_vector>>(X[raw]);
        } else if (kind[raw] == 1) {
            nbrs[f].is_monopole = true;
            nbrs[f].data.m = std::make_shared<std::vector<double>>(mon[raw]);
        }
    }
    std::vector<std::shared_ptr<std::vector<space_vector>>> com_ptr{std::make_shared<std::vector<space_vector>>(X[13])};
    std::array<bool, 26> empty{};
    const std::array<double, 3> xbase{xb[0], xb[1], xb[2]};
    const gsolve_type type = hdr[1] == 0 ? RHO : DRHODT;
    try {
        octofmm_adapter::context ctx(0.5, 0);
        if (hdr[0] == 0) {
            octofmm_adapter::multipole_interaction_interface<grid> mi(ctx);
            mi.set_grid_ptr(g);
            mi.compute_multipole_interactions(mon[13], M[13], com_ptr, nbrs, type, dx, empty, xbase, hdr[2] != 0);
        } else {
            octofmm_adapter::monopole_interaction_interface<grid> mo(ctx);
            mo.compute_interactions(mon[13], com_ptr, nbrs, type, dx, empty, g, hdr[3] != 0, xbase);
        }
        std::ofstream out(argv[2], std::ios::binary);
        out.write((const char*)g->L.data(), 512 * 20 * 8);
        out.write((const char*)g->L_c.data(), 512 * 3 * 8);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
