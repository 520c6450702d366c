// octofmm.hip -- liboctofmm: gfx950 kernels and the C-ABI of include/octofmm.h.
//
// Data path (one level pass):
//   1. stage_records: the caller's AoS state (M[cell][20], com[cell][3],
//      mon[cell]) becomes one 24-double partner record per cell in HBM,
//      ordered coarse-cell-major / child-minor inside each 8^3 slot.  The
//      record already holds what a partner contributes under the reference's
//      staging rules (multipole_interaction_interface.hpp:88-281): refined
//      cells give (M, com), leaf cells give a monopole at the geometric cell
//      centre.
//   2. level_kernel: one 512-lane workgroup per target sub-grid.  Wave w owns
//      a 2x2x2 block of coarse cells, lane = (coarse cell, child), so the 8
//      lanes of one coarse cell share every partner record they load.  Each
//      lane walks the coarse offsets of the opening-criterion stencil and the
//      8 children of each partner coarse cell, accumulating its 20-component
//      local expansion (and 3-component correction) in registers, and adds
//      them to L / writes L_c once at the end.
//   3. root: root_kernel splits the 512 x 512 all-pairs sweep of the root
//      sub-grid over partner chunks; root_reduce sums the chunks in a fixed
//      order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "kernels.hpp"
#include "octofmm.h"
#include "octofmm_internal.hpp"

using namespace octofmm;

// ===========================================================================
// Kernels
// ===========================================================================
namespace {

// Workgroups are handed to the 8 XCDs round-robin; give each XCD a contiguous
// run of items so neighbouring sub-grids share an L2.
__device__ __forceinline__ int xcd_item(int b, int n) {
    const int q = n >> 3, r = n & 7;
    const int x = b & 7, i = b >> 3;
    return x * q + (x < r ? x : r) + i;
}

// A store of data nothing in the same pass or solve reads through L2 again
// (staged records, which the level kernels read once per image from HBM, and
// a leaf's final rows): non-temporal, so the lines stream out without
// displacing what the concurrent kernels reuse.
__device__ __forceinline__ void store_stream(double2* p, double2 v) {
    __builtin_nontemporal_store(v.x, &p->x);
    __builtin_nontemporal_store(v.y, &p->y);
}

// One block per 128 cells of a slot (two x-planes).  Those cells' records are
// the 128 consecutive records of the block's coarse x-layer, so the block
// loads the cells' M and centre-of-mass rows (contiguous) into LDS in 16-byte
// pieces, then writes the 128 records as one contiguous range, consecutive
// threads in consecutive 16-byte chunks.  Loads and stores are both
// coalesced, where a lane per record spread every instruction over 64 rows.
__global__ __launch_bounds__(256) void stage_records(StageArgs a) {
    // leaf-cell centres are rounded operation by operation, as the reference does
#pragma clang fp contract(off)
    __shared__ double sM[128 * 20];
    __shared__ double sC[128 * 3];
    __shared__ double sm[128];
    const int si = blockIdx.x >> 2, q = blockIdx.x & 3;
    if (si >= a.nslots) return;                // uniform per block
    const int slot = a.list ? a.list[si] : si;
    const int k = a.kind[slot];
    const int c0 = q * 128;
    const int t = threadIdx.x;
    const size_t row0 = (size_t)slot * kCells + c0;
    if (k == OCTOFMM_SLOT_NONLEAF) {
        const double2* srcM = reinterpret_cast<const double2*>(a.M + row0 * 20);
        for (int i = t; i < 128 * 10; i += 256) reinterpret_cast<double2*>(sM)[i] = srcM[i];
        const double* srcC = a.com + row0 * 3;
        for (int i = t; i < 128 * 3; i += 256) sC[i] = srcC[i];
    } else if (k == OCTOFMM_SLOT_LEAF) {
        if (t < 128) sm[t] = a.mon[row0 + t];
    }
    __syncthreads();
    double2* dst = reinterpret_cast<double2*>(a.rec + row0 * kRecDoubles);
    for (int ch = t; ch < 128 * 12; ch += 256) {
        const int r = c0 + ch / 12, part = ch % 12;
        // the cell of record r: coarse cell r >> 3 (x-major), child r & 7
        const int co = r >> 3, ci = r & 7;
        const int x = 2 * (co >> 4) + ((ci >> 2) & 1), y = 2 * ((co >> 2) & 3) + ((ci >> 1) & 1),
                  z = 2 * (co & 3) + (ci & 1);
        const int cl = x * 64 + y * 8 + z - c0;
        double2 v = make_double2(0.0, 0.0);
        if (k == OCTOFMM_SLOT_NONLEAF) {
            if (part < 10) {
                v.x = sM[cl * 20 + 2 * part] * record_scale(2 * part);
                v.y = sM[cl * 20 + 2 * part + 1] * record_scale(2 * part + 1);
            } else if (part == 10) {
                v = make_double2(sC[cl * 3], sC[cl * 3 + 1]);
            } else {
                v = make_double2(sC[cl * 3 + 2], 0.0);
            }
        } else if (k == OCTOFMM_SLOT_LEAF) {
            const double m = sm[cl];
            // geometric centre in the target's frame (multipole_interaction_interface.hpp:216-235)
            const int* fo = a.foff + 3 * slot;
            const double* fb = a.fxbase + 3 * slot;
            if (part == 0) {
                v.x = m;
            } else if (part == 10) {
                v.x = (double)(x + fo[0]) * a.dx + fb[0] - kInx * a.dx;
                v.y = (double)(y + fo[1]) * a.dx + fb[1] - kInx * a.dx;
            } else if (part == 11) {
                v.x = (double)(z + fo[2]) * a.dx + fb[2] - kInx * a.dx;
                v.y = m;
            }
        }
        store_stream(dst + ch, v);
        if (a.recm) {
            // the mixed-precision record of the same cell: m0..m19 rounded to
            // FP32 (4 per chunk), (Y0, Y1), (Y2, m0), the octupole in FP64
            double2 w = make_double2(0.0, 0.0);
            if (k == OCTOFMM_SLOT_NONLEAF) {
                if (part < 5) {
                    float4 f;
                    f.x = (float)(sM[cl * 20 + 4 * part] * record_scale(4 * part));
                    f.y = (float)(sM[cl * 20 + 4 * part + 1] * record_scale(4 * part + 1));
                    f.z = (float)(sM[cl * 20 + 4 * part + 2] * record_scale(4 * part + 2));
                    f.w = (float)(sM[cl * 20 + 4 * part + 3] * record_scale(4 * part + 3));
                    w = __builtin_bit_cast(double2, f);
                } else if (part == kRecmY01) {
                    w = make_double2(sC[cl * 3], sC[cl * 3 + 1]);
                } else if (part == kRecmY2M0) {
                    w = make_double2(sC[cl * 3 + 2], sM[cl * 20]);
                } else {
                    const int c = 10 + 2 * (part - kRecmOct);
                    w = make_double2(sM[cl * 20 + c] * record_scale(c), sM[cl * 20 + c + 1] * record_scale(c + 1));
                }
            } else if (k == OCTOFMM_SLOT_LEAF) {
                const double m = sm[cl];
                const int* fo = a.foff + 3 * slot;
                const double* fb = a.fxbase + 3 * slot;
                if (part == 0) {
                    w = __builtin_bit_cast(double2, make_float4((float)m, 0.0f, 0.0f, 0.0f));
                } else if (part == kRecmY01) {
                    w.x = (double)(x + fo[0]) * a.dx + fb[0] - kInx * a.dx;
                    w.y = (double)(y + fo[1]) * a.dx + fb[1] - kInx * a.dx;
                } else if (part == kRecmY2M0) {
                    w.x = (double)(z + fo[2]) * a.dx + fb[2] - kInx * a.dx;
                    w.y = m;
                }
            }
            store_stream(reinterpret_cast<double2*>(a.recm + row0 * kRecDoubles) + ch, w);
        }
    }
}

struct LaneGeom {
    int Cx, Cy, Cz;   // coarse cell in the sub-grid (0..3)
    int k;            // child index
    int cell;         // x*64 + y*8 + z
    int rec;          // record index inside the slot
};

// wave w (0..7 across the sub-grid) owns coarse block ((w>>2)&1, (w>>1)&1, w&1);
// t is the thread's index in the workgroup
__device__ __forceinline__ LaneGeom lane_geom(int w, int t = threadIdx.x) {
    LaneGeom g;
    const int l = t & 63;
    const int c = l >> 3;
    g.k = l & 7;
    g.Cx = ((w >> 2) & 1) * 2 + ((c >> 2) & 1);
    g.Cy = ((w >> 1) & 1) * 2 + ((c >> 1) & 1);
    g.Cz = (w & 1) * 2 + (c & 1);
    const int fx = 2 * g.Cx + ((g.k >> 2) & 1), fy = 2 * g.Cy + ((g.k >> 1) & 1), fz = 2 * g.Cz + (g.k & 1);
    g.cell = fx * 64 + fy * 8 + fz;
    g.rec = (g.Cx * 16 + g.Cy * 4 + g.Cz) * 8 + g.k;
    return g;
}

// Partner coarse cell C + delta -> (slot, coarse index in the slot); slot < 0
// when the direction is empty.
__device__ __forceinline__ int partner_base(const LaneGeom& g, int ax, int ay, int az, const int* s_nbr,
    int& nslot, int& raw) {
    const int px = g.Cx + ax, py = g.Cy + ay, pz = g.Cz + az;
    const int rx = px < 0 ? 0 : (px < 4 ? 1 : 2);
    const int ry = py < 0 ? 0 : (py < 4 ? 1 : 2);
    const int rz = pz < 0 ? 0 : (pz < 4 ? 1 : 2);
    raw = rx + 3 * ry + 9 * rz;
    nslot = s_nbr[raw];
    return ((px & 3) * 16 + (py & 3) * 4 + (pz & 3)) * 8;
}

// Result of one target cell.  Level plans accumulate L (ncomp components:
// 20 for M2L targets, 4 for leaves) and overwrite L_c when the kernel forms
// it (write_c).  In a whole-tree solve the cell's final expansion is written
// directly: interactions + the parent cell's expansion re-expanded about the
// cell's centre of mass (taylor::operator<<=, grid::compute_expansions,
// src/grid_fmm.cpp:1524-1637), L_c + the parent's L_c, and for leaf cells
// phi = G L0 and g = -G (L_1..3 + L_c).
// The final expansion of a cell in a whole-tree solve: its interactions plus
// the parent cell's expansion re-expanded about the cell's centre of mass, and
// L_c plus the parent's L_c.
template <bool RHO>
__device__ __forceinline__ void tree_final_row(const LevelArgs& a, int slot, int cell, const double* L, int ncomp,
    const double* C, bool write_c, double* Lf, double* Cf) {
    double l[20], lc[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < 20; ++j) l[j] = 0.0;
    const int pn = a.parent[slot];
    if (pn >= 0) {
        const int pc = parent_cell(a.octant[slot], cell >> 7, (cell >> 4) & 3, (cell >> 1) & 3);
        double A[20];
        const double2* src = reinterpret_cast<const double2*>(a.pL + ((size_t)pn * kCells + pc) * 20);
#pragma unroll
        for (int j = 0; j < 10; ++j) {
            const double2 v = src[j];
            A[2 * j] = v.x;
            A[2 * j + 1] = v.y;
        }
        const double* Y = a.pcom + ((size_t)pn * kCells + pc) * 3;
        const double* X = a.com + ((size_t)slot * kCells + cell) * 3;
        const double dX[3] = {X[0] - Y[0], X[1] - Y[1], X[2] - Y[2]};
        local_shift(A, dX, l);
        if (RHO) {
            const double* q = a.pLc + ((size_t)pn * kCells + pc) * 3;
            lc[0] = q[0];
            lc[1] = q[1];
            lc[2] = q[2];
        }
    }
#pragma unroll
    for (int j = 0; j < 20; ++j) Lf[j] = (j < ncomp ? L[j] : 0.0) + l[j];
#pragma unroll
    for (int d = 0; d < 3; ++d) Cf[d] = RHO ? (write_c ? C[d] : 0.0) + lc[d] : 0.0;
}

// Leaf force assembly: phi = G L0, g = -G (L_1..3 + L_c).
template <bool RHO>
__device__ __forceinline__ void leaf_force(const LevelArgs& a, const double* Lf, const double* Cf, double* g4) {
    g4[0] = a.G * Lf[0];
#pragma unroll
    for (int d = 0; d < 3; ++d) g4[1 + d] = RHO ? -a.G * Lf[1 + d] - (a.correct_am ? a.G * Cf[d] : 0.0) : 0.0;
}

// Expansion epilogue of one target cell.  Outside a whole-tree solve:
// accumulate the cell's L (the ncomp leading components: 20 for M2L
// targets, 4 for leaves) and overwrite L_c when the kernel forms it
// (write_c).  In a whole-tree solve the cell's final expansion is written
// directly: interactions + the parent cell's expansion re-expanded about the
// cell's centre of mass (taylor::operator<<=, grid::compute_expansions,
// src/grid_fmm.cpp:1524-1637), L_c + the parent's L_c, and for leaf cells
// phi = G L0 and g = -G (L_1..3 + L_c).
template <bool RHO>
__device__ __forceinline__ void finish_cell(const LevelArgs& a, int slot, int cell, const double* L, int ncomp,
    const double* C, bool write_c, bool leaf_target) {
    // expansion rows are 160 bytes and 16-byte aligned: moved as 10 double2
    double2* Lo = reinterpret_cast<double2*>(a.L + ((size_t)slot * kCells + cell) * 20);
    double* Co = a.Lc + ((size_t)slot * kCells + cell) * 3;
    if (!a.l2l) {
        for (int i = 0; i < ncomp / 2; ++i) {
            double2 v = Lo[i];
            v.x += L[2 * i];
            v.y += L[2 * i + 1];
            Lo[i] = v;
        }
        if (RHO && write_c) {
            Co[0] = C[0];
            Co[1] = C[1];
            Co[2] = C[2];
        }
        return;
    }
    double Lf[20], Cf[3];
    tree_final_row<RHO>(a, slot, cell, L, ncomp, C, write_c, Lf, Cf);
    if (!leaf_target || a.leaf_rows) {
#pragma unroll
        for (int j = 0; j < 10; ++j) Lo[j] = make_double2(Lf[2 * j], Lf[2 * j + 1]);
    }
    if (RHO) {
#pragma unroll
        for (int d = 0; d < 3; ++d) Co[d] = Cf[d];
    }
    if (a.g && leaf_target) {
        double2* go = reinterpret_cast<double2*>(a.g + ((size_t)slot * kCells + cell) * 4);
        double g4[4];
        leaf_force<RHO>(a, Lf, Cf, g4);
        if (RHO) {
            go[0] = make_double2(g4[0], g4[1]);
            go[1] = make_double2(g4[2], g4[3]);
        } else {
            reinterpret_cast<double*>(go)[0] = g4[0];   // DRHODT: phi slot only
        }
    }
}

// Leaf epilogue of a whole-tree RHO solve, with the rows staged through LDS so
// that the sub-grid's L, L_c and g blocks (contiguous: 512 rows of 20, 3 and 4
// doubles) are written by consecutive lanes in consecutive 16-byte chunks.
// Written straight from the class waves, each store instruction would touch
// 64 rows two cells apart.  L goes in two halves of 256 cells (40 KiB of
// LDS), then L_c and g together (28 KiB); without a.leaf_rows only L_c and
// g.  Every thread of the workgroup calls it; buf is the kernel's dynamic
// LDS, no longer in use.
__device__ __forceinline__ void leaf_rows_coalesced(const LevelArgs& a, int slot, int cell, const double* Lf,
    const double* Cf, const double* g4, double* __restrict__ buf) {
    // Each staged row, L_c triple and g quadruple is shifted by one more
    // element every 16 cells (c >> 4): the lanes of a wave hold cells 2 apart
    // in z, 16 in y and 128 in x, and without the shift their rows fall on a
    // quarter of the LDS banks (8-way conflicts on the row writes, 4-way on
    // L_c, 8-way on g); with it each half-wave spreads over all of them.
    const int t = threadIdx.x;
    double2* b2 = reinterpret_cast<double2*>(buf);
    for (int half = 0; half < (a.leaf_rows ? 2 : 0); ++half) {   // uniform
        __syncthreads();
        if ((cell >> 8) == half) {
            const int c = cell & 255;
            double2* row = b2 + c * 10 + (c >> 4);
#pragma unroll
            for (int j = 0; j < 10; ++j) row[j] = make_double2(Lf[2 * j], Lf[2 * j + 1]);
        }
        __syncthreads();
        double2* dst = reinterpret_cast<double2*>(a.L + ((size_t)slot * kCells + half * 256) * 20);
#pragma unroll
        for (int r = 0; r < 5; ++r) {   // 256 rows x 10 chunks
            const int ch = t + 512 * r;
            store_stream(dst + ch, b2[ch + (ch / 10 >> 4)]);
        }
    }
    __syncthreads();
    buf[cell * 3 + (cell >> 4) + 0] = Cf[0];
    buf[cell * 3 + (cell >> 4) + 1] = Cf[1];
    buf[cell * 3 + (cell >> 4) + 2] = Cf[2];
    double2* bg = b2 + 784;   // after the 1536 + 32 doubles of L_c
    bg[cell * 2 + (cell >> 4)] = make_double2(g4[0], g4[1]);
    bg[cell * 2 + (cell >> 4) + 1] = make_double2(g4[2], g4[3]);
    __syncthreads();
    double2* lc = reinterpret_cast<double2*>(a.Lc + (size_t)slot * kCells * 3);
    double2* gg = reinterpret_cast<double2*>(a.g + (size_t)slot * kCells * 4);
    auto lc_pair = [&](int ch) {   // doubles 2 ch and 2 ch + 1 of the L_c block
        const int d0 = 2 * ch, d1 = d0 + 1;
        return make_double2(buf[d0 + (d0 / 3 >> 4)], buf[d1 + (d1 / 3 >> 4)]);
    };
    store_stream(lc + t, lc_pair(t));                 // 768 chunks of L_c
    if (t < 256) store_stream(lc + 512 + t, lc_pair(512 + t));
    store_stream(gg + t, bg[t + (t >> 5)]);           // 1024 chunks of g (cell = chunk / 2)
    store_stream(gg + 512 + t, bg[512 + t + ((512 + t) >> 5)]);
}

template <bool RHO, bool MIXED>
__device__ void m2l_item(const LevelArgs& a, int slot, int flags, const int* s_nbr, const int* s_kind, int w) {
    const LaneGeom g = lane_geom(w);
    const Tables* __restrict__ tab = a.tab;
    const double* trec = a.rec + ((size_t)slot * kCells + g.rec) * kRecDoubles;
    const double X0 = trec[kRecY], X1 = trec[kRecY + 1], X2 = trec[kRecY + 2];
    const double inv_mc0 = RHO ? 1.0 / trec[0] : 0.0;
    double mc[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) mc[i] = RHO ? trec[10 + i] : 0.0;
    double L[20], C[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 20; ++i) L[i] = 0.0;

    const int nd = tab->ndelta;
    // far pass, software-pipelined: the record of the next pair is loaded
    // (by every lane, from a valid address) while the current one is computed
    auto far_base = [&](int di, unsigned& fb) -> const double* {
        const int ax = tab->delta[di][0], ay = tab->delta[di][1], az = tab->delta[di][2];
        fb = (unsigned)(tab->farbits[di] >> (8 * g.k)) & 0xffu;
        int ns, raw;
        const int cb = partner_base(g, ax, ay, az, s_nbr, ns, raw);
        if (ns < 0) fb = 0;
        return a.rec + ((size_t)(ns < 0 ? 0 : ns) * kCells + cb) * kRecDoubles;
    };
    unsigned fb;
    const double* pb = far_base(0, fb);
    Rec nxt;
    if constexpr (!MIXED) load_rec(pb, nxt);
    float Lf[20];
    if constexpr (MIXED) {
#pragma unroll
        for (int i = 0; i < 20; ++i) Lf[i] = 0.0f;
    }
    for (int di = 0; di < nd; ++di) {
        unsigned fbn = 0;
        const double* pbn = pb;
        const unsigned mb = MIXED ? (unsigned)(a.mixbits[di] >> (8 * g.k)) & 0xffu : 0u;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            // FP64: the next record is prefetched while this pair computes;
            // MIXED keeps a single record live (its FP32 partial sums need
            // the registers)
            Rec cur;
            if constexpr (MIXED) {
                load_rec(pb + j * kRecDoubles, cur);
                if (j == 7 && di + 1 < nd) pbn = far_base(di + 1, fbn);
            } else {
                cur = nxt;
                if (j < 7) {
                    load_rec(pb + (j + 1) * kRecDoubles, nxt);
                } else if (di + 1 < nd) {
                    pbn = far_base(di + 1, fbn);
                    load_rec(pbn, nxt);
                }
            }
            if (fb & (1u << j)) {
                if (MIXED && (mb & (1u << j)))
                    m2l_far_mixed<RHO>(cur, X0, X1, X2, inv_mc0, mc, Lf, C);
                else
                    m2l_far_rec<RHO>(cur, X0, X1, X2, inv_mc0, mc, L, C);
            }
        }
        if constexpr (MIXED) {
            // the coarse offset's FP32 partial sums join the FP64 expansion
#pragma unroll
            for (int i = 0; i < 20; ++i) {
                L[i] += (double)Lf[i];
                Lf[i] = 0.0f;
            }
        }
        pb = pbn;
        fb = fbn;
    }
    if (flags & kItemNear) {
        for (int di = 0; di < nd; ++di) {
            const int ax = tab->delta[di][0], ay = tab->delta[di][1], az = tab->delta[di][2];
            unsigned nb = (unsigned)(tab->nearbits[di] >> (8 * g.k)) & 0xffu;
            int ns, raw;
            const int cb = partner_base(g, ax, ay, az, s_nbr, ns, raw);
            if (ns < 0 || s_kind[raw] != OCTOFMM_SLOT_LEAF) nb = 0;
            const double* pb = a.rec + ((size_t)(ns < 0 ? 0 : ns) * kCells + cb) * kRecDoubles;
#pragma unroll 1
            for (int j = 0; j < 8; ++j) {
                if (nb & (1u << j)) {
                    const double* p = pb + j * kRecDoubles;
                    m2l_near<RHO>(p[kRecMon], p[kRecY], p[kRecY + 1], p[kRecY + 2], X0, X1, X2, inv_mc0, mc, L, C);
                }
            }
        }
    }
    finish_cell<RHO>(a, slot, g.cell, L, 20, C, true, false);
}

// Mapping 5: the coarse offsets of one 2x2x2 coarse block are split over
// SPLIT waves of a workgroup (wave s takes offsets s, s + SPLIT, ...), so a
// sub-grid runs on 8 * SPLIT waves; this shortens the serial time of one
// item for levels with few sub-grids.  Waves s > 0 leave their partial
// expansions in LDS and wave 0 adds them in the order s = 1, 2, ... before
// writing the result.
template <bool RHO, int SPLIT>
__device__ void m2l_item_split(const LevelArgs& a, int slot, int flags, const int* s_nbr, const int* s_kind, int w,
    int sl, double* __restrict__ s_part) {
    const LaneGeom g = lane_geom(w);
    const int lane = threadIdx.x & 63;
    const Tables* __restrict__ tab = a.tab;
    const double* trec = a.rec + ((size_t)slot * kCells + g.rec) * kRecDoubles;
    const double X0 = trec[kRecY], X1 = trec[kRecY + 1], X2 = trec[kRecY + 2];
    const double inv_mc0 = RHO ? 1.0 / trec[0] : 0.0;
    double mc[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) mc[i] = RHO ? trec[10 + i] : 0.0;
    double L[20], C[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 20; ++i) L[i] = 0.0;
    const int nd = tab->ndelta;
    for (int di = sl; di < nd; di += SPLIT) {
        const int ax = tab->delta[di][0], ay = tab->delta[di][1], az = tab->delta[di][2];
        unsigned fb = (unsigned)(tab->farbits[di] >> (8 * g.k)) & 0xffu;
        int ns, raw;
        const int cb = partner_base(g, ax, ay, az, s_nbr, ns, raw);
        if (ns < 0) fb = 0;
        const double* pb = a.rec + ((size_t)(ns < 0 ? 0 : ns) * kCells + cb) * kRecDoubles;
        Rec nxt;
        load_rec(pb, nxt);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const Rec cur = nxt;
            if (j < 7) load_rec(pb + (j + 1) * kRecDoubles, nxt);
            if (fb & (1u << j)) m2l_far_rec<RHO>(cur, X0, X1, X2, inv_mc0, mc, L, C);
        }
        if (flags & kItemNear) {
            unsigned nb = (unsigned)(tab->nearbits[di] >> (8 * g.k)) & 0xffu;
            if (ns < 0 || s_kind[raw] != OCTOFMM_SLOT_LEAF) nb = 0;
#pragma unroll 1
            for (int j = 0; j < 8; ++j) {
                if (nb & (1u << j)) {
                    const double* p = pb + j * kRecDoubles;
                    m2l_near<RHO>(p[kRecMon], p[kRecY], p[kRecY + 1], p[kRecY + 2], X0, X1, X2, inv_mc0, mc, L, C);
                }
            }
        }
    }
    if (sl > 0) {
        double* o = s_part + ((size_t)(sl - 1) * 64 + lane) * 23;
#pragma unroll
        for (int i = 0; i < 20; ++i) o[i] = L[i];
        o[20] = C[0];
        o[21] = C[1];
        o[22] = C[2];
    }
    __syncthreads();
    if (sl == 0) {
        for (int q = 0; q < SPLIT - 1; ++q) {
            const double* o = s_part + ((size_t)q * 64 + lane) * 23;
#pragma unroll
            for (int i = 0; i < 20; ++i) L[i] += o[i];
            C[0] += o[20];
            C[1] += o[21];
            C[2] += o[22];
        }
        finish_cell<RHO>(a, slot, g.cell, L, 20, C, true, false);
    }
}

// Mapping-5 kernel: one workgroup of SPLIT waves per 2x2x2 coarse block.
template <bool RHO, int SPLIT>
__global__ __launch_bounds__(64 * SPLIT) void level_kernel_split(LevelArgs a) {
    __shared__ int s_nbr[27];
    __shared__ int s_kind[27];
    __shared__ double s_part[(SPLIT - 1) * 64 * 23];
    const int v = xcd_item(blockIdx.x, gridDim.x);
    const int item = v >> 3, w = v & 7;
    const int2 it = a.items[a.item0 + item];
    if (threadIdx.x < 27) {
        const int n = a.nbr[(size_t)it.x * 27 + threadIdx.x];
        s_nbr[threadIdx.x] = n;
        s_kind[threadIdx.x] = n >= 0 ? a.kind[n] : -1;
    }
    __syncthreads();
    if (it.y & kItemLeaf) return;
    m2l_item_split<RHO, SPLIT>(a, it.x, it.y, s_nbr, s_kind, w, (int)(threadIdx.x >> 6), s_part);
}

// Waits for this wave's LDS writes (lgkmcnt(0), global loads left in flight)
// and keeps the compiler from moving LDS accesses across the point.  The
// lanes of a wave execute each LDS instruction together and in order, so the
// written records are then visible to every lane of the wave.
__device__ __forceinline__ void lds_wave_sync() {
    __builtin_amdgcn_s_waitcnt(0xC07F);   // vmcnt = 63, expcnt = 7, lgkmcnt = 0
    __builtin_amdgcn_wave_barrier();
}

// Mapping 3: the partner records of each coarse offset staged once per wave
// in LDS by direct global->LDS loads
// (global_load_lds_dwordx4): the 64 partner records a wave needs for one
// coarse offset (8 partner coarse cells x 8 children, 12 KiB) go from L2
// into the wave's LDS image without passing through registers, 16 bytes per
// lane and instruction; the 8 lanes of a coarse cell then read each record as
// an LDS broadcast.  Each group of 96 chunks is followed by one unused pad
// chunk so the 8 groups read together fall into different banks.
constexpr int kDirGroup = 97;              // 96 chunks + 1 pad
constexpr int kDirWave = 8 * kDirGroup;    // 776 double2 = 12.1 KiB per wave

typedef __attribute__((address_space(3))) void* lds_ptr;
// Read-only data seen through the constant address space: uniform loads stay
// scalar even after the kernel has issued LDS-direct loads.
// Forms the 8 values in scalar registers at this point, so that the scalar
// loads they depend on are in flight together (the constraint exists only in
// the device pass).
__device__ __forceinline__ void form_sgprs(const unsigned (&v)[8]) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" ::"s"(v[0]), "s"(v[1]), "s"(v[2]), "s"(v[3]), "s"(v[4]), "s"(v[5]), "s"(v[6]), "s"(v[7]));
#else
    (void)v;
#endif
}

// A uniform address formed in scalar registers at this point: a load that
// adds a 32-bit lane offset to it then takes the scalar-base form, with no
// per-lane 64-bit address arithmetic (the constraint exists only in the
// device pass).
__device__ __forceinline__ const char* uniform_ptr(const char* p) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+s"(p));
#endif
    return p;
}

template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* as_const(const T* p) {
    return (const __attribute__((address_space(4))) T*)p;
}

// SPLIT = 1: mapping 3, the wave walks every offset of its coarse block.
// SPLIT = 2, 4 or 8: mapping 7, the SPLIT waves of the workgroup share one
// coarse block and wave sl walks offsets sl, sl + SPLIT, ... (far and near);
// the other waves then hand their partial sums to wave 0 through their own
// images, and wave 0 adds them in wave order, so the sum is the same on every
// run.
//
// MIXED (mixed-precision plans): a far offset all of whose far pairs are
// evaluated in FP32 (a.mixoff) stages the mixed-precision records instead and
// runs m2l_far_m32 on them, with the offset's FP32 partial sums added to the
// FP64 expansion after its 8 children; the other far offsets run in FP64.
// The precision is uniform over the wave for each offset, so no lane waits
// for the other precision's arithmetic.
template <bool RHO, int SPLIT = 1, bool MIXED = false>
__device__ void m2l_item_direct(const LevelArgs& a, int slot, int flags, const int* s_nbr, const int* s_kind,
    int w, double2* __restrict__ s_stage, int sl = 0) {
    // this wave's image (wave-uniform address)
    double2* stage = s_stage + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) * kDirWave;
    // an opaque copy of the lane index: lane values formed from it are formed
    // where they are used instead of being held across the pair loop, where
    // the RHO kernels run at their register limit
    auto lane = []() -> int {
        int t;
        asm volatile("v_mov_b32 %0, %1" : "=v"(t) : "v"((int)threadIdx.x));
        return t & 63;
    };
    const auto* tab = as_const(a.tab);
    const int wu = __builtin_amdgcn_readfirstlane(w);
    const double* trec = a.rec + ((size_t)slot * kCells + lane_geom(wu).rec) * kRecDoubles;
    const double X0 = trec[kRecY], X1 = trec[kRecY + 1], X2 = trec[kRecY + 2];
    const double inv_mc0 = RHO ? 1.0 / trec[0] : 0.0;
    double mc[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) mc[i] = RHO ? trec[10 + i] : 0.0;
    double L[20], C[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 20; ++i) L[i] = 0.0;
    const int nd = tab->ndelta;
    const int bx = ((wu >> 2) & 1) * 2, by = ((wu >> 1) & 1) * 2, bz = (wu & 1) * 2;
    const auto* nbr_row = as_const(a.nbr + (size_t)__builtin_amdgcn_readfirstlane(slot) * 27);
    // record index (slot * 512 + first child) of the partner coarse cell of
    // group gi (coarse cell gi of the wave's block) at offset a; empty
    // directions are marked in bit gi of `empty`.  32-bit indices keep the 8
    // bases of the next image in 8 scalar registers across the pair loop.
    auto group_rec = [&](int gi, int ax, int ay, int az, unsigned& empty) -> unsigned {
        const int px = bx + ((gi >> 2) & 1) + ax, py = by + ((gi >> 1) & 1) + ay, pz = bz + (gi & 1) + az;
        const int raw = (px < 0 ? 0 : (px < 4 ? 1 : 2)) + 3 * (py < 0 ? 0 : (py < 4 ? 1 : 2)) +
            9 * (pz < 0 ? 0 : (pz < 4 ? 1 : 2));
        const int ns = nbr_row[raw];
        const int cb = ((px & 3) * 16 + (py & 3) * 4 + (pz & 3)) * 8;
        if (ns < 0) empty |= 1u << gi;
        return (unsigned)(ns < 0 ? 0 : ns) * (unsigned)kCells + (unsigned)cb;
    };
    // Group c (8 records = 96 chunks) is loaded by one full instruction
    // (chunks 0..63) and one on the lower 32 lanes (chunks 64..95): both are a
    // uniform base plus the lane's 16-byte offset, and the LDS destination is
    // uniform too, so a refill costs one address add per load instruction.
    // The 8 partner bases of an offset are formed ahead of its refill, their
    // neighbour-slot loads in flight together (one wait, taken while the
    // current image is still arriving, instead of one wait per group).
    unsigned nrec[8];
    const double* src_next = a.rec;   // the records the entry bases() last formed is staged from
    unsigned empty_next = 0;    // groups of that entry whose partner (a) slot is empty
    unsigned emptyb_next = 0;   // the same for offset b of a split entry (formed by issue())
    int entry_next = 0;         // the schedule entry bases() last formed
    auto bases = [&](int q, const double* src) {
        const int e = tab->sched[q];
        const int di = e & 0xff;
        const int ax = tab->delta[di][0], ay = tab->delta[di][1], az = tab->delta[di][2];
        unsigned empty = 0;
#pragma unroll
        for (int gi = 0; gi < 8; ++gi) nrec[gi] = group_rec(gi, ax, ay, az, empty);
        form_sgprs(nrec);
        empty_next = empty;
        entry_next = e;
        src_next = src;
    };
    // The full-wave loads of all groups come first and the half-wave loads
    // follow in one branch of their own: with a full and a half load of one
    // group on the two sides of a branch, the compiler may merge them into one
    // load whose LDS base differs by lane and is then taken from lane 0, which
    // sends the other lanes' chunks to the wrong place.
    //
    // A split entry (kinds 1 and 2) stages 4 records of offset a's partner
    // and 4 of offset b's per group, in two loads on lanes 0..47: lane l
    // moves chunk l % 12 of the half's (l / 12)-th child with the given bit
    // on the entry's axis.  Offset b is -a; its partner bases are formed once
    // a's loads are issued (split entries
    // always stage the FP64 records).  The LDS bases of all loads, and the lane
    // offsets of the split loads, come from opaque copies, so that they are
    // formed here and not held across the pair loop, where the kernel runs at
    // its register limit.
    auto issue = [&]() {
        const int kind = (entry_next >> 16) & 0xff;
        double2* st = stage;
        asm volatile("" : "+s"(st));
        const int lo = lane();
        if (kind == 0) {
            const char* rb = reinterpret_cast<const char*>(src_next);
            const unsigned loff = (unsigned)lo * 16u;
#pragma unroll
            for (int gi = 0; gi < 8; ++gi)
                __builtin_amdgcn_global_load_lds(uniform_ptr(rb + (size_t)nrec[gi] * kRecBytes) + loff,
                                                 (lds_ptr)(st + gi * kDirGroup), 16, 0, 0);
            if (lo < 32) {
#pragma unroll
                for (int gi = 0; gi < 8; ++gi)
                    __builtin_amdgcn_global_load_lds(uniform_ptr(rb + (size_t)nrec[gi] * kRecBytes + 1024) + loff,
                                                     (lds_ptr)(st + gi * kDirGroup + 64), 16, 0, 0);
            }
        } else {
            const char* rb = reinterpret_cast<const char*>(a.rec);
            const int axis = entry_next >> 24;
            const int r = lo / 12, part = lo - 12 * r;
            const int low = (1 << axis) - 1;
            const int jr = ((r & ~low) << 1) | (r & low);   // r-th child with a 0 bit on the axis
            const int va = kind == 1 ? 1 : 0;               // the axis bit of a's staged children
            const unsigned offa = (unsigned)(((jr | (va << axis)) * 12 + part) * 16);
            const unsigned offb = (unsigned)(((jr | ((1 - va) << axis)) * 12 + part) * 16);
            if (lo < 48) {
#pragma unroll
                for (int gi = 0; gi < 8; ++gi)
                    __builtin_amdgcn_global_load_lds(uniform_ptr(rb + (size_t)nrec[gi] * kRecBytes) + offa,
                                                     (lds_ptr)(st + gi * kDirGroup), 16, 0, 0);
            }
            const int di = entry_next & 0xff;
            const int ax = tab->delta[di][0], ay = tab->delta[di][1], az = tab->delta[di][2];
            unsigned brec[8];
            unsigned emptyb = 0;
#pragma unroll
            for (int gi = 0; gi < 8; ++gi) brec[gi] = group_rec(gi, -ax, -ay, -az, emptyb);
            form_sgprs(brec);
            emptyb_next = emptyb;
            if (lo < 48) {
#pragma unroll
                for (int gi = 0; gi < 8; ++gi)
                    __builtin_amdgcn_global_load_lds(uniform_ptr(rb + (size_t)brec[gi] * kRecBytes) + offb,
                                                     (lds_ptr)(st + gi * kDirGroup + 48), 16, 0, 0);
            }
        }
    };
    // only the offsets with a far pair are staged (delta 0 is all near / self),
    // in the order of the far schedule
    const int nfd = tab->nsched;
    const int q0 = SPLIT == 1 ? 0 : __builtin_amdgcn_readfirstlane(sl);
    const auto* mixoff = as_const(a.mixoff);
    // the records entry q is staged from
    auto source = [&](int q) -> const double* { return (MIXED && mixoff[q]) ? a.recm : a.rec; };
    if (q0 < nfd) {
        bases(q0, source(q0));
        issue();
    }
    // the lane's group of 8 records in the image
    const double2* group = stage + (lane() >> 3) * kDirGroup;
    for (int q = q0; q < nfd; q += SPLIT) {
        // the lanes that read record j of the image: those of the classes in
        // byte j of the entry's step masks, unless their group's partner slot
        // is empty; kind 2 classes on the far side of the axis read records
        // 4..7 as their records 0..3.  The masks are formed in scalar
        // registers, one 64-bit lane mask per step (lane = group * 8 + class).
        const uint64_t ra = tab->sched_rowa[q], rb = tab->sched_rowb[q];
        // (emptyb_next is formed between the branches of issue(), after which
        // the compiler no longer knows it to be uniform)
        const uint64_t ea = spread_groups(empty_next);
        const uint64_t eb = spread_groups(__builtin_amdgcn_readfirstlane(emptyb_next));
        auto step_mask = [&](int j) -> uint64_t { return step_lanes(ra, rb, ea, eb, j); };
        const double2* recs =
            __builtin_amdgcn_inverse_ballot_w64(repeat_bytes(tab->sched_half[q] & 0xffu)) ? group + 48 : group;
        const bool f32 = MIXED && mixoff[q];   // uniform
        if (q + SPLIT < nfd) bases(q + SPLIT, source(q + SPLIT));
        __builtin_amdgcn_s_waitcnt(0x3F70);   // vmcnt(0): this wave's image is in LDS
        __builtin_amdgcn_wave_barrier();
        if (MIXED && f32) {
            float Lf[20];
#pragma unroll
            for (int i = 0; i < 20; ++i) Lf[i] = 0.0f;
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (__builtin_amdgcn_inverse_ballot_w64(step_mask(j)))
                    m2l_far_m32<RHO>(recs + j * 12, X0, X1, X2, inv_mc0, mc, Lf, C);
            // the offset's FP32 partial sums join the FP64 expansion
#pragma unroll
            for (int i = 0; i < 20; ++i) L[i] += (double)Lf[i];
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                if (__builtin_amdgcn_inverse_ballot_w64(step_mask(j))) {
                    Rec cur;
#pragma unroll
                    for (int i = 0; i < 12; ++i) cur.v[i] = recs[j * 12 + i];
                    m2l_far_rec<RHO>(cur, X0, X1, X2, inv_mc0, mc, L, C);
                }
            }
        }
        lds_wave_sync();   // every read of the image done before it is refilled
        if (q + SPLIT < nfd) issue();
    }
    // The lane's geometry again, from an opaque copy of the thread index: the
    // far loop needs none of it, and formed here it is not held in registers
    // across that loop (where the RHO kernels run at the register limit).
    int t2;
    asm volatile("v_mov_b32 %0, %1" : "=v"(t2) : "v"((int)threadIdx.x));
    const LaneGeom g = lane_geom(wu, t2);
    if (flags & kItemNear) {
        for (int di = q0; di < nd; di += SPLIT) {
            const int ax = tab->delta[di][0], ay = tab->delta[di][1], az = tab->delta[di][2];
            unsigned nb = (unsigned)(tab->nearbits[di] >> (8 * g.k)) & 0xffu;
            int ns, raw;
            const int cb = partner_base(g, ax, ay, az, s_nbr, ns, raw);
            if (ns < 0 || s_kind[raw] != OCTOFMM_SLOT_LEAF) nb = 0;
            const double* pb2 = a.rec + ((size_t)(ns < 0 ? 0 : ns) * kCells + cb) * kRecDoubles;
#pragma unroll 1
            for (int j = 0; j < 8; ++j) {
                if (nb & (1u << j)) {
                    const double* p = pb2 + j * kRecDoubles;
                    m2l_near<RHO>(p[kRecMon], p[kRecY], p[kRecY + 1], p[kRecY + 2], X0, X1, X2, inv_mc0, mc, L, C);
                }
            }
        }
    }
    if constexpr (SPLIT > 1) {
        // no refill is in flight and every read of the image is done, so the
        // image can hold this wave's 64 x 23 partial sums (11.5 KiB)
        double* part = reinterpret_cast<double*>(stage) + (threadIdx.x & 63) * 23;
        if (q0 != 0) {
#pragma unroll
            for (int i = 0; i < 20; ++i) part[i] = L[i];
            part[20] = C[0];
            part[21] = C[1];
            part[22] = C[2];
        }
        __syncthreads();
        if (q0 != 0) return;
        for (int o = 1; o < SPLIT; ++o) {
            const double* po = reinterpret_cast<const double*>(s_stage + o * kDirWave) + (threadIdx.x & 63) * 23;
#pragma unroll
            for (int i = 0; i < 20; ++i) L[i] += po[i];
            C[0] += po[20];
            C[1] += po[21];
            C[2] += po[22];
        }
    }
    finish_cell<RHO>(a, slot, g.cell, L, 20, C, true, false);
}

// Leaf targets (monopole kernels): one 512-lane workgroup per leaf sub-grid.
// The monopoles of the sub-grid and of its leaf neighbours within the stencil
// reach are staged once into an LDS window (zeros for refined or missing
// neighbours, whose cells the monopole kernel does not see).  Wave w takes
// child class k = w over the 64 coarse cells, so the partner offsets of its
// p2p loop -- and the four constants (-1/r, d/r^3) per offset -- are the same
// for every lane: each pair is one LDS read and four FMAs with scalar
// operands (monopole_kernel_templates.hpp:16-23, kokkos_kernel.hpp:255-284;
// the 1/dx and -1/dx^2 factors are applied once per cell).  Refined
// neighbours contribute through p2m with their full multipoles over the same
// stencil, far and near alike (monopole_kernel_templates.hpp:69-239): those
// sums come from p2m_kernel, launched before this kernel on the same stream.

// LDS layout of the 14^3 monopole window of the theta = 0.5 leaf kernel: 8
// parity planes (x & 1, y & 1, z & 1) of 7^3 cells each, row stride 12 and
// plane stride 80 doubles.  At each partner offset a wave's 64 lanes read
// cells 2C + const over the 4^3 coarse cells C, i.e. consecutive cells of
// one parity plane: with strides 1, 12 and 80 (mod 32: 1, 12, 16) the 32
// lanes of each half-wave hit 32 distinct 8-byte bank pairs.  In the plain
// x-slowest layout the lanes' stride of 2 cells put 4 of them on every bank
// pair in use.
constexpr int kWinSY = 12, kWinSX = 80, kWinPlane = 7 * kWinSX;
__device__ __forceinline__ int cube_win(int wx, int wy, int wz) {
    return (((wx & 1) << 2) | ((wy & 1) << 1) | (wz & 1)) * kWinPlane + (wx >> 1) * kWinSX + (wy >> 1) * kWinSY + (wz >> 1);
}

template <bool RHO, bool CUBE>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void leaf_kernel(LevelArgs a) {
    extern __shared__ double s_win[];
    __shared__ int s_nbr[27];
    __shared__ int s_kind[27];
    const int2 it = a.items[a.item0 + xcd_item(blockIdx.x, gridDim.x)];
    const int slot = it.x;
    if (threadIdx.x < 27) {
        const int n = a.nbr[(size_t)slot * 27 + threadIdx.x];
        s_nbr[threadIdx.x] = n;
        s_kind[threadIdx.x] = n >= 0 ? a.kind[n] : -1;
    }
    __syncthreads();
    const Tables* __restrict__ tab = a.tab;
    // window edge: 14 (halo 3) on the cube path, known at compile time, so the
    // gather's index arithmetic folds and its loads are all in flight at once
    const int h = CUBE ? 3 : tab->p2p_halo, W = kInx + 2 * h;
    auto gather = [&](int i) {
        const int x = i / (W * W) - h, y = (i / W) % W - h, z = i % W - h;
        const int raw = (x < 0 ? 0 : x < kInx ? 1 : 2) + 3 * (y < 0 ? 0 : y < kInx ? 1 : 2) +
            9 * (z < 0 ? 0 : z < kInx ? 1 : 2);
        const int ns = s_nbr[raw];
        double m = 0.0;
        if (ns >= 0 && s_kind[raw] == OCTOFMM_SLOT_LEAF)
            m = a.mon[(size_t)ns * kCells + ((x + kInx) & 7) * 64 + ((y + kInx) & 7) * 8 + ((z + kInx) & 7)];
        if constexpr (CUBE) {
            const int wx = x + h, wy = y + h, wz = z + h;
            s_win[cube_win(wx, wy, wz)] = m;
        } else {
            s_win[i] = m;
        }
    };
    if constexpr (CUBE) {
        constexpr int kWin = 14 * 14 * 14;
#pragma unroll
        for (int r = 0; r < (kWin + 511) / 512; ++r)
            if (threadIdx.x + r * 512 < kWin) gather(threadIdx.x + r * 512);
    } else {
        for (int i = threadIdx.x; i < W * W * W; i += 512) gather(i);
    }
    __syncthreads();
    const int k = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int l = threadIdx.x & 63;
    const int Cx = l >> 4, Cy = (l >> 2) & 3, Cz = l & 3;
    const int kx = (k >> 2) & 1, ky = (k >> 1) & 1, kz = k & 1;
    const int tx = 2 * Cx + kx, ty = 2 * Cy + ky, tz = 2 * Cz + kz;
    const int cell = tx * 64 + ty * 8 + tz;
    double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
    if constexpr (CUBE) {
        // the 6^3 block 2C - 2 .. 2C + 3 around the target's coarse cell C
        // (the target itself included: four of offset 0 is zero), with the
        // window offsets as LDS immediates and the constants of offset
        // d = p - k at fixed offsets from one class base:
        // box(d) = box(p) + 3 * 133 - (121 kx + 11 ky + kz) for p in [0, 6)^3
        // window cell 2C + 1 + p lies in parity plane (1 + p) & 1 at index
        // C + ((1 + p) >> 1): the lane's base plus a compile-time offset
        const double* wc = s_win + Cx * kWinSX + Cy * kWinSY + Cz;
        const double(*__restrict__ fb)[4] = tab->four + (3 * 133 - (121 * kx + 11 * ky + kz));
#pragma unroll
        for (int px = 0; px < 6; ++px)
#pragma unroll
            for (int py = 0; py < 6; ++py)
#pragma unroll
                for (int pz = 0; pz < 6; ++pz) {
                    const double m = wc[cube_win(1 + px, 1 + py, 1 + pz)];
                    const double* f = fb[px * 121 + py * 11 + pz];
                    p0 = fma(f[0], m, p0);
                    p1 = fma(f[1], m, p1);
                    p2 = fma(f[2], m, p2);
                    p3 = fma(f[3], m, p3);
                }
    } else {
        const double* wp = s_win + ((tx + h) * W + (ty + h)) * W + (tz + h) - 2048;
        const int32_t* __restrict__ list = tab->p2plist[k];
        const char* __restrict__ fc = reinterpret_cast<const char*>(tab->p2pconst);
        const int n = tab->np2p_k[k];
#pragma unroll 8
        for (int q = 0; q < n; ++q) {
            const int e = list[q];
            const double* f = reinterpret_cast<const double*>(fc + (e & 0xffff));
            const double m = wp[e >> 16];
            p0 = fma(f[0], m, p0);
            p1 = fma(f[1], m, p1);
            p2 = fma(f[2], m, p2);
            p3 = fma(f[3], m, p3);
        }
    }
    const double s1 = a.inv_dx * a.neg_inv_dx;
    double L[4] = {p0 * a.inv_dx, p1 * s1, p2 * s1, p3 * s1};
    double C[3] = {0.0, 0.0, 0.0};
    const bool p2m = (it.y & kItemP2M) != 0;
    if (p2m) {
        // refined neighbours: the p2m sums p2m_kernel formed for this leaf
        // (leaf items with p2m come first among the leaf items, so the item's
        // place among them indexes p2m_out)
        const double2* ps = reinterpret_cast<const double2*>(
            a.p2m_out + ((size_t)xcd_item(blockIdx.x, gridDim.x) * kCells + cell) * 8);
        const double2 v0 = ps[0], v1 = ps[1], v2 = ps[2], v3 = ps[3];
        L[0] += v0.x;
        L[1] += v0.y;
        L[2] += v1.x;
        L[3] += v1.y;
        C[0] = v2.x;
        C[1] = v2.y;
        C[2] = v3.x;
    }
    if constexpr (RHO) {
        if (a.l2l && a.g) {   // uniform: whole-tree RHO solve
            double Lf[20], Cf[3], g4[4];
            tree_final_row<RHO>(a, slot, cell, L, 4, C, p2m, Lf, Cf);
            leaf_force<RHO>(a, Lf, Cf, g4);
            leaf_rows_coalesced(a, slot, cell, Lf, Cf, g4, s_win);
            return;
        }
    }
    finish_cell<RHO>(a, slot, cell, L, 4, C, p2m, true);
}

// p2m on the staging of lane mapping 7: for leaf targets with refined
// neighbours, the pairs whose partner cell lies in a refined neighbour slot
// (far and near alike, monopole_kernel_templates.hpp:69-239) are the M2L pair
// set of the target's cells restricted to those slots.  So a workgroup of
// SPLIT waves takes one coarse block of one such leaf (lane = coarse cell x
// child, as in the M2L kernels) and walks only the coarse offsets that reach
// a refined neighbour from that block; each offset's 64 partner records are
// staged into the wave's LDS image by direct loads and read as 8-lane
// broadcasts.  The waves split the offsets and hand their sums to wave 0,
// which adds them in wave order and writes each cell's (L0..3, L_c, 0) to
// p2m_out[p2m item][cell].
template <bool RHO, int SPLIT>
__device__ void p2m_item(const LevelArgs& a, int slot, int ordinal, int w, const int* s_nbr, const int* s_kind,
    const int* s_live, int nlive, double2* __restrict__ s_stage, int sl) {
    double2* stage = s_stage + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) * kDirWave;
    const int l = threadIdx.x & 63, c = l >> 3;
    const auto* tab = as_const(a.tab);
    const int wu = __builtin_amdgcn_readfirstlane(w);
    const LaneGeom g = lane_geom(wu);
    const double* xc = a.com + ((size_t)slot * kCells + g.cell) * 3;
    const double X0 = xc[0], X1 = xc[1], X2 = xc[2];
    double L[4] = {0.0, 0.0, 0.0, 0.0}, C[3] = {0.0, 0.0, 0.0};
    const int bx = ((wu >> 2) & 1) * 2, by = ((wu >> 1) & 1) * 2, bz = (wu & 1) * 2;
    // partner coarse cell of group gi at offset (ax, ay, az): its records, and
    // bit gi of `absent` where it is not in a refined neighbour slot
    auto group_base = [&](int gi, int ax, int ay, int az, unsigned& absent) -> const char* {
        const int px = bx + ((gi >> 2) & 1) + ax, py = by + ((gi >> 1) & 1) + ay, pz = bz + (gi & 1) + az;
        const int raw = (px < 0 ? 0 : (px < 4 ? 1 : 2)) + 3 * (py < 0 ? 0 : (py < 4 ? 1 : 2)) +
            9 * (pz < 0 ? 0 : (pz < 4 ? 1 : 2));
        const int ns = s_nbr[raw];
        const bool refined = raw != 13 && ns >= 0 && s_kind[raw] == OCTOFMM_SLOT_NONLEAF;
        if (!refined) absent |= 1u << gi;
        const int cb = ((px & 3) * 16 + (py & 3) * 4 + (pz & 3)) * 8;
        return reinterpret_cast<const char*>(a.rec + ((size_t)(refined ? ns : 0) * kCells + cb) * kRecDoubles);
    };
    const unsigned loff = (unsigned)l * 16u;
    const char* nbase[8];
    unsigned absent_next = 0;
    auto bases = [&](int di) {
        const int ax = tab->delta[di][0], ay = tab->delta[di][1], az = tab->delta[di][2];
        unsigned absent = 0;
#pragma unroll
        for (int gi = 0; gi < 8; ++gi) nbase[gi] = group_base(gi, ax, ay, az, absent);
        asm volatile("" ::"s"(nbase[0]), "s"(nbase[1]), "s"(nbase[2]), "s"(nbase[3]), "s"(nbase[4]),
                     "s"(nbase[5]), "s"(nbase[6]), "s"(nbase[7]));
        absent_next = absent;
    };
    // the full-wave loads of every group first, the half-wave loads after
    // them in a branch of their own (as in m2l_item_direct)
    auto issue = [&]() {
#pragma unroll
        for (int gi = 0; gi < 8; ++gi)
            __builtin_amdgcn_global_load_lds(nbase[gi] + loff, (lds_ptr)(stage + gi * kDirGroup), 16, 0, 0);
        if (l < 32) {
#pragma unroll
            for (int gi = 0; gi < 8; ++gi)
                __builtin_amdgcn_global_load_lds(nbase[gi] + 1024 + loff, (lds_ptr)(stage + gi * kDirGroup + 64), 16, 0, 0);
        }
    };
    const double2* group = stage + c * kDirGroup;
    const int q0 = __builtin_amdgcn_readfirstlane(sl);
    if (q0 < nlive) {
        bases(s_live[q0]);
        issue();
    }
    for (int q = q0; q < nlive; q += SPLIT) {
        // lanes of step j: the classes in byte j of the offset's pair rows,
        // in the groups whose partner is in a refined slot (lane mask formed
        // in scalar registers, lane = group * 8 + class)
        const int di = s_live[q];
        const uint64_t rows = tab->allrows[di];
        const uint64_t gone = spread_groups(absent_next);
        if (q + SPLIT < nlive) bases(s_live[q + SPLIT]);
        __builtin_amdgcn_s_waitcnt(0x3F70);   // vmcnt(0): this wave's image is in LDS
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (__builtin_amdgcn_inverse_ballot_w64(step_lanes(rows, 0, gone, 0, j))) {
                Rec cur;
#pragma unroll
                for (int i = 0; i < 12; ++i) cur.v[i] = group[j * 12 + i];
                p2m_pair_rec<RHO>(cur, X0, X1, X2, L, C);
            }
        }
        lds_wave_sync();   // every read of the image done before it is refilled
        if (q + SPLIT < nlive) issue();
    }
    // waves 1..SPLIT-1 hand their sums to wave 0 through their own images
    double* part = reinterpret_cast<double*>(stage) + l * 7;
    if (q0 != 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) part[i] = L[i];
        part[4] = C[0];
        part[5] = C[1];
        part[6] = C[2];
    }
    __syncthreads();
    if (q0 != 0) return;
    for (int o = 1; o < SPLIT; ++o) {
        const double* po = reinterpret_cast<const double*>(s_stage + o * kDirWave) + l * 7;
#pragma unroll
        for (int i = 0; i < 4; ++i) L[i] += po[i];
        C[0] += po[4];
        C[1] += po[5];
        C[2] += po[6];
    }
    double2* out = reinterpret_cast<double2*>(a.p2m_out + ((size_t)ordinal * kCells + g.cell) * 8);
    out[0] = make_double2(L[0], L[1]);
    out[1] = make_double2(L[2], L[3]);
    out[2] = make_double2(C[0], C[1]);
    out[3] = make_double2(C[2], 0.0);
}

// One workgroup of SPLIT waves per (leaf item with p2m, coarse block); the
// items are a.items[a.item0 ...], in the order their sums are stored.
template <bool RHO, int SPLIT>
__global__ __launch_bounds__(64 * SPLIT) void p2m_kernel(LevelArgs a) {
    __shared__ int s_nbr[27];
    __shared__ int s_kind[27];
    __shared__ int s_live[kMaxDelta];
    __shared__ int s_nlive;
    __shared__ double2 s_stage[SPLIT * kDirWave];
    const int v = xcd_item(blockIdx.x, gridDim.x);
    const int ordinal = v >> 3, w = v & 7;
    const int2 it = a.items[a.item0 + ordinal];
    if (threadIdx.x < 27) {
        const int n = a.nbr[(size_t)it.x * 27 + threadIdx.x];
        s_nbr[threadIdx.x] = n;
        s_kind[threadIdx.x] = n >= 0 ? a.kind[n] : -1;
    }
    __syncthreads();
    // the coarse offsets with a pair whose partner coarse cell, seen from this
    // block, lies in a refined neighbour slot (wave 0, one offset per lane,
    // compacted in offset order by a ballot)
    if (threadIdx.x < 64) {
        const Tables* tab = a.tab;
        const int lane = threadIdx.x;
        const int bx = ((w >> 2) & 1) * 2, by = ((w >> 1) & 1) * 2, bz = (w & 1) * 2;
        int n = 0;
        for (int base = 0; base < tab->ndelta; base += 64) {
            const int di = base + lane;
            bool live = false;
            if (di < tab->ndelta && (tab->farbits[di] | tab->nearbits[di])) {
                const int ax = tab->delta[di][0], ay = tab->delta[di][1], az = tab->delta[di][2];
                for (int gi = 0; gi < 8; ++gi) {
                    const int px = bx + ((gi >> 2) & 1) + ax, py = by + ((gi >> 1) & 1) + ay, pz = bz + (gi & 1) + az;
                    const int raw = (px < 0 ? 0 : (px < 4 ? 1 : 2)) + 3 * (py < 0 ? 0 : (py < 4 ? 1 : 2)) +
                        9 * (pz < 0 ? 0 : (pz < 4 ? 1 : 2));
                    live = live || (raw != 13 && s_nbr[raw] >= 0 && s_kind[raw] == OCTOFMM_SLOT_NONLEAF);
                }
            }
            const uint64_t m = __ballot(live);
            if (live) s_live[n + __popcll(m & ((uint64_t(1) << lane) - 1))] = di;
            n += __popcll(m);
        }
        if (lane == 0) s_nlive = n;
    }
    __syncthreads();
    p2m_item<RHO, SPLIT>(a, it.x, ordinal, w, s_nbr, s_kind, s_live, s_nlive, s_stage, (int)(threadIdx.x >> 6));
}

template <bool RHO>
void launch_p2m(int nitems, hipStream_t st, const LevelArgs& la) {
    hipLaunchKernelGGL((p2m_kernel<RHO, 4>), dim3(nitems * 8), dim3(256), 0, st, la);
}

template <bool RHO>
void launch_leaf(int nitems, hipStream_t st, const LevelArgs& la, const Tables& tab) {
    const int W = kInx + 2 * tab.p2p_halo;
    const bool cube = tab.p2p_cube != 0;
    // window (parity planes on the cube path), and the staged L rows of half
    // a sub-grid (256 x 20, shifted by one chunk every 16 rows)
    const size_t win = cube ? (size_t)8 * kWinPlane : (size_t)W * W * W;
    const size_t lds = sizeof(double) * std::max<size_t>(win, (size_t)256 * 20 + 32);
    if (cube)
        hipLaunchKernelGGL((leaf_kernel<RHO, true>), dim3(nitems), dim3(512), lds, st, la);
    else
        hipLaunchKernelGGL((leaf_kernel<RHO, false>), dim3(nitems), dim3(512), lds, st, la);
}

// One target sub-grid is split over 8 / WAVES workgroups of WAVES waves.  The
// workgroup index is first spread over the XCDs (contiguous runs per XCD),
// so the parts of one sub-grid and its neighbours share an L2.
template <bool RHO, int WAVES, bool MIXED = false>
__global__ __launch_bounds__(64 * WAVES) void level_kernel(LevelArgs a) {
    constexpr int kParts = 8 / WAVES;
    __shared__ int s_nbr[27];
    __shared__ int s_kind[27];
    const int v = xcd_item(blockIdx.x, gridDim.x);
    const int item = v / kParts;
    const int w = (v % kParts) * WAVES + (threadIdx.x >> 6);
    const int2 it = a.items[a.item0 + item];
    if (threadIdx.x < 27) {
        const int n = a.nbr[(size_t)it.x * 27 + threadIdx.x];
        s_nbr[threadIdx.x] = n;
        s_kind[threadIdx.x] = n >= 0 ? a.kind[n] : -1;
    }
    __syncthreads();
    if (it.y & kItemLeaf)
        return;   // leaf items run in leaf_kernel
    m2l_item<RHO, MIXED>(a, it.x, it.y, s_nbr, s_kind, w);
}

// Mapping-3 kernel (direct global->LDS staging); one wave per 2x2x2 coarse block.
template <bool RHO, int WAVES>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(3))) void level_kernel_direct(LevelArgs a) {
    constexpr int kParts = 8 / WAVES;
    __shared__ int s_nbr[27];
    __shared__ int s_kind[27];
    __shared__ double2 s_stage[WAVES * kDirWave];
    const int v = xcd_item(blockIdx.x, gridDim.x);
    const int item = v / kParts;
    const int w = (v % kParts) * WAVES + (threadIdx.x >> 6);
    const int2 it = a.items[a.item0 + item];
    if (threadIdx.x < 27) {
        const int n = a.nbr[(size_t)it.x * 27 + threadIdx.x];
        s_nbr[threadIdx.x] = n;
        s_kind[threadIdx.x] = n >= 0 ? a.kind[n] : -1;
    }
    __syncthreads();
    if (it.y & kItemLeaf) return;
    m2l_item_direct<RHO>(a, it.x, it.y, s_nbr, s_kind, w, s_stage);
}

// Mapping-7 kernel: one workgroup of SPLIT waves per 2x2x2 coarse block, the
// coarse offsets split over the waves (mapping 3's staging in each).
template <bool RHO, int SPLIT, bool MIXED = false>
__global__ __launch_bounds__(64 * SPLIT) __attribute__((amdgpu_waves_per_eu(3))) void level_kernel_dsplit(LevelArgs a) {
    __shared__ int s_nbr[27];
    __shared__ int s_kind[27];
    __shared__ double2 s_stage[SPLIT * kDirWave];
    const int v = xcd_item(blockIdx.x, gridDim.x);
    const int item = v >> 3, w = v & 7;
    const int2 it = a.items[a.item0 + item];
    if (threadIdx.x < 27) {
        const int n = a.nbr[(size_t)it.x * 27 + threadIdx.x];
        s_nbr[threadIdx.x] = n;
        s_kind[threadIdx.x] = n >= 0 ? a.kind[n] : -1;
    }
    __syncthreads();
    if (it.y & kItemLeaf) return;   // uniform per workgroup
    m2l_item_direct<RHO, SPLIT, MIXED>(a, it.x, it.y, s_nbr, s_kind, w, s_stage, (int)(threadIdx.x >> 6));
}

// Persistent form of the mapping-7 kernel (variants 804, 802): a grid of
// the workgroups the chip holds at once, each taking work units (sub-grid,
// coarse block) from a counter until none are left.  Workgroups are handed
// to the 8 XCDs round-robin, so workgroup b runs on XCD b % 8 and takes its
// units from that XCD's counter, over the XCD's contiguous share of the units
// (the same split xcd_item makes): neighbouring sub-grids still share an L2.
// Every workgroup leaves the loop once its XCD's counter has passed the end
// of the share, so the grid drains.
__global__ void queue_reset(uint32_t* q) {
    if (threadIdx.x < 8) q[threadIdx.x] = 0u;
}

template <bool RHO, int SPLIT, bool MIXED = false>
__global__ __launch_bounds__(64 * SPLIT) __attribute__((amdgpu_waves_per_eu(3))) void level_kernel_dqueue(LevelArgs a) {
    __shared__ int s_nbr[27];
    __shared__ int s_kind[27];
    __shared__ int s_unit;
    __shared__ double2 s_stage[SPLIT * kDirWave];
    const int n = 8 * a.nitems;                 // units: 8 coarse blocks per item
    const int x = blockIdx.x & 7;
    const int q = n >> 3, r = n & 7;
    const int lo = x * q + (x < r ? x : r), hi = lo + q + (x < r ? 1 : 0);
    for (;;) {
        if (threadIdx.x == 0) s_unit = lo + (int)atomicAdd(a.queue + x, 1u);
        __syncthreads();
        const int v = s_unit;
        if (v >= hi) break;                     // uniform
        const int2 it = a.items[a.item0 + (v >> 3)];
        if (threadIdx.x < 27) {
            const int nb = a.nbr[(size_t)it.x * 27 + threadIdx.x];
            s_nbr[threadIdx.x] = nb;
            s_kind[threadIdx.x] = nb >= 0 ? a.kind[nb] : -1;
        }
        __syncthreads();
        if (!(it.y & kItemLeaf))   // uniform per workgroup
            m2l_item_direct<RHO, SPLIT, MIXED>(a, it.x, it.y, s_nbr, s_kind, v & 7, s_stage, (int)(threadIdx.x >> 6));
        __syncthreads();   // the unit's LDS (images, tables, s_unit) is free again
    }
}

template <bool RHO, int SPLIT, bool MIXED>
void launch_dqueue(int nitems, hipStream_t st, const LevelArgs& la) {
    // workgroups the device holds at once (occupancy of this instantiation
    // times CUs), found once per process
    static const int capacity = [] {
        int dev = 0, per = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, level_kernel_dqueue<RHO, SPLIT, MIXED>, 64 * SPLIT, 0) ==
                hipSuccess && per > 0)
            return per * prop.multiProcessorCount;
        return 256;
    }();
    const int units = 8 * nitems;
    int grid = std::min(units, capacity);
    grid = std::max(8, (grid + 7) & ~7);   // every XCD's counter has workgroups
    hipLaunchKernelGGL(queue_reset, dim3(1), dim3(64), 0, st, la.queue);
    hipLaunchKernelGGL((level_kernel_dqueue<RHO, SPLIT, MIXED>), dim3(grid), dim3(64 * SPLIT), 0, st, la);
}

template <bool RHO>
void launch_level(int variant, int nitems, hipStream_t st, const LevelArgs& la, bool mixed = false) {
    const int waves = variant % 100;
    const int parts = 8 / waves;
    const dim3 grid(nitems * parts), block(64 * waves);
    if (mixed) {
        switch (variant) {
        case 8: hipLaunchKernelGGL((level_kernel<RHO, 8, true>), grid, block, 0, st, la); break;
        case 4: hipLaunchKernelGGL((level_kernel<RHO, 4, true>), grid, block, 0, st, la); break;
        case 704:
            hipLaunchKernelGGL((level_kernel_dsplit<RHO, 4, true>), dim3(nitems * 8), dim3(256), 0, st, la);
            break;
        case 702:
            hipLaunchKernelGGL((level_kernel_dsplit<RHO, 2, true>), dim3(nitems * 8), dim3(128), 0, st, la);
            break;
        case 708:
            hipLaunchKernelGGL((level_kernel_dsplit<RHO, 8, true>), dim3(nitems * 8), dim3(512), 0, st, la);
            break;
        case 804: launch_dqueue<RHO, 4, true>(nitems, st, la); break;
        case 802: launch_dqueue<RHO, 2, true>(nitems, st, la); break;
        default: hipLaunchKernelGGL((level_kernel<RHO, 2, true>), grid, block, 0, st, la); break;
        }
        return;
    }
    switch (variant) {
    case 4: hipLaunchKernelGGL((level_kernel<RHO, 4>), grid, block, 0, st, la); break;
    case 2: hipLaunchKernelGGL((level_kernel<RHO, 2>), grid, block, 0, st, la); break;
    case 304: hipLaunchKernelGGL((level_kernel_direct<RHO, 4>), grid, block, 0, st, la); break;
    case 302: hipLaunchKernelGGL((level_kernel_direct<RHO, 2>), grid, block, 0, st, la); break;
    case 504:
        hipLaunchKernelGGL((level_kernel_split<RHO, 4>), dim3(nitems * 8), dim3(256), 0, st, la);
        break;
    case 704:
        hipLaunchKernelGGL((level_kernel_dsplit<RHO, 4>), dim3(nitems * 8), dim3(256), 0, st, la);
        break;
    case 702:
        hipLaunchKernelGGL((level_kernel_dsplit<RHO, 2>), dim3(nitems * 8), dim3(128), 0, st, la);
        break;
    case 708:
        hipLaunchKernelGGL((level_kernel_dsplit<RHO, 8>), dim3(nitems * 8), dim3(512), 0, st, la);
        break;
    case 804: launch_dqueue<RHO, 4, false>(nitems, st, la); break;
    case 802: launch_dqueue<RHO, 2, false>(nitems, st, la); break;
    default: hipLaunchKernelGGL((level_kernel<RHO, 8>), grid, block, 0, st, la); break;
    }
}

// Root sub-grid: every pair of its 512 cells except self and near offsets
// (kokkos_kernel.hpp:324-471).  Workgroup q takes partners [q*per, (q+1)*per);
// all lanes read the same partner, one target cell per lane.
template <bool RHO>
__global__ __launch_bounds__(512) void root_kernel(RootArgs a) {
    const int t = threadIdx.x;
    const int tx = t >> 6, ty = (t >> 3) & 7, tz = t & 7;
    const double* trec = a.rec + (size_t)record_of_cell(tx, ty, tz) * kRecDoubles;
    const double X0 = trec[kRecY], X1 = trec[kRecY + 1], X2 = trec[kRecY + 2];
    const double inv_mc0 = RHO ? 1.0 / trec[0] : 0.0;
    double mc[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) mc[i] = RHO ? trec[10 + i] : 0.0;
    double L[20], C[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 20; ++i) L[i] = 0.0;
    const int p0 = blockIdx.x * a.per_chunk;
    for (int p = p0; p < p0 + a.per_chunk; ++p) {
        const int px = p >> 6, py = (p >> 3) & 7, pz = p & 7;
        const int dx = px - tx, dy = py - ty, dz = pz - tz;
        const bool inbox = dx >= -5 && dx <= 5 && dy >= -5 && dy <= 5 && dz >= -5 && dz <= 5;
        if (inbox && a.tab->root_excluded[(dx + 5) * 121 + (dy + 5) * 11 + (dz + 5)]) continue;
        m2l_far<RHO>(a.rec + (size_t)record_of_cell(px, py, pz) * kRecDoubles, X0, X1, X2, inv_mc0, mc, L, C);
    }
    double* o = a.partial + ((size_t)blockIdx.x * kCells + t) * kRecDoubles;
#pragma unroll
    for (int i = 0; i < 20; ++i) o[i] = L[i];
    o[20] = C[0];
    o[21] = C[1];
    o[22] = C[2];
}

// One thread per output value (cell, component); the chunks are summed in
// chunk order, so the result does not depend on scheduling.  Sixteen chunks'
// loads are issued before their adds, so the thread waits on one round trip
// per sixteen chunks rather than one per chunk.
__global__ __launch_bounds__(256) void root_reduce(RootReduceArgs a) {
    const int gid = blockIdx.x * 256 + threadIdx.x;
    if (gid >= kCells * 23) return;
    const int t = gid / 23, c = gid - 23 * (gid / 23);
    const double* src = a.partial + (size_t)t * kRecDoubles + c;
    const size_t stride = (size_t)kCells * kRecDoubles;
    double s = 0.0;
    for (int q0 = 0; q0 < a.nchunks; q0 += 16) {
        double v[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = q0 + i < a.nchunks ? src[(size_t)(q0 + i) * stride] : 0.0;
#pragma unroll
        for (int i = 0; i < 16; ++i)
            if (q0 + i < a.nchunks) s += v[i];
    }
    if (c < 20)
        a.L[(size_t)t * 20 + c] = a.overwrite ? s : a.L[(size_t)t * 20 + c] + s;
    else if (a.rho)
        a.Lc[(size_t)t * 3 + (c - 20)] = s;
}

// Fill kernel used instead of hipMemsetAsync so the passes stay graph-safe.
// Dense FMA rate probe: 8 independent chains per lane, so the issue rate and
// not the FMA latency bounds it.  b < 1 keeps the chains bounded, away from
// overflow and denormals (they converge to c / (1 - b)).
template <typename T>
__global__ __launch_bounds__(256) void fma_rate_kernel(T* __restrict__ out, T b, T c, int iters) {
    T a[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = (T)(threadIdx.x & 63) + (T)i;
#pragma unroll 4
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 8; ++i) a[i] = fma(a[i], b, c);
    }
    T s = a[0];
#pragma unroll
    for (int i = 1; i < 8; ++i) s += a[i];
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ void zero_fill(double* p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Partial halos of the per-sub-grid entry points: a neighbour that hands over
// only some of its cells (grid::get_ilist_n_bnd order,
// multipole_interaction_interface.hpp:177,257, p2m_interaction_interface.hpp:142,239)
// becomes a full slot with zeros in every cell it did not hand over, which is
// what update_input's zero padding gives.  One block per such neighbour: it
// zeroes the slot, then writes each listed cell from its position in the
// handed-over arrays.  The host lists each cell once (its last occurrence),
// so no two threads write one cell.
struct ScatterArgs {
    const double* cM;       // [26][512][20] handed-over multipoles, by partial direction q
    const double* ccom;     // [26][512][3]
    const double* cmon;     // [26][512]
    const int2* cidx;       // [26][512] (cell, position)
    double* M;              // [27][512][20] slots
    double* com;            // [27][512][3]
    double* mon;            // [27][512]
    int32_t slot[26];       // slot of partial direction q
    int32_t count[26];      // listed cells of q
    int32_t multipole[26];  // 1: multipoles and centres of mass, 0: monopoles
};

__global__ __launch_bounds__(256) void scatter_partial(ScatterArgs a) {
    const int q = blockIdx.x, t = threadIdx.x;
    const int s = a.slot[q], n = a.count[q];
    const bool mp = a.multipole[q] != 0;
    double* M = a.M + (size_t)s * kCells * 20;
    double* C = a.com + (size_t)s * kCells * 3;
    double* m = a.mon + (size_t)s * kCells;
    if (mp) {
        for (int i = t; i < kCells * 20; i += 256) M[i] = 0.0;
        for (int i = t; i < kCells * 3; i += 256) C[i] = 0.0;
    } else {
        for (int i = t; i < kCells; i += 256) m[i] = 0.0;
    }
    __syncthreads();
    const int2* idx = a.cidx + (size_t)q * kCells;
    for (int i = t; i < n; i += 256) {
        const int2 e = idx[i];
        if (mp) {
            const double* src = a.cM + ((size_t)q * kCells + e.y) * 20;
#pragma unroll
            for (int c = 0; c < 20; ++c) M[e.x * 20 + c] = src[c];
            const double* sc = a.ccom + ((size_t)q * kCells + e.y) * 3;
            C[e.x * 3] = sc[0];
            C[e.x * 3 + 1] = sc[1];
            C[e.x * 3 + 2] = sc[2];
        } else {
            m[e.x] = a.cmon[(size_t)q * kCells + e.y];
        }
    }
}

// Partial halo slots of a level plan (octofmm_partial_slot): the same scatter
// as scatter_partial, one block per slot, from descriptors in device memory.
struct PartialSlot {
    const double* M;        // [count][20] handed-over multipoles (non-leaf slot)
    const double* com;      // [count][3]
    const double* mon;      // [count]     (leaf slot)
    const int2* idx;        // [n] (cell, position in the handed-over arrays), each cell once
    int32_t slot;
    int32_t n;
    int32_t multipole;      // 1: multipoles and centres of mass, 0: monopoles
    int32_t pad;
};

__global__ __launch_bounds__(256) void scatter_slots(const PartialSlot* __restrict__ d, double* M, double* com,
    double* mon) {
    const PartialSlot e = d[blockIdx.x];
    const int t = threadIdx.x;
    double* Ms = M + (size_t)e.slot * kCells * 20;
    double* Cs = com + (size_t)e.slot * kCells * 3;
    double* ms = mon + (size_t)e.slot * kCells;
    if (e.multipole) {
        for (int i = t; i < kCells * 20; i += 256) Ms[i] = 0.0;
        for (int i = t; i < kCells * 3; i += 256) Cs[i] = 0.0;
    } else {
        for (int i = t; i < kCells; i += 256) ms[i] = 0.0;
    }
    __syncthreads();
    for (int i = t; i < e.n; i += 256) {
        const int2 q = e.idx[i];
        if (e.multipole) {
            const double* src = e.M + (size_t)q.y * 20;
#pragma unroll
            for (int c = 0; c < 20; ++c) Ms[q.x * 20 + c] = src[c];
            const double* sc = e.com + (size_t)q.y * 3;
            Cs[q.x * 3] = sc[0];
            Cs[q.x * 3 + 1] = sc[1];
            Cs[q.x * 3 + 2] = sc[2];
        } else {
            ms[q.x] = e.mon[q.y];
        }
    }
}

// geometric centre xbase + i*dx, rounded as written (grid_fmm.cpp:1669-1671)
__device__ __forceinline__ double cell_centre(double xb, int i, double dx) {
#pragma clang fp contract(off)
    return xb + (double)i * dx;
}

// Upward pass of one level (grid::compute_multipoles, src/grid_fmm.cpp:1639-1803):
// one lane per child cell, 8 consecutive lanes per coarse cell (512 per
// node).  For RHO the leaf nodes' cell centres are written first; the 8 lanes
// form the coarse cell's mass and mass-weighted centre by a fixed butterfly,
// shift their child's multipole to it, and sum the shifted multipoles in
// child order through LDS, each lane summing and storing 2 or 3 components of
// the parent's cell.
__global__ __launch_bounds__(256) void m2m_kernel(M2MArgs a) {
    __shared__ double s_m2m[256 * 11];
    const int gid = blockIdx.x * 256 + threadIdx.x;
    if (gid >= a.nnodes * kCells) return;     // whole 8-lane groups exit together
    const int n = gid >> 9, q = (gid >> 3) & 63, ci = gid & 7;
    const int ip = q >> 4, jp = (q >> 2) & 3, kp = q & 3;
    const bool leaf = a.kind[n] == OCTOFMM_SLOT_LEAF;   // uniform per node, so per wave
    const int u = (2 * ip + (ci & 1)) * 64 + (2 * jp + ((ci >> 1) & 1)) * 8 + (2 * kp + ((ci >> 2) & 1));
    const double* xb = a.xbase + 3 * n;
    if (leaf && a.rho) {
        // the leaf cells' centres, consecutive lanes on consecutive cells so
        // the stores coalesce
        const int v = gid & (kCells - 1);
        double* Xv = a.com + ((size_t)n * kCells + v) * 3;
        Xv[0] = cell_centre(xb[0], v >> 6, a.dx);
        Xv[1] = cell_centre(xb[1], (v >> 3) & 7, a.dx);
        Xv[2] = cell_centre(xb[2], v & 7, a.dx);
    }
    const int pn = a.parent[n];
    if (pn < 0) return;                        // uniform per node
    const int pc = parent_cell(a.octant[n], ip, jp, kp);
    double* Yp = a.pcom + ((size_t)pn * kCells + pc) * 3;
    double A[20];
    double x0, x1, x2;
    if (leaf) {
        A[0] = a.mon[(size_t)n * kCells + u];
#pragma unroll
        for (int j = 1; j < 20; ++j) A[j] = 0.0;
    } else {
        const double2* src = reinterpret_cast<const double2*>(a.M + ((size_t)n * kCells + u) * 20);
#pragma unroll
        for (int j = 0; j < 10; ++j) {
            const double2 v = src[j];
            A[2 * j] = v.x;
            A[2 * j + 1] = v.y;
        }
    }
    if (leaf && a.rho) {
        // the centre stored above for cell u, formed again rather than read back
        x0 = cell_centre(xb[0], u >> 6, a.dx);
        x1 = cell_centre(xb[1], (u >> 3) & 7, a.dx);
        x2 = cell_centre(xb[2], u & 7, a.dx);
    } else {
        const double* X = a.com + ((size_t)n * kCells + u) * 3;
        x0 = X[0];
        x1 = X[1];
        x2 = X[2];
    }
    double Y[3];
    if (a.rho) {
        const double mtot = sum8(A[0]);
        Y[0] = sum8(x0 * A[0]) / mtot;
        Y[1] = sum8(x1 * A[0]) / mtot;
        Y[2] = sum8(x2 * A[0]) / mtot;
        if (ci == 0) {
            Yp[0] = Y[0];
            Yp[1] = Y[1];
            Yp[2] = Y[2];
        }
    } else {
        Y[0] = Yp[0];
        Y[1] = Yp[1];
        Y[2] = Yp[2];
    }
    const double dX[3] = {x0 - Y[0], x1 - Y[1], x2 - Y[2]};
    double B[20];
    if (leaf)
        monopole_shift(A[0], dX, B);
    else
        multipole_shift(A, dX, B);
    // the 8 children's shifted multipoles meet in the wave's LDS rows, 10
    // components at a time (11 doubles apart, so the 8 lanes reading one
    // component hit distinct banks; two passes of 10 halve the LDS a
    // workgroup holds, so 5 waves per SIMD fit instead of 3); lane ci of the
    // group sums components ci (and ci + 8) of each half over the children
    // in order and stores them, so the group writes the parent's 160-byte
    // row as consecutive doubles
    double* row = s_m2m + threadIdx.x * 11;
    const double* grp = s_m2m + (threadIdx.x & ~7) * 11;
    double* Mo = a.pM + ((size_t)pn * kCells + pc) * 20;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if (h == 1) lds_wave_sync();   // the first half's rows are read before they are overwritten
#pragma unroll
        for (int j = 0; j < 10; ++j) row[j] = B[10 * h + j];
        lds_wave_sync();
#pragma unroll
        for (int j = ci; j < 10; j += 8) {
            double v = grp[j];
#pragma unroll
            for (int c = 1; c < 8; ++c) v += grp[c * 11 + j];
            Mo[10 * h + j] = v;
        }
    }
}


}   // namespace

// ===========================================================================
// Host side
// ===========================================================================
namespace {

constexpr int kRootChunks = 64;
// record buffers end with a tail the direct-to-LDS staging may read into
// (one 16-byte pad chunk past the last record)
constexpr size_t kRecTailBytes = 4096;
// layout of the per-sub-grid workspace metadata (int32 words); the int2 work
// item comes first so it stays 8-byte aligned
constexpr int kMetaItems = 0;
constexpr int kMetaNbr = 2;
constexpr int kMetaKind = kMetaNbr + 27 * 27;
constexpr int kMetaFoff = kMetaKind + 27;
constexpr int kMetaSize = kMetaFoff + 81;

// Device workspace of the per-sub-grid entry points: a 27-slot level (slot =
// raw direction index, 13 = the target sub-grid), the neighbour data as the
// caller hands it over, and the root kernel's chunk sums.  Every region starts
// on a 256-byte boundary of one block.
struct SubgridWorkspace {
    double* M = nullptr;        // [27][512][20]
    double* com = nullptr;      // [27][512][3]
    double* mon = nullptr;      // [27][512]
    double* rec = nullptr;      // [27][512][24] + tail
    double* L = nullptr;        // [27][512][20]: slot 13 holds the target's L
    double* Lc = nullptr;       // [27][512][3]
    double* partial = nullptr;  // [kRootChunks][512][24]
    double* cM = nullptr;       // [26][512][20] partial halos as handed over, by partial direction
    double* ccom = nullptr;     // [26][512][3]
    double* cmon = nullptr;     // [26][512]
    int2* cidx = nullptr;       // [26][512] (cell, position in the handed-over arrays)
    int32_t* meta = nullptr;    // kMetaSize
    double* fx = nullptr;       // [27][3]
    double* p2m = nullptr;      // [512][8] p2m sums of a leaf target
};

constexpr size_t ws_align(size_t n) { return (n + 255) & ~size_t(255); }
constexpr size_t kWsM = 0;
constexpr size_t kWsCom = kWsM + ws_align(sizeof(double) * 27 * kCells * 20);
constexpr size_t kWsMon = kWsCom + ws_align(sizeof(double) * 27 * kCells * 3);
constexpr size_t kWsRec = kWsMon + ws_align(sizeof(double) * 27 * kCells);
constexpr size_t kWsL = kWsRec + ws_align(sizeof(double) * 27 * kCells * kRecDoubles + kRecTailBytes);
constexpr size_t kWsLc = kWsL + ws_align(sizeof(double) * 27 * kCells * 20);
constexpr size_t kWsPartial = kWsLc + ws_align(sizeof(double) * 27 * kCells * 3);
constexpr size_t kWsCM = kWsPartial + ws_align(sizeof(double) * kRootChunks * kCells * kRecDoubles);
constexpr size_t kWsCCom = kWsCM + ws_align(sizeof(double) * 26 * kCells * 20);
constexpr size_t kWsCMon = kWsCCom + ws_align(sizeof(double) * 26 * kCells * 3);
constexpr size_t kWsCIdx = kWsCMon + ws_align(sizeof(double) * 26 * kCells);
constexpr size_t kWsMeta = kWsCIdx + ws_align(sizeof(int2) * 26 * kCells);
constexpr size_t kWsFx = kWsMeta + ws_align(sizeof(int32_t) * kMetaSize);
constexpr size_t kWsP2m = kWsFx + ws_align(sizeof(double) * 81);
constexpr size_t kWsBytes = kWsP2m + ws_align(sizeof(double) * kCells * 8);

SubgridWorkspace carve_workspace(char* base) {
    SubgridWorkspace w;
    w.M = reinterpret_cast<double*>(base + kWsM);
    w.com = reinterpret_cast<double*>(base + kWsCom);
    w.mon = reinterpret_cast<double*>(base + kWsMon);
    w.rec = reinterpret_cast<double*>(base + kWsRec);
    w.L = reinterpret_cast<double*>(base + kWsL);
    w.Lc = reinterpret_cast<double*>(base + kWsLc);
    w.partial = reinterpret_cast<double*>(base + kWsPartial);
    w.cM = reinterpret_cast<double*>(base + kWsCM);
    w.ccom = reinterpret_cast<double*>(base + kWsCCom);
    w.cmon = reinterpret_cast<double*>(base + kWsCMon);
    w.cidx = reinterpret_cast<int2*>(base + kWsCIdx);
    w.meta = reinterpret_cast<int32_t*>(base + kWsMeta);
    w.fx = reinterpret_cast<double*>(base + kWsFx);
    w.p2m = reinterpret_cast<double*>(base + kWsP2m);
    return w;
}

}   // namespace

// Workspace of octofmm_compute_multipole_interactions_noalloc: one per-sub-grid
// workspace (kWsBytes, 16.0 MiB) in the code object's device memory, so that
// entry point never allocates.  Each GPU the library runs on holds its own
// copy until the process ends.  In host code `scratch` names a host-side
// stand-in, not a device address: host code uses only the address
// hipGetSymbolAddress returns for the current device (scratch_address).
static __device__ __align__(256) char scratch[kWsBytes];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_setup_mu;
char* g_scratch_ptr[kScratchDevices] = {};
// calls of the no-allocation entry point on one device share its scratch:
// they are serialised per device within the process
std::mutex g_scratch_call_mu[kScratchDevices];

// The device address of `scratch` on `device`, looked up once per device with
// that device current (the caller has made it current); null on failure.
char* scratch_address(int device) {
    if (device < 0 || device >= kScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_setup_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_scratch_ptr[device] = static_cast<char*>(p);
    }
    return g_scratch_ptr[device];
}
}   // namespace

struct octofmm_context {
    int device = 0;
    double theta = 0.5;
    int flags = 0;                 // OCTOFMM_CHECK_FINITE
    Tables tab;
    Tables* d_tab = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    // per-sub-grid entry points: the device workspace (one allocation on the
    // first call) and the host-side metadata they upload, built at creation
    SubgridWorkspace ws;
    char* d_ws = nullptr;
    char* d_scratch = nullptr;     // this device's `scratch` (set at creation)
    int32_t* h_meta = nullptr;     // items[2] nbr[27*27] kind[27] foff[27*3]
    double* h_fx = nullptr;        // [27][3]
    int2* h_idx = nullptr;         // [26][512] (cell, position) of partial halos
};

struct octofmm_plan {
    octofmm_context* ctx = nullptr;
    int32_t nslots = 0, nitems = 0, type = 0, precision = 0, flags = 0;
    int32_t n_m2l = 0;             // items [0, n_m2l) are non-leaf, the rest leaf targets
    int32_t n_stage = 0;           // slots whose partner records some item reads
    int32_t* d_stage_list = nullptr;   // those slots, or null when all are read
    double dx = 0.0;
    const double* d_M = nullptr;
    const double* d_com = nullptr;
    const double* d_mon = nullptr;
    double* d_L = nullptr;
    double* d_Lc = nullptr;
    double* d_rec = nullptr;
    int32_t* d_nbr = nullptr;
    int32_t* d_kind = nullptr;
    int32_t* d_foff = nullptr;
    double* d_fxbase = nullptr;
    int2* d_items = nullptr;
    octofmm_plan_stats stats;
    int variant = 8;           // level-kernel variant (see kVariants)
    int32_t mixed_far_d2 = 4;
    std::vector<int2> h_items;     // the work items (host copy, for the tuning key)
    uint64_t* d_mixbits = nullptr;
    // mixed precision, lane mappings 3 and 7: the mixed-precision records and
    // the far-schedule entries evaluated in FP32 as a whole (Tables::sched order)
    double* d_recm = nullptr;
    uint8_t* d_mixoff = nullptr;
    uint32_t* d_queue = nullptr;   // per-XCD work counters of the persistent variants
    int32_t n_p2m = 0;             // leaf items with p2m: items [n_m2l, n_m2l + n_p2m)
    double* d_p2m_out = nullptr;   // their p2m sums [n_p2m][512][8]
    int64_t mixed_far_pairs_by_offset = 0;   // FP32 far pairs under that rule
    int64_t mixed_far_pairs_by_pair = 0;     // and under the per-pair rule of mapping 0
    bool root = false;             // root sub-grid: root stencil over slot 0
    // set by a whole-tree plan: write final expansions (fused downward pass)
    bool l2l = false;
    const int32_t* t_parent = nullptr;
    const int32_t* t_octant = nullptr;
    const double* t_pL = nullptr;
    const double* t_pLc = nullptr;
    const double* t_pcom = nullptr;
    double* t_g = nullptr;
    double t_G = 1.0;
    int32_t t_correct_am = 1;
    double* d_partial = nullptr;   // root: per-chunk partial expansions
    int32_t leaf_rows = 1;         // whole-tree plans: 0 = leaf cells get g (and L_c) only
    // partial halo slots: rebuilt in the caller's slot rows before staging
    int32_t npartial = 0;
    PartialSlot* d_pslots = nullptr;
    int2* d_pidx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // a level with both kinds of items runs its leaf kernel on this stream,
    // forked after staging and joined before the pass ends
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

namespace {

octofmm_status fail(octofmm_context* c, octofmm_status s, const char* fmt, ...) {
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        std::vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        c->err = buf;
    }
    return s;
}

#define HIPC(ctx, call)                                                                         \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail((ctx), e_ == hipErrorOutOfMemory ? OCTOFMM_ERR_ALLOCATION : OCTOFMM_ERR_HIP, \
                "%s failed: %s", #call, hipGetErrorString(e_));                                 \
    } while (0)

int raw_of_flat(int f) { return f < 13 ? f : f + 1; }

// The per-sub-grid workspace of a context, allocated on its first call.
octofmm_status ensure_workspace(octofmm_context* c) {
    if (c->d_ws) return OCTOFMM_OK;
    HIPC(c, hipMalloc((void**)&c->d_ws, kWsBytes));
    c->ws = carve_workspace(c->d_ws);
    return OCTOFMM_OK;
}

bool check_neighbor(const octofmm_neighbor& nb, bool want_multipole) {
    if (nb.count < 0 || nb.count > kCells) return false;
    if (nb.count != kCells && nb.count > 0 && !nb.cells) return false;
    if (nb.cells)
        for (int i = 0; i < nb.count; ++i)
            if (nb.cells[i] < 0 || nb.cells[i] >= kCells) return false;
    (void)want_multipole;
    return true;
}

bool all_finite(const double* p, size_t n) {
    if (!p) return true;
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

// OCTOFMM_CHECK_FINITE: the sub-grid's own arrays and every neighbour's.
octofmm_status check_finite_inputs(octofmm_context* c, const octofmm_subgrid* in, bool leaf_target) {
    if (!(c->flags & OCTOFMM_CHECK_FINITE)) return OCTOFMM_OK;
    if (!leaf_target && !all_finite(in->M, (size_t)kCells * 20))
        return fail(c, OCTOFMM_ERR_NONFINITE, "non-finite multipole in the sub-grid");
    if (!all_finite(in->com, (size_t)kCells * 3))
        return fail(c, OCTOFMM_ERR_NONFINITE, "non-finite centre of mass in the sub-grid");
    if (leaf_target && !all_finite(in->mon, kCells))
        return fail(c, OCTOFMM_ERR_NONFINITE, "non-finite monopole in the sub-grid");
    if (!std::isfinite(in->dx) || !std::isfinite(in->xbase[0]) || !std::isfinite(in->xbase[1]) ||
        !std::isfinite(in->xbase[2]))
        return fail(c, OCTOFMM_ERR_NONFINITE, "non-finite dx or xbase");
    for (int f = 0; f < 26; ++f) {
        const octofmm_neighbor& nb = in->neighbors[f];
        if (nb.kind == OCTOFMM_NEIGHBOR_EMPTY) continue;
        const size_t n = (size_t)nb.count;
        if (!all_finite(nb.M, n * 20) || !all_finite(nb.com, n * 3) || !all_finite(nb.mon, n))
            return fail(c, OCTOFMM_ERR_NONFINITE, "non-finite data in neighbour %d", f);
    }
    return OCTOFMM_OK;
}

// Loads one sub-grid descriptor into a workspace, on the context's stream:
// the target's own arrays and every full neighbour are copied from the
// caller's memory straight into their slots; a partial halo is copied as
// handed over into its compact region and scatter_partial writes it into its
// zeroed slot; then the 27-slot metadata.  Slots the metadata does not list
// are never read, and staging reads only M and com of non-leaf slots and mon
// of leaf slots, so nothing else needs clearing between calls.  The host
// buffers read by the copies (c->h_meta, h_fx, h_idx) stay untouched until
// the caller synchronises the stream.
octofmm_status load_subgrid(octofmm_context* c, const SubgridWorkspace& w, const octofmm_subgrid* in,
    bool leaf_target, bool root, int item_flags) {
    hipStream_t st = c->stream;
    int32_t* meta = c->h_meta;
    int32_t* items = meta + kMetaItems;
    int32_t* nbr = meta + kMetaNbr;  // only row 13 is used
    int32_t* kind = meta + kMetaKind;
    int32_t* foff = meta + kMetaFoff;
    items[0] = 13;
    items[1] = item_flags;
    for (int i = 0; i < 27 * 27; ++i) nbr[i] = -1;
    for (int sl = 0; sl < 27; ++sl) {
        kind[sl] = -1;
        const int d[3] = {sl % 3 - 1, (sl / 3) % 3 - 1, sl / 9 - 1};
        for (int k = 0; k < 3; ++k) {
            foff[3 * sl + k] = kInx + kInx * d[k];
            c->h_fx[3 * sl + k] = in->xbase[k];
        }
    }
    const size_t nM = (size_t)kCells * 20, nC = (size_t)kCells * 3;
    if (leaf_target) {
        HIPC(c, hipMemcpyAsync(w.mon + 13 * kCells, in->mon, sizeof(double) * kCells, hipMemcpyHostToDevice, st));
        HIPC(c, hipMemcpyAsync(w.com + 13 * nC, in->com, sizeof(double) * nC, hipMemcpyHostToDevice, st));
        kind[13] = OCTOFMM_SLOT_LEAF;
    } else {
        HIPC(c, hipMemcpyAsync(w.M + 13 * nM, in->M, sizeof(double) * nM, hipMemcpyHostToDevice, st));
        HIPC(c, hipMemcpyAsync(w.com + 13 * nC, in->com, sizeof(double) * nC, hipMemcpyHostToDevice, st));
        kind[13] = OCTOFMM_SLOT_NONLEAF;
    }
    nbr[13 * 27 + 13] = 13;
    ScatterArgs sa;
    std::memset(&sa, 0, sizeof(sa));
    int np = 0;
    for (int f = 0; f < (root ? 0 : 26); ++f) {
        const octofmm_neighbor& nb = in->neighbors[f];
        const int sl = raw_of_flat(f);
        const bool multipole = nb.kind == OCTOFMM_NEIGHBOR_MULTIPOLE && nb.M && nb.com;
        const bool monopole = nb.kind == OCTOFMM_NEIGHBOR_MONOPOLE && nb.mon;
        if (!multipole && !monopole) continue;   // empty, or refined without data: the reference skips it
        if (nb.count == kCells && !nb.cells) {
            if (multipole) {
                HIPC(c, hipMemcpyAsync(w.M + sl * nM, nb.M, sizeof(double) * nM, hipMemcpyHostToDevice, st));
                HIPC(c, hipMemcpyAsync(w.com + sl * nC, nb.com, sizeof(double) * nC, hipMemcpyHostToDevice, st));
            } else {
                HIPC(c, hipMemcpyAsync(w.mon + sl * kCells, nb.mon, sizeof(double) * kCells, hipMemcpyHostToDevice, st));
            }
        } else {
            // each listed cell once, from its last occurrence (later entries win,
            // as in update_input's loop over the list)
            int16_t last[kCells];
            for (int i = 0; i < kCells; ++i) last[i] = -1;
            for (int i = 0; i < nb.count; ++i) last[nb.cells[i]] = (int16_t)i;
            int2* idx = c->h_idx + (size_t)np * kCells;
            int n = 0;
            for (int cell = 0; cell < kCells; ++cell)
                if (last[cell] >= 0) idx[n++] = make_int2(cell, last[cell]);
            if (nb.count > 0) {
                if (multipole) {
                    HIPC(c, hipMemcpyAsync(w.cM + np * nM, nb.M, sizeof(double) * 20 * nb.count,
                        hipMemcpyHostToDevice, st));
                    HIPC(c, hipMemcpyAsync(w.ccom + np * nC, nb.com, sizeof(double) * 3 * nb.count,
                        hipMemcpyHostToDevice, st));
                } else {
                    HIPC(c, hipMemcpyAsync(w.cmon + (size_t)np * kCells, nb.mon, sizeof(double) * nb.count,
                        hipMemcpyHostToDevice, st));
                }
            }
            sa.slot[np] = sl;
            sa.count[np] = n;
            sa.multipole[np] = multipole ? 1 : 0;
            ++np;
        }
        kind[sl] = multipole ? OCTOFMM_SLOT_NONLEAF : OCTOFMM_SLOT_LEAF;
        nbr[13 * 27 + sl] = sl;
    }
    HIPC(c, hipMemcpyAsync(w.meta, meta, sizeof(int32_t) * kMetaSize, hipMemcpyHostToDevice, st));
    HIPC(c, hipMemcpyAsync(w.fx, c->h_fx, sizeof(double) * 81, hipMemcpyHostToDevice, st));
    if (np > 0) {
        HIPC(c, hipMemcpyAsync(w.cidx, c->h_idx, sizeof(int2) * kCells * np, hipMemcpyHostToDevice, st));
        sa.cM = w.cM;
        sa.ccom = w.ccom;
        sa.cmon = w.cmon;
        sa.cidx = w.cidx;
        sa.M = w.M;
        sa.com = w.com;
        sa.mon = w.mon;
        hipLaunchKernelGGL(scatter_partial, dim3(np), dim3(256), 0, st, sa);
        HIPC(c, hipGetLastError());
    }
    return OCTOFMM_OK;
}

}   // namespace

// Runs one loaded sub-grid through a workspace: the caller's L is copied in
// and the kernels add the sub-grid's interactions to it (the same single
// addition per component a host-side L += result makes), then L comes back,
// and L_c (when the call writes it) is copied out; synchronous.
static octofmm_status launch_subgrid(octofmm_context* c, const SubgridWorkspace& w, double dx, int type,
    int item_flags, bool root, double* L, double* L_c) {
    hipStream_t st = c->stream;
    const int rho = type == OCTOFMM_RHO;
    StageArgs sa;
    sa.M = w.M;
    sa.com = w.com;
    sa.mon = w.mon;
    sa.kind = w.meta + kMetaKind;
    sa.fxbase = w.fx;
    sa.foff = w.meta + kMetaFoff;
    sa.rec = w.rec;
    sa.recm = nullptr;
    sa.list = nullptr;
    sa.dx = dx;
    sa.nslots = 27;
    hipLaunchKernelGGL(stage_records, dim3(27 * 4), dim3(256), 0, st, sa);
    HIPC(c, hipGetLastError());
    const size_t nL = (size_t)kCells * 20, nC = (size_t)kCells * 3;
    double* L13 = w.L + 13 * nL;
    double* C13 = w.Lc + 13 * nC;
    HIPC(c, hipMemcpyAsync(L13, L, sizeof(double) * nL, hipMemcpyHostToDevice, st));
    if (root) {
        RootArgs ra;
        ra.rec = w.rec + (size_t)13 * kCells * kRecDoubles;
        ra.tab = c->d_tab;
        ra.partial = w.partial;
        ra.per_chunk = kCells / kRootChunks;
        ra.nchunks = kRootChunks;
        if (rho)
            hipLaunchKernelGGL(root_kernel<true>, dim3(kRootChunks), dim3(512), 0, st, ra);
        else
            hipLaunchKernelGGL(root_kernel<false>, dim3(kRootChunks), dim3(512), 0, st, ra);
        HIPC(c, hipGetLastError());
        RootReduceArgs rr;
        rr.partial = w.partial;
        rr.L = L13;
        rr.Lc = C13;
        rr.nchunks = kRootChunks;
        rr.rho = rho;
        rr.overwrite = 0;
        hipLaunchKernelGGL(root_reduce, dim3((kCells * 23 + 255) / 256), dim3(256), 0, st, rr);
        HIPC(c, hipGetLastError());
    } else {
        LevelArgs la;
        la.rec = w.rec;
        la.com = w.com;
        la.L = w.L;
        la.Lc = w.Lc;
        la.nbr = w.meta + kMetaNbr;
        la.kind = w.meta + kMetaKind;
        la.items = reinterpret_cast<const int2*>(w.meta + kMetaItems);
        la.tab = c->d_tab;
        la.mixbits = nullptr;
        la.recm = nullptr;
        la.mixoff = nullptr;
        la.queue = nullptr;
        la.mon = w.mon;
        la.p2m_out = w.p2m;
        la.parent = nullptr;
        la.octant = nullptr;
        la.pL = la.pLc = la.pcom = nullptr;
        la.g = nullptr;
        la.G = 1.0;
        la.l2l = 0;
        la.correct_am = 0;
        la.leaf_rows = 1;
        la.nitems = 1;
        la.item0 = 0;
        la.inv_dx = 1.0 / dx;
        la.neg_inv_dx = -1.0 / dx;
        if (item_flags & kItemLeaf) {
            if (item_flags & kItemP2M) {
                if (rho)
                    launch_p2m<true>(1, st, la);
                else
                    launch_p2m<false>(1, st, la);
                HIPC(c, hipGetLastError());
            }
            if (rho)
                launch_leaf<true>(1, st, la, c->tab);
            else
                launch_leaf<false>(1, st, la, c->tab);
        } else if (rho) {
            launch_level<true>(8, 1, st, la);
        } else {
            launch_level<false>(8, 1, st, la);
        }
        HIPC(c, hipGetLastError());
    }
    HIPC(c, hipMemcpyAsync(L, L13, sizeof(double) * nL, hipMemcpyDeviceToHost, st));
    if (L_c) HIPC(c, hipMemcpyAsync(L_c, C13, sizeof(double) * nC, hipMemcpyDeviceToHost, st));
    HIPC(c, hipStreamSynchronize(st));
    return OCTOFMM_OK;
}

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

int32_t octofmm_abi_version(void) { return OCTOFMM_ABI_VERSION; }

const char* octofmm_status_string(octofmm_status s) {
    switch (s) {
    case OCTOFMM_OK: return "ok";
    case OCTOFMM_ERR_INVALID_ARGUMENT: return "invalid argument";
    case OCTOFMM_ERR_THETA: return "invalid theta / stencil";
    case OCTOFMM_ERR_NO_DEVICE: return "no HIP device";
    case OCTOFMM_ERR_HIP: return "HIP runtime error";
    case OCTOFMM_ERR_ALLOCATION: return "allocation failed";
    case OCTOFMM_ERR_SIZE: return "size not supported";
    case OCTOFMM_ERR_NONFINITE: return "non-finite input";
    }
    return "unknown status";
}

const char* octofmm_last_error(const octofmm_context* ctx) { return ctx ? ctx->err.c_str() : ""; }

octofmm_status octofmm_create(const octofmm_config* cfg, octofmm_context** out) {
    if (!cfg || !out) return OCTOFMM_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return OCTOFMM_ERR_NO_DEVICE;
    if (cfg->device < 0 || cfg->device >= n) return OCTOFMM_ERR_NO_DEVICE;
    octofmm_context* c = new octofmm_context();
    c->device = cfg->device;
    c->theta = cfg->theta;
    c->flags = cfg->flags;
    c->err.reserve(600);   // error texts (at most 511 characters) are stored without allocating
    c->h_meta = new int32_t[kMetaSize];
    c->h_fx = new double[81];
    c->h_idx = new int2[26 * kCells];
    char err[256];
    if (!build_tables(cfg->theta, c->tab, err, sizeof(err))) {
        delete c;
        return OCTOFMM_ERR_THETA;
    }
    if (hipSetDevice(c->device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void**)&c->d_tab, sizeof(Tables)) != hipSuccess ||
        hipMemcpy(c->d_tab, &c->tab, sizeof(Tables), hipMemcpyHostToDevice) != hipSuccess) {
        octofmm_destroy(c);
        return OCTOFMM_ERR_HIP;
    }
    c->d_scratch = scratch_address(c->device);
    if (!c->d_scratch) {
        octofmm_destroy(c);
        return OCTOFMM_ERR_HIP;
    }
    *out = c;
    return OCTOFMM_OK;
}

octofmm_status octofmm_destroy(octofmm_context* c) {
    if (!c) return OCTOFMM_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    delete[] c->h_meta;
    delete[] c->h_fx;
    delete[] c->h_idx;
    (void)hipFree(c->d_ws);
    (void)hipFree(c->d_tab);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return OCTOFMM_OK;
}

}   // extern "C"

// Argument checks of the multipole drop-ins; sets the item flags and the root
// choice of the call.
static octofmm_status check_multipole_call(octofmm_context* c, const octofmm_subgrid* in, double* L, double* L_c,
    int& flags, bool& root) {
    if (!in || !in->M || !in->com || !L) return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "null input or output");
    if (in->type != OCTOFMM_RHO && in->type != OCTOFMM_DRHODT) return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "bad type");
    if (in->type == OCTOFMM_RHO && !L_c) return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "RHO needs L_c");
    root = in->use_root_stencil != 0;
    flags = 0;
    if (!root)
        for (int f = 0; f < 26; ++f) {
            const octofmm_neighbor& nb = in->neighbors[f];
            if (nb.kind < 0 || nb.kind > 2) return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "bad neighbour kind %d", f);
            if (nb.kind != OCTOFMM_NEIGHBOR_EMPTY && !check_neighbor(nb, nb.kind == OCTOFMM_NEIGHBOR_MULTIPOLE))
                return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "bad neighbour %d", f);
            if (nb.kind == OCTOFMM_NEIGHBOR_MONOPOLE && nb.mon) flags |= kItemNear;
        }
    return check_finite_inputs(c, in, false);
}

extern "C" {

octofmm_status octofmm_compute_multipole_interactions(octofmm_context* c, const octofmm_subgrid* in,
    double* L, double* L_c) {
    if (!c) return OCTOFMM_ERR_INVALID_ARGUMENT;
    int flags = 0;
    bool root = false;
    octofmm_status s = check_multipole_call(c, in, L, L_c, flags, root);
    if (s != OCTOFMM_OK) return s;
    HIPC(c, hipSetDevice(c->device));
    (void)hipGetLastError();   // launch checks below report this call's launches only
    s = ensure_workspace(c);
    if (s != OCTOFMM_OK) return s;
    s = load_subgrid(c, c->ws, in, false, root, flags);
    if (s != OCTOFMM_OK) return s;
    return launch_subgrid(c, c->ws, in->dx, in->type, flags, root, L, in->type == OCTOFMM_RHO ? L_c : nullptr);
}

// The same call with the workspace in `scratch` instead of the context's
// allocation: the same kernels and copies, no device allocation.  Every
// region the kernels read is written by this call first (the target and
// handed-over neighbour slots, partial halos zeroed before their scatter, the
// metadata; empty directions are marked absent and never read), so what a
// previous call left in the scratch does not enter the result.
octofmm_status octofmm_compute_multipole_interactions_noalloc(octofmm_context* c, const octofmm_subgrid* in,
    double* L, double* L_c) {
    if (!c) return OCTOFMM_ERR_INVALID_ARGUMENT;
    int flags = 0;
    bool root = false;
    octofmm_status s = check_multipole_call(c, in, L, L_c, flags, root);
    if (s != OCTOFMM_OK) return s;
    if (!c->d_scratch) return fail(c, OCTOFMM_ERR_HIP, "no static workspace on device %d", c->device);
    HIPC(c, hipSetDevice(c->device));
    (void)hipGetLastError();   // launch checks below report this call's launches only
    std::lock_guard<std::mutex> lock(g_scratch_call_mu[c->device]);
    const SubgridWorkspace w = carve_workspace(c->d_scratch);
    s = load_subgrid(c, w, in, false, root, flags);
    if (s != OCTOFMM_OK) return s;
    return launch_subgrid(c, w, in->dx, in->type, flags, root, L, in->type == OCTOFMM_RHO ? L_c : nullptr);
}

octofmm_status octofmm_compute_monopole_interactions(octofmm_context* c, const octofmm_subgrid* in,
    int32_t contains_multipole_neighbor, double* L, double* L_c) {
    if (!c) return OCTOFMM_ERR_INVALID_ARGUMENT;
    if (!in || !in->mon || !in->com || !L) return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "null input or output");
    if (in->type != OCTOFMM_RHO && in->type != OCTOFMM_DRHODT) return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "bad type");
    const bool p2m = contains_multipole_neighbor != 0;
    if (in->type == OCTOFMM_RHO && p2m && !L_c) return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "RHO with p2m needs L_c");
    for (int f = 0; f < 26; ++f) {
        const octofmm_neighbor& nb = in->neighbors[f];
        if (nb.kind < 0 || nb.kind > 2) return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "bad neighbour kind %d", f);
        if (nb.kind != OCTOFMM_NEIGHBOR_EMPTY && !check_neighbor(nb, nb.kind == OCTOFMM_NEIGHBOR_MULTIPOLE))
            return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "bad neighbour %d", f);
    }
    octofmm_status s = check_finite_inputs(c, in, true);
    if (s != OCTOFMM_OK) return s;
    HIPC(c, hipSetDevice(c->device));
    (void)hipGetLastError();   // launch checks below report this call's launches only
    s = ensure_workspace(c);
    if (s != OCTOFMM_OK) return s;
    const int flags = kItemLeaf | (p2m ? kItemP2M : 0);
    s = load_subgrid(c, c->ws, in, true, false, flags);
    if (s != OCTOFMM_OK) return s;
    return launch_subgrid(c, c->ws, in->dx, in->type, flags, false, L,
        in->type == OCTOFMM_RHO && p2m ? L_c : nullptr);
}

// ---------------------------------------------------------------------------
// Level plans
// ---------------------------------------------------------------------------
static void plan_free(octofmm_plan* p) {
    if (!p) return;
    if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
    if (p->ev_join) (void)hipEventDestroy(p->ev_join);
    if (p->side) (void)hipStreamDestroy(p->side);
    (void)hipFree(p->d_rec);
    (void)hipFree(p->d_nbr);
    (void)hipFree(p->d_kind);
    (void)hipFree(p->d_foff);
    (void)hipFree(p->d_fxbase);
    (void)hipFree(p->d_items);
    (void)hipFree(p->d_mixbits);
    (void)hipFree(p->d_recm);
    (void)hipFree(p->d_mixoff);
    (void)hipFree(p->d_queue);
    (void)hipFree(p->d_p2m_out);
    (void)hipFree(p->d_partial);
    (void)hipFree(p->d_pslots);
    (void)hipFree(p->d_pidx);
    (void)hipFree(p->d_stage_list);
    delete p;
}

// stage_mask bit 0: stage the partner records; bit 1: the root kernel, or the
// level kernel over the multipole-target items; bit 2: the leaf kernel over
// the monopole-target items; bit 3: the p2m kernel of the leaf items with
// refined neighbours, whose sums the leaf kernel reads (so bit 3 goes before
// bit 2, on the same stream or ordered before it).  Bits 1 and 2 write
// disjoint sub-grids.
static octofmm_status plan_launch(octofmm_plan* p, hipStream_t st, int stage_mask) {
    octofmm_context* c = p->ctx;
    if ((stage_mask & 1) && p->npartial > 0) {
        hipLaunchKernelGGL(scatter_slots, dim3(p->npartial), dim3(256), 0, st, p->d_pslots,
            const_cast<double*>(p->d_M), const_cast<double*>(p->d_com), const_cast<double*>(p->d_mon));
        HIPC(c, hipGetLastError());
    }
    if (stage_mask & 1) {
        StageArgs sa;
        sa.M = p->d_M;
        sa.com = p->d_com;
        sa.mon = p->d_mon;
        sa.kind = p->d_kind;
        sa.fxbase = p->d_fxbase;
        sa.foff = p->d_foff;
        sa.rec = p->d_rec;
        sa.recm = p->d_recm;
        sa.list = p->d_stage_list;
        sa.dx = p->dx;
        sa.nslots = p->d_stage_list ? p->n_stage : p->nslots;
        if (sa.nslots > 0) {
            hipLaunchKernelGGL(stage_records, dim3(sa.nslots * 4), dim3(256), 0, st, sa);
            HIPC(c, hipGetLastError());
        }
    }
    if ((stage_mask & 2) && p->root) {
        RootArgs ra;
        ra.rec = p->d_rec;
        ra.tab = c->d_tab;
        ra.partial = p->d_partial;
        ra.per_chunk = kCells / kRootChunks;
        ra.nchunks = kRootChunks;
        if (p->type == OCTOFMM_RHO)
            hipLaunchKernelGGL(root_kernel<true>, dim3(kRootChunks), dim3(512), 0, st, ra);
        else
            hipLaunchKernelGGL(root_kernel<false>, dim3(kRootChunks), dim3(512), 0, st, ra);
        HIPC(c, hipGetLastError());
        RootReduceArgs rr;
        rr.partial = p->d_partial;
        rr.L = p->d_L;
        rr.Lc = p->d_Lc;
        rr.nchunks = kRootChunks;
        rr.rho = p->type == OCTOFMM_RHO;
        rr.overwrite = p->l2l ? 1 : 0;
        hipLaunchKernelGGL(root_reduce, dim3((kCells * 23 + 255) / 256), dim3(256), 0, st, rr);
        HIPC(c, hipGetLastError());
    } else if (!p->root && (stage_mask & 14) && p->nitems > 0) {
        LevelArgs la;
        la.rec = p->d_rec;
        la.com = p->d_com;
        la.L = p->d_L;
        la.Lc = p->d_Lc;
        la.nbr = p->d_nbr;
        la.kind = p->d_kind;
        la.items = p->d_items;
        la.tab = c->d_tab;
        la.mixbits = p->d_mixbits;
        la.recm = p->d_recm;
        la.mixoff = p->d_mixoff;
        la.queue = p->d_queue;
        la.mon = p->d_mon;
        la.p2m_out = p->d_p2m_out;
        la.parent = p->t_parent;
        la.octant = p->t_octant;
        la.pL = p->t_pL;
        la.pLc = p->t_pLc;
        la.pcom = p->t_pcom;
        la.g = p->t_g;
        la.G = p->t_G;
        la.l2l = p->l2l ? 1 : 0;
        la.correct_am = p->t_correct_am;
        la.leaf_rows = p->leaf_rows;
        la.inv_dx = 1.0 / p->dx;
        la.neg_inv_dx = -1.0 / p->dx;
        const bool rho = p->type == OCTOFMM_RHO;
        if ((stage_mask & 2) && p->n_m2l > 0) {
            la.nitems = p->n_m2l;
            la.item0 = 0;
            if (rho)
                launch_level<true>(p->variant, p->n_m2l, st, la, p->precision == OCTOFMM_PRECISION_MIXED);
            else
                launch_level<false>(p->variant, p->n_m2l, st, la, p->precision == OCTOFMM_PRECISION_MIXED);
            HIPC(c, hipGetLastError());
        }
        if ((stage_mask & 8) && p->n_p2m > 0) {
            // the leaf items with p2m lead the leaf items
            la.nitems = p->n_p2m;
            la.item0 = p->n_m2l;
            if (rho)
                launch_p2m<true>(p->n_p2m, st, la);
            else
                launch_p2m<false>(p->n_p2m, st, la);
            HIPC(c, hipGetLastError());
        }
        if ((stage_mask & 4) && p->nitems > p->n_m2l) {
            la.nitems = p->nitems - p->n_m2l;
            la.item0 = p->n_m2l;
            if (rho)
                launch_leaf<true>(la.nitems, st, la, c->tab);
            else
                launch_leaf<false>(la.nitems, st, la, c->tab);
            HIPC(c, hipGetLastError());
        }
    }
    return OCTOFMM_OK;
}

// The kernels of one pass over the masked stages.  A level with both kinds of
// items runs its leaf kernel (bit 2) on the plan's side stream, forked after
// the staging and joined back before returning: the two kernels write
// disjoint sub-grids, so the leaf kernel overlaps the multipole-target
// kernel instead of queueing behind it.
static bool plan_forks(const octofmm_plan* p, int mask) {
    return !p->root && p->n_m2l > 0 && p->nitems > p->n_m2l && (mask & 6) == 6;
}

static octofmm_status plan_fork_resources(octofmm_plan* p) {
    octofmm_context* c = p->ctx;
    if (p->side || !plan_forks(p, 15)) return OCTOFMM_OK;
    HIPC(c, hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking));
    HIPC(c, hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming));
    HIPC(c, hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming));
    return OCTOFMM_OK;
}

static octofmm_status plan_run(octofmm_plan* p, hipStream_t st, int mask) {
    octofmm_context* c = p->ctx;
    if (!plan_forks(p, mask) || !p->side) return plan_launch(p, st, mask);
    octofmm_status s = plan_launch(p, st, mask & 1);
    if (s != OCTOFMM_OK) return s;
    HIPC(c, hipEventRecord(p->ev_fork, st));
    HIPC(c, hipStreamWaitEvent(p->side, p->ev_fork, 0));
    s = plan_launch(p, p->side, mask & 12);
    if (s != OCTOFMM_OK) return s;
    s = plan_launch(p, st, 2);
    if (s != OCTOFMM_OK) return s;
    HIPC(c, hipEventRecord(p->ev_join, p->side));
    HIPC(c, hipStreamWaitEvent(st, p->ev_join, 0));
    return OCTOFMM_OK;
}

}   // extern "C"

// Kernel variants a plan may run: FP64 plans every variant, mixed-precision
// plans the mapping-0 forms (8, 4, 2 waves per workgroup; FP32 chosen per
// pair) and mapping 7 with its persistent form (704, 702, 708, 804, 802;
// FP32 chosen per coarse offset).
static bool variant_allowed(const octofmm_plan* p, int v) {
    bool known = false;
    for (int k : kVariants) known = known || k == v;
    const int mapping = v / 100;
    return known && !(p->precision == OCTOFMM_PRECISION_MIXED && mapping != 0 && mapping != 7 && mapping != 8);
}

// The plan's count of FP32 far pairs follows the chosen variant's rule.
static void set_mixed_count(octofmm_plan* p) {
    if (p->precision == OCTOFMM_PRECISION_MIXED)
        p->stats.mixed_far_pairs = p->variant >= 700 ? p->mixed_far_pairs_by_offset : p->mixed_far_pairs_by_pair;
}

// Tuning cache: the variant the autotuner chose, keyed by level shape (device
// architecture, solve type, precision, opening angle, counts of multipole-target
// items with and without a near pass, fused downward pass or not).  A plan of a
// shape seen before takes the cached choice without timing anything, so plans
// of one shape run the same kernel, and give the same bits, within a process.
// When the environment variable OCTOFMM_TUNE_CACHE names a file, the cache is
// read from it at the first lookup and every new choice is appended to it
// (one "key variant" line each), so the choices also hold across processes.
namespace {
std::mutex g_tune_mu;
std::vector<std::pair<std::string, int>> g_tune;
bool g_tune_loaded = false;

std::string tune_key(const octofmm_plan* p) {
    hipDeviceProp_t prop;
    const char* arch = hipGetDeviceProperties(&prop, p->ctx->device) == hipSuccess ? prop.gcnArchName : "unknown";
    int near = 0;
    for (const int2& it : p->h_items)
        if (!(it.y & kItemLeaf) && (it.y & kItemNear)) ++near;
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%s/cu%d/t%d/p%d/d%d/th%.17g/m%d/n%d/l%d", arch,
        hipGetDeviceProperties(&prop, p->ctx->device) == hipSuccess ? prop.multiProcessorCount : 0, p->type,
        p->precision, p->precision == OCTOFMM_PRECISION_MIXED ? p->mixed_far_d2 : 0, p->ctx->theta, p->n_m2l, near,
        p->l2l ? 1 : 0);
    for (char* q = buf; *q; ++q)
        if (*q == ' ') *q = '_';
    return buf;
}

// the variant cached for key, or 0
int tune_lookup(const std::string& key) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    if (!g_tune_loaded) {
        g_tune_loaded = true;
        if (const char* path = std::getenv("OCTOFMM_TUNE_CACHE")) {
            if (FILE* f = std::fopen(path, "r")) {
                char k[256];
                int v = 0;
                while (std::fscanf(f, "%255s %d", k, &v) == 2) g_tune.emplace_back(k, v);
                std::fclose(f);
            }
        }
    }
    int v = 0;
    for (const auto& e : g_tune)
        if (e.first == key) v = e.second;   // the last entry of a key wins
    return v;
}

void tune_store(const std::string& key, int v) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune.emplace_back(key, v);
    if (const char* path = std::getenv("OCTOFMM_TUNE_CACHE")) {
        if (FILE* f = std::fopen(path, "a")) {
            std::fprintf(f, "%s %d\n", key.c_str(), v);
            std::fclose(f);
        }
    }
}
}   // namespace

extern "C" {

// Times each level-kernel mapping on the live level's multipole-target items
// (the leaf kernel does not depend on it) and keeps the fastest.  The trial
// passes write into private scratch outputs, never the caller's L / L_c.
static octofmm_status plan_autotune_run(octofmm_plan* p);

static octofmm_status plan_autotune(octofmm_plan* p) {
    octofmm_context* c = p->ctx;
    double* userL = p->d_L;
    double* userLc = p->d_Lc;
    double *tL = nullptr, *tLc = nullptr;
    const size_t n = (size_t)p->nslots * kCells;
    if (hipMalloc((void**)&tL, n * 20 * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&tLc, n * 3 * sizeof(double)) != hipSuccess) {
        (void)hipFree(tL);
        return fail(c, OCTOFMM_ERR_ALLOCATION, "autotune scratch allocation failed");
    }
    p->d_L = tL;
    p->d_Lc = tLc;
    const octofmm_status s = plan_autotune_run(p);
    (void)hipStreamSynchronize(c->stream);
    p->d_L = userL;
    p->d_Lc = userLc;
    (void)hipFree(tL);
    (void)hipFree(tLc);
    return s;
}

static octofmm_status plan_autotune_run(octofmm_plan* p) {
    octofmm_context* c = p->ctx;
    hipStream_t st = c->stream;
    HIPC(c, hipStreamSynchronize(st));
    octofmm_status s = plan_launch(p, st, 1);
    if (s != OCTOFMM_OK) return s;
    hipEvent_t e0, e1;
    HIPC(c, hipEventCreate(&e0));
    HIPC(c, hipEventCreate(&e1));
    // Two rounds, the variants in list order and then in reverse, each
    // variant's times summed: the clock of a GPU that has just come out of
    // idle still rises over the first tens of milliseconds, which in one
    // round in list order favours the variants timed last.
    std::vector<std::pair<int, float>> timed;
    for (int w : kTunedVariants)
        if (variant_allowed(p, w)) timed.emplace_back(w, 0.f);
    const int nv = (int)timed.size();
    for (int round = 0; round < 2; ++round) {
        for (int i = 0; i < nv; ++i) {
            auto& t = timed[round == 0 ? i : nv - 1 - i];
            p->variant = t.first;
            s = plan_launch(p, st, 2);   // warm-up
            if (s != OCTOFMM_OK) return s;
            HIPC(c, hipEventRecord(e0, st));
            for (int l = 0; l < 3; ++l) {
                s = plan_launch(p, st, 2);
                if (s != OCTOFMM_OK) return s;
            }
            HIPC(c, hipEventRecord(e1, st));
            HIPC(c, hipEventSynchronize(e1));
            float ms = 0.f;
            HIPC(c, hipEventElapsedTime(&ms, e0, e1));
            t.second += ms;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    float best = 1e30f;
    int best_w = 8;
    for (const auto& t : timed) {
        if (t.second < best) {
            best = t.second;
            best_w = t.first;
        }
    }
    // Within 2% of the fastest the variant that re-reads the fewest bytes
    // from HBM wins (the timings of one plan creation spread by about 1%; in
    // interleaved A/B runs of 20 launches each on a 4096-sub-grid level,
    // 704 ran 1.3-1.8% faster than 702 every time, while the two-round
    // timing at plan creation still picked 702 in one process of six with a
    // 1% window): mapping 7 over 4 waves,
    // whose waves of one block stage overlapping neighbourhoods from one CU
    // at the same time (1.4x the algorithmic bytes on a 4096-sub-grid level,
    // 1.9x over 2 waves, 4.3x for mapping 3), then the other splits, mapping
    // 5, 3 and 0.
    auto rank = [](int w) {
        return w == 704 ? 0 : w / 100 == 7 ? 1 : w / 100 == 5 ? 2 : w / 100 == 3 ? 3 : 4;
    };
    float best_near = best;
    for (const auto& t : timed) {
        if (t.second > best * 1.02f) continue;
        if (rank(t.first) < rank(best_w) || (rank(t.first) == rank(best_w) && t.second < best_near)) {
            best_w = t.first;
            best_near = t.second;
        }
    }
    p->variant = best_w;
    return OCTOFMM_OK;
}

}   // extern "C"

// First half of level plan creation: checks, work items, pair counts, device
// tables.  The kernel variant and the graph come in level_plan_finish.
static octofmm_status level_plan_build(octofmm_context* c, const octofmm_level* lv, octofmm_plan** out) {
    if (lv->nslots <= 0 || !lv->neighbors || !lv->kind || !lv->xbase || lv->ntargets < 0 ||
        (lv->ntargets > 0 && !lv->targets))
        return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "bad level description");
    if (!lv->d_M || !lv->d_com || !lv->d_mon || !lv->d_L || (lv->type == OCTOFMM_RHO && !lv->d_L_c))
        return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "null device array");
    if (lv->type != OCTOFMM_RHO && lv->type != OCTOFMM_DRHODT) return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "bad type");
    if (lv->precision != OCTOFMM_PRECISION_FP64 && lv->precision != OCTOFMM_PRECISION_MIXED)
        return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "bad precision");
    if (!(lv->dx > 0.0)) return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "dx must be positive");
    if ((size_t)lv->nslots * kCells * kRecDoubles * sizeof(double) >= ((size_t)1 << 40))
        return fail(c, OCTOFMM_ERR_SIZE, "level too large");
    const int ns = lv->nslots;
    for (int s = 0; s < ns; ++s) {
        if (lv->kind[s] != OCTOFMM_SLOT_NONLEAF && lv->kind[s] != OCTOFMM_SLOT_LEAF)
            return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "bad kind of slot %d", s);
        for (int r = 0; r < 27; ++r) {
            const int n = lv->neighbors[(size_t)s * 27 + r];
            if (n < -1 || n >= ns) return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "bad neighbour of slot %d", s);
        }
        if (lv->neighbors[(size_t)s * 27 + 13] != s)
            return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "neighbour entry 13 of slot %d must be the slot", s);
    }
    const bool root = (lv->flags & OCTOFMM_PLAN_ROOT) != 0;
    if (root && (ns != 1 || lv->ntargets != 1 || lv->targets[0] != 0 || lv->kind[0] != OCTOFMM_SLOT_NONLEAF))
        return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "a root plan takes exactly one non-leaf slot as its target");
    // work items, heaviest kinds first: non-leaf with leaf neighbours, non-leaf,
    // leaf with refined neighbours, pure p2p leaves
    std::vector<int2> buckets[4];
    octofmm_plan_stats st;
    std::memset(&st, 0, sizeof(st));
    const Tables& T = c->tab;
    std::vector<char> seen(ns, 0);
    if (root) {
        // every pair of the root's cells except self and near offsets
        for (int t = 0; t < kCells; ++t)
            for (int q = 0; q < kCells; ++q) {
                const int dx = (q >> 6) - (t >> 6), dy = ((q >> 3) & 7) - ((t >> 3) & 7), dz = (q & 7) - (t & 7);
                const bool inbox = dx >= -5 && dx <= 5 && dy >= -5 && dy <= 5 && dz >= -5 && dz <= 5;
                if (inbox && T.root_excluded[(dx + 5) * 121 + (dy + 5) * 11 + (dz + 5)]) continue;
                st.far_pairs++;
            }
        st.n_m2l_items = 1;
    }
    for (int t = 0; t < (root ? 0 : lv->ntargets); ++t) {
        const int s = lv->targets[t];
        if (s < 0 || s >= ns) return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "bad target %d", t);
        if (seen[s]) return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "duplicate target slot %d", s);
        seen[s] = 1;
        bool has_leaf = false, has_nonleaf = false;
        for (int r = 0; r < 27; ++r) {
            if (r == 13) continue;
            const int n = lv->neighbors[(size_t)s * 27 + r];
            if (n < 0) continue;
            (lv->kind[n] == OCTOFMM_SLOT_LEAF ? has_leaf : has_nonleaf) = true;
        }
        if (lv->kind[s] == OCTOFMM_SLOT_NONLEAF) {
            buckets[has_leaf ? 0 : 1].push_back(make_int2(s, has_leaf ? kItemNear : 0));
            st.n_m2l_items++;
        } else {
            buckets[has_nonleaf ? 2 : 3].push_back(make_int2(s, kItemLeaf | (has_nonleaf ? kItemP2M : 0)));
            st.n_leaf_items++;
        }
    }
    // partial halo slots: each cell once, from its last entry (later entries
    // win, as in update_input's loop over the list)
    if (lv->npartial < 0 || (lv->npartial > 0 && !lv->partial))
        return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "bad partial halo list");
    if (root && lv->npartial > 0) return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "a root plan has no halo");
    std::vector<PartialSlot> pslots;
    std::vector<size_t> poff;   // first (cell, position) pair of each partial slot in pidx
    std::vector<int2> pidx;
    {
        std::vector<char> listed(ns, 0);
        for (int q = 0; q < lv->npartial; ++q) {
            const octofmm_partial_slot& e = lv->partial[q];
            if (e.slot < 0 || e.slot >= ns) return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "bad partial slot %d", q);
            if (seen[e.slot]) return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "partial slot %d is a target", e.slot);
            if (listed[e.slot]) return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "slot %d listed twice as partial", e.slot);
            listed[e.slot] = 1;
            if (e.count < 0 || e.count > kCells || (e.count > 0 && !e.cells))
                return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "bad cell list of partial slot %d", e.slot);
            const bool mp = lv->kind[e.slot] == OCTOFMM_SLOT_NONLEAF;
            if (e.count > 0 && (mp ? (!e.d_M || !e.d_com) : !e.d_mon))
                return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "null data of partial slot %d", e.slot);
            int last[kCells];
            for (int i = 0; i < kCells; ++i) last[i] = -1;
            for (int i = 0; i < e.count; ++i) {
                if (e.cells[i] < 0 || e.cells[i] >= kCells)
                    return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "bad cell in partial slot %d", e.slot);
                last[e.cells[i]] = i;
            }
            PartialSlot d;
            d.M = e.d_M;
            d.com = e.d_com;
            d.mon = e.d_mon;
            d.idx = nullptr;   // set once pidx is on the device
            poff.push_back(pidx.size());
            d.slot = e.slot;
            d.multipole = mp ? 1 : 0;
            d.pad = 0;
            d.n = 0;
            for (int cell = 0; cell < kCells; ++cell)
                if (last[cell] >= 0) {
                    pidx.push_back(make_int2(cell, last[cell]));
                    ++d.n;
                }
            pslots.push_back(d);
        }
    }
    // pair counts: per cell per coarse offset
    for (int b = 0; b < 4; ++b)
        for (const int2& it : buckets[b]) {
            const int s = it.x;
            for (int cx = 0; cx < 4; ++cx)
                for (int cy = 0; cy < 4; ++cy)
                    for (int cz = 0; cz < 4; ++cz)
                        for (int di = 0; di < T.ndelta; ++di) {
                            const int px = cx + T.delta[di][0], py = cy + T.delta[di][1], pz = cz + T.delta[di][2];
                            const int raw = (px < 0 ? 0 : px < 4 ? 1 : 2) + 3 * (py < 0 ? 0 : py < 4 ? 1 : 2) +
                                9 * (pz < 0 ? 0 : pz < 4 ? 1 : 2);
                            const int n = lv->neighbors[(size_t)s * 27 + raw];
                            if (n < 0) continue;
                            const int64_t nf = __builtin_popcountll(T.farbits[di]);
                            const int64_t nn = __builtin_popcountll(T.nearbits[di]);
                            if (!(it.y & kItemLeaf)) {
                                st.far_pairs += nf;
                                if (lv->kind[n] == OCTOFMM_SLOT_LEAF) st.near_pairs += nn;
                            } else if (lv->kind[n] == OCTOFMM_SLOT_LEAF) {
                                st.p2p_pairs += nf + nn;
                            } else if (it.y & kItemP2M) {
                                st.p2m_pairs += nf + nn;
                            }
                        }
        }
    const bool rho = lv->type == OCTOFMM_RHO;
    st.flops = st.far_pairs * (rho ? 407 : 273) + st.near_pairs * (rho ? 407 : 273) + st.p2p_pairs * 12 +
        st.p2m_pairs * (rho ? 318 : 205);
    std::vector<int2> items;
    for (int b = 0; b < 4; ++b) items.insert(items.end(), buckets[b].begin(), buckets[b].end());
    const int32_t n_m2l = (int32_t)(buckets[0].size() + buckets[1].size());
    // partner records are read by non-leaf targets (their own slot and every
    // neighbour) and by leaf targets with refined neighbours (those
    // neighbours); pure p2p leaves read monopoles directly
    std::vector<int32_t> stage_list;
    {
        std::vector<char> need(ns, 0);
        for (const int2& it : items) {
            const int32_t* nb = lv->neighbors + (size_t)it.x * 27;
            if (!(it.y & kItemLeaf)) {
                for (int r = 0; r < 27; ++r)
                    if (nb[r] >= 0) need[nb[r]] = 1;
            } else if (it.y & kItemP2M) {
                for (int r = 0; r < 27; ++r)
                    if (nb[r] >= 0 && lv->kind[nb[r]] == OCTOFMM_SLOT_NONLEAF) need[nb[r]] = 1;
            }
        }
        if (root) need[0] = 1;
        for (int s2 = 0; s2 < ns; ++s2)
            if (need[s2]) stage_list.push_back(s2);
    }

    HIPC(c, hipSetDevice(c->device));
    (void)hipGetLastError();   // launch checks below report this call's launches only
    octofmm_plan* p = new octofmm_plan();
    p->ctx = c;
    p->nslots = ns;
    p->nitems = (int32_t)items.size();
    p->n_m2l = n_m2l;
    p->n_stage = (int32_t)stage_list.size();
    p->type = lv->type;
    p->precision = lv->precision;
    p->flags = lv->flags;
    p->dx = lv->dx;
    p->d_M = lv->d_M;
    p->d_com = lv->d_com;
    p->d_mon = lv->d_mon;
    p->d_L = lv->d_L;
    p->d_Lc = lv->d_L_c;
    p->root = root;
    if (root) p->precision = OCTOFMM_PRECISION_FP64;
    p->stats = st;
    std::vector<int32_t> foff(3 * (size_t)ns, kInx);
    auto cleanup = [&](octofmm_status s, const char* what) {
        plan_free(p);
        return fail(c, s, "%s", what);
    };
    if (hipMalloc((void**)&p->d_rec, (size_t)ns * kCells * kRecDoubles * sizeof(double) + kRecTailBytes) != hipSuccess)
        return cleanup(OCTOFMM_ERR_ALLOCATION, "record buffer allocation failed");
    if (hipMalloc((void**)&p->d_nbr, sizeof(int32_t) * 27 * ns) != hipSuccess ||
        hipMalloc((void**)&p->d_kind, sizeof(int32_t) * ns) != hipSuccess ||
        hipMalloc((void**)&p->d_foff, sizeof(int32_t) * 3 * ns) != hipSuccess ||
        hipMalloc((void**)&p->d_fxbase, sizeof(double) * 3 * ns) != hipSuccess ||
        hipMalloc((void**)&p->d_items, sizeof(int2) * std::max<size_t>(1, items.size())) != hipSuccess)
        return cleanup(OCTOFMM_ERR_ALLOCATION, "plan table allocation failed");
    if (hipMemcpy(p->d_nbr, lv->neighbors, sizeof(int32_t) * 27 * ns, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->d_kind, lv->kind, sizeof(int32_t) * ns, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->d_foff, foff.data(), sizeof(int32_t) * 3 * ns, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->d_fxbase, lv->xbase, sizeof(double) * 3 * ns, hipMemcpyHostToDevice) != hipSuccess ||
        (!items.empty() &&
            hipMemcpy(p->d_items, items.data(), sizeof(int2) * items.size(), hipMemcpyHostToDevice) != hipSuccess))
        return cleanup(OCTOFMM_ERR_HIP, "plan table upload failed");
    if (root && hipMalloc((void**)&p->d_partial, sizeof(double) * kRootChunks * kCells * kRecDoubles) != hipSuccess)
        return cleanup(OCTOFMM_ERR_ALLOCATION, "root partial buffer allocation failed");
    if (!root && hipMalloc((void**)&p->d_queue, sizeof(uint32_t) * 8) != hipSuccess)
        return cleanup(OCTOFMM_ERR_ALLOCATION, "work counter allocation failed");
    p->n_p2m = (int32_t)buckets[2].size();
    if (p->n_p2m > 0 &&
        hipMalloc((void**)&p->d_p2m_out, sizeof(double) * (size_t)p->n_p2m * kCells * 8) != hipSuccess)
        return cleanup(OCTOFMM_ERR_ALLOCATION, "p2m sum buffer allocation failed");
    if (!pslots.empty()) {
        if (hipMalloc((void**)&p->d_pidx, sizeof(int2) * std::max<size_t>(1, pidx.size())) != hipSuccess ||
            hipMalloc((void**)&p->d_pslots, sizeof(PartialSlot) * pslots.size()) != hipSuccess)
            return cleanup(OCTOFMM_ERR_ALLOCATION, "partial halo table allocation failed");
        for (size_t q = 0; q < pslots.size(); ++q) pslots[q].idx = p->d_pidx + poff[q];
        if ((!pidx.empty() &&
                hipMemcpy(p->d_pidx, pidx.data(), sizeof(int2) * pidx.size(), hipMemcpyHostToDevice) != hipSuccess) ||
            hipMemcpy(p->d_pslots, pslots.data(), sizeof(PartialSlot) * pslots.size(), hipMemcpyHostToDevice) !=
                hipSuccess)
            return cleanup(OCTOFMM_ERR_HIP, "partial halo table upload failed");
        p->npartial = (int32_t)pslots.size();
    }
    if ((int)stage_list.size() < ns) {
        if (hipMalloc((void**)&p->d_stage_list, sizeof(int32_t) * std::max<size_t>(1, stage_list.size())) != hipSuccess ||
            (!stage_list.empty() && hipMemcpy(p->d_stage_list, stage_list.data(), sizeof(int32_t) * stage_list.size(),
                                        hipMemcpyHostToDevice) != hipSuccess))
            return cleanup(OCTOFMM_ERR_ALLOCATION, "staging list upload failed");
    }
    if (p->precision == OCTOFMM_PRECISION_MIXED) {
        // far pairs with |d|^2 > mixed_far_d2 (cell units) are evaluated in FP32
        std::vector<uint64_t> mb(T.ndelta > 0 ? T.ndelta : 1, 0);
        for (int di = 0; di < T.ndelta; ++di)
            for (int k = 0; k < 8; ++k)
                for (int j = 0; j < 8; ++j) {
                    const int bit = 8 * k + j;
                    if (!((T.farbits[di] >> bit) & 1)) continue;
                    const int dx = 2 * T.delta[di][0] + ((j >> 2) & 1) - ((k >> 2) & 1);
                    const int dy = 2 * T.delta[di][1] + ((j >> 1) & 1) - ((k >> 1) & 1);
                    const int dz = 2 * T.delta[di][2] + (j & 1) - (k & 1);
                    if (dx * dx + dy * dy + dz * dz > lv->mixed_far_d2) mb[di] |= uint64_t(1) << bit;
                }
        if (hipMalloc((void**)&p->d_mixbits, sizeof(uint64_t) * mb.size()) != hipSuccess ||
            hipMemcpy(p->d_mixbits, mb.data(), sizeof(uint64_t) * mb.size(), hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(OCTOFMM_ERR_ALLOCATION, "mixed-precision table upload failed");
        // lane mappings 3 and 7 choose the precision per far coarse offset:
        // FP32 where every far pair of the offset is an FP32 pair above.  The
        // split entries of opposite face offsets stage FP64 records only, so
        // their offsets stay FP64 (at mixed_far_d2 >= 4 they hold |d| = 2
        // pairs and would anyway)
        std::vector<uint8_t> mo(T.nsched > 0 ? T.nsched : 1, 0);
        std::vector<char> whole(T.ndelta > 0 ? T.ndelta : 1, 0);
        for (int q = 0; q < T.nsched; ++q) {
            if ((T.sched[q] >> 16) & 0xff) continue;
            const int di = T.sched[q] & 0xff;
            whole[di] = T.farbits[di] != 0 && mb[di] == T.farbits[di];
            mo[q] = whole[di] ? 1 : 0;
        }
        if (hipMalloc((void**)&p->d_mixoff, mo.size()) != hipSuccess ||
            hipMemcpy(p->d_mixoff, mo.data(), mo.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMalloc((void**)&p->d_recm, (size_t)ns * kCells * kRecDoubles * sizeof(double) + kRecTailBytes) !=
                hipSuccess)
            return cleanup(OCTOFMM_ERR_ALLOCATION, "mixed-precision record allocation failed");
        // FP32 pair counts under both rules, summed over the M2L items
        p->mixed_far_pairs_by_pair = 0;
        p->mixed_far_pairs_by_offset = 0;
        for (const int2& it : items) {
            if (it.y & kItemLeaf) continue;
            for (int cx = 0; cx < 4; ++cx)
                for (int cy = 0; cy < 4; ++cy)
                    for (int cz = 0; cz < 4; ++cz)
                        for (int di = 0; di < T.ndelta; ++di) {
                            const int px = cx + T.delta[di][0], py = cy + T.delta[di][1], pz = cz + T.delta[di][2];
                            const int raw = (px < 0 ? 0 : px < 4 ? 1 : 2) + 3 * (py < 0 ? 0 : py < 4 ? 1 : 2) +
                                9 * (pz < 0 ? 0 : pz < 4 ? 1 : 2);
                            if (lv->neighbors[(size_t)it.x * 27 + raw] < 0) continue;
                            p->mixed_far_pairs_by_pair += __builtin_popcountll(mb[di]);
                            if (whole[di]) p->mixed_far_pairs_by_offset += __builtin_popcountll(T.farbits[di]);
                        }
        }
        set_mixed_count(p);
    }
    p->mixed_far_d2 = lv->mixed_far_d2;
    p->h_items = items;
    *out = p;
    return OCTOFMM_OK;
}

// Second half of plan creation, once the plan's epilogue is wired: the kernel
// variant (OCTOFMM_VARIANT, else with OCTOFMM_PLAN_AUTOTUNE the tuning cache
// or a timing of every variant on the live level, else the default) and, with
// OCTOFMM_PLAN_USE_GRAPH, the captured graph.  Frees the plan on failure.
static octofmm_status level_plan_finish(octofmm_plan* p, int flags) {
    octofmm_context* c = p->ctx;
    const char* forced = std::getenv("OCTOFMM_VARIANT");
    const bool forced_ok = forced && variant_allowed(p, std::atoi(forced));
    if (forced_ok) {
        p->variant = std::atoi(forced);
    } else if ((flags & OCTOFMM_PLAN_AUTOTUNE) && p->n_m2l > 0 && !p->root) {
        const std::string key = tune_key(p);
        const int cached = tune_lookup(key);
        if (cached && variant_allowed(p, cached)) {
            p->variant = cached;
        } else {
            // time each workgroup shape on the live level (outputs are scratch
            // here: the caller zeroes L before a real pass) and keep the fastest
            const octofmm_status ts = plan_autotune(p);
            if (ts != OCTOFMM_OK) {
                plan_free(p);
                return ts;
            }
            tune_store(key, p->variant);
        }
    }
    p->stats.variant = p->variant;
    set_mixed_count(p);
    if (flags & OCTOFMM_PLAN_USE_GRAPH) {
        // the side stream must exist before the capture starts
        octofmm_status fs = plan_fork_resources(p);
        if (fs != OCTOFMM_OK) {
            plan_free(p);
            return fs;
        }
        hipStream_t cs = c->stream;
        if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            plan_free(p);
            return fail(c, OCTOFMM_ERR_HIP, "graph capture failed to start");
        }
        octofmm_status ls = plan_run(p, cs, 15);
        hipGraph_t g = nullptr;
        const hipError_t ee = hipStreamEndCapture(cs, &g);
        if (ls != OCTOFMM_OK || ee != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            plan_free(p);
            return fail(c, OCTOFMM_ERR_HIP, "graph capture failed");
        }
        p->graph = g;
        if (hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0) != hipSuccess) {
            plan_free(p);
            return fail(c, OCTOFMM_ERR_HIP, "graph instantiation failed");
        }
    }
    return OCTOFMM_OK;
}

extern "C" {

octofmm_status octofmm_level_plan_create(octofmm_context* c, const octofmm_level* lv, octofmm_plan** out) {
    if (!c) return OCTOFMM_ERR_INVALID_ARGUMENT;
    if (!lv || !out) return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "null level or plan");
    *out = nullptr;
    octofmm_plan* p = nullptr;
    octofmm_status s = level_plan_build(c, lv, &p);
    if (s != OCTOFMM_OK) return s;
    s = level_plan_finish(p, lv->flags);
    if (s != OCTOFMM_OK) return s;
    *out = p;
    return OCTOFMM_OK;
}

octofmm_status octofmm_level_plan_execute(octofmm_plan* p, void* stream) {
    if (!p) return OCTOFMM_ERR_INVALID_ARGUMENT;
    octofmm_context* c = p->ctx;
    HIPC(c, hipSetDevice(c->device));
    (void)hipGetLastError();   // launch checks below report this call's launches only
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (p->gexec) {
        HIPC(c, hipGraphLaunch(p->gexec, st));
        return OCTOFMM_OK;
    }
    const octofmm_status fs = plan_fork_resources(p);   // first eager execute
    if (fs != OCTOFMM_OK) return fs;
    return plan_run(p, st, 15);
}

octofmm_status octofmm_level_plan_stats(const octofmm_plan* p, octofmm_plan_stats* s) {
    if (!p || !s) return OCTOFMM_ERR_INVALID_ARGUMENT;
    *s = p->stats;
    return OCTOFMM_OK;
}

octofmm_status octofmm_level_plan_time_stages(octofmm_plan* p, void* stream, int32_t iters, double* ms) {
    if (!p || !ms || iters <= 0) return OCTOFMM_ERR_INVALID_ARGUMENT;
    octofmm_context* c = p->ctx;
    HIPC(c, hipSetDevice(c->device));
    (void)hipGetLastError();   // launch checks below report this call's launches only
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const octofmm_status fs = plan_fork_resources(p);
    if (fs != OCTOFMM_OK) return fs;
    hipEvent_t e0, e1;
    HIPC(c, hipEventCreate(&e0));
    HIPC(c, hipEventCreate(&e1));
    for (int stage = 0; stage < 2; ++stage) {
        HIPC(c, hipEventRecord(e0, st));
        for (int i = 0; i < iters; ++i) {
            octofmm_status s = plan_run(p, st, stage == 0 ? 1 : 14);
            if (s != OCTOFMM_OK) return s;
        }
        HIPC(c, hipEventRecord(e1, st));
        HIPC(c, hipEventSynchronize(e1));
        float f = 0.f;
        HIPC(c, hipEventElapsedTime(&f, e0, e1));
        ms[stage] = (double)f / iters;
    }
    ms[2] = 0.0;   // write-back is fused into the interaction kernel
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return OCTOFMM_OK;
}

octofmm_status octofmm_level_plan_destroy(octofmm_plan* p) {
    if (!p) return OCTOFMM_OK;
    (void)hipSetDevice(p->ctx->device);
    plan_free(p);
    return OCTOFMM_OK;
}


}   // extern "C"

// ---------------------------------------------------------------------------
// Whole-tree solve
// ---------------------------------------------------------------------------
struct octofmm_tree_plan {
    struct Level {
        int32_t nnodes = 0;
        double dx = 0.0;
        double* M = nullptr;
        double* com = nullptr;
        const double* mon = nullptr;
        double* L = nullptr;
        double* Lc = nullptr;
        double* g = nullptr;
        int32_t* d_kind = nullptr;
        int32_t* d_parent = nullptr;
        int32_t* d_octant = nullptr;
        double* d_xbase = nullptr;
        octofmm_plan* plan = nullptr;
    };
    octofmm_context* ctx = nullptr;
    std::vector<Level> lv;
    int32_t type = 0, flags = 0, correct_am = 1;
    double G = 1.0;
    octofmm_tree_stats stats;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // the interactions phase forks onto three more streams: one stages the
    // partner records of every level, coarsest first, the other two run the
    // leaf kernels of alternate levels (the leaf kernel of level l + 1 need not
    // wait for that of level l); the level kernels stay on the caller's stream
    hipStream_t side_stage = nullptr, side_leaf[2] = {nullptr, nullptr};
    hipEvent_t ev_up = nullptr, ev_join_stage = nullptr, ev_join_leaf[2] = {nullptr, nullptr};
    std::vector<hipEvent_t> ev_stage, ev_level;
};

namespace {

void tree_free(octofmm_tree_plan* t) {
    if (!t) return;
    if (t->gexec) (void)hipGraphExecDestroy(t->gexec);
    if (t->graph) (void)hipGraphDestroy(t->graph);
    for (hipEvent_t e : {t->ev_up, t->ev_join_stage, t->ev_join_leaf[0], t->ev_join_leaf[1]})
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : t->ev_stage)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : t->ev_level)
        if (e) (void)hipEventDestroy(e);
    if (t->side_stage) (void)hipStreamDestroy(t->side_stage);
    for (hipStream_t s : t->side_leaf)
        if (s) (void)hipStreamDestroy(s);
    for (auto& l : t->lv) {
        plan_free(l.plan);
        (void)hipFree(l.d_kind);
        (void)hipFree(l.d_parent);
        (void)hipFree(l.d_octant);
        (void)hipFree(l.d_xbase);
    }
    delete t;
}

// phase bit 0: upward pass; bit 1: interactions with the fused downward pass
octofmm_status tree_launch(octofmm_tree_plan* t, hipStream_t st, int phases) {
    octofmm_context* c = t->ctx;
    const int nl = (int)t->lv.size();
    const int rho = t->type == OCTOFMM_RHO;
    if (phases & 1)
        for (int l = nl - 1; l >= 0; --l) {
            const auto& L = t->lv[l];
            M2MArgs a;
            a.M = L.M;
            a.com = L.com;
            a.mon = L.mon;
            a.kind = L.d_kind;
            a.parent = L.d_parent;
            a.octant = L.d_octant;
            a.xbase = L.d_xbase;
            a.pM = l > 0 ? t->lv[l - 1].M : nullptr;
            a.pcom = l > 0 ? t->lv[l - 1].com : nullptr;
            a.dx = L.dx;
            a.nnodes = L.nnodes;
            a.rho = rho;
            hipLaunchKernelGGL(m2m_kernel, dim3((L.nnodes * kCells + 255) / 256), dim3(256), 0, st, a);
            HIPC(c, hipGetLastError());
        }
    // Interactions of every level, root first.  Each level's kernels also
    // write the final expansions (the downward pass is fused into them), so
    // both kernels of level l read the finished non-leaf sub-grids of level
    // l - 1, which only the level kernel of l - 1 writes: the leaf kernel of
    // level l waits for that kernel and for its own level's records, and
    // overlaps the level kernels that follow.  Staging depends only on the
    // upward pass.
    if (phases & 2) {
        hipStream_t ss = t->side_stage;
        HIPC(c, hipEventRecord(t->ev_up, st));
        HIPC(c, hipStreamWaitEvent(ss, t->ev_up, 0));
        for (hipStream_t sl : t->side_leaf) HIPC(c, hipStreamWaitEvent(sl, t->ev_up, 0));
        for (int l = 0; l < nl; ++l) {
            const octofmm_status s = plan_launch(t->lv[l].plan, ss, 1);
            if (s != OCTOFMM_OK) return s;
            HIPC(c, hipEventRecord(t->ev_stage[l], ss));
        }
        for (int l = 0; l < nl; ++l) {
            octofmm_plan* p = t->lv[l].plan;
            HIPC(c, hipStreamWaitEvent(st, t->ev_stage[l], 0));
            octofmm_status s = plan_launch(p, st, 2);
            if (s != OCTOFMM_OK) return s;
            HIPC(c, hipEventRecord(t->ev_level[l], st));
            if (!p->root && p->nitems > p->n_m2l) {
                // the p2m sums need only this level's records; the leaf
                // kernel also needs the finished level above
                hipStream_t sl = t->side_leaf[l & 1];
                HIPC(c, hipStreamWaitEvent(sl, t->ev_stage[l], 0));
                s = plan_launch(p, sl, 8);
                if (s != OCTOFMM_OK) return s;
                if (l > 0) HIPC(c, hipStreamWaitEvent(sl, t->ev_level[l - 1], 0));
                s = plan_launch(p, sl, 4);
                if (s != OCTOFMM_OK) return s;
            }
        }
        HIPC(c, hipEventRecord(t->ev_join_stage, ss));
        HIPC(c, hipStreamWaitEvent(st, t->ev_join_stage, 0));
        for (int k = 0; k < 2; ++k) {
            HIPC(c, hipEventRecord(t->ev_join_leaf[k], t->side_leaf[k]));
            HIPC(c, hipStreamWaitEvent(st, t->ev_join_leaf[k], 0));
        }
    }
    return OCTOFMM_OK;
}

}   // namespace

extern "C" {

octofmm_status octofmm_tree_plan_create(octofmm_context* c, const octofmm_tree* tr, octofmm_tree_plan** out) {
    if (!c) return OCTOFMM_ERR_INVALID_ARGUMENT;
    if (!tr || !out || tr->nlevels <= 0 || !tr->levels) return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "bad tree");
    *out = nullptr;
    if (tr->type != OCTOFMM_RHO && tr->type != OCTOFMM_DRHODT) return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "bad type");
    const bool rho = tr->type == OCTOFMM_RHO;
    const int nl = tr->nlevels;
    // structure: root alone on level 0, 8 children per non-leaf node, leaves at the bottom
    for (int l = 0; l < nl; ++l) {
        const octofmm_tree_level& L = tr->levels[l];
        if (L.nnodes <= 0 || !L.neighbors || !L.kind || !L.parent || !L.octant || !L.xbase || !(L.dx > 0.0))
            return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "bad description of level %d", l);
        if (!L.d_M || !L.d_com || !L.d_mon || !L.d_L || (rho && !L.d_L_c))
            return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "null device array on level %d", l);
        // expansion, correction and force rows are moved in 16-byte pieces
        if (((uintptr_t)L.d_L | (uintptr_t)L.d_L_c | (uintptr_t)L.d_g) & 15)
            return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "L, L_c and g of level %d must be 16-byte aligned", l);
        if (l == 0 && (L.nnodes != 1 || L.parent[0] != -1))
            return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "level 0 must hold the root alone, without a parent");
        std::vector<int> children(l + 1 < nl ? tr->levels[l + 1].nnodes : 0, 0);
        for (int n = 0; n < L.nnodes; ++n) {
            if (L.kind[n] != OCTOFMM_SLOT_NONLEAF && L.kind[n] != OCTOFMM_SLOT_LEAF)
                return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "bad kind of node %d on level %d", n, l);
            if (l == nl - 1 && L.kind[n] == OCTOFMM_SLOT_NONLEAF)
                return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "non-leaf node %d on the finest level", n);
            if (L.octant[n] < 0 || L.octant[n] > 7)
                return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "bad octant of node %d on level %d", n, l);
            if ((tr->flags & OCTOFMM_TREE_LEAF_G_ONLY) && L.kind[n] == OCTOFMM_SLOT_LEAF && !L.d_g)
                return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT,
                    "OCTOFMM_TREE_LEAF_G_ONLY needs d_g on level %d, which holds leaves", l);
        }
        if (l > 0) {
            const octofmm_tree_level& U = tr->levels[l - 1];
            std::vector<uint8_t> used((size_t)U.nnodes * 8, 0);
            for (int n = 0; n < L.nnodes; ++n) {
                const int pn = L.parent[n];
                if (pn < 0 || pn >= U.nnodes || U.kind[pn] != OCTOFMM_SLOT_NONLEAF)
                    return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "node %d on level %d has no non-leaf parent", n, l);
                uint8_t& u = used[(size_t)pn * 8 + L.octant[n]];
                if (u) return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "two children in one octant on level %d", l);
                u = 1;
            }
            for (int pn = 0; pn < U.nnodes; ++pn) {
                if (U.kind[pn] != OCTOFMM_SLOT_NONLEAF) continue;
                for (int o = 0; o < 8; ++o)
                    if (!used[(size_t)pn * 8 + o])
                        return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "non-leaf node %d on level %d lacks child %d",
                            pn, l - 1, o);
            }
        }
    }
    HIPC(c, hipSetDevice(c->device));
    (void)hipGetLastError();   // launch checks below report this call's launches only
    octofmm_tree_plan* t = new octofmm_tree_plan();
    t->ctx = c;
    t->type = tr->type;
    t->flags = tr->flags;
    t->G = tr->G;
    t->correct_am = tr->correct_am;
    std::memset(&t->stats, 0, sizeof(t->stats));
    t->stats.nlevels = nl;
    t->lv.resize(nl);
    t->ev_stage.assign(nl, nullptr);
    t->ev_level.assign(nl, nullptr);
    auto cleanup = [&](octofmm_status s, const char* what) {
        tree_free(t);
        return fail(c, s, "%s", what);
    };
    {
        bool ok = hipStreamCreateWithFlags(&t->side_stage, hipStreamNonBlocking) == hipSuccess &&
                  hipStreamCreateWithFlags(&t->side_leaf[0], hipStreamNonBlocking) == hipSuccess &&
                  hipStreamCreateWithFlags(&t->side_leaf[1], hipStreamNonBlocking) == hipSuccess;
        for (hipEvent_t* e : {&t->ev_up, &t->ev_join_stage, &t->ev_join_leaf[0], &t->ev_join_leaf[1]})
            ok = ok && hipEventCreateWithFlags(e, hipEventDisableTiming) == hipSuccess;
        for (int l = 0; l < nl; ++l)
            ok = ok && hipEventCreateWithFlags(&t->ev_stage[l], hipEventDisableTiming) == hipSuccess &&
                 hipEventCreateWithFlags(&t->ev_level[l], hipEventDisableTiming) == hipSuccess;
        if (!ok) return cleanup(OCTOFMM_ERR_HIP, "tree stream or event creation failed");
    }
    for (int l = 0; l < nl; ++l) {
        const octofmm_tree_level& L = tr->levels[l];
        auto& D = t->lv[l];
        D.nnodes = L.nnodes;
        D.dx = L.dx;
        D.M = L.d_M;
        D.com = L.d_com;
        D.mon = L.d_mon;
        D.L = L.d_L;
        D.Lc = L.d_L_c;
        D.g = L.d_g;
        const size_t n = (size_t)L.nnodes;
        if (hipMalloc((void**)&D.d_kind, sizeof(int32_t) * n) != hipSuccess ||
            hipMalloc((void**)&D.d_parent, sizeof(int32_t) * n) != hipSuccess ||
            hipMalloc((void**)&D.d_octant, sizeof(int32_t) * n) != hipSuccess ||
            hipMalloc((void**)&D.d_xbase, sizeof(double) * 3 * n) != hipSuccess)
            return cleanup(OCTOFMM_ERR_ALLOCATION, "tree table allocation failed");
        if (hipMemcpy(D.d_kind, L.kind, sizeof(int32_t) * n, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(D.d_parent, L.parent, sizeof(int32_t) * n, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(D.d_octant, L.octant, sizeof(int32_t) * n, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(D.d_xbase, L.xbase, sizeof(double) * 3 * n, hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(OCTOFMM_ERR_HIP, "tree table upload failed");
        t->stats.nodes += L.nnodes;
        for (int i = 0; i < L.nnodes; ++i) t->stats.leaves += L.kind[i] == OCTOFMM_SLOT_LEAF;
    }
    // one upward pass first, so that autotuning times the interactions on real multipoles
    {
        const octofmm_status s = tree_launch(t, c->stream, 1);
        if (s != OCTOFMM_OK) {
            tree_free(t);
            return s;
        }
        if (hipStreamSynchronize(c->stream) != hipSuccess) return cleanup(OCTOFMM_ERR_HIP, "upward pass failed");
    }
    for (int l = 0; l < nl; ++l) {
        const octofmm_tree_level& L = tr->levels[l];
        std::vector<int32_t> targets(L.nnodes);
        for (int i = 0; i < L.nnodes; ++i) targets[i] = i;
        octofmm_level lv;
        std::memset(&lv, 0, sizeof(lv));
        lv.nslots = L.nnodes;
        lv.neighbors = L.neighbors;
        lv.kind = L.kind;
        lv.xbase = L.xbase;
        lv.dx = L.dx;
        lv.ntargets = L.nnodes;
        lv.targets = targets.data();
        lv.type = tr->type;
        lv.precision = tr->precision;
        lv.mixed_far_d2 = tr->mixed_far_d2;
        lv.flags = tr->flags & OCTOFMM_PLAN_AUTOTUNE;
        if (l == 0 && L.kind[0] == OCTOFMM_SLOT_NONLEAF) {
            lv.flags = OCTOFMM_PLAN_ROOT;
            lv.precision = OCTOFMM_PRECISION_FP64;
        }
        lv.d_M = L.d_M;
        lv.d_com = L.d_com;
        lv.d_mon = L.d_mon;
        lv.d_L = L.d_L;
        lv.d_L_c = rho ? L.d_L_c : nullptr;
        // the level's plan is tuned only once its epilogue is the fused
        // downward pass it runs with in the solve
        octofmm_plan* lp = nullptr;
        octofmm_status s = level_plan_build(c, &lv, &lp);
        if (s == OCTOFMM_OK) {
            lp->l2l = true;
            lp->t_parent = t->lv[l].d_parent;
            lp->t_octant = t->lv[l].d_octant;
            lp->t_pL = l > 0 ? t->lv[l - 1].L : nullptr;
            lp->t_pLc = l > 0 ? t->lv[l - 1].Lc : nullptr;
            lp->t_pcom = l > 0 ? t->lv[l - 1].com : nullptr;
            lp->t_g = t->lv[l].g;
            lp->t_G = t->G;
            lp->t_correct_am = t->correct_am;
            lp->leaf_rows = (tr->flags & OCTOFMM_TREE_LEAF_G_ONLY) ? 0 : 1;
            s = level_plan_finish(lp, lv.flags);
        }
        if (s != OCTOFMM_OK) {
            const std::string msg = c->err;
            tree_free(t);
            return fail(c, s, "level %d: %s", l, msg.c_str());
        }
        t->lv[l].plan = lp;
        const octofmm_plan_stats& ps = t->lv[l].plan->stats;
        if (t->lv[l].plan->root)
            t->stats.root_pairs += ps.far_pairs;
        else
            t->stats.far_pairs += ps.far_pairs;
        t->stats.near_pairs += ps.near_pairs;
        t->stats.p2p_pairs += ps.p2p_pairs;
        t->stats.p2m_pairs += ps.p2m_pairs;
        t->stats.flops += ps.flops;
        t->stats.mixed_far_pairs += ps.mixed_far_pairs;
    }
    if (tr->flags & OCTOFMM_PLAN_USE_GRAPH) {
        hipStream_t cs = c->stream;
        if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return cleanup(OCTOFMM_ERR_HIP, "graph capture failed to start");
        const octofmm_status ls = tree_launch(t, cs, 7);
        hipGraph_t g = nullptr;
        const hipError_t ee = hipStreamEndCapture(cs, &g);
        if (ls != OCTOFMM_OK || ee != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return cleanup(OCTOFMM_ERR_HIP, "graph capture failed");
        }
        t->graph = g;
        if (hipGraphInstantiate(&t->gexec, g, nullptr, nullptr, 0) != hipSuccess)
            return cleanup(OCTOFMM_ERR_HIP, "graph instantiation failed");
    }
    *out = t;
    return OCTOFMM_OK;
}

octofmm_status octofmm_tree_plan_execute(octofmm_tree_plan* t, void* stream) {
    if (!t) return OCTOFMM_ERR_INVALID_ARGUMENT;
    octofmm_context* c = t->ctx;
    HIPC(c, hipSetDevice(c->device));
    (void)hipGetLastError();   // launch checks below report this call's launches only
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (t->gexec) {
        HIPC(c, hipGraphLaunch(t->gexec, st));
        return OCTOFMM_OK;
    }
    return tree_launch(t, st, 7);
}

octofmm_status octofmm_tree_plan_stats(const octofmm_tree_plan* t, octofmm_tree_stats* s) {
    if (!t || !s) return OCTOFMM_ERR_INVALID_ARGUMENT;
    *s = t->stats;
    return OCTOFMM_OK;
}

octofmm_status octofmm_tree_plan_time_phases(octofmm_tree_plan* t, void* stream, int32_t iters, double* ms) {
    if (!t || !ms || iters <= 0) return OCTOFMM_ERR_INVALID_ARGUMENT;
    octofmm_context* c = t->ctx;
    HIPC(c, hipSetDevice(c->device));
    (void)hipGetLastError();   // launch checks below report this call's launches only
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipEvent_t e0, e1;
    HIPC(c, hipEventCreate(&e0));
    HIPC(c, hipEventCreate(&e1));
    for (int ph = 0; ph < 3; ++ph) {
        HIPC(c, hipEventRecord(e0, st));
        for (int i = 0; i < iters; ++i) {
            const octofmm_status s = tree_launch(t, st, 1 << ph);
            if (s != OCTOFMM_OK) return s;
        }
        HIPC(c, hipEventRecord(e1, st));
        HIPC(c, hipEventSynchronize(e1));
        float f = 0.f;
        HIPC(c, hipEventElapsedTime(&f, e0, e1));
        ms[ph] = (double)f / iters;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return OCTOFMM_OK;
}

octofmm_status octofmm_measure_fma_rate(octofmm_context* c, int32_t precision, double* tflops) {
    if (!c || !tflops) return OCTOFMM_ERR_INVALID_ARGUMENT;
    if (precision != OCTOFMM_PRECISION_FP64 && precision != OCTOFMM_PRECISION_MIXED)
        return fail(c, OCTOFMM_ERR_INVALID_ARGUMENT, "precision must be FP64 or MIXED (FP32)");
    HIPC(c, hipSetDevice(c->device));
    (void)hipGetLastError();   // launch checks below report this call's launches only
    // 4096 workgroups of 256 lanes (16 per CU on 256 CUs), 8 chains x 2048 steps each
    constexpr int kBlocks = 4096, kIters = 2048;
    void* out = nullptr;
    HIPC(c, hipMalloc(&out, sizeof(double) * (size_t)kBlocks * 256));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto run = [&]() {
        if (precision == OCTOFMM_PRECISION_FP64)
            hipLaunchKernelGGL(fma_rate_kernel<double>, dim3(kBlocks), dim3(256), 0, c->stream,
                static_cast<double*>(out), 0.999999, 1.0e-7, kIters);
        else
            hipLaunchKernelGGL(fma_rate_kernel<float>, dim3(kBlocks), dim3(256), 0, c->stream,
                static_cast<float*>(out), 0.999999f, 1.0e-7f, kIters);
    };
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        run();
        e = hipEventRecord(e0, c->stream);
    }
    if (e == hipSuccess) {
        run();
        e = hipEventRecord(e1, c->stream);
    }
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    HIPC(c, e);
    const double fmas = (double)kBlocks * 256.0 * 8.0 * kIters;
    *tflops = ms > 0.f ? 2.0 * fmas / (ms * 1e-3) / 1e12 : 0.0;
    return OCTOFMM_OK;
}

octofmm_status octofmm_tree_plan_destroy(octofmm_tree_plan* t) {
    if (!t) return OCTOFMM_OK;
    (void)hipSetDevice(t->ctx->device);
    tree_free(t);
    return OCTOFMM_OK;
}

}   // extern "C"
