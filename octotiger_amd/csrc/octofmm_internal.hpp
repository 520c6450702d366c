// octofmm_internal.hpp -- definitions shared by the host-side plan builder and
// the gfx950 kernels of liboctofmm.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace octofmm {

constexpr int kInx = 8;
constexpr int kCells = 512;
constexpr int kRecDoubles = 24;       // partner record: m0, M1..M19, Y0..Y2, mon
constexpr int kRecBytes = kRecDoubles * 8;
constexpr int kRecY = 20;
constexpr int kRecMon = 23;
constexpr int kMaxDelta = 125;        // coarse offsets in [-2, 2]^3 (theta >= 0.34)
constexpr int kStencilInx = 11;       // stencil box edge (offsets -5..5)
constexpr int kStencilBox = kStencilInx * kStencilInx * kStencilInx;
constexpr double kThetaFloor = 0.34;  // OCTOTIGER_THETA_MINIMUM default (CMakeLists.txt:66)

// Item flags of a level plan
constexpr int kItemLeaf = 1;          // leaf target: p2p (+ p2m)
constexpr int kItemNear = 2;          // non-leaf target with a leaf neighbour: near pass
constexpr int kItemP2M = 4;           // leaf target with a refined neighbour: p2m + L_c

// Interaction tables derived from theta by the reference's stencil rule.
// A target cell t = 2C + k (C its coarse cell, k its child index
// (x&1)*4 + (y&1)*2 + (z&1)) interacts with partner p = 2(C + delta) + j for
// every coarse offset delta that passes the opening criterion; the pair is
// "far" (multipole partner) or "near" (monopole-only partner) by the fine
// offset d = 2 delta + j - k.
struct Tables {
    int32_t ndelta;
    int32_t max_delta;                // largest |delta| component
    int32_t nfar_per_cell;            // far partners of one cell
    int32_t nnear_per_cell;           // near partners of one cell
    int8_t delta[kMaxDelta][4];       // coarse offsets (x, y, z, 0)
    uint64_t farbits[kMaxDelta];      // byte k: bit j set if (delta, j, k) is far
    uint64_t nearbits[kMaxDelta];     // byte k: bit j set if (delta, j, k) is near (self excluded)
    uint64_t allrows[kMaxDelta];      // farbits | nearbits transposed: byte j, bit k
    int32_t nfar_delta;               // coarse offsets with a far pair for some child class,
    int32_t far_delta[kMaxDelta];     // in delta order (delta 0 has none: it is all near / self)
    // Far schedule of the direct-staging kernels: one staged image of 8
    // partner records per coarse cell for each entry, packed as
    // a | b << 8 | kind << 16 | axis << 24 (a, b delta indices; axis = the
    // child-index bit of x, y or z: 2, 1, 0).
    //   kind 0: the 8 children of offset a, in child order.
    //   kinds 1 and 2 take a pair of opposite face offsets a = +e, b = -e
    //   when its far bits are those of theta = 0.5: far(a, k, j) iff
    //   j_axis = 1 or k_axis = 0, and far(b, k, j) iff j_axis = 0 or
    //   k_axis = 1.  Walked as two plain images, a class would idle on 4 of
    //   the 8 children of one of them.  Instead,
    //   kind 1 stages a's children with j_axis = 1 and b's with j_axis = 0
    //   (records 0..3 and 4..7, each in child order), far for every class;
    //   kind 2 stages a's children with j_axis = 0 and b's with j_axis = 1,
    //   and class k reads only the half of its own side, k_axis = 0 the
    //   first, k_axis = 1 the second: 12 steps for the pair instead of 16,
    //   each far pair once.
    // Per entry and class k (byte k): the records class k reads, as bits of
    // the image's records 0..7 taken from offset a's partner (sched_fara)
    // and from b's (sched_farb), and bit k of sched_half set where class k
    // reads records 4..7 as its records 0..3 (kind 2, k_axis = 1).
    // sched_rowa / sched_rowb hold the same bits transposed: byte j, bit k
    // set where class k's step j reads a record of a's (b's) partner, so a
    // wave forms the lane mask of step j from scalar registers.
    int32_t nsched;
    int32_t sched[kMaxDelta];
    uint64_t sched_fara[kMaxDelta];
    uint64_t sched_farb[kMaxDelta];
    uint64_t sched_rowa[kMaxDelta];
    uint64_t sched_rowb[kMaxDelta];
    uint32_t sched_half[kMaxDelta];
    double four[kStencilBox][4];      // p2p constants (-1/r, x/r^3, y/r^3, z/r^3) by fine offset
    uint8_t root_excluded[kStencilBox];   // root kernel: near offsets and self
    uint8_t phase[kStencilBox];       // 1 far, 0 near (inside the stencil), 1 elsewhere
    uint8_t mask[kStencilBox];        // union stencil mask
    // per child class k: the (delta, j) pairs of its far and near partners,
    // packed as delta_index * 8 + j, ordered by delta
    int32_t nfar_k[8];
    int32_t nnear_k[8];
    uint16_t farlist[8][kMaxDelta * 8];
    uint16_t nearlist[8][kMaxDelta * 8];
    // p2p of a leaf target of class k: every stencil partner (far and near),
    // packed as const_byte_offset | (window_offset + 2048) << 16, where the
    // window is the sub-grid plus a halo of p2p_halo cells (edge 8 + 2
    // p2p_halo, x slowest), window_offset = (dx * W + dy) * W + dz, and
    // const_byte_offset locates the pair's four constants in p2pconst
    int32_t p2p_halo;
    // 1 when every class list is the whole 6^3 fine block around the target's
    // coarse cell minus the target (max |delta| = 1, i.e. theta = 0.5): the
    // leaf kernel then walks that block with compile-time offsets and takes
    // the constants straight from `four`
    int32_t p2p_cube;
    int32_t np2p_k[8];
    int32_t p2plist[8][kMaxDelta * 8];
    // the distinct four-constant rows the eight lists use (342 at theta 0.5),
    // in order of first use: one compact table that every class shares, so
    // the scalar cache holds it whole while the eight class waves stream
    // their lists (per-class copies of the rows are five times larger)
    int32_t np2pconst;
    double p2pconst[kStencilBox][4];
};

// Level-kernel variants: lane mapping x 100 + waves per workgroup.
//   mapping 0: a wave owns a 2x2x2 block of coarse cells, lane = (coarse cell,
//              child); the 8 lanes of a coarse cell share every partner load
//              (8, 4, 2 waves per workgroup).  The per-sub-grid drop-ins run
//              variant 8.
//   mapping 3: the partner records of each coarse offset staged once per wave
//              in LDS by direct global->LDS loads and read back as 8-lane
//              broadcasts; 4 or 2 waves per workgroup so that 3 waves per
//              SIMD fit the 12 KiB LDS image of each wave.
//   mapping 5: the coarse offsets of each 2x2x2 coarse block split over 4
//              waves (variant 504), for levels with few sub-grids.
//   mapping 7: mapping 3's staging with the coarse offsets of a block split
//              over 4, 2 or 8 waves (variants 704, 702, 708); the partial sums
//              meet in the waves' images.  The 8-wave split (one workgroup per
//              CU by its LDS) is for levels with very few sub-grids.
//   mapping 8: mapping 7 as one persistent grid over per-XCD work counters
//              (variants 804, 802).
// Mappings 1, 2 and 4 were measured and are not built (slower than mapping 3
// or 7 at every level size: one child class per wave, register-staged LDS
// images, double-buffered images).
constexpr int kVariants[11] = {8, 4, 2, 304, 302, 504, 704, 702, 708, 804, 802};
// the variants the autotuner times: mapping 8 is 5-9% slower than mapping 7
// on every level measured (each work unit waits for its counter's atomic
// before staging, where the hardware dispatcher overlaps a new workgroup's
// staging with the others' arithmetic), so it is built and tested for
// measurement (OCTOFMM_VARIANT) but not timed at plan creation
constexpr int kTunedVariants[9] = {8, 4, 2, 304, 302, 504, 704, 702, 708};

// Builds the tables; returns false when theta is below the floor or the
// stencil fails the reference's symmetry / width checks.
bool build_tables(double theta, Tables& t, char* err, size_t errlen);

// Lane masks of the direct-staging kernels, whose lane is group * 8 + class
// (group = the partner coarse cell's slot in the wave's block).
//
// The 8 low bits of e spread to the 8 bytes of a 64-bit lane mask (bit g ->
// byte g all ones): x * 0x204081 places bit i of a nibble at bit 8i with no
// carries (the four shifted copies do not overlap), the mask keeps those
// bits and * 0xff fills their bytes.
__host__ __device__ inline uint64_t spread_groups(unsigned e) {
    const unsigned lo = (((e & 15u) * 0x204081u) & 0x01010101u) * 0xffu;
    const unsigned hi = ((((e >> 4) & 15u) * 0x204081u) & 0x01010101u) * 0xffu;
    return ((uint64_t)hi << 32) | lo;
}

// byte b repeated over the 8 bytes of a 64-bit lane mask
__host__ __device__ inline uint64_t repeat_bytes(unsigned b) {
    const unsigned w = b * 0x01010101u;
    return ((uint64_t)w << 32) | w;
}

// The lanes that evaluate step j of an image: the classes in byte j of the
// step rows of the records taken from offset a's partners (rowa) and from
// b's (rowb), in every group except those whose partner is gone (gone_a,
// gone_b: spread_groups of the groups' bits).
__host__ __device__ inline uint64_t step_lanes(uint64_t rowa, uint64_t rowb, uint64_t gone_a, uint64_t gone_b, int j) {
    return (repeat_bytes((unsigned)(rowa >> (8 * j)) & 0xffu) & ~gone_a) |
        (repeat_bytes((unsigned)(rowb >> (8 * j)) & 0xffu) & ~gone_b);
}

__host__ __device__ inline int child_of(int x, int y, int z) { return ((x & 1) << 2) | ((y & 1) << 1) | (z & 1); }
__host__ __device__ inline int coarse_of(int x, int y, int z) { return ((x >> 1) << 4) | ((y >> 1) << 2) | (z >> 1); }
// record index of cell (x,y,z) inside its slot: coarse-major, child-minor
__host__ __device__ inline int record_of_cell(int x, int y, int z) { return coarse_of(x, y, z) * 8 + child_of(x, y, z); }

}   // namespace octofmm
