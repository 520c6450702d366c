/*
 * octofmm.h -- C-ABI of the MI355X (gfx950) FMM gravity interaction kernels.
 *
 * This is the drop-in boundary for Octo-Tiger's cell-to-cell gravity
 * interaction path.  Every entry point takes plain pointers and sizes; no
 * C++ or PyTorch types cross it.  The reference functions each entry point
 * replaces are cited next to it (paths relative to the Octo-Tiger source
 * tree).
 *
 * Conventions (shared by every entry point):
 *   - A sub-grid is 8x8x8 cells.  A cell's flat index is x*64 + y*8 + z.
 *   - Multipoles / expansions are 20 doubles per cell in Octo-Tiger's
 *     taylor<4> component order: [0] mass, [1..3] x y z, [4..9] xx xy xz yy yz
 *     zz, [10..19] xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz
 *     (octotiger/taylor.hpp:346-420, src/taylor.cpp:10-58).
 *   - Neighbour directions use geo::direction's flat order: raw index
 *     i = (dx+1) + 3*(dy+1) + 9*(dz+1) in 0..26, the centre (13) skipped, so
 *     flat index = i < 13 ? i : i - 1 (octotiger/geometry.hpp:91-194).
 *   - L is accumulated (+=), L_c is overwritten, exactly as the reference
 *     writes grid::get_L() and grid::get_L_c()
 *     (octotiger/multipole_interactions/kernel/kokkos_kernel.hpp:966-981).
 *   - Errors are reported as octofmm_status codes; nothing aborts.
 */
#ifndef OCTOFMM_H
#define OCTOFMM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OCTOFMM_INX 8
#define OCTOFMM_CELLS 512
#define OCTOFMM_NCOMP 20
#define OCTOFMM_NDIR 26
#define OCTOFMM_ABI_VERSION 1

typedef enum octofmm_status {
    OCTOFMM_OK = 0,
    OCTOFMM_ERR_INVALID_ARGUMENT = 1, /* null pointer, bad count, bad kind */
    OCTOFMM_ERR_THETA = 2,            /* theta outside [theta_min, 1/2] or stencil check failed */
    OCTOFMM_ERR_NO_DEVICE = 3,        /* no HIP device / bad ordinal */
    OCTOFMM_ERR_HIP = 4,              /* a HIP runtime call failed */
    OCTOFMM_ERR_ALLOCATION = 5,       /* device or pinned host allocation failed */
    OCTOFMM_ERR_SIZE = 6,             /* a size exceeds what the plan / library supports */
    OCTOFMM_ERR_NONFINITE = 7         /* non-finite input detected by a checking entry point */
} octofmm_status;

/* gsolve_type (octotiger/defs.hpp:73-75) */
typedef enum octofmm_solve_type { OCTOFMM_RHO = 0, OCTOFMM_DRHODT = 1 } octofmm_solve_type;

/* What a neighbour direction holds (neighbor_gravity_type,
 * octotiger/interaction_types.hpp:27-64): a refined (non-leaf) neighbour hands
 * over multipoles and centres of mass, a leaf neighbour hands over monopoles,
 * an empty direction (physical boundary) contributes zeros. */
typedef enum octofmm_neighbor_kind {
    OCTOFMM_NEIGHBOR_EMPTY = 0,
    OCTOFMM_NEIGHBOR_MULTIPOLE = 1,
    OCTOFMM_NEIGHBOR_MONOPOLE = 2
} octofmm_neighbor_kind;

typedef enum octofmm_precision {
    OCTOFMM_PRECISION_FP64 = 0,
    /* FP32 pair arithmetic for far offsets with |d|^2 > mixed_far_d2 (cell
     * units); FP64 accumulation, FP64 angular-momentum correction, FP64 near
     * offsets, root kernel always FP64. */
    OCTOFMM_PRECISION_MIXED = 1
} octofmm_precision;

typedef struct octofmm_neighbor {
    int32_t kind;          /* octofmm_neighbor_kind */
    int32_t count;         /* cells provided: 512 = full sub-grid, fewer = partial halo */
    const double* M;       /* [count][20] multipoles (MULTIPOLE), else NULL */
    const double* com;     /* [count][3]  centres of mass (MULTIPOLE), else NULL */
    const double* mon;     /* [count]     monopoles (MONOPOLE), else NULL */
    const int32_t* cells;  /* [count] neighbour-local flat indices of a partial halo, in the
                              order the data arrays hold them (grid::get_ilist_n_bnd order);
                              NULL when count == 512 */
} octofmm_neighbor;

/* One sub-grid's inputs, as multipole_interaction_interface::update_input and
 * monopole_interactions::update_input receive them
 * (octotiger/multipole_interactions/legacy/multipole_interaction_interface.hpp:88-96,
 *  octotiger/monopole_interactions/legacy/monopole_interaction_interface.hpp:64-68). */
typedef struct octofmm_subgrid {
    const double* M;       /* [512][20] own multipoles (non-leaf targets and root) */
    const double* com;     /* [512][3]  own centres of mass, com_ptr[0] */
    const double* mon;     /* [512]     own monopoles (leaf targets) */
    octofmm_neighbor neighbors[OCTOFMM_NDIR];
    double dx;             /* cell width */
    double xbase[3];       /* centre of own cell (0,0,0) */
    int32_t type;          /* octofmm_solve_type */
    int32_t use_root_stencil;
} octofmm_subgrid;

typedef struct octofmm_config {
    double theta;          /* opening criterion, Octo-Tiger default 0.5 */
    int32_t device;        /* HIP device ordinal */
    int32_t flags;         /* OCTOFMM_CHECK_FINITE or 0 */
} octofmm_config;

/* The per-sub-grid entry points reject non-finite inputs (M, com, mon of the
 * sub-grid and of every neighbour handed over) with OCTOFMM_ERR_NONFINITE
 * before any device work.  No counterpart in the reference, which does not
 * check; off by default (one host pass over ~2.6 MB per call). */
#define OCTOFMM_CHECK_FINITE 1

typedef struct octofmm_context octofmm_context;

/* Creates a context on one GPU and builds the stencils by the reference's
 * rule (src/multipole_interactions/util/calculate_stencil.cpp:23-161,
 * src/monopole_interactions/util/calculate_stencil.cpp:19-177).  Replaces the
 * start-up stencil initialisation of frontend/init_methods.cpp:189-192 and
 * init_fmm_constants (src/cuda_util/cuda_scheduler.cpp:31-90).  A context may be
 * used by one host thread at a time. */
octofmm_status octofmm_create(const octofmm_config* config, octofmm_context** ctx);
/* Destroy every level and tree plan made from a context before the context:
 * plans use it until they are destroyed. */
octofmm_status octofmm_destroy(octofmm_context* ctx);
const char* octofmm_status_string(octofmm_status status);
/* Text of the last error recorded on the context (never NULL). */
const char* octofmm_last_error(const octofmm_context* ctx);
int32_t octofmm_abi_version(void);

/* Drop-in for
 *   multipole_interaction_interface::compute_multipole_interactions(monopoles, M_ptr, com_ptr,
 *       neighbors, type, dx, is_direction_empty, xbase, use_root_stencil)
 *   (octotiger/multipole_interactions/legacy/multipole_interaction_interface.hpp:38-43,
 *    src/multipole_interactions/legacy/multipole_interaction_interface.cpp:82-179,
 *    its CUDA twin src/multipole_interactions/legacy/cuda_multipole_interaction_interface.cpp:56-279,
 *    and the free function multipole_kernel_interface,
 *    src/multipole_interactions/multipole_kernel_interface.cpp:73-169).
 * Non-leaf target: M2L over the theta stencil (non-root) or over all pairs of
 * the root sub-grid (use_root_stencil).  Host pointers in, synchronous.
 *   L   [512][20]  accumulated
 *   L_c [512][3]   overwritten (RHO only; untouched for DRHODT) */
octofmm_status octofmm_compute_multipole_interactions(octofmm_context* ctx,
    const octofmm_subgrid* subgrid, double* L, double* L_c);

/* Drop-in for
 *   monopole_interaction_interface::compute_interactions(monopoles, com_ptr, neighbors, type, dx,
 *       is_direction_empty, grid, contains_multipole_neighbor)
 *   (octotiger/monopole_interactions/legacy/monopole_interaction_interface.hpp:35-39,
 *    src/monopole_interactions/legacy/monopole_interaction_interface.cpp,
 *    and monopole_kernel_interface, src/monopole_interactions/monopole_kernel_interface.cpp:97-200).
 * Leaf target: p2p against own and leaf-neighbour monopoles, plus p2m against
 * every refined (MULTIPOLE) neighbour when contains_multipole_neighbor != 0.
 *   L   [512][20]  components 0..3 accumulated
 *   L_c [512][3]   overwritten when RHO and contains_multipole_neighbor, else untouched */
octofmm_status octofmm_compute_monopole_interactions(octofmm_context* ctx,
    const octofmm_subgrid* subgrid, int32_t contains_multipole_neighbor, double* L, double* L_c);

/* octofmm_compute_multipole_interactions without any device allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (tight per-node loops, code that must not call hipMalloc).  Same
 * arguments, kernels, results (bit for bit) and synchronous behaviour; its
 * device workspace is a static array in the library's code object, 16 MiB per
 * GPU, held from the first octofmm_create on that GPU until the process
 * ends.  The maximum is one call's inputs: a sub-grid and its 26 neighbours,
 * each of at most 512 cells (larger counts are rejected with
 * OCTOFMM_ERR_INVALID_ARGUMENT).  Every call on one GPU uses that GPU's one
 * workspace, which keeps the last call's contents (each call rewrites what it
 * reads); calls on one GPU are serialised within the process, and callers
 * must not expect them to overlap. */
octofmm_status octofmm_compute_multipole_interactions_noalloc(octofmm_context* ctx,
    const octofmm_subgrid* subgrid, double* L, double* L_c);

/* ---------------------------------------------------------------------------
 * Level batch: every sub-grid of one octree level in one device pass.
 * This has no single counterpart in the reference; it batches what
 * node_server::compute_fmm (src/node_server.cpp:589-798) issues as one call of
 * the two entry points above per octree node, and generalises the reference's
 * aggregated p2p launch (octotiger/monopole_interactions/kernel/kokkos_kernel.hpp:321-514).
 * All data stays on the device.  A plan is built once per level shape and
 * executed every solve.
 * ------------------------------------------------------------------------- */
typedef enum octofmm_slot_kind {
    OCTOFMM_SLOT_NONLEAF = 0,  /* refined: holds multipoles (M, com) */
    OCTOFMM_SLOT_LEAF = 1      /* leaf: holds monopoles (mon); seen by neighbours at cell centres */
} octofmm_slot_kind;

/* A halo slot of a level handed over partially, as a remote neighbour hands
 * it over: only `count` cells, in the order grid::get_ilist_n_bnd lists them
 * (octotiger/multipole_interactions/legacy/multipole_interaction_interface.hpp:177,257,
 *  octotiger/monopole_interactions/legacy/p2m_interaction_interface.hpp:142,239).
 * At every execute the plan writes the slot's rows of the level's d_M and
 * d_com (non-leaf slot) or d_mon (leaf slot): zeros, then each listed cell
 * from the compact arrays, a cell listed twice taking its later entry, which
 * is what update_input stages from such a list. */
typedef struct octofmm_partial_slot {
    int32_t slot;
    int32_t count;             /* entries, 0..512 */
    const int32_t* cells;      /* host [count] flat cell indices x*64 + y*8 + z */
    const double* d_M;         /* device [count][20] (non-leaf slot) */
    const double* d_com;       /* device [count][3]  (non-leaf slot) */
    const double* d_mon;       /* device [count]     (leaf slot) */
} octofmm_partial_slot;

typedef struct octofmm_level {
    int32_t nslots;            /* sub-grid slots in the device arrays */
    const int32_t* neighbors;  /* host [nslots][27] slot of the neighbour in raw direction
                                  (dx+1)+3(dy+1)+9(dz+1); entry 13 is the slot itself;
                                  -1 = empty direction */
    const int32_t* kind;       /* host [nslots] octofmm_slot_kind */
    const double* xbase;       /* host [nslots][3] centre of each slot's cell (0,0,0) */
    double dx;
    int32_t ntargets;
    const int32_t* targets;    /* host [ntargets] slots whose interactions are computed */
    int32_t type;              /* octofmm_solve_type */
    int32_t precision;         /* octofmm_precision */
    int32_t mixed_far_d2;      /* MIXED: FP32 for far offsets with |d|^2 > mixed_far_d2 (4 = "beyond 2 cells") */
    int32_t flags;             /* OCTOFMM_PLAN_* */
    /* device arrays, AoS per slot, caller-owned */
    const double* d_M;         /* [nslots][512][20] multipoles (zeros in leaf slots) */
    const double* d_com;       /* [nslots][512][3]  centres of mass */
    const double* d_mon;       /* [nslots][512]     monopoles (zeros in non-leaf slots) */
    double* d_L;               /* [nslots][512][20] accumulated for target slots */
    double* d_L_c;             /* [nslots][512][3]  overwritten for target slots (RHO) */
    /* halo slots handed over partially (none: npartial = 0); each names a
       slot once and no target; their rows of d_M / d_com / d_mon are
       written by the plan (octofmm_partial_slot) */
    int32_t npartial;
    int32_t reserved;
    const octofmm_partial_slot* partial;   /* host [npartial] */
} octofmm_level;

#define OCTOFMM_PLAN_USE_GRAPH 1   /* capture the level pass in one hipGraph */
#define OCTOFMM_PLAN_AUTOTUNE 2    /* time the kernel variants once and keep the fastest */
/* The environment variable OCTOFMM_VARIANT, when it names a variant of the
 * multipole-target kernel (lane mapping x 100 + waves per workgroup: 8, 4, 2,
 * 304, 302, 504, 704, 702, 708, and 804, 802 for mapping 7 as one persistent
 * grid over per-XCD work counters; mixed precision: 8, 4, 2, 704, 702, 708,
 * 804, 802), fixes that variant
 * for every plan created while it is set, in place of the default or the
 * autotuner's choice.  All variants give the same results to rounding; it
 * exists for measurement.
 * With OCTOFMM_PLAN_AUTOTUNE the choice is cached by level shape (device
 * architecture and CU count, solve type, precision, opening angle, counts of
 * multipole-target items with and without leaf neighbours, fused downward pass
 * or not): a plan of a shape seen before in the process takes the cached
 * variant without timing.  When the environment variable OCTOFMM_TUNE_CACHE
 * names a file, the cache is read from it and each new choice appended to it. */
/* The level is the root sub-grid alone (nslots = ntargets = 1, non-leaf):
 * every pair of its cells except self and near offsets, the root stencil of
 * octotiger/multipole_interactions/kernel/kokkos_kernel.hpp:324-471, 637-764.
 * Always FP64. */
#define OCTOFMM_PLAN_ROOT 4

typedef struct octofmm_plan octofmm_plan;

typedef struct octofmm_plan_stats {
    int64_t far_pairs;         /* M2L pairs with far (multipole) partners */
    int64_t near_pairs;        /* M2L near pairs with monopole partners (evaluated) */
    int64_t p2p_pairs;         /* leaf p2p pairs */
    int64_t p2m_pairs;         /* leaf p2m pairs */
    int64_t flops;             /* nominal FLOPs of one execution */
    int64_t mixed_far_pairs;   /* far pairs evaluated in FP32 (MIXED) */
    int32_t n_m2l_items;
    int32_t n_leaf_items;
    int32_t variant;           /* kernel variant chosen */
    int32_t reserved;
} octofmm_plan_stats;

octofmm_status octofmm_level_plan_create(octofmm_context* ctx, const octofmm_level* level,
    octofmm_plan** plan);
/* Executes the plan on a HIP stream (hipStream_t, NULL = default stream); asynchronous.
 * A level with both multipole and leaf targets runs its leaf kernel on a
 * plan-owned stream, forked after staging and joined back into `stream`, so
 * the pass is still ordered like one launch on `stream`.  Executes of one plan
 * must not overlap. */
octofmm_status octofmm_level_plan_execute(octofmm_plan* plan, void* stream);
octofmm_status octofmm_level_plan_stats(const octofmm_plan* plan, octofmm_plan_stats* stats);
/* Runs each stage of the plan `iters` times on `stream`, bracketed by HIP
 * events recorded on that stream, and writes the average milliseconds per
 * stage: ms[0] AoS->SoA staging, ms[1] interaction kernels (both kinds,
 * overlapped as in execute), ms[2] = 0 (write-back is fused). */
octofmm_status octofmm_level_plan_time_stages(octofmm_plan* plan, void* stream, int32_t iters,
    double* ms);
octofmm_status octofmm_level_plan_destroy(octofmm_plan* plan);

/* ---------------------------------------------------------------------------
 * Whole-tree gravity solve: the passes node_server::compute_fmm runs per node
 * (src/node_server.cpp:589-798), batched per octree level and captured in
 * one hipGraph:
 *   1. upward M2M, finest level first (grid::compute_multipoles,
 *      src/grid_fmm.cpp:1639-1803, taylor::operator>>=, octotiger/taylor.hpp:222-255);
 *   2. interactions of every level, root first (level plans; the root
 *      sub-grid with the root stencil), each level's kernels also applying
 *      the downward L2L from the finished level above and the leaf force
 *      assembly (grid::compute_expansions, src/grid_fmm.cpp:1524-1637,
 *      taylor::operator<<=, octotiger/taylor.hpp:263-311).
 * Level 0 holds the root alone.  Every non-leaf node of level l has exactly 8
 * children on level l+1; the neighbours of a non-leaf node exist on its level
 * unless outside the domain (proper nesting).
 * ------------------------------------------------------------------------- */
typedef struct octofmm_tree_level {
    int32_t nnodes;
    const int32_t* neighbors;  /* host [nnodes][27], as octofmm_level.neighbors */
    const int32_t* kind;       /* host [nnodes] octofmm_slot_kind */
    const int32_t* parent;     /* host [nnodes] index of the parent on the level above; -1 on level 0 */
    const int32_t* octant;     /* host [nnodes] 0..7, which 4^3 cell block of the parent:
                                  bit 0 x, bit 1 y, bit 2 z (geo::octant, octotiger/geometry.hpp:333-335) */
    const double* xbase;       /* host [nnodes][3] centre of each node's cell (0,0,0) */
    double dx;                 /* cell width on this level */
    /* device arrays, AoS per node, caller-owned; d_L, d_L_c and d_g 16-byte aligned
       (hipMalloc's are) */
    double* d_M;               /* [nnodes][512][20] written by the upward pass (non-leaf nodes) */
    double* d_com;             /* [nnodes][512][3]  written by the upward pass for RHO (leaf cells:
                                  geometric centres); DRHODT reuses the RHO pass's values */
    const double* d_mon;       /* [nnodes][512]     leaf cell masses (input; unused in non-leaf nodes) */
    double* d_L;               /* [nnodes][512][20] output: expansion of every cell */
    double* d_L_c;             /* [nnodes][512][3]  output (RHO): angular-momentum correction */
    double* d_g;               /* [nnodes][512][4]  output, may be NULL: leaf cells' phi = G L0 and
                                  g = -G (L_1..3 + L_c) (L_c only with correct_am); DRHODT: phi only */
} octofmm_tree_level;

/* Tree flag: leaf cells get only g (and L_c for RHO); their L rows are left
 * as they are.  grid::compute_expansions uses a leaf's L only to form its
 * potential and gravity (src/grid_fmm.cpp:1613-1629), so callers that read g
 * alone skip most of the leaf rows' write traffic.  Needs d_g on every level
 * with leaves.  No counterpart in the reference, which keeps leaf L. */
#define OCTOFMM_TREE_LEAF_G_ONLY 8
typedef struct octofmm_tree {
    int32_t nlevels;
    const octofmm_tree_level* levels;
    int32_t type;              /* octofmm_solve_type */
    int32_t precision;         /* octofmm_precision of the non-root levels */
    int32_t mixed_far_d2;
    int32_t flags;             /* OCTOFMM_PLAN_USE_GRAPH, OCTOFMM_PLAN_AUTOTUNE, OCTOFMM_TREE_LEAF_G_ONLY */
    double G;                  /* gravitational constant of the force assembly */
    int32_t correct_am;        /* add L_c to g (opts().correct_am_grav) */
    int32_t reserved;
} octofmm_tree;

typedef struct octofmm_tree_plan octofmm_tree_plan;

typedef struct octofmm_tree_stats {
    int64_t far_pairs;         /* M2L far pairs over all non-root levels */
    int64_t near_pairs;
    int64_t p2p_pairs;
    int64_t p2m_pairs;
    int64_t root_pairs;        /* M2L pairs of the root sub-grid */
    int64_t flops;             /* nominal FLOPs of the interactions of one solve */
    int64_t mixed_far_pairs;
    int32_t nodes;
    int32_t leaves;
    int32_t nlevels;
    int32_t reserved;
} octofmm_tree_stats;

octofmm_status octofmm_tree_plan_create(octofmm_context* ctx, const octofmm_tree* tree,
    octofmm_tree_plan** plan);
/* One full solve on a HIP stream (NULL = default stream); asynchronous.
 * L and L_c are cleared by the solve itself. */
octofmm_status octofmm_tree_plan_execute(octofmm_tree_plan* plan, void* stream);
octofmm_status octofmm_tree_plan_stats(const octofmm_tree_plan* plan, octofmm_tree_stats* stats);
/* Runs each phase `iters` times between HIP events on `stream`; ms[0] upward
 * pass, ms[1] interactions of all levels with the downward pass and the leaf
 * force assembly fused into their kernels, ms[2] = 0 (kept for layout). */
octofmm_status octofmm_tree_plan_time_phases(octofmm_tree_plan* plan, void* stream, int32_t iters,
    double* ms);
octofmm_status octofmm_tree_plan_destroy(octofmm_tree_plan* plan);

/* ---------------------------------------------------------------------------
 * Device FMA rate (diagnostic).  No counterpart in the reference: it measures
 * the dense vector FMA rate of the context's device, the practical peak to
 * set the interaction kernels' rates against.  Every lane runs 8 independent FMA chains
 * (precision OCTOFMM_PRECISION_FP64: v_fma_f64; OCTOFMM_PRECISION_MIXED: the
 * FP32 form, unpacked as all device code here is built) over enough
 * workgroups to fill every CU; *tflops = 2 x FMAs / time of the timed launch,
 * run on the context's stream after one warm-up.  Allocates and frees an
 * 8 MiB scratch output; synchronous.
 * ------------------------------------------------------------------------- */
octofmm_status octofmm_measure_fma_rate(octofmm_context* ctx, int32_t precision, double* tflops);

#ifdef __cplusplus
}
#endif

#endif /* OCTOFMM_H */
