"""Deterministic synthetic gravity states for the interaction kernels.

Every workload is a *level*: a block of 8x8x8-cell sub-grid slots with a
27-entry neighbour table per slot (raw geo::direction order, centre = the
slot itself, -1 = empty direction), a slot kind (non-leaf: multipoles, leaf:
monopoles) and the per-cell state in Octo-Tiger's AoS layout:

    M   [nslots, 512, 20]  multipoles (taylor<4> order), zeros in leaf slots
    com [nslots, 512, 3]   centres of mass (geometric cell centres in leaves)
    mon [nslots, 512]      monopoles, zeros in non-leaf slots

Random numbers come from SplitMix64 counters turned into uniforms on [0, 1)
and Box-Muller normals, so any config is reproducible from its seed alone.
Multipoles are built from 8 point masses per cell at the child-cell centres
(x +- dx/4): mass = sum, centre of mass = mass-weighted mean, and the order
1..3 moments sum m (x - com)^alpha, which is what the M2M shift operator
taylor::operator>>= (octotiger/taylor.hpp:222-255) produces for point masses.
"""
from dataclasses import dataclass, field

import numpy as np

from . import _abi

SEED_BASE = 20261202

# order-1..3 multi-indices in taylor<4> component order (octotiger/taylor.hpp:346-420)
_ALPHA = [
    (1, 0, 0), (0, 1, 0), (0, 0, 1),
    (2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1), (0, 0, 2),
    (3, 0, 0), (2, 1, 0), (2, 0, 1), (1, 2, 0), (1, 1, 1), (1, 0, 2),
    (0, 3, 0), (0, 2, 1), (0, 1, 2), (0, 0, 3),
]

_U64 = np.uint64


def splitmix64(seed, counters):
    """SplitMix64 output for the given uint64 counters (vectorised)."""
    with np.errstate(over="ignore"):
        z = _U64(seed) + (counters.astype(np.uint64) + _U64(1)) * _U64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> _U64(30))) * _U64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> _U64(27))) * _U64(0x94D049BB133111EB)
        return z ^ (z >> _U64(31))


def uniform(seed, counters):
    return (splitmix64(seed, counters) >> _U64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, counters):
    """Box-Muller standard normal from counters 2k and 2k+1."""
    c = counters.astype(np.uint64) * _U64(2)
    u1 = 1.0 - uniform(seed, c)
    u2 = uniform(seed, c + _U64(1))
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


@dataclass
class Level:
    name: str
    dx: float
    shape: tuple                 # slots per dimension
    neighbors: np.ndarray        # int32 [nslots, 27]
    kind: np.ndarray             # int32 [nslots]
    xbase: np.ndarray            # float64 [nslots, 3]
    targets: np.ndarray          # int32 [ntargets]
    M: np.ndarray                # float64 [nslots, 512, 20]
    com: np.ndarray              # float64 [nslots, 512, 3]
    mon: np.ndarray              # float64 [nslots, 512]
    info: dict = field(default_factory=dict)

    @property
    def nslots(self):
        return self.kind.shape[0]


def _lattice_moments(seed, shape_cells, dx, origin, density=None, noise_sigma=1.0):
    """Per-cell mass, centre of mass and 20-component multipole from 8 child
    point masses.  Child mass = (dx/2)^3 * rho(child) * LogNormal(0, sigma)."""
    nx, ny, nz = shape_cells
    ncell = nx * ny * nz
    gi = np.arange(ncell, dtype=np.int64)
    ix, iy, iz = gi // (ny * nz), (gi // nz) % ny, gi % nz
    centre = np.stack([origin[0] + (ix + 0.5) * dx, origin[1] + (iy + 0.5) * dx,
                       origin[2] + (iz + 0.5) * dx], axis=1)
    off = np.array([[(c >> 2) & 1, (c >> 1) & 1, c & 1] for c in range(8)], dtype=np.float64)
    off = (off - 0.5) * (0.5 * dx)                           # +- dx/4
    pos = centre[:, None, :] + off[None, :, :]               # [ncell, 8, 3]
    counters = (gi[:, None] * 8 + np.arange(8)[None, :]).astype(np.uint64)
    m = np.exp(noise_sigma * normal(seed, counters)) * (0.5 * dx) ** 3
    if density is not None:
        m = m * density(pos)
    mass = m.sum(axis=1)
    com = (m[:, :, None] * pos).sum(axis=1) / mass[:, None]
    rel = pos - com[:, None, :]
    M = np.empty((ncell, 20))
    M[:, 0] = mass
    for k, (a, b, c) in enumerate(_ALPHA):
        M[:, k + 1] = (m * rel[:, :, 0] ** a * rel[:, :, 1] ** b * rel[:, :, 2] ** c).sum(axis=1)
    return centre, mass, com, M


def _to_slots(arr, shape_slots, comps):
    """[NX*NY*NZ cells, comps] (x-major lattice) -> [nslots, 512, comps]."""
    sx, sy, sz = shape_slots
    a = arr.reshape(sx, 8, sy, 8, sz, 8, comps)
    a = a.transpose(0, 2, 4, 1, 3, 5, 6)
    return np.ascontiguousarray(a.reshape(sx * sy * sz, 512, comps))


def _neighbor_table(shape_slots, present=None):
    sx, sy, sz = shape_slots
    n = sx * sy * sz
    nb = np.full((n, 27), -1, dtype=np.int32)
    for s in range(n):
        a, b, c = s // (sy * sz), (s // sz) % sy, s % sz
        for raw in range(27):
            dxs, dys, dzs = _abi.dir_components(raw)
            aa, bb, cc = a + dxs, b + dys, c + dzs
            if 0 <= aa < sx and 0 <= bb < sy and 0 <= cc < sz:
                t = (aa * sy + bb) * sz + cc
                if present is None or present[t]:
                    nb[s, raw] = t
    return nb


def _slot_xbase(shape_slots, dx, origin):
    sx, sy, sz = shape_slots
    s = np.arange(sx * sy * sz)
    a, b, c = s // (sy * sz), (s // sz) % sy, s % sz
    return np.stack([origin[0] + (8 * a + 0.5) * dx, origin[1] + (8 * b + 0.5) * dx,
                     origin[2] + (8 * c + 0.5) * dx], axis=1)


def uniform_level(shape_slots, target_box, seed, dx=None, name="uniform"):
    """Non-leaf slots on a full block; targets = slots inside target_box
    ((lo, hi) per dimension, half open)."""
    sx, sy, sz = shape_slots
    if dx is None:
        dx = 1.0 / (8 * max(shape_slots))
    origin = (0.0, 0.0, 0.0)
    _, _, com, M = _lattice_moments(seed, (8 * sx, 8 * sy, 8 * sz), dx, origin)
    Ms = _to_slots(M, shape_slots, 20)
    coms = _to_slots(com, shape_slots, 3)
    nsl = sx * sy * sz
    kind = np.full(nsl, _abi.SLOT_NONLEAF, dtype=np.int32)
    mon = np.zeros((nsl, 512))
    tg = []
    for s in range(nsl):
        a, b, c = s // (sy * sz), (s // sz) % sy, s % sz
        if all(lo <= v < hi for v, (lo, hi) in zip((a, b, c), target_box)):
            tg.append(s)
    return Level(name=name, dx=dx, shape=tuple(shape_slots), neighbors=_neighbor_table(shape_slots),
                 kind=kind, xbase=_slot_xbase(shape_slots, dx, origin),
                 targets=np.array(tg, dtype=np.int32), M=Ms, com=coms, mon=mon,
                 info={"seed": seed})


def config1(seed=SEED_BASE + 1):
    """One 8^3 target sub-grid with its 26 non-leaf neighbours (24^3 frame,
    dx = 1/24)."""
    lv = uniform_level((3, 3, 3), ((1, 2), (1, 2), (1, 2)), seed, dx=1.0 / 24, name="single-subgrid")
    return lv


def config2(seed=SEED_BASE + 2):
    """512 non-leaf target sub-grids (8x8x8 block) inside a one-slot halo
    shell: 10^3 slots, 80^3 cells, dx = 1/80."""
    return uniform_level((10, 10, 10), ((1, 9),) * 3, seed, name="uniform-level-512")


def config4(seed=SEED_BASE + 4):
    """4096 non-leaf target sub-grids (16^3 block) inside a one-slot halo
    shell: 18^3 slots, 144^3 cells, dx = 1/144."""
    return uniform_level((18, 18, 18), ((1, 17),) * 3, seed, name="uniform-level-4096")


def mixed_level(shape_slots=(18, 18, 10), seed=SEED_BASE + 3, leaf_fraction=0.4, name="mixed-level"):
    """AMR-like level with leaf and non-leaf slots.  Density
    rho ~ (1 + r^2/0.1^2)^-1.5 about the box centre with 5% log-normal noise;
    the 40% of target slots with the largest noisy distance from the centre
    are leaves (monopoles at cell centres), and so is every halo slot beyond
    the same distance; the rest is non-leaf.  Targets are the slots inside a
    one-slot halo shell."""
    sx, sy, sz = shape_slots
    dx = 1.0 / (8 * max(shape_slots))
    ext = np.array([8 * sx * dx, 8 * sy * dx, 8 * sz * dx])
    ctr = ext / 2.0

    def density(pos):
        r2 = ((pos - ctr) ** 2).sum(axis=-1)
        return (1.0 + r2 / 0.01) ** -1.5

    centre, mass, com, M = _lattice_moments(seed, (8 * sx, 8 * sy, 8 * sz), dx, (0.0, 0.0, 0.0),
                                            density=density, noise_sigma=0.05)
    nsl = sx * sy * sz
    xb = _slot_xbase(shape_slots, dx, (0.0, 0.0, 0.0))
    slot_ctr = xb + 3.5 * dx
    rnoise = 1.0 + 0.05 * normal(seed + 7, np.arange(nsl, dtype=np.uint64))
    r = np.sqrt(((slot_ctr - ctr) ** 2).sum(axis=1)) * rnoise
    interior = np.array([1 <= s // (sy * sz) < sx - 1 and 1 <= (s // sz) % sy < sy - 1 and 1 <= s % sz < sz - 1
                         for s in range(nsl)])
    rt = np.sort(r[interior])[::-1]
    nleaf = int(round(leaf_fraction * rt.size))
    kind = np.full(nsl, _abi.SLOT_NONLEAF, dtype=np.int32)
    if nleaf > 0:
        kind[r >= rt[nleaf - 1]] = _abi.SLOT_LEAF
    Ms = _to_slots(M, shape_slots, 20)
    coms = _to_slots(com, shape_slots, 3)
    mons = _to_slots(mass[:, None], shape_slots, 1)[:, :, 0]
    cen = _to_slots(centre, shape_slots, 3)
    leaf = kind == _abi.SLOT_LEAF
    Ms[leaf] = 0.0
    coms[leaf] = cen[leaf]
    mons[~leaf] = 0.0
    nb = _neighbor_table(shape_slots)
    tg = []
    for s in range(nsl):
        a, b, c = s // (sy * sz), (s // sz) % sy, s % sz
        if 1 <= a < sx - 1 and 1 <= b < sy - 1 and 1 <= c < sz - 1:
            tg.append(s)
    tg = np.array(tg, dtype=np.int32)
    nbk = np.where(nb >= 0, kind[np.maximum(nb, 0)], -1)
    is_nonleaf = kind[tg] == _abi.SLOT_NONLEAF
    has_leaf_nb = (nbk[tg] == _abi.SLOT_LEAF).any(axis=1)
    info = {
        "seed": seed,
        "targets": int(tg.size),
        "leaf": int((~is_nonleaf).sum()),
        "nonleaf_interior": int((is_nonleaf & ~has_leaf_nb).sum()),
        "nonleaf_boundary": int((is_nonleaf & has_leaf_nb).sum()),
    }
    return Level(name=name, dx=dx, shape=tuple(shape_slots), neighbors=nb, kind=kind, xbase=xb,
                 targets=tg, M=Ms, com=coms, mon=mons, info=info)


def config3(seed=SEED_BASE + 3):
    """2048 target sub-grids (16x16x8) of mixed kind inside a halo shell."""
    return mixed_level((18, 18, 10), seed=seed, name="mixed-level-2048")


def subgrid_view(level, slot, root=False, type_=_abi.RHO):
    """octofmm_subgrid descriptor of one slot (keeps numpy views alive in
    the returned holder)."""
    sg = _abi.Subgrid()
    keep = []

    def arr(a):
        a = np.ascontiguousarray(a)
        keep.append(a)
        return a

    M = arr(level.M[slot])
    com = arr(level.com[slot])
    mon = arr(level.mon[slot])
    sg.M = _abi.dptr(M)
    sg.com = _abi.dptr(com)
    sg.mon = _abi.dptr(mon)
    sg.dx = level.dx
    for k in range(3):
        sg.xbase[k] = level.xbase[slot, k]
    sg.type = type_
    sg.use_root_stencil = 1 if root else 0
    for f in range(26):
        raw = _abi.flat_to_raw(f)
        n = int(level.neighbors[slot, raw])
        nbd = sg.neighbors[f]
        if n < 0 or root:
            nbd.kind = _abi.NEIGHBOR_EMPTY
            continue
        nbd.count = 512
        if level.kind[n] == _abi.SLOT_LEAF:
            nbd.kind = _abi.NEIGHBOR_MONOPOLE
            nbd.mon = _abi.dptr(arr(level.mon[n]))
        else:
            nbd.kind = _abi.NEIGHBOR_MULTIPOLE
            nbd.M = _abi.dptr(arr(level.M[n]))
            nbd.com = _abi.dptr(arr(level.com[n]))
    return sg, keep


# ---------------------------------------------------------------------------
# Adaptive trees for the full gravity solve
# ---------------------------------------------------------------------------
def lane_emden(n=1.5, h=1.0e-3):
    """Lane-Emden solution theta(xi) of index n by classical RK4 from the
    series start theta = 1 - xi^2/6 + n xi^4/120; returns (xi, theta, xi_1)
    with xi_1 the first zero (3.65375 for n = 1.5)."""
    def f(xi, y):
        th, dth = y
        return np.array([dth, -(max(th, 0.0) ** n) - 2.0 * dth / xi])

    xi = 1.0e-4
    y = np.array([1.0 - xi * xi / 6.0 + n * xi ** 4 / 120.0, -xi / 3.0 + n * xi ** 3 / 30.0])
    xs, ts = [0.0, xi], [1.0, y[0]]
    while y[0] > 0.0:
        k1 = f(xi, y)
        k2 = f(xi + 0.5 * h, y + 0.5 * h * k1)
        k3 = f(xi + 0.5 * h, y + 0.5 * h * k2)
        k4 = f(xi + h, y + h * k3)
        y_new = y + (h / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)
        if y_new[0] <= 0.0:
            xi1 = xi + h * y[0] / (y[0] - y_new[0])
            xs.append(xi1)
            ts.append(0.0)
            break
        xi += h
        y = y_new
        xs.append(xi)
        ts.append(y[0])
    return np.array(xs), np.array(ts), xs[-1]


def polytrope_density(r, radius=0.3, n=1.5, rho_c=1.0, floor=1.0e-10):
    xi_t, th_t, xi1 = _lane_emden_table(n)
    xi = np.asarray(r, dtype=np.float64) * (xi1 / radius)
    th = np.interp(xi, xi_t, th_t, right=0.0)
    return np.maximum(rho_c * np.power(np.maximum(th, 0.0), n), floor * rho_c)


_LE_CACHE = {}


def _lane_emden_table(n):
    if n not in _LE_CACHE:
        _LE_CACHE[n] = lane_emden(n)
    return _LE_CACHE[n]


def _assign_leaf_masses(tree, density, seed, noise_sigma):
    """Leaf cell masses rho(centre) dx^3 x LogNormal(0, noise_sigma); the
    noise counter is the cell's global index in level-major, Morton node order."""
    base = 0
    for lv in tree.levels:
        cen = tree.cell_centres(lv)
        mon = np.zeros((lv.nnodes, 512))
        leaf = lv.kind == _abi.SLOT_LEAF
        if leaf.any():
            rho = density(cen[leaf])
            cnt = base + np.arange(lv.nnodes * 512, dtype=np.uint64).reshape(lv.nnodes, 512)[leaf]
            noise = np.exp(noise_sigma * normal(seed, cnt)) if noise_sigma > 0 else 1.0
            mon[leaf] = rho * lv.dx ** 3 * noise
        lv.mon = np.ascontiguousarray(mon)
        base += lv.nnodes * 512
    return tree


def polytrope_tree(max_level=5, mass_threshold=2.0e-6, radius=0.3, n=1.5, seed=SEED_BASE + 5,
                   noise_sigma=0.01, name="polytrope-tree"):
    """Adaptive tree over the unit box [-1/2, 1/2]^3 for an index-n polytrope
    of the given radius (centre at the origin): a node is refined where one of
    its cells holds more mass than mass_threshold (rho(centre) dx^3), up to
    max_level, then properly nested.  Leaf cells carry rho dx^3 with seeded
    LogNormal(0, noise_sigma) noise.  The star's rotation changes velocities
    only, so the gravity state does not depend on it."""
    from . import tree as _tree

    dens = lambda x: polytrope_density(np.sqrt((x * x).sum(axis=-1)), radius=radius, n=n)

    def refine(lev, cs, xb, dxl):
        i = np.arange(8, dtype=np.float64) * dxl
        g = np.stack(np.meshgrid(i, i, i, indexing="ij"), axis=-1).reshape(512, 3)
        cen = xb[:, None, :] + g[None]
        return (dens(cen) * dxl ** 3).max(axis=1) > mass_threshold

    t = _tree.build_tree(refine, max_level, name=name)
    _assign_leaf_masses(t, dens, seed, noise_sigma)
    t.info = {"seed": seed, "max_level": max_level, "mass_threshold": mass_threshold, "radius": radius,
              "polytrope_index": n, "noise_sigma": noise_sigma,
              "nodes_per_level": [lv.nnodes for lv in t.levels],
              "leaves_per_level": [lv.nleaves for lv in t.levels], "nodes": t.nnodes}
    return t


def config5(seed=SEED_BASE + 5):
    """Full adaptive-tree solve: n = 1.5 polytrope of radius 0.3 refined on
    cell mass > 2.5e-8 up to level 6 -- 10,697 sub-grids on levels 0..6
    (levels 0..2 refine fully, leaves sit on levels 3..6)."""
    return polytrope_tree(max_level=6, mass_threshold=2.5e-8, seed=seed, name="polytrope-tree-c5")


def small_tree(seed=SEED_BASE + 50, max_level=2, name="small-tree"):
    """A few dozen nodes over 3 levels with leaves at every level: the root is
    refined, 2 of its children are refined (so level 2 has 16 nodes, leaves and
    refined-neighbour boundaries at level 1).  Masses log-normal (sigma 1)."""
    from . import tree as _tree

    def refine(lev, cs, xb, dxl):
        if lev == 0:
            return np.ones(len(cs), bool)
        key = cs[:, 0] * 4 + cs[:, 1] * 2 + cs[:, 2]
        return (key == 0) | (key == 1) if lev == 1 else np.zeros(len(cs), bool)

    t = _tree.build_tree(refine, max_level, name=name)
    _assign_leaf_masses(t, lambda x: np.ones(x.shape[:-1]), seed, 1.0)
    t.info = {"seed": seed, "nodes_per_level": [lv.nnodes for lv in t.levels],
              "leaves_per_level": [lv.nleaves for lv in t.levels], "nodes": t.nnodes}
    return t


# ---------------------------------------------------------------------------
# Octo-Tiger's solid-sphere regression problem (test_problems/sphere):
# problem=solid_sphere, max_level=3, hydro=off, stop_step=1
# ---------------------------------------------------------------------------
SPHERE_CENTRE = (0.25, 0.0, 0.0)     # solid_sphere_{x,y,z}center defaults (src/options_processing.cpp:80-82)
SPHERE_RADIUS = 1.0 / 3.0            # solid_sphere_radius (:83)
SPHERE_MASS = 1.0                    # solid_sphere_mass (:84)
SPHERE_RHO_MIN = 1.0e-12             # solid_sphere_rho_min (:85)
SPHERE_SUBSAMPLES = 25               # N of solid_sphere (src/problem.cpp:404)
SPHERE_REFINE_R2 = 1.0 / 3.0         # ssr0 of refine_test_center (src/problem.cpp:30)
SPHERE_POT_NORMS = (4.470774e-01, 3.176879e+00)   # test_problems/sphere/CMakeLists.txt:18 (INX = 8)
INX_SPHERE = 8


def solid_sphere_density(X, dx):
    """Cell densities of solid_sphere (src/problem.cpp:403-452) for cell
    centres X[..., 3] of edge dx, with the same floating-point steps: cells
    wholly inside get M / V, cells the surface cuts count their N^3
    sub-cell centres strictly inside (adding drho once per centre), every
    cell is floored at rho_min."""
    X = np.asarray(X, dtype=np.float64)
    shape = X.shape[:-1]
    X = X.reshape(-1, 3)
    N = SPHERE_SUBSAMPLES
    r0 = SPHERE_RADIUS
    V = 4.0 / 3.0 * np.pi * r0 * r0 * r0
    drho = SPHERE_MASS / float(N * N * N) / V
    acc = np.concatenate([[0.0], np.cumsum(np.full(N * N * N, drho))])   # drho added k times, in order
    full = drho * N * N * N
    h = dx / 2.0
    rel = X - np.array(SPHERE_CENTRE)
    lo, hi = rel - h, rel + h   # x0 - dx/2, x0 + dx/2
    vmax = np.maximum(np.abs(hi), np.abs(lo))
    vmin = np.where(hi * lo < 0.0, 0.0, np.where(hi > 0.0, np.minimum(hi, lo), np.maximum(hi, lo)))
    r2 = r0 * r0
    inside = (vmax[:, 0] * vmax[:, 0] + vmax[:, 1] * vmax[:, 1] + vmax[:, 2] * vmax[:, 2]) <= r2
    cut = ~inside & ((vmin[:, 0] * vmin[:, 0] + vmin[:, 1] * vmin[:, 1] + vmin[:, 2] * vmin[:, 2]) <= r2)
    rho = np.zeros(len(X))
    rho[inside] = full
    d = dx / N
    k = np.arange(N, dtype=np.float64)
    for c in np.nonzero(cut)[0]:
        o = (rel[c] - h) + d / 2.0
        sx = (o[0] + k * d) ** 2
        sy = (o[1] + k * d) ** 2
        sz = (o[2] + k * d) ** 2
        cnt = int(((sx[:, None, None] + sy[None, :, None]) + sz[None, None, :] < r2).sum())
        rho[c] = acc[cnt]
    return np.maximum(rho, SPHERE_RHO_MIN).reshape(shape)


def solid_sphere_tree(max_level=3, passes=None, name="solid-sphere"):
    """The octree and leaf masses of Octo-Tiger's solid-sphere regression
    run.  Domain [-1, 1]^3 (xscale = 1), 8^3 cells per node.  Refinement is
    refine_test_center (src/problem.cpp:227-233): a node below max_level
    refines when a cell it scans lies within r^2 < 1/3 of the origin, the
    root always (min_level = 1, grid::refine_me, src/grid.cpp:1493-1548,
    which scans the interior plus a REFINE_BW = 2 face band, skipping band
    cells on edges and corners).  The tree is what max_level + 1 regrid
    passes leave (tree.regrid_tree; passes overrides the count).  Leaf cells carry the solid_sphere
    density at their centres and mass rho dx^3 (grid::compute_multipoles,
    src/grid_fmm.cpp:1643, 1774)."""
    from . import tree as _tree

    rel = np.arange(-2, INX_SPHERE + 2)
    R = np.stack(np.meshgrid(rel, rel, rel, indexing="ij"), axis=-1).reshape(-1, 3)
    R = R[((R < 0) | (R >= INX_SPHERE)).sum(axis=1) <= 1].astype(np.float64)

    def flag(lev, cs):
        if lev < 1:
            return np.ones(len(cs), bool)
        if lev >= max_level:
            return np.zeros(len(cs), bool)
        dx = 2.0 / float(INX_SPHERE << lev)
        xmin = cs.astype(np.float64) * (2.0 / float(1 << lev)) - 1.0
        X = (R[None, :, :] + 0.5) * dx + xmin[:, None, :]
        r2 = X[..., 0] * X[..., 0] + X[..., 1] * X[..., 1] + X[..., 2] * X[..., 2]
        return (r2 < SPHERE_REFINE_R2).any(axis=1)

    npass = max_level + 1 if passes is None else passes
    t = _tree.regrid_tree(flag, max_level, npass, lo=-1.0, size=2.0, name=name)
    for lv in t.levels:
        leaf = lv.kind == _abi.SLOT_LEAF
        rho = np.zeros((lv.nnodes, 512))
        if leaf.any():
            i = np.arange(INX_SPHERE, dtype=np.float64)
            g = np.stack(np.meshgrid(i, i, i, indexing="ij"), axis=-1).reshape(512, 3)
            xmin = lv.coords[leaf].astype(np.float64) * (2.0 / float(1 << lv.level)) - 1.0
            rho[leaf] = solid_sphere_density((g[None] + 0.5) * lv.dx + xmin[:, None, :], lv.dx)
        lv.rho = np.ascontiguousarray(rho)
        lv.mon = np.ascontiguousarray(rho * (lv.dx * lv.dx * lv.dx))
    t.info = {"problem": "solid_sphere", "max_level": max_level,
              "nodes_per_level": [lv.nnodes for lv in t.levels],
              "leaves_per_level": [lv.nleaves for lv in t.levels], "nodes": t.nnodes}
    return t


def sphere_pot_norms(tree, phi_levels):
    """The 'pot' line of node_server::compare_analytic
    (src/node_server_actions_2.cpp:154-182): the analytic pot field of the
    solid sphere is zero (solid_sphere never sets it), so grid::compute_analytic
    (src/grid.cpp:1796-1800) sums |phi rho| dV and (phi rho)^2 dV over the
    interior cells of the leaves, with pot = phi rho (src/grid_fmm.cpp:1628);
    the root divides by the domain volume (2 xscale)^3 = 8 and takes the root
    of the L2 sum.  phi_levels[l] is [n, 512] (leaf rows are used)."""
    l1 = l2 = 0.0
    for lv, phi in zip(tree.levels, phi_levels):
        leaf = lv.kind == _abi.SLOT_LEAF
        if not leaf.any():
            continue
        dv = lv.dx * lv.dx * lv.dx
        pot = np.asarray(phi)[leaf] * lv.rho[leaf]
        l1 += float(np.abs(pot).sum()) * dv
        l2 += float((pot * pot).sum()) * dv
    vol = 8.0
    return l1 / vol, float(np.sqrt(l2 / vol))
