"""Minimal adaptive octree of 8^3 sub-grids for a full gravity solve.

Host-side mirror of the part of Octo-Tiger's node hierarchy that the
gravity solver walks (node_server::compute_fmm, src/node_server.cpp:589-798):
every node is an 8^3 sub-grid, a refined node has all 8 children, and the
refinement is properly nested -- the 26 same-level neighbours of a refined node
exist unless they fall outside the domain -- so that every interaction a fine
cell misses at its own level has been taken by its parent level (near pairs
with monopole partners, p2m at coarse/fine boundaries).

Nodes of a level are stored in Morton order of their integer coordinates, so
neighbouring sub-grids get nearby slot indices.  Per level the tree provides
exactly what the level interaction plans and the M2M / L2L passes take:
neighbour table [n][27] (raw direction order, entry 13 = self, -1 = empty),
slot kind, parent index, octant in the parent (bit 0 x, bit 1 y, bit 2 z, as
geo::octant, octotiger/geometry.hpp:333-335), centre of cell (0,0,0), dx.
"""
from dataclasses import dataclass, field

import numpy as np

from . import _abi

INX = 8


def _morton_key(c):
    """Interleaved bits of integer coordinates (x highest within each triple)."""
    c = np.asarray(c, dtype=np.int64)
    key = np.zeros(len(c), dtype=np.int64)
    for b in range(21):
        for d in range(3):
            key |= ((c[:, d] >> b) & 1) << (3 * b + (2 - d))
    return key


@dataclass
class TreeLevel:
    level: int
    dx: float
    coords: np.ndarray      # [n, 3] int64 node coordinates at this level
    neighbors: np.ndarray   # [n, 27] int32
    kind: np.ndarray        # [n] int32, _abi.SLOT_NONLEAF / SLOT_LEAF
    parent: np.ndarray      # [n] int32, -1 at level 0
    octant: np.ndarray      # [n] int32
    xbase: np.ndarray       # [n, 3] centre of cell (0,0,0)
    mon: np.ndarray = None  # [n, 512] cell masses of leaf nodes (zeros elsewhere)
    rho: np.ndarray = None  # [n, 512] cell densities of leaf nodes, where a workload defines them

    @property
    def nnodes(self):
        return int(self.coords.shape[0])

    @property
    def nleaves(self):
        return int((self.kind == _abi.SLOT_LEAF).sum())


@dataclass
class Tree:
    name: str
    levels: list
    lo: float = -0.5
    size: float = 1.0
    info: dict = field(default_factory=dict)

    @property
    def nnodes(self):
        return sum(lv.nnodes for lv in self.levels)

    def cell_centres(self, lv):
        """[n, 512, 3] geometric cell centres of a level (x*64 + y*8 + z order)."""
        i = np.arange(INX, dtype=np.float64)
        g = np.stack(np.meshgrid(i, i, i, indexing="ij"), axis=-1).reshape(512, 3)
        return lv.xbase[:, None, :] + g[None, :, :] * lv.dx


def build_tree(refine, max_level, lo=-0.5, size=1.0, name="tree"):
    """Builds a properly nested octree.

    refine(level, coords[n,3], xbase[n,3], dx) -> bool[n] says which existing
    nodes of a level the caller wants refined (only levels < max_level are
    asked).  Refinements that nesting needs are then added at coarser levels:
    a refined node's 26 neighbours must exist, so their parents are refined
    too, which may create further leaves but never removes a refinement.
    """
    # desired refinement, top-down
    refined = [set() for _ in range(max_level + 1)]
    exist = [{(0, 0, 0)}]
    for lev in range(max_level):
        cs = np.array(sorted(exist[lev]), dtype=np.int64).reshape(-1, 3)
        dxl = size / (INX * 2 ** lev)
        xb = lo + cs * (INX * dxl) + 0.5 * dxl
        want = np.asarray(refine(lev, cs, xb, dxl), dtype=bool) if len(cs) else np.zeros(0, bool)
        refined[lev] = {tuple(int(v) for v in c) for c, w in zip(cs, want) if w}
        exist.append(_children(refined[lev]))
    # nesting, bottom-up: neighbours of refined nodes must exist
    for lev in range(max_level - 1, 0, -1):
        n_side = 2 ** lev
        need = set()
        for c in refined[lev]:
            for dz in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dx_ in (-1, 0, 1):
                        q = (c[0] + dx_, c[1] + dy, c[2] + dz)
                        if min(q) < 0 or max(q) >= n_side:
                            continue
                        need.add((q[0] >> 1, q[1] >> 1, q[2] >> 1))
            need.add((c[0] >> 1, c[1] >> 1, c[2] >> 1))
        refined[lev - 1] |= need
    # existence, top-down from the final refinement sets
    exist = [{(0, 0, 0)}]
    for lev in range(max_level):
        refined[lev] &= exist[lev]
        exist.append(_children(refined[lev]))
    return _assemble(exist, refined, lo, size, name)


def _children(nodes):
    return {(2 * c[0] + (o & 1), 2 * c[1] + ((o >> 1) & 1), 2 * c[2] + ((o >> 2) & 1)) for c in nodes for o in range(8)}


def regrid_tree(flag, max_level, passes, lo=-1.0, size=2.0, name="tree"):
    """The tree Octo-Tiger's start-up regrid passes leave behind.

    Each pass (node_server::regrid, src/node_server_actions_1.cpp:244-279)
    asks every existing node whether it wants refinement
    (check_for_refinement, src/node_server_actions_2.cpp:36-73):
    flag(level, coords[n,3]) -> bool[n].  A flagged node forces its 26
    in-domain same-level neighbours to exist, i.e. flags their parents, and
    so on up the tree (force_nodes_to_exist, :387-451).  The flagged set is
    the new refinement: refined nodes without a flag lose their children,
    unrefined flagged nodes get 8 new leaf children (regrid_gather,
    src/node_server_actions_1.cpp:68-124), which the next pass examines.
    The frontend runs max_level + 1 such passes before the first solve
    (frontend/frontend-helper.cpp:190-200).
    """
    refined = [set() for _ in range(max_level + 1)]
    for _ in range(passes):
        exist = [{(0, 0, 0)}]
        for lev in range(max_level):
            exist.append(_children(refined[lev]))
        flagged = [set() for _ in range(max_level + 1)]
        for lev in range(max_level + 1):
            cs = sorted(exist[lev])
            if cs:
                want = np.asarray(flag(lev, np.array(cs, dtype=np.int64)), dtype=bool)
                flagged[lev] = {c for c, w in zip(cs, want) if w}
        for lev in range(max_level, 0, -1):
            n_side = 1 << lev
            for c in flagged[lev]:
                for dz in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dx_ in (-1, 0, 1):
                            q = (c[0] + dx_, c[1] + dy, c[2] + dz)
                            if min(q) < 0 or max(q) >= n_side:
                                continue
                            p = (q[0] >> 1, q[1] >> 1, q[2] >> 1)
                            if p not in exist[lev - 1]:
                                raise ValueError("regrid pass left a refined node without neighbours")
                            flagged[lev - 1].add(p)
        refined = flagged
    exist = [{(0, 0, 0)}]
    for lev in range(max_level):
        exist.append(_children(refined[lev]))
    return _assemble(exist, refined, lo, size, name)


def _assemble(exist, refined, lo, size, name):
    while len(exist) > 1 and not exist[-1]:
        exist.pop()
    levels = []
    prev_index = None
    for lev, nodes in enumerate(exist):
        cs = np.array(sorted(nodes), dtype=np.int64).reshape(-1, 3)
        cs = cs[np.argsort(_morton_key(cs), kind="stable")]
        index = {tuple(int(v) for v in c): i for i, c in enumerate(cs)}
        n = len(cs)
        nbr = np.full((n, 27), -1, dtype=np.int32)
        for i, c in enumerate(cs):
            for raw in range(27):
                d = _abi.dir_components(raw)
                q = (int(c[0]) + d[0], int(c[1]) + d[1], int(c[2]) + d[2])
                nbr[i, raw] = index.get(q, -1)
        kind = np.array([_abi.SLOT_NONLEAF if (lev < len(refined) and tuple(int(v) for v in c) in refined[lev])
                         else _abi.SLOT_LEAF for c in cs], dtype=np.int32)
        if lev == 0:
            parent = np.full(n, -1, dtype=np.int32)
        else:
            parent = np.array([prev_index[(int(c[0]) >> 1, int(c[1]) >> 1, int(c[2]) >> 1)] for c in cs],
                              dtype=np.int32)
        octant = ((cs[:, 0] & 1) | ((cs[:, 1] & 1) << 1) | ((cs[:, 2] & 1) << 2)).astype(np.int32)
        dxl = size / (INX * 2 ** lev)
        xbase = lo + cs.astype(np.float64) * (INX * dxl) + 0.5 * dxl
        levels.append(TreeLevel(level=lev, dx=dxl, coords=cs, neighbors=nbr, kind=kind, parent=parent,
                                octant=octant, xbase=np.ascontiguousarray(xbase)))
        prev_index = index
    return Tree(name=name, levels=levels, lo=lo, size=size)


def check_tree(tree):
    """Structural invariants the solver relies on; raises ValueError."""
    for lev, lv in enumerate(tree.levels):
        if lev == 0 and lv.nnodes != 1:
            raise ValueError("level 0 must hold the root only")
        if np.any(lv.neighbors[:, 13] != np.arange(lv.nnodes)):
            raise ValueError(f"level {lev}: neighbour entry 13 must be the node itself")
        if lev > 0:
            up = tree.levels[lev - 1]
            if np.any(up.kind[lv.parent] != _abi.SLOT_NONLEAF):
                raise ValueError(f"level {lev}: parent is a leaf")
            pairs = lv.parent.astype(np.int64) * 8 + lv.octant
            if len(np.unique(pairs)) != lv.nnodes:
                raise ValueError(f"level {lev}: two nodes in one octant")
            if np.bincount(lv.parent, minlength=up.nnodes)[up.kind == _abi.SLOT_NONLEAF].min(initial=8) != 8:
                raise ValueError(f"level {lev}: a refined node lacks children")
        for i in np.nonzero(lv.kind == _abi.SLOT_NONLEAF)[0]:
            c = lv.coords[i]
            for raw in range(27):
                d = _abi.dir_components(raw)
                q = c + np.array(d)
                if q.min() < 0 or q.max() >= 2 ** lev:
                    continue
                if lv.neighbors[i, raw] < 0:
                    raise ValueError(f"level {lev}: refined node {i} lacks neighbour {raw}")
    return True
