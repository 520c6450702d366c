"""Sub-grid test cases shared by the golden-fixture generator and the tests.

Every case is rebuilt from its seed, so fixtures only hold outputs (plus a
digest of the inputs that guards against generator drift).
"""
import hashlib

import numpy as np

from octotiger_amd import _abi, workloads

# (name, kind of target, solve types)
CASES = [
    ("c1_nonroot", "multipole"),
    ("c1_root", "root"),
    ("mixed_halo", "multipole"),
    ("partial_halo", "multipole"),
    ("leaf_p2p", "monopole"),
    ("leaf_p2m", "monopole"),
]

SEED = workloads.SEED_BASE + 100


def _leafify(level, slots):
    """Turns slots into leaves: monopole = cell mass, centre = geometric."""
    for s in slots:
        level.mon[s] = level.M[s, :, 0]
        level.M[s] = 0.0
        idx = np.arange(512)
        x, y, z = idx // 64, (idx // 8) % 8, idx % 8
        level.com[s] = level.xbase[s] + np.stack([x, y, z], axis=1) * level.dx
        level.kind[s] = _abi.SLOT_LEAF


def _drop(level, slot, raws):
    for r in raws:
        level.neighbors[slot, r] = -1


FACES = [_abi.raw_index(*d) for d in [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]]


def build_level(name):
    """The 27-slot level behind a case (target slot 13)."""
    if name in ("c1_nonroot", "c1_root"):
        return workloads.config1()
    lv = workloads.uniform_level((3, 3, 3), ((1, 2),) * 3, SEED + hash_name(name), dx=1.0 / 24, name=name)
    if name == "mixed_halo":
        _leafify(lv, FACES)                     # six leaf (monopole) face neighbours
        _drop(lv, 13, [0, 2, 26])               # three empty corner directions
    elif name == "partial_halo":
        _leafify(lv, [_abi.raw_index(0, 0, 1)])
    elif name == "leaf_p2p":
        _leafify(lv, list(range(27)))
        _drop(lv, 13, [_abi.raw_index(1, 1, 1)])
    elif name == "leaf_p2m":
        leaves = [s for s in range(27) if s not in (_abi.raw_index(1, 0, 0), _abi.raw_index(0, -1, 0))]
        _leafify(lv, leaves)                    # two refined face neighbours (+x, -y)
        _drop(lv, 13, [_abi.raw_index(-1, -1, -1)])
    return lv


def hash_name(name):
    return int(hashlib.sha256(name.encode()).hexdigest()[:6], 16)


def subgrid(name, type_=_abi.RHO):
    """(octofmm_subgrid, keep-alive list, target kind, contains_multipole_neighbor)."""
    lv = build_level(name)
    kind = dict(CASES)[name]
    sg, keep = workloads.subgrid_view(lv, 13, root=(kind == "root"), type_=type_)
    cmp_ = any(lv.neighbors[13, r] >= 0 and lv.kind[lv.neighbors[13, r]] == _abi.SLOT_NONLEAF
               for r in range(27) if r != 13)
    if name == "partial_halo":
        rng = np.random.default_rng(hash_name(name))
        # a refined neighbour handed over partially, cells in shuffled order
        f = _abi.raw_to_flat(_abi.raw_index(1, 0, 0))
        cells = rng.permutation(512)[:300].astype(np.int32)
        n = int(lv.neighbors[13, _abi.raw_index(1, 0, 0)])
        M = np.ascontiguousarray(lv.M[n][cells])
        x = np.ascontiguousarray(lv.com[n][cells])
        keep += [cells, M, x]
        nb = sg.neighbors[f]
        nb.count, nb.M, nb.com, nb.cells = cells.size, _abi.dptr(M), _abi.dptr(x), _abi.iptr(cells)
        # a leaf neighbour handed over partially
        f2 = _abi.raw_to_flat(_abi.raw_index(0, 0, 1))
        n2 = int(lv.neighbors[13, _abi.raw_index(0, 0, 1)])
        cells2 = rng.permutation(512)[:200].astype(np.int32)
        m2 = np.ascontiguousarray(lv.mon[n2][cells2])
        keep += [cells2, m2]
        nb2 = sg.neighbors[f2]
        nb2.count, nb2.mon, nb2.cells = cells2.size, _abi.dptr(m2), _abi.iptr(cells2)
    keep.append(lv)
    return sg, keep, kind, cmp_


def input_digest(name):
    lv = build_level(name)
    h = hashlib.sha256()
    for a in (lv.M, lv.com, lv.mon, lv.neighbors, lv.kind, lv.xbase):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def partial_subgrid(level, spec, type_=_abi.RHO, slot=13, stale=None):
    """Sub-grid descriptor of a slot in which the neighbours named in spec
    (raw direction -> int32 cell list, in the order the handed-over arrays
    hold them; repeats allowed) hand over only those cells, as a remote
    neighbour does with grid::get_ilist_n_bnd.  stale: raw direction -> the
    number of leading entries whose data are scaled by 3, so that only a
    staging in which later entries overwrite earlier ones gives the level's
    values
    (octotiger/multipole_interactions/legacy/multipole_interaction_interface.hpp:177,257,
    octotiger/monopole_interactions/legacy/p2m_interaction_interface.hpp:142,239)."""
    sg, keep = workloads.subgrid_view(level, slot, type_=type_)
    for raw, cells in spec.items():
        n = int(level.neighbors[slot, raw])
        assert n >= 0, "partial halo of an empty direction"
        cells = np.ascontiguousarray(cells, dtype=np.int32)
        nb = sg.neighbors[_abi.raw_to_flat(raw)]
        nb.count, nb.cells = int(cells.size), _abi.iptr(cells)
        keep.append(cells)
        k = (stale or {}).get(raw, 0)
        if level.kind[n] == _abi.SLOT_LEAF:
            m = np.ascontiguousarray(level.mon[n][cells])
            m[:k] *= 3.0
            keep.append(m)
            nb.mon = _abi.dptr(m)
        else:
            M = np.ascontiguousarray(level.M[n][cells])
            x = np.ascontiguousarray(level.com[n][cells])
            M[:k] *= 3.0
            x[:k] *= 3.0
            keep += [M, x]
            nb.M, nb.com = _abi.dptr(M), _abi.dptr(x)
    keep.append(level)
    return sg, keep


def zeroed_level(level, spec, slot=13):
    """A copy of the level in which every neighbour named in spec holds zeros
    in the cells its list leaves out (multipoles and centres of mass of
    refined neighbours, monopoles of leaf neighbours): what the reference's
    staging builds from a partial halo."""
    import copy
    lv = copy.deepcopy(level)
    for raw, cells in spec.items():
        n = int(lv.neighbors[slot, raw])
        absent = np.ones(512, bool)
        absent[np.asarray(cells)] = False
        if lv.kind[n] == _abi.SLOT_LEAF:
            lv.mon[n][absent] = 0.0
        else:
            lv.M[n][absent] = 0.0
            lv.com[n][absent] = 0.0
    return lv


def partial_specs(name):
    """Partial-halo patterns over a case's level: every cell of two neighbours
    in shuffled order; subsets of them; the subsets preceded by stale repeats
    of some of their cells (pass stale=stale_prefix(spec) to partial_subgrid)."""
    lv = build_level(name)
    rng = np.random.default_rng(hash_name(name) + 7)
    refined = [r for r in range(27) if r != 13 and lv.neighbors[13, r] >= 0
               and lv.kind[lv.neighbors[13, r]] == _abi.SLOT_NONLEAF]
    leaves = [r for r in range(27) if r != 13 and lv.neighbors[13, r] >= 0
              and lv.kind[lv.neighbors[13, r]] == _abi.SLOT_LEAF]
    pick = (refined[:1] + leaves[:1]) if leaves and refined else (refined or leaves)[:2]
    full = {r: rng.permutation(512).astype(np.int32) for r in pick}
    sub = {r: rng.permutation(512)[: 150 + 100 * i].astype(np.int32) for i, r in enumerate(pick)}
    rep = {r: np.concatenate([c[::3], c]).astype(np.int32) for r, c in sub.items()}
    return {"all_shuffled": full, "subset": sub, "repeats": rep}


def stale_prefix(spec_name, spec):
    """Leading stale entries of the 'repeats' pattern (none for the others)."""
    if spec_name != "repeats":
        return None
    return {r: int(c.size - (c.size * 3) // 4) for r, c in spec.items()}
