"""Level plans with partial halo slots (octofmm_partial_slot): halo slots
handed over as compact cell lists in grid::get_ilist_n_bnd order
(multipole_interaction_interface.hpp:177,257; p2m_interaction_interface.hpp:142,239).

What the mathematics fixes: a partial slot stages exactly what a full slot
with zeros in the cells the list leaves out stages, a cell listed twice taking
its later entry.  So the plan with partial slots must give, bit for bit, the
result of the same plan over that zeroed level, and the oracle's result on it
to 1e-12 normwise.  The lists carry stale leading repeats (data scaled by 3),
and the slots' own device rows start out holding the full data, so a plan
that staged the wrong entry or left a row unwritten fails.
"""
import copy

import numpy as np
import pytest

import oracle
from metrics import normwise
from octotiger_amd import _abi, hip, native, workloads

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(theta=0.5, device=0)
    yield c
    c.close()


def partial_lists(lv, seed):
    """Halo slots next to targets, both kinds, with shuffled subsets of their
    cells preceded by stale repeats; and the level with the left-out cells
    zeroed."""
    rng = np.random.default_rng(seed)
    tg = set(int(t) for t in lv.targets)
    halo = sorted({int(n) for t in tg for n in lv.neighbors[t] if n >= 0 and int(n) not in tg})
    leaf = [s for s in halo if lv.kind[s] == _abi.SLOT_LEAF]
    nonleaf = [s for s in halo if lv.kind[s] == _abi.SLOT_NONLEAF]
    pick = nonleaf[:4] + leaf[:4]
    assert nonleaf and leaf
    zeroed = copy.deepcopy(lv)
    lists = []
    for i, s in enumerate(pick):
        if i == 0:
            keep, stale = rng.permutation(512), np.zeros(0, np.int64)   # every cell, shuffled
        else:
            keep = rng.permutation(512)[: 64 + 45 * i]
            stale = keep[::4]
        cells = np.concatenate([stale, keep]).astype(np.int32)
        k = stale.size
        absent = np.ones(512, bool)
        absent[keep] = False
        if lv.kind[s] == _abi.SLOT_NONLEAF:
            M = np.ascontiguousarray(lv.M[s][cells])
            com = np.ascontiguousarray(lv.com[s][cells])
            M[:k] *= 3.0
            com[:k] *= 3.0
            lists.append({"slot": s, "cells": cells, "M": M, "com": com})
            zeroed.M[s][absent] = 0.0
            zeroed.com[s][absent] = 0.0
        else:
            mon = np.ascontiguousarray(lv.mon[s][cells])
            mon[:k] *= 3.0
            lists.append({"slot": s, "cells": cells, "mon": mon})
            zeroed.mon[s][absent] = 0.0
    return lists, zeroed


def run(ctx, lv, type_, partial=None, reps=1):
    dl = hip.LevelOnDevice(lv)
    keep = []
    spec = None
    if partial:
        spec = []
        for e in partial:
            d = {"slot": e["slot"], "cells": e["cells"]}
            for k in ("M", "com", "mon"):
                if k in e:
                    a = hip.DeviceArray.from_numpy(e[k])
                    keep.append(a)
                    d[k] = a.addr
            spec.append(d)
    plan = dl.plan(ctx, type_=type_, flags=_abi.PLAN_USE_GRAPH | _abi.PLAN_AUTOTUNE, partial=spec)
    s = hip.Stream()
    outs = []
    for _ in range(reps):
        dl.zero_outputs()
        plan.execute(s.handle)
        s.synchronize()
        outs.append((dl.L.numpy(), dl.L_c.numpy()))
    rows = (dl.M.numpy(), dl.com.numpy(), dl.mon.numpy())
    plan.close()
    s.destroy()
    dl.free()
    for a in keep:
        a.free()
    return outs, rows


def halo_level():
    """A 4^3-slot uniform level (8 non-leaf targets inside a one-slot halo
    shell) whose halo slots alternate between refined and leaf, so targets
    see both kinds of neighbour (leaf slots: monopoles at cell centres)."""
    lv = workloads.uniform_level((4, 4, 4), ((1, 3),) * 3, seed=91)
    tg = set(int(t) for t in lv.targets)
    idx = np.arange(512)
    geo = np.stack([idx // 64, (idx // 8) % 8, idx % 8], axis=1)
    for s in range(lv.nslots):
        if s in tg or s % 2:
            continue
        lv.mon[s] = lv.M[s, :, 0]
        lv.M[s] = 0.0
        lv.com[s] = lv.xbase[s] + geo * lv.dx
        lv.kind[s] = _abi.SLOT_LEAF
    return lv


@pytest.mark.parametrize("t", [_abi.RHO, _abi.DRHODT])
def test_level_plan_partial_halo_equals_zeroed_full_halo(ctx, t):
    lv = halo_level()
    lists, zeroed = partial_lists(lv, 17)
    print(f"partial level type {t}: slots {[e['slot'] for e in lists]}", flush=True)
    outs, rows = run(ctx, lv, t, partial=lists, reps=2)
    (L, Lc), (L2, Lc2) = outs
    assert np.array_equal(L, L2) and np.array_equal(Lc, Lc2)          # replays
    (Lz, Lcz), = run(ctx, zeroed, t)[0]
    tg = lv.targets
    assert np.array_equal(L[tg], Lz[tg])
    if t == _abi.RHO:
        assert np.array_equal(Lc[tg], Lcz[tg])
    # the plan rebuilt the partial slots' rows: zeros and the latest entries
    for e in lists:
        s = e["slot"]
        if "M" in e:
            assert np.array_equal(rows[0][s], zeroed.M[s]) and np.array_equal(rows[1][s], zeroed.com[s])
        else:
            assert np.array_equal(rows[2][s], zeroed.mon[s])
    Lr, Lcr = oracle.level_interactions(zeroed, type_=t)
    assert normwise(L[tg], Lr[tg]) <= TOL
    if t == _abi.RHO:
        assert normwise(Lc[tg], Lcr[tg]) <= TOL
    # the left-out cells mattered
    Lf, _ = oracle.level_interactions(lv, type_=t)
    assert normwise(L[tg], Lf[tg]) > 1e-6


def test_level_plan_rejects_bad_partial_lists(ctx):
    lv = workloads.uniform_level((3, 3, 3), ((1, 2),) * 3, seed=5)
    dl = hip.LevelOnDevice(lv)
    buf = hip.DeviceArray.from_numpy(np.zeros((4, 20)))
    try:
        target = int(lv.targets[0])
        halo = next(s for s in range(lv.nslots) if s != target)
        bad = [
            [{"slot": target, "cells": np.arange(4), "M": buf.addr, "com": buf.addr}],      # a target
            [{"slot": halo, "cells": np.array([0, 512]), "M": buf.addr, "com": buf.addr}],  # cell out of range
            [{"slot": halo, "cells": np.arange(4)}],                                        # no data
            [{"slot": halo, "cells": np.arange(4), "M": buf.addr, "com": buf.addr}] * 2,    # slot twice
        ]
        for spec in bad:
            with pytest.raises(native.OctofmmError) as e:
                dl.plan(ctx, partial=spec)
            assert e.value.status == _abi.ERR_INVALID_ARGUMENT
    finally:
        buf.free()
        dl.free()
