"""GPU parity: liboctofmm against the CPU FP64 oracle and the golden fixtures.

Gate (FP64 path): per expansion component, max over cells of |gpu - ref|
divided by max over cells of |ref| <= 1e-12 (reordered summation only), and
total force / torque of the target cells within 1e-12 of the reference's own
totals (|F - F_ref| / |F_ref|), or within their rounding floor 4 eps sum_i |F_i|
where the totals cancel further than that (tests/metrics.py).
"""
import os

import numpy as np
import pytest

import fmm_cases
import oracle
from metrics import describe, force_torque_gate, normwise
from octotiger_amd import _abi, hip, interface, native, workloads

pytestmark = pytest.mark.gpu

TOL = 1e-12
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(theta=0.5, device=0)
    yield c
    c.close()


def run_case(ctx, name, t):
    sg, keep, kind, cmp_ = fmm_cases.subgrid(name, t)
    L = np.zeros((512, 20))
    Lc = np.zeros((512, 3))
    if kind == "monopole":
        ctx.monopole_interactions(sg, cmp_, L, Lc)
    else:
        ctx.multipole_interactions(sg, L, Lc)
    return L, Lc, keep[-1]


@pytest.mark.parametrize("name", [c[0] for c in fmm_cases.CASES])
@pytest.mark.parametrize("t", [_abi.RHO, _abi.DRHODT])
def test_subgrid_dropin_matches_golden(ctx, name, t):
    print(f"case {name} type {t}", flush=True)
    L, Lc, lv = run_case(ctx, name, t)
    g = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    tn = "rho" if t == _abi.RHO else "drhodt"
    Lr, Lcr = g[f"L_{tn}"], g[f"Lc_{tn}"]
    assert normwise(L, Lr) <= TOL
    if t == _abi.RHO:
        assert normwise(Lc, Lcr) <= TOL
        mass = lv.M[13, :, 0] if lv.kind[13] == _abi.SLOT_NONLEAF else lv.mon[13]
        ok, e = force_torque_gate(L, Lc, Lr, Lcr, mass, lv.com[13], TOL)
        print(describe(e), flush=True)
        assert ok


def test_dropin_accumulates_L_and_overwrites_Lc(ctx):
    sg, keep, kind, _ = fmm_cases.subgrid("c1_nonroot", _abi.RHO)
    L0 = np.full((512, 20), 3.0)
    Lc0 = np.full((512, 3), 7.0)
    ctx.multipole_interactions(sg, L0, Lc0)
    L1, Lc1 = np.zeros((512, 20)), np.zeros((512, 3))
    ctx.multipole_interactions(sg, L1, Lc1)
    assert np.array_equal(L0, L1 + 3.0)
    assert np.array_equal(Lc0, Lc1)


def test_interface_mirror_matches_oracle(ctx):
    lv = fmm_cases.build_level("mixed_halo")
    mi = interface.multipole_interaction_interface(ctx)
    grid = interface.Grid()
    mi.set_grid_ptr(grid)
    nbrs = interface.neighbors_from_level(lv, 13)
    for f in (0, 2, 25):    # the empty corners of this case
        nbrs[f] = interface.NeighborGravity(is_monopole=False)
    mi.compute_multipole_interactions(None, lv.M[13], [lv.com[13]], nbrs, interface.RHO, lv.dx,
                                      [False] * 26, lv.xbase[13], False)
    sg, keep = workloads.subgrid_view(lv, 13)
    Lr, Lcr = oracle.multipole_interactions(sg)
    assert normwise(grid.L, Lr) <= TOL and normwise(grid.L_c, Lcr) <= TOL


def run_level(ctx, lv, t=_abi.RHO, flags=_abi.PLAN_USE_GRAPH, repeats=1):
    assert hip.device_count() > 0, "GPU tests need a HIP device"
    dl = hip.LevelOnDevice(lv)
    stream = hip.Stream()
    plan = dl.plan(ctx, type_=t, flags=flags)
    outs = []
    for _ in range(repeats):
        dl.zero_outputs()
        plan.execute(stream.handle)
        stream.synchronize()
        outs.append((dl.L.numpy(), dl.L_c.numpy()))
    stats = plan.stats()
    plan.close()
    stream.destroy()
    dl.free()
    return outs, stats


@pytest.mark.parametrize("t", [_abi.RHO, _abi.DRHODT])
@pytest.mark.parametrize("graph", [True, False])
def test_level_uniform_small(ctx, t, graph):
    lv = workloads.uniform_level((4, 4, 4), ((1, 3),) * 3, seed=99, name="uniform-small")
    print(f"uniform small type {t} graph {graph}", flush=True)
    outs, _ = run_level(ctx, lv, t, _abi.PLAN_USE_GRAPH if graph else 0)
    L, Lc = outs[0]
    Lr, Lcr = oracle.level_interactions(lv, type_=t)
    tg = lv.targets
    assert normwise(L[tg], Lr[tg]) <= TOL
    if t == _abi.RHO:
        assert normwise(Lc[tg], Lcr[tg]) <= TOL


@pytest.mark.parametrize("t", [_abi.RHO, _abi.DRHODT])
@pytest.mark.parametrize("graph", [True, False])
def test_level_mixed_kinds(ctx, t, graph):
    """Both kinds of items: the leaf kernel runs on the plan's side stream
    beside the multipole-target kernel, in the graph and in direct launches;
    two executes are bit-identical."""
    lv = workloads.mixed_level((6, 6, 6), seed=11, leaf_fraction=0.5)
    print(f"mixed level {lv.info} type {t} graph {graph}", flush=True)
    outs, stats = run_level(ctx, lv, t, _abi.PLAN_USE_GRAPH if graph else 0, repeats=2)
    assert all(np.array_equal(a, b) for a, b in zip(outs[0], outs[1]))
    L, Lc = outs[0]
    Lr, Lcr = oracle.level_interactions(lv, type_=t)
    tg = lv.targets
    assert stats["n_leaf_items"] > 0 and stats["n_m2l_items"] > 0
    assert normwise(L[tg], Lr[tg]) <= TOL
    if t == _abi.RHO:
        assert normwise(Lc[tg], Lcr[tg]) <= TOL


@pytest.mark.parametrize("theta", [0.34, 0.45])
@pytest.mark.parametrize("variant", [8, 304, 504, 704, 708])
def test_level_mixed_kinds_other_theta(theta, variant, monkeypatch):
    """Wider stencils (theta 0.34: coarse offsets up to 2, p2p window 18^3)
    under every lane mapping, the staged ones included."""
    monkeypatch.setenv("OCTOFMM_VARIANT", str(variant))
    c = native.Context(theta=theta, device=0)
    lv = workloads.mixed_level((5, 5, 5), seed=13, leaf_fraction=0.5)
    print(f"mixed level theta {theta} variant {variant} {lv.info}", flush=True)
    outs, stats = run_level(c, lv, _abi.RHO)
    c.close()
    assert stats["variant"] == variant
    L, Lc = outs[0]
    Lr, Lcr = oracle.level_interactions(lv, theta=theta, type_=_abi.RHO)
    tg = lv.targets
    assert stats["p2m_pairs"] > 0 and stats["p2p_pairs"] > 0 and stats["near_pairs"] > 0
    assert normwise(L[tg], Lr[tg]) <= TOL
    assert normwise(Lc[tg], Lcr[tg]) <= TOL


def test_level_graph_replays_identical(ctx):
    lv = workloads.mixed_level((5, 5, 5), seed=5, leaf_fraction=0.4)
    outs, _ = run_level(ctx, lv, _abi.RHO, repeats=3)
    for L, Lc in outs[1:]:
        assert np.array_equal(L, outs[0][0]) and np.array_equal(Lc, outs[0][1])


def test_level_config2_full_size(ctx):
    """512 non-leaf sub-grids in one pass against the oracle, plus the total
    force and torque of all target cells."""
    lv = workloads.config2()
    print("config2", flush=True)
    outs, stats = run_level(ctx, lv, _abi.RHO)
    L, Lc = outs[0]
    assert stats["far_pairs"] == 512 * 96768 and stats["near_pairs"] == 0
    Lr, Lcr = oracle.level_interactions(lv)
    tg = lv.targets
    assert normwise(L[tg], Lr[tg]) <= TOL and normwise(Lc[tg], Lcr[tg]) <= TOL
    ok, e = force_torque_gate(L[tg], Lc[tg], Lr[tg], Lcr[tg], lv.M[tg, :, 0], lv.com[tg], TOL)
    print(describe(e), flush=True)
    assert ok


@pytest.mark.parametrize("t", [_abi.RHO, _abi.DRHODT])
@pytest.mark.parametrize("variant", [8, 4, 2, 304, 302, 504, 704, 702, 708, 804, 802])
def test_level_kernel_variants(ctx, variant, t, monkeypatch):
    """Every lane mapping and workgroup shape of the level kernel gives the
    oracle's result (mixed level: M2L far + near passes, p2p, p2m), for both
    solve types, on every replay of its graph (the persistent variants clear
    their work counters in each one)."""
    monkeypatch.setenv("OCTOFMM_VARIANT", str(variant))
    lv = workloads.mixed_level((6, 6, 6), seed=13, leaf_fraction=0.5)
    print(f"variant {variant} type {t}", flush=True)
    outs, stats = run_level(ctx, lv, t, repeats=2)
    assert stats["variant"] == variant
    L, Lc = outs[0]
    assert np.array_equal(L, outs[1][0]) and np.array_equal(Lc, outs[1][1])
    Lr, Lcr = oracle.level_interactions(lv, type_=t)
    tg = lv.targets
    assert normwise(L[tg], Lr[tg]) <= TOL
    if t == _abi.RHO:
        assert normwise(Lc[tg], Lcr[tg]) <= TOL


@pytest.mark.parametrize("theta", [0.5, 0.34])
@pytest.mark.parametrize("variant", [8, 304, 504, 704, 708, 804])
def test_level_edge_targets_missing_neighbours(theta, variant, monkeypatch):
    """Targets on the faces, edges and corners of the domain, whose partner
    coarse cells beyond it do not exist: every lane mapping must leave those
    pairs out."""
    monkeypatch.setenv("OCTOFMM_VARIANT", str(variant))
    c = native.Context(theta=theta, device=0)
    lv = workloads.uniform_level((3, 3, 3), ((0, 3),) * 3, seed=31)
    print(f"edge targets theta {theta} variant {variant}", flush=True)
    outs, stats = run_level(c, lv, _abi.RHO)
    c.close()
    assert stats["variant"] == variant
    assert (lv.neighbors[lv.targets] < 0).any()
    L, Lc = outs[0]
    Lr, Lcr = oracle.level_interactions(lv, theta=theta, type_=_abi.RHO)
    tg = lv.targets
    assert normwise(L[tg], Lr[tg]) <= TOL
    assert normwise(Lc[tg], Lcr[tg]) <= TOL


def test_autotuned_plan_leaves_outputs_untouched_until_executed(ctx):
    lv = workloads.uniform_level((4, 4, 4), ((1, 3),) * 3, seed=21)
    dl = hip.LevelOnDevice(lv)
    plan = dl.plan(ctx, type_=_abi.RHO, flags=_abi.PLAN_USE_GRAPH | _abi.PLAN_AUTOTUNE)
    assert plan.stats()["variant"] in (8, 4, 2, 304, 302, 504, 704, 702, 708)
    assert not dl.L.numpy().any() and not dl.L_c.numpy().any()
    stream = hip.Stream()
    plan.execute(stream.handle)
    stream.synchronize()
    Lr, Lcr = oracle.level_interactions(lv)
    tg = lv.targets
    assert normwise(dl.L.numpy()[tg], Lr[tg]) <= TOL
    plan.close()
    stream.destroy()
    dl.free()


MIXED_FORCE_TOL = 1e-6      # total force and torque, relative to the oracle's totals
MIXED_COMPONENT_TOL = 1e-5  # per component normwise
MIXED_VARIANTS = [8, 4, 2, 704, 702, 708, 804, 802]
# FP32 share of the far pairs at theta = 0.5, mixed_far_d2 = 4: mapping 0
# chooses per pair (|d|^2 > 4: 183 of 189 per cell), mapping 7 per coarse
# offset (the 20 edge and corner offsets: 153 of 189; the 6 face offsets,
# which hold the |d| = 2 pairs, stay FP64)
MIXED_SHARE = {8: 0.96, 4: 0.96, 2: 0.96, 704: 0.8, 702: 0.8, 708: 0.8, 804: 0.8, 802: 0.8}


@pytest.mark.parametrize("t", [_abi.RHO, _abi.DRHODT])
@pytest.mark.parametrize("variant", MIXED_VARIANTS)
def test_mixed_precision_far_field(ctx, t, variant, monkeypatch):
    """FP32 pair arithmetic for far offsets beyond 2 cells, FP64 accumulation
    and FP64 angular-momentum correction: errors against the FP64 oracle,
    under every kernel variant the autotuner may pick for this precision."""
    monkeypatch.setenv("OCTOFMM_VARIANT", str(variant))
    lv = workloads.uniform_level((4, 4, 4), ((1, 3),) * 3, seed=77)
    dl = hip.LevelOnDevice(lv)
    stream = hip.Stream()
    plan = dl.plan(ctx, type_=t, precision=_abi.PRECISION_MIXED, mixed_far_d2=4, flags=_abi.PLAN_USE_GRAPH)
    st = plan.stats()
    assert st["variant"] == variant
    assert st["mixed_far_pairs"] > MIXED_SHARE[variant] * st["far_pairs"]
    plan.execute(stream.handle)
    stream.synchronize()
    L, Lc = dl.L.numpy(), dl.L_c.numpy()
    plan.close()
    stream.destroy()
    dl.free()
    Lr, Lcr = oracle.level_interactions(lv, type_=t)
    tg = lv.targets
    e = normwise(L[tg], Lr[tg])
    print(f"mixed variant {variant} type {t}: normwise {e:.3e}", flush=True)
    assert 0 < e <= MIXED_COMPONENT_TOL
    if t == _abi.RHO:
        ok, ft = force_torque_gate(L[tg], Lc[tg], Lr[tg], Lcr[tg], lv.M[tg, :, 0], lv.com[tg], MIXED_FORCE_TOL)
        print(f"mixed {describe(ft)}; Lc {normwise(Lc[tg], Lcr[tg]):.3e}", flush=True)
        assert ok
        assert normwise(Lc[tg], Lcr[tg]) <= TOL


@pytest.mark.parametrize("variant", [8, 704])
def test_mixed_precision_other_theta(variant, monkeypatch):
    """Mixed precision on the wider stencil of theta = 0.34 (coarse offsets
    up to 2: 124 offsets, most of whose far pairs lie beyond 2 cells), per pair
    (mapping 0) and per coarse offset (mapping 7), against the FP64 oracle."""
    monkeypatch.setenv("OCTOFMM_VARIANT", str(variant))
    c = native.Context(theta=0.34, device=0)
    lv = workloads.uniform_level((4, 4, 4), ((1, 3),) * 3, seed=78)
    dl = hip.LevelOnDevice(lv)
    stream = hip.Stream()
    plan = dl.plan(c, type_=_abi.RHO, precision=_abi.PRECISION_MIXED, mixed_far_d2=4, flags=_abi.PLAN_USE_GRAPH)
    st = plan.stats()
    plan.execute(stream.handle)
    stream.synchronize()
    L, Lc = dl.L.numpy(), dl.L_c.numpy()
    plan.close()
    stream.destroy()
    dl.free()
    c.close()
    assert st["variant"] == variant and st["mixed_far_pairs"] > 0.8 * st["far_pairs"]
    Lr, Lcr = oracle.level_interactions(lv, theta=0.34, type_=_abi.RHO)
    tg = lv.targets
    e = normwise(L[tg], Lr[tg])
    ok, ft = force_torque_gate(L[tg], Lc[tg], Lr[tg], Lcr[tg], lv.M[tg, :, 0], lv.com[tg], MIXED_FORCE_TOL)
    print(f"mixed theta 0.34 variant {variant}: FP32 share {st['mixed_far_pairs'] / st['far_pairs']:.3f} "
          f"normwise {e:.3e} {describe(ft)}", flush=True)
    assert 0 < e <= MIXED_COMPONENT_TOL and ok
    assert normwise(Lc[tg], Lcr[tg]) <= TOL


def test_level_config3_full_size(ctx):
    """2048 sub-grids of mixed kind (leaf p2p/p2m, non-leaf interior, non-leaf
    next to leaves) in one pass against the oracle."""
    lv = workloads.config3()
    print(f"config3 {lv.info}", flush=True)
    outs, stats = run_level(ctx, lv, _abi.RHO)
    L, Lc = outs[0]
    assert stats["n_m2l_items"] + stats["n_leaf_items"] == 2048
    assert stats["near_pairs"] > 0 and stats["p2p_pairs"] > 0 and stats["p2m_pairs"] > 0
    Lr, Lcr = oracle.level_interactions(lv)
    tg = lv.targets
    assert normwise(L[tg], Lr[tg]) <= TOL and normwise(Lc[tg], Lcr[tg]) <= TOL
    mass = np.where(lv.kind[tg, None] == _abi.SLOT_LEAF, lv.mon[tg], lv.M[tg, :, 0])
    ok, e = force_torque_gate(L[tg], Lc[tg], Lr[tg], Lcr[tg], mass, lv.com[tg], TOL)
    print(describe(e), flush=True)
    assert ok


def test_dropin_checks_finiteness_when_asked():
    """With OCTOFMM_CHECK_FINITE the drop-in rejects a NaN in a neighbour's
    multipoles with ERR_NONFINITE before any device work; clean inputs pass."""
    c = native.Context(theta=0.5, device=0, check_finite=True)
    sg, keep, kind, _ = fmm_cases.subgrid("mixed_halo", _abi.RHO)
    L = np.zeros((512, 20))
    Lc = np.zeros((512, 3))
    c.multipole_interactions(sg, L, Lc)                       # clean: accepted
    f = next(i for i in range(26) if sg.neighbors[i].kind == _abi.NEIGHBOR_MULTIPOLE)
    sg.neighbors[f].M[7] = float("nan")
    L2 = np.zeros((512, 20))
    with pytest.raises(native.OctofmmError) as e:
        c.multipole_interactions(sg, L2, Lc)
    assert e.value.status == _abi.ERR_NONFINITE and not L2.any()
    c.close()


def _gpu_case(ctx, kind, lv, sg):
    cmp_ = any(lv.neighbors[13, r] >= 0 and lv.kind[lv.neighbors[13, r]] == _abi.SLOT_NONLEAF
               for r in range(27) if r != 13)
    L, Lc = np.zeros((512, 20)), np.zeros((512, 3))
    if kind == "monopole":
        ctx.monopole_interactions(sg, cmp_, L, Lc)
        Lr, Lcr = oracle.monopole_interactions(sg, cmp_)
    else:
        ctx.multipole_interactions(sg, L, Lc)
        Lr, Lcr = oracle.multipole_interactions(sg)
    return L, Lc, Lr, Lcr


@pytest.mark.parametrize("name", ["c1_nonroot", "mixed_halo", "leaf_p2p", "leaf_p2m"])
@pytest.mark.parametrize("pattern", ["all_shuffled", "subset", "repeats"])
def test_partial_halo_scattered_on_device(ctx, name, pattern):
    """Partial halos handed over in grid::get_ilist_n_bnd order are scattered
    into their slots on the device: the result equals the full-halo call on a
    level whose left-out cells hold zeros, bit for bit (every cell in shuffled
    order: the full-halo result itself; repeated cells: the later entry wins),
    and the oracle's to 1e-12."""
    lv = fmm_cases.build_level(name)
    kind = dict(fmm_cases.CASES)[name]
    spec = fmm_cases.partial_specs(name)[pattern]
    print(f"partial halo {name} {pattern}", flush=True)
    sg, keep = fmm_cases.partial_subgrid(lv, spec, stale=fmm_cases.stale_prefix(pattern, spec))
    L, Lc, Lr, Lcr = _gpu_case(ctx, kind, lv, sg)
    full = lv if pattern == "all_shuffled" else fmm_cases.zeroed_level(lv, spec)
    sgf, keepf = workloads.subgrid_view(full, 13)
    Lf, Lcf, _, _ = _gpu_case(ctx, kind, full, sgf)
    assert np.array_equal(L, Lf) and np.array_equal(Lc, Lcf)
    assert normwise(L, Lr) <= TOL and normwise(Lc, Lcr) <= TOL


def test_tuning_cache_by_level_shape(ctx, tmp_path, monkeypatch):
    """Autotuned plans of one level shape take one variant: the first plan
    times the variants and records its choice (appended to the file
    OCTOFMM_TUNE_CACHE names), later plans of the shape reuse it."""
    path = tmp_path / "tune.txt"
    monkeypatch.setenv("OCTOFMM_TUNE_CACHE", str(path))
    lv = workloads.uniform_level((5, 5, 3), ((1, 4), (1, 4), (1, 2)), seed=57, name="tune-cache")
    dl = hip.LevelOnDevice(lv)
    variants = []
    for _ in range(3):
        plan = dl.plan(ctx, type_=_abi.RHO, flags=_abi.PLAN_AUTOTUNE)
        variants.append(plan.stats()["variant"])
        plan.close()
    dl.free()
    lines = path.read_text().split("\n")
    keys = [ln.split() for ln in lines if f"/m{len(lv.targets)}/" in ln]
    print(f"tuning cache: {variants} {keys}", flush=True)
    assert len(set(variants)) == 1
    assert len(keys) == 1 and int(keys[0][1]) == variants[0]
