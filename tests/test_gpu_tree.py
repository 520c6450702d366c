"""GPU parity of the whole adaptive-tree gravity solve (upward M2M, level
interactions with the root stencil on level 0, downward L2L and the leaf force
assembly) against the oracle's tree solve on the same tree.

Gate: per level and component, max over cells |gpu - ref| / max |ref| <= 1e-12
for L, L_c and the leaf phi / g (FP64); multipoles of non-leaf cells within
1e-12 of the size of their terms (m0 dx^order: moments about the centre of mass
cancel towards the noise level of the masses, so a relative gate on the result
would measure that cancellation; DRHODT sources are signed, and there the
multipoles are gated normwise per component like L); the GPU's total force and
torque over all leaf cells vanish like the oracle's.
"""
import numpy as np
import pytest

import oracle
from metrics import normwise
from octotiger_amd import _abi, hip, native, tree as octree, workloads
from workloads_helpers import leaf_cells

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(theta=0.5, device=0)
    yield c
    c.close()


def run(ctx, tree, dev=None, type_=_abi.RHO, precision=_abi.PRECISION_FP64,
        flags=_abi.PLAN_USE_GRAPH | _abi.PLAN_AUTOTUNE, correct_am=True, reps=1):
    dev = dev or hip.TreeOnDevice(tree)
    plan = dev.plan(ctx, type_=type_, precision=precision, flags=flags, correct_am=correct_am)
    s = hip.Stream()
    for _ in range(reps):
        plan.execute(s.handle)
    s.synchronize()
    return plan, dev, dev.fetch("M", "com", "L", "L_c", "g")


def compare(tree, got, ref, rho=True, tol=TOL):
    worst = {}
    for li, (lv, g, r) in enumerate(zip(tree.levels, got, ref)):
        nl = lv.kind == _abi.SLOT_NONLEAF
        if nl.any():
            if rho:
                # moments about the centre of mass cancel towards the noise level,
                # so they are gated against the size of their terms, m0 dx^order
                order = np.array([0] + [1] * 3 + [2] * 6 + [3] * 10)
                scale = np.abs(r["M"][nl][..., :1]) * lv.dx ** order
                worst[f"M{li}"] = float((np.abs(g["M"][nl] - r["M"][nl]) / scale).max())
            else:
                # DRHODT sources are signed, so a cell's m0 can cancel while its
                # higher moments do not: per component, normwise over the level
                worst[f"M{li}"] = normwise(g["M"][nl], r["M"][nl])
        worst[f"com{li}"] = np.abs(g["com"] - r["com"]).max() / tree.size
        worst[f"L{li}"] = normwise(g["L"], r["L"])
        if rho:
            worst[f"Lc{li}"] = normwise(g["L_c"], r["L_c"])
        lf = lv.kind == _abi.SLOT_LEAF
        if lf.any():
            comps = [0, 1, 2, 3] if rho else [0]
            worst[f"g{li}"] = normwise(g["g"][lf][..., comps], r["g"][lf][..., comps])
    bad = {k: v for k, v in worst.items() if not v <= tol}
    assert not bad, f"beyond {tol}: {bad}"
    return worst


def totals(tree, res):
    pos, m = leaf_cells(tree)
    g = np.concatenate([res[i]["g"][lv.kind == _abi.SLOT_LEAF].reshape(-1, 4)
                        for i, lv in enumerate(tree.levels) if (lv.kind == _abi.SLOT_LEAF).any()])
    F = m[:, None] * g[:, 1:]
    tq = np.cross(pos, F)
    return np.abs(F.sum(0)).max() / np.abs(F).sum(), np.abs(tq.sum(0)).max() / np.abs(tq).sum()


def test_small_tree_matches_oracle(ctx):
    t = workloads.small_tree()
    print(f"tree {t.name} nodes {t.info['nodes_per_level']}", flush=True)
    plan, dev, got = run(ctx, t)
    ref = oracle.tree_solve(t)
    compare(t, got, ref)
    st = plan.stats()
    assert st["nodes"] == t.nnodes and st["root_pairs"] == 251496
    assert st["p2p_pairs"] > 0 and st["p2m_pairs"] > 0 and st["near_pairs"] > 0
    ef, et = totals(t, got)
    assert ef < 1e-13 and et < 1e-13
    dev.free()


@pytest.mark.parametrize("theta", [0.4, 0.34])
def test_small_tree_other_theta_matches_oracle(theta):
    """Below theta = 0.5 the stencil reaches two coarse cells: the leaf kernel
    takes its list-driven p2p path over a wider window, with the same
    contiguous tree epilogue."""
    c = native.Context(theta=theta, device=0)
    t = workloads.small_tree()
    print(f"tree {t.name} theta {theta}", flush=True)
    plan, dev, got = run(c, t)
    compare(t, got, oracle.tree_solve(t, theta=theta))
    ef, et = totals(t, got)
    assert ef < 1e-13 and et < 1e-13
    plan.close()
    dev.free()
    c.close()


def test_tree_replays_clear_and_repeat(ctx):
    t = workloads.small_tree()
    plan, dev, first = run(ctx, t)
    s = hip.Stream()
    for rep in range(3):
        dev.levels[-1]["L"].zero()   # eager work between the replays
        plan.execute(s.handle)
        s.synchronize()
        again = dev.fetch("L", "L_c", "g")
        for a, b in zip(first, again):
            for k in ("L", "L_c", "g"):
                assert np.array_equal(a[k], b[k]), f"replay {rep} differs in {k}"
    dev.free()


@pytest.mark.parametrize("name", ["small", "c5"])
def test_tree_without_graph_matches_graph(ctx, name):
    """Eager launches on the plan's streams give the graph's bits, on the
    small tree and on C5 (four streams, p2m and leaf kernels of every level)."""
    t = workloads.small_tree() if name == "small" else workloads.config5()
    print(f"tree {t.name}", flush=True)
    _, d1, a = run(ctx, t, flags=_abi.PLAN_USE_GRAPH)
    _, d2, b = run(ctx, t, flags=0)
    for x, y in zip(a, b):
        for k in ("L", "L_c", "g"):
            assert np.array_equal(x[k], y[k])
    d1.free()
    d2.free()


@pytest.mark.parametrize("name", ["small", "c5"])
def test_tree_drhodt_matches_oracle(ctx, name):
    """DRHODT solve (monopoles of the leaves as the source, the RHO solve's
    centres of mass), on the small tree and at full C5 size."""
    t = workloads.small_tree() if name == "small" else workloads.config5()
    print(f"tree {t.name} DRHODT", flush=True)
    nth = 16 if name == "c5" else None
    ref_rho = oracle.tree_solve(t, nthreads=nth)
    _, dev, _ = run(ctx, t)                      # leaves the RHO centres of mass on the device
    rng = np.random.default_rng(5)
    for lv, d in zip(t.levels, dev.levels):
        lv.mon = np.ascontiguousarray(rng.normal(size=lv.mon.shape) * lv.dx ** 3 *
                                      (lv.kind == _abi.SLOT_LEAF)[:, None])
        d["mon"].upload(lv.mon)
    _, _, got = run(ctx, t, dev=dev, type_=_abi.DRHODT)
    ref = oracle.tree_solve(t, type_=_abi.DRHODT, coms=[o["com"] for o in ref_rho], nthreads=nth)
    compare(t, got, ref, rho=False)
    dev.free()


def test_root_only_tree_matches_oracle(ctx):
    t = octree.build_tree(lambda lev, cs, xb, dx: np.zeros(len(cs), bool), 0, name="root-leaf")
    workloads._assign_leaf_masses(t, lambda x: np.ones(x.shape[:-1]), 99, 1.0)
    _, dev, got = run(ctx, t)
    compare(t, got, oracle.tree_solve(t))
    dev.free()


def test_config5_full_size_matches_oracle(ctx):
    t = workloads.config5()
    print(f"tree {t.name} nodes {t.info['nodes_per_level']}", flush=True)
    plan, dev, got = run(ctx, t)
    ref = oracle.tree_solve(t, nthreads=16)
    worst = compare(t, got, ref)
    print("worst", max(worst.values()), flush=True)
    ef, et = totals(t, got)
    assert ef < 1e-13 and et < 1e-13
    st = plan.stats()
    assert st["nodes"] == 10697 and st["far_pairs"] > 0 and st["p2m_pairs"] > 0
    dev.free()


def test_config5_mixed_precision_force_error(ctx):
    t = workloads.config5()
    _, d64, a = run(ctx, t)
    _, d32, b = run(ctx, t, precision=_abi.PRECISION_MIXED)
    ga = np.concatenate([r["g"][lv.kind == _abi.SLOT_LEAF].reshape(-1, 4) for lv, r in zip(t.levels, a)])
    gb = np.concatenate([r["g"][lv.kind == _abi.SLOT_LEAF].reshape(-1, 4) for lv, r in zip(t.levels, b)])
    ef = np.sqrt(((gb[:, 1:] - ga[:, 1:]) ** 2).sum() / (ga[:, 1:] ** 2).sum())
    ephi = np.abs(gb[:, 0] - ga[:, 0]).max() / np.abs(ga[:, 0]).max()
    print(f"mixed vs fp64: force rms {ef:.3e} phi max {ephi:.3e}", flush=True)
    assert ef <= 1e-6 and ephi <= 1e-6
    d64.free()
    d32.free()


def test_solid_sphere_reproduces_reference_regression(ctx):
    """Octo-Tiger's solid-sphere regression (test_problems/sphere, max_level 3):
    the GPU solve matches the oracle and prints the reference's pot norms."""
    t = workloads.solid_sphere_tree()
    print(f"tree {t.name} nodes {t.info['nodes_per_level']}", flush=True)
    _, dev, got = run(ctx, t)
    compare(t, got, oracle.tree_solve(t))
    l1, l2 = workloads.sphere_pot_norms(t, [r["g"][..., 0] for r in got])
    assert f"pot {l1:e} {l2:e}" == "pot 4.470774e-01 3.176879e+00"
    dev.free()


@pytest.mark.parametrize("name", ["small", "c5"])
def test_leaf_g_only_writes_g_and_leaves_leaf_rows_alone(ctx, name):
    """OCTOFMM_TREE_LEAF_G_ONLY: g and L_c of every cell, and L of non-leaf
    cells, are bit for bit those of the full solve; leaf cells' L rows keep
    what they held before the solve."""
    t = workloads.small_tree() if name == "small" else workloads.config5()
    print(f"leaf g-only tree {t.name}", flush=True)
    _, d1, full = run(ctx, t)
    dev = hip.TreeOnDevice(t)
    for lv, d in zip(t.levels, dev.levels):
        d["L"].upload(np.full((lv.nnodes, 512, 20), 7.0))
    _, d2, got = run(ctx, t, dev=dev, flags=_abi.PLAN_USE_GRAPH | _abi.PLAN_AUTOTUNE | _abi.TREE_LEAF_G_ONLY)
    for lv, a, b in zip(t.levels, full, got):
        leaf = lv.kind == _abi.SLOT_LEAF
        assert np.array_equal(a["g"], b["g"]) and np.array_equal(a["L_c"], b["L_c"])
        assert np.array_equal(a["L"][~leaf], b["L"][~leaf])
        assert (b["L"][leaf] == 7.0).all()
    d1.free()
    d2.free()


def test_leaf_g_only_needs_g():
    """The flag without a g array on a level with leaves is refused at plan
    creation."""
    c = native.Context(theta=0.5, device=0)
    t = workloads.small_tree()
    dev = hip.TreeOnDevice(t)
    addr = dev.addresses()
    for a in addr:
        a.pop("g")
    with pytest.raises(native.OctofmmError) as e:
        c.tree_plan(t, addr, flags=_abi.PLAN_USE_GRAPH | _abi.TREE_LEAF_G_ONLY)
    assert e.value.status == _abi.ERR_INVALID_ARGUMENT
    dev.free()
    c.close()
