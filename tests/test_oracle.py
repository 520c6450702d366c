"""Pins the CPU oracle (oracle/octofmm_oracle.cpp) without a GPU.

  * golden fixtures written from the reference's own templates (bit for bit);
  * the reference templates themselves, when oracle/_ref is built (bit for bit);
  * closed forms: stencil counts, derivatives of -1/r, point-mass sums over an
    independently enumerated pair set, Newton's third law, multipole
    expansions against direct summation (truncation order), linearity.
"""
import os

import numpy as np
import pytest

import fmm_cases
import oracle
from octotiger_amd import _abi, workloads

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RNG = np.random.default_rng(1234)


# ---------------------------------------------------------------- stencils
def test_stencil_counts_theta_half():
    mm, mi, pm, four = oracle.stencils(0.5)
    assert mm.sum() == 342                      # super-imposed stencil
    assert (mm & mi).sum() == 316               # far (multipole) offsets
    assert (mm & (1 - mi)).sum() == 26          # near offsets
    assert (pm == mm).all()                     # p2p uses the same union
    el, ph = oracle.multipole_stencil(0.5)
    assert len(el) == 342 and np.abs(el).max() == 3
    assert (ph == (np.sum(el.astype(float) ** 2, axis=1) >= 4)).all()


def _coarse(i):
    return np.floor_divide(i, 2)


def brute_pairs(theta=0.5):
    """Independent enumeration of (target parity, offset) pairs: partners whose
    parent cells are closer than 1/theta, excluding the cell itself."""
    out = {}
    for p in range(8):
        t = np.array([(p >> 2) & 1, (p >> 1) & 1, p & 1])
        far, near = 0, 0
        for d in np.ndindex(11, 11, 11):
            d = np.array(d) - 5
            if not d.any():
                continue
            dc = _coarse(t + d) - _coarse(t)
            if float((dc ** 2).sum()) < (1.0 / theta) ** 2:
                if 1.0 / np.sqrt((d ** 2).sum()) <= theta:
                    far += 1
                else:
                    near += 1
        out[p] = (far, near)
    return out


def test_per_cell_partner_counts():
    pairs = brute_pairs(0.5)
    assert all(v == (189, 26) for v in pairs.values())


@pytest.mark.parametrize("theta", [0.34, 0.4, 0.5, 0.6])
def test_stencil_union_matches_brute_force(theta):
    mm, mi, pm, _ = oracle.stencils(theta)
    union = np.zeros(1331, np.uint8)
    for p in range(8):
        t = np.array([(p >> 2) & 1, (p >> 1) & 1, p & 1])
        for idx, d in enumerate(np.ndindex(11, 11, 11)):
            d = np.array(d) - 5
            if not d.any():
                continue
            dc = _coarse(t + d) - _coarse(t)
            if 1.0 / np.sqrt(float((dc ** 2).sum())) > theta if dc.any() else True:
                union[idx] = 1
    assert (union == mm).all()


def test_theta_below_floor_rejected():
    with pytest.raises(ValueError):
        oracle.stencils(0.3)


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("theta", [0.34, 0.45, 0.5])
def test_stencils_equal_reference(theta):
    a = oracle.stencils(theta)
    b = oracle.stencils(theta, use_ref=True)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


# ---------------------------------------------------------------- pair arithmetic
def analytic_D(x):
    """Derivatives of -1/r up to order 3 in taylor<4> component order."""
    r = np.sqrt((x ** 2).sum())
    out = np.zeros(20)
    out[0] = -1.0 / r
    out[1:4] = x / r ** 3
    k = 4
    for i in range(3):
        for j in range(i, 3):
            out[k] = (1.0 if i == j else 0.0) / r ** 3 - 3 * x[i] * x[j] / r ** 5
            k += 1
    for i in range(3):
        for j in range(i, 3):
            for l in range(j, 3):
                dl = lambda a, b: 1.0 if a == b else 0.0
                out[k] = (-3 * (dl(i, j) * x[l] + dl(i, l) * x[j] + dl(j, l) * x[i]) / r ** 5
                          + 15 * x[i] * x[j] * x[l] / r ** 7)
                k += 1
    return out


def test_d_factors_against_analytic_derivatives():
    for _ in range(50):
        x = RNG.normal(size=3) * 10 ** RNG.uniform(-2, 1)
        D = oracle.d_factors(x)[5:]
        ref = analytic_D(x)
        assert np.allclose(D, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())


def test_d_factors_floor_on_coincident_points():
    out = oracle.d_factors(np.zeros(3))
    assert np.isfinite(out).all() and out[5] == -1.0e10


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
def test_pair_functions_bitwise_equal_reference():
    for _ in range(200):
        X, Y = RNG.normal(size=3), RNG.normal(size=3)
        mp, mc = RNG.normal(size=20), RNG.normal(size=20)
        mc[0] = abs(mc[0]) + 0.1
        assert np.array_equal(oracle.d_factors(X - Y), oracle.d_factors(X - Y, use_ref=True))
        for rho in (True, False):
            La, Ca = RNG.normal(size=20), RNG.normal(size=3)
            Lb, Cb = La.copy(), Ca.copy()
            oracle.pair_m2l(X, Y, mp, mc, La, Ca, rho)
            oracle.pair_m2l(X, Y, mp, mc, Lb, Cb, rho, use_ref=True)
            assert np.array_equal(La, Lb) and np.array_equal(Ca, Cb)
            La, Ca = RNG.normal(size=4), RNG.normal(size=3)
            Lb, Cb = La.copy(), Ca.copy()
            oracle.pair_p2m(X, Y, mp, La, Ca, rho)
            oracle.pair_p2m(X, Y, mp, Lb, Cb, rho, use_ref=True)
            assert np.array_equal(La, Lb) and np.array_equal(Ca, Cb)


def test_monopole_partner_closed_form():
    """A pure monopole partner gives L = m * D (all 20 components)."""
    X, Y = np.array([0.3, -0.2, 0.1]), np.array([-0.5, 0.4, 0.9])
    m = 2.5
    mp = np.zeros(20)
    mp[0] = m
    for rho in (True, False):
        L, C = np.zeros(20), np.zeros(3)
        mc = np.zeros(20)
        mc[0] = 1.0
        oracle.pair_m2l(X, Y, mp, mc, L, C, rho)
        assert np.allclose(L, m * analytic_D(X - Y), rtol=1e-14, atol=1e-15)


def test_newton_third_law_monopoles():
    XA, XB = np.array([0.1, 0.2, 0.3]), np.array([-0.7, 0.5, 1.1])
    mA, mB = 1.7, 0.6
    LA, LB, C = np.zeros(20), np.zeros(20), np.zeros(3)
    pa, pb = np.zeros(20), np.zeros(20)
    pa[0], pb[0] = mA, mB
    oracle.pair_m2l(XA, XB, pb, pa, LA, C, False)
    oracle.pair_m2l(XB, XA, pa, pb, LB, C, False)
    FA, FB = -mA * LA[1:4], -mB * LB[1:4]
    assert np.allclose(FA, -FB, rtol=1e-14)
    # attractive: force on A points towards B
    assert np.dot(FA, XB - XA) > 0


def cluster(seed, h, n=8):
    """n point masses within +-h of the origin, their COM and raw moments."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-h, h, size=(n, 3))
    m = rng.uniform(0.5, 1.5, size=n)
    com = (m[:, None] * pos).sum(0) / m.sum()
    rel = pos - com
    M = np.zeros(20)
    M[0] = m.sum()
    for k, (a, b, c) in enumerate(workloads._ALPHA):
        M[k + 1] = (m * rel[:, 0] ** a * rel[:, 1] ** b * rel[:, 2] ** c).sum()
    return pos, m, com, M


def direct_expansion(X, pos, m):
    """Taylor coefficients (orders 0..3) of phi(x) = -sum m/|x - y| at X."""
    out = np.zeros(20)
    for y, mm in zip(pos, m):
        out += mm * analytic_D(X - y)
    return out


@pytest.mark.parametrize("rho", [True, False])
def test_m2l_converges_to_direct_sum(rho):
    """Far field of a point-mass cluster: the 20-component local expansion
    matches direct summation with errors shrinking by the truncation order."""
    pos, m, com, M = cluster(7, 0.05)
    errs = []
    for dist in (0.5, 1.0, 2.0):
        X = com + dist * np.array([0.6, -0.48, 0.64])
        L, C = np.zeros(20), np.zeros(3)
        mc = np.zeros(20)
        mc[0] = 1.0
        oracle.pair_m2l(X, com, M, mc, L, C, rho)
        ref = direct_expansion(X, pos, m)
        errs.append([abs(L[0] - ref[0]) / abs(ref[0]), np.abs(L[1:4] - ref[1:4]).max() / np.abs(ref[1:4]).max()])
    errs = np.array(errs)
    assert errs[0, 0] < 1e-4 and errs[0, 1] < 1e-3
    # potential: order-3 multipoles leave an O((h/r)^4) relative error
    assert errs[0, 0] / errs[2, 0] > 2 ** 3.5 * 2 ** 3.5 / 4
    assert errs[0, 1] / errs[2, 1] > 2 ** 2.5 * 2 ** 2.5 / 4


def test_m2l_linear_in_partner():
    X, Y = np.array([0.9, 0.1, -0.4]), np.zeros(3)
    m1, m2 = RNG.normal(size=20), RNG.normal(size=20)
    mc = np.zeros(20)
    mc[0] = 1.0
    out = []
    for mp in (m1, m2, 2.0 * m1 + 3.0 * m2):
        L, C = np.zeros(20), np.zeros(3)
        oracle.pair_m2l(X, Y, mp, mc, L, C, False)
        out.append(L)
    assert np.allclose(out[2], 2.0 * out[0] + 3.0 * out[1], rtol=1e-12, atol=1e-14)


def test_correction_vanishes_for_proportional_octupoles():
    """The angular-momentum correction acts on m_p - m_c (m_p0/m_c0): a partner
    whose octupole is the target's scaled by the mass ratio gives none."""
    X, Y = np.array([0.9, 0.1, -0.4]), np.zeros(3)
    mc = RNG.normal(size=20)
    mc[0] = 2.0
    mp = RNG.normal(size=20)
    mp[0] = 3.0
    mp[10:] = mc[10:] * (mp[0] / mc[0])
    L, C = np.zeros(20), np.zeros(3)
    oracle.pair_m2l(X, Y, mp, mc, L, C, True)
    assert np.abs(C).max() < 1e-14


# ---------------------------------------------------------------- whole sub-grids
@pytest.mark.parametrize("name", [c[0] for c in fmm_cases.CASES])
def test_oracle_matches_golden_fixtures(name):
    g = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    assert str(g["digest"]) == fmm_cases.input_digest(name), "input generator drifted"
    for tname, t in (("rho", _abi.RHO), ("drhodt", _abi.DRHODT)):
        sg, keep, kind, cmp_ = fmm_cases.subgrid(name, t)
        if kind == "monopole":
            L, Lc = oracle.monopole_interactions(sg, cmp_)
        else:
            L, Lc = oracle.multipole_interactions(sg)
        assert np.array_equal(L, g[f"L_{tname}"]), (name, tname)
        assert np.array_equal(Lc, g[f"Lc_{tname}"]), (name, tname)


def _cells_xyz():
    i = np.arange(512)
    return np.stack([i // 64, (i // 8) % 8, i % 8], axis=1)


def test_p2p_against_brute_force_point_masses():
    """Leaf sub-grid whose neighbours are leaves: the p2p result equals direct
    point-mass sums over the independently enumerated pair set."""
    sg, keep, kind, cmp_ = fmm_cases.subgrid("leaf_p2p", _abi.RHO)
    lv = keep[-1]
    L, _ = oracle.monopole_interactions(sg, False)
    frame = np.zeros((24, 24, 24))
    for raw in range(27):
        n = lv.neighbors[13, raw]
        if n < 0:
            continue
        d = np.array(_abi.dir_components(raw))
        o = 8 + 8 * d
        frame[o[0]:o[0] + 8, o[1]:o[1] + 8, o[2]:o[2] + 8] = lv.mon[n].reshape(8, 8, 8)
    ref = np.zeros((512, 4))
    dx = lv.dx
    for u, t in enumerate(_cells_xyz() + 8):
        for dd in np.ndindex(11, 11, 11):
            d = np.array(dd) - 5
            if not d.any():
                continue
            p = t + d
            dc = _coarse(p) - _coarse(t)
            if not float((dc ** 2).sum()) < 4.0:
                continue
            m = frame[tuple(p)]
            r = np.sqrt(float((d ** 2).sum()))
            ref[u, 0] += -m / (r * dx)
            ref[u, 1:] += -(d / r ** 3) * m / dx ** 2
    assert np.abs(L[:, :4] - ref).max() <= 1e-12 * np.abs(ref).max()
    assert (L[:, 4:] == 0).all()


def test_multipole_pair_set_against_brute_force():
    """Non-leaf target whose own cells are pure monopoles and whose neighbours
    are leaves: L_0 = sum over far own-cell partners and over every
    (near or far) leaf partner of -m/|X - Y| -- the near/far split and the
    pair set are enumerated independently here."""
    lv = fmm_cases.build_level("c1_nonroot")
    fmm_cases._leafify(lv, [s for s in range(27) if s != 13])
    lv.M[13, :, 1:] = 0.0
    sg, keep = workloads.subgrid_view(lv, 13, type_=_abi.DRHODT)
    L, _ = oracle.multipole_interactions(sg)
    dx = lv.dx
    pos = np.zeros((24, 24, 24, 3))
    mass = np.zeros((24, 24, 24))
    own = np.zeros((24, 24, 24), bool)
    for raw in range(27):
        n = lv.neighbors[13, raw]
        d = np.array(_abi.dir_components(raw))
        o = 8 + 8 * d
        sl = (slice(o[0], o[0] + 8), slice(o[1], o[1] + 8), slice(o[2], o[2] + 8))
        if raw == 13:
            mass[sl] = lv.M[13, :, 0].reshape(8, 8, 8)
            pos[sl] = lv.com[13].reshape(8, 8, 8, 3)
            own[sl] = True
        else:
            mass[sl] = lv.mon[n].reshape(8, 8, 8)
            idx = np.indices((8, 8, 8)).reshape(3, -1).T
            # monopole neighbours sit at geometric centres in the target's frame
            pos[sl] = ((idx + o) * dx + lv.xbase[13] - 8 * dx).reshape(8, 8, 8, 3)
    ref = np.zeros(512)
    for u, t in enumerate(_cells_xyz() + 8):
        X = lv.com[13, u]
        for dd in np.ndindex(11, 11, 11):
            d = np.array(dd) - 5
            if not d.any():
                continue
            p = t + d
            dc = _coarse(p) - _coarse(t)
            if not float((dc ** 2).sum()) < 4.0:
                continue
            far = float((d ** 2).sum()) >= 4.0
            if own[tuple(p)] and not far:
                continue
            ref[u] += -mass[tuple(p)] / np.linalg.norm(X - pos[tuple(p)])
    assert np.abs(L[:, 0] - ref).max() <= 1e-12 * np.abs(ref).max()


def test_root_pair_set_against_brute_force():
    lv = fmm_cases.build_level("c1_root")
    lv.M[13, :, 1:] = 0.0
    sg, keep = workloads.subgrid_view(lv, 13, root=True, type_=_abi.DRHODT)
    L, _ = oracle.multipole_interactions(sg)
    xyz = _cells_xyz()
    m, X = lv.M[13, :, 0], lv.com[13]
    d = xyz[None, :, :] - xyz[:, None, :]
    d2 = (d ** 2).sum(-1)
    keep_pair = d2 >= 4                         # self and the 26 near offsets excluded
    r = np.linalg.norm(X[:, None, :] - X[None, :, :], axis=-1)
    ref = -(np.where(keep_pair, m[None, :] / np.where(r > 0, r, 1.0), 0.0)).sum(1)
    assert keep_pair.sum() == 251496
    assert np.abs(L[:, 0] - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.slow
def test_level_driver_equals_per_subgrid_calls():
    lv = workloads.mixed_level((5, 5, 5), seed=3, leaf_fraction=0.5)
    L, Lc = oracle.level_interactions(lv, nthreads=4)
    for s in lv.targets[:6]:
        sg, keep = workloads.subgrid_view(lv, int(s))
        if lv.kind[s] == _abi.SLOT_LEAF:
            cmp_ = any(lv.kind[n] == _abi.SLOT_NONLEAF for n in lv.neighbors[s] if n >= 0 and n != s)
            a, b = oracle.monopole_interactions(sg, cmp_)
        else:
            a, b = oracle.multipole_interactions(sg)
        assert np.array_equal(a, L[s])


@pytest.mark.skipif(not oracle.ref_simd_available(), reason="vector build of the reference harness not built")
@pytest.mark.parametrize("t", [_abi.RHO, _abi.DRHODT])
def test_vector_reference_baseline_matches_oracle(t):
    """The timed CPU baseline (the reference templates on host-width vectors,
    -O3, z-row walk of the Kokkos host kernel) gives the oracle's result to
    rounding, including targets on the domain edge (empty neighbours)."""
    from metrics import normwise
    lv = workloads.uniform_level((3, 3, 3), ((0, 3),) * 3, seed=41)
    Lr, Lcr = oracle.level_interactions(lv, type_=t)
    Ls, Lcs = oracle.level_interactions_simd(lv, type_=t)
    tg = lv.targets
    assert (lv.neighbors[tg] < 0).any()
    assert normwise(Ls[tg], Lr[tg]) <= 1e-12
    if t == _abi.RHO:
        assert normwise(Lcs[tg], Lcr[tg]) <= 1e-12


def _run_case(kind, lv, sg, t, use_ref=False):
    cmp_ = any(lv.neighbors[13, r] >= 0 and lv.kind[lv.neighbors[13, r]] == _abi.SLOT_NONLEAF
               for r in range(27) if r != 13)
    if kind == "monopole":
        return oracle.monopole_interactions(sg, cmp_, use_ref=use_ref)
    return oracle.multipole_interactions(sg, use_ref=use_ref)


@pytest.mark.parametrize("name", ["c1_nonroot", "mixed_halo", "leaf_p2p", "leaf_p2m"])
@pytest.mark.parametrize("pattern", ["all_shuffled", "subset", "repeats"])
def test_partial_halo_equals_full_halo_with_absent_cells_zeroed(name, pattern):
    """A partial halo (cells handed over in a list, as grid::get_ilist_n_bnd
    orders them) stages exactly what a full halo whose left-out cells hold
    zeros stages: every cell in shuffled order gives the full-halo result bit
    for bit, a subset gives the zeroed full halo's, and with repeated cells
    the later entry wins (update_input overwrites in list order)."""
    lv = fmm_cases.build_level(name)
    kind = dict(fmm_cases.CASES)[name]
    spec = fmm_cases.partial_specs(name)[pattern]
    sg, keep = fmm_cases.partial_subgrid(lv, spec, stale=fmm_cases.stale_prefix(pattern, spec))
    full = lv if pattern == "all_shuffled" else fmm_cases.zeroed_level(lv, spec)
    sgf, keepf = workloads.subgrid_view(full, 13)
    for use_ref in ([False, True] if oracle.ref_available() else [False]):
        L, Lc = _run_case(kind, lv, sg, _abi.RHO, use_ref)
        Lf, Lcf = _run_case(kind, full, sgf, _abi.RHO, use_ref)
        assert np.array_equal(L, Lf) and np.array_equal(Lc, Lcf)
        if pattern != "all_shuffled":
            L0, _ = _run_case(kind, lv, workloads.subgrid_view(lv, 13)[0], _abi.RHO, use_ref)
            assert not np.array_equal(L, L0)      # the left-out cells mattered


@pytest.mark.skipif(not oracle.ref_simd_available(), reason="vector build of the reference harness not built")
def test_force_torque_gate_separates_reordering_from_mistakes():
    """The total force / torque gate (relative to the reference totals, with
    the rounding floor 4 eps sum_i |F_i|) accepts a second FP64 evaluation that
    differs only in summation order and contraction (the -O3 vector build of
    the reference templates against the scalar oracle) on a level whose
    totals cancel 10^3-fold, and rejects planted mistakes of the size a
    dropped or mis-signed term makes."""
    from metrics import force_torque_gate
    lv = workloads.uniform_level((4, 4, 4), ((1, 3),) * 3, seed=77)
    Lr, Lcr = oracle.level_interactions(lv)
    Ls, Lcs = oracle.level_interactions_simd(lv)
    tg = lv.targets
    args = (lv.M[tg, :, 0], lv.com[tg])
    ok, e = force_torque_gate(Ls[tg], Lcs[tg], Lr[tg], Lcr[tg], *args, 1e-12)
    assert ok and e["force_cancellation"] > 100
    assert e["force"] > 0 or e["torque"] > 0        # the two evaluations do differ
    # the angular-momentum correction left out
    ok, _ = force_torque_gate(Ls[tg], 0 * Lcs[tg], Lr[tg], Lcr[tg], *args, 1e-12)
    assert not ok
    # one correction component with its sign flipped
    bad = Lcs[tg].copy()
    bad[..., 2] *= -1
    ok, _ = force_torque_gate(Ls[tg], bad, Lr[tg], Lcr[tg], *args, 1e-12)
    assert not ok
    # one cell's gradient off by one part in 10^9
    badL = Ls[tg].copy()
    badL[0, 100, 1] *= 1 + 1e-9
    ok, _ = force_torque_gate(badL, Lcs[tg], Lr[tg], Lcr[tg], *args, 1e-12)
    assert not ok
