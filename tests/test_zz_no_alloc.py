"""octofmm_compute_multipole_interactions_noalloc (static device workspace, no
allocation) against octofmm_compute_multipole_interactions, both called
through the C-ABI: the same kernels, so the same bits.

Calls run back to back on one context, each later one with less input than
the one before (full halos; then a mixed halo with three empty directions and
leaf neighbours; then partial halos; then the root sub-grid, its own cells
alone), so a region the previous call wrote and this one does not must not
reach the result.  Each result is also checked against the golden fixture
written from the reference templates (normwise 1e-12).
"""
import os

import numpy as np
import pytest

import fmm_cases
from metrics import normwise
from octotiger_amd import _abi, native

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SEQUENCE = ["c1_nonroot", "mixed_halo", "partial_halo", "c1_root"]


def _run(fn, sg, L0, Lc0):
    L, Lc = L0.copy(), Lc0.copy()
    fn(sg, L, Lc)
    return L, Lc


@pytest.mark.parametrize("t", [_abi.RHO, _abi.DRHODT])
def test_noalloc_equals_allocating_entry_point(t):
    ctx = native.Context(theta=0.5, device=0)
    rng = np.random.default_rng(5)
    try:
        for rep in range(2):
            for name in SEQUENCE:
                print(f"noalloc case {name} type {t} pass {rep}", flush=True)
                sg, keep, kind, _ = fmm_cases.subgrid(name, t)
                # L is accumulated: start from a non-zero expansion
                L0 = rng.standard_normal((512, 20)) * 1e-3
                Lc0 = np.full((512, 3), 7.0)
                La, Lca = _run(ctx.multipole_interactions, sg, L0, Lc0)
                Ln, Lcn = _run(ctx.multipole_interactions_noalloc, sg, L0, Lc0)
                assert np.array_equal(Ln, La), name
                assert np.array_equal(Lcn, Lca), name
                g = np.load(os.path.join(GOLDEN, f"{name}.npz"))
                tn = "rho" if t == _abi.RHO else "drhodt"
                assert normwise(Ln - L0, g[f"L_{tn}"]) <= 1e-12
                if t == _abi.RHO:
                    assert normwise(Lcn, g[f"Lc_{tn}"]) <= 1e-12
                else:
                    assert np.array_equal(Lcn, Lc0)      # DRHODT leaves L_c untouched
    finally:
        ctx.close()


def test_noalloc_two_contexts_share_the_device_workspace():
    """Two contexts on one GPU use the one static workspace in turn; each
    call's result is still its own."""
    a = native.Context(theta=0.5, device=0)
    b = native.Context(theta=0.5, device=0)
    try:
        out = {}
        for name, c in (("c1_nonroot", a), ("mixed_halo", b), ("c1_nonroot", b), ("mixed_halo", a)):
            print(f"noalloc two contexts {name}", flush=True)
            sg, keep, kind, _ = fmm_cases.subgrid(name, _abi.RHO)
            L, Lc = np.zeros((512, 20)), np.zeros((512, 3))
            c.multipole_interactions_noalloc(sg, L, Lc)
            if name in out:
                assert np.array_equal(L, out[name][0]) and np.array_equal(Lc, out[name][1])
            out[name] = (L, Lc)
    finally:
        a.close()
        b.close()


def test_noalloc_rejects_oversized_neighbour():
    """A neighbour handing over more cells than a sub-grid holds is refused
    before any device work."""
    ctx = native.Context(theta=0.5, device=0)
    try:
        sg, keep, kind, _ = fmm_cases.subgrid("c1_nonroot", _abi.RHO)
        sg.neighbors[0].count = 513
        with pytest.raises(native.OctofmmError) as e:
            ctx.multipole_interactions_noalloc(sg, np.zeros((512, 20)), np.zeros((512, 3)))
        assert e.value.status == _abi.ERR_INVALID_ARGUMENT
    finally:
        ctx.close()
