"""The 4096-sub-grid uniform level (16^3 non-leaf targets in an 18^3-slot
block, 144^3 cells, theta = 0.5, RHO) at full size against the CPU FP64
oracle, in both precisions, each plan autotuned as the bench runs it.

Gates:
  FP64   per component normwise <= 1e-12; total force and torque of all target
         cells within 1e-12 of the oracle's totals (|F - F_ref| / |F_ref|),
         or within the rounding floor 4 eps sum_i |F_i| where that is larger
         (tests/metrics.py).
  mixed  FP32 pair arithmetic for far pairs with |d|^2 > 4 (cell units; the
         mapping-7 kernels take whole coarse offsets whose far pairs all
         qualify), FP64 accumulation and FP64 angular-momentum correction: total force
         and torque within 1e-6 of the oracle's totals, per component
         normwise <= 1e-5, L_c normwise <= 1e-12.
"""
import numpy as np
import pytest

import oracle
from metrics import describe, force_torque_gate, normwise
from octotiger_amd import _abi, hip, native, workloads

pytestmark = pytest.mark.gpu

TOL = 1e-12
MIXED_FORCE_TOL = 1e-6
MIXED_COMPONENT_TOL = 1e-5


@pytest.fixture(scope="module")
def c4():
    lv = workloads.config4()
    print(f"config4: computing the oracle over {len(lv.targets)} sub-grids", flush=True)
    Lr, Lcr = oracle.level_interactions(lv)
    yield lv, Lr, Lcr


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(theta=0.5, device=0)
    yield c
    c.close()


def run(ctx, lv, precision):
    dl = hip.LevelOnDevice(lv)
    stream = hip.Stream()
    plan = dl.plan(ctx, type_=_abi.RHO, precision=precision, mixed_far_d2=4,
                   flags=_abi.PLAN_USE_GRAPH | _abi.PLAN_AUTOTUNE)
    st = plan.stats()
    plan.execute(stream.handle)
    stream.synchronize()
    L, Lc = dl.L.numpy(), dl.L_c.numpy()
    plan.close()
    stream.destroy()
    dl.free()
    return L, Lc, st


@pytest.mark.parametrize("precision", [_abi.PRECISION_FP64, _abi.PRECISION_MIXED])
def test_level_config4_full_size(ctx, c4, precision):
    lv, Lr, Lcr = c4
    print(f"config4 precision {precision}", flush=True)
    L, Lc, st = run(ctx, lv, precision)
    assert st["far_pairs"] == 4096 * 96768 and st["near_pairs"] == 0
    tg = lv.targets
    en = normwise(L[tg], Lr[tg])
    ec = normwise(Lc[tg], Lcr[tg])
    tol = TOL if precision == _abi.PRECISION_FP64 else MIXED_FORCE_TOL
    ok, ft = force_torque_gate(L[tg], Lc[tg], Lr[tg], Lcr[tg], lv.M[tg, :, 0], lv.com[tg], tol)
    print(f"config4 precision {precision} variant {st['variant']}: normwise {en:.3e} L_c {ec:.3e} "
          f"{describe(ft)}", flush=True)
    assert ec <= TOL
    if precision == _abi.PRECISION_FP64:
        assert en <= TOL
        assert ok
    else:
        # per pair (mapping 0) 96.8% of the far pairs are FP32, per coarse
        # offset (mapping 7: the face offsets stay FP64) 81%
        assert st["mixed_far_pairs"] > (0.8 if st["variant"] // 100 == 7 else 0.96) * st["far_pairs"]
        assert 0 < en <= MIXED_COMPONENT_TOL
        assert ok


def test_level_config4_drhodt_full_size(ctx, c4):
    """DRHODT, the first half of the production step (compute_fmm(DRHODT) then
    compute_fmm(RHO), src/node_server_actions_3.cpp:591-592), at full size:
    per component normwise <= 1e-12 against the oracle, autotuned."""
    lv = c4[0]
    print("config4 DRHODT: computing the oracle", flush=True)
    Lr, _ = oracle.level_interactions(lv, type_=_abi.DRHODT)
    dl = hip.LevelOnDevice(lv)
    stream = hip.Stream()
    plan = dl.plan(ctx, type_=_abi.DRHODT, flags=_abi.PLAN_USE_GRAPH | _abi.PLAN_AUTOTUNE)
    st = plan.stats()
    plan.execute(stream.handle)
    stream.synchronize()
    L = dl.L.numpy()
    plan.close()
    stream.destroy()
    dl.free()
    tg = lv.targets
    en = normwise(L[tg], Lr[tg])
    print(f"config4 DRHODT variant {st['variant']}: normwise {en:.3e}", flush=True)
    assert en <= TOL
