"""Benchmark of the FMM gravity interaction pass on one MI355X.

One step = one pass of the level interaction path over the batched uniform
level of 512 non-leaf 8^3 sub-grids (80^3 cells, theta = 0.5, RHO with the
angular-momentum correction): AoS -> partner-record staging plus the
interaction kernel, replayed as one captured hipGraph, kernel variant chosen
by the plan's autotuner.  Inputs are resident in HBM before timing starts.

Prints one JSON line: cell-interactions/s (M2L pairs per second) of the FP64
pass, the roofline of the dominant kernel (FP64 VALU), a CPU baseline (the
reference's own interaction templates -- or the oracle when the reference
tree was not available at build time -- on the host cores, over the same
level), and the mixed-precision form on the same level and on the
4096-sub-grid level with its error against the FP64 path.

It also reports the mixed-kind 2048-sub-grid level and the full adaptive-tree
solve of a polytrope (10,697 sub-grids, M2M + interactions + L2L + forces).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-c3] [--no-c4] [--no-c5] [--cpu-repeats R]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-interactions/s FP64 vs mixed-prec, % VALU peak, rel. force err vs FP64"
FP64_VALU_PEAK_TFLOPS = 78.6     # MI355X FP64 vector, dense
FLOPS_PER_PAIR = {"rho": 407, "drhodt": 273}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-c4", action="store_true", help="skip the 4096-sub-grid FP64/mixed comparison")
    ap.add_argument("--no-c3", action="store_true", help="skip the mixed-kind 2048-sub-grid level")
    ap.add_argument("--no-c5", action="store_true", help="skip the full adaptive-tree solve")
    ap.add_argument("--cpu-repeats", type=int, default=3, help="passes of the CPU baseline over the level (0 = skip)")
    ap.add_argument("--no-check", action="store_true", help="skip the post-run parity spot check")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run 3 device-to-device copies of this many MiB (HBM counter calibration)")
    return ap.parse_args()


def kernel_name(variant):
    """Device kernel the plan's interaction launch runs for a variant
    (lane mapping x 100 + waves per workgroup)."""
    mapping, waves = divmod(variant, 100)
    if mapping == 5:
        return f"level_kernel_split<RHO, {waves}>"
    if mapping == 7:
        return f"level_kernel_dsplit<RHO, {waves}>"
    if mapping >= 3:
        return f"level_kernel_direct<RHO, {waves}, {mapping}>"
    return f"level_kernel<RHO, {waves}, {mapping}>"


def time_plan(plan, stream, steps, warmup, hip):
    for _ in range(warmup):
        plan.execute(stream.handle)
    stream.synchronize()
    e0, e1 = hip.Event(), hip.Event()
    e0.record(stream)
    for _ in range(steps):
        plan.execute(stream.handle)
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / steps
    e0.destroy()
    e1.destroy()
    return ms


def one_pass(plan, dl, stream):
    dl.zero_outputs()
    plan.execute(stream.handle)
    stream.synchronize()
    return dl.L.numpy(), dl.L_c.numpy()


def mixed_errors(lv, Lm, Lcm, Lr, Lcr):
    """Error of the mixed-precision pass against the FP64 pass over all
    target cells: per component normwise (max cell error / max |ref|),
    elementwise relative, and total force / torque."""
    import numpy as np

    from metrics import force_torque_error
    tg = lv.targets
    a = np.concatenate([Lm[tg], Lcm[tg]], axis=-1).reshape(-1, 23)
    b = np.concatenate([Lr[tg], Lcr[tg]], axis=-1).reshape(-1, 23)
    scale = np.abs(b).max(axis=0)
    normwise = np.abs(a - b).max(axis=0) / np.where(scale > 0, scale, 1.0)
    elem = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    ef, et = force_torque_error(Lm[tg], Lcm[tg], Lr[tg], Lcr[tg], lv.M[tg, :, 0], lv.com[tg])
    return {
        "max_normwise_rel_err": float(normwise.max()),
        "normwise_rel_err_per_component": [float(x) for x in normwise],
        "max_elementwise_rel_err": float(elem.max()),
        "median_elementwise_rel_err": float(np.median(elem)),
        "force_rel_err": ef,
        "torque_rel_err": et,
    }


def run_level(ctx, lv, args, hip, _abi, with_mixed=True):
    dl = hip.LevelOnDevice(lv)
    stream = hip.Stream()
    out = {}
    p64 = dl.plan(ctx, type_=_abi.RHO, flags=_abi.PLAN_USE_GRAPH | _abi.PLAN_AUTOTUNE)
    st64 = p64.stats()
    ms64 = time_plan(p64, stream, args.steps, args.warmup, hip)
    stage_ms = p64.time_stages(stream.handle, iters=max(3, min(args.steps, 20)))
    L64, Lc64 = one_pass(p64, dl, stream)
    out["fp64"] = {"ms_per_step": ms64, "stage_ms": stage_ms, "stats": st64}
    if with_mixed:
        pm = dl.plan(ctx, type_=_abi.RHO, precision=_abi.PRECISION_MIXED, mixed_far_d2=4,
                     flags=_abi.PLAN_USE_GRAPH | _abi.PLAN_AUTOTUNE)
        stm = pm.stats()
        msm = time_plan(pm, stream, args.steps, args.warmup, hip)
        stage_m = pm.time_stages(stream.handle, iters=max(3, min(args.steps, 20)))
        Lm, Lcm = one_pass(pm, dl, stream)
        out["mixed"] = {"ms_per_step": msm, "stage_ms": stage_m, "stats": stm,
                        "errors": mixed_errors(lv, Lm, Lcm, L64, Lc64)}
        pm.close()
    p64.close()
    stream.destroy()
    dl.free()
    return out, (L64, Lc64)


def run_tree(ctx, args, hip, _abi, np):
    """Full adaptive-tree solve (upward M2M, interactions of every level, root
    stencil on level 0, downward L2L and leaf force assembly) as one hipGraph,
    FP64 and mixed precision."""
    from octotiger_amd import workloads

    t = workloads.config5()
    dev = hip.TreeOnDevice(t)
    stream = hip.Stream()
    res = {}
    for name, prec in (("fp64", _abi.PRECISION_FP64), ("mixed", _abi.PRECISION_MIXED)):
        plan = dev.plan(ctx, precision=prec, flags=_abi.PLAN_USE_GRAPH | _abi.PLAN_AUTOTUNE)
        steps = max(5, args.steps // 5)
        ms = time_plan(plan, stream, steps, max(1, args.warmup), hip)
        phases = plan.time_phases(stream.handle, iters=3)
        plan.execute(stream.handle)
        stream.synchronize()
        g = np.concatenate([d["g"].numpy()[lv.kind == _abi.SLOT_LEAF].reshape(-1, 4)
                            for lv, d in zip(t.levels, dev.levels)])
        res[name] = {"ms_per_solve": ms, "phases_ms": {"upward_m2m": phases[0],
                                                       "interactions_l2l_forces": phases[1]},
                     "stats": plan.stats(), "g": g}
        plan.close()
    dev.free()
    # Octo-Tiger's own end-to-end regression value: the solid-sphere run
    # (test_problems/sphere/CMakeLists.txt:18) prints "pot 4.470774e-01 3.176879e+00"
    ts = workloads.solid_sphere_tree()
    dsp = hip.TreeOnDevice(ts)
    plan = dsp.plan(ctx, flags=_abi.PLAN_USE_GRAPH | _abi.PLAN_AUTOTUNE)
    ms_sphere = time_plan(plan, stream, max(5, args.steps // 5), max(1, args.warmup), hip)
    plan.execute(stream.handle)
    stream.synchronize()
    l1, l2 = workloads.sphere_pot_norms(ts, [d["g"].numpy()[..., 0] for d in dsp.levels])
    plan.close()
    dsp.free()
    got = f"pot {l1:e} {l2:e}"
    want = "pot {:e} {:e}".format(*workloads.SPHERE_POT_NORMS)
    sphere = {"workload": f"solid sphere, max_level 3, {ts.nnodes} sub-grids, nodes per level "
                          f"{ts.info['nodes_per_level']}", "ms_per_solve": ms_sphere,
              "pot_norms": got, "reference": want, "matches_reference": got == want}
    st = res["fp64"]["stats"]
    pairs = st["far_pairs"] + st["near_pairs"] + st["p2p_pairs"] + st["p2m_pairs"] + st["root_pairs"]
    ga, gb = res["fp64"]["g"], res["mixed"]["g"]
    ef = float(np.sqrt(((gb[:, 1:] - ga[:, 1:]) ** 2).sum() / (ga[:, 1:] ** 2).sum()))
    ephi = float(np.abs(gb[:, 0] - ga[:, 0]).max() / np.abs(ga[:, 0]).max())
    inter = res["fp64"]["phases_ms"]["interactions_l2l_forces"]
    return {
        "workload": f"{t.name}: n=1.5 polytrope (radius 0.3, 1% log-normal noise), {t.nnodes} sub-grids, "
                    f"nodes per level {t.info['nodes_per_level']}, leaves per level {t.info['leaves_per_level']}",
        "fp64": {"ms_per_solve": res["fp64"]["ms_per_solve"], "phases_ms": res["fp64"]["phases_ms"],
                 "cell_interactions_per_s": pairs / (res["fp64"]["ms_per_solve"] * 1e-3),
                 "interaction_valu_frac": st["flops"] / (inter * 1e-3) / 1e12 / FP64_VALU_PEAK_TFLOPS},
        "mixed": {"ms_per_solve": res["mixed"]["ms_per_solve"], "phases_ms": res["mixed"]["phases_ms"],
                  "speedup_vs_fp64": res["fp64"]["ms_per_solve"] / res["mixed"]["ms_per_solve"],
                  "fp32_far_pairs": int(res["mixed"]["stats"]["mixed_far_pairs"]),
                  "leaf_force_rms_rel_err_vs_fp64": ef, "leaf_phi_max_rel_err_vs_fp64": ephi},
        "pairs": {k: int(st[k]) for k in ("far_pairs", "near_pairs", "p2p_pairs", "p2m_pairs", "root_pairs")},
        "nodes": int(st["nodes"]), "leaves": int(st["leaves"]),
        "solid_sphere_regression": sphere,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the level pass is single-GPU "
              "in this build", file=sys.stderr)
        sys.exit(2)
    import numpy as np

    from octotiger_amd import _abi, hip, native, workloads

    sys.path.insert(0, os.path.join(ROOT, "tests"))

    if hip.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        sys.exit(3)
    ctx = native.Context(theta=0.5, device=0)
    if args.copy_calibration_mib > 0:
        nb = args.copy_calibration_mib << 20
        src, dst = hip.DeviceArray((nb // 8,)), hip.DeviceArray((nb // 8,))
        src.zero()
        for _ in range(3):
            hip.check(hip.rt().hipMemcpy(dst.ptr, src.ptr, nb, hip.D2D), "hipMemcpy D2D")
        hip.synchronize()
        src.free()
        dst.free()

    lv = workloads.config2()
    t0 = time.perf_counter()
    c2, (L64, Lc64) = run_level(ctx, lv, args, hip, _abi)
    wall = time.perf_counter() - t0
    st = c2["fp64"]["stats"]
    pairs = st["far_pairs"] + st["near_pairs"]
    ms_step = c2["fp64"]["ms_per_step"]
    value = pairs / (ms_step * 1e-3)
    k_ms = c2["fp64"]["stage_ms"][1]
    achieved = st["flops"] / (k_ms * 1e-3) / 1e12

    traffic = None
    # HBM bytes per launch of this kernel variant, from the calibrated
    # FETCH_SIZE / WRITE_SIZE counter runs kept beside the profiles
    tpath = os.path.join(ROOT, "profiles", f"traffic_{lv.name}_v{int(st['variant'])}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    check = None
    if not args.no_check:
        import oracle
        sample = lv.targets[:: max(1, len(lv.targets) // 8)][:8]
        Lr, Lcr = oracle.level_interactions(lv, targets=sample)
        err = 0.0
        for a, b in ((L64[sample], Lr[sample]), (Lc64[sample], Lcr[sample])):
            a2, b2 = a.reshape(-1, a.shape[-1]), b.reshape(-1, b.shape[-1])
            for c in range(b2.shape[1]):
                err = max(err, float(np.abs(a2[:, c] - b2[:, c]).max() / np.abs(b2[:, c]).max()))
        check = {"targets": int(len(sample)), "max_normwise_rel_err_vs_oracle": err}

    cpu = None
    if args.cpu_repeats > 0:
        import oracle
        use_ref = oracle.ref_available()
        ncores = min(16, os.cpu_count() or 1)
        t = time.perf_counter()
        for _ in range(args.cpu_repeats):
            oracle.level_interactions(lv, nthreads=ncores, use_ref=use_ref)
        dt = (time.perf_counter() - t) / args.cpu_repeats
        cpu = {"value": pairs / dt, "unit": "cell-interactions/s", "cores": ncores,
               "kind": "reference" if use_ref else "port",
               "sample": (f"the whole level ({len(lv.targets)} sub-grids, {pairs} pairs) x {args.cpu_repeats} passes, "
                          f"{dt:.3f} s per pass on {ncores} threads = {dt * ncores * args.cpu_repeats:.1f} core-s; "
                          + ("Octo-Tiger's own FP64 interaction templates (compute_kernel_templates.hpp) compiled "
                             "with g++ -O2, walked in the Kokkos host-kernel order with SIMD width 1"
                             if use_ref else "FP64 scalar oracle (restated Kokkos host-kernel walk)")
                          + ", one sub-grid per OpenMP task")}
        # one core: the first 32 targets of the level on one thread
        one = lv.targets[:32]
        t = time.perf_counter()
        oracle.level_interactions(lv, targets=one, nthreads=1, use_ref=use_ref)
        dt1 = time.perf_counter() - t
        pairs1 = pairs * len(one) / len(lv.targets)
        cpu["single_core_value"] = pairs1 / dt1
        cpu["single_core_sample"] = f"{len(one)} sub-grids ({int(pairs1)} pairs) on 1 thread, {dt1:.3f} s"

    mixed = {}
    m2 = c2["mixed"]
    mixed["uniform-level-512"] = {
        "value": pairs / (m2["ms_per_step"] * 1e-3), "ms_per_step": m2["ms_per_step"],
        "speedup_vs_fp64": ms_step / m2["ms_per_step"], "kernel_ms": m2["stage_ms"][1],
        "fp32_far_pairs": m2["stats"]["mixed_far_pairs"], "errors": m2["errors"]}
    if not args.no_c4:
        lv4 = workloads.config4()
        c4, _ = run_level(ctx, lv4, args, hip, _abi)
        s4 = c4["fp64"]["stats"]
        p4 = s4["far_pairs"] + s4["near_pairs"]
        mixed["uniform-level-4096"] = {
            "fp64_value": p4 / (c4["fp64"]["ms_per_step"] * 1e-3), "fp64_ms_per_step": c4["fp64"]["ms_per_step"],
            "fp64_kernel_ms": c4["fp64"]["stage_ms"][1],
            "fp64_valu_frac": s4["flops"] / (c4["fp64"]["stage_ms"][1] * 1e-3) / 1e12 / FP64_VALU_PEAK_TFLOPS,
            "value": p4 / (c4["mixed"]["ms_per_step"] * 1e-3), "ms_per_step": c4["mixed"]["ms_per_step"],
            "kernel_ms": c4["mixed"]["stage_ms"][1],
            "speedup_vs_fp64": c4["fp64"]["ms_per_step"] / c4["mixed"]["ms_per_step"],
            "fp32_far_pairs": c4["mixed"]["stats"]["mixed_far_pairs"], "errors": c4["mixed"]["errors"]}

    kinds = None
    if not args.no_c3:
        lv3 = workloads.config3()
        c3, _ = run_level(ctx, lv3, args, hip, _abi, with_mixed=False)
        s3 = c3["fp64"]["stats"]
        ms3 = c3["fp64"]["ms_per_step"]
        p3 = s3["far_pairs"] + s3["near_pairs"] + s3["p2p_pairs"] + s3["p2m_pairs"]
        kinds = {"workload": f"{lv3.name}: {len(lv3.targets)} sub-grids, {lv3.info}",
                 "ms_per_step": ms3, "kernel_ms": c3["fp64"]["stage_ms"][1],
                 "cell_interactions_per_s": p3 / (ms3 * 1e-3),
                 "pairs": {k: int(s3[k]) for k in ("far_pairs", "near_pairs", "p2p_pairs", "p2m_pairs")},
                 "valu_frac": s3["flops"] / (c3["fp64"]["stage_ms"][1] * 1e-3) / 1e12 / FP64_VALU_PEAK_TFLOPS,
                 "kernel_variant": int(s3["variant"])}

    # dense vector FMA rates of this device: FP64, and FP32 as all device code
    # here is built (packed FP32 off), the rate the mixed-precision pairs get
    fma_rate = {"fp64": ctx.measure_fma_rate(_abi.PRECISION_FP64),
                "fp32_unpacked": ctx.measure_fma_rate(_abi.PRECISION_MIXED)}

    tree = None
    if not args.no_c5:
        tree = run_tree(ctx, args, hip, _abi, np)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{lv.name}: {len(lv.targets)} non-leaf 8^3 sub-grids with 26-neighbour halos "
                        f"({lv.nslots} slots), theta=0.5, RHO (angular-momentum correction on), one "
                        f"hipGraph per pass",
            "subgrids": int(len(lv.targets)),
            "far_pairs_per_pass": int(st["far_pairs"]),
            "flops_per_pair": FLOPS_PER_PAIR["rho"],
            "precision": "fp64",
            "kernel_variant": int(st["variant"]),
        },
        "roofline": {
            "bound": "valu",
            "kernel": kernel_name(int(st["variant"])),
            "achieved": achieved,
            "peak": FP64_VALU_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_VALU_PEAK_TFLOPS,
            "traffic": traffic,
            "kernel_ms": k_ms,
            "staging_ms": c2["fp64"]["stage_ms"][0],
            # the dense FP64 FMA rate of this box, re-measured (octofmm_measure_fma_rate)
            "measured_peak": fma_rate["fp64"],
            "frac_of_measured_peak": achieved / fma_rate["fp64"],
        },
        "cpu_baseline": cpu,
        "mixed_precision": mixed,
        "fma_rate_tflops": fma_rate,
        "mixed_kind_level": kinds,
        "tree_solve": tree,
        "parity_check": check,
        "wall_s_c2": wall,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
