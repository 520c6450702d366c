"""Benchmark of the FMM gravity interaction pass on one MI355X.

One step = one pass of the level interaction path over the batched uniform
level of 4096 non-leaf 8^3 sub-grids (a 16^3 block inside a one-slot halo
shell: 18^3 slots, 144^3 cells, theta = 0.5, RHO with the angular-momentum
correction): AoS -> partner-record staging plus the interaction kernel,
replayed as one captured hipGraph, kernel variant chosen by the plan's
autotuner.  Inputs are resident in HBM before timing starts.

Prints one JSON line: cell-interactions/s (M2L pairs per second) of the FP64
pass, the roofline of the dominant kernel (FP64 VALU), a CPU baseline (the
reference's own interaction templates on host-width vectors, -O3, walked as
the Kokkos host kernel walks them, on the host cores over a sample of the same
level), the mixed-precision form on the same level with its errors against
the FP64 oracle, and the production step (DRHODT then RHO, as
node_server::compute_fmm is called each step).

Extra keys: the 512-sub-grid level, the mixed-kind 2048-sub-grid level and
the full adaptive-tree solve of a polytrope (10,697 sub-grids, M2M +
interactions + L2L + forces), each with its production step.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-c2] [--no-c3] [--no-c5]
                    [--cpu-repeats R] [--no-check]
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
HBM_PEAK_TBS = 8.0               # MI355X HBM3E
FLOPS_PER_PAIR = {"rho": 407, "drhodt": 273}
CPU_THREADS_MAX = 16             # host threads one GPU's share of the box gives a process


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-c2", action="store_true", help="skip the 512-sub-grid level")
    ap.add_argument("--no-c3", action="store_true", help="skip the mixed-kind 2048-sub-grid level")
    ap.add_argument("--no-c5", action="store_true", help="skip the full adaptive-tree solve")
    ap.add_argument("--cpu-repeats", type=int, default=2, help="passes of the CPU baseline over its sample (0 = skip)")
    ap.add_argument("--no-check", action="store_true", help="skip the oracle comparison of the headline level")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run 3 device-to-device copies of this many MiB (HBM counter calibration)")
    return ap.parse_args()


def staging_bytes(lv, stats):
    """Algorithmic bytes of one FP64 staging pass over every slot of a
    uniform level: per cell 184 bytes read (M, com) and a 192-byte record
    written."""
    return lv.nslots * 512 * (184 + 192)


def kernel_name(variant, rho=True):
    """Device kernel the plan's interaction launch runs for a variant
    (lane mapping x 100 + waves per workgroup)."""
    t = "RHO" if rho else "DRHODT"
    mapping, waves = divmod(variant, 100)
    if mapping == 5:
        return f"level_kernel_split<{t}, {waves}>"
    if mapping == 7:
        return f"level_kernel_dsplit<{t}, {waves}>"
    if mapping == 8:
        return f"level_kernel_dqueue<{t}, {waves}>"
    if mapping == 3:
        return f"level_kernel_direct<{t}, {waves}>"
    return f"level_kernel<{t}, {waves}, {mapping}>"


# untimed back-to-back replays of a tree solve before it is timed (its
# steady rate; the headline pass keeps the --warmup steps of the contract)
TREE_WARM_MS = 150.0


def time_plans(plans, stream, steps, warmup, hip, warm_ms=0.0):
    """Milliseconds per step of the plans executed back to back on one stream,
    after `warmup` untimed steps and, with warm_ms, further untimed steps
    until at least warm_ms of them have run (a solve of a few ms reaches its
    steady rate only after some tens of ms of back-to-back replays)."""
    for _ in range(warmup):
        for p in plans:
            p.execute(stream.handle)
    stream.synchronize()
    if warm_ms > 0.0:
        w0, w1 = hip.Event(), hip.Event()
        spent = 0.0
        while spent < warm_ms:
            w0.record(stream)
            for _ in range(10):
                for p in plans:
                    p.execute(stream.handle)
            w1.record(stream)
            w1.synchronize()
            spent += w0.elapsed_ms(w1)
        w0.destroy()
        w1.destroy()
    e0, e1 = hip.Event(), hip.Event()
    e0.record(stream)
    for _ in range(steps):
        for p in plans:
            p.execute(stream.handle)
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


def errors_vs(lv, L, Lc, Lr, Lcr):
    """Errors of a pass against a reference over all target cells: per
    component normwise (max cell error / max |ref|), elementwise relative, and
    the total force / torque relative to the reference's totals."""
    import numpy as np

    from metrics import force_torque_errors, normwise
    tg = lv.targets
    a = np.concatenate([L[tg], Lc[tg]], axis=-1).reshape(-1, 23)
    b = np.concatenate([Lr[tg], Lcr[tg]], axis=-1).reshape(-1, 23)
    scale = np.abs(b).max(axis=0)
    nw = np.abs(a - b).max(axis=0) / np.where(scale > 0, scale, 1.0)
    elem = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    ft = force_torque_errors(L[tg], Lc[tg], Lr[tg], Lcr[tg], lv.M[tg, :, 0], lv.com[tg])
    # per-cell acceleration g_i = -(L_i[1..3] + L_c,i): RMS error over RMS value
    ga = -(L[tg][..., 1:4] + Lc[tg])
    gb = -(Lr[tg][..., 1:4] + Lcr[tg])
    g_rms = float(np.sqrt(((ga - gb) ** 2).sum() / max((gb ** 2).sum(), 1e-300)))
    return {
        "max_normwise_rel_err": float(nw.max()),
        "normwise_rel_err_L": float(nw[:20].max()),
        "normwise_rel_err_L_c": float(nw[20:].max()),
        "normwise_rel_err_per_component": [float(x) for x in nw],
        "max_elementwise_rel_err": float(elem.max()),
        "median_elementwise_rel_err": float(np.median(elem)),
        "force_rel_err": ft["force"],
        "torque_rel_err": ft["torque"],
        "force_rel_err_rounding_floor": ft["force_floor"],
        "torque_rel_err_rounding_floor": ft["torque_floor"],
        "force_rel_err_per_cell_sum": ft["force_per_cell_sum"],
        "torque_rel_err_per_cell_sum": ft["torque_per_cell_sum"],
        "per_cell_force_rms_rel_err": g_rms,
    }


def run_level(ctx, lv, args, hip, _abi, ref=None, with_mixed=True, with_drhodt=True):
    """FP64 (and mixed) RHO passes of a level, the DRHODT pass, and the
    production step (DRHODT then RHO); errors against ref = (L, L_c) of the
    oracle when given."""
    dl = hip.LevelOnDevice(lv)
    stream = hip.Stream()
    out = {}
    flags = _abi.PLAN_USE_GRAPH | _abi.PLAN_AUTOTUNE
    iters = max(3, min(args.steps, 20))
    t0 = time.perf_counter()
    p64 = dl.plan(ctx, type_=_abi.RHO, flags=flags)
    create_s = time.perf_counter() - t0
    st64 = p64.stats()
    ms64 = time_plans([p64], stream, args.steps, args.warmup, hip)
    stage_ms = p64.time_stages(stream.handle, iters=iters)
    L64, Lc64 = one_pass(p64, dl, stream)
    out["fp64"] = {"ms_per_step": ms64, "stage_ms": stage_ms, "stats": st64, "plan_create_s": create_s}
    if ref is not None:
        out["fp64"]["errors_vs_oracle"] = errors_vs(lv, L64, Lc64, *ref)
    plans = {"fp64": p64}
    if with_mixed:
        pm = dl.plan(ctx, type_=_abi.RHO, precision=_abi.PRECISION_MIXED, mixed_far_d2=4, flags=flags)
        stm = pm.stats()
        msm = time_plans([pm], stream, args.steps, args.warmup, hip)
        stage_m = pm.time_stages(stream.handle, iters=iters)
        Lm, Lcm = one_pass(pm, dl, stream)
        out["mixed"] = {"ms_per_step": msm, "stage_ms": stage_m, "stats": stm,
                        "errors_vs_fp64_pass": errors_vs(lv, Lm, Lcm, L64, Lc64)}
        if ref is not None:
            out["mixed"]["errors_vs_oracle"] = errors_vs(lv, Lm, Lcm, *ref)
        plans["mixed"] = pm
    if with_drhodt:
        # node_server::compute_fmm(DRHODT) then compute_fmm(RHO) every step
        # (src/node_server_actions_3.cpp:591-592)
        step = {}
        for name, prec in (("fp64", _abi.PRECISION_FP64), ("mixed", _abi.PRECISION_MIXED)):
            if name not in plans:
                continue
            pd = dl.plan(ctx, type_=_abi.DRHODT, precision=prec, mixed_far_d2=4, flags=flags)
            std = pd.stats()
            msd = time_plans([pd], stream, args.steps, args.warmup, hip)
            ms2 = time_plans([pd, plans[name]], stream, args.steps, args.warmup, hip)
            pairs = sum(int(sx[k]) for sx in (std, plans[name].stats())
                        for k in ("far_pairs", "near_pairs", "p2p_pairs", "p2m_pairs"))
            step[name] = {"drhodt_ms": msd, "drhodt_kernel_variant": int(std["variant"]),
                          "drhodt_then_rho_ms": ms2, "cell_interactions_per_s": pairs / (ms2 * 1e-3),
                          "flops": int(std["flops"] + plans[name].stats()["flops"])}
            pd.close()
        if "mixed" in step:
            step["mixed"]["speedup_vs_fp64"] = step["fp64"]["drhodt_then_rho_ms"] / step["mixed"]["drhodt_then_rho_ms"]
            step["mixed"]["drhodt_speedup_vs_fp64"] = step["fp64"]["drhodt_ms"] / step["mixed"]["drhodt_ms"]
        out["production_step"] = step
    for p in plans.values():
        p.close()
    stream.destroy()
    dl.free()
    return out


def run_tree(ctx, args, hip, _abi, np):
    """Full adaptive-tree solve (upward M2M, interactions of every level, root
    stencil on level 0, downward L2L and leaf force assembly) as one hipGraph,
    FP64 and mixed precision, RHO and the production step DRHODT + RHO."""
    from octotiger_amd import workloads

    t = workloads.config5()
    dev = hip.TreeOnDevice(t)
    stream = hip.Stream()
    flags = _abi.PLAN_USE_GRAPH | _abi.PLAN_AUTOTUNE
    steps = max(10, args.steps // 2)
    res = {}
    for name, prec in (("fp64", _abi.PRECISION_FP64), ("mixed", _abi.PRECISION_MIXED)):
        t0 = time.perf_counter()
        plan = dev.plan(ctx, precision=prec, flags=flags)
        create_s = time.perf_counter() - t0
        ms = time_plans([plan], stream, steps, max(1, args.warmup), hip, warm_ms=TREE_WARM_MS)
        phases = plan.time_phases(stream.handle, iters=3)
        plan.execute(stream.handle)
        stream.synchronize()
        g = np.concatenate([d["g"].numpy()[lv.kind == _abi.SLOT_LEAF].reshape(-1, 4)
                            for lv, d in zip(t.levels, dev.levels)])
        # production step: DRHODT (reusing this RHO solve's centres of mass), then RHO
        pd = dev.plan(ctx, type_=_abi.DRHODT, precision=prec, flags=flags)
        msd = time_plans([pd], stream, steps, 1, hip, warm_ms=TREE_WARM_MS)
        ms2 = time_plans([pd, plan], stream, steps, 1, hip, warm_ms=TREE_WARM_MS)
        pd.close()
        res[name] = {"ms_per_solve": ms, "phases_ms": {"upward_m2m": phases[0],
                                                       "interactions_l2l_forces": phases[1]},
                     "stats": plan.stats(), "g": g, "plan_create_s": create_s,
                     "production_step": {"drhodt_ms": msd, "drhodt_then_rho_ms": ms2}}
        plan.close()
    # the FP64 solve with leaf cells writing g (and L_c) only, no L rows
    # (OCTOFMM_TREE_LEAF_G_ONLY)
    plan = dev.plan(ctx, flags=flags | _abi.TREE_LEAF_G_ONLY)
    res["fp64"]["leaf_g_only_ms_per_solve"] = time_plans([plan], stream, steps, max(1, args.warmup), hip,
                                                         warm_ms=TREE_WARM_MS)
    plan.close()
    dev.free()
    # Octo-Tiger's own end-to-end regression value: the solid-sphere run
    # (test_problems/sphere/CMakeLists.txt:18) prints "pot 4.470774e-01 3.176879e+00"
    ts = workloads.solid_sphere_tree()
    dsp = hip.TreeOnDevice(ts)
    plan = dsp.plan(ctx, flags=flags)
    ms_sphere = time_plans([plan], stream, steps, max(1, args.warmup), hip, warm_ms=TREE_WARM_MS)
    plan.execute(stream.handle)
    stream.synchronize()
    l1, l2 = workloads.sphere_pot_norms(ts, [d["g"].numpy()[..., 0] for d in dsp.levels])
    plan.close()
    dsp.free()
    stream.destroy()
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
                 "leaf_g_only_ms_per_solve": res["fp64"]["leaf_g_only_ms_per_solve"],
                 "cell_interactions_per_s": pairs / (res["fp64"]["ms_per_solve"] * 1e-3),
                 "interaction_valu_frac": st["flops"] / (inter * 1e-3) / 1e12 / FP64_VALU_PEAK_TFLOPS,
                 "plan_create_s": res["fp64"]["plan_create_s"],
                 "production_step": res["fp64"]["production_step"]},
        "mixed": {"ms_per_solve": res["mixed"]["ms_per_solve"], "phases_ms": res["mixed"]["phases_ms"],
                  "speedup_vs_fp64": res["fp64"]["ms_per_solve"] / res["mixed"]["ms_per_solve"],
                  "fp32_far_pairs": int(res["mixed"]["stats"]["mixed_far_pairs"]),
                  "leaf_force_rms_rel_err_vs_fp64": ef, "leaf_phi_max_rel_err_vs_fp64": ephi,
                  "plan_create_s": res["mixed"]["plan_create_s"],
                  "production_step": res["mixed"]["production_step"]},
        "pairs": {k: int(st[k]) for k in ("far_pairs", "near_pairs", "p2p_pairs", "p2m_pairs", "root_pairs")},
        "nodes": int(st["nodes"]), "leaves": int(st["leaves"]),
        "solid_sphere_regression": sphere,
    }


def host_info():
    """Logical CPUs, physical cores and model of the host, from /proc/cpuinfo."""
    info = {"logical_cpus": os.cpu_count(), "usable_cpus": len(os.sched_getaffinity(0))}
    cores, model = set(), None
    try:
        phys = core = None
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("physical id"):
                    phys = ln.split(":")[1].strip()
                elif ln.startswith("core id"):
                    core = ln.split(":")[1].strip()
                elif ln.startswith("model name") and model is None:
                    model = ln.split(":", 1)[1].strip()
                elif not ln.strip():
                    if core is not None:
                        cores.add((phys, core))
                    phys = core = None
    except OSError:
        pass
    info["physical_cores"] = len(cores) or None
    info["model"] = model
    return info


def cpu_baseline(lv, pairs_per_subgrid, args):
    """The reference's own interaction templates, vectorised across z-rows as
    the Kokkos host kernel does, -O3, one sub-grid per OpenMP task, timed on
    the host over a sample of the headline level."""
    import numpy as np

    import oracle
    hi = host_info()
    threads = max(1, min(CPU_THREADS_MAX, hi["usable_cpus"]))
    simd = oracle.ref_simd_available()
    if simd:
        run = lambda tg, nt: oracle.level_interactions_simd(lv, targets=tg, nthreads=nt)  # noqa: E731
        kind, width = "reference", oracle.host_simd_width()
        how = (f"Octo-Tiger's own FP64 interaction templates (compute_kernel_templates.hpp) instantiated on "
               f"{width}-wide double vectors, walked in the Kokkos host-kernel order (z-rows of 8 cells as "
               f"vectors), g++ -O3 -march=x86-64-v{4 if width == 8 else 3}")
    else:
        use_ref = oracle.ref_available()
        run = lambda tg, nt: oracle.level_interactions(lv, targets=tg, nthreads=nt, use_ref=use_ref)  # noqa: E731
        kind = "reference" if use_ref else "port"
        how = "scalar walk, -O2" + (" (reference templates)" if use_ref else " (oracle)")
    # sample: the whole level, every sub-grid once per pass
    sample = np.ascontiguousarray(lv.targets)
    pairs = pairs_per_subgrid * len(sample)
    run(sample[:threads], threads)                       # warm-up
    times = []
    for _ in range(max(1, args.cpu_repeats)):
        t = time.perf_counter()
        run(sample, threads)
        times.append(time.perf_counter() - t)
    dt = float(np.median(times))
    one = np.ascontiguousarray(lv.targets[:32])
    t = time.perf_counter()
    run(one, 1)
    dt1 = time.perf_counter() - t
    return {
        "value": pairs / dt, "unit": "cell-interactions/s", "cores": threads, "kind": kind,
        "sample": (f"all {len(sample)} sub-grids of the level, {pairs} pairs, "
                   f"median of {len(times)} passes: {dt:.3f} s on {threads} threads = {dt * threads:.1f} core-s "
                   f"per pass; {how}; one sub-grid per OpenMP task"),
        "host": hi,
        "threads_note": f"{threads} threads: the host CPUs a one-GPU process is given on this box "
                        f"(at most {CPU_THREADS_MAX})",
        "single_core_value": pairs_per_subgrid * len(one) / dt1,
        "single_core_sample": f"{len(one)} sub-grids ({pairs_per_subgrid * len(one)} pairs) on 1 thread, {dt1:.3f} s",
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

    # headline: the 4096-sub-grid level
    lv = workloads.config4()
    ref = None
    if not args.no_check:
        import oracle
        t = time.perf_counter()
        ref = oracle.level_interactions(lv, nthreads=min(CPU_THREADS_MAX, len(os.sched_getaffinity(0))))
        oracle_s = time.perf_counter() - t
    t0 = time.perf_counter()
    c4 = run_level(ctx, lv, args, hip, _abi, ref=ref)
    wall = time.perf_counter() - t0
    st = c4["fp64"]["stats"]
    pairs = st["far_pairs"] + st["near_pairs"]
    ms_step = c4["fp64"]["ms_per_step"]
    value = pairs / (ms_step * 1e-3)
    k_ms = c4["fp64"]["stage_ms"][1]
    achieved = st["flops"] / (k_ms * 1e-3) / 1e12

    traffic = None
    # HBM bytes per launch of this kernel variant, from the calibrated
    # FETCH_SIZE / WRITE_SIZE counter runs kept beside the profiles
    tpath = os.path.join(ROOT, "profiles", f"traffic_{lv.name}_v{int(st['variant'])}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    cpu = cpu_baseline(lv, pairs // len(lv.targets), args) if args.cpu_repeats > 0 else None

    m4 = c4["mixed"]
    mixed = {"uniform-level-4096": {
        "value": pairs / (m4["ms_per_step"] * 1e-3), "ms_per_step": m4["ms_per_step"],
        "speedup_vs_fp64": ms_step / m4["ms_per_step"], "kernel_ms": m4["stage_ms"][1],
        "kernel_speedup_vs_fp64": k_ms / m4["stage_ms"][1], "kernel_variant": int(m4["stats"]["variant"]),
        "fp32_far_pairs": int(m4["stats"]["mixed_far_pairs"]),
        "errors_vs_oracle": m4.get("errors_vs_oracle"), "errors_vs_fp64_pass": m4["errors_vs_fp64_pass"]}}

    extra = {}
    def oracle_ref(level):
        if args.no_check:
            return None
        import oracle
        return oracle.level_interactions(level, nthreads=min(CPU_THREADS_MAX, len(os.sched_getaffinity(0))))

    def mixed_summary(res):
        # the mixed pass of a level against the oracle (or, without one, the
        # level's FP64 pass), with its speed against the FP64 pass
        m = res["mixed"]
        return {"speedup_vs_fp64": res["fp64"]["ms_per_step"] / m["ms_per_step"],
                "kernel_speedup_vs_fp64": res["fp64"]["stage_ms"][1] / m["stage_ms"][1],
                "fp32_far_pair_share": m["stats"]["mixed_far_pairs"] / max(1, m["stats"]["far_pairs"]),
                "kernel_variant": int(m["stats"]["variant"]),
                "errors": {k: ([float(f"{x:.3g}") for x in v] if k == "normwise_rel_err_per_component" else v)
                           for k, v in (m.get("errors_vs_oracle") or m["errors_vs_fp64_pass"]).items()},
                "errors_against": "oracle" if m.get("errors_vs_oracle") else "fp64 pass"}

    per_config = {}
    # the single sub-grid of the CPU-runnable configuration, as a one-target level
    lv1 = workloads.config1()
    c1 = run_level(ctx, lv1, args, hip, _abi, ref=oracle_ref(lv1), with_drhodt=False)
    per_config["single-subgrid (C1)"] = mixed_summary(c1)
    per_config["uniform-level-4096 (C4)"] = mixed_summary(c4)

    if not args.no_c2:
        lv2 = workloads.config2()
        c2 = run_level(ctx, lv2, args, hip, _abi, ref=oracle_ref(lv2))
        per_config["uniform-level-512 (C2)"] = mixed_summary(c2)
        s2 = c2["fp64"]["stats"]
        p2 = s2["far_pairs"] + s2["near_pairs"]
        k2 = c2["fp64"]["stage_ms"][1]
        extra["uniform-level-512"] = {
            "value": p2 / (c2["fp64"]["ms_per_step"] * 1e-3), "ms_per_step": c2["fp64"]["ms_per_step"],
            "kernel_ms": k2, "staging_ms": c2["fp64"]["stage_ms"][0], "kernel_variant": int(s2["variant"]),
            "kernel": kernel_name(int(s2["variant"])),
            "valu_frac": s2["flops"] / (k2 * 1e-3) / 1e12 / FP64_VALU_PEAK_TFLOPS,
            "mixed": {"ms_per_step": c2["mixed"]["ms_per_step"], "kernel_ms": c2["mixed"]["stage_ms"][1],
                      "speedup_vs_fp64": c2["fp64"]["ms_per_step"] / c2["mixed"]["ms_per_step"],
                      "kernel_variant": int(c2["mixed"]["stats"]["variant"]),
                      "errors_vs_fp64_pass": c2["mixed"]["errors_vs_fp64_pass"]},
            "production_step": c2["production_step"]}

    kinds = None
    if not args.no_c3:
        lv3 = workloads.config3()
        c3 = run_level(ctx, lv3, args, hip, _abi, ref=oracle_ref(lv3))
        per_config["mixed-level-2048 (C3)"] = mixed_summary(c3)
        s3 = c3["fp64"]["stats"]
        ms3 = c3["fp64"]["ms_per_step"]
        p3 = s3["far_pairs"] + s3["near_pairs"] + s3["p2p_pairs"] + s3["p2m_pairs"]
        kinds = {"workload": f"{lv3.name}: {len(lv3.targets)} sub-grids, {lv3.info}",
                 "ms_per_step": ms3, "kernel_ms": c3["fp64"]["stage_ms"][1],
                 "cell_interactions_per_s": p3 / (ms3 * 1e-3),
                 "pairs": {k: int(s3[k]) for k in ("far_pairs", "near_pairs", "p2p_pairs", "p2m_pairs")},
                 "valu_frac": s3["flops"] / (c3["fp64"]["stage_ms"][1] * 1e-3) / 1e12 / FP64_VALU_PEAK_TFLOPS,
                 "kernel_variant": int(s3["variant"]), "production_step": c3["production_step"]}

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
                        f"({lv.nslots} slots, 144^3 cells), theta=0.5, RHO (angular-momentum correction on), "
                        f"one hipGraph per pass",
            "subgrids": int(len(lv.targets)),
            "far_pairs_per_pass": int(st["far_pairs"]),
            "flops_per_pair": FLOPS_PER_PAIR["rho"],
            "precision": "fp64",
            "kernel_variant": int(st["variant"]),
            "plan_create_s": c4["fp64"]["plan_create_s"],
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
            "staging_ms": c4["fp64"]["stage_ms"][0],
            # HBM rates: the kernel's measured bytes, and the staging kernel's
            # algorithmic bytes (AoS M, com / mon read, records written)
            "hbm_tbs": (traffic / (k_ms * 1e-3) / 1e12) if traffic else None,
            "hbm_frac": (traffic / (k_ms * 1e-3) / 1e12 / HBM_PEAK_TBS) if traffic else None,
            "staging_hbm_tbs": staging_bytes(lv, st) / (c4["fp64"]["stage_ms"][0] * 1e-3) / 1e12,
            "staging_hbm_frac": staging_bytes(lv, st) / (c4["fp64"]["stage_ms"][0] * 1e-3) / 1e12 / HBM_PEAK_TBS,
            # the dense FP64 FMA rate of this box, re-measured (octofmm_measure_fma_rate)
            "measured_peak": fma_rate["fp64"],
            "frac_of_measured_peak": achieved / fma_rate["fp64"],
        },
        "cpu_baseline": cpu,
        "parity_check": ({"targets": int(len(lv.targets)), "oracle_s": oracle_s,
                          **{k: c4["fp64"]["errors_vs_oracle"][k] for k in
                             ("max_normwise_rel_err", "force_rel_err", "torque_rel_err")}}
                         if ref is not None else None),
        "mixed_precision": mixed,
        # FP32 far-pair arithmetic, FP64 accumulation and correction, on every
        # configuration: speed and errors (the tree's are in tree_solve.mixed)
        "mixed_precision_per_config": per_config,
        "production_step": c4["production_step"],
        "fma_rate_tflops": fma_rate,
        "uniform_level_512": extra.get("uniform-level-512"),
        "mixed_kind_level": kinds,
        "tree_solve": tree,
        "wall_s_headline": wall,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
