"""Benchmark of the FMM gravity interaction pass on one MI355X.

One step = one pass of the level interaction path over the batched uniform
level of 512 non-leaf 8^3 sub-grids (80^3 cells, theta = 0.5, RHO with the
angular-momentum correction): AoS -> partner-record staging plus the
interaction kernel, replayed as one captured hipGraph.  Inputs are resident in
HBM before timing starts.

Prints one JSON line: cell-interactions/s (far M2L pairs per second), the
roofline of the dominant kernel (FP64 VALU), and a CPU baseline (the FP64
oracle on the host cores, on a bounded sample of the same level).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--config c2|c4] [--precision fp64]
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
    ap.add_argument("--config", default="c2", choices=["c2", "c4"])
    ap.add_argument("--cpu-sample", type=int, default=64, help="target sub-grids timed on the CPU (0 = skip)")
    ap.add_argument("--no-check", action="store_true", help="skip the post-run parity spot check")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run 3 device-to-device copies of this many MiB (HBM counter calibration)")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the level pass is single-GPU "
              "in this build", file=sys.stderr)
        sys.exit(2)
    import numpy as np

    from octotiger_amd import _abi, hip, native, workloads

    if hip.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        sys.exit(3)
    lv = workloads.config2() if args.config == "c2" else workloads.config4()
    ctx = native.Context(theta=0.5, device=0)
    dl = hip.LevelOnDevice(lv)
    plan = dl.plan(ctx, type_=_abi.RHO, flags=_abi.PLAN_USE_GRAPH | _abi.PLAN_AUTOTUNE)
    stats = plan.stats()
    if args.copy_calibration_mib > 0:
        nb = args.copy_calibration_mib << 20
        src, dst = hip.DeviceArray((nb // 8,)), hip.DeviceArray((nb // 8,))
        src.zero()
        for _ in range(3):
            hip.check(hip.rt().hipMemcpy(dst.ptr, src.ptr, nb, hip.D2D), "hipMemcpy D2D")
        hip.synchronize()
        src.free()
        dst.free()
    stream = hip.Stream()
    sh = stream.handle
    for _ in range(args.warmup):
        plan.execute(sh)
    stream.synchronize()
    e0, e1 = hip.Event(), hip.Event()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        plan.execute(sh)
    e1.record(stream)
    e1.synchronize()
    wall = time.perf_counter() - t0
    ms_step = e0.elapsed_ms(e1) / args.steps
    pairs = stats["far_pairs"] + stats["near_pairs"]
    value = pairs / (ms_step * 1e-3)

    # dominant kernel: the level interaction kernel, timed alone with HIP
    # events on the stream it is launched on
    stage_ms = plan.time_stages(sh, iters=max(3, min(args.steps, 20)))
    k_ms = stage_ms[1]
    flops = stats["flops"]
    achieved = flops / (k_ms * 1e-3) / 1e12
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{lv.name}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    check = None
    if not args.no_check:
        import oracle
        # the accumulated L is (warmup + steps + stage timing) times one pass;
        # one fresh pass on a few targets is compared with the oracle
        dl.zero_outputs()
        plan.execute(sh)
        stream.synchronize()
        sample = lv.targets[:: max(1, len(lv.targets) // 8)][:8]
        Lr, Lcr = oracle.level_interactions(lv, targets=sample)
        Lg, Lcg = dl.L.numpy(), dl.L_c.numpy()
        err = 0.0
        for a, b in ((Lg[sample], Lr[sample]), (Lcg[sample], Lcr[sample])):
            a2, b2 = a.reshape(-1, a.shape[-1]), b.reshape(-1, b.shape[-1])
            for c in range(b2.shape[1]):
                err = max(err, float(np.abs(a2[:, c] - b2[:, c]).max() / np.abs(b2[:, c]).max()))
        check = {"targets": int(len(sample)), "max_normwise_rel_err": err}

    cpu = None
    if args.cpu_sample > 0:
        import oracle
        ncores = min(16, os.cpu_count() or 1)
        sample = lv.targets[: args.cpu_sample]
        t = time.perf_counter()
        oracle.level_interactions(lv, targets=sample, nthreads=ncores)
        dt = time.perf_counter() - t
        cpu = {"value": len(sample) * (pairs / len(lv.targets)) / dt, "unit": "cell-interactions/s",
               "cores": ncores, "kind": "port",
               "sample": f"{len(sample)} of the {len(lv.targets)} target sub-grids of the same level, "
                         f"FP64 scalar oracle (Kokkos host-kernel walk), OpenMP over sub-grids, {dt:.2f} s"}

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
            "far_pairs_per_pass": int(stats["far_pairs"]),
            "flops_per_pair": FLOPS_PER_PAIR["rho"],
            "precision": "fp64",
            "workgroup_waves": int(stats["variant"]),
        },
        "roofline": {
            "bound": "valu",
            "kernel": "level_kernel<RHO>",
            "achieved": achieved,
            "peak": FP64_VALU_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_VALU_PEAK_TFLOPS,
            "traffic": traffic,
            "kernel_ms": k_ms,
            "staging_ms": stage_ms[0],
        },
        "cpu_baseline": cpu,
        "parity_check": check,
        "wall_s_timed": wall,
    }
    print(json.dumps(out), flush=True)
    plan.close()
    dl.free()
    stream.destroy()
    ctx.close()


if __name__ == "__main__":
    main()
