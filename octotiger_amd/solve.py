"""Run the whole-tree gravity solve on one GPU and print its timings as JSON.

    python -m octotiger_amd.solve [--tree c5|small|sphere] [--steps N] [--precision fp64|mixed] [--no-graph]
                                  [--leaf-g-only]

Builds the tree on the host (octotiger_amd.workloads), uploads the leaf masses,
creates the tree plan (one hipGraph per solve, autotuned level kernels; with
--no-graph the kernels are launched on each execute instead) and runs it N
times between HIP events; then times its two phases separately:
the upward M2M pass, and the interactions of every level with the downward
L2L pass and the leaf force assembly.  --leaf-g-only writes only g (and L_c)
for leaf cells (OCTOFMM_TREE_LEAF_G_ONLY).  For the
solid-sphere tree it also prints the 'pot' norms Octo-Tiger's regression
test checks (workloads.sphere_pot_norms).
"""
import argparse
import json
import sys


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--tree", choices=("c5", "small", "sphere"), default="c5")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    # a solve of a few ms reaches its steady rate only after some tens of ms
    # of back-to-back replays: further untimed replays until this many ms
    ap.add_argument("--warm-ms", type=float, default=150.0)
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--leaf-g-only", action="store_true")
    args = ap.parse_args(argv)

    from . import _abi, hip, native, workloads

    if hip.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        return 3
    tree = {"c5": workloads.config5, "small": workloads.small_tree, "sphere": workloads.solid_sphere_tree}[args.tree]()
    ctx = native.Context(theta=0.5, device=0)
    dev = hip.TreeOnDevice(tree)
    flags = (0 if args.no_graph else _abi.PLAN_USE_GRAPH) | (0 if args.no_autotune else _abi.PLAN_AUTOTUNE)
    if args.leaf_g_only:
        flags |= _abi.TREE_LEAF_G_ONLY
    prec = _abi.PRECISION_MIXED if args.precision == "mixed" else _abi.PRECISION_FP64
    plan = dev.plan(ctx, precision=prec, flags=flags)
    s = hip.Stream()
    for _ in range(args.warmup):
        plan.execute(s.handle)
    s.synchronize()
    e0, e1 = hip.Event(), hip.Event()
    spent = 0.0
    while spent < args.warm_ms:
        e0.record(s)
        for _ in range(10):
            plan.execute(s.handle)
        e1.record(s)
        e1.synchronize()
        spent += e0.elapsed_ms(e1)
    e0.record(s)
    for _ in range(args.steps):
        plan.execute(s.handle)
    e1.record(s)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / args.steps
    phases = plan.time_phases(s.handle, iters=3)
    st = plan.stats()
    extra = {}
    if args.tree == "sphere":
        l1, l2 = workloads.sphere_pot_norms(tree, [d["g"][..., 0] for d in dev.fetch("g")])
        extra["pot_norms"] = f"pot {l1:e} {l2:e}"
    pairs = sum(st[k] for k in ("far_pairs", "near_pairs", "p2p_pairs", "p2m_pairs", "root_pairs"))
    print(json.dumps({"tree": tree.name, "nodes_per_level": tree.info["nodes_per_level"],
                      "precision": args.precision, "graph": not args.no_graph,
                      "leaf_g_only": args.leaf_g_only, "ms_per_solve": ms,
                      "phases_ms": {"upward_m2m": phases[0], "interactions_l2l_forces": phases[1]},
                      "cell_interactions_per_s": pairs / (ms * 1e-3), "stats": st, **extra}), flush=True)
    plan.close()
    dev.free()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
