"""PD-IPC step benchmark on one MI355X.

Workload: a 256x256-vertex cloth (130,050 triangles) dropped onto a static
noise-displaced icosphere (subdivision 6, 81,920 triangles) above a ground
plane; dt = 1/100, 10 local/global passes x 15 Chebyshev sweeps per outer
iteration, dHat = 1e-3, kappa = 1e3.  A "step" is one full PD-IPC time step
(prediction, outer CCD iterations, velocity update) with all data resident in
HBM.  By default 20 warm-up steps and 180 timed steps: steps 21-200 of the
200-step run from rest.  The FP32-split step is the headline value; the plain
FP64 step is timed from the same state, and the per-step error of FP32-split
against FP64 is measured from identical states.

Prints one JSON line.  --gpus above 1 exits: this build runs on one GPU.
"""

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pd-ipc-armadillodemo_amd"))

import pdipc  # noqa: E402
import scenes  # noqa: E402

METRIC = "PD-IPC steps/s, 256² cloth: FP32-split vs FP64 at 1e-5 err; adjoint/fwd ratio"
HBM_PEAK_GBS = 8000.0
# Bytes one FP32 sweep launch moves past L2 at the default workload, from two
# rocprofv3 --pmc runs (FETCH_SIZE x 2.000 and WRITE_SIZE x 1.000, the factors
# a plain 1 GiB copy gives): profiles/k_sweep_config1_hbm_traffic.md.
SWEEP32_TRAFFIC_BYTES = {(256, 6, 8): 29.86e6}
# steps of each scene's run from rest; by default the warm-up steps and the
# timed steps together make up the whole run
RUN_STEPS = {"cloth": 200, "config0": 100, "tet_block": 100, "stacked": 50}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def timed_steps(sim, steps):
    t0 = time.perf_counter()
    outer = 0
    for _ in range(steps):
        sim.step(1)
        outer += sim.counters()["outer_iters"]
    return time.perf_counter() - t0, outer


def stacked_bench(args):
    """The scale run: three stacked 1024^2 cloths (3,145,728 vertices) over the
    sphere and the ground, FP32-split (FP64 CCD and barrier), `args.steps`
    steps from rest; reports steps/s, the candidate records K and patch pairs
    Q of the broad phases, and the FP64 triangle-triangle intersection count
    before and after (must stay 0)."""
    sc = scenes.config4()
    sizes = scenes.scene_sizes(sc)
    t0 = time.perf_counter()
    sim = pdipc.Simulator.from_scene(sc, precision=pdipc.FP32_SPLIT, sell_lanes=args.sell_lanes,
                                     use_graphs=args.graphs, skip_static_pairs=1)
    setup_s = time.perf_counter() - t0
    c0 = sim.counters()
    log(f"setup {setup_s:.1f} s; {sizes}; nnz_r2 {c0['nnz_r2']}, 2-paths {c0['n_2paths']}")
    n_start = sim.count_intersections()
    log(f"intersections at start: {n_start}")
    secs, outer, ks, qs = 0.0, 0, [], []
    for k in range(args.steps):
        t1 = time.perf_counter()
        sim.step(1)
        secs += time.perf_counter() - t1
        c = sim.counters()
        outer += c["outer_iters"]
        ks.append(c["candidates"])
        qs.append(c["patch_pairs"])
        if k % 5 == 4 or k == args.steps - 1:
            log(f"step {k + 1}: {secs / (k + 1) * 1000:.0f} ms/step, outer {c['outer_iters']}, "
                f"K {c['candidates']}, Q {c['patch_pairs']}, active {c['active_pairs']}")
    n_end = sim.count_intersections()
    log(f"intersections at end: {n_end}")
    return {
        "metric": "PD-IPC steps/s, scale run (3 stacked 1024^2 cloths)",
        "value": args.steps / secs,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 0,
        "ms_per_step": 1000.0 * secs / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": "3 stacked 1024x1024 cloths (2 m, 5 mm apart, seeded shear |s| <= 0.1) over icosphere "
                        "subdiv 6 + ground, dt=1/100, 10 LG x 15 Chebyshev, dHat=1e-3, kappa=1e3, from rest",
            "precision": "FP32-split (FP64 CCD/barrier)",
            "sim_points": sizes["sim_points"], "triangles": sizes["surface_triangles"],
            "nnz_r2": c0["nnz_r2"], "n_2paths": c0["n_2paths"], "setup_s": setup_s,
            "outer_iters_per_step": outer / args.steps,
            "candidates_K": {"first": ks[0], "last": ks[-1], "max": max(ks)},
            "patch_pairs_Q": {"first": qs[0], "last": qs[-1], "max": max(qs)},
            "intersections_fp64": {"start": n_start, "end": n_end},
            "graphs": bool(args.graphs),
        },
    }


def _host_threads():
    n = os.environ.get("OMP_NUM_THREADS", "")
    if n.isdigit() and int(n) > 0:
        return int(n)
    return len(os.sched_getaffinity(0))


def cpu_baseline(scene, x, v, seconds_budget=20.0, label="256x256 cloth"):
    """The FP64 oracle (the C++ restatement of the reference step; OpenMP over
    the solve, operator update, TOI and barrier-activity loops) on this host, on
    a bounded sample: the full-size scene from the same warmed-up state the GPU
    timing starts from, as many steps as fit the budget, on all the host
    threads this process may use (`value`) and on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the checker, timed here as the CPU baseline

    def sample(threads, budget):
        oracle.set_threads(threads)
        t0 = time.perf_counter()
        ref = oracle.Oracle(scene)
        ref.set_x(x)
        ref.set_v(v)
        init_s = time.perf_counter() - t0
        n, t, outer = 0, 0.0, 0
        while t < budget and n < 50:
            t1 = time.perf_counter()
            ref.step(1)
            t += time.perf_counter() - t1
            n += 1
            outer += int(ref.counters()["outer"])
        return n, t, init_s, outer

    threads = _host_threads()
    n, t, init_s, outer = sample(threads, 0.5 * seconds_budget)
    n1, t1, _, _ = sample(1, seconds_budget)
    oracle.set_threads(0)
    return dict(value=n / t, unit="steps/s", cores=threads, kind="port",
                single_thread_steps_per_s=n1 / t1, outer_iters_per_step=outer / n,
                sample=f"FP64 oracle steps of the same scene ({label}) from the GPU timing's warmed-up state: "
                       f"{n} steps on {threads} threads ({t:.1f} s, {outer / n:.1f} outer iterations per step), "
                       f"{n1} steps on 1 thread ({t1:.1f} s); setup {init_s:.1f} s each")


def adjoint_bench(graphs, fd_dirs, log_every=8):
    """The differentiable run: 128 x 128 cloth over the config-0 obstacle, 50 FP64
    steps, loss |x_T - target|^2, gradient w.r.t. per-triangle stiffness and
    initial velocity by the library's adjoint, in its two modes: the rollout
    keeps the whole run's tape in HBM (checkpoint_every = 0; the backward runs
    no forward step), or keeps (x, v) every 10 steps (the backward re-runs each
    segment recording it).  Timed forward (rollout) and backward of each, then
    central differences (eps 1e-6) on `fd_dirs` random directions."""
    steps, every = 50, 10
    sc, v0, target = scenes.config3(0, n_cloth=128, steps=steps)
    sim = pdipc.Simulator.from_scene(sc, use_graphs=graphs)
    x0 = sim.x()[: sim.sim].copy()
    k0 = sc["cloths"][0]["tri_k"].copy()
    ntri = len(k0)

    def forward(k, v, ck):
        sim.set_cloth_stiffness(0, k)
        sim.set_x(x0)
        sim.set_v(v)
        t0 = time.perf_counter()
        sim.rollout(steps, ck)
        t = time.perf_counter() - t0
        x = sim.x()[: sim.sim]
        return float(np.sum((x - target) ** 2)), x, t

    res = {}
    grads = {}
    for mode, ck in (("recorded", 0), ("checkpointed", every)):
        forward(k0, v0, ck)  # warm: allocations, graphs
        _, xT, _ = forward(k0, v0, ck)
        sim.backward(dx=2.0 * (xT - target), n_tris=ntri)
        fw, bw = [], []
        for _ in range(2):
            loss, xT, tf = forward(k0, v0, ck)
            t0 = time.perf_counter()
            _, dv0, dk = sim.backward(dx=2.0 * (xT - target), n_tris=ntri)
            bw.append(time.perf_counter() - t0)
            fw.append(tf)
        stats = sim.adjoint_stats()
        tf, tb = min(fw), min(bw)
        grads[mode] = (dk, dv0)
        res[mode] = {"checkpoint_every": ck, "forward_s": tf, "backward_s": tb, "backward_fwd_ratio": tb / tf,
                     "tape_bytes": stats["tape_bytes"], "recorded_outer_iters": stats["outer_iters"]}
        log(f"adjoint ({mode}): forward {tf:.3f} s, backward {tb:.3f} s (ratio {tb / tf:.2f}); {stats}")
    (dk, dv0), (dkc, dvc) = grads["recorded"], grads["checkpointed"]
    same = max(float(np.abs(dk - dkc).max() / np.abs(dkc).max()), float(np.abs(dv0 - dvc).max() / np.abs(dvc).max()))
    rng = np.random.default_rng(12345)
    errs = []
    eps = 1e-6
    for q in range(fd_dirs):
        dkq = rng.normal(size=ntri) * k0
        dvq = rng.normal(size=v0.shape) * 0.05
        lp, _, _ = forward(k0 + eps * dkq, v0 + eps * dvq, every)
        lm, _, _ = forward(k0 - eps * dkq, v0 - eps * dvq, every)
        fd = (lp - lm) / (2 * eps)
        ad = float(dk @ dkq + np.sum(dv0 * dvq))
        errs.append(abs(fd - ad) / max(abs(ad), 1e-300))
        if q % log_every == 0:
            log(f"  direction {q}: fd {fd:.6e} adjoint {ad:.6e} rel {errs[-1]:.2e}")
    errs = np.array(errs)
    r, c = res["recorded"], res["checkpointed"]
    return {
        "workload": "cloth 128x128 (32,258 triangles) over icosphere subdiv 4 (config-0 obstacle), 50 FP64 steps, "
                    "loss |x_T - target|^2, d/d(per-triangle stiffness, v0)",
        "mode": "whole-run tape kept in HBM (checkpoint_every = 0)",
        "forward_s": r["forward_s"], "backward_s": r["backward_s"], "adjoint_fwd_ratio": r["backward_fwd_ratio"],
        "gradient_cost_vs_checkpointed_forward": (r["forward_s"] + r["backward_s"]) / c["forward_s"],
        "recorded": r, "checkpointed": c, "modes_max_rel_diff": same,
        "fd_directions": int(fd_dirs), "fd_eps": eps,
        "fd_rel_err_median": float(np.median(errs)) if fd_dirs else None,
        "fd_rel_err_max": float(errs.max()) if fd_dirs else None,
        "fd_dirs_within_1e-3": int(np.sum(errs < 1e-3)) if fd_dirs else None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: the rest of the scene's run after the warm-up: 200 steps in all "
                         "for the 256^2 cloth, 100 for config0 and tet_block, 50 for stacked)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--scene", choices=("cloth", "config0", "tet_block", "stacked"), default="cloth",
                    help="cloth: the n x n cloth over the sphere (default); config0: the 64x64 cloth over the "
                         "subdivision-4 sphere without ground; tet_block: the 40^3-cell block "
                         "(320,000 tets) under a 128x128 cloth, dt = 1/200, 15 local/global iterations; "
                         "stacked: the scale run, 3 stacked 1024^2 cloths, FP32-split only, --steps steps")
    ap.add_argument("--n", type=int, default=256, help="cloth vertices per side")
    ap.add_argument("--subdiv", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--sell-lanes", type=int, default=8)
    ap.add_argument("--graphs", type=int, default=1, help="replay the local/global loop from a HIP graph")
    ap.add_argument("--keep-static-pairs", action="store_true",
                    help="keep static-static patch pairs (the reference's full candidate set)")
    ap.add_argument("--adjoint-fd", type=int, default=32, help="finite-difference directions of the adjoint check")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device copy (3 times): the reference for the HBM byte counters")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--only-adjoint", action="store_true", help="run only the differentiable-run measurement")
    args = ap.parse_args()
    run_steps = RUN_STEPS[args.scene]
    if args.steps is None:
        args.steps = run_steps if args.scene == "stacked" else max(1, run_steps - args.warmup)
    if args.copy_probe:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read_per_launch": a.numel() * 4,
                          "copy_probe_bytes_written_per_launch": b.numel() * 4}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs the PD-IPC step on one GPU; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)
    if args.scene == "stacked":
        print(json.dumps(stacked_bench(args)), flush=True)
        return
    if args.only_adjoint:
        print(json.dumps(adjoint_bench(args.graphs, args.adjoint_fd)), flush=True)
        return

    if args.scene == "config0":
        scene = scenes.config0()
        workload = ("cloth 64x64 (7,938 triangles) over icosphere subdiv 4 (5,120 triangles), no ground, "
                    "dt=1/100, 10 LG x 15 Chebyshev, dHat=1e-3, kappa=1e3")
    elif args.scene == "tet_block":
        scene = scenes.config2()
        workload = ("tet block 40^3 cells (320,000 tets) + 128x128 cloth on the ground, dt=1/200, "
                    "15 LG x 15 Chebyshev, dHat=1e-3, kappa=1e3")
    else:
        scene = scenes.cloth_over_sphere(args.n, args.subdiv, ground=True, seed=0, patch_dim=args.n,
                                         sphere_patch_dim=args.n)
        workload = None
    t0 = time.perf_counter()
    sims = {}
    for name, prec in (("fp64", pdipc.FP64), ("fp32_split", pdipc.FP32_SPLIT)):
        sims[name] = pdipc.Simulator.from_scene(scene, precision=prec, sell_lanes=args.sell_lanes,
                                                use_graphs=args.graphs,
                                                skip_static_pairs=0 if args.keep_static_pairs else 1)
    log(f"setup {time.perf_counter() - t0:.1f} s; counters {sims['fp64'].counters()}")

    # warm-up on the FP64 handle, then both precisions time K steps from that state
    base = sims["fp64"]
    for k in range(args.warmup):
        base.step(1)
    x_w, v_w = base.x(), base.v()
    sims["fp32_split"].set_x(x_w)
    sims["fp32_split"].set_v(v_w)
    sims["fp32_split"].step(1)  # warm its kernels, then reset
    sims["fp32_split"].set_x(x_w)
    sims["fp32_split"].set_v(v_w)
    log(f"warm-up done ({args.warmup} steps)")

    res = {}
    for name in ("fp64", "fp32_split"):
        sim = sims[name]
        sim.set_x(x_w)
        sim.set_v(v_w)
        secs, outer = timed_steps(sim, args.steps)
        res[name] = dict(seconds=secs, outer=outer)
        log(f"{name}: {args.steps} steps in {secs:.3f} s ({args.steps / secs:.2f} steps/s, {outer} outer iterations)")

    # per-step error of FP32-split vs FP64 from identical states (teacher forced)
    errs = []
    a, b = sims["fp64"], sims["fp32_split"]
    for k in range(5):
        x, v = a.x(), a.v()
        b.set_x(x)
        b.set_v(v)
        a.step(1)
        b.step(1)
        xa, xb = a.x()[: a.sim], b.x()[: b.sim]
        errs.append(float(np.abs(xa - xb).max() / np.abs(xa).max()))
    log(f"FP32-split per-step error vs FP64: {errs}")

    # dominant kernel: the Chebyshev sweep (150 launches per outer iteration)
    us64, bytes64 = a.time_kernel("sweep_fp64", 200)
    us32, bytes32 = b.time_kernel("sweep_fp32", 200)
    c = a.counters()
    ach = bytes32 / (us32 * 1e-6) / 1e9
    log(f"sweep fp64 {us64:.2f} us ({bytes64 / (us64 * 1e-6) / 1e9:.0f} GB/s), fp32 {us32:.2f} us ({ach:.0f} GB/s)")

    adj = None
    if not args.no_adjoint:
        adj = adjoint_bench(args.graphs, args.adjoint_fd)

    cpu = None
    if not args.no_cpu:
        label = {"config0": "64x64 cloth", "tet_block": "tet block + 128x128 cloth"}.get(args.scene,
                                                                                    f"{args.n}x{args.n} cloth")
        cpu = cpu_baseline(scene, x_w, v_w, args.cpu_seconds, label)
        log(f"cpu baseline {cpu}")

    steps = args.steps
    t32 = res["fp32_split"]["seconds"]
    out = {
        "metric": METRIC,
        "value": steps / t32,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * t32 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": workload or f"cloth {args.n}x{args.n} ({c['cloth_points']} vertices) over icosphere subdiv "
                                    f"{args.subdiv} + ground, dt=1/100, 10 LG x 15 Chebyshev, dHat=1e-3, kappa=1e3",
            "precision": "FP32-split (FP32 local step + sweeps on an FP64 residual; FP64 CCD/barrier)",
            "timed_steps": f"steps {args.warmup + 1}-{args.warmup + steps} of a run from rest "
                           f"({RUN_STEPS[args.scene]} steps configured), each precision from the same state",
            "fp64_steps_per_s": steps / res["fp64"]["seconds"],
            "fp32_split_speedup_vs_fp64": res["fp64"]["seconds"] / t32,
            "fp32_split_step_error_vs_fp64": max(errs),
            "outer_iters_per_step": res["fp32_split"]["outer"] / steps,
            "nnz_r2": c["nnz_r2"], "candidates": c["candidates"], "patch_pairs": c["patch_pairs"],
            "patches": c["patches"],
            "static_static_pairs": "kept" if args.keep_static_pairs else
                                   "skipped (inert: they only write rows the step never reads)",
            "graphs": bool(args.graphs),
            "parallelism": "single GPU",
            "adjoint": adj,
        },
        "roofline": {
            "kernel": "k_sweep32 (FP32 Chebyshev sweep, SELL M^2)",
            "bound": "hbm",
            "achieved": ach,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS,
            "traffic": None if workload else SWEEP32_TRAFFIC_BYTES.get((args.n, args.subdiv, args.sell_lanes)),
            "avg_us": us32,
            "bytes_per_launch": bytes32,
            "fp64_sweep": {"avg_us": us64, "achieved": bytes64 / (us64 * 1e-6) / 1e9},
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
