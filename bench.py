"""PD-IPC step benchmark on one MI355X.

Workload: a 256x256-vertex cloth (130,050 triangles) dropped onto a static
noise-displaced icosphere (subdivision 6, 81,920 triangles) above a ground
plane; dt = 1/100, 10 local/global passes x 15 Chebyshev sweeps per outer
iteration, dHat = 1e-3, kappa = 1e3, a 200-step run from rest.  A "step" is
one full PD-IPC time step (prediction, outer CCD iterations, velocity update)
with all data resident in HBM.

Measured, each precision from the same state:
  * the timed window: steps W+1 .. W+K of the run (W warm-up steps, K timed;
    default W = 20, K = 180, i.e. steps 21-200), graphs on -- `value` is the
    FP32-split steps/s there;
  * the settled window: 10 steps from step 150, once the cloth rests on the
    sphere (the outer iterations per step, hence the cost, roughly double);
  * the per-step error of FP32-split against FP64 from identical states;
  * in-step kernel durations (HIP events on every launch, over the timed
    window re-run with direct launches): the roofline of the FP32 sweep uses
    the algorithmic bytes (8 B per M^2 entry: an FP32 value and a 32-bit
    column, plus the row vectors) over that mean.

Prints one JSON line.  --gpus above 1 exits: the step runs on one GPU.
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
# a plain 1 GiB copy gives), by cloth size, sphere subdivision, lanes per row
# and how the sweep reads its operand (1: 16-bit staged indices, 2: narrow
# staging): profiles/k_sweep_config1_hbm_traffic.md.
SWEEP32_TRAFFIC_BYTES = {(256, 6, 8, 1): 31.63e6, (256, 6, 8, 2): 28.55e6}
# the same for the three stacked 1024^2 cloths at 8 lanes per row, by how the
# sweep reads its operand (1: 16-bit staged indices, 2: narrow staging):
# profiles/config4_stacked_kernel_stats.md
STACKED_SWEEP32_TRAFFIC_BYTES = {(8, 1): 1518.6e6, (8, 2): 1370.8e6}
# the FP32 sweep kernel by pdipc_sweep_staging's answer, and its bytes per
# stored slot (FP32 value + union index or column)
SWEEP32_KERNELS = {
    0: ("k_sweep32 (FP32 Chebyshev sweep, SELL M^2, operand gathered from L2)", 6.0),
    1: ("k_sweep32s (FP32 Chebyshev sweep, SELL M^2, operand staged in LDS)", 6.0),
    2: ("k_sweep32n (FP32 Chebyshev sweep, SELL M^2, operand staged in LDS, one-byte union indices)", 5.0),
}
# steps of each scene's run from rest; by default the warm-up steps and the
# timed steps together make up the whole run
RUN_STEPS = {"cloth": 200, "config0": 100, "tet_block": 100, "stacked": 50}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def timed_steps(sim, steps):
    """Wall time of `steps` steps (the counter reads between steps are outside
    the step but inside the window); outer iterations and the candidate
    records of each step's broad phase."""
    t0 = time.perf_counter()
    outer, recs = 0, []
    for _ in range(steps):
        sim.step(1)
        c = sim.counters()
        outer += c["outer_iters"]
        recs.append(c["candidates"])
    return time.perf_counter() - t0, outer, recs


def cloth_ring_slots(scene):
    """N + 2E of the simulated cloths: the local step's ring slots."""
    n = 0
    for c in scene["cloths"]:
        f = np.asarray(c["f"], np.int64)
        e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
        n += len(c["v"]) + 2 * len(np.unique(e, axis=0))
    return n


def algorithmic_step_bytes(c, ring_slots, outer_per_step, recs_per_step, p_vec, lg=10, cheb=15):
    """Algorithmic HBM bytes of one step (FP32 values,
    int32 indices, p_vec bytes per vector element in the solve and local step,
    FP64 CCD and barrier): per outer iteration the operator update, lg x
    (local step + cheb sweeps), CCD and barrier over the step's candidate
    records K, plus prediction and velocity (4 and 3 vector passes; the broad
    phase is left out)."""
    N, nnz1, nnz2, n2p = c["sim_points"], c["nnz_r1"], c["nnz_r2"], c["n_2paths"]
    D, tot, K = 3 * N, c["total_points"], recs_per_step
    b_spmv = 8 * nnz2 + 4 * D * p_vec
    b_local = ring_slots * 20 + N * 4 + 2 * D * p_vec
    b_ops = nnz1 * 12 + n2p * 16 + nnz2 * 16 + N * p_vec
    b_ccd = K * (20 + 24 * 8)
    b_bar = K * (20 + 12 * 8) + 2 * (tot * 8 + 3 * tot * 8)
    per_outer = b_ops + lg * (b_local + cheb * b_spmv) + b_ccd + b_bar
    return outer_per_step * per_outer + 7 * D * p_vec, b_spmv


def kernel_means(sim, x, v, steps):
    """Mean in-step duration (us) and launches per timed class over `steps`
    steps from (x, v), every launch timed with HIP events (direct launches)."""
    sim.set_x(x)
    sim.set_v(v)
    sim.kernel_timing(True)
    sim.step(steps)
    t = sim.kernel_times()
    sim.kernel_timing(False)
    return {k: {"mean_us": us / n, "launches": n} for k, (us, n) in t.items() if n}


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
    # in-step kernel durations over two more steps (direct launches, every launch timed)
    sim.kernel_timing(True)
    sim.step(2)
    kt = {k: {"mean_us": us / n, "launches": n} for k, (us, n) in sim.kernel_times().items() if n}
    sim.kernel_timing(False)
    log(f"in-step kernel means: {json.dumps(kt)}")
    sweep_bytes = 8 * c0["nnz_r2"] + 4 * 3 * c0["sim_points"] * 4
    sweep_us = kt["sweep_fp32"]["mean_us"]
    staging = sim.sweep_staging()
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
            "kernel_means_us": kt,
        },
        "roofline": {
            "kernel": SWEEP32_KERNELS[staging][0],
            "bound": "hbm",
            "achieved": sweep_bytes / (sweep_us * 1e-6) / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": sweep_bytes / (sweep_us * 1e-6) / 1e9 / HBM_PEAK_GBS,
            "traffic": STACKED_SWEEP32_TRAFFIC_BYTES.get((args.sell_lanes, staging)),
            "avg_us": sweep_us,
            "bytes_per_launch": sweep_bytes,
            "bytes_formula": "8 B per M^2 entry (FP32 value + int32 column) + 4 vectors x 3N x 4 B",
            "mall_resident": c0["sell_entries"] * SWEEP32_KERNELS[staging][1] < 256e6,
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
    # the oracle built for this host, -O3 -march=native (into a temporary
    # folder); the prebuilt -O2 portable build if that fails
    build_note = "-O3 -march=native, built on this host"
    try:
        import tempfile
        ndir = tempfile.mkdtemp(prefix="pdipc_oracle_native_")
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native", f"NATIVEDIR={ndir}"], check=True,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=300)
        os.environ["ORACLE_LIB"] = os.path.join(ndir, "liboracle.so")
    except Exception as e:  # noqa: BLE001
        build_note = f"-O2 portable build (native build failed: {type(e).__name__})"
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
    return dict(value=n / t, unit="steps/s", cores=threads, kind="port", build=build_note,
                single_thread_steps_per_s=n1 / t1, outer_iters_per_step=outer / n,
                sample=f"FP64 oracle steps of the same scene ({label}) from the GPU timing's warmed-up state: "
                       f"{n} steps on {threads} threads ({t:.1f} s, {outer / n:.1f} outer iterations per step), "
                       f"{n1} steps on 1 thread ({t1:.1f} s); setup {init_s:.1f} s each")


def adjoint_bench(graphs, fd_dirs, log_every=8):
    """The differentiable run: 128 x 128 cloth over the subdivision-4 icosphere, 50 FP64
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

    # plain steps of the same run (bit for bit the rollout's) on a handle that
    # also reports the record attaining the minimum TOI at every outer iteration
    sim_fd = pdipc.Simulator.from_scene(sc, use_graphs=graphs, track_toi_argmin=1)

    def plain_forward(k, v):
        """The loss, and the choices the adjoint freezes, read after every step."""
        sim_fd.set_cloth_stiffness(0, k)
        sim_fd.set_x(x0)
        sim_fd.set_v(v)
        sig = []
        for _ in range(steps):
            sim_fd.step(1)
            cn = sim_fd.counters()
            sig.append((cn["outer_iters"], cn["candidates"], cn["active_pairs"], cn["toi_argmins"],
                        cn["active_sets"]))
        return float(np.sum((sim_fd.x()[: sim_fd.sim] - target) ** 2)), sig

    def switch_of(s0, sp, sm):
        """Why the +/- eps runs leave the adjoint's frozen choices, or None: a
        different outer-iteration, candidate-record or active-pair count, a
        different record attaining the minimum TOI or ACCD exit path of that
        record, or a different set of active barrier records, at some step."""
        names = ("outer iterations", "candidate records", "active barrier pairs",
                 "minimum-TOI records or their ACCD exit paths", "active barrier record sets")
        for st, (a, b, c0) in enumerate(zip(sp, sm, s0)):
            for idx, name in enumerate(names):
                if not (a[idx] == b[idx] == c0[idx]):
                    if idx >= 3:
                        return f"step {st + 1}: {name} differ"
                    return f"step {st + 1}: {name} {b[idx]} / {c0[idx]} / {a[idx]} (-eps / 0 / +eps)"
        return None

    l0, sig0 = plain_forward(k0, v0)
    rng = np.random.default_rng(12345)
    errs, switches, outliers = [], [], []
    eps = 1e-6
    for q in range(fd_dirs):
        dkq = rng.normal(size=ntri) * k0
        dvq = rng.normal(size=v0.shape) * 0.05
        lp, sp = plain_forward(k0 + eps * dkq, v0 + eps * dvq)
        lm, sm = plain_forward(k0 - eps * dkq, v0 - eps * dvq)
        fd = (lp - lm) / (2 * eps)
        ad = float(dk @ dkq + np.sum(dv0 * dvq))
        errs.append(abs(fd - ad) / max(abs(ad), 1e-300))
        why = switch_of(sig0, sp, sm)
        switches.append(why)
        if errs[-1] > 1e-4:
            outliers.append({"direction": q, "rel_err": errs[-1], "switch": why})
        if q % log_every == 0 or errs[-1] > 1e-4:
            log(f"  direction {q}: fd {fd:.6e} adjoint {ad:.6e} rel {errs[-1]:.2e} switch: {why}")
    errs = np.array(errs)
    smooth = np.array([w is None for w in switches], bool)
    r, c = res["recorded"], res["checkpointed"]
    return {
        "workload": "cloth 128x128 (32,258 triangles) over icosphere subdiv 4, 50 FP64 steps, "
                    "loss |x_T - target|^2, d/d(per-triangle stiffness, v0)",
        "mode": "whole-run tape kept in HBM (checkpoint_every = 0)",
        "forward_s": r["forward_s"], "backward_s": r["backward_s"], "adjoint_fwd_ratio": r["backward_fwd_ratio"],
        "gradient_cost_vs_checkpointed_forward": (r["forward_s"] + r["backward_s"]) / c["forward_s"],
        "recorded": r, "checkpointed": c, "modes_max_rel_diff": same,
        "fd_directions": int(fd_dirs), "fd_eps": eps,
        "fd_rel_err_median": float(np.median(errs)) if fd_dirs else None,
        "fd_rel_err_max": float(errs.max()) if fd_dirs else None,
        "fd_dirs_within_1e-3": int(np.sum(errs < 1e-3)) if fd_dirs else None,
        "fd_non_switching": {"directions": int(smooth.sum()),
                             "rel_err_max": float(errs[smooth].max()) if smooth.any() else None,
                             "rel_err_median": float(np.median(errs[smooth])) if smooth.any() else None},
        "fd_switching": {"directions": int((~smooth).sum()),
                         "rel_err_max": float(errs[~smooth].max()) if (~smooth).any() else None},
        "fd_outliers_above_1e-4": outliers,
        "fd_switch_rule": "a direction switches when its -eps, 0 and +eps runs differ at some step in outer "
                          "iterations, candidate records, active barrier pairs, the record attaining the minimum "
                          "TOI of an outer iteration or the iteration and test at which its ACCD loop ended, or "
                          "the set of active barrier records (track_toi_argmin)",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: the rest of the scene's run after the warm-up: 200 steps in all "
                         "for the 256^2 cloth, 100 for config0 and tet_block, 50 for stacked)")
    ap.add_argument("--settle-step", type=int, default=None,
                    help="step from which the settled window is timed (default: 150 of the 200-step cloth run, "
                         "75 of the 100-step runs; 0: no settled window)")
    ap.add_argument("--settle-steps", type=int, default=10, help="steps of the settled window")
    ap.add_argument("--replay", action="store_true",
                    help="also time 200 back-to-back replays of each sweep (pdipc_time_kernel)")
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
    ap.add_argument("--state-out", default=None,
                    help="write the state after the warm-up steps to this file (pdipc_save_state)")
    ap.add_argument("--state-in", default=None,
                    help="start from this pdipc_save_state file instead of running the warm-up steps (written "
                         "by --state-out with the same --warmup: the timed steps keep their step numbers)")
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
        print(f"bench.py: the PD-IPC step runs on one GPU; --gpus {args.gpus} is not supported",
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
    if args.state_in:
        base.load_state(args.state_in)
    else:
        for k in range(args.warmup):
            base.step(1)
    if args.state_out:
        base.save_state(args.state_out)
    x_w, v_w = base.x(), base.v()
    sims["fp32_split"].set_x(x_w)
    sims["fp32_split"].set_v(v_w)
    sims["fp32_split"].step(1)  # warm its kernels, then reset
    sims["fp32_split"].set_x(x_w)
    sims["fp32_split"].set_v(v_w)
    log(f"warm-up done ({args.warmup} steps{', restored from ' + args.state_in if args.state_in else ''})")

    res = {}
    for name in ("fp64", "fp32_split"):
        sim = sims[name]
        sim.set_x(x_w)
        sim.set_v(v_w)
        secs, outer, recs = timed_steps(sim, args.steps)
        res[name] = dict(seconds=secs, outer=outer, recs=recs)
        log(f"{name}: steps {args.warmup + 1}-{args.warmup + args.steps} in {secs:.3f} s "
            f"({args.steps / secs:.2f} steps/s, {outer} outer iterations)")

    # per-step error of FP32-split vs FP64 from identical states (teacher forced)
    def step_errors(n):
        errs = []
        a, b = sims["fp64"], sims["fp32_split"]
        for _ in range(n):
            x, v = a.x(), a.v()
            b.set_x(x)
            b.set_v(v)
            a.step(1)
            b.step(1)
            xa, xb = a.x()[: a.sim], b.x()[: b.sim]
            errs.append(float(np.abs(xa - xb).max() / np.abs(xa).max()))
        return errs

    errs = step_errors(5)
    log(f"FP32-split per-step error vs FP64: {errs}")

    # in-step kernel durations: the timed window again, every launch timed
    kt = {name: kernel_means(sims[name], x_w, v_w, args.steps) for name in ("fp64", "fp32_split")}
    log(f"in-step kernel means: {json.dumps(kt)}")

    # settled window: the FP64 run continued to the settle step, both precisions from there
    settle_step = args.settle_step if args.settle_step is not None else (150 if run_steps >= 200 else 75)
    settled = None
    if settle_step > 0:
        a = sims["fp64"]
        a.set_x(x_w)
        a.set_v(v_w)
        a.step(max(0, settle_step - args.warmup))  # the run from the warm-up state to the settle step
        x_s, v_s = a.x(), a.v()
        st = {}
        for name in ("fp64", "fp32_split"):
            sim = sims[name]
            sim.set_x(x_s)
            sim.set_v(v_s)
            secs, outer, recs = timed_steps(sim, args.settle_steps)
            st[name] = dict(seconds=secs, outer=outer, recs=recs)
            log(f"settled {name}: steps {settle_step + 1}-{settle_step + args.settle_steps} in {secs:.3f} s "
                f"({args.settle_steps / secs:.2f} steps/s, {outer} outer iterations)")
        kt_settled = {"fp32_split": kernel_means(sims["fp32_split"], x_s, v_s, args.settle_steps)}
        log(f"settled in-step kernel means: {json.dumps(kt_settled)}")
        a.set_x(x_s)
        a.set_v(v_s)
        serr = step_errors(3)
        settled = {
            "timed_steps": f"steps {settle_step + 1}-{settle_step + args.settle_steps} of the same run from rest "
                           f"(the FP64 run continued to step {settle_step}), each precision from that state, graphs "
                           f"{'on' if args.graphs else 'off'}",
            "fp32_split_steps_per_s": args.settle_steps / st["fp32_split"]["seconds"],
            "fp64_steps_per_s": args.settle_steps / st["fp64"]["seconds"],
            "fp32_split_speedup_vs_fp64": st["fp64"]["seconds"] / st["fp32_split"]["seconds"],
            "outer_iters_per_step": st["fp32_split"]["outer"] / args.settle_steps,
            "fp32_split_step_error_vs_fp64": max(serr),
            "kernel_means_us": kt_settled,
        }

    c = sims["fp64"].counters()
    # optional: back-to-back replays of one sweep (caches warm on one input)
    replay = None
    if args.replay:
        us64, bytes64 = sims["fp64"].time_kernel("sweep_fp64", 200)
        us32, bytes32 = sims["fp32_split"].time_kernel("sweep_fp32", 200)
        replay = {"sweep_fp64_us": us64, "sweep_fp32_us": us32, "note": "200 back-to-back launches, HIP events"}

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
    outer32 = res["fp32_split"]["outer"] / steps
    recs32 = float(np.mean(res["fp32_split"]["recs"]))
    slots = cloth_ring_slots(scene)
    step_bytes, spmv_bytes = algorithmic_step_bytes(c, slots, outer32, recs32, 4, scene["params"]["lg_iters"],
                                               scene["params"]["cheb_iters"])
    sweep = kt["fp32_split"]["sweep_fp32"]
    ach = spmv_bytes / (sweep["mean_us"] * 1e-6) / 1e9
    staging = sims["fp32_split"].sweep_staging()
    kname, slot_bytes = SWEEP32_KERNELS[staging]
    stored = c["nnz_r2"] * slot_bytes + 3 * c["sim_points"] * 4.0 * 4
    matrix_bytes = c["sell_entries"] * slot_bytes
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
            "timed_steps": f"steps {args.warmup + 1}-{args.warmup + steps} of a {run_steps}-step run from rest "
                           f"({args.warmup} warm-up steps, FP64), each precision from the same state; the steps "
                           f"after them were not run in this window (see settled)",
            "fp64_steps_per_s": steps / res["fp64"]["seconds"],
            "fp32_split_speedup_vs_fp64": res["fp64"]["seconds"] / t32,
            "fp32_split_step_error_vs_fp64": max(errs),
            "outer_iters_per_step": outer32,
            "settled": settled,
            "algorithmic_bytes_per_step": step_bytes,
            "algorithmic_GBs": step_bytes * steps / t32 / 1e9,
            "nnz_r2": c["nnz_r2"], "candidates_per_step": recs32, "patch_pairs": c["patch_pairs"],
            "patches": c["patches"],
            "static_static_pairs": "kept" if args.keep_static_pairs else
                                   "skipped (inert: they only write rows the step never reads)",
            "graphs": bool(args.graphs),
            "parallelism": "single GPU",
            "kernel_means_us": kt,
            "adjoint": adj,
        },
        "roofline": {
            "kernel": kname,
            "bound": "hbm",
            "achieved": ach,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS,
            "traffic": None if workload else
                       SWEEP32_TRAFFIC_BYTES.get((args.n, args.subdiv, args.sell_lanes, staging)),
            "traffic_source": "profiles/k_sweep_config1_hbm_traffic.md: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE "
                              "per launch, scaled by a 1 GiB copy; a recorded constant, not measured in this run",
            "avg_us": sweep["mean_us"],
            "avg_us_source": f"mean of {sweep['launches']} in-step launches, HIP events on each launch "
                             "(pdipc_kernel_timing, the timed window with direct launches)",
            "bytes_per_launch": spmv_bytes,
            "bytes_formula": "8 B per M^2 entry (FP32 value + int32 column) + 4 vectors x 3N x 4 B",
            "bytes_stored_format": stored,
            "frac_stored_format": stored / (sweep["mean_us"] * 1e-6) / 1e9 / HBM_PEAK_GBS,
            "mall_resident": matrix_bytes < 256e6,
            "mall_note": (f"the M^2 matrix (FP32 values + {int(slot_bytes - 4) * 8}-bit indices, "
                          f"{matrix_bytes / 1e6:.1f} MB) fits the 256 MB Infinity Cache: after the first sweep it is "
                          "served from MALL, so this fraction of HBM peak is not evidence of HBM-bound execution")
                         if matrix_bytes < 256e6 else
                         (f"the M^2 matrix (FP32 values + {int(slot_bytes - 4) * 8}-bit indices, "
                          f"{matrix_bytes / 1e6:.1f} MB) exceeds the 256 MB Infinity Cache: every sweep streams it "
                          "from HBM"),
            "fp64_sweep": {"avg_us": kt["fp64"]["sweep_fp64"]["mean_us"],
                           "achieved": (8 * c["nnz_r2"] + 4 * 3 * c["sim_points"] * 8)
                           / (kt["fp64"]["sweep_fp64"]["mean_us"] * 1e-6) / 1e9},
            "replay": replay,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
