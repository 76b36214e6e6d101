#!/usr/bin/env python3
"""Benchmark of the 3D TV-L1 flow path on one MI355X.

Workload (one "step"): the full-sequence batch — a synthetic 35-frame 80^3
FP32 phantom sequence, 34 forward + 34 backward consecutive-pair solves
(4 levels x 10 warps x 50 primal-dual iterations, conf/flow.yaml parameters
otherwise) run as one batched plan replayed as a hipGraph, the 3x3x3 median
post-filter of every flow (conf/flow.yaml enables it, processor.py:116-117),
then the propagation of the ventricle masks of frames 0 and 12 along the
forward and backward chains with the tricubic warp.  Frames are resident in
HBM before timing starts.

Printed as one JSON line: batched solves/s (the value), the plain per-pair
drop-in's solves/s, voxel-iterations/s, the HBM roofline fraction of the
fused sweep kernel (algorithmic bytes over its event-timed average launch
duration), the CPU oracle timed on the host cores, the reference's
step-by-step driver over the drop-in operators on the 80^3 pair (speed and
difference from the fused solve), and a serving leg: Poisson requests of the
five-size mix through the continuous-batching server, with p50/p99 latency,
offered load against measured capacity, and the plain per-request drop-in.
"""
import argparse
import json
import os
import sys
import time

METRIC = "80^3 TV-L1 flow solves/s (plain vs batched/served), p99 latency, HBM GB/s"
HBM_PEAK_GBPS = 8000.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=35)
    ap.add_argument("--size", type=int, default=80)
    ap.add_argument("--plain-pairs", type=int, default=6, help="pairs timed through the plain drop-in")
    ap.add_argument("--operator-pairs", type=int, default=1,
                    help="pairs timed through the reference's step-by-step driver over the drop-in operators")
    ap.add_argument("--sweep-launches", type=int, default=100, help="sweeps per level timed for the roofline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-stats", action="store_true", help="skip the per-iteration statistics leg")
    ap.add_argument("--sweep-shapes", default="",
                    help="fixed fused-sweep shapes 'bx,by,cz;...' per level instead of autotuning "
                         "(reproduces a measured configuration, e.g. under a profiler)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (reference kernel for PMC byte counters)")
    ap.add_argument("--workload", default="sequence",
                    choices=["sequence", "uneven", "serve", "unet", "augment", "train"],
                    help="sequence: 35-frame batch (the headline line); uneven: 96 pairs of mixed sizes; "
                         "serve: Poisson request service; unet: refinement training with the warp adjoint; "
                         "augment: flow-consistent rotate + flips of one 80^3 training sample; "
                         "train: mask-propagation training steps (lit_unet.py) on GPU-augmented batches")
    ap.add_argument("--pairs", type=int, default=96, help="uneven: number of pairs")
    ap.add_argument("--requests", type=int, default=100, help="serve: requests per arrival rate")
    ap.add_argument("--rates", default="5,20,50", help="serve: arrival rates (req/s)")
    ap.add_argument("--serve-requests", type=int, default=200,
                    help="sequence workload: requests of the serving leg (0 skips it)")
    ap.add_argument("--serve-rate", type=float, default=4.0, help="serving leg: Poisson arrival rate (req/s)")
    ap.add_argument("--serve-burst", type=int, default=30, help="serving leg: requests of the capacity burst")
    ap.add_argument("--serve-plain-requests", type=int, default=30,
                    help="serving leg: requests run through the plain per-request drop-in")
    ap.add_argument("--no-live-traffic", action="store_true",
                    help="take roofline.traffic from the committed counter summary instead of measuring it")
    ap.add_argument("--pmc-timeout", type=int, default=90,
                    help="seconds allowed for each child rocprofv3 counter run (the roofline traffic)")
    ap.add_argument("--pmc-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--quiet", action="store_true")
    return ap.parse_args()


PMC_SUMMARY = os.path.join("profiles", "k_sweep_68x80cubed_pmc.md")


def _run_child(cmd, timeout_s, cwd):
    """Runs cmd in its own process group; kills the group at the time limit.
    Returns (returncode or None on timeout, combined output)."""
    import signal
    import subprocess
    env = dict(os.environ, TMPDIR="/tmp")
    proc = subprocess.Popen(cmd, cwd=cwd, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                            start_new_session=True)
    try:
        out, _ = proc.communicate(timeout=timeout_s)
        return proc.returncode, out
    except subprocess.TimeoutExpired:
        os.killpg(proc.pid, signal.SIGKILL)
        out, _ = proc.communicate()
        return None, out


def _counter_rows(outdir, counter):
    """(dispatch id, kernel name, value) of one counter from a rocprofv3 CSV
    tree, and the CSV header seen (for the error message when nothing matches)."""
    import csv
    import glob
    rows, header = [], None
    for path in glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as f:
            reader = csv.DictReader(f)
            header = reader.fieldnames
            for r in reader:
                if r.get("Counter_Name") == counter:
                    rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"])))
    rows.sort()
    return rows, header


def per_level_counter(rows, resident, n):
    """Per-iteration counter value of every level and the 1 GiB copy's value,
    from (dispatch id, kernel name, value) rows in dispatch order: the last
    sweep dispatches are the probe's n sweeps per level in level order (one
    n-iteration launch at a resident level); the last copy is the calibration
    copy.  None when the rows hold fewer."""
    sweeps = [r for r in rows if "k_sweep" in r[1]]
    copies = [r for r in rows if "copyBuffer" in r[1]]
    counts = [1 if res else n for res in resident]
    if len(sweeps) < sum(counts) or not copies:
        return None
    sweeps = sweeps[len(sweeps) - sum(counts):]
    per_level, i = [], 0
    for lv, c in enumerate(counts):
        vals = [v for _, _, v in sweeps[i:i + c]]
        i += c
        per_level.append(sum(vals) / len(vals) / (n if resident[lv] else 1))
    return per_level, copies[-1][2]


def live_pmc_traffic(args, shapes, resident, algorithmic):
    """HBM bytes per fused-sweep iteration of every level, measured on the tuned
    shapes: this script runs twice as a child under rocprofv3 --pmc (FETCH_SIZE,
    then WRITE_SIZE) in its --pmc-probe mode, which builds the same plan, runs
    --sweep-launches sweeps of each level (one launch of that many iterations at
    a resident level) and a 1 GiB device copy.  The copy calibrates both
    counters (bytes = counter x 1 GiB / the copy's counter).  Returns a dict, or
    None and the reason."""
    import shutil
    import tempfile
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None, "rocprofv3 not found"
    n = args.sweep_launches
    spec = ";".join(",".join(str(v) for v in t) for t in shapes)
    me = os.path.abspath(__file__)
    per_counter = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        outdir = tempfile.mkdtemp(prefix="svs_pmc_", dir="/tmp")
        cmd = [prof, "--pmc", counter, "-d", outdir, "-o", "run", "--output-format", "csv", "--",
               sys.executable, me, "--pmc-probe", "--quiet", "--size", str(args.size), "--frames", str(args.frames),
               "--sweep-launches", str(n), "--sweep-shapes", spec]
        log(args, f"traffic: rocprofv3 --pmc {counter} on the sweeps of every level")
        rc, out = _run_child(cmd, args.pmc_timeout, outdir)
        rows, header = _counter_rows(outdir, counter)
        shutil.rmtree(outdir, ignore_errors=True)
        if "pmc-probe done" not in (out or ""):
            return None, f"rocprofv3 --pmc {counter} run did not complete (rc {rc})"
        got = per_level_counter(rows, resident, n)
        if got is None:
            return None, f"{counter}: too few sweep or copy dispatches in the counter output (columns {header})"
        per_counter[counter] = got
    gib = float(1 << 30)
    fetch, f_copy = per_counter["FETCH_SIZE"]
    write, w_copy = per_counter["WRITE_SIZE"]
    levels = [f * gib / f_copy + w * gib / w_copy for f, w in zip(fetch, write)]
    return {"traffic": int(sum(levels) / len(levels)), "per_level": [int(v) for v in levels],
            "per_level_over_algorithmic": [round(v / a, 3) for v, a in zip(levels, algorithmic)],
            "FETCH_SIZE_per_level": [round(v, 1) for v in fetch], "WRITE_SIZE_per_level": [round(v, 1) for v in write],
            "calibration": {"copy_bytes": int(gib), "FETCH_SIZE_copy": f_copy, "WRITE_SIZE_copy": w_copy}}, None


def pmc_probe(args):
    """--pmc-probe: the sweeps live_pmc_traffic counts (run under rocprofv3)."""
    import torch
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from singleventriclesegmentation_amd.solver import Plan, SolverConfig
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    shape = (args.size,) * 3
    npairs = 2 * (args.frames - 1)
    cfg = SolverConfig(num_scales=4, max_warps=1, max_outer_iterations=2)
    frames = torch.rand((2,) + shape, device=dev)
    plan = Plan(cfg, [shape] * npairs, device=dev)
    for k in range(npairs):
        plan.bind(k, frames[0], frames[1])
    plan.run(use_graph=False)
    plan.set_shapes([tuple(int(v) for v in t.split(",")) for t in args.sweep_shapes.split(";")])
    for lv in range(cfg.num_scales):
        plan.sweeps(lv, args.sweep_launches)
    src = torch.empty(1 << 28, dtype=torch.float32, device=dev).uniform_()
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    plan.close()
    print("pmc-probe done", flush=True)


def pmc_traffic(size, npairs):
    """HBM bytes per fused-sweep launch (mean over the levels) measured with
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on this workload, scaled against a
    device copy; read from the committed summary (counters cannot be read from
    inside the timed process).  None for any other workload."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), PMC_SUMMARY)
    if size != 80 or npairs != 68 or not os.path.exists(path):
        return None, None
    vals = {}
    with open(path) as f:
        for line in f:
            key, _, val = line.partition(":")
            if key in ("traffic_bytes_per_launch_avg", "traffic_shapes"):
                vals[key] = val.strip()
    if "traffic_bytes_per_launch_avg" not in vals:
        return None, None
    return int(vals["traffic_bytes_per_launch_avg"]), vals.get("traffic_shapes")


def log(args, *a):
    if not args.quiet:
        print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.pmc_probe:
        return pmc_probe(args)
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on exactly one GPU (got --gpus {args.gpus}, WORLD_SIZE {world}): frame pairs are "
              "independent, so more GPUs would each run this same batch on their own share.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from singleventriclesegmentation_amd import TVL13DOpticalFlow, Warp, phantom, propagate
    from singleventriclesegmentation_amd.flow_utils import apply_median_filter, median_filter3d
    from singleventriclesegmentation_amd.solver import (BYTES_PER_VOXEL_ITERATION, Plan, SolverConfig,
                                                        algorithmic_bytes)

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.workload != "sequence":
        return OTHER[args.workload](args, dev)
    shape = (args.size,) * 3
    T = args.frames
    cfg = SolverConfig(num_scales=4, max_warps=10, max_outer_iterations=50)

    log(args, f"synthesising {T} frames of {shape}")
    seq_np = phantom.sequence(shape, T)
    masks_np = phantom.sequence_masks(shape, T, which=(0, 12))
    seq = torch.from_numpy(seq_np).to(dev)
    m0 = torch.from_numpy(masks_np[0].astype(np.float32)).to(dev)[None, None]
    m12 = torch.from_numpy(masks_np[12].astype(np.float32)).to(dev)[None, None]

    npairs = 2 * (T - 1)
    plan = Plan(cfg, [shape] * npairs, device=dev)
    u_all = torch.empty((npairs,) + shape + (3,), device=dev)
    for t in range(T - 1):
        plan.bind(t, seq[t + 1], seq[t], u_out=u_all[t])                   # forward: I0 = f[t+1], I1 = f[t]
        plan.bind(T - 1 + t, seq[t], seq[t + 1], u_out=u_all[T - 1 + t])    # backward: I0 = f[t], I1 = f[t+1]
    warper = Warp("CUBIC_HERMITESPLINE", "MIRROR", "numDofs", (1, 1, 1))

    filtered = {}

    def step():
        plan.run(use_graph=True)
        uf = median_filter3d(u_all, 3)    # every flow, one launch
        filtered["u"] = uf
        fwd = uf[0:12]                    # flows 0->1 ... 11->12
        bwd = uf[T - 1:T - 1 + 12].flip(0)  # flows 12->11 ... 1->0
        return propagate(warper, m0, m12, fwd, bwd)

    step()
    if args.sweep_shapes:
        tuned = [tuple(int(v) for v in t.split(",")) for t in args.sweep_shapes.split(";")]
        plan.set_shapes(tuned)
        log(args, "fixed sweep shapes (bx, by, cz) per level:", tuned)
    else:
        log(args, "warm-up and sweep-shape autotuning")
        tuned = plan.autotune(reps=5)
        log(args, "tuned sweep shapes (bx, by, cz) per level:", tuned)
    for _ in range(args.warmup):
        step()  # the first step after tuning re-captures the graph
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        fm, bm = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / max(args.steps, 1)
    solves_per_s = npairs * args.steps / elapsed
    vox_it = cfg.voxel_iterations(shape)
    log(args, f"batched: {ms_per_step:.1f} ms/step, {solves_per_s:.1f} solves/s")

    # The same steps with per-iteration statistics on (the residual norms the
    # reference computes each iteration, and the energy): their cost, and the
    # finest level's values at the first and last iteration of its last warp.
    stats = None
    if not args.no_stats:
        plan.set_stats(True)
        step()  # captures the graph again, with the statistics launches
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        st_ms = 1e3 * (time.perf_counter() - t0) / max(args.steps, 1)
        last_warp = plan.stats[0, -1].double()  # [iterations, instances, 3]
        stats = {"ms_per_step": round(st_ms, 3), "overhead": round(st_ms / ms_per_step - 1.0, 4)}
        for name, it in (("first", 0), ("last", -1)):
            v = last_warp[it].mean(0)
            stats[f"finest_last_warp_{name}_iteration_mean"] = {
                "primal_residual": float(v[0]), "dual_residual": float(v[1]), "energy": float(v[2])}
        plan.set_stats(False)
        log(args, f"with per-iteration statistics: {st_ms:.1f} ms/step ({100 * stats['overhead']:+.1f} %)")

    # Fused sweep kernel on its own: event-timed launches per pyramid level on
    # the stream the kernel runs on (torch's current stream).
    levels = cfg.level_shapes(shape)
    sweep_bytes = []
    sweep_ms = []
    stream = torch.cuda.current_stream()
    for lv, (a, b, c) in enumerate(levels):
        n = args.sweep_launches
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        plan.sweeps(lv, 4)
        e0.record(stream)
        plan.sweeps(lv, n)
        e1.record(stream)
        e1.synchronize()
        sweep_ms.append(e0.elapsed_time(e1) / n)
        sweep_bytes.append(BYTES_PER_VOXEL_ITERATION * a * b * c * npairs)
    avg_bytes = sum(sweep_bytes) / len(levels)
    avg_ms = sum(sweep_ms) / len(levels)
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
    traffic = None
    traffic_detail = None
    traffic_source = None
    if not args.no_live_traffic:
        resident = [plan.info(lv)["resident"] > 0 for lv in range(len(levels))]
        live, why = live_pmc_traffic(args, tuned, resident, [b for b in sweep_bytes])
        if live is not None:
            traffic = live["traffic"]
            traffic_detail = live
            traffic_source = ("measured in this run: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE on the tuned sweep "
                              "shapes, each scaled by what it reads for a 1 GiB device copy; per iteration, mean "
                              "of the levels")
            log(args, "traffic per level (bytes per iteration):", live["per_level"],
                "over algorithmic:", live["per_level_over_algorithmic"], "; counters per iteration FETCH_SIZE",
                live["FETCH_SIZE_per_level"], "WRITE_SIZE", live["WRITE_SIZE_per_level"], "; 1 GiB copy",
                live["calibration"])
        else:
            log(args, f"traffic: not measured live ({why})")
    if traffic is None:
        traffic, traffic_shapes = pmc_traffic(args.size, npairs)
        if traffic:
            traffic_source = (f"{PMC_SUMMARY}: 2 x FETCH_SIZE + WRITE_SIZE per launch, mean of the levels, "
                              f"sweep shapes {traffic_shapes}")
    level0_gbps = sweep_bytes[0] / (sweep_ms[0] * 1e-3) / 1e9
    log(args, "sweep per level (us/launch):", [round(1e3 * v, 1) for v in sweep_ms])

    # Plain per-pair drop-in (TVL13DOpticalFlow.compute, one pair at a time).
    plain = TVL13DOpticalFlow(4, 10, 50, 1.0, 0.025, 2, (0.5, 0.5, 0.25, 1.0), True, (5.0, 0.5), True, 1.0,
                              "CUBIC_HERMITESPLINE", "MIRROR", "numDofs", (1, 1, 1), dev, autotune=True)
    plain_sps = None
    if args.plain_pairs > 0:
        plain.compute(seq[1], seq[0], None, None)  # first call also autotunes the single-pair plan
        plain.compute(seq[1], seq[0], None, None)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for j in range(args.plain_pairs):
            u, _ = plain.compute(seq[j + 1], seq[j], None, None)
            u = apply_median_filter(u, 3, dev)
        torch.cuda.synchronize()
        plain_sps = args.plain_pairs / (time.perf_counter() - t0)
        # the batched instance computes the same pair: identical bits
        same = bool(torch.equal(u, filtered["u"][args.plain_pairs - 1]))
        log(args, f"plain drop-in: {plain_sps:.2f} solves/s; equal to batched result: {same}")

    if args.copy_mib > 0:
        src = torch.empty(args.copy_mib << 18, dtype=torch.float32, device=dev).uniform_()
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()

    # The single-pair workload with a known displacement (frame1 = frame0
    # warped by d, fresh noise): the drop-in's flow against d, and against the
    # FP32 CPU oracle solving the same pair (which is also the timed baseline).
    f0, f1, disp = phantom.frame_pair(shape)
    I0_pair, I1_pair = torch.from_numpy(f1).to(dev), torch.from_numpy(f0).to(dev)
    u_pair_dev, _ = plain.compute(I0_pair, I1_pair, None, None)
    u_pair = u_pair_dev.cpu().numpy()
    accuracy = {"pair_mean_epe_vs_true_displacement": round(float(np.linalg.norm(u_pair - disp, axis=-1).mean()), 4),
                "pair_max_displacement": round(float(np.linalg.norm(disp, axis=-1).max()), 3)}

    # The reference's own driver structure (tvl13d.py:122-255: one operator
    # call per reference call, torch's elementwise and sum kernels between
    # them, ~15 launches per iteration) over the drop-in operators, on the
    # same pair: the per-pair path a user gets by swapping the import only.
    op_sps = None
    if args.operator_pairs > 0:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.operator_pairs):
            uo, _ = plain.compute_unfused(I0_pair, I1_pair, None, None)
        torch.cuda.synchronize()
        op_sps = args.operator_pairs / (time.perf_counter() - t0)
        d = (uo.double() - u_pair_dev.double())
        accuracy["operator_path_max_abs_vs_fused"] = float(d.abs().max().item())
        accuracy["operator_path_mean_epe_vs_fused"] = float(d.pow(2).sum(-1).sqrt().mean().item())
        accuracy["operator_path_voxels_above_1e-3"] = int((d.abs().amax(-1) > 1e-3).sum().item())
        log(args, f"operator path: {op_sps:.3f} solves/s; vs fused: max |du| = "
                  f"{accuracy['operator_path_max_abs_vs_fused']:.2e}, mean EPE "
                  f"{accuracy['operator_path_mean_epe_vs_fused']:.2e}, voxels above 1e-3: "
                  f"{accuracy['operator_path_voxels_above_1e-3']}")
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        from singleventriclesegmentation_amd.flow_utils import gaussian_weights3d
        cores = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
        O.lib()
        w = gaussian_weights3d(1.0)
        params = O.solve_params(num_scales=4, max_warps=10, max_outer_iterations=50)
        # The timed sample: three of the batch's own 68 pair solves (a forward
        # pair, a later forward pair, a backward pair), each checked against
        # the batched GPU flow of that pair.
        sample = [0, 17, T - 1 + 16]
        seq_cpu = seq.cpu().numpy()
        cpu_diff = 0.0
        t0 = time.perf_counter()
        for k in sample:
            t = k if k < T - 1 else k - (T - 1)
            a_, b_ = (seq_cpu[t + 1], seq_cpu[t]) if k < T - 1 else (seq_cpu[t], seq_cpu[t + 1])
            u_cpu, _ = O.solve(O.blur(a_, w), O.blur(b_, w), params)
            cpu_diff = max(cpu_diff, float(np.abs(u_all[k].cpu().numpy() - u_cpu).max()))
        cpu_s = (time.perf_counter() - t0) / len(sample)
        accuracy["sequence_pairs_max_abs_vs_fp32_oracle"] = cpu_diff
        u_cpu, _ = O.solve(O.blur(f1, w), O.blur(f0, w), params)
        accuracy["pair_max_abs_vs_fp32_oracle"] = float(np.abs(u_pair - u_cpu).max())
        model = ""
        try:
            with open("/proc/cpuinfo") as f:
                model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
        except OSError:
            pass
        cpu = {"value": round(1.0 / cpu_s, 4), "unit": "solves/s", "cores": cores, "kind": "port",
               "voxel_iterations_per_s": round(vox_it / cpu_s, 1), "cpu_model": model,
               "sample": f"3 of the batch's 68 pair solves (forward pairs 0->1 and 17->18, backward pair 17->16; "
                         f"4x10x50 schedule, blur included) solved in turn by the FP32 C++ oracle on {cores} "
                         f"OpenMP threads: {cpu_s:.2f} s per solve; each equals the batched GPU flow "
                         f"(max |du| = {cpu_diff:.1e})"}
        log(args, f"cpu oracle: {cpu_s:.2f} s per solve on {cores} threads; GPU vs oracle max |du| = "
                  f"{cpu_diff:.2e} on 3 batch pairs, {accuracy['pair_max_abs_vs_fp32_oracle']:.2e} on the "
                  f"known-displacement pair")
    log(args, f"pair flow vs true displacement: mean EPE {accuracy['pair_mean_epe_vs_true_displacement']} voxel")

    out = {
        "metric": METRIC,
        "value": round(solves_per_s, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic phantom sequence (blob ventricle, periodic contraction, Gaussian noise)",
        "config": {"workload": f"full-sequence batch: {T} frames of {shape[0]}^3, {T - 1} forward + {T - 1} "
                               f"backward pair solves in one batched hipGraph plan + 3^3 median filter + mask propagation 0<->12",
                   "solves_per_step": npairs, "levels": cfg.num_scales, "warps": cfg.max_warps,
                   "iterations": cfg.max_outer_iterations, "shape_zyx": list(shape),
                   "sweep_shapes_bx_by_cz": [list(t) for t in tuned]},
        "plain_solves_per_s": round(plain_sps, 3) if plain_sps else None,
        "batched_over_plain": round(solves_per_s / plain_sps, 3) if plain_sps else None,
        "operator_path_solves_per_s": round(op_sps, 4) if op_sps else None,
        "batched_over_operator_path": round(solves_per_s / op_sps, 1) if op_sps else None,
        "voxel_iterations_per_s": round(solves_per_s * vox_it, 1),
        "algorithmic_GBps": round(solves_per_s * algorithmic_bytes(cfg, shape) / 1e9, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                     "traffic_source": traffic_source,
                     "traffic_per_level": traffic_detail["per_level"] if traffic_detail else None,
                     "traffic_over_algorithmic_per_level": (traffic_detail["per_level_over_algorithmic"]
                                                            if traffic_detail else None),
                     "kernel": "k_sweep (fused primal-dual sweep; k_sweep_resident, all iterations of a warp in "
                               "one launch, at levels of at most 2048 voxels per instance)",
                     "bytes_per_launch_avg": int(avg_bytes), "us_per_launch_avg": round(avg_ms * 1e3, 2),
                     "level0_GBps": round(level0_gbps, 1),
                     "us_per_launch_by_level": [round(v * 1e3, 2) for v in sweep_ms]},
        "cpu_baseline": cpu,
        "accuracy": accuracy,
        "iteration_stats": stats,
    }
    if args.serve_requests > 0:
        out["served"] = serve_leg(args, dev)
    print(json.dumps(out), flush=True)


def serve_leg(args, dev):
    """Continuous batching under load: Poisson arrivals of 70% single pairs and
    30% 35-frame forward + backward sequences with mask propagation, sizes
    from the five-size mix, a 64 GB HBM budget.  Reports the served solves/s
    and p50/p99 latency at one offered load below capacity, the capacity
    (all of a burst of requests submitted at once), and the plain per-request
    drop-in's service times on the first requests of the same stream."""
    import numpy as np
    import torch
    from singleventriclesegmentation_amd import TVL13DOpticalFlow, phantom
    from singleventriclesegmentation_amd.serve import FlowServer, make_pool, poisson_requests, run_plain, run_service
    from singleventriclesegmentation_amd.solver import SolverConfig
    cfg = SolverConfig(num_scales=4, max_warps=10, max_outer_iterations=50)
    sizes = list(phantom.UNEVEN_SIZES)
    log(args, "serve: building the request pool (35-frame sequences of the five sizes)")
    pool = make_pool(sizes, args.frames, dev)
    buckets = (1, 2, 4)
    server = FlowServer(cfg, device=dev, hbm_budget_bytes=64 << 30, slots=4, pair_buckets=buckets)
    server.warm({sh: (pool[sh][0][1], pool[sh][0][0]) for sh in sizes}, frames=args.frames, autotune=False)
    run_service(server, poisson_requests(0, 6, sizes, pool, args.frames, seed=99))

    burst = poisson_requests(0, args.serve_burst, sizes, pool, args.frames, seed=5)
    _, span = run_service(server, burst)
    capacity = sum(r.solves for r in burst) / span
    log(args, f"serve: capacity {capacity:.1f} solves/s ({len(burst)} requests submitted at once, {span:.1f} s)")

    reqs = poisson_requests(args.serve_rate, args.serve_requests, sizes, pool, args.frames, seed=4)
    lat, span = run_service(server, reqs)
    solves = sum(r.solves for r in reqs)
    arrival_span = max(r.arrival for r in reqs)
    offered = solves / arrival_span
    def ms(values, q):
        return round(1e3 * float(np.percentile(values, q)), 1) if len(values) else None

    pairs = [v for v, r in zip(lat, reqs) if r.kind == "pair"]
    seqs = [v for v, r in zip(lat, reqs) if r.kind != "pair"]
    served = {"rate_req_per_s": args.serve_rate, "requests": len(reqs), "solves": solves,
              "served_solves_per_s": round(solves / span, 2), "p50_ms": ms(lat, 50), "p99_ms": ms(lat, 99),
              "p50_ms_pairs": ms(pairs, 50), "p99_ms_pairs": ms(pairs, 99),
              "p50_ms_sequences": ms(seqs, 50), "p99_ms_sequences": ms(seqs, 99),
              "offered_solves_per_s": round(offered, 2), "capacity_solves_per_s": round(capacity, 2),
              "offered_over_capacity": round(offered / capacity, 3),
              "policy": {"slots": len(server.streams), "sequence_slots": server.sequence_slots,
                         "pairs_first": server.pairs_first, "pair_buckets": list(buckets)}}
    server.close()
    log(args, f"serve: {served}")

    # The plain per-request drop-in on the first requests of the same stream,
    # with its per-size plans created before timing.  Its service time per
    # request is reported; at this arrival rate its queue would grow without
    # bound when its capacity is below the offered load.
    plain = TVL13DOpticalFlow(4, 10, 50, 1.0, 0.025, 2, (0.5, 0.5, 0.25, 1.0), True, (5.0, 0.5), True, 1.0,
                              "CUBIC_HERMITESPLINE", "MIRROR", "numDofs", (1, 1, 1), dev)
    for sh in sizes:
        plain.compute(pool[sh][0][1], pool[sh][0][0], None, None)
    torch.cuda.synchronize()
    first = poisson_requests(0, args.serve_plain_requests, sizes, pool, args.frames, seed=4)  # back to back
    lat, span = run_plain(plain, first)
    service = np.diff(np.concatenate([[0.0], np.sort(lat)]))   # requests ran one after another from t = 0
    plain_sps = sum(r.solves for r in first) / span
    served["plain_requests"] = len(first)
    served["plain_solves_per_s"] = round(plain_sps, 2)
    served["plain_service_p50_ms"] = ms(service, 50)
    served["plain_service_p99_ms"] = ms(service, 99)
    served["plain_offered_over_capacity"] = round(offered / plain_sps, 3)
    served["served_over_plain"] = round(served["served_solves_per_s"] / plain_sps, 3)
    served["capacity_over_plain"] = round(capacity / plain_sps, 3)
    log(args, f"serve: plain drop-in {plain_sps:.1f} solves/s on {len(first)} requests back to back")
    return served


def _line(args, metric, value, unit, ms_per_step, config, extra, higher=True):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": ms_per_step, "higher_is_better": higher, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic phantoms", "config": config}
    out.update(extra)
    print(json.dumps(out), flush=True)


def bench_uneven(args, dev):
    """96 frame pairs of sizes drawn from the five-size mix, one batched plan."""
    import numpy as np
    import torch
    from singleventriclesegmentation_amd import opticalFlow as OF
    from singleventriclesegmentation_amd import phantom
    from singleventriclesegmentation_amd.solver import (BYTES_PER_VOXEL_ITERATION, Plan, SolverConfig,
                                                        algorithmic_bytes)
    cfg = SolverConfig(num_scales=4, max_warps=10, max_outer_iterations=50)
    shapes = phantom.uneven_shapes(args.pairs, seed=3)
    frames = {}
    for sh in sorted(set(shapes)):
        clean = phantom.clean_frame(sh)
        d = torch.from_numpy(phantom.displacement(sh, peak=3.0)).to(dev)
        op = OF.WarpingCNN3D(OF.MeshInfo3D(*sh, sh[0] - 1, sh[1] - 1, sh[2] - 1), 2, 1)
        moved = op.forward(torch.from_numpy(clean).to(dev)[None, None], d[None].contiguous())[0, 0]
        f0 = torch.from_numpy(phantom.add_noise(clean, 0)).to(dev)
        f1 = (moved + torch.from_numpy(np.random.default_rng(1).normal(0, 0.02, sh).astype(np.float32)).to(dev))
        frames[sh] = (f1.clamp(0, 1).contiguous(), f0)
    plan = Plan(cfg, shapes, device=dev)
    outs = [torch.empty(sh + (3,), device=dev) for sh in shapes]
    for k, sh in enumerate(shapes):
        plan.bind(k, frames[sh][0], frames[sh][1], u_out=outs[k])
    plan.run(use_graph=True)
    if args.sweep_shapes:
        tuned = [tuple(int(v) for v in t.split(",")) for t in args.sweep_shapes.split(";")]
        plan.set_shapes(tuned)
    else:
        tuned = plan.autotune(reps=5)
    for _ in range(args.warmup):
        plan.run(use_graph=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.run(use_graph=True)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    vox_it = sum(cfg.voxel_iterations(sh) for sh in shapes)
    abytes = sum(algorithmic_bytes(cfg, sh) for sh in shapes)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    plan.sweeps(0, args.sweep_launches)
    e1.record()
    e1.synchronize()
    l0_us = e0.elapsed_time(e1) * 1e3 / args.sweep_launches
    l0_bytes = BYTES_PER_VOXEL_ITERATION * sum(a * b * c for a, b, c in shapes)
    _line(args, "uneven-size batch: TV-L1 solves/s", round(len(shapes) * args.steps / el, 3), "solves/s",
          round(1e3 * el / args.steps, 3),
          {"workload": f"{len(shapes)} pairs, sizes from {{48^3, 64x64x40, 80^3, 96x96x64, 128x128x48}}, 4x10x50",
           "sizes_zyx": sorted(set(shapes)), "sweep_shapes_bx_by_cz": [list(t) for t in tuned]},
          {"voxel_iterations_per_s": round(vox_it * args.steps / el, 1),
           "algorithmic_GBps": round(abytes * args.steps / el / 1e9, 1),
           "level0_sweep_GBps": round(l0_bytes / (l0_us * 1e-6) / 1e9, 1), "level0_sweep_us": round(l0_us, 1)})


def bench_serve(args, dev):
    """Poisson arrivals (70% pairs, 30% 35-frame sequences) against a 64 GB budget."""
    import numpy as np
    import torch
    from singleventriclesegmentation_amd import TVL13DOpticalFlow, phantom
    from singleventriclesegmentation_amd.serve import FlowServer, make_pool, poisson_requests, run_plain, run_service
    from singleventriclesegmentation_amd.solver import SolverConfig
    cfg = SolverConfig(num_scales=4, max_warps=10, max_outer_iterations=50)
    sizes = list(phantom.UNEVEN_SIZES)
    log(args, "building request pool")
    pool = make_pool(sizes, args.frames, dev)
    server = FlowServer(cfg, device=dev, hbm_budget_bytes=64 << 30, slots=4)
    log(args, "warm-up: creating and tuning the plans of every size and bucket")
    server.warm({sh: (pool[sh][0][1], pool[sh][0][0]) for sh in sizes}, frames=args.frames)
    warm = poisson_requests(0, 6, sizes, pool, args.frames, seed=99)
    run_service(server, warm)
    burst = poisson_requests(0, args.serve_burst, sizes, pool, args.frames, seed=5)
    _, bspan = run_service(server, burst)
    capacity = sum(r.solves for r in burst) / bspan
    log(args, f"capacity {capacity:.1f} solves/s ({len(burst)} requests submitted at once)")
    res = {}
    total_solves, total_time = 0, 0.0
    for rate in [float(r) for r in args.rates.split(",")]:
        reqs = poisson_requests(rate, args.requests, sizes, pool, args.frames, seed=4)
        lat, span = run_service(server, reqs)
        solves = sum(r.solves for r in reqs)
        total_solves += solves
        total_time += span
        offered = solves / max(r.arrival for r in reqs)
        res[f"{rate:g}"] = {"p50_s": round(float(np.percentile(lat, 50)), 3),
                            "p99_s": round(float(np.percentile(lat, 99)), 3),
                            "solves_per_s": round(solves / span, 2), "requests": len(reqs), "solves": solves,
                            "offered_solves_per_s": round(offered, 1),
                            "offered_over_capacity": round(offered / capacity, 2)}
        log(args, f"rate {rate}: {res[f'{rate:g}']}")
    server.close()
    # The plain per-request drop-in on the same request set, its per-size plans
    # created (and their graphs captured) before timing, as a long-running plain
    # service would have them.
    plain = TVL13DOpticalFlow(4, 10, 50, 1.0, 0.025, 2, (0.5, 0.5, 0.25, 1.0), True, (5.0, 0.5), True, 1.0,
                              "CUBIC_HERMITESPLINE", "MIRROR", "numDofs", (1, 1, 1), dev)
    for sh in sizes:
        plain.compute(pool[sh][0][1], pool[sh][0][0], None, None)
    torch.cuda.synchronize()
    log(args, "plain drop-in: one request at a time")
    reqs = poisson_requests(5.0, args.requests, sizes, pool, args.frames, seed=4)
    lat, span = run_plain(plain, reqs)
    plain_res = {"rate": 5.0, "requests": len(reqs), "p50_s": round(float(np.percentile(lat, 50)), 3),
                 "p99_s": round(float(np.percentile(lat, 99)), 3),
                 "solves_per_s": round(sum(r.solves for r in reqs) / span, 2)}
    _line(args, "served TV-L1 flow: solves/s with p50/p99 latency", round(total_solves / total_time, 3), "solves/s",
          None, {"workload": f"Poisson arrivals at {args.rates} req/s, {args.requests} requests each; 70% pairs, "
                             f"30% {args.frames}-frame fwd+bwd sequences with mask propagation; sizes from the "
                             "five-size mix; 64 GB budget; 4 stream slots"},
          {"by_rate": res, "capacity_solves_per_s": round(capacity, 2), "plain_drop_in": plain_res})


def bench_unet(args, dev):
    """Refinement training: batch 4 of 80^3, U-Net 4 levels / 16 base channels, warp adjoint in the loss."""
    import torch
    from singleventriclesegmentation_amd.refinement import Refiner, RefinementConfig, synthetic_batch
    from singleventriclesegmentation_amd.unet import parameter_count
    shape = (args.size,) * 3
    # RefinementConfig.conv_search (on) lets MIOpen search its convolution
    # algorithms during the warm-up steps; FP32 throughout.
    batch = synthetic_batch(shape, 4, dev)
    ref = Refiner(shape, 4, dev, RefinementConfig())
    for _ in range(args.warmup):
        ref.step(*batch)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = ref.step(*batch)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    _line(args, "U-Net refinement training steps/s (tricubic warp + adjoint in the loss)",
          round(args.steps / el, 3), "steps/s", round(1e3 * el / args.steps, 3),
          {"workload": f"batch 4 x 2 channels x {shape[0]}^3, 4-class softmax, consistency loss through the warp",
           "params": parameter_count(ref.model)},
          {"loss": round(float(out["total"].item()), 4)})


AUG_PMC_SUMMARY = os.path.join("profiles", "k_spatial_rotate_35frames_80cubed_pmc.md")


def _summary_value(path, key):
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), path)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            k, _, v = line.partition(":")
            if k.strip() == key:
                return v.strip()
    return None


def bench_augment(args, dev):
    """Flow-consistent augmentation of one training sample (TransformsWarper's
    spatial transforms): RandomRotate + the three RandomFlips on the image
    sequence, both masks and both flow sequences, fused into one launch."""
    import numpy as np
    import torch
    from singleventriclesegmentation_amd import transforms as T

    shape = (args.size,) * 3
    F = args.frames
    g = torch.Generator(device=dev).manual_seed(0)
    data = {"image": torch.rand((F, 1) + shape, device=dev, generator=g),
            "med": (torch.rand((1,) + shape, device=dev, generator=g) > 0.5).float(),
            "mes": (torch.rand((1,) + shape, device=dev, generator=g) > 0.5).float(),
            "fwd_flow": torch.randn((F - 1, 3) + shape, device=dev, generator=g),
            "bwd_flow": torch.randn((F - 1, 3) + shape, device=dev, generator=g)}
    keys = list(data)
    modes = {"image": "bilinear", "med": "nearest", "mes": "nearest", "fwd_flow": "bilinear",
             "bwd_flow": "bilinear"}
    channels = sum(int(np.prod(v.shape[:-3])) for v in data.values())
    angles = (37.0, 211.0, 118.0)
    rates = {}
    for grid in ("device", "host"):
        rot = T.RandomRotate(keys, 1.0, shape, [(0, 360)] * 3, "zeros", modes, grid=grid, align_corners=False)
        rot.R = torch.from_numpy(T.rotz(angles[2]) @ T.roty(angles[1]) @ T.rotx(angles[0])).float()
        step = lambda: T.spatial_pass(dict(data), keys, rotate=rot, flips=7)   # noqa: E731
        for _ in range(max(1, args.warmup)):
            step()
        torch.cuda.synchronize()
        n = max(args.steps, 10)
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        start.record()
        for _ in range(n):
            step()
        end.record()
        torch.cuda.synchronize()
        rates[grid] = (n / (time.perf_counter() - t0), start.elapsed_time(end) / n)
        log(args, f"augment grid={grid}: {rates[grid][0]:.1f} samples/s, {rates[grid][1]:.3f} ms per sample (events)")
    if args.copy_mib > 0:
        src = torch.empty(args.copy_mib << 18, dtype=torch.float32, device=dev).uniform_()
        torch.empty_like(src).copy_(src)
        torch.cuda.synchronize()
    # Algorithmic bytes of one launch: every output channel-voxel written once,
    # and read once where its sample point lies inside the source volume.
    R = (T.rotz(angles[2]) @ T.roty(angles[1]) @ T.rotx(angles[0])).astype(np.float64)
    lin = [np.linspace(-1, 1, s) for s in shape]
    zz, yy, xx = np.meshgrid(*lin, indexing="ij")
    pts = np.stack([xx, yy, zz], -1) @ R
    size = np.array(shape[::-1], np.float64)
    src = ((pts + 1.0) * size - 1.0) / 2.0
    inside = int(np.all((src > -1.0) & (src < size), axis=-1).sum())
    nvox = int(np.prod(shape))
    alg_bytes = channels * 4 * (nvox + inside)
    ms = rates["device"][1]
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    # Counters cannot be read inside the timed process: traffic comes from the
    # committed summary, measured on this workload (80^3, 35 frames).
    traffic = _summary_value(AUG_PMC_SUMMARY, "traffic_bytes_per_launch") if (shape[0], F) == (80, 35) else None
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        xs = {k: (v[None] if v.ndim == 4 else v).cpu().numpy() for k, v in data.items()}
        axes = [torch.linspace(-1, 1, s).numpy() for s in shape]
        Rf = np.ascontiguousarray(rot.R.numpy(), dtype=np.float32)
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            for k, x in xs.items():
                O.spatial_augment(x, Rf, (2, 3, 4), axes, modes[k], "zeros", False, vector=k in T.FLOWS_KEYS)
            times.append(time.perf_counter() - t0)
        el = sorted(times)[1]
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = {"value": round(1.0 / el, 4), "unit": "samples/s", "cores": cores, "kind": "port",
               "sample": f"the whole sample ({channels} channels) through the C++ oracle's rotate + 3 flips (the "
                         f"reference's sequence), median of 3 runs on {cores} OpenMP threads: {el:.2f} s"}
        log(args, f"cpu oracle: {el:.2f} s per sample")
    _line(args, "flow-consistent augmentation samples/s (rotate + 3 flips, all keys of a TransformsWarper sample)",
          round(rates["device"][0], 2), "samples/s", round(rates["device"][1], 4),
          {"workload": f"{F} frames + 2 masks + 2 x {F - 1} flows of {shape[0]}^3 ({channels} channels), "
                       f"RandomRotate (bilinear/nearest, zeros, align_corners False) + flips of Z, Y, X in one "
                       f"launch", "grid": "device", "rotation_deg_xyz": list(angles)},
          {"host_grid_samples_per_s": round(rates["host"][0], 2),
           "host_grid_ms_per_sample": round(rates["host"][1], 4),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBPS, 4),
                        "traffic": int(traffic) if traffic else None,
                        "traffic_source": f"{AUG_PMC_SUMMARY}: 2 x FETCH_SIZE + WRITE_SIZE" if traffic else None,
                        "kernel": "k_spatial_rotate", "bytes_per_launch": alg_bytes,
                        "inside_fraction": round(inside / nvox, 4), "us_per_launch": round(ms * 1e3, 2)},
           "cpu_baseline": cpu})


def bench_train(args, dev):
    """Mask-propagation training (svs/models/lit_unet.py with conf/model/unet.yaml):
    batch 4 of 80^3 patients with ED->ES chains of 8-11 flows (flows from the
    batched TV-L1 solver), every sample augmented on the GPU (rotation + 3
    flips) and collated, then both propagation chains, the loss, backward and
    the Adam step."""
    import numpy as np
    import torch
    from singleventriclesegmentation_amd import training as TR
    from singleventriclesegmentation_amd import transforms as T
    from singleventriclesegmentation_amd.solver import SolverConfig

    shape = (args.size,) * 3
    t0 = time.perf_counter()
    data = TR.SyntheticPropagationData(shape, frames=12, n=4, device=dev, gaps=(8, 9, 10, 11),
                                       solver_cfg=SolverConfig(num_scales=4, max_warps=10, max_outer_iterations=50))
    torch.cuda.synchronize()
    log(args, f"train: 76 flows of {shape} solved in {time.perf_counter() - t0:.2f} s")
    keys = ["image", "mi", "mf", "fwd_flow", "bwd_flow"]
    modes = {"image": "bilinear", "mi": "nearest", "mf": "nearest", "fwd_flow": "bilinear", "bwd_flow": "bilinear"}
    aug = T.Compose([T.RandomRotate(keys, 1.0, shape, [(0, 360)] * 3, "zeros", modes, grid="device",
                                    align_corners=False)] + [T.RandomFlip(keys, 0.5, a) for a in (2, 3, 4)])
    module = TR.PropagationModule(epochs=10, lambda_u=1.0, gamma_p=0.1, lr=1e-3, betas=(0.9, 0.999), weight_decay=1e-5,
                                  model_kwargs=dict(num_layers=4, num_classes=1, input_channels=2, features_start=8,
                                                    trilinear=False, padding=1, kernel_size=[3, 3, 3]),
                                  warp_kwargs=dict(interpolation_type="LINEAR", boundary_type="MIRROR",
                                                   mesh_length_type="numDofs", lenghts=(1, 1, 1))).to(dev)
    opt, _ = module.configure_optimizers()
    np.random.seed(0)

    def make_batch():
        return TR.collate_fn([data.item(k, aug) for k in range(len(data))])

    def train_step(batch):
        opt.zero_grad(set_to_none=True)
        loss, _ = module.forward_and_loss(batch)
        loss["t"].backward()
        opt.step()
        return loss

    for _ in range(max(1, args.warmup)):
        train_step(make_batch())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss = train_step(make_batch())
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    batch = make_batch()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        train_step(batch)
    torch.cuda.synchronize()
    el_train = time.perf_counter() - t1
    ts = int(batch["fwd_flow"].shape[1])
    _line(args, "mask-propagation training steps/s (batch 4 of 80^3, GPU augmentation + both chains + Adam)",
          round(args.steps / el, 3), "steps/s", round(1e3 * el / args.steps, 2),
          {"workload": f"batch 4 x {shape[0]}^3, chains of {ts} flows (ED->ES 8..11, padded), U-Net 4 levels / "
                       f"8 channels, LINEAR warp, rotation + 3 flips per sample on the GPU",
           "unet_calls_per_step": ts, "unet_batch": 2 * int(batch["fwd_flow"].shape[0])},
          {"train_only_steps_per_s": round(args.steps / el_train, 3),
           "loss": round(float(loss["t"].detach()), 4), "samples_per_s": round(4 * args.steps / el, 2)})


OTHER = {"uneven": bench_uneven, "serve": bench_serve, "unet": bench_unet, "augment": bench_augment,
         "train": bench_train}


if __name__ == "__main__":
    main()
