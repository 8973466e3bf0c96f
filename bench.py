#!/usr/bin/env python3
"""Throughput of the MPM substep on one MI355X, in particle-substeps/s.

Default workload: the 1,048,576-particle dam break (weakly compressible
neo-Hookean column, dx = 1/256 sparse hashed grid, dt = 2e-5, separating
walls), seeded synthetic input resident in HBM before timing starts.

Timed region: W warm-up substeps, then exactly K substeps captured once into a
HIP graph and replayed, bracketed by stream synchronizations and HIP events on
the launch stream. value = N * K / t.

Roofline: K more substeps launched directly, each pass timed by HIP events that
its kernels' dispatch packets write (hipExtLaunchKernelGGL, on the launch
stream); the pass with the largest time is the dominant kernel and its achieved
HBM rate is its algorithmic bytes per launch divided by its mean launch time.

roofline.traffic: HBM bytes per launch of the dominant kernel from the rocprofv3
FETCH_SIZE / WRITE_SIZE summary committed under profiles/ for the default
workload (null for other workloads or when the summary is absent).

cpu_baseline: the CPU oracle (a C restatement of the reference substep, OpenMP)
timed on the host cores on the first S substeps of the same scene.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-substeps/s (1M dam break; 256-scene batch) + p50/p99 scene latency"
HBM_PEAK_GBS = 8000.0


def pass_bytes(n, n_blk, n_dp=0):
    """Algorithmic HBM bytes per launch of each pass (FP32 fields, each moved once):
    grid_sort: block key write + read, 8 B/particle;
    p2g: read x, v, m, C (64 B/particle) + write 64 nodes x 16 B per block;
    grid_update: read + write 64 nodes x 16 B per block;
    g2p (fused with the particle update): read x, m, F, V0, tag, lambda, mu (68 B)
      and write x, v, C, F, force_dt (108 B) per particle + read 64 nodes x 16 B
      per block; Drucker-Prager particles add 16 B of parameters and 12 + 12 B of state."""
    return {
        "grid_sort": 8.0 * n,
        "p2g": 64.0 * n + 1024.0 * n_blk,
        "grid_update": 2048.0 * n_blk,
        "g2p": 176.0 * n + 40.0 * n_dp + 1024.0 * n_blk,
    }


def measured_traffic(workload, pass_name):
    """HBM bytes per launch of `pass_name`'s kernel from the committed PMC summary."""
    if workload != "dam_break":
        return None
    path = os.path.join(ROOT, "profiles", "dam_break_1M_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for ln in open(path):
        f = ln.split()
        if len(f) == 4 and f[0] == "traffic" and f[1] == pass_name:
            return float(f[3])
    return None


def make_scene(workload):
    from slosh_amd import scenes
    if workload == "dam_break":
        return scenes.dam_break()
    if workload == "dam_break_small":
        return scenes.dam_break(counts=(32, 128, 32), grid_capacity=4096)
    raise ValueError(workload)


def cpu_baseline(scene, steps, threads):
    from oracle import oracle as O  # the checker, timed as the CPU baseline only
    orc = O.Oracle(scene, svd_mode=1, threads=threads)
    t0 = time.perf_counter()
    orc.step(steps)
    dt = time.perf_counter() - t0
    return {
        "value": scene.n * steps / dt,
        "unit": "particle-substeps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {steps} substeps of the same {scene.n}-particle scene (C oracle, OpenMP, {dt:.1f} s)",
    }


def run_plain(args):
    from slosh_amd import pipeline as P
    scene = make_scene(args.workload)
    n = scene.n
    pl = P.MpmPipeline()
    st = P.Stream()
    data = P.MpmData.from_scene(pl, scene)
    # warm-up (also sorts particle storage by block)
    pl.launch_step(data, st)
    pl.launch_steps_graph(data, max(args.warmup - 1, 0), st)
    st.synchronize()
    data.check_errors(st)
    g_plain = P.Graph(pl, data, args.steps)
    e0, e1 = P.Event(), P.Event()
    st.synchronize()
    e0.record(st)
    g_plain.launch(st)
    e1.record(st)
    st.synchronize()
    ms = e0.elapsed_ms(e1)
    data.check_errors(st)
    n_blk = data.num_blocks(st)
    value = n * args.steps / (ms * 1e-3)

    # Per-pass times: K more substeps with packet-written events around every pass.
    ts = P.Timestamps()
    ts.reset()
    for _ in range(args.steps):
        pl.launch_step(data, st, timestamps=ts)
    per_pass = ts.read()
    data.check_errors(st)
    nbytes = pass_bytes(n, n_blk)
    passes = {k: {"mean_us": 1e3 * v[0] / max(v[1], 1), "launches": v[1]} for k, v in per_pass.items() if v[1]}
    for k, v in passes.items():
        if k in nbytes:
            v["alg_bytes"] = nbytes[k]
            v["GBps"] = nbytes[k] / (v["mean_us"] * 1e-6) / 1e9
    dom = max((k for k in passes if k in nbytes), key=lambda k: passes[k]["mean_us"])
    achieved = passes[dom]["GBps"]
    timed_ms = sum(v[0] for v in per_pass.values()) / args.steps
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload} ({n} particles)",
            "particles": n,
            "cell_width": float(scene.cell_width),
            "dt": float(scene.params["dt"]),
            "active_blocks": int(n_blk),
            "launch": f"hipGraph of {args.steps} substeps",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(args.workload, dom),
        },
        "passes": passes,
        "sum_of_pass_ms_per_step": timed_ms,
        "substep_alg_bytes": sum(nbytes.values()),
        "substep_alg_GBps": sum(nbytes.values()) / (ms / args.steps * 1e-3) / 1e9,
    }
    if args.cpu_steps > 0:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline(scene, args.cpu_steps, threads)
    return line


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="dam_break", choices=["dam_break", "dam_break_small"])
    ap.add_argument("--cpu-steps", type=int, default=120, help="oracle substeps for cpu_baseline (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    args.steps += args.steps & 1  # the captured graph holds an even number of substeps
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only (--gpus must be 1); sharding across "
                                   "GPUs is not implemented", "n_gpus": args.gpus}))
        sys.exit(2)
    line = run_plain(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
