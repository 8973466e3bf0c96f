#!/usr/bin/env python3
"""Throughput of the MPM substep on one MI355X, in particle-substeps/s.

Default workload: the 1,048,576-particle dam break (weakly compressible
neo-Hookean column, dx = 1/256 sparse hashed grid, dt = 2e-5, separating
walls), seeded synthetic input resident in HBM before timing starts.

Timed region: W warm-up substeps, then exactly K substeps captured once into a
HIP graph and replayed (an odd K: K - 1 in the graph and one direct launch),
bracketed by stream synchronizations and HIP events on the launch stream.
value = N * K / t.

Roofline: min(K, 20) more substeps launched directly behind one more replay of
the timed graph (so that they run back to back), each pass timed by HIP events
that its kernels' dispatch packets write (hipExtLaunchKernelGGL, on the launch
stream); the pass with the largest time is the dominant kernel and its achieved
HBM rate is its algorithmic bytes per launch divided by its mean launch time.

roofline.traffic: HBM bytes per launch of the dominant kernel from the rocprofv3
FETCH_SIZE / WRITE_SIZE summary committed under profiles/ for the default
workload (null for other workloads or when the summary is absent).

cpu_baseline: the CPU oracle (a C restatement of the reference substep, OpenMP)
timed on the host cores on the first S substeps of the same scene.

The same line carries two more workloads of the metric: `batched` (256
independent scenes, N log-uniform in [2e3, 2e5], neo-Hookean / fixed-corotated /
Drucker-Prager, 200 substeps in one launch sequence, against stepping each scene
alone) and `served` (32 scene requests, N log-uniform in [5e3, 2e6], 100-2000
substeps each, Poisson arrivals at each rate of --serve-rates under a 200 GB HBM
budget: p50 / p99 completion latency and aggregate particle-substeps/s, against
serving the same requests one at a time).
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


def measured_traffic(workload, pass_name, tuning):
    """HBM bytes per launch of the kernel instance `pass_name` runs under `tuning`
    (steady-state substeps), from the committed PMC summary of the workload: the
    dam break runs the neo-Hookean-only instance, the batch the all-materials
    instance with the refined SVD path."""
    files = {"dam_break": "dam_break_1M_hbm_traffic.md", "batched": "batched_256_hbm_traffic.md"}
    if workload not in files:
        return None
    path = os.path.join(ROOT, "profiles", files[workload])
    if not os.path.exists(path):
        return None
    g2p = "k_g2p_update<2u,%d,false,true>" if workload == "dam_break" else "k_g2p_update<15u,%d,true,true>"
    kernel = {"g2p": g2p % tuning["g2p_threads"], "p2g": "k_p2g<%d,true>" % tuning["p2g_chunk"],
              "grid_update": "k_grid_update<false>"}.get(pass_name)
    for ln in open(path):
        f = ln.split()
        if len(f) == 4 and f[0] == "traffic" and f[1] == pass_name and f[2] == kernel:
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
    # 40-substep trials (~4 ms each): shorter ones let timing noise pick among
    # launch shapes within a few per cent of each other.
    tuning = data.autotune(pl, st, trial_steps=40) if args.autotune else data.tuning
    pl.launch_steps_graph(data, max(args.warmup - 1, 0), st)
    st.synchronize()
    data.check_errors(st)
    # A captured graph holds an even number of substeps (a pair returns the
    # alternating buffers to where they started); an odd K adds one direct launch.
    k_graph = args.steps - (args.steps & 1)
    g_plain = P.Graph(pl, data, k_graph) if k_graph else None
    e0, e1 = P.Event(), P.Event()
    st.synchronize()
    e0.record(st)
    if g_plain is not None:
        g_plain.launch(st)
    if args.steps & 1:
        pl.launch_step(data, st)
    e1.record(st)
    st.synchronize()
    ms = e0.elapsed_ms(e1)
    data.check_errors(st)
    n_blk = data.num_blocks(st)
    value = n * args.steps / (ms * 1e-3)

    # Per-pass times: up to 20 more substeps with packet-written events around every
    # pass, queued behind one more replay of the timed graph so that they run back to
    # back as in the timed region (a host-paced loop leaves the GPU idle between
    # launches, and its kernels then run ~20 % slower at lower clocks).
    ts = P.Timestamps()
    ts.reset()
    rsteps = max(2, min(args.steps, 20))
    if args.steps & 1:
        pl.launch_step(data, st)  # back to the parity the graph was captured at
    if g_plain is not None:
        g_plain.launch(st)
    for _ in range(rsteps):
        pl.launch_step(data, st, timestamps=ts)
    per_pass = ts.read()
    data.check_errors(st)
    nbytes = pass_bytes(n, n_blk)
    passes = {k: {"mean_us": 1e3 * v[0] / max(v[1], 1), "launches": v[1]} for k, v in per_pass.items() if v[1]}
    for k, v in passes.items():
        if k in nbytes:
            v["alg_bytes"] = nbytes[k]
            v["GBps"] = nbytes[k] / (v["mean_us"] * 1e-6) / 1e9
    # The dominant kernel: the grid-sort pass is a chain of short kernels (each far
    # below the single-kernel passes), so the candidates are P2G, the grid update
    # and G2P + particle update.
    dom = max((k for k in passes if k in nbytes and k != "grid_sort"), key=lambda k: passes[k]["mean_us"])
    achieved = passes[dom]["GBps"]
    timed_ms = sum(v[0] for v in per_pass.values()) / rsteps
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
            "launch": " + ".join(([f"hipGraph of {k_graph} substeps"] if k_graph else []) +
                                 (["1 direct launch"] if args.steps & 1 else [])),
            "tuning": tuning,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(args.workload, dom, tuning),
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


def run_batched(args):
    """256 independent scenes in one batch vs the plain per-scene loop (each scene
    its own MpmData stepped by its own captured graphs, one after another)."""
    from slosh_amd import pipeline as P
    from slosh_amd import scenes
    scs = [scenes.batched_scene(i) for i in range(args.batch_scenes)]
    batch = scenes.Batch(scs)
    pl = P.MpmPipeline()
    st = P.Stream()
    data = P.MpmData.from_batch(pl, batch)
    pl.launch_step(data, st)
    tuning = data.autotune(pl, st, trial_steps=10) if args.autotune else data.tuning
    pl.launch_steps_graph(data, max(args.warmup - 1, 0), st)
    st.synchronize()
    data.check_errors(st)
    g = P.Graph(pl, data, args.batch_steps)
    e0, e1 = P.Event(), P.Event()
    st.synchronize()
    e0.record(st)
    g.launch(st)
    e1.record(st)
    st.synchronize()
    ms = e0.elapsed_ms(e1)
    data.check_errors(st)
    n_blk = data.num_blocks(st)
    batched = batch.n * args.batch_steps / (ms * 1e-3)
    g = None
    # Roofline of the batch's dominant kernel, as for the single scene: more
    # substeps launched directly with packet-written events around every pass.
    ts = P.Timestamps()
    ts.reset()
    rsteps = max(2, min(args.steps, 20))
    for _ in range(rsteps):
        pl.launch_step(data, st, timestamps=ts)
    per_pass = ts.read()
    data.check_errors(st)
    data.close()
    tags = batch.particles["models"][:, 0]
    n_dp = int(np.count_nonzero((tags == 2) | (tags == 3)))
    nbytes = pass_bytes(batch.n, n_blk, n_dp)
    passes = {k: {"mean_us": 1e3 * v[0] / max(v[1], 1), "launches": v[1]} for k, v in per_pass.items() if v[1]}
    for k, v in passes.items():
        if k in nbytes:
            v["alg_bytes"] = nbytes[k]
            v["GBps"] = nbytes[k] / (v["mean_us"] * 1e-6) / 1e9
    dom = max((k for k in passes if k in nbytes and k != "grid_sort"), key=lambda k: passes[k]["mean_us"])
    out = {
        "workload": f"{len(scs)} scenes, {batch.n} particles (N log-uniform [2e3, 2e5]; NH / FC / DP)",
        "value": batched,
        "unit": "particle-substeps/s",
        "steps": args.batch_steps,
        "ms_per_step": ms / args.batch_steps,
        "active_blocks": int(n_blk),
        "drucker_prager_particles": n_dp,
        "tuning": tuning,
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": passes[dom]["GBps"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": passes[dom]["GBps"] / HBM_PEAK_GBS,
            "traffic": measured_traffic("batched", dom, tuning),
        },
        "passes": passes,
    }
    if not args.batch_loop:
        return out
    # Plain per-scene loop over the same scenes and substeps.
    loop_ms = 0.0
    for sc in scs:
        cap = scenes.pow2_at_least(scenes.estimate_blocks(sc.particles["pos4"], sc.cell_width) * 2 + 256)
        d = P.MpmData(pl, sc.params, sc.particles, sc.bodies, sc.materials, sc.cell_width, cap)
        pl.launch_step(d, st)
        pl.launch_steps_graph(d, max(args.warmup - 1, 0), st)
        gs = P.Graph(pl, d, args.batch_steps)
        st.synchronize()
        e0.record(st)
        gs.launch(st)
        e1.record(st)
        st.synchronize()
        loop_ms += e0.elapsed_ms(e1)
        d.check_errors(st)
        gs = None
        d.close()
    loop = batch.n * args.batch_steps / (loop_ms * 1e-3)
    out.update({"per_scene_loop": loop, "per_scene_loop_ms_per_step": loop_ms / args.batch_steps,
                "speedup_vs_per_scene_loop": batched / loop})
    return out


def run_served(args):
    """Poisson arrivals of scene requests served by continuous batching, against
    the same requests served one at a time."""
    from slosh_amd import pipeline as P
    from slosh_amd import scenes
    from slosh_amd import serving as SV
    pl = P.MpmPipeline()
    st = P.Stream()
    out = []
    gen = [scenes.served_request(k) for k in range(args.serve_requests)]
    for rate in args.serve_rates:
        arr = scenes.poisson_arrivals(rate, args.serve_requests)
        reqs = [SV.Request(k, sc, steps, arrival=float(a)) for k, ((sc, steps), a) in enumerate(zip(gen, arr))]
        srv = SV.SceneServer(pl, reqs[0].scene, hbm_budget=args.hbm_budget, stream=st)
        done = srv.serve(reqs, realtime=True)
        summ = SV.latency_summary(done)
        summ.update({"rate_per_s": rate, "peak_particles_in_flight": srv.peak_in_flight,
                     "device_bytes": srv.device_bytes, "chunks": srv.chunks,
                     "host_seconds": {k: round(v, 4) for k, v in srv.seconds.items()}})
        srv.data.close()
        out.append(summ)
    # The plain loop at the highest rate (the most overlap to gain).
    rate = max(args.serve_rates)
    arr = scenes.poisson_arrivals(rate, args.serve_requests)
    reqs = [SV.Request(k, sc, steps, arrival=float(a)) for k, ((sc, steps), a) in enumerate(zip(gen, arr))]
    SV.serve_one_at_a_time(pl, reqs, stream=st, realtime=True)
    plain = SV.latency_summary(reqs)
    plain["rate_per_s"] = rate
    best = [o for o in out if o["rate_per_s"] == rate][0]
    return {"requests": args.serve_requests, "hbm_budget_bytes": args.hbm_budget, "rates": out,
            "one_at_a_time": plain, "speedup_vs_one_at_a_time": best["throughput"] / plain["throughput"]}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="dam_break", choices=["dam_break", "dam_break_small"])
    ap.add_argument("--cpu-steps", type=int, default=120, help="oracle substeps for cpu_baseline (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--batch-scenes", type=int, default=256, help="scenes of the batched leg (0 = skip)")
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--no-batch-loop", dest="batch_loop", action="store_false",
                    help="skip the per-scene loop the batched leg is compared with")
    ap.add_argument("--serve-requests", type=int, default=32, help="requests per served run (0 = skip)")
    ap.add_argument("--serve-rates", type=float, nargs="*", default=[10.0, 20.0, 50.0, 100.0, 200.0])
    ap.add_argument("--hbm-budget", type=float, default=200e9)
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the default launch shape instead of autotuning it before timing")
    args = ap.parse_args()
    if args.steps < 1:
        ap.error("--steps must be at least 1")
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only (--gpus must be 1); sharding across "
                                   "GPUs is not implemented", "n_gpus": args.gpus}))
        sys.exit(2)
    line = run_plain(args)
    if args.batch_scenes > 0:
        args.batch_steps += args.batch_steps & 1
        line["batched"] = run_batched(args)
    if args.serve_requests > 0 and args.serve_rates:
        line["served"] = run_served(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
