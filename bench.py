"""Benchmark of the MI355X PPHPC step.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

One "step" is one iteration of the model (move, sort, actions, statistics)
over the whole population. The default workload is the 1600x1600 dense
v2-style run: 409,600 sheep + 204,800 wolves, 4000 iterations from the seeded
initial state (seed 0), capacity 4x the initial population. The timed region
is exactly K iterations from iteration 0 with the state resident in HBM,
bracketed by device synchronisation; value = agent-steps/s, i.e. the sum over
the K iterations of the agents alive at the start of each, divided by the
wall time.

Also reported (one JSON line on stdout, progress on stderr):
  roofline      dominant kernel of a profiled run of the same workload:
                algorithmic bytes per launch (exact, from the run's event
                counters) / average launch duration (HIP events on the
                simulation's stream), against 8 TB/s
  roofline_12800  the same for the 12800x12800 run (HBM-resident working
                set), every iteration sorted by the radix sort (the
                bandwidth-bound path)
  roofline_12800_auto  the same with the automatic sort (row buckets: fewer
                bytes, latency-bound), and its time per iteration
  cpu_baseline  the OpenMP restatement of the step (row bands, as the
                reference's CPU simulator) on this host's cores, and the
                single-threaded C oracle: the OpenMP one over the whole
                workload, the oracle over its first iterations
  batched_1600  N (8 and 30) replicas of the 1600x1600 workload (seeds 1..N)
                as one batch against the same N run one at a time as plain
                simulations: agent-steps/s and replicas/hour both ways
  batched       64 replicas of grid sides drawn from {100 .. 3200}, v1 or v2s
                parameters 50/50, seeds 1..64, 4000 iterations each: one
                batch (one set of launches per iteration, sort shape from the
                tuner) against the same 64 run one at a time as plain
                simulations; replicas/hour both ways, and every statistics row
                compared between the two
  serving       a Poisson stream of requests (grid side log-uniform over
                {100 .. 6400}, v1 or v2s, 1000-4000 iterations, random seeds)
                served by continuous batching with admission by device
                footprint: p50/p99 completion latency and agent-steps/s
"""
import argparse
import collections
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("agent-steps/s @1600² v2 plain vs batched; replicas/hr; "
          "p99 run latency")
HBM_PEAK_GBS = 8000.0
STEP_KERNELS = ("bucket_scan", "bucket_sort", "sort_up", "sort_scan",
                "sort_down", "act", "act_scan", "emit_move")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def v2s_params(L, iters=4000):
    cells = L * L
    return dict(init_sheep=cells * 1600 // 10000, sheep_gain_from_food=4,
                sheep_reproduce_threshold=2, sheep_reproduce_prob=10,
                init_wolves=cells * 800 // 10000, wolves_gain_from_food=20,
                wolves_reproduce_threshold=2, wolves_reproduce_prob=5,
                grass_restart=5, grid_x=L, grid_y=L, grid_xy=cells,
                iters=iters)


def v1_params(L, iters=4000):
    cells = L * L
    return dict(init_sheep=cells * 400 // 10000, sheep_gain_from_food=4,
                sheep_reproduce_threshold=2, sheep_reproduce_prob=4,
                init_wolves=cells * 200 // 10000, wolves_gain_from_food=20,
                wolves_reproduce_threshold=2, wolves_reproduce_prob=5,
                grass_restart=10, grid_x=L, grid_y=L, grid_xy=cells,
                iters=iters)


WORKLOADS = {
    "v2s1600": lambda: v2s_params(1600),
    "v2s12800": lambda: v2s_params(12800),
    "v1_100": lambda: v1_params(100),
}


def kernel_bytes(name, c, passes, radix_all=False):
    """Algorithmic bytes of all launches of one kernel over a run of
    iterations, from the exact event sums c of those iterations (N agents at
    the start of each iteration, L alive after the move, N' after actions, O
    occupied cells, E grass cells eaten, B births), with the per-unit figures
    of the step's byte model (a 9-byte agent record: cell, energy, type):
      act          the actions: read each agent and write its acted half, a
                   cell-index and sheep-list read per agent (22 B per agent),
                   read and write each occupied cell's grass (8 B), write each
                   child (9 B): 22N + 8O + 9B
      emit_move    compaction and spawn of this iteration (19 B per agent, 9
                   per child) and the move of the next (read and write cell,
                   energy and flags of each agent of the new state, 18 B):
                   19N + 9B + 18N'
      bucket_sort  one read/write pass of the 8-byte (cell, value) records:
                   read the N moved records, write the L survivors in cell
                   order: 8N + 8L
      sort_up      radix pass 0 reads the N moved keys, later passes the L
                   survivors' keys: 4N + 4L(P-1)
      sort_down    pass 0 reads the N moved pairs and writes the L survivors,
                   later passes read and write L pairs: 8N + 8L + 16L(P-1)
    The radix kernels get a count only when every iteration of the run
    sorted by radix (radix_all); otherwise they ran only in the first
    iteration after a change of the replicas.
    """
    N, L, Np, O, B = (c["agent_steps"], c["live"], c["out"], c["occupied"],
                      c["births"])
    if name == "bucket_sort":
        return 8 * N + 8 * L
    if name == "sort_up" and radix_all:
        return 4 * N + 4 * L * (passes - 1)
    if name == "sort_down" and radix_all:
        return 8 * N + 8 * L + 16 * L * (passes - 1)
    if name == "act":
        return 22 * N + 8 * O + 9 * B
    if name == "emit_move":
        return 19 * N + 9 * B + 18 * Np
    return None


def compulsory_bytes(name, c):
    """The bytes this implementation cannot avoid moving (its own packed
    layout: 8-byte records, 1-byte grass as for grass restarts up to 191),
    for comparison with kernel_bytes:
      act        read L records, write L values, read each occupied cell's
                 grass, write each eaten one: 12L + O + E
      emit_move  read L records, write each output's moved record: 8L + 8N'
      bucket_sort  8N + 8L"""
    N, L, Np, O, E = (c["agent_steps"], c["live"], c["out"], c["occupied"],
                      c["eaten"])
    return {"act": 12 * L + O + E, "emit_move": 8 * L + 8 * Np,
            "bucket_sort": 8 * N + 8 * L}.get(name)


def counted_traffic(kernel, which):
    """Bytes per dispatch of a kernel from the committed rocprofv3 counter
    runs (profiles/hbm_traffic.json: FETCH_SIZE and WRITE_SIZE, scaled by a
    1 GiB copy; summaries in profiles/hbm_traffic_v2s12800.md and
    profiles/counters_v2s12800.md): which = "radix" or "auto" at
    12800x12800, "v2s1600" at 1600x1600. None when not measured."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        d = json.load(open(path))
        table = (d["per_dispatch_v2s1600"] if which == "v2s1600"
                 else d["per_dispatch"][which])
        return table["k_" + kernel]
    except (OSError, KeyError, ValueError):
        return None


def measure_roofline(params, capacity, steps, seed=0, sort_mode=None):
    """Profiled run: per-kernel event timing on the simulation's stream, and
    the unprofiled time per iteration of the same iterations (graphs)."""
    from pphpc_amd import Params, Simulation
    sim = Simulation(Params.from_values(**params), seed=seed,
                     capacity=capacity, device=0)
    if sort_mode is not None:
        sim.set_sort(sort_mode)
    passes = sim.info()["sort_passes"]
    sim.init()
    sim.run(1, collect=False)
    sim.sync()
    t0 = time.perf_counter()
    sim.run(steps, collect=False)
    sim.sync()
    ms_tick = 1e3 * (time.perf_counter() - t0) / steps
    sim.init()
    sim.run(min(steps, 4), collect=False)      # warm the code paths
    sim.init()
    sim.run(1, collect=False)                  # the radix-sorted iteration
    c0 = sim.counters()
    sim.set_profiling(True)
    sim.profile_reset()
    sim.run(steps, collect=False)
    prof = sim.profile()
    c1 = sim.counters()
    sim.close()
    c = {k: c1[k] - c0[k] for k in c1}
    step_kernels = {k: v for k, v in prof.items()
                    if k in STEP_KERNELS and v[0] > 0}
    radix_all = step_kernels.get("sort_down", (0, 0))[0] == passes * steps
    dom = max((k for k in step_kernels
               if kernel_bytes(k, c, passes, radix_all)),
              key=lambda k: step_kernels[k][1])
    launches, ns = step_kernels[dom]
    total_bytes = kernel_bytes(dom, c, passes, radix_all)
    per_kernel = {}
    for k, (nl, t) in step_kernels.items():
        b = kernel_bytes(k, c, passes, radix_all)
        bc = compulsory_bytes(k, c)
        per_kernel[k] = {
            "launches": nl, "avg_us": round(t / max(nl, 1) / 1e3, 3),
            "GBps": round(b / t, 1) if b and t else None,
            "frac": round(b / t / HBM_PEAK_GBS, 4) if b and t else None,
            "compulsory_GBps": round(bc / t, 1) if bc and t else None,
            "share": round(t / sum(v[1] for v in step_kernels.values()), 3)}
    achieved = total_bytes / ns  # bytes/ns == GB/s
    return {
        "kernel": dom,
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "bytes_per_launch": int(total_bytes / launches),
        "avg_launch_us": round(ns / launches / 1e3, 3),
        "profiled_steps": steps,
        "ms_per_iteration": round(ms_tick, 4),
        "sort": "radix" if radix_all else "row buckets",
        "per_kernel": per_kernel,
        "events_per_iteration": {k: round(v / steps, 1) for k, v in c.items()},
    }


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _run_cpu(exe, params, capacity, sample_steps, extra=()):
    keys = ("INIT_SHEEP", "SHEEP_GAIN_FROM_FOOD", "SHEEP_REPRODUCE_THRESHOLD",
            "SHEEP_REPRODUCE_PROB", "INIT_WOLVES", "WOLVES_GAIN_FROM_FOOD",
            "WOLVES_REPRODUCE_THRESHOLD", "WOLVES_REPRODUCE_PROB",
            "GRASS_RESTART", "GRID_X", "GRID_Y", "ITERS")
    fields = ("init_sheep", "sheep_gain_from_food",
              "sheep_reproduce_threshold", "sheep_reproduce_prob",
              "init_wolves", "wolves_gain_from_food",
              "wolves_reproduce_threshold", "wolves_reproduce_prob",
              "grass_restart", "grid_x", "grid_y", "iters")
    with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as f:
        for k, fl in zip(keys, fields):
            f.write(f"{k} = {params[fl]}\n")
        path = f.name
    try:
        out = subprocess.run([exe, "-p", path, "-n", str(sample_steps), "-q",
                              "-m", str(capacity), *extra],
                             capture_output=True, text=True, timeout=600,
                             check=True).stdout
    finally:
        os.unlink(path)
    return dict(tok.split("=") for tok in out.split())


def probe_reference_cpu_path():
    """Whether the reference's OpenCL CPU simulator (pp_cpu) could be built
    and run on this host: an OpenCL CPU device (clinfo) and the cf4ocl2,
    cl_ops and glib-2.0 headers its build needs
    (opencl/CMakeLists.txt:22-27). Returns what was found, verbatim where
    short."""
    out = {}
    try:
        r = subprocess.run(["clinfo", "-l"], capture_output=True, text=True,
                           timeout=60)
        out["clinfo_l"] = (r.stdout + r.stderr).strip()[-600:]
        r = subprocess.run(["clinfo"], capture_output=True, text=True,
                           timeout=60)
        cpu_dev = [ln.strip() for ln in r.stdout.splitlines()
                   if "Device Type" in ln and "CPU" in ln]
        out["opencl_cpu_devices"] = len(cpu_dev)
    except (OSError, subprocess.SubprocessError) as e:
        out["clinfo_l"] = f"clinfo not runnable: {e}"
        out["opencl_cpu_devices"] = 0
    incs = ["/usr/include", "/usr/local/include", "/opt/rocm/include"]
    for name, rel in (("cf4ocl2", "cf4ocl2.h"), ("cl_ops", "cl_ops.h"),
                      ("glib-2.0", "glib-2.0/glib.h")):
        hits = [d for d in incs if os.path.exists(os.path.join(d, rel))
                or os.path.exists(os.path.join(d, name, rel))]
        out[name + "_headers"] = hits or None
    out["pkg-config"] = bool(subprocess.run(
        ["sh", "-c", "command -v pkg-config"], capture_output=True).stdout)
    out["buildable"] = bool(out["opencl_cpu_devices"] and out["cf4ocl2_headers"]
                            and out["cl_ops_headers"] and out["glib-2.0_headers"])
    return out


def cpu_baseline(params, capacity, sample_steps, single_steps, threads):
    """CPU restatements of the step on this host, on the first iterations of
    the same workload: the OpenMP row-band restatement (oracle/pp_omp.c) on
    `threads` threads, and the single-threaded C oracle. Both compute the
    same results as the GPU (same semantics and random streams)."""
    omp = os.path.join(ROOT, "oracle", "build", "pp_omp")
    one = os.path.join(ROOT, "oracle", "build", "pp_oracle")
    if not (os.path.exists(omp) and os.path.exists(one)):
        return None
    sample_steps = min(sample_steps, params["iters"])
    span = (f"all {sample_steps} iterations" if sample_steps == params["iters"]
            else f"first {sample_steps} iterations")
    kv = _run_cpu(omp, params, capacity, sample_steps, ("-t", str(threads)))
    kv1 = _run_cpu(one, params, capacity, single_steps)
    return {"value": float(kv["agent_steps_per_s"]), "unit": "agent-steps/s",
            "cores": int(kv["threads"]), "kind": "port",
            "sample": f"{span} of the same workload "
                      f"(seed 0), OpenMP row-band restatement of the step "
                      f"(oracle/pp_omp.c) on {kv['threads']} threads: "
                      f"{float(kv['seconds']):.1f} s",
            "cpu_model": _cpu_model(),
            "host_cpus": os.cpu_count(),
            "single_thread": {
                "value": float(kv1["agent_steps_per_s"]),
                "unit": "agent-steps/s", "cores": 1, "kind": "port",
                "sample": f"first {single_steps} iterations, single-threaded "
                          f"C oracle: {float(kv1['seconds']):.1f} s"},
            "reference_cpu_path": probe_reference_cpu_path()}


def batched_same_leg(counts, iters, side=1600):
    """N replicas of the same workload (side x side, dense v2-style, seeds
    1..N) as one batch against the same N run one at a time as plain
    simulations: agent-steps/s and replicas/hour both ways, rows compared."""
    import numpy as np
    from pphpc_amd import Batch, Params, Simulation, stats_rows_to_array
    d = v2s_params(side, iters)
    n0 = d["init_sheep"] + d["init_wolves"]
    res = 4 * n0
    out = []
    for n in counts:
        t0 = time.perf_counter()
        b = Batch(n * res, n * side * side, n, d["grass_restart"], device=0)
        ids = [b.add(Params.from_values(**d), seed, res)
               for seed in range(1, n + 1)]
        while b.resources()["active"]:
            b.run(512)
        b.sync()
        batched_s = time.perf_counter() - t0
        steps_b = b.counters()["agent_steps"]
        rows_b = [b.rows_array(r) for r in ids]
        b.close()
        t1 = time.perf_counter()
        steps_p = 0
        equal = True
        for seed, rb in zip(range(1, n + 1), rows_b):
            sim = Simulation(Params.from_values(**d), seed=seed, capacity=res,
                             device=0)
            rows = sim.simulate()
            steps_p += sim.agent_steps
            sim.close()
            equal &= bool(np.array_equal(stats_rows_to_array(rows), rb))
        plain_s = time.perf_counter() - t1
        log(f"{n} x {side}^2 v2s: batched {batched_s:.2f} s, plain "
            f"{plain_s:.2f} s, rows equal {equal}")
        out.append({
            "replicas": n, "side": side, "iters": iters, "seeds": f"1..{n}",
            "batched_s": round(batched_s, 3), "plain_s": round(plain_s, 3),
            "agent_steps_per_s_batched": steps_b / batched_s,
            "agent_steps_per_s_plain": steps_p / plain_s,
            "replicas_per_hour_batched": round(n * 3600 / batched_s, 1),
            "replicas_per_hour_plain": round(n * 3600 / plain_s, 1),
            "speedup": round(plain_s / batched_s, 3),
            "rows_equal_plain": equal})
    return out


def batched_leg(n, iters, draw_key=4, tune_ticks=16):
    """The batched-replica workload, batched and then plain, same GPU."""
    import numpy as np
    from pphpc_amd import Batch, Params, Simulation, stats_rows_to_array
    from pphpc_amd.workloads import batch_replicas, reserve_for
    reps = batch_replicas(n, draw_key, iters)
    cells = sum(L * L for _, L, _, _ in reps)
    agents = sum(reserve_for(d) for _, _, d, _ in reps)
    log(f"batched: {n} replicas, {cells} cells, {agents} reserved agents")
    t0 = time.perf_counter()
    b = Batch(agents, cells, n, 10, device=0)
    ids = [b.add(Params.from_values(**d), seed, reserve_for(d))
           for _, _, d, seed in reps]
    tuned = b.tune(tune_ticks) if tune_ticks else None
    while b.resources()["active"]:
        b.run(512)
    b.sync()
    t1 = time.perf_counter()
    batched_s = t1 - t0
    agent_steps = b.counters()["agent_steps"]
    info = b.info()
    rows_b = [b.rows_array(r) for r in ids]
    states = [b.replica(r)["state"] for r in ids]
    b.close()
    log(f"batched: {batched_s:.2f} s ({agent_steps / batched_s:.3e} "
        f"agent-steps/s), shape {tuned}")
    plain_s = 0.0
    equal = True
    for (kind, L, d, seed), rb in zip(reps, rows_b):
        t2 = time.perf_counter()
        sim = Simulation(Params.from_values(**d), seed=seed,
                         capacity=reserve_for(d), device=0)
        rows = sim.simulate()
        sim.close()
        plain_s += time.perf_counter() - t2
        equal &= bool(np.array_equal(stats_rows_to_array(rows), rb))
    log(f"plain: {plain_s:.2f} s, rows equal: {equal}")
    return {
        "replicas": n, "iters": iters, "cells": cells,
        "sides": sorted(collections.Counter(L for _, L, _, _ in reps).items()),
        "kinds": dict(collections.Counter(k for k, _, _, _ in reps)),
        "seeds": "1..%d" % n,
        "batched_s": round(batched_s, 3), "plain_s": round(plain_s, 3),
        "replicas_per_hour_batched": round(n * 3600 / batched_s, 1),
        "replicas_per_hour_plain": round(n * 3600 / plain_s, 1),
        "speedup": round(plain_s / batched_s, 3),
        "agent_steps": int(agent_steps),
        "agent_steps_per_s_batched": agent_steps / batched_s,
        "agent_steps_per_s_plain": agent_steps / plain_s,
        "all_done": all(st == 2 for st in states),
        "rows_equal_plain": equal,
        "sort_shape": tuned, "sort_passes": info["sort_passes"],
        "note": "wall time from batch creation to the last replica's "
                "completion vs the sum of create+simulate+destroy of the "
                "plain runs; the tuner's trial iterations count as progress",
    }


def serving_leg(seconds, rate, hbm_gb, chunk, draw_key=5):
    from pphpc_amd.serve import Server
    from pphpc_amd.workloads import serve_requests
    stream = serve_requests(rate, seconds, draw_key)
    log(f"serving: {len(stream)} requests over {seconds} s at {rate}/s")
    with Server(int(hbm_gb * (1 << 30)), chunk=chunk) as srv:
        reqs, summary = srv.run(stream, log=log)
    summary.update({"arrival_rate_per_s": rate, "window_s": seconds,
                    "window_note": "arrivals over a 60 s window (the workload "
                                   "names 10 minutes; the bench keeps its run "
                                   "to minutes); the loop then serves the "
                                   "queue to its end",
                    "chunk_ticks": chunk, "hbm_budget_gib": hbm_gb,
                    "sides": "log-uniform over 100..6400",
                    "iters": "uniform 1000..4000"})
    log(f"serving: p50 {summary['p50_latency_s']} s, p99 "
        f"{summary['p99_latency_s']} s, {summary['agent_steps_per_s']:.3e} "
        "agent-steps/s")
    return summary


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--workload", default="v2s1600", choices=list(WORKLOADS))
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--roofline-steps", type=int, default=400)
    ap.add_argument("--big-roofline-steps", type=int, default=16)
    ap.add_argument("--cpu-steps", type=int, default=4000)
    ap.add_argument("--cpu-single-steps", type=int, default=150)
    ap.add_argument("--cpu-threads", type=int,
                    default=min(16, int(os.environ.get("OMP_NUM_THREADS", 0))
                                or os.cpu_count() or 1))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-big", action="store_true")
    ap.add_argument("--batch-replicas", type=int, default=64)
    ap.add_argument("--batch-iters", type=int, default=4000)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-same", default="8,30",
                    help="replica counts of the 1600x1600 batched-vs-plain "
                         "leg (empty: skip)")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--serve-hbm-gb", type=float, default=48.0)
    ap.add_argument("--serve-chunk", type=int, default=32)
    ap.add_argument("--no-serve", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build runs the step on a "
              "single MI355X (no multi-GPU path); use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    from pphpc_amd import Params, Simulation

    params = WORKLOADS[args.workload]()
    n0 = params["init_sheep"] + params["init_wolves"]
    capacity = 4 * n0
    log(f"workload {args.workload}: {params['grid_x']}x{params['grid_y']}, "
        f"{n0} agents, capacity {capacity}")

    sim = Simulation(Params.from_values(**params), seed=args.seed,
                     capacity=capacity, device=0)
    info = sim.info()
    sim.init()
    if args.warmup:
        sim.run(args.warmup, collect=False)
    sim.init()
    sim.sync()
    t0 = time.perf_counter()
    sim.run(args.steps, collect=False)
    sim.sync()
    t1 = time.perf_counter()
    agent_steps = sim.agent_steps
    n_end = sim.num_agents
    sim.close()
    wall = t1 - t0
    value = agent_steps / wall
    log(f"{args.steps} steps in {wall:.3f} s: {value:.4e} agent-steps/s "
        f"({n_end} agents at the end)")

    log("profiled run for the roofline")
    roof = measure_roofline(params, capacity, min(args.roofline_steps,
                                                  args.steps), args.seed)
    if args.workload == "v2s1600":
        roof["traffic"] = counted_traffic(roof["kernel"], "v2s1600")
        roof["traffic_source"] = ("profiles/counters_v2s12800.md (FETCH_SIZE "
                                  "+ WRITE_SIZE; at this size mostly "
                                  "Infinity Cache, not HBM)")
    log(f"dominant kernel {roof['kernel']}: {roof['achieved']} GB/s")

    roof_big = roof_big_b = None
    if not args.no_big:
        big = v2s_params(12800)
        nb = big["init_sheep"] + big["init_wolves"]
        log("profiled 12800x12800 runs for the HBM roofline: radix sort")
        roof_big = measure_roofline(big, 4 * nb, args.big_roofline_steps,
                                    args.seed, sort_mode=0)
        roof_big["traffic"] = counted_traffic(roof_big["kernel"], "radix")
        roof_big["traffic_source"] = "profiles/hbm_traffic_v2s12800.md"
        log(f"12800^2 radix: dominant kernel {roof_big['kernel']}: "
            f"{roof_big['achieved']} GB/s, {roof_big['ms_per_iteration']} ms")
        roof_big_b = measure_roofline(big, 4 * nb, args.big_roofline_steps,
                                      args.seed)
        roof_big_b["traffic"] = counted_traffic(roof_big_b["kernel"], "auto")
        roof_big_b["traffic_source"] = "profiles/counters_v2s12800.md"
        log(f"12800^2 auto (buckets): dominant kernel {roof_big_b['kernel']}: "
            f"{roof_big_b['achieved']} GB/s, "
            f"{roof_big_b['ms_per_iteration']} ms")

    batched = None
    if not args.no_batch:
        batched = batched_leg(args.batch_replicas, args.batch_iters)

    batched_1600 = None
    if not args.no_batch and args.batch_same:
        batched_1600 = batched_same_leg(
            [int(x) for x in args.batch_same.split(",")], args.batch_iters)

    serving = None
    if not args.no_serve:
        serving = serving_leg(args.serve_seconds, args.serve_rate,
                              args.serve_hbm_gb, args.serve_chunk)

    cpu = None
    if not args.no_cpu:
        log(f"cpu baseline: {args.cpu_threads} threads, {args.cpu_steps} "
            f"iterations; single thread, {args.cpu_single_steps}")
        cpu = cpu_baseline(params, capacity, args.cpu_steps,
                           args.cpu_single_steps, args.cpu_threads)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "agent-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic: seeded initial state (Philox, seed "
                f"{args.seed}), no dataset",
        "config": {
            "workload": f"{args.workload}: {params['grid_x']}x"
                        f"{params['grid_y']} torus, {params['init_sheep']} "
                        f"sheep + {params['init_wolves']} wolves, "
                        f"{args.steps} iterations, plain (one replica)",
            "grid": [params["grid_x"], params["grid_y"]],
            "capacity": capacity,
            "seed": args.seed,
            "sort_passes": info["sort_passes"],
            "sort_tile": info["sort_tile"],
            "agent_steps": agent_steps,
            "parallelism": "single GPU",
        },
        "roofline": roof,
        "roofline_12800": roof_big,
        "roofline_12800_auto": roof_big_b,
        "batched": batched,
        "batched_1600": batched_1600,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
