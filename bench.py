#!/usr/bin/env python3
"""Throughput of the plain colour sweep on one MI355X.

Workload (default): the large linear melt (512^3 periodic box, 16384 chains of
N=512, 8,388,608 monomers, volume fraction 0.5, 2 colours, bit-packed
compact-Morton lattice), synthetic seeded initial state; --workload small_melt
(64^3, 512 x 32) and --workload network (256^3 diamond network, 1,064,960
monomers) time the other single-system cases the same way.  A step is one
Monte-Carlo step (every monomer attempts one move).  Inputs are resident in
HBM before the timed region.  Prints one JSON line.

The cpu_baseline leg times the oracle's random-sequential updater (the
LeMonADE-style CPU updater, one core) on a bounded number of MCS of the same
system, with the serial colour-sweep replay (one core) and the OpenMP colour
sweep (all threads) beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    # 512^3 periodic melt, 16384 linear chains of 512 (8,388,608 monomers, 2 colours)
    "large_melt": dict(kind="melt", box=512, n_chains=16384, chain_len=512, seed=0x5EED000000000002),
    # 64^3 periodic melt, 512 linear chains of 32 (16,384 monomers, 2 colours)
    "small_melt": dict(kind="melt", box=64, n_chains=512, chain_len=32, seed=0x5EED000000000001),
    # 256^3 periodic end-linked diamond network: 8 x 8 x 32 unit cells, 16,384
    # 4-functional crosslinkers joined by 32,768 strands of 32 (1,064,960 monomers)
    "network": dict(kind="network", box=256, cells=(8, 8, 32), strand_len=32, seed=0x5EED000000000003),
}


def build_workload(w: dict) -> dict:
    from pscbfm_amd import systems
    if w["kind"] == "melt":
        return systems.linear_melt(w["box"], w["n_chains"], w["chain_len"], w["seed"])
    return systems.diamond_network(w["box"], w["cells"], w["strand_len"], w["seed"])


def workload_config(name: str, w: dict) -> dict:
    if w["kind"] == "melt":
        return {"workload": name, "box": w["box"], "chains": w["n_chains"], "chain_len": w["chain_len"]}
    return {"workload": name, "box": w["box"], "unit_cells": list(w["cells"]), "strand_len": w["strand_len"]}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


TRAFFIC_SUMMARIES = {
    "large_melt": "k_check_large_melt_hbm_traffic.md",
    "small_melt": "k_check_small_melt_trace_and_traffic.md",
    "network": "k_check_network_trace_and_traffic.md",
}


def measured_traffic(workload: str):
    """HBM bytes per k_check launch from the committed rocprofv3 PMC summary
    of the workload (FETCH_SIZE and WRITE_SIZE, scaled by the copy kernel)."""
    import re
    name = TRAFFIC_SUMMARIES.get(workload)
    f = ROOT / "profiles" / name if name else None
    if f is None or not f.exists():
        return None
    m = re.search(r"```json\n(.*?)\n```", f.read_text(), flags=re.S)
    return json.loads(m.group(1))["traffic_bytes"] if m else None


def cpu_baseline(system: dict, seed: int, target_s: float) -> dict:
    """Three CPU legs on the same system, filling about target_s together:
    half for the random-sequential updater (the reference's CPU path is
    LeMonADE's serial updater; this is its restatement, one core; the reported
    value), a quarter each for the serial colour-sweep replay (one core) and
    the OpenMP colour sweep (all threads OpenMP is given)."""
    from oracle import oracle as orc
    osys = orc.System(system, seed)
    n = len(system["xyz"])

    def timed(run, share):
        t = time.perf_counter()
        xyz = run(system["xyz"], 0, 1)
        one = time.perf_counter() - t
        m = max(1, int(share / max(one, 1e-6)) - 1)
        t = time.perf_counter()
        run(xyz, 1, m)
        dt = time.perf_counter() - t
        return m, dt, m * n / dt

    m_rs, dt_rs, rs = timed(lambda x, a, k: orc.random_sequential(osys, x, a, k)[0], target_s / 2)
    m_sw, dt_sw, sw = timed(lambda x, a, k: orc.sweep(osys, x, a, k, 1)[0], target_s / 4)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    m_omp, dt_omp, omp = timed(lambda x, a, k: orc.sweep_parallel(osys, x, a, k, threads)[0], target_s / 4)
    log(f"CPU: random-sequential {rs:.3g}, sweep replay {sw:.3g} (1 core), OpenMP sweep {omp:.3g} "
        f"({threads} threads) attempts/s")
    return {"value": rs, "unit": "attempts/s", "cores": 1, "kind": "port",
            "sample": f"oracle random-sequential updater, {m_rs} MCS x {n} monomers of the same system "
                      f"({dt_rs:.1f} s on 1 host core)",
            "legs": {
                "sweep_replay": {"value": sw, "cores": 1,
                                 "sample": f"{m_sw} MCS ({dt_sw:.1f} s), serial colour sweep, bit-exact oracle"},
                "sweep_openmp": {"value": omp, "cores": threads,
                                 "sample": f"{m_omp} MCS ({dt_omp:.1f} s), OpenMP colour sweep"},
            }}


def batched_leg(args) -> dict:
    """Mixed-size batch: every system swept in the same launches (hipGraph per
    chunk) versus a plain per-system launch sequence, with bit-exact
    comparison of every final state."""
    import numpy as np
    from pscbfm_amd import Batch, updater_from_system, workloads
    t = time.perf_counter()
    if args.batch_uniform:
        count, box, chain_len, phi = args.batch_uniform.split(",")
        specs = workloads.uniform_batch(int(count), int(box), int(chain_len), float(phi), seed=0xBA7C4)
        composition = f"{int(count)} x (box {int(box)}, chain length {int(chain_len)}, phi {float(phi)})"
    else:
        specs = workloads.mixed_batch(args.batch_systems, seed=0xBA7C4)
        composition = "mixed"
    built = [workloads.build(sp) for sp in specs]
    n_tot = sum(len(x["xyz"]) for x in built)
    log(f"batched: {len(built)} systems, {n_tot} monomers, built in {time.perf_counter() - t:.1f} s")

    # plain: one initialized updater per system, run one after the other
    # (--plain-mcs 0 skips it, for profiled runs of the batched kernels)
    plain = [updater_from_system(x, seed=sp["seed"], gpu=0) for x, sp in zip(built, specs)] if args.plain_mcs else []
    for u in plain:
        u.getStats()
    t0 = time.perf_counter()
    for u in plain:
        u.runResident(args.plain_mcs)
    for u in plain:
        u.getStats()                  # synchronizes each updater's stream
    t_plain = time.perf_counter() - t0
    plain_rate = n_tot * args.plain_mcs / t_plain if args.plain_mcs else None
    if plain:
        log(f"plain per-system: {args.plain_mcs} MCS each in {t_plain:.2f} s -> {plain_rate:.4g} attempts/s")

    b = Batch(0)
    b.set_graphs(not args.no_graphs)
    if args.working_set_mib >= 0:
        b.set_working_set_bytes(args.working_set_mib << 20)
    ids = [b.add(updater_from_system(x, seed=sp["seed"], validate_only=True), len(x["xyz"]))
           for x, sp in zip(built, specs)]
    b.synchronize()
    t0 = time.perf_counter()
    b.run(args.plain_mcs, chunk=args.chunk)
    b.synchronize()
    t_a = time.perf_counter() - t0
    exact = 0
    for u, m, x in zip(plain, ids, built):
        u.copyBack()
        exact += int(np.array_equal(u.getMonomerPositions(len(x["xyz"])), b.positions(m)))
        u.cleanup()
    rest = args.batch_mcs - args.plain_mcs
    t0 = time.perf_counter()
    if rest > 0:
        b.run(rest, chunk=args.chunk)
    b.synchronize()
    t_batched = t_a + time.perf_counter() - t0
    rate = n_tot * args.batch_mcs / t_batched
    check = f"{exact}/{len(built)} states bit-exact vs plain" if plain else "no plain comparison"
    log(f"batched: {args.batch_mcs} MCS in {t_batched:.2f} s -> {rate:.4g} attempts/s, {check}")
    return {"systems": len(built), "composition": composition, "monomers": n_tot, "mcs": args.batch_mcs,
            "chunk": args.chunk, "graphs": not args.no_graphs,
            "working_set_mib": 128 if args.working_set_mib < 0 else args.working_set_mib,
            "working_set_groups": b.working_set_groups(),
            "value": rate, "unit": "attempts/s", "plain_value": plain_rate, "plain_mcs": args.plain_mcs,
            "speedup_vs_plain": rate / plain_rate if plain_rate else None, "bit_exact_systems": exact}


def served_leg(args, capacity: float) -> dict:
    """Poisson job arrivals at rates swept around the batched capacity;
    continuous batching admitted by free HBM; p50/p99 completion latency.

    A job counts as finished when its client holds the result: the loop copies
    each job's positions out (unsorted and unwrapped on the device at
    retirement) and releases the job as soon as the server reports it done, on
    the same thread that drives step().  Jobs are submitted with the
    validated system handed over (bfm_server_submit_take); their updaters are
    built and validated before the clock starts, as clients would.
    Throughput is given twice: total work over the makespan (first arrival to
    last result, so the drain at the end is included) and over the steady
    window between the arrivals of the 20th-percentile and the last job (the
    server's attempt counter at the step boundaries closest to both ends)."""
    import numpy as np
    from pscbfm_amd import Server, updater_from_system, workloads
    points = []
    runs = [(factor, args.jobs) for factor in args.load_factors]
    if args.steady_jobs > 0:
        runs.append((1.0, args.steady_jobs))
    for k, (factor, n_jobs) in enumerate(runs):
        jobs = workloads.job_stream(n_jobs, 1.0, seed=0x5E4E + k)   # arrival times rescaled below
        built = [workloads.build(j) for j in jobs]
        work = np.array([len(x["xyz"]) * j["mcs"] for x, j in zip(built, jobs)], dtype=float)
        mean_work = float(work.mean())
        rate = factor * capacity / mean_work                          # jobs per second
        arrivals = np.array([j["arrival"] for j in jobs]) / rate       # unit-rate stream rescaled
        upd = [updater_from_system(x, seed=j["seed"], validate_only=True) for x, j in zip(built, jobs)]
        srv = Server(0, hbm_budget=0, chunk=args.chunk)
        w0, w1 = float(np.percentile(arrivals, 20)), float(arrivals[-1])
        win = {}
        t0 = time.perf_counter()
        ids = [None] * len(jobs)
        finish = [None] * len(jobs)
        results = [None] * len(jobs)
        running = set()
        nxt = 0
        while True:
            now = time.perf_counter() - t0
            while nxt < len(jobs) and arrivals[nxt] <= now:
                ids[nxt] = srv.submit(upd[nxt], jobs[nxt]["mcs"], len(built[nxt]["xyz"]), take=True)
                running.add(nxt)
                nxt += 1
            pending = srv.step()
            for j in list(running):
                if srv.state(ids[j])[0] >= 2:      # done, failed or cancelled
                    results[j] = srv.result(ids[j])
                    srv.release(ids[j])            # the client holds its copy
                    finish[j] = time.perf_counter() - t0
                    running.discard(j)
            now = time.perf_counter() - t0
            for edge, t_edge in (("start", w0), ("end", w1)):
                if edge not in win and now >= t_edge:
                    win[edge] = (now, srv.info()["attempts"])
            if pending == 0:
                if nxt == len(jobs):
                    break
                time.sleep(max(0.0, arrivals[nxt] - (time.perf_counter() - t0)))
        graphs = srv.graph_stats()
        lat = np.array([finish[j] - arrivals[j] for j in range(len(jobs))])
        makespan = max(finish) - arrivals[0]
        (ta, aa), (tb, ab) = win.get("start", (0.0, 0)), win.get("end", (makespan, int(work.sum())))
        steady = (ab - aa) / (tb - ta) if tb > ta else None
        exact = 0
        for j in range(len(jobs)):
            u = updater_from_system(built[j], seed=jobs[j]["seed"], gpu=0)
            u.runSimulationOnGPU(jobs[j]["mcs"])
            exact += int(np.array_equal(u.getMonomerPositions(len(built[j]["xyz"])), results[j]))
            u.cleanup()
        pt = {"load_factor": factor, "rate_jobs_per_s": round(rate, 2), "jobs": len(jobs),
              "mean_job_attempts": mean_work,
              "throughput": float(work.sum() / makespan),
              "throughput_steady": None if steady is None else float(steady),
              "steady_window_s": [round(ta, 4), round(tb, 4)],
              "p50_latency_s": float(np.percentile(lat, 50)),
              "p99_latency_s": float(np.percentile(lat, 99)), "graphs": graphs,
              "bit_exact_jobs": f"{exact}/{len(jobs)}"}
        log(f"served at {factor}x capacity ({rate:.1f} jobs/s, {len(jobs)} jobs): {pt['throughput']:.4g} attempts/s "
            f"over the makespan, {steady or 0:.4g} in the steady window, "
            f"p50 {pt['p50_latency_s']*1e3:.0f} ms, p99 {pt['p99_latency_s']*1e3:.0f} ms, exact {pt['bit_exact_jobs']}, "
            f"chunk graphs {graphs['launches']} ({graphs['updates']} updated, {graphs['instantiations']} instantiated)")
        points.append(pt)
    return {"unit": "attempts/s", "chunk": args.chunk, "capacity": capacity, "points": points}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000, help="timed Monte-Carlo steps")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--autotune", type=int, default=20,
                    help="least MCS per autotuner probe after the relaxation, more where a probe would last "
                         "under 4 ms (0: the default launch configuration)")
    ap.add_argument("--equilibrate", type=int, default=1000,
                    help="untimed MCS before the warm-up: the generated start (chains on the even sublattice, "
                         "every bond |b|^2 = 4) relaxes toward the melt's bond-vector distribution")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="large_melt")
    ap.add_argument("--layout", choices=["morton", "id", "shuffled"], default="morton",
                    help="monomer order within a colour: Morton order of position (the default), original id, "
                         "or original id after a random relabelling of the monomers (input numbering "
                         "without locality)")
    ap.add_argument("--no-pcie", action="store_true",
                    help="skip the PCIe-inclusive call after the timed region (for counter runs)")
    ap.add_argument("--cpu-seconds", type=float, default=24.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--time-every", type=int, default=0,
                    help="time the kernels of every n-th substep with HIP events (0: 8 for systems of a million "
                         "monomers or more, 128 for smaller ones, whose kernels last a few microseconds and whose "
                         "launch pipeline a timed launch would stall)")
    ap.add_argument("--batch-systems", type=int, default=512)
    ap.add_argument("--batch-uniform", metavar="COUNT,BOX,CHAIN_LEN,PHI", default=None,
                    help="batch COUNT melts of one box edge, chain length and volume fraction "
                         "instead of the mixed batch (for kernel traces of single compositions)")
    ap.add_argument("--batch-mcs", type=int, default=1000)
    ap.add_argument("--plain-mcs", type=int, default=200)
    ap.add_argument("--chunk", type=int, default=100)
    ap.add_argument("--jobs", type=int, default=96, help="jobs per served-load point")
    ap.add_argument("--load-factors", type=float, nargs="*", default=[0.5, 1.0, 2.0])
    ap.add_argument("--steady-jobs", type=int, default=1024,
                    help="jobs of one more served point at capacity (0 = skip)")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-graphs", action="store_true",
                    help="batched leg without chunk graphs (kernel by kernel; the same kernels, for profilers)")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--sort-every", type=int, default=0,
                    help="setStepsBetweenSortings of the plain leg (0 = the updater's default, 5000 as the "
                         "reference's); trajectories do not depend on it, the generator being keyed on the id")
    ap.add_argument("--served-capacity", type=float, default=0.0,
                    help="attempts/s that the served load factors scale (0 = the batched leg's rate)")
    ap.add_argument("--working-set-mib", type=int, default=-1,
                    help="working-set budget of the batched leg in MiB (-1 = the engine's "
                         "default of 128, 0 = never split the systems into groups)")
    ap.add_argument("--copy-mib", type=int, default=1024,
                    help="size of the plain copy kernel that calibrates HBM counters (0 = skip)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this engine runs every system on a single device; "
              "use --gpus 1", file=sys.stderr)
        return 2

    from pscbfm_amd import device_count, updater_from_system
    if device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1

    copy_gbs = None
    if args.copy_mib > 0:
        from pscbfm_amd import _capi
        nbytes = args.copy_mib << 20
        ms = _capi.calibration_copy(0, nbytes, 10)
        copy_gbs = 2.0 * nbytes / (ms * 1e-3) / 1e9
        log(f"plain copy of {args.copy_mib} MiB: {ms:.3f} ms -> {copy_gbs:.0f} GB/s (read + write)")

    w = WORKLOADS[args.workload]
    t = time.perf_counter()
    system = build_workload(w)
    n = len(system["xyz"])
    if args.layout == "shuffled":
        import numpy as np
        perm = np.random.default_rng(0x5407).permutation(n)        # new id of monomer i
        xyz = np.empty_like(system["xyz"])
        xyz[perm] = system["xyz"]
        system = dict(system, xyz=xyz, bonds=perm[system["bonds"]])
    log(f"generated {args.workload}: {n} monomers in {time.perf_counter() - t:.1f} s")
    t = time.perf_counter()
    u = updater_from_system(system, seed=w["seed"], gpu=0, spatial_sorting=args.layout == "morton",
                            steps_between_sortings=args.sort_every or None)
    colors, n_colors = u.getColors(n)
    log(f"initialized in {time.perf_counter() - t:.1f} s, {n_colors} colours")

    # HIP events around the kernels of every 8th substep of the timed region:
    # live kernel durations without slowing the launch pipeline of the rest
    # (short runs time every substep, so that some launches are timed)
    every = args.time_every or (8 if n >= 1_000_000 else 128)
    every = every if args.steps * n_colors >= 4 * every else 1
    t = time.perf_counter()
    u.runResident(args.equilibrate)
    u.getStats()
    t_relax = time.perf_counter() - t
    log(f"relaxed for {args.equilibrate} MCS in {t_relax:.2f} s")
    # launch configuration chosen by the autotuner (untimed; the system goes on
    # relaxing through its probes).  Probes of at least --autotune MCS, and
    # long enough to last about 4 ms, so that run-to-run noise of short
    # kernels does not decide
    probe_mcs = args.autotune
    if args.autotune > 0 and args.equilibrate > 0:
        probe_mcs = max(args.autotune, -int(-0.004 // (t_relax / args.equilibrate)))
    if args.autotune > 0:
        tuned = u.autotune(probe_mcs, n_colors)
    else:
        cfg = [u.launchConfig(c) for c in range(n_colors)]
        tuned = {"perform_group": cfg[0]["perform_group"], "xcd_ranges": cfg[0]["xcd_ranges"],
                 "check_block": [c["check_block"] for c in cfg], "check_items": [c["check_items"] for c in cfg],
                 "perform_groups": [c["perform_group"] for c in cfg],
                 "morton_order": cfg[0]["morton_order"]}
    log(f"launch configuration: {tuned}")
    u.setKernelTiming(every)
    u.runResident(args.warmup)
    s0 = u.getStats()  # synchronizes the updater's stream
    t0 = time.perf_counter()
    u.runResident(args.steps)
    s1 = u.getStats()
    t1 = time.perf_counter()
    dt = t1 - t0
    attempts = args.steps * n
    value = attempts / dt
    log(f"timed {args.steps} MCS in {dt:.3f} s -> {value:.4g} attempts/s")

    launches = s1["timed_launches"] - s0["timed_launches"]
    check_ms = (s1["check_ms"] - s0["check_ms"]) / max(launches, 1)
    perform_ms = (s1["perform_ms"] - s0["perform_ms"]) / max(launches, 1)
    accepted = s1["accepted"] - s0["accepted"]
    proposals = s1["proposals"] - s0["proposals"]
    p_acc = accepted / attempts
    kbar = 2.0 * len(system["bonds"]) / n
    per_launch = n / n_colors                       # attempts in one check launch
    p_prop = proposals / attempts
    clear = (w["box"] ** 3 // 8) / per_launch       # claim lattice cleared per launch, per attempt
    # canonical algorithmic bytes per attempt of the check phase: own int32
    # xyz (12), per bond a 4-byte partner id + 12-byte partner xyz, 4 front bits
    b_check = 12.5 + 16.0 * kbar
    achieved = per_launch * b_check / (check_ms * 1e-3) / 1e9
    # whole path adds 13 B per accepted move (12-byte xyz write + 8 bit flips)
    b_attempt = b_check + 13.0 * p_acc
    # what this layout needs instead: packed position + id word (8), per bond
    # index + packed position (8), 9 front bits, per survivor its 8-byte entry
    # and claim bit, and the clear of the other claim lattice
    b_layout = 8.0 + 8.0 * kbar + 9.0 / 8.0 + p_prop * (8.0 + 1.0 / 8.0) + clear
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(args.workload) if args.layout == "morton" else None,
                "kernel": "k_check", "kernel_avg_ms": round(check_ms, 4),
                "perform_avg_ms": round(perform_ms, 4), "bytes_per_attempt_check": round(b_check, 3),
                "layout_bytes_per_attempt_check": round(b_layout, 3),
                "path_frac": round(value * b_attempt / 1e9 / HBM_PEAK_GBS, 4),
                "copy_kernel_gbs": None if copy_gbs is None else round(copy_gbs, 1)}

    out = {
        "metric": "monomer move attempts/s (plain colour sweep, bit-exact)",
        "value": value,
        "unit": "attempts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": ("synthetic: seeded melt grown as self-avoiding walks on the even sublattice with (2,0,0)-class "
                 "steps (every initial bond |b|^2 = 4), relaxed for the equilibrate MCS before timing"
                 if w["kind"] == "melt" else
                 "synthetic: seeded end-linked diamond network ((2,0,0)-class bonds on the even sublattice), "
                 "relaxed for the equilibrate MCS before timing"),
        "config": {**workload_config(args.workload, w), "layout": args.layout, "monomers": n, "colours": n_colors,
                   "mcs": args.steps, "equilibrated_mcs": args.equilibrate,
                   "steps_between_sortings": u.getStepsBetweenSortings(),
                   "p_acc": round(p_acc, 4), "p_proposal": round(p_prop, 4),
                   "rng": "int_hash (the reference's default generator)",
                   "launch": {"check_workgroup": tuned["check_block"], "check_monomers_per_lane": tuned["check_items"],
                              "perform_group": tuned["perform_groups"],
                              "xcd_ranges": tuned["xcd_ranges"],
                              "order_within_colour": "morton" if tuned["morton_order"] else "id",
                              "autotune_probe_mcs": probe_mcs}},
        "roofline": roofline,
    }
    # the drop-in call as a LeMonADE host makes it: MCS plus the copy-back of
    # every position to host memory (PCIe-inclusive; reported beside value)
    if not args.no_pcie:
        pcie_mcs = 100
        u.runSimulationOnGPU(pcie_mcs)
        t = time.perf_counter()
        u.runSimulationOnGPU(pcie_mcs)
        dt_pcie = time.perf_counter() - t
        out["pcie_inclusive"] = {"call": f"runSimulationOnGPU({pcie_mcs}) with copy-back",
                                 "value": pcie_mcs * n / dt_pcie, "unit": "attempts/s", "ms_per_call": dt_pcie * 1e3}
        log(f"runSimulationOnGPU({pcie_mcs}) with copy-back: {dt_pcie * 1e3:.1f} ms -> "
            f"{pcie_mcs * n / dt_pcie:.4g} attempts/s")
    u.cleanup()
    out["roofline"]["timed_every_substeps"] = every
    if not args.no_batched:
        out["batched"] = batched_leg(args)
        if args.batch_uniform:
            log("served leg skipped: its jobs follow the mixed distribution, not --batch-uniform")
        elif not args.no_served:
            out["served"] = served_leg(args, args.served_capacity or out["batched"]["value"])
    if not args.no_cpu:
        log("timing the CPU baseline")
        out["cpu_baseline"] = cpu_baseline(system, w["seed"], args.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
