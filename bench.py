#!/usr/bin/env python3
"""Benchmark of the MI355X PRIME20 DMD engine.

Workload (default): 64 uneven PRIME20 systems (chain counts log-uniform in
[4,192], chain lengths uniform in {8,12,16,24,32} residues, densities
rho* in [0.005,0.02], T* in [0.15,0.50], random sequences and coils from a
seed).  A step advances every system by --per-step committed collisions.

  * batched: all systems in one persistent launch (one workgroup each, W
             wavefronts per system picked by the autotuner unless --waves)
  * plain:   the same systems advanced one at a time (one launch each, same W)
  * cpu_baseline: the CPU oracle on a bounded sample of the same systems,
    one system per host core

Prints one JSON line (see the keys below).
"""
import argparse
import json
import math
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "parallel-dmd-for-biomolecules_amd")
sys.path.insert(0, PKG)

METRIC = "aggregate collisions/s + p50/p99 trajectory latency, batched vs plain, 1 GPU"


def uneven64(seed=2024, nsys=64):
    rng = np.random.default_rng(seed)
    specs = []
    for s in range(nsys):
        nch = int(round(math.exp(rng.uniform(math.log(4), math.log(192)))))
        L = int(rng.choice([8, 12, 16, 24, 32]))
        dens = float(rng.uniform(0.005, 0.02))
        tstar = float(rng.uniform(0.15, 0.50))
        specs.append(dict(nchains=nch, L=L, density=dens, tstar=tstar, seed=seed * 1000 + s))
    return specs


def build_systems(specs):
    import pdmd
    return [pdmd.generate([sp["L"]] * sp["nchains"], sp["density"], sp["tstar"], sp["seed"]) for sp in specs]


def bytes_per_event(k_up, k_dn):
    # per committed event: 2 (K_up + K_dn) pair predictions x 53 B (partner
    # state 48 B + index 4 B + code 1 B), 6 calendar link updates x 4 B and
    # the 2 x 48 B state write-back of the colliding pair
    return 2.0 * (k_up + k_dn) * 53.0 + 6 * 4 + 2 * 48


# FP64 operations of one pair prediction (free flight, minimum image, dot
# products, discriminant, root; a sqrt and a division count as one each)
FLOPS_PER_PREDICTION = 37.0


def _cpu_worker(args):
    sysd, seed, nev, param_dir = args
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    o = Oracle(sysd, seed=seed, param_dir=param_dir)
    t0 = time.perf_counter()
    o.run(nev)
    return nev, time.perf_counter() - t0


def host_cores():
    """(os.cpu_count(), CPUs this process may run on, the cgroup CPU quota
    or None): the pool below uses the smallest."""
    nproc = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = nproc
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = max(1, int(int(q) // int(per)))
    except (OSError, ValueError):
        pass
    return nproc, aff, quota


def cpu_baseline(systems, seeds, per_sys, param_dir, workers):
    """Every system of the workload as an independent serial process of the
    CPU oracle, `workers` at a time; the aggregate is the collisions over the
    busy time per core (set-up excluded, perfect packing assumed)."""
    work = [(systems[k], seeds[k], per_sys, param_dir) for k in range(len(systems))]
    ctx = mp.get_context("spawn")
    t0 = time.perf_counter()
    with ctx.Pool(workers) as pool:
        res = list(pool.imap_unordered(_cpu_worker, work, chunksize=1))
    wall = time.perf_counter() - t0
    ncoll = sum(r[0] for r in res)
    busy = sum(r[1] for r in res)
    per_core = np.array([r[0] / r[1] for r in res])
    return dict(rate=ncoll / (busy / workers), wall=wall, ncoll=ncoll, per_core_median=float(np.median(per_core)),
                per_core_min=float(per_core.min()), per_core_max=float(per_core.max()))


PROFILES = os.path.join(ROOT, "profiles")


def profile_figure(name, key):
    """A figure (`key: value` line) of a counter summary in profiles/
    (None if absent)."""
    try:
        for line in open(os.path.join(PROFILES, name)):
            if line.startswith(key + ":"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def measured_bytes_per_entry():
    """HBM bytes per committed entry of k_run from the FETCH_SIZE/WRITE_SIZE
    counter passes."""
    return profile_figure("k_run_uneven64_hbm_counters.md", "bytes_per_entry")


def measured_fp64_per_entry():
    """FP64 operations per committed entry of k_run from the FP64 VALU
    instruction counters, weighted by the measured VALU lane utilisation
    (an FMA counts two)."""
    return profile_figure("k_run_uneven64_fp64_counters.md", "fp64_flops_per_entry")


def offered_load(rate, lo, hi):
    """Collisions per second that Poisson arrivals at `rate` ask for, each
    log-uniform in [lo, hi] (mean (hi - lo) / ln(hi / lo))."""
    return rate * (hi - lo) / math.log(hi / lo)


# the serving configuration as specified: Poisson arrivals at 20/s for 10
# minutes, 1e5 to 1e8 collisions per request
CONFIG5 = dict(rate=20.0, seconds=600.0, lo=1e5, hi=1e8)


def serve_run(systems, seeds, args, waves):
    """Poisson trajectory requests against the continuous-batching server.

    Requests arrive at rate --serve-rate for --serve-seconds; each picks one of
    the workload's systems, its own thermostat seed, and a log-uniform number
    of collisions in [--serve-min, --serve-max].  Latency is completion time
    minus arrival time on the server's clock.
    """
    import pdmd
    rng = np.random.default_rng(7)
    n_req = max(1, int(rng.poisson(args.serve_rate * args.serve_seconds)))
    arr = np.sort(rng.uniform(0.0, args.serve_seconds, n_req))
    pick = rng.integers(0, len(systems), n_req)
    ncoll = np.exp(rng.uniform(math.log(args.serve_min), math.log(args.serve_max), n_req)).astype(np.int64)
    eng = pdmd.Engine(waves=waves)
    srv = pdmd.Server(eng, mem_budget_bytes=int(args.serve_budget_gib * (1 << 30)), quantum=args.serve_quantum,
                      slice_ms=args.serve_slice_ms)
    t0 = time.perf_counter()
    rids = []
    k = 0
    steps = 0
    while True:
        now = time.perf_counter() - t0
        while k < n_req and arr[k] <= now:
            rids.append(srv.submit(systems[pick[k]], int(ncoll[k]), seed=seeds[pick[k]] + 1000003 * (k + 1)))
            k += 1
        srv.step()
        steps += 1
        q, r, committed = srv.pending()
        if k == n_req and q == 0 and r == 0:
            break
        if q == 0 and r == 0 and k < n_req:
            time.sleep(max(0.0, arr[k] - (time.perf_counter() - t0)))
    wall = time.perf_counter() - t0
    lat, done_coll, failed = [], 0, 0
    first_sub, last_done = None, 0.0
    for rid in rids:
        info, _ = srv.result(rid)
        if info.state != 2:
            failed += 1
            continue
        lat.append(info.t_done - info.t_submit)
        done_coll += info.stats.coll
        first_sub = info.t_submit if first_sub is None else min(first_sub, info.t_submit)
        last_done = max(last_done, info.t_done)
    srv.close()
    eng.close()
    lat = np.array(lat)
    span = last_done - (first_sub or 0.0)
    return {
        "requests": n_req, "failed": failed, "steps": steps, "wall_s": round(wall, 2),
        "collisions": int(done_coll), "collisions_per_s": round(done_coll / span, 1) if span > 0 else None,
        "p50_latency_s": round(float(np.percentile(lat, 50)), 4) if lat.size else None,
        "p99_latency_s": round(float(np.percentile(lat, 99)), 4) if lat.size else None,
        "arrival_rate_per_s": args.serve_rate, "arrival_seconds": args.serve_seconds,
        "collisions_per_request": [args.serve_min, args.serve_max], "quantum": args.serve_quantum,
        "slice_ms": args.serve_slice_ms,
        "offered_load_collisions_per_s": round(offered_load(args.serve_rate, args.serve_min, args.serve_max), 1),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--per-step", type=int, default=20000, help="collisions per system per step")
    ap.add_argument("--nsys", type=int, default=64)
    ap.add_argument("--plain-per-sys", type=int, default=0, help="collisions per system in the plain run (0 = per-step)")
    ap.add_argument("--cpu-per-sys", type=int, default=600000, help="collisions per system in the CPU baseline")
    ap.add_argument("--waves", type=int, default=0, help="wavefronts per system (0: autotune)")
    ap.add_argument("--tune-events", type=int, default=3000, help="entries per system per autotuner trial")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the 6144-bead plain system and the 11-system temperature ladder")
    ap.add_argument("--capacity-nsys", type=int, default=256,
                    help="systems of the same distribution in the capacity run (0: skip it)")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson request arrivals per second")
    ap.add_argument("--serve-seconds", type=float, default=20.0, help="arrival window")
    ap.add_argument("--serve-min", type=float, default=1e4, help="collisions per request, log-uniform lower end")
    ap.add_argument("--serve-max", type=float, default=3e5, help="collisions per request, log-uniform upper end")
    ap.add_argument("--serve-quantum", type=int, default=200000, help="most collisions per request per launch")
    ap.add_argument("--serve-slice-ms", type=float, default=25.0, help="device time per launch (0: event quanta)")
    ap.add_argument("--serve-budget-gib", type=float, default=64.0, help="device memory for resident requests")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (scale for the HBM byte counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20)
        a = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": n, "read_bytes": n, "write_bytes": n}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); systems shard "
              "across GPUs as independent batches, which this single-GPU build does not launch.", file=sys.stderr)
        sys.exit(2)
    import pdmd

    specs = uneven64(nsys=args.nsys)
    t0 = time.perf_counter()
    systems = build_systems(specs)
    seeds = [sp["seed"] + 17 for sp in specs]
    gen_s = time.perf_counter() - t0
    eng = pdmd.Engine()
    t0 = time.perf_counter()
    ids = [eng.add_system(s, seed=sd) for s, sd in zip(systems, seeds)]
    setup_s = time.perf_counter() - t0
    beads = [eng.nbeads[i] for i in ids]

    # batched: wavefronts per system (given, or autotuned after the warm-up),
    # warm-up, then K timed steps, each one launch over all systems
    if args.waves > 0:
        eng.set_waves(args.waves)
        waves = args.waves
    for _ in range(args.warmup):
        eng.advance(ids, [args.per_step] * len(ids))
    tuning = None
    if args.waves <= 0:
        # per system: wavefronts (1, 2, 4, the most that fit) and the list
        # rebuild's coarse cell factor (2..4), each system timed by its own
        # device clock; probe entries are ordinary committed entries
        wsys, cfsys, rates = eng.autotune_systems(ids, args.tune_events)
        waves = int(wsys.max())
        tuning = {
            "waves_per_system": {str(int(v)): int((wsys == v).sum()) for v in np.unique(wsys)},
            "cell_factor_per_system": {str(int(v)): int((cfsys == v).sum()) for v in np.unique(cfsys)},
            # a trial the tuner skipped (the most that fit equals a probed
            # count) reads as null
            "trial_rates_sum_per_s": {str(t): (round(float(rates[:, q].sum()), 1) if rates[:, q].any() else None)
                                      for q, t in enumerate(("1", "2", "4", "max"))},
            "probe_entries_per_trial": args.tune_events,
        }
    coll0 = sum(eng.stats(i).coll for i in ids)
    launch_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.advance(ids, [args.per_step] * len(ids))
        launch_ms.append(eng.last_launch_ms())
    dt = time.perf_counter() - t0
    st = [eng.stats(i) for i in ids]
    ncoll = sum(s.coll for s in st) - coll0
    value = ncoll / dt
    if any(s.status for s in st):
        raise SystemExit(f"device status error: {[s.status for s in st]}")
    k_up = float(np.average([s.avg_up for s in st], weights=beads))
    k_dn = float(np.average([s.avg_dn for s in st], weights=beads))
    bpe = bytes_per_event(k_up, k_dn)
    ev_per_launch = args.per_step * len(ids)
    avg_launch_s = float(np.mean(launch_ms)) / 1e3
    achieved = ev_per_launch * bpe / avg_launch_s / 1e9
    # FP64 rate: the counted FP64 operations per entry over this run's
    # launch time, beside the algorithmic 2 (K_up + K_dn) root solves
    fp64_alg_gflops = ev_per_launch * 2.0 * (k_up + k_dn) * FLOPS_PER_PREDICTION / avg_launch_s / 1e9
    fpe_meas = measured_fp64_per_entry()
    fp64_gflops = None if fpe_meas is None else ev_per_launch * fpe_meas / avg_launch_s / 1e9
    bpe_meas = measured_bytes_per_entry()

    # plain: the same systems one at a time
    plain = None
    if not args.no_plain:
        pps = args.plain_per_sys or args.per_step
        eng.advance(ids[:1], [1000])
        t0 = time.perf_counter()
        for i in ids:
            eng.advance([i], [pps])
        pdt = time.perf_counter() - t0
        plain = pps * len(ids) / pdt

    # the other GPU configurations at their shapes: one 6144-bead system
    # (96 x 16 residues, rho* = 0.01, T* = 0.20) alone, with its list
    # rebuilds timed by its own device clock, and the 11-copy temperature
    # ladder of it in one launch
    configs = None
    if not args.no_configs:
        configs = {}
        ladder = [0.18, 0.20, 0.22, 0.24, 0.26, 0.28, 0.30, 0.35, 0.40, 0.45, 0.50]
        for name, temps in (("plain_6144", [0.20]), ("ladder_11x6144", ladder)):
            geng = pdmd.Engine(waves=waves)
            try:
                gids = [geng.add_system(pdmd.generate([16] * 96, density=0.01, tstar=t, seed=6144 + q), seed=61 + q)
                        for q, t in enumerate(temps)]
                geng.advance(gids, [args.per_step] * len(gids))
                g0 = [geng.stats(i) for i in gids]
                t0 = time.perf_counter()
                geng.advance(gids, [2 * args.per_step] * len(gids))
                gdt = time.perf_counter() - t0
                g1 = [geng.stats(i) for i in gids]
                rec = {"systems": len(gids), "beads_each": 6144,
                       "collisions_per_s": round(2 * args.per_step * len(gids) / gdt, 1)}
                if name == "plain_6144":
                    st0, st1 = g0[0], g1[0]
                    nreb = st1.rebuilds - st0.rebuilds
                    if nreb > 0:
                        # lists (16 B per bead, 4 B per entry, 4 B per search
                        # cell) and events() (53 B per up-pair, 16 B per leaf)
                        n = st1.nbeads
                        cells = (st1.cells_axis // st1.cell_factor) ** 3
                        b = 16 * n + 4 * n * (st1.avg_up + st1.avg_dn) + 4 * cells + 53 * n * st1.avg_up + 16 * n
                        per = (st1.rebuild_ms - st0.rebuild_ms) / nreb
                        rec.update(list_rebuilds=int(nreb), list_rebuild_ms=round(per, 3),
                                   list_rebuild_algorithmic_gbps=round(b / (per / 1e3) / 1e9, 2),
                                   list_rebuild_share=round((st1.rebuild_ms - st0.rebuild_ms) /
                                                            (st1.device_ms - st0.device_ms), 3))
                configs[name] = rec
            finally:
                geng.close()

    # capacity: more systems of the same distribution resident at once (one
    # workgroup each, up to one per CU), same per-step collisions and W;
    # reported beside the headline, which stays the 64-system workload
    capacity = None
    if args.capacity_nsys > 0:
        cspecs = uneven64(nsys=args.capacity_nsys)
        csys = build_systems(cspecs)
        ceng = pdmd.Engine()
        try:
            cids = [ceng.add_system(s, seed=sp["seed"] + 17) for s, sp in zip(csys, cspecs)]
            ceng.set_waves(waves)
            ceng.advance(cids, [args.per_step] * len(cids))
            csteps = 2
            t0 = time.perf_counter()
            for _ in range(csteps):
                ceng.advance(cids, [args.per_step] * len(cids))
            cdt = time.perf_counter() - t0
            capacity = {
                "systems": len(cids), "beads_total": int(sum(ceng.nbeads[i] for i in cids)),
                "collisions_per_s": round(args.per_step * len(cids) * csteps / cdt, 1),
                "ms_per_step": round(cdt / csteps * 1e3, 3), "waves_per_system": waves,
                "note": "systems 0-63 are the headline workload; the rest are further draws of its distribution",
            }
        finally:
            ceng.close()

    serve = None
    if not args.no_serve:
        serve = serve_run(systems, seeds, args, waves)
        c5 = offered_load(CONFIG5["rate"], CONFIG5["lo"], CONFIG5["hi"])
        serve["device_capacity_collisions_per_s"] = round(value, 1)
        serve["load_factor"] = round(serve["offered_load_collisions_per_s"] / value, 3)
        serve["reduced_from_config"] = {
            "arrival_rate_per_s": CONFIG5["rate"], "window_s": [args.serve_seconds, CONFIG5["seconds"]],
            "collisions_per_request": [[args.serve_min, args.serve_max], [CONFIG5["lo"], CONFIG5["hi"]]],
            "offered_load_collisions_per_s": round(c5, 1), "load_factor": round(c5 / value, 2),
            "note": "the serving configuration as specified offers load_factor times the device's capacity, so "
                    "its queue grows for the whole window; this run keeps its arrival process and system "
                    "distribution and scales window and request sizes down to fit minutes",
        }

    cpu = None
    if not args.no_cpu:
        nproc, aff, quota = host_cores()
        workers = min(len(systems), aff, quota or aff)
        cb = cpu_baseline(systems, seeds, args.cpu_per_sys, pdmd.PARAM_DIR, workers)
        cpu = {"value": round(cb["rate"], 1), "unit": "collisions/s", "cores": workers, "kind": "port",
               "host_nproc": nproc, "affinity_cpus": aff, "cgroup_cpu_quota": quota,
               "per_core_median": round(cb["per_core_median"], 1), "per_core_min": round(cb["per_core_min"], 1),
               "per_core_max": round(cb["per_core_max"], 1),
               "gpu_batched_over_cpu": round(value / cb["rate"], 3),
               "sample": f"CPU oracle (gcc -O2, serial restatement), all {len(systems)} systems of the same workload "
                         f"as independent processes, {args.cpu_per_sys} collisions each, {workers} at a time "
                         f"({cb['wall']:.1f} s wall, set-up included); value = collisions / (busy seconds / cores)"}

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "collisions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator: random sequences, self-avoiding coils, Maxwell-Boltzmann velocities)",
        "config": {
            "workload": "uneven64: 64 PRIME20 systems, chains log-uniform [4,192] x {8,12,16,24,32} residues, "
                        "rho* U[0.005,0.02], T* U[0.15,0.50], Andersen ghosts on",
            "systems": len(ids),
            "beads_total": int(sum(beads)),
            "beads_min": int(min(beads)),
            "beads_max": int(max(beads)),
            "collisions_per_system_per_step": args.per_step,
            "collisions_per_system_timed": args.per_step * args.steps,
            "config_collisions_per_system": 10 ** 7,
            "mode": f"batched: one persistent launch, one workgroup of {waves} wavefronts per system "
                    "(each system limited to its own tuned count)",
            "waves_per_launch": waves,
            "autotune": tuning,
        },
        "plain_collisions_per_s": None if plain is None else round(plain, 1),
        "batched_over_plain": None if plain is None else round(value / plain, 2),
        "p50_latency_s": None if serve is None else serve["p50_latency_s"],
        "p99_latency_s": None if serve is None else serve["p99_latency_s"],
        "serving": serve,
        "capacity": capacity,
        "configs": configs,
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 3),
            "peak": 8000.0,
            "unit": "GB/s",
            "frac": achieved / 8000.0,
            "traffic": None if bpe_meas is None else round(bpe_meas * ev_per_launch),
            "traffic_per_entry": bpe_meas,
            "kernel": "k_run",
            "bytes_per_event": round(bpe, 1),
            "k_up": round(k_up, 2),
            "k_dn": round(k_dn, 2),
            "avg_launch_ms": round(avg_launch_s * 1e3, 3),
        },
        "fp64": {
            "kernel": "k_run",
            "achieved": None if fp64_gflops is None else round(fp64_gflops, 3),
            "peak": 78600.0,
            "unit": "GFLOP/s",
            "frac": None if fp64_gflops is None else fp64_gflops / 78600.0,
            "flops_per_entry_counted": fpe_meas,
            "source": "FP64 VALU instruction counters x 64 x VALU lane utilisation (FMA = 2) per committed "
                      "entry, over this run's launch time",
            "algorithmic_root_solves": {
                "achieved": round(fp64_alg_gflops, 3),
                "frac": fp64_alg_gflops / 78600.0,
                "flops_per_prediction": FLOPS_PER_PREDICTION,
                "predictions_per_entry": round(2.0 * (k_up + k_dn), 2),
            },
        },
        "list_and_predictor_kernels": {
            "source": "standalone dispatches on this workload (kernel_report.py under rocprofv3), "
                      "profiles/k_rebuild_uneven64_counters.md and profiles/k_predict_uneven64_counters.md",
            "unit": "GB/s of algorithmic bytes",
            "k_rebuild_lists": profile_figure("k_rebuild_uneven64_counters.md", "lists_algorithmic_gbps"),
            "k_rebuild_events": profile_figure("k_rebuild_uneven64_counters.md", "events_algorithmic_gbps"),
            "k_predict_largest": profile_figure("k_predict_uneven64_counters.md", "predict_largest_algorithmic_gbps"),
            "k_predict_largest_fp64_frac": profile_figure("k_predict_uneven64_counters.md",
                                                          "predict_largest_fp64_frac"),
        },
        "cpu_baseline": cpu,
        "setup_s": round(gen_s + setup_s, 2),
    }
    print(json.dumps(out), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
