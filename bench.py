"""Benchmark: cell-steps/s of the LCS per-step hot path on one MI355X.

A step is LCS.timestep! + LCS.evolve! (CFL and particle dt, two RK stages of
ghost fill + particles + RHS/update + HSMAC, linear-power forcing) over the
whole box, with all state resident in HBM.  Default workload: one 1024^3 FP32
box, linear-power P = 0.5, Re = 1333 (Re_lambda ~ 200 at u' = 1, eps = P),
6.4e7 uniform particles with St = 1, 5 warm-up + 20 timed steps.

The default run adds three sub-objects to its line: ``ensemble`` (the batched
96-box ensemble against the same boxes stepped one at a time), ``served`` (the
served request stream) and ``reference_case`` (750^3 FP64 with 325^3 particles,
the one-device case of the reference's published weak-scaling runs).

Prints one JSON line.  ``roofline`` describes the kernel with the largest total
time in the timed region, timed live with HIP events on the box stream;
``cpu_baseline`` times the CPU oracle (the C restatement of the reference
path) on a bounded sample of the same workload on this host.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "lcs.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

PEAK_HBM_GBS = 8000.0
METRIC = "cell-steps/s at 1024³ (%HBM peak); ensemble box-steps/s + p99 vs serial"


def workload(name, n_override=None, particles_override=None):
    from lcs_mi355x.workloads import single_box
    try:
        return single_box(name, n_override, particles_override)
    except ValueError as e:
        raise SystemExit(str(e))


def copy_calibration(nbytes):
    """One device-to-device copy of nbytes through the HIP runtime (its copy
    kernel reads and writes exactly nbytes), for scaling traffic counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) != 0 or hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipMemset failed")
        if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) != 0 or hip.hipDeviceSynchronize() != 0:  # DtoD
            raise RuntimeError("hipMemcpy failed")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    return nbytes


def hbm_ceiling(nbytes=4 << 30, reps=5):
    """Best rates of hand-written streaming kernels (read, write, copy,
    triad; liblcs_stream.so) over nbytes-sized arrays on this GPU: what a pure
    stream reaches, for reading the kernels' fractions of the 8 TB/s peak."""
    import ctypes as C
    L = C.CDLL(os.path.join(ROOT, "lcs.jl_amd", "lcs_mi355x", "liblcs_stream.so"))
    L.lcs_stream_probe.argtypes = [C.c_int, C.c_int, C.c_size_t, C.c_int, C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]
    nv = L.lcs_stream_variants()
    names = [f"U{u}" for u in (1, 2, 4, 8)] + [f"U{u}-nt" for u in (1, 2, 4, 8)] + ["dword", "dword-nt", "dword-U4", "dword-U16"]
    out = {"unit": "GB/s", "bytes_per_array": nbytes, "peak": PEAK_HBM_GBS}
    for kind, name in enumerate(("read", "write", "copy", "triad")):
        g, m = (C.c_double * nv)(), (C.c_double * nv)()
        rc = L.lcs_stream_probe(0, kind, nbytes, reps, g, m)
        if rc != 0:
            raise RuntimeError(f"stream probe {name}: hip error {rc}")
        best = max(range(nv), key=lambda v: g[v])
        out[name] = {"GBs": g[best], "frac": g[best] / PEAK_HBM_GBS, "variant": names[best],
                     "all": {names[v]: round(g[v], 1) for v in range(nv) if g[v] > 0}}
        print(f"# ceiling {name:5s} {g[best]:7.0f} GB/s ({names[best]})", file=sys.stderr, flush=True)
    return out


def ensemble_configs(nbox, seed=12345):
    from lcs_mi355x.workloads import ensemble_configs as make
    return make(nbox, seed)


def bench_ensemble(args):
    """Box-steps/s of the batched ensemble against the same boxes stepped one
    at a time with lcs_step on the same GPU."""
    import lcs_mi355x as L
    cfgs = ensemble_configs(args.boxes)
    t0 = time.perf_counter()
    boxes = [L.build(c) for c in cfgs]
    for b in boxes:
        b.sync()
    print(f"# built {len(boxes)} boxes in {time.perf_counter() - t0:.1f} s, "
          f"{sum(b.nbytes() for b in boxes) / 2**30:.1f} GiB", file=sys.stderr, flush=True)
    scfgs = [L.step_config(c, localize_interval=args.localize_interval) for c in cfgs]
    cells = sum(math.prod(c.grid.dims) for c in cfgs)
    if not args.no_autotune:  # the serial baseline runs each box with its best single-box tiles
        t0 = time.perf_counter()
        for b in boxes:
            L.autotune(b)
        print(f"# autotuned the serial boxes in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    # serial: every box alone, one lcs_step at a time
    for b, s in zip(boxes, scfgs):
        L.step(b, s, want_log=False)
        b.sync()
    t0 = time.perf_counter()
    for _ in range(args.serial_steps):
        for b, s in zip(boxes, scfgs):
            L.step(b, s, want_log=False)
    for b in boxes:
        b.sync()
    serial_s = time.perf_counter() - t0
    serial = len(boxes) * args.serial_steps / serial_s
    print(f"# serial: {serial:.1f} box-steps/s", file=sys.stderr, flush=True)
    batch = L.Batch()
    if args.no_graphs:
        batch.set_graphs(False)
    if args.persistent:
        batch.set_persistent(True)
    for b, s in zip(boxes, scfgs):
        batch.add(b, s)
    for _ in range(args.warmup):
        batch.step(want_log=False)
    batch.sync()
    tiling = None
    if not args.no_autotune:  # the batch's sweep tiling per size class, timed on the boxes as they are
        t0 = time.perf_counter()
        tiles = batch.autotune()
        rtiles = batch.rhs_tiles()
        tiling = {"x".join(map(str, c.grid.dims)): f"{t[0]}x{t[1]}, z-chunks of {t[2]}; RHS {r[0]}x{r[1]}"
                  for c, t, r in zip(cfgs, tiles, rtiles)}
        print(f"# batch tiling ({time.perf_counter() - t0:.1f} s): {tiling}", file=sys.stderr, flush=True)
        for _ in range(max(1, args.warmup)):  # tables rebuilt, graphs captured again
            batch.step(want_log=False)
        batch.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.step(want_log=False)
    batch.sync()
    el = time.perf_counter() - t0
    value = len(boxes) * args.steps / el
    gstats = batch.graph_stats()
    print(f"# batched: {value:.1f} box-steps/s ({value / serial:.2f}x serial); steps by path {gstats}",
          file=sys.stderr, flush=True)
    ns = sorted({c.grid.dims[0] for c in cfgs})
    out = {
        "metric": METRIC, "value": value, "unit": "box-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded spectra and particles per box)",
        "config": {"workload": f"batched ensemble of {len(boxes)} FP32 boxes, N in {ns}, particles log-uniform "
                               f"[1e3, 1e6], St log-uniform [0.1, 10], k0 in {{2, 4, 6}}, linear-power P=0.5",
                   "boxes": len(boxes), "cells": cells,
                   "particles": sum(c.particles[0].population.valid for c in cfgs),
                   "particle_localize_interval": args.localize_interval, "graphs": not args.no_graphs,
                   "persistent_sweeps": args.persistent,
                   "sweep_tiling": tiling or "32x8, z-chunks of about 28 planes; RHS 32x8 (default)",
                   "steps_by_path": gstats},
        "cell_steps_per_s": cells * args.steps / el,
        "serial": {"value": serial, "unit": "box-steps/s", "steps_per_box": args.serial_steps,
                   "how": "the same boxes, one lcs_step call per box at a time, same GPU"},
        "speedup_vs_serial": value / serial,
    }
    batch.close()
    for b in boxes:
        b.close()
    return out


def bench_reference_case(args, steps=6, warmup=3):
    """The one-device case of the reference's weak-scaling runs (workload c6:
    750^3 FP64, rc-energy, 325^3 particles with nonlinear drag), timed beside
    the published figure it corresponds to (data/gpu-weak.csv:2: 0.9237 s on
    an H100 in FP64; the csv does not say whether that is one step)."""
    import lcs_mi355x as L
    cfg, desc = workload("c6")
    box = L.build(cfg)
    tiles = L.get_tiles(box) if args.no_autotune else L.autotune(box)
    scfg = L.step_config(cfg, localize_interval=args.localize_interval)
    for _ in range(warmup):
        L.step(box, scfg)
    box.sync()
    sweeps = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        log = L.step(box, scfg)
        sweeps += log["flow"]["couple1"]["itr"] + log["flow"]["couple2"]["itr"] + 2
    box.sync()
    el = time.perf_counter() - t0
    box.close()
    n3 = math.prod(cfg.grid.dims)
    ms = el / steps * 1e3
    print(f"# reference case: {desc}: {ms:.1f} ms/step", file=sys.stderr, flush=True)
    return {"workload": desc, "dtype": cfg.dtype, "steps": steps, "warmup": warmup, "ms_per_step": ms,
            "cell_steps_per_s": n3 * steps / el, "hsmac_sweeps_per_step": sweeps / steps,
            "tiles": {k: list(v) for k, v in tiles.items()},
            "reference": {"seconds": 0.9237, "hardware": "H100 SXM5, FP64", "source": "data/gpu-weak.csv:2",
                          "note": "unit not stated in the csv; the reference starts from a restart file of "
                                  "developed turbulence, this run from the seeded spectrum"},
            "ratio_if_per_step": 923.7 / ms}


def serve_requests(nreq, seed=4242):
    from lcs_mi355x.workloads import serve_requests as make
    return make(nreq, seed)


def _serve_stream(L, reqs, arrivals, max_live, localize_interval, tag):
    """Feeds reqs to a server at the given arrival times (seconds from the
    start) and steps it until every run has finished.  Returns (box-steps,
    wall seconds, completion latencies)."""
    import numpy as np
    S = L.Server(max_live=max_live, localize_interval=localize_interval)
    done, lat, steps_total, live_ids = 0, [], 0, set()
    nxt = 0
    t0 = time.perf_counter()
    last_report = t0
    while done < len(reqs):
        now = time.perf_counter() - t0
        while nxt < len(reqs) and arrivals[nxt] <= now:
            live_ids.add(S.submit(**reqs[nxt]))
            nxt += 1
        if not live_ids:
            time.sleep(max(0.0, min(0.01, arrivals[nxt] - now)))
            continue
        S.step()
        for rid in list(live_ids):
            st = S.status(rid)
            if st["status"] in ("done", "failed"):
                live_ids.discard(rid)
                done += 1
                if st["status"] == "done":
                    lat.append(st["finish_s"] - st["submit_s"])
                    steps_total += st["steps_done"]
                S.release(rid)
        if time.perf_counter() - last_report > 20:
            last_report = time.perf_counter()
            c = S.counts()
            print(f"# {tag} t={last_report - t0:.0f}s submitted {nxt} done {done} live {c['live']} "
                  f"queued {c['queued']}", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    S.close()
    return steps_total, wall, np.array(lat)


def bench_serve(args):
    """Continuous batching of a request stream, in two runs of the same
    requests: every request submitted at once (the server's capacity in
    box-steps/s, against the same requests' serial rate), then Poisson
    arrivals at a rate the server sustains (completion latency percentiles
    that measure service rather than a growing queue).  --rate fixes the
    arrival rate; by default it is --load times the capacity."""
    import numpy as np
    import lcs_mi355x as L
    reqs = serve_requests(args.requests)
    mean_steps = float(np.mean([r["steps"] for r in reqs]))
    cap_steps, cap_wall, cap_lat = _serve_stream(L, reqs, np.zeros(len(reqs)), args.max_live,
                                                 args.localize_interval, "capacity")
    capacity = cap_steps / cap_wall
    print(f"# served capacity: {capacity:.1f} box-steps/s ({capacity / mean_steps:.2f} requests/s)", file=sys.stderr,
          flush=True)
    rate = args.rate if args.rate else args.load * capacity / mean_steps
    rng = np.random.default_rng(99)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(reqs)))
    steps_total, wall, lat = _serve_stream(L, reqs, arrivals, args.max_live, args.localize_interval, "poisson")
    delivered = steps_total / wall
    # serial reference on a sample: the first few requests, a few steps each, one at a time
    sample = reqs[: args.serial_sample]
    boxes = []
    from lcs_mi355x.workloads import request_config
    for r in sample:
        cfg = request_config(r)
        b = L.build(cfg)
        if not args.no_autotune:
            L.autotune(b)
        boxes.append((b, L.step_config(cfg, localize_interval=args.localize_interval)))
    for b, s in boxes:
        L.step(b, s, want_log=False)
        b.sync()
    ts = time.perf_counter()
    for b, s in boxes:
        for _ in range(args.serial_steps):
            L.step(b, s, want_log=False)
        b.sync()
    serial = len(boxes) * args.serial_steps / (time.perf_counter() - ts)
    for b, _ in boxes:
        b.close()
    out = {
        "metric": METRIC, "value": capacity, "unit": "box-steps/s", "n_gpus": 1, "steps": int(cap_steps),
        "warmup": 0, "ms_per_step": 1e3 / capacity, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (per-request seeds)",
        "config": {"workload": f"served stream of {len(reqs)} requests, N in 32..128 step 16, particles "
                               f"log-uniform [1e3, 2e6], St log-uniform [0.1, 10], forcing linear-power or "
                               f"rc-energy, steps uniform [50, 500]",
                   "max_live": args.max_live, "particle_localize_interval": args.localize_interval,
                   "mean_steps_per_request": mean_steps},
        "capacity": {"box_steps_per_s": capacity, "wall_s": cap_wall, "how": "every request submitted at once; "
                     "box-steps of the whole stream over its makespan (drain included)",
                     "latency_s": {"p50": float(np.percentile(cap_lat, 50)),
                                   "p99": float(np.percentile(cap_lat, 99))}},
        "poisson": {"rate_requests_per_s": rate, "load": rate * mean_steps / capacity,
                    "delivered_box_steps_per_s": delivered, "wall_s": wall, "completed": int(len(lat)),
                    "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                                  "mean": float(lat.mean()), "max": float(lat.max())}},
        "serial": {"value": serial, "unit": "box-steps/s", "sample": f"first {len(boxes)} requests, "
                                                                      f"{args.serial_steps} steps each, one lcs_step at a time"},
        "speedup_vs_serial": capacity / serial,
    }
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(name, n, steps, n1, steps1):
    """Time the CPU oracle on a bounded sample (n^3 box, same parameters) with
    every OpenMP thread, and on a smaller sample (n1^3) with one thread."""
    from oracle import oracle as O
    threads = O.set_threads(0)
    out = oracle_rate(name, n, steps)
    out.update({"cores": threads, "kind": "port", "cpu_model": cpu_model(),
                "affinity_cores": len(os.sched_getaffinity(0))})
    O.set_threads(1)
    try:
        one = oracle_rate(name, n1, steps1)
    finally:
        O.set_threads(threads)
    out["one_core"] = {"value": one["value"], "unit": one["unit"], "cores": 1, "sample": one["sample"]}
    return out


def oracle_rate(name, n, steps):
    import numpy as np
    import lcs_mi355x as L
    from oracle.driver import OracleSim, ParticleSet
    cfg, desc = workload(name, n_override=n)
    box = L.build(cfg)  # the device only prepares the initial state
    dt = np.float32 if cfg.dtype == "f32" else np.float64
    sets = []
    for q, prof in enumerate(box.profiles):
        nv = prof["nvalid"]
        sets.append(ParticleSet([box.download_prop(q, k)[:nv] for k in ("x", "y", "z")],
                                [box.download_prop(q, k)[:nv] for k in ("u", "v", "w")],
                                box.download_prop(q, "diams")[:nv], box.download_prop(q, "ids")[:nv],
                                prof["ntotal"]))
    sim = OracleSim(cfg, dt, [box.download(k) for k in ("U", "V", "W")], box.download("P"), sets)
    box.close()
    sim.timestep()
    sim.evolve()  # warm-up (first-touch, thread pool)
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.timestep()
        sim.evolve()
    el = time.perf_counter() - t0
    return {"value": n ** 3 * steps / el, "unit": "cell-steps/s",
            "sample": f"{steps} timed steps (after 1 warm-up) of the same pipeline at {n}^3 ({desc}); "
                      f"{el:.1f} s wall"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c3", choices=["c1", "c2", "c3", "c4", "c5", "c6"])
    ap.add_argument("--requests", type=int, default=2000, help="served requests (workload c5)")
    ap.add_argument("--rate", type=float, default=0.0,
                    help="Poisson arrival rate, requests/s (c5); 0: --load times the measured capacity")
    ap.add_argument("--load", type=float, default=0.75, help="arrival rate as a fraction of capacity (c5)")
    ap.add_argument("--max-live", type=int, default=128, help="runs in the batch at once (c5)")
    ap.add_argument("--serial-sample", type=int, default=16, help="requests of the serial reference (c5)")
    ap.add_argument("--boxes", type=int, default=96, help="ensemble size (workload c4)")
    ap.add_argument("--serial-steps", type=int, default=3, help="steps per box of the serial baseline (c4)")
    ap.add_argument("--size", type=int, default=None, help="override the grid size (testing only)")
    ap.add_argument("--cpu-size", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=30)
    ap.add_argument("--cpu1-size", type=int, default=128, help="grid of the one-thread CPU sample")
    ap.add_argument("--cpu1-steps", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ceiling", action="store_true", help="skip the streaming-kernel HBM ceiling")
    ap.add_argument("--no-ensemble", action="store_true", help="default run: skip the 96-box ensemble sub-run")
    ap.add_argument("--ensemble-steps", type=int, default=30, help="timed batch steps of the ensemble sub-run")
    ap.add_argument("--no-serve", action="store_true", help="default run: skip the served-stream sub-run")
    ap.add_argument("--serve-requests", type=int, default=200, help="requests of the served-stream sub-run")
    ap.add_argument("--no-reference-case", action="store_true",
                    help="default run: skip the 750^3 FP64 sub-run of the reference's one-device case")
    ap.add_argument("--no-autotune", action="store_true", help="use the size rule for tile shapes")
    ap.add_argument("--persistent", action="store_true",
                    help="c4: batched HSMAC sweeps as persistent grids walking the tile table")
    ap.add_argument("--no-graphs", action="store_true",
                    help="c4: launch the batched step eagerly instead of replaying hipGraphs (kernel work is the "
                         "same; used for profiling, where rocprofv3 ended graph replays with a host fault)")
    ap.add_argument("--localize-interval", type=int, default=30,
                    help="reorder particle storage by position every N steps (0: never; changes no result)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the run, copy a MIB-sized device buffer once (a plain copy kernel of known size "
                         "for scaling hardware traffic counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (asked for {args.gpus}); ensembles would shard by box across "
              f"GPUs and a single box by slabs with halo exchange, neither of which is implemented", file=sys.stderr)
        sys.exit(2)

    if args.workload == "c4":
        out = bench_ensemble(args)
        if args.copy_calibration:
            out["copy_calibration_bytes"] = copy_calibration(args.copy_calibration << 20)
        print(json.dumps(out), flush=True)
        return
    if args.workload == "c5":
        print(json.dumps(bench_serve(args)), flush=True)
        return
    ceiling = None if args.no_ceiling else hbm_ceiling()
    import lcs_mi355x as L
    cfg, desc = workload(args.workload, n_override=args.size)
    t0 = time.perf_counter()
    box = L.build(cfg)
    init_s = time.perf_counter() - t0
    print(f"# built {desc} in {init_s:.1f} s, {box.nbytes() / 2**30:.1f} GiB on the device", file=sys.stderr,
          flush=True)
    t0 = time.perf_counter()
    tiles = L.get_tiles(box) if args.no_autotune else L.autotune(box)
    tune_s = time.perf_counter() - t0
    print(f"# tiles {tiles} ({'size rule' if args.no_autotune else f'autotuned in {tune_s:.1f} s'})",
          file=sys.stderr, flush=True)
    scfg = L.step_config(cfg, localize_interval=args.localize_interval)
    for s in range(args.warmup):
        log = L.step(box, scfg)
        print(f"# warmup {s}: itr {log['flow']['couple1']['itr']}/{log['flow']['couple2']['itr']} "
              f"dt {log['dt']:.3e}", file=sys.stderr, flush=True)
    box.sync()
    box.profile(True)
    sweeps = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        log = L.step(box, scfg)
        sweeps.append((log["flow"]["couple1"]["itr"] + 1, log["flow"]["couple2"]["itr"] + 1))
    box.sync()
    el = time.perf_counter() - t0
    prof = box.profile_read()
    box.profile(False)
    n3 = math.prod(cfg.grid.dims)
    value = n3 * args.steps / el
    esize = 4 if cfg.dtype == "f32" else 8
    s_mean = sum(a + b for a, b in sweeps) / len(sweeps)
    b_cell = esize * (29 + 8 * s_mean)  # algorithmic bytes per cell-step
    dom = max(prof.items(), key=lambda kv: kv[1][0])
    name, (ms, launches, byts) = dom
    achieved = byts / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"{name}_traffic.json")
    if os.path.exists(tf):
        with open(tf) as fh:
            t = json.load(fh)
        own = {"hsmac_sweep": "sweep", "rhs_update": "rhs"}.get(name)  # the kernel's own tile shape
        if (t.get("workload") == args.workload and t.get("size") == cfg.grid.dims[0]
                and own and t.get("tile") == list(tiles[own])):
            traffic = t.get("bytes_per_launch")
    kernels = {k: {"ms_per_launch": v[0] / max(v[1], 1), "launches": v[1],
                   "algo_GBs": (v[2] / (v[0] * 1e-3) / 1e9) if v[0] > 0 else None} for k, v in prof.items()}
    for k, v in sorted(kernels.items(), key=lambda kv: -kv[1]["ms_per_launch"] * kv[1]["launches"]):
        print(f"# {k:18s} {v['launches']:6d} launches  {v['ms_per_launch']:.3f} ms/launch  "
              f"{(v['algo_GBs'] or 0):8.1f} GB/s algorithmic", file=sys.stderr, flush=True)
    out = {
        "metric": METRIC, "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": cfg.dtype.replace("f", "f"), "data": "synthetic (seeded spectrum init)",
        "config": {"workload": desc, "grid": list(cfg.grid.dims), "particles": sum(p.population.valid for p in
                                                                                     cfg.particles),
                   "hsmac_sweeps_per_step": s_mean, "particle_localize_interval": args.localize_interval,
                   "tiles": {k: list(v) for k, v in tiles.items()}, "autotuned": not args.no_autotune},
        "step_hbm": {"bytes_per_cell_step": b_cell, "achieved_GBs": value * b_cell / 1e9,
                     "frac": value * b_cell / 1e9 / PEAK_HBM_GBS},
        "roofline": {"kernel": name, "bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                     "ms_per_launch": ms / max(launches, 1), "launches": launches},
        "kernels": kernels,
    }
    if args.localize_interval > 0 and box.profiles:
        # the timed steps fall between two reorders of particle storage
        # (steps 0 and 30), so none of them pays for one: time one here and
        # report the per-step share it would add
        box.sync()
        t0 = time.perf_counter()
        for q in range(len(box.profiles)):
            L.localize(box, q)
        box.sync()
        loc_ms = (time.perf_counter() - t0) * 1e3
        per_step = el / args.steps + loc_ms * 1e-3 / args.localize_interval
        out["particle_localize"] = {"ms": loc_ms, "interval_steps": args.localize_interval,
                                    "amortized_ms_per_step": loc_ms / args.localize_interval,
                                    "value_with_localize": n3 / per_step}
    box.close()
    if args.copy_calibration:
        out["copy_calibration_bytes"] = copy_calibration(args.copy_calibration << 20)
    if ceiling:
        out["hbm_ceiling"] = ceiling
        for k in ("roofline", "step_hbm"):
            a = out[k]["achieved"] if k == "roofline" else out[k]["achieved_GBs"]
            out[k]["frac_of_copy_ceiling"] = a / ceiling["copy"]["GBs"]
    if args.workload == "c3" and args.size is None and not args.no_ensemble:
        ea = argparse.Namespace(**vars(args))
        ea.steps, ea.warmup, ea.serial_steps = args.ensemble_steps, 3, 3
        out["ensemble"] = bench_ensemble(ea)
    if args.workload == "c3" and args.size is None and not args.no_serve:
        sa = argparse.Namespace(**vars(args))
        sa.requests = args.serve_requests
        out["served"] = bench_serve(sa)
    if args.workload == "c3" and args.size is None and not args.no_reference_case:
        out["reference_case"] = bench_reference_case(args)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_size, args.cpu_steps, args.cpu1_size,
                                           args.cpu1_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
