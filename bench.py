#!/usr/bin/env python3
"""Benchmark: LULESH Sedov Lagrange leapfrog on one MI355X.

A step is one leapfrog cycle (TimeIncrement + LagrangeLeapFrog) of the
Sedov blast problem, -s 100 -r 11 -b 1 -c 1 (10^6 hex elements, FP64), with
the whole mesh resident in HBM before the timed region starts.  The metric
is true zone-cycles per second (elements x cycles / wall time); the
reference's printed FOM is that figure / 1000.

Default run: --steps 200 timed cycles after --warmup 20, then
  * a per-kernel pass timed with HIP events on the library's stream,
    giving the dominant kernel's achieved HBM bandwidth against 8 TB/s;
  * the reference's CPU code (oracle/_ref/lulesh_ref, built from the
    unmodified omp_4.0 sources) timed on the host cores on a bounded sample
    of the same problem, or the C restatement in oracle/ when that binary is
    absent.

The headline line also carries the batched and served halves of the metric
as sub-objects ("batched": the 64-problem ensemble to stoptime vs sequential
plain runs of the same mix; "served": a Poisson stream of --serve-requests
requests at --serve-rate per second, a load the GPU keeps up with, for
latency; "served_saturated": --serve-sat-requests requests arriving faster
than the GPU serves them, for throughput against sequential plain runs;
"multidomain": 2x2x2 subdomains of -s 45 on the one GPU, halo exchange
included, on each halo path); --no-extra skips them.

Other modes (one JSON line each):
  --mode batched   64-problem ensemble vs sequential plain runs
  --mode serve     Poisson-arrival serving loop, p50/p99 latency
  --mode multidomain  2x2x2 subdomains of -s 45 on one GPU
  --mode umesh     the unstructured-mesh (.lmesh) path: the -s 100 block as
                   an unstructured mesh, and the reference's sedov15oct mesh
                   to stoptime
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "cuda-ipc-benchmarks_amd"))

import lulesh_amd  # noqa: E402

HBM_PEAK_GBS = 8000.0

# algorithmic (compulsory) bytes per launch of each kernel class, per
# element E and node N of the launch (E elements, N nodes of the mesh)
KERNEL_NAMES = ["time_increment", "force_gather", "node_gather_update", "node_update",
                "kinematics", "qeos", "halo_exchange"]


TWOPASS = os.environ.get("LULESH_FORCE_PATH") == "twopass"


def kernel_bytes(k, E, N, fused_update=False):
    if k == 1 and not TWOPASS and fused_update:
        # force + gather + node update: x..zd, p q v volo hgcoef, nodalMass in;
        # the new x..zd out (forces stay on chip)
        return 48 * N + 40 * E + 8 * N + 48 * N
    if k == 1 and not TWOPASS:   # fused force+gather: x..zd, p q v volo hgcoef in, f out
        return 48 * N + 40 * E + 24 * N
    if k == 1:   # x..zd of 8 nodes (once per node), 6 elem doubles in, 48 corner doubles out
        return 48 * N + 48 * E + 384 * E
    if k == 2:   # 48 corner doubles in, nodalMass + x..zd in, x..zd out
        return 384 * E + 8 * N + 48 * N + 48 * N
    if k == 4:   # x..zd in, volo v in, 10 doubles out
        return 48 * N + 16 * E + 80 * E
    if k == 5:   # 15 doubles + elemRep in; p e q ss v hgcoef out
        return 120 * E + 4 * E + 48 * E
    if k == 3:
        return 24 * N + 8 * N + 48 * N + 48 * N
    return 0


KERNEL_SYMBOLS = {1: "k_force_march", 3: "k_node_update", 4: "k_kinematics", 5: "k_qeos"}


def measured_traffic(k, nx, fused_update=False):
    """HBM bytes per launch of kernel class k at -s nx, as measured by the
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs summarised in profiles/ (copy-
    calibrated); None when no such measurement is committed."""
    sym = KERNEL_SYMBOLS.get(k)
    if sym is None or TWOPASS:
        return None
    if k == 1 and fused_update:
        sym += "+update"
    pdir = os.path.join(REPO, "profiles")
    if not os.path.isdir(pdir):
        return None
    for fn in sorted(os.listdir(pdir)):
        if not fn.endswith(".md"):
            continue
        with open(os.path.join(pdir, fn)) as f:
            for line in f:
                w = line.split()
                if len(w) == 4 and w[0] == "traffic_bytes_per_launch" and w[1] == sym \
                        and w[2] == f"s{nx}":
                    return int(w[3])
    return None


FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector (an FMA counts as 2)


def profiled_value(key, k, nx, fused_update=False):
    """A per-launch figure of kernel class k at -s nx from the rocprofv3
    summaries in profiles/ (lines '<key> <kernel> s<nx> <value>')."""
    sym = KERNEL_SYMBOLS.get(k)
    if sym is None or TWOPASS:
        return None
    if k == 1 and fused_update:
        sym += "+update"
    pdir = os.path.join(REPO, "profiles")
    if not os.path.isdir(pdir):
        return None
    for fn in sorted(os.listdir(pdir)):
        if not fn.endswith(".md"):
            continue
        with open(os.path.join(pdir, fn)) as f:
            for line in f:
                w = line.split()
                if len(w) == 4 and w[0] == key and w[1] == sym and w[2] == f"s{nx}":
                    return float(w[3])
    return None


def host_cpu_info():
    """CPU model, physical cores and logical CPUs of this host, and the
    CPUs this process may run on."""
    model, cores = None, set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    logical = os.cpu_count() or 1
    return {"cpu_model": model, "host_cores": len(cores) or logical, "host_logical_cpus": logical,
            "affinity_cpus": len(os.sched_getaffinity(0))}


# host threads a benchmark process may use on the GPU machines: one GPU's
# job is allotted 16 CPUs whatever the host has
CPU_SHARE = 16


def cpu_baseline(nx, target_s=8.0, samples=3):
    """The reference's own CPU code on the host cores: median of `samples`
    bounded runs of the same problem (-s nx -i K, K sized to ~target_s)."""
    ref = os.path.join(REPO, "oracle", "_ref", "lulesh_ref")
    info = host_cpu_info()
    threads = max(2, min(info["affinity_cpus"], CPU_SHARE))
    why = (f"{threads} threads: the CPUs this job may use (min of its affinity set, "
           f"{info['affinity_cpus']}, and the {CPU_SHARE}-CPU share of one GPU's job)")
    if os.path.exists(ref):
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
                   OMP_PLACES="cores")

        def run(its):
            t0 = time.time()
            out = subprocess.run([ref, "-s", str(nx), "-i", str(its)], env=env,
                                 capture_output=True, text=True, timeout=600).stdout
            el = time.time() - t0
            for line in out.splitlines():
                if line.startswith("Grind time"):
                    return float(line.split("=")[1].split()[0]), el
            return None, el

        # probe a few cycles, then size each sample to ~target_s
        its = 3
        grind, el = run(its)
        its = max(its + 1, min(400, int(its * target_s / max(el, 1e-3))))
        rates = []
        for _ in range(samples):
            grind, _ = run(its)
            if grind:
                rates.append(1e6 / grind)
        if rates:
            return dict({"value": float(np.median(rates)), "unit": "zone-cycles/s",
                         "cores": threads, "kind": "reference",
                         "sample": f"omp_4.0 lulesh -s {nx} -i {its}, median of {len(rates)} runs "
                                   f"(unmodified reference sources, g++ -O3 -fopenmp, "
                                   f"OMP_NUM_THREADS={threads}, OMP_PROC_BIND=close)",
                         "samples": rates, "threads_note": why}, **info)
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle
    o = pyoracle.Oracle(nx)
    t0 = time.time()
    its = 0
    while time.time() - t0 < target_s and its < 50:
        o.run(1)
        its += 1
    el = time.time() - t0
    return dict({"value": nx ** 3 * its / el, "unit": "zone-cycles/s", "cores": 1, "kind": "port",
                 "sample": f"oracle C restatement -s {nx}, {its} cycles, 1 thread"}, **info)


# Cycles of a Sedov -s 100 run before its stop time (3517); the spin-up
# below never lets warm-up + spin-up + timed cycles reach it.
S100_CYCLES = 3517


def clock_spinup(g, chunk, cycle, steps, busy_s=0.15, stop_cycles=S100_CYCLES):
    """Keeps the GPU busy with real cycles of the run for ~busy_s before the
    timed region.  An MI355X coming out of idle raises its clocks over the
    first ~0.1 s of load: the force kernel takes ~108-111 us per launch over
    the first 20 cycles of a fresh process and ~104 us from about cycle 35
    on (rocprofv3 kernel trace), so a timed region that starts a few cycles
    after start-up measures the clock ramp, not the kernels.  Returns the
    cycles run."""
    n = 0
    room = max(0, stop_cycles - cycle - steps - chunk - 10)
    t0 = time.perf_counter()
    while n + chunk <= room and time.perf_counter() - t0 < busy_s:
        g.enqueue(chunk)
        g.sync()
        n += chunk
    return n


def bench_plain(args):
    nx = args.size
    E, N = nx ** 3, (nx + 1) ** 3
    g = lulesh_amd.Group(nx, 1, 11, 1, 1)
    g.enqueue(args.warmup)
    g.sync()
    # the timed loop enqueues graphs of `chunk` cycles and one of the
    # remainder: both shapes are captured (and run) before the timed region
    chunk = max(1, min(args.chunk, args.steps))
    g.enqueue(chunk)
    if args.steps % chunk:
        g.enqueue(args.steps % chunk)
    rc, c0, _, _ = g.sync()
    spin = clock_spinup(g, chunk, c0, args.steps)
    rc, c0, _, _ = g.sync()
    t0 = time.perf_counter()
    done = 0
    while done < args.steps:
        n = min(chunk, args.steps - done)
        g.enqueue(n)
        done += n
    rc, c1, _, _ = g.sync()
    t1 = time.perf_counter()
    if rc != 0:
        raise SystemExit(f"run failed: {lulesh_amd.status_string(rc)}")
    cycles = c1 - c0
    if cycles != args.steps:
        raise SystemExit(f"expected {args.steps} cycles, device ran {cycles}")
    el = t1 - t0
    value = E * cycles / el

    # per-kernel live timing (HIP events on the library stream)
    kcycles = 10
    ms = g.profile(kcycles)
    per = ms / kcycles
    fu = per[3] == 0 and per[2] == 0   # node update folded into the force kernel
    dom = int(np.argmax(per))
    b = kernel_bytes(dom, E, N, fu)
    achieved = b / (per[dom] * 1e-3) / 1e9
    kern = {KERNEL_NAMES[k]: {"ms": round(float(per[k]), 4),
                              "GBs": round(kernel_bytes(k, E, N, fu) / (per[k] * 1e-3) / 1e9, 1)
                              if per[k] > 0 and kernel_bytes(k, E, N, fu) else None}
            for k in range(7) if per[k] > 0}
    line = {
        "metric": "LULESH FOM zones/s at -s 100 on 1 GPU (true zone-cycles/s; printed FOM = value/1000)",
        "value": value,
        "unit": "zone-cycles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / cycles,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: Sedov blast initial condition (deterministic set-up of the reference)",
        "config": {"workload": f"LULESH Sedov -s {nx} -r 11 -b 1 -c 1, single domain, plain path",
                   "elements": E, "nodes": N, "graph_cycles_per_launch": chunk,
                   "clock_spinup_cycles": spin},
        "roofline": {"bound": "hbm", "kernel": KERNEL_NAMES[dom], "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(dom, nx, fu), "algorithmic_bytes": b},
        "kernels": kern,
    }
    # the dominant kernel is FP64-issue bound: its FP64 rate against the
    # vector peak, from the FLOPs per launch rocprofv3 counted (profiles/)
    flops = profiled_value("fp64_flops_per_launch", dom, nx, fu)
    if flops:
        tf = flops / (per[dom] * 1e-3) / 1e12
        line["roofline"]["compute"] = {
            "bound": "fp64 valu", "achieved": tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": tf / FP64_PEAK_TFLOPS, "flops_per_launch": flops,
            "valu_issue_busy": profiled_value("valu_busy", dom, nx, fu)}
        # what any code can issue at the force kernel's occupancy (232
        # VGPRs: two waves per SIMD), measured now on this GPU: eight
        # independent FP64 multiply / add chains per lane, no memory traffic
        ceil_tf, mhz = lulesh_amd.probe_fp64(2)
        line["roofline"]["compute"]["issue_ceiling"] = {
            "waves_per_simd": 2, "achieved": ceil_tf, "unit": "TFLOP/s", "clock_mhz": round(mhz),
            "frac": tf / ceil_tf}
    g.close()
    if not args.no_cpu:
        progress("reference CPU baseline (3 bounded samples)")
        line["cpu_baseline"] = cpu_baseline(nx)
    return line


def problem_mix(n, seed, s_lo, s_hi, r_lo, r_hi, balances=(1,), costs=(1,)):
    """Seeded problem mix (numpy PCG64)."""
    rng = np.random.default_rng(seed)
    nx = rng.integers(s_lo, s_hi + 1, n)
    nr = rng.integers(r_lo, r_hi + 1, n)
    b = rng.choice(np.array(balances), n)
    c = rng.choice(np.array(costs), n)
    return nx.astype(np.int32), nr.astype(np.int32), b.astype(np.int32), c.astype(np.int32)


def run_sequential(mix, idx):
    """Plain single-problem runs, one after the other (the comparison
    baseline of the batched and served forms).  Returns (zone-cycles, s,
    per-problem (cycles, origin energy))."""
    nx, nr, b, c = mix
    zc, t, res = 0, 0.0, {}
    for i in idx:
        g = lulesh_amd.Group(int(nx[i]), 1, int(nr[i]), int(b[i]), int(c[i]))
        # the engine is built and its graph captured by a first, untimed
        # chunk of real cycles, as the batched and served forms' are before
        # their timed regions; the timed part is the rest of the run
        g.enqueue(64)
        rc0, c0, _, _ = g.sync()
        t0 = time.perf_counter()
        rc = g.run(64) if rc0 == 0 else rc0
        t += time.perf_counter() - t0
        if rc != 0:
            raise SystemExit(f"sequential run {i} failed: {rc}")
        cyc = g.sync()[1]
        res[int(i)] = (cyc, g.verify()["origin_energy"])
        zc += int(nx[i]) ** 3 * (cyc - c0)
        g.close()
    return zc, t, res


def bench_batched(args):
    mix = problem_mix(64, args.seed, 16, 64, 1, 16, (1, 2, 4), (1, 2))
    nx = mix[0]
    # autotune on a throwaway copy of the ensemble
    tb = lulesh_amd.Batch(*mix)
    tune = tb.autotune(5)
    tb.close()
    b = lulesh_amd.Batch(*mix)
    b.set_tiles(tune["tile_elems"], tune["tile_nodes"])
    b.set_kchunk(tune["kchunk"])
    b.enqueue(args.chunk)          # capture outside the timed region (real cycles)
    b.synchronize()
    c0, _ = b.results()
    t0 = time.perf_counter()
    rc = b.run(args.chunk)
    t1 = time.perf_counter()
    cyc, st = b.results()
    if rc != 0 or np.any(st != 0):
        raise SystemExit(f"batched run failed: {rc} {st}")
    zc = int(np.sum(nx.astype(np.int64) ** 3 * (cyc - c0)))
    value = zc / (t1 - t0)
    energies = [b.group(i).verify()["origin_energy"] for i in range(len(nx))]
    # sequential plain runs of the same mix: every problem, to stoptime
    szc, st_s, seq = run_sequential(mix, range(len(nx)))
    seq_rate = szc / st_s
    worst = max(abs(energies[i] - seq[i][1]) / abs(seq[i][1]) for i in range(len(nx)))
    same_cycles = all(int(cyc[i]) == seq[i][0] for i in range(len(nx)))
    return {
        "metric": "batched Sedov ensemble aggregate zone-cycles/s vs sequential plain runs",
        "value": value, "unit": "zone-cycles/s", "n_gpus": 1, "steps": int(np.max(cyc)),
        "warmup": args.chunk, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": f"synthetic: 64 Sedov problems, seed {args.seed}",
        "config": {"workload": "64 problems, s~U{16..64}, r~U{1..16}, b in {1,2,4}, c in {1,2}, "
                               "to stoptime", "zone_cycles": zc, "autotune": tune},
        "sequential_plain": {"value": seq_rate, "zone_cycles": szc, "seconds": st_s},
        "speedup_vs_sequential": value / seq_rate,
        "max_origin_energy_rel_diff_vs_plain": worst, "cycles_match_plain": same_cycles,
    }


def bench_serve(args):
    n = args.requests
    mix = problem_mix(n, args.seed, 20, 80, 1, 11)
    rng = np.random.default_rng(args.seed + 1)
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, n))
    srv = lulesh_amd.Server(0, args.slots)
    t_start = srv.now()
    ids = [srv.submit(int(mix[0][i]), int(mix[1][i]), 1, 1, float(t_start + arrivals[i]))
           for i in range(n)]
    done = 0
    last = time.time()
    while done < n:
        done += srv.poll(args.chunk)
        if time.time() - last > 20:
            print(f"# serve: {done}/{n} done, {srv.stats()}", file=sys.stderr, flush=True)
            last = time.time()
    t_end = srv.now()
    res = [srv.result(i) for i in ids]
    lat = np.array([r["t_done"] - (t_start + arrivals[i]) for i, r in enumerate(res)])
    zc = int(sum(int(mix[0][i]) ** 3 * r["cycles"] for i, r in enumerate(res)))
    bad = [i for i, r in enumerate(res) if r["status"] != 0]
    srv.close()
    # sequential plain runs of a sample of the same mix (rate per zone-cycle)
    sample = list(range(0, n, max(1, n // args.seq_sample)))[: args.seq_sample]
    szc, st_s, seq = run_sequential(mix, sample)
    worst = max(abs(res[i]["origin_energy"] - seq[i][1]) / abs(seq[i][1]) for i in sample)
    # The stream's throughput is bounded by its arrivals (offered load), so
    # the server is compared on latency with a one-at-a-time FIFO server fed
    # the same arrivals, each request taking its zone-cycles at the
    # sequential sample's rate (a model from measured rates, not a run).
    seq_rate = szc / st_s
    t_free, fifo = 0.0, []
    for i, r in enumerate(res):
        start = max(float(arrivals[i]), t_free)
        t_free = start + int(mix[0][i]) ** 3 * r["cycles"] / seq_rate
        fifo.append(t_free - float(arrivals[i]))
    fifo = np.array(fifo)
    return {
        "metric": "served Sedov stream: aggregate zone-cycles/s, p50/p99 completion latency",
        "value": zc / (t_end - t_start), "unit": "zone-cycles/s", "n_gpus": 1, "steps": n,
        "warmup": 0, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": f"synthetic: {n} Sedov requests, Poisson rate {args.rate}/s, seed {args.seed}",
        "config": {"workload": "s~U{20..80}, r~U{1..11}, to stoptime, continuous batching",
                   "slots": args.slots, "zone_cycles": zc},
        "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                      "mean": float(lat.mean())},
        "wall_s": t_end - t_start, "failed": len(bad),
        "offered_load": zc / float(arrivals[-1]),
        "sequential_plain_sample": {"requests": len(sample), "value": seq_rate},
        # below 1 while the stream is arrival-bound (value ~ offered_load)
        "speedup_vs_sequential": (zc / (t_end - t_start)) / seq_rate,
        "fifo_sequential_model_latency_s": {"p50": float(np.percentile(fifo, 50)),
                                            "p99": float(np.percentile(fifo, 99)),
                                            "mean": float(fifo.mean())},
        "max_origin_energy_rel_diff_vs_plain": worst,
    }


def multidomain_run(args, path):
    """ms per cycle of 2x2x2 subdomains of -s 45 on one halo path:
    "direct" (device_local, the in-place halo steps), "exchange"
    (device_local, pack and unpack kernels through the Comm* layout) or
    "device_copy" (Mode A backend: local pack, one copy per message)."""
    if path == "exchange":
        os.environ["LULESH_NO_DIRECT"] = "1"   # read when the group's engine is built
    try:
        g = lulesh_amd.Group(45, 2, 11, 1, 1)
        if path == "device_copy":
            g.set_comm_backend("device_copy")
        g.enqueue(args.warmup)
        g.enqueue(args.chunk)
        if args.steps % args.chunk:
            g.enqueue(args.steps % args.chunk)
        rc, c0, _, _ = g.sync()
        clock_spinup(g, args.chunk, c0, args.steps, stop_cycles=3145)
        rc, c0, _, _ = g.sync()
        t0 = time.perf_counter()
        done = 0
        while done < args.steps:
            k = min(args.chunk, args.steps - done)
            g.enqueue(k)
            done += k
        rc, c1, _, _ = g.sync()
        t1 = time.perf_counter()
        if rc != 0 or c1 - c0 != args.steps:
            raise SystemExit(f"multidomain {path} run failed: rc {rc}, {c1 - c0} cycles")
        ms = g.profile(5) / 5
        g.close()
    finally:
        os.environ.pop("LULESH_NO_DIRECT", None)
    return (t1 - t0) * 1e3 / (c1 - c0), {KERNEL_NAMES[k]: round(float(ms[k]), 4) for k in range(7) if ms[k] > 0}


def bench_multidomain(args):
    E = 8 * 45 ** 3
    paths = {}
    for path in ("direct", "exchange", "device_copy"):
        if args.halo_path not in ("all", path):
            continue
        progress(f"multidomain {path}")
        ms, kern = multidomain_run(args, path)
        paths[path] = {"ms_per_step": ms, "value": E / (ms * 1e-3), "kernels_ms": kern}
    best = paths["direct" if args.halo_path == "all" else args.halo_path]
    return {
        "metric": "8 subdomains (2x2x2 of -s 45) on one GPU, zone-cycles/s incl. on-device halo exchange",
        "value": best["value"], "unit": "zone-cycles/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": best["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: Sedov blast, 8-rank decomposition on one GPU",
        "config": {"workload": "2x2x2 subdomains, -s 45 each; value = "
                               + ("direct halo steps" if args.halo_path == "all" else args.halo_path)},
        "halo_paths": paths,
    }


def bench_umesh(args):
    nx = args.size
    E = nx ** 3
    m = lulesh_amd.Mesh.from_arrays(lulesh_amd.block_mesh_arrays(nx))
    scale = nx / 45.0
    m.set_energy(3.948746e+7 * scale * scale * scale)   # the block's own deposit
    g = lulesh_amd.Group.from_mesh(m)
    m.close()
    g.enqueue(args.warmup)
    g.enqueue(args.chunk)
    if args.steps % args.chunk:
        g.enqueue(args.steps % args.chunk)
    rc, c0, _, _ = g.sync()
    t0 = time.perf_counter()
    done = 0
    while done < args.steps:
        k = min(args.chunk, args.steps - done)
        g.enqueue(k)
        done += k
    rc, c1, _, _ = g.sync()
    t1 = time.perf_counter()
    if rc != 0 or c1 - c0 != args.steps:
        raise SystemExit(f"umesh run failed: rc {rc}, {c1 - c0} cycles")
    ms = g.profile(10) / 10
    g.close()
    # the reference's own unstructured mesh, 875 elements, to stoptime
    small = {}
    path = os.path.join(REPO, "tests", "golden", "sedov15oct.lmesh")
    if os.path.exists(path):
        gs = lulesh_amd.Group.from_mesh(lulesh_amd.Mesh.read(path))
        gs.set_max_cycles(20)
        gs.run()
        gs.set_max_cycles(-1)
        s0 = time.perf_counter()
        rc2 = gs.run(100)
        s1 = time.perf_counter()
        _, cyc, _, _ = gs.sync()
        small = {"mesh": "sedov15oct.lmesh (875 elements)", "status": rc2, "cycles": cyc - 20,
                 "ms_per_cycle": (s1 - s0) * 1e3 / max(1, cyc - 20)}
        gs.close()
    return {
        "metric": "unstructured-mesh path: zone-cycles/s on the -s 100 block given as an .lmesh-style mesh",
        "value": E * (c1 - c0) / (t1 - t0), "unit": "zone-cycles/s", "n_gpus": 1,
        "steps": c1 - c0, "warmup": args.warmup, "ms_per_step": (t1 - t0) * 1e3 / (c1 - c0),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: Sedov block written as an unstructured mesh (explicit connectivity)",
        "config": {"workload": f"-s {nx} block as an unstructured mesh, -r 11 -b 1 -c 1",
                   "elements": E},
        "kernels_ms": {KERNEL_NAMES[k]: round(float(ms[k]), 4) for k in range(7) if ms[k] > 0},
        "sedov15oct": small,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=100)
    ap.add_argument("--chunk", type=int, default=50, help="cycles per captured graph launch")
    ap.add_argument("--mode", default="plain",
                    choices=["plain", "copy", "batched", "serve", "multidomain", "umesh"])
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--halo-path", default="all", choices=["all", "direct", "exchange", "device_copy"],
                    help="multidomain mode: the halo paths to time")
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--rate", type=float, default=8.0, help="serve: Poisson arrivals per second")
    ap.add_argument("--slots", type=int, default=64, help="serve: max resident problems")
    ap.add_argument("--seq-sample", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="plain mode: skip the batched and served sub-objects")
    ap.add_argument("--serve-requests", type=int, default=256,
                    help="plain mode: requests of the served sub-object")
    ap.add_argument("--serve-rate", type=float, default=12.0,
                    help="plain mode: Poisson arrivals per second of the served sub-object")
    ap.add_argument("--serve-sat-requests", type=int, default=400,
                    help="plain mode: requests of the saturated served sub-object (0: none)")
    ap.add_argument("--serve-sat-rate", type=float, default=200.0,
                    help="plain mode: arrivals per second of the saturated served sub-object")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(subdomains of a decomposition share one GPU's HBM instead)", file=sys.stderr)
        raise SystemExit(2)
    lulesh_amd.init(0)
    if args.mode == "copy":
        # streaming-copy yardstick: 2^26 doubles (512 MiB) read + written
        n = 1 << 26
        ms = lulesh_amd.probe_copy_ms(n, max(1, args.steps))
        print(json.dumps({"mode": "copy", "bytes_read": 8 * n, "bytes_written": 8 * n,
                          "ms": ms, "GBs": 16 * n / (ms * 1e-3) / 1e9}), flush=True)
        return
    if args.mode == "batched":
        line = bench_batched(args)
    elif args.mode == "serve":
        line = bench_serve(args)
    elif args.mode == "multidomain":
        line = bench_multidomain(args)
    elif args.mode == "umesh":
        line = bench_umesh(args)
    else:
        line = bench_plain(args)
        if not args.no_extra:
            line.update(headline_extras(args))
    print(json.dumps(line), flush=True)


def progress(msg):
    print(f"# bench: {msg}", file=sys.stderr, flush=True)


def headline_extras(args):
    """The batched and served halves of the metric, as sub-objects of the
    headline line: the 64-problem ensemble to stoptime against sequential
    plain runs of the same mix, and a served Poisson stream (fewer requests
    than the 2,000 of the full stream, stated, so the default run stays
    within minutes)."""
    progress("batched ensemble (64 problems to stoptime) and its sequential plain runs")
    b = bench_batched(args)
    batched = {"value": b["value"], "unit": b["unit"], "workload": b["config"]["workload"],
               "problems": 64, "seed": args.seed, "zone_cycles": b["config"]["zone_cycles"],
               "autotune": b["config"]["autotune"], "sequential_plain_value": b["sequential_plain"]["value"],
               "speedup_vs_sequential": b["speedup_vs_sequential"],
               "max_origin_energy_rel_diff_vs_plain": b["max_origin_energy_rel_diff_vs_plain"],
               "cycles_match_plain": b["cycles_match_plain"]}
    def served(requests, rate):
        sa = argparse.Namespace(**vars(args))
        sa.requests, sa.rate = requests, rate
        sa.seq_sample = min(args.seq_sample, requests)
        progress(f"served stream ({sa.requests} requests at {sa.rate}/s)")
        v = bench_serve(sa)
        return {"value": v["value"], "unit": v["unit"], "workload": v["config"]["workload"],
                "requests": sa.requests, "full_stream_requests": 2000, "rate_per_s": sa.rate,
                "slots": sa.slots, "seed": args.seed, "offered_load": v["offered_load"],
                "latency_s": v["latency_s"], "wall_s": v["wall_s"], "failed": v["failed"],
                "sequential_plain_sample_value": v["sequential_plain_sample"]["value"],
                "speedup_vs_sequential": v["speedup_vs_sequential"],
                "fifo_sequential_model_latency_s": v["fifo_sequential_model_latency_s"],
                "max_origin_energy_rel_diff_vs_plain": v["max_origin_energy_rel_diff_vs_plain"]}
    # latency under a load the GPU keeps up with, and throughput with
    # arrivals faster than it can serve (the served form against
    # sequential plain runs)
    out = {"batched": batched, "served": served(args.serve_requests, args.serve_rate)}
    if args.serve_sat_requests > 0:
        out["served_saturated"] = served(args.serve_sat_requests, args.serve_sat_rate)
    # eight subdomains (2x2x2 of -s 45) on the one GPU, halo exchange
    # included, on each halo path
    ma = argparse.Namespace(**vars(args))
    ma.halo_path = "all"
    md = bench_multidomain(ma)
    out["multidomain"] = {"unit": md["unit"], "workload": md["config"]["workload"],
                          "steps": md["steps"],
                          "halo_paths": {k: {"value": v["value"], "ms_per_step": v["ms_per_step"]}
                                         for k, v in md["halo_paths"].items()}}
    return out


if __name__ == "__main__":
    main()
