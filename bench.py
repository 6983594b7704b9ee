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

Other modes (extra JSON lines are not emitted by default):
  --mode batched   64-problem ensemble vs sequential plain runs
  --mode serve     Poisson-arrival serving loop, p50/p99 latency
  --mode multidomain  2x2x2 subdomains of -s 45 on one GPU
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
# element E and node N of the launch; see DESIGN.md
KERNEL_NAMES = ["time_increment", "force_gather", "node_gather_update", "node_update",
                "kinematics", "qeos", "halo_exchange"]


TWOPASS = os.environ.get("LULESH_FORCE_PATH") == "twopass"


def kernel_bytes(k, E, N):
    if k == 1 and not TWOPASS:   # fused force+gather: x..zd, 6 elem doubles in, f out
        return 48 * N + 48 * E + 24 * N
    if k == 1:   # x..zd of 8 nodes (once per node), 6 elem doubles in, 48 corner doubles out
        return 48 * N + 48 * E + 384 * E
    if k == 2:   # 48 corner doubles in, nodalMass + x..zd in, x..zd out
        return 384 * E + 8 * N + 48 * N + 48 * N
    if k == 4:   # x..zd in, volo v in, 10 doubles out
        return 48 * N + 16 * E + 80 * E
    if k == 5:   # 15 doubles + elemRep in, 5 doubles out
        return 120 * E + 4 * E + 40 * E
    if k == 3:
        return 24 * N + 8 * N + 48 * N + 48 * N
    return 0


def cpu_baseline(nx, target_s=15.0):
    """The reference's own CPU code on the host cores, bounded sample."""
    ref = os.path.join(REPO, "oracle", "_ref", "lulesh_ref")
    cores = len(os.sched_getaffinity(0))
    threads = max(2, min(cores, 16))
    if os.path.exists(ref):
        # probe a few cycles, then size the sample to ~target_s
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
                   OMP_PLACES="cores")
        its = 3
        for _ in range(2):
            t0 = time.time()
            out = subprocess.run([ref, "-s", str(nx), "-i", str(its)], env=env,
                                 capture_output=True, text=True, timeout=600).stdout
            el = time.time() - t0
            if el > target_s / 3 or its >= 200:
                break
            its = max(its + 1, min(200, int(its * target_s / max(el, 1e-3))))
        grind = None
        for line in out.splitlines():
            if line.startswith("Grind time"):
                grind = float(line.split("=")[1].split()[0])
        if grind:
            return {"value": 1e6 / grind, "unit": "zone-cycles/s", "cores": threads,
                    "kind": "reference",
                    "sample": f"omp_4.0 lulesh -s {nx} -i {its} (unmodified reference sources, "
                              f"g++ -O3 -fopenmp, OMP_NUM_THREADS={threads})"}
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle
    o = pyoracle.Oracle(nx)
    t0 = time.time()
    its = 0
    while time.time() - t0 < target_s and its < 50:
        o.run(1)
        its += 1
    el = time.time() - t0
    return {"value": nx ** 3 * its / el, "unit": "zone-cycles/s", "cores": 1, "kind": "port",
            "sample": f"oracle C restatement -s {nx}, {its} cycles, 1 thread"}


def bench_plain(args):
    nx = args.size
    E, N = nx ** 3, (nx + 1) ** 3
    g = lulesh_amd.Group(nx, 1, 11, 1, 1)
    # warmup (also builds and caches the captured graph)
    g.enqueue(args.warmup)
    g.sync()
    chunk = args.chunk
    g.enqueue(chunk)      # capture the timed graph shape outside the timed region
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
    dom = int(np.argmax(per))
    b = kernel_bytes(dom, E, N)
    achieved = b / (per[dom] * 1e-3) / 1e9
    kern = {KERNEL_NAMES[k]: {"ms": round(float(per[k]), 4),
                              "GBs": round(kernel_bytes(k, E, N) / (per[k] * 1e-3) / 1e9, 1)
                              if per[k] > 0 and kernel_bytes(k, E, N) else None}
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
                   "elements": E, "nodes": N, "graph_cycles_per_launch": chunk},
        "roofline": {"bound": "hbm", "kernel": KERNEL_NAMES[dom], "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "algorithmic_bytes": b},
        "kernels": kern,
    }
    g.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(nx)
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=100)
    ap.add_argument("--chunk", type=int, default=50, help="cycles per captured graph launch")
    ap.add_argument("--mode", default="plain", choices=["plain", "copy"])
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
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
    line = bench_plain(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
