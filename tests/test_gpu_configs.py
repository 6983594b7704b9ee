"""The benchmark configurations at their own workloads.

* Eight subdomains (2x2x2 of -s 45) on one GPU, through both halo paths
  (the in-place direct steps of device_local, and the pack/copy/unpack of
  the device_copy backend): after 200 cycles and at the stop time every
  field of every subdomain equals the oracle's (exact cube root) bit for
  bit, and rank 0's origin energy is within 1e-10 of the reference
  arithmetic's (libm cube root), printing as the reference's published
  8-rank results do: 2.188600e+07 after 200 cycles
  (ucx_lulesh_experiments_46508.csv:11) and 1.482403e+06 at the stop time,
  3145 cycles (results/lulesh_results.csv:12,19-27).  Fixtures:
  tests/golden/multidomain_s45.json (make_multidomain_golden.py).
* The 64-problem batched ensemble of the benchmark (seeded mix, s in
  16..64, r in 1..16, b in {1,2,4}, c in {1,2}), autotuned, to stoptime:
  every member's final state is bit-identical to the plain run of its size
  (results do not depend on the region layout), and the smallest sizes to
  the oracle's.
* The served stream (2,000 seeded requests, s in 20..80, r in 1..11) through
  the continuous-batching server in four slices of 500 arriving at once:
  every result's cycle count, origin energy and symmetry figure is
  bit-identical to the plain run of its size, and the smallest sizes to the
  oracle's.
"""
import json
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, REPO, assert_bitwise

sys.path.insert(0, REPO)
from bench import problem_mix  # noqa: E402

pytestmark = pytest.mark.gpu

FIELDS = ["e", "p", "q", "v", "ss", "x", "y", "z", "xd", "yd", "zd"]
MD = json.load(open(os.path.join(GOLDEN, "multidomain_s45.json")))


def sha(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def check_md(g, key, printed):
    exact = MD.get(f"{key}_cbrt1")
    libm = MD.get(f"{key}_cbrt0")
    if exact is None or libm is None:
        pytest.skip(f"fixture {key} not generated")
    rc, cyc, t, _ = g.sync()
    assert rc == 0 and cyc == exact["cycle"] == libm["cycle"]
    assert t == float.fromhex(exact["time"])
    for r in range(8):
        for f in FIELDS:
            assert sha(g.field(f, r)) == exact["sha256"][f"{f}{r}"], f"{key} rank {r} {f}"
    e0 = g.field("e", 0)[0]
    want = float.fromhex(libm["origin_energy"])
    assert abs(e0 - want) / abs(want) <= 1e-10
    assert "%12.6e" % e0 == printed
    v = g.verify(0)
    assert abs(v["total_abs_diff"] - float.fromhex(libm["verify"]["total_abs_diff"])) <= 1e-8


@pytest.mark.parametrize("backend", ["device_local", "device_copy"])
def test_8_subdomains_s45_200_cycles(amd, backend):
    g = amd.Group(45, tp=2)
    if backend != "device_local":
        g.set_comm_backend(backend)
    g.set_max_cycles(200)
    assert g.run() == 0
    check_md(g, "nx45_its200", "2.188600e+07")   # ucx_lulesh_experiments_46508.csv:11
    g.close()


@pytest.mark.parametrize("backend", ["device_local", "device_copy"])
def test_8_subdomains_s45_to_stoptime(amd, backend):
    g = amd.Group(45, tp=2)
    if backend != "device_local":
        g.set_comm_backend(backend)
    assert g.run(200) == 0
    check_md(g, "nx45_its-1", "1.482403e+06")   # results/lulesh_results.csv:12,19-27
    g.close()


_plain = {}


def plain(amd, s):
    """cycles, verify figures and final fields of the plain -s s run"""
    if s not in _plain:
        g = amd.Group(int(s))
        assert g.run() == 0
        _plain[s] = (g.sync()[1], g.verify(), {f: g.field(f) for f in ("e", "x", "xd", "v")})
        g.close()
    return _plain[s]


def oracle_run(oracle_mod, s):
    o = oracle_mod.Oracle(int(s), cbrt_mode=oracle_mod.CBRT_EXACT)
    assert o.run() == 0
    return o


def test_batched_64_problem_ensemble(amd, oracle_mod):
    mix = problem_mix(64, 2024, 16, 64, 1, 16, (1, 2, 4), (1, 2))
    tb = amd.Batch(*mix)
    tune = tb.autotune(5)
    tb.close()
    b = amd.Batch(*mix)
    b.set_tiles(tune["tile_elems"], tune["tile_nodes"])
    b.set_kchunk(tune["kchunk"])
    assert b.run(50) == 0
    cyc, st = b.results()
    assert np.all(st == 0)
    nx = mix[0]
    for i in range(64):
        pc, pv, pf = plain(amd, int(nx[i]))
        assert cyc[i] == pc, f"problem {i} (-s {nx[i]})"
        g = b.group(i)
        v = g.verify()
        assert v["origin_energy"] == pv["origin_energy"] and v["total_abs_diff"] == pv["total_abs_diff"]
        for f, a in pf.items():
            assert_bitwise(g.field(f), a, f"problem {i} (-s {nx[i]}) {f}")
    # the smallest members against the oracle (exact cube root)
    for s in sorted(set(int(v) for v in nx))[:2]:
        i = int(np.nonzero(nx == s)[0][0])
        o = oracle_run(oracle_mod, s)
        assert cyc[i] == o.cycle
        for f in FIELDS:
            assert_bitwise(b.group(i).field(f), o.field(f), f"batched -s {s} vs oracle {f}")
    b.close()


SERVE_MIX = problem_mix(2000, 2024, 20, 80, 1, 11)


@pytest.mark.parametrize("part", range(4))
def test_served_stream_2000_requests(amd, oracle_mod, part):
    nx, nr = SERVE_MIX[0][500 * part:500 * (part + 1)], SERVE_MIX[1][500 * part:500 * (part + 1)]
    print(f"served requests {500 * part}..{500 * part + 499}", flush=True)
    srv = amd.Server(0, 64)
    t0 = srv.now()
    ids = [srv.submit(int(nx[i]), int(nr[i]), 1, 1, t0) for i in range(len(nx))]
    done = 0
    while done < len(ids):
        done += srv.poll(50)
    for i, rid in enumerate(ids):
        r = srv.result(rid)
        pc, pv, _ = plain(amd, int(nx[i]))
        assert r["status"] == 0
        assert r["cycles"] == pc, f"request {500 * part + i} (-s {nx[i]})"
        assert r["origin_energy"] == pv["origin_energy"], f"request {500 * part + i}"
        assert r["total_abs_diff"] == pv["total_abs_diff"], f"request {500 * part + i}"
    if part == 0:
        for s in sorted(set(int(v) for v in nx))[:1]:
            o = oracle_run(oracle_mod, s)
            ov = o.verify()
            i = int(np.nonzero(nx == s)[0][0])
            r = srv.result(ids[i])
            assert r["cycles"] == o.cycle
            assert r["origin_energy"] == ov["origin_energy"]
            assert r["total_abs_diff"] == ov["total_abs_diff"]
    srv.close()
