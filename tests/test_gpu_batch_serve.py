"""Batched ensemble and continuous-batching server: every problem's result
is the same alone and in the mix, bit for bit -- whatever the other members,
tile shapes, column chunks, admission order or slot and HBM limits."""
import numpy as np
import pytest

from conftest import assert_bitwise

pytestmark = pytest.mark.gpu

FIELDS = ["x", "y", "z", "xd", "yd", "zd", "e", "p", "q", "v", "ss"]

# (s, regions, balance, cost): uneven sizes and region mixes
MIX = [(6, 11, 1, 1), (9, 4, 2, 2), (12, 1, 1, 1), (16, 16, 4, 2), (7, 7, 1, 2), (10, 11, 2, 1)]


def plain(amd, s, nr, b, c, cycles=-1):
    g = amd.Group(s, 1, nr, b, c)
    if cycles >= 0:
        g.set_max_cycles(cycles)
    assert g.run() == 0
    return g


def test_batch_full_runs_match_plain_runs(amd, oracle_mod):
    print("case: batch of 6 problems to stoptime", flush=True)
    nx, nr, bal, cost = (list(v) for v in zip(*MIX))
    b = amd.Batch(nx, nr, bal, cost)
    assert b.run() == 0
    cycles, status = b.results()
    assert (status == 0).all()
    for i, (s, r, ba, c) in enumerate(MIX):
        g = plain(amd, s, r, ba, c)
        assert cycles[i] == g.sync()[1]
        m = b.group(i)
        for f in FIELDS:
            assert_bitwise(m.field(f), g.field(f), f"problem {i} (s={s}) {f}")
        g.close()
    # and the smallest against the oracle
    o = oracle_mod.Oracle(6, cbrt_mode=oracle_mod.CBRT_EXACT)
    assert o.run() == 0
    assert_bitwise(b.group(0).field("e"), o.field("e"), "problem 0 vs oracle e")


def test_batch_tiny_problems_vs_oracle(amd, oracle_mod):
    # problems of 1, 2 and 3 elements per edge batched with a larger one
    # (tiles far smaller than a workgroup, 1 to 23 cycles to stoptime while
    # the larger one keeps stepping): each ends in the state the oracle
    # reaches alone, bit for bit
    print("case: batch of tiny problems to stoptime", flush=True)
    mix = [(1, 11, 1, 1), (2, 11, 1, 1), (3, 1, 1, 0), (20, 11, 1, 1)]
    nx, nr, bal, cost = (list(v) for v in zip(*mix))
    b = amd.Batch(nx, nr, bal, cost)
    assert b.run() == 0
    cycles, status = b.results()
    assert (status == 0).all()
    for i, s in enumerate(nx):
        o = oracle_mod.Oracle(s, cbrt_mode=oracle_mod.CBRT_EXACT)
        assert o.run() == 0
        assert cycles[i] == o.cycle
        for f in ["x", "y", "z", "xd", "yd", "zd", "e", "p", "q", "v", "ss"]:
            assert_bitwise(b.group(i).field(f), o.field(f), f"problem {i} (s={s}) {f}")
        o.close()


def test_batch_members_rearmed_between_runs(amd):
    """Members stop at their caps.  In the first run the two small members
    stop at cycle 10 and the engine shrinks to the large one mid-run (its
    cycle count taken from a chunk's snapshot); raising a stopped member's
    cap through its group handle makes the next run rebuild over it (its
    cycle count read afresh by the run's first poll); raising the batch's
    cap resumes all three.  Every state equals a plain run to the same
    cycle."""
    print("case: batch members re-armed between runs", flush=True)
    mix = MIX[:3]                          # s 6, 9, 12
    nx, nr, bal, cost = (list(v) for v in zip(*mix))
    b = amd.Batch(nx, nr, bal, cost)
    b.set_max_cycles(10)
    b.group(2).set_max_cycles(20)
    assert b.run(4) == 0
    assert list(b.results()[0]) == [10, 10, 20]
    b.group(0).set_max_cycles(27)          # a member the engine no longer holds
    assert b.run(4) == 0
    assert list(b.results()[0]) == [27, 10, 20]
    b.set_max_cycles(33)                   # every member
    assert b.run(4) == 0
    assert list(b.results()[0]) == [33, 33, 33]
    for i, (s, r, ba, c) in enumerate(mix):
        g = plain(amd, s, r, ba, c, 33)
        for f in FIELDS:
            assert_bitwise(b.group(i).field(f), g.field(f), f"re-armed problem {i} (s={s}) {f}")
        g.close()
    b.close()


@pytest.mark.parametrize("tiles,kchunk", [((64, 128), 4), ((256, 256), 1), ((128, 64), 0)])
def test_batch_tiles_do_not_change_results(amd, tiles, kchunk):
    print(f"case: batch tiles {tiles} kchunk {kchunk}", flush=True)
    nx, nr, bal, cost = (list(v) for v in zip(*MIX[:4]))
    b = amd.Batch(nx, nr, bal, cost)
    b.set_tiles(*tiles)
    b.set_kchunk(kchunk)
    b.set_max_cycles(40)
    assert b.run() == 0
    for i, (s, r, ba, c) in enumerate(MIX[:4]):
        g = plain(amd, s, r, ba, c, 40)
        for f in FIELDS:
            assert_bitwise(b.group(i).field(f), g.field(f), f"problem {i} {f}")


def test_batch_autotune_keeps_results(amd):
    # the tuner's probe cycles are real cycles of the ensemble, run under
    # every candidate shape in turn; the problems still end bit for bit as
    # their plain runs do, and the chosen shapes on a fresh ensemble too
    print("case: batch autotune", flush=True)
    nx, nr, bal, cost = (list(v) for v in zip(*MIX))
    b = amd.Batch(nx, nr, bal, cost)
    best = b.autotune(probe_cycles=3)
    assert best["tile_elems"] in (64, 128, 256) and best["tile_nodes"] in (128, 256)
    assert best["ms_per_cycle"] > 0
    probed, _ = b.results()
    assert (probed > 0).all()
    b.set_max_cycles(150)
    assert b.run() == 0
    fresh = amd.Batch(nx, nr, bal, cost)
    fresh.set_tiles(best["tile_elems"], best["tile_nodes"])
    fresh.set_kchunk(best["kchunk"])
    fresh.set_max_cycles(150)
    assert fresh.run() == 0
    for i, (s, r, ba, c) in enumerate(MIX):
        g = plain(amd, s, r, ba, c, 150)
        for f in ["e", "x", "xd"]:
            assert_bitwise(b.group(i).field(f), g.field(f), f"tuned problem {i} {f}")
            assert_bitwise(fresh.group(i).field(f), g.field(f), f"fresh problem {i} {f}")


def _serve(amd, server, reqs, max_active):
    ids = [server.submit(s, r, b, c, t) for (s, r, b, c, t) in reqs]
    seen_active = 0
    done = 0
    for _ in range(100000):
        done += server.poll(16)
        st = server.stats()
        seen_active = max(seen_active, st["active"])
        assert st["active"] <= max_active
        if st["done"] == len(reqs):
            break
    assert server.stats()["done"] == len(reqs)
    return ids, seen_active


def test_server_results_same_alone_and_in_mix(amd):
    # arrivals spread over time, three slots: requests queue, are admitted
    # as slots free up and run beside whatever else is resident
    print("case: server, 8 requests, 3 slots", flush=True)
    reqs = [(8, 11, 1, 1, 0.0), (12, 5, 2, 1, 0.0), (6, 1, 1, 2, 0.0), (10, 11, 1, 1, 0.02),
            (14, 8, 4, 2, 0.02), (7, 3, 1, 1, 0.05), (9, 11, 2, 2, 0.05), (11, 6, 1, 1, 0.1)]
    sv = amd.Server(0, 3)
    ids, seen = _serve(amd, sv, reqs, 3)
    assert seen == 3
    for rid, (s, r, b, c, t) in zip(ids, reqs):
        res = sv.result(rid)
        g = plain(amd, s, r, b, c)
        v = g.verify()
        assert res["status"] == 0
        assert res["cycles"] == g.sync()[1]
        assert np.float64(res["origin_energy"]).view(np.int64) == np.float64(v["origin_energy"]).view(np.int64)
        assert res["total_abs_diff"] == v["total_abs_diff"]
        assert res["t_admit"] >= t - 1e-9 and res["t_done"] >= res["t_admit"]
        g.close()
    sv.close()


def test_server_admission_by_hbm_budget(amd):
    # a budget that holds one of these problems at a time: they run one
    # after another and still give the same results
    print("case: server, HBM budget for one problem", flush=True)
    reqs = [(10, 11, 1, 1, 0.0), (10, 11, 1, 1, 0.0), (9, 4, 1, 1, 0.0)]
    sv = amd.Server(512 << 10, 8)   # s=10: 393 KB, s=9 -r 4: 271 KB
    ids, seen = _serve(amd, sv, reqs, 8)
    assert seen == 1
    energies = [sv.result(i)["origin_energy"] for i in ids]
    assert energies[0] == energies[1]
    g = plain(amd, 9, 4, 1, 1)
    assert sv.result(ids[2])["origin_energy"] == g.verify()["origin_energy"]
    sv.close()


def test_server_rejects_bad_requests(amd):
    sv = amd.Server(0, 2)
    L = amd.load()
    assert L.lulesh_server_submit(sv.h, 0, 11, 1, 1, 0.0) == amd.EINVAL
    assert L.lulesh_server_submit(sv.h, 8, 0, 1, 1, 0.0) == amd.EINVAL
    sv.close()
