"""Several subdomains on one GPU, exchanging halos through the reference's
comm interface (device_local backend), against the oracle's loopback
exchange.  Halos are copies and ordered sums, so they must match bit for
bit; whole runs match bit for bit with the oracle in exact-cbrt mode."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, assert_bitwise

pytestmark = pytest.mark.gpu

FIELDS = ["e", "p", "q", "v", "ss", "x", "y", "z", "xd", "yd", "zd"]


@pytest.mark.parametrize("nx,cycles", [(3, 5), (4, -1), (6, 40)])
def test_decomposed_run_bitwise(amd, oracle_mod, nx, cycles):
    o = oracle_mod.Oracle(nx, tp=2, cbrt_mode=oracle_mod.CBRT_EXACT)
    assert o.run(cycles) == 0
    g = amd.Group(nx, tp=2)
    if cycles >= 0:
        g.set_max_cycles(cycles)
    assert g.run() == 0
    assert g.sync()[1] == o.cycle
    for r in range(8):
        for f in FIELDS:
            assert_bitwise(g.field(f, r), o.field(f, r), f"nx={nx} rank {r} {f}")


def test_tp3_short_run_bitwise(amd, oracle_mod):
    o = oracle_mod.Oracle(3, tp=3, cbrt_mode=oracle_mod.CBRT_EXACT)
    o.run(8)
    g = amd.Group(3, tp=3)
    g.set_max_cycles(8)
    assert g.run() == 0
    for r in range(27):
        for f in ["e", "x", "xd"]:
            assert_bitwise(g.field(f, r), o.field(f, r), f"tp3 rank {r} {f}")


def _seed(g, o, rng, fields):
    for r in range(g.ndom):
        for f in fields:
            n = g.domains[r]._len(amd_fid(f))
            v = rng.standard_normal(n)
            g.domains[r].set_field(f, v)
            o.set_field(f, v, r)


def amd_fid(name):
    import lulesh_amd
    return lulesh_amd.FIELDS[name]


@pytest.mark.parametrize("msg,fields", [
    (1024, ["fx", "fy", "fz"]),
    (2048, ["x", "y", "z", "xd", "yd", "zd"]),
    (3072, ["delv_xi", "delv_eta", "delv_zeta"]),
])
def test_comm_api_halo_bitwise(amd, oracle_mod, msg, fields):
    nx, tp = 4, 2
    g = amd.Group(nx, tp=tp)
    o = oracle_mod.Oracle(nx, tp=tp, cbrt_mode=oracle_mod.CBRT_EXACT)
    rng = np.random.default_rng(msg)
    _seed(g, o, rng, fields)
    plane_only = msg == 3072
    do_flag = msg != 2048
    dims = nx if plane_only else nx + 1
    # the reference's call sequence: CommRecv + CommSend on every rank, then
    # the unpack on every rank
    for d in g.domains:
        assert d.comm_recv(msg, len(fields), dims, dims, dims, do_flag, plane_only) == 0
    for d in g.domains:
        assert d.comm_send(msg, fields, dims, dims, dims, do_flag, plane_only) == 0
    for d in g.domains:
        if msg == 1024:
            assert d.comm_sbn(fields) == 0
        elif msg == 2048:
            assert d.comm_sync_pos_vel() == 0
        else:
            assert d.comm_monoq() == 0
    assert o.exchange(msg, fields) == 0
    for r in range(8):
        for f in fields:
            assert_bitwise(g.field(f, r), o.field(f, r), f"msg {msg} rank {r} {f}")
        assert_bitwise(g.domains[r].recv_buffer(), o.recv_buffer(r), f"msg {msg} recv buffer {r}")


def test_group_exchange_matches_per_rank_calls(amd):
    g1 = amd.Group(3, tp=2)
    g2 = amd.Group(3, tp=2)
    rng = np.random.default_rng(5)
    for r in range(8):
        for f in ["fx", "fy", "fz"]:
            v = rng.standard_normal(64)
            g1.domains[r].set_field(f, v)
            g2.domains[r].set_field(f, v)
    assert g1.exchange(1024, ["fx", "fy", "fz"]) == 0
    for d in g2.domains:
        d.comm_send(1024, ["fx", "fy", "fz"], 4, 4, 4, True, False)
    for d in g2.domains:
        d.comm_sbn(["fx", "fy", "fz"])
    for r in range(8):
        assert_bitwise(g1.field("fx", r), g2.field("fx", r), "fx")


def test_recv_layout_matches_reference(amd, oracle_mod):
    g = amd.Group(5, tp=3)
    o = oracle_mod.Oracle(5, tp=3)
    dirs = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)
            if (a, b, c) != (0, 0, 0)]
    for r in range(27):
        assert g.domains[r].com_buf_size() == o.com_buf_size(r)
        for d in dirs:
            for xfer, dr, po in [(3, 1, 0), (6, 0, 0), (3, 1, 1)]:
                assert g.domains[r].recv_offset(*d, xfer, dr, po) == o.recv_offset(r, *d, xfer, dr, po)


def test_comm_api_rejects_bad_arguments(amd):
    g = amd.Group(4, tp=2)
    d = g.domains[0]
    assert d.comm_send(1024, ["fx"], 4, 4, 4, True, False) == amd.EINVAL   # dims are nx+1
    assert d.comm_recv(3072, 3, 4, 4, 4, True, False) == amd.EINVAL       # MonoQ is planeOnly
    assert d.comm_recv(999, 3, 5, 5, 5, True, False) == amd.EINVAL


def test_s45_8_subdomains_200_cycles(amd):
    path = os.path.join(GOLDEN, "multidomain_s45.json")
    if not os.path.exists(path) or "nx45_its200_cbrt0" not in json.load(open(path)):
        pytest.skip("multidomain fixture not generated")
    rec = json.load(open(path))["nx45_its200_cbrt0"]
    g = amd.Group(45, tp=2)
    g.set_max_cycles(200)
    assert g.run() == 0
    assert g.sync()[1] == 200
    e0 = g.field("e", 0)[0]
    want = float.fromhex(rec["origin_energy"])
    assert abs(e0 - want) / want <= 1e-10
    assert "%12.6e" % e0 == "2.188600e+07"   # ucx_lulesh_experiments_46508.csv:11


def test_tp4_short_run_bitwise(amd, oracle_mod):
    # 64 subdomains, the largest decomposition the library takes
    o = oracle_mod.Oracle(3, tp=4, cbrt_mode=oracle_mod.CBRT_EXACT)
    o.run(6)
    g = amd.Group(3, tp=4)
    g.set_max_cycles(6)
    assert g.run() == 0
    for r in range(64):
        for f in ["e", "p", "x", "xd", "delv_zeta"]:
            assert_bitwise(g.field(f, r), o.field(f, r), f"tp4 rank {r} {f}")
