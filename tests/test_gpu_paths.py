"""Every stepping path of the library gives the oracle's state bit for bit:
the fused force + node-update march (default for one-subdomain groups),
which also computes the cycle's time increment, the same with a separate
time-increment launch, the same with a separate node update, the two-pass
corner-array path, other z-chunk depths, eager launches instead of graphs, other EOS tile
sizes (the s=12 default is 64 elements per tile); node-state buffer
alternation across separate run calls and host field writes; the forces
kept on request."""
import numpy as np
import pytest

from conftest import assert_bitwise

pytestmark = pytest.mark.gpu

NODE = ["x", "y", "z", "xd", "yd", "zd"]
ELEM = ["e", "p", "q", "v", "ss", "delv", "vdov", "arealg", "vnew"]


def _compare(g, o, what):
    for f in NODE + ELEM:
        assert_bitwise(g.field(f), o.field(f), f"{what} {f}")


@pytest.mark.parametrize("env", [
    {"LULESH_NO_FUSED_UPDATE": "1"},
    {"LULESH_FORCE_PATH": "twopass"},
    {"LULESH_KCHUNK": "1"},
    {"LULESH_KCHUNK": "7"},
    {"LULESH_NO_GRAPH": "1"},
    {"LULESH_QEOS_TILE": "128"},
    {"LULESH_QEOS_TILE": "256"},
    {"LULESH_NO_FUSED_TI": "1"},
    {"LULESH_NO_FUSED_TI": "1", "LULESH_NO_GRAPH": "1"},
])
def test_paths_bitwise(amd, oracle_mod, monkeypatch, env):
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    o = oracle_mod.Oracle(12, cbrt_mode=oracle_mod.CBRT_EXACT)
    assert o.run() == 0
    g = amd.Group(12)
    assert g.run() == 0
    _compare(g, o, str(env))
    g.close()
    o.close()


def test_buffers_across_runs_and_field_writes(amd, oracle_mod):
    o = oracle_mod.Oracle(10, cbrt_mode=oracle_mod.CBRT_EXACT)
    g = amd.Group(10)
    o.run(7)
    g.set_max_cycles(7)
    assert g.run() == 0
    _compare(g, o, "after 7 cycles")
    # perturb the velocities on an odd cycle: the write must land in the
    # buffer the next cycle starts from
    for f in ["xd", "zd"]:
        v = o.field(f) * 1.5
        o.set_field(f, v)
        g.domains[0].set_field(f, v)
    o.run(6)
    g.set_max_cycles(13)
    assert g.run() == 0
    _compare(g, o, "after 13 cycles")


def test_keep_forces(amd, oracle_mod):
    o = oracle_mod.Oracle(16, cbrt_mode=oracle_mod.CBRT_EXACT)
    o.run(25)
    g = amd.Group(16)
    g.set_keep_forces(True)
    g.set_max_cycles(25)
    assert g.run() == 0
    for f in ["fx", "fy", "fz"]:
        assert_bitwise(g.field(f), o.field(f), f)
    _compare(g, o, "keep_forces")


def test_forces_not_kept_rejected(amd):
    g = amd.Group(8)
    g.set_max_cycles(3)
    assert g.run() == 0
    with pytest.raises(amd.LuleshError):
        g.field("fx")


@pytest.mark.parametrize("env", [{}, {"LULESH_NO_DIRECT": "1"}, {"LULESH_NO_DIRECT_POSVEL": "1"},
                                 {"LULESH_NO_FUSED_TI": "1"},
                                 {"LULESH_FORCE_PATH": "twopass"}])
def test_decomposed_paths_bitwise(amd, oracle_mod, monkeypatch, env):
    """2x2x2 subdomains: the direct halo steps (CommSBN summed in the node
    update, CommSyncPosVel written by each shared node's final writer,
    CommMonoQ written by the kinematics kernel), the pack/unpack exchanges,
    and the two-pass force path all give the oracle's state."""
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    o = oracle_mod.Oracle(5, tp=2, cbrt_mode=oracle_mod.CBRT_EXACT)
    assert o.run(30) == 0
    g = amd.Group(5, tp=2)
    g.set_max_cycles(30)
    assert g.run() == 0
    for r in range(8):
        for f in NODE + ["e", "p", "q", "v", "ss", "delv_xi", "delv_eta", "delv_zeta"]:
            assert_bitwise(g.field(f, r), o.field(f, r), f"{env} rank {r} {f}")
    g.set_keep_forces(True)
    g.set_max_cycles(31)
    o.run(1)
    assert g.run() == 0
    for r in range(8):
        for f in ["fx", "fy", "fz"]:
            assert_bitwise(g.field(f, r), o.field(f, r), f"{env} rank {r} {f}")


@pytest.mark.parametrize("env", [{}, {"LULESH_NO_DIRECT_POSVEL": "1"}])
def test_decomposed_inconsistent_copies(amd, oracle_mod, monkeypatch, env):
    """Host writes that leave the copies of shared nodes different (one
    subdomain's velocities scaled): CommSyncPosVel's final writers then
    carry values the other copies did not compute, and every copy must
    still end as the exchange leaves it."""
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    o = oracle_mod.Oracle(5, tp=2, cbrt_mode=oracle_mod.CBRT_EXACT)
    g = amd.Group(5, tp=2)
    o.run(40)
    g.set_max_cycles(40)
    assert g.run() == 0
    for r in (0, 5):
        for f in ["xd", "y"]:
            v = o.field(f, r) * (1.0 + 1e-9 * (r + 1))
            o.set_field(f, v, r)
            g.domains[r].set_field(f, v)
    o.run(25)
    g.set_max_cycles(65)
    assert g.run() == 0
    for r in range(8):
        for f in NODE + ["e", "p", "q", "v"]:
            assert_bitwise(g.field(f, r), o.field(f, r), f"{env} rank {r} {f}")


def test_paired_runs_bitwise(amd, oracle_mod, monkeypatch):
    """The force kernel's column runs are split by the paired plan (longer
    runs for the first workgroup of each CU) where it is faster, by the
    even split otherwise: the split decides which workgroup completes which
    node planes, never the order of any sum, so both give the oracle's
    state bit for bit."""
    o = oracle_mod.Oracle(64, cbrt_mode=oracle_mod.CBRT_EXACT)
    o.run(25)
    states = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("LULESH_PAIRED_RUNS", flag)
        g = amd.Group(64)
        wg, longest, paired = g.force_plan()
        assert paired == (flag == "1"), (wg, longest, paired)
        g.set_max_cycles(25)
        assert g.run() == 0
        _compare(g, o, f"paired={flag}")
        states[flag] = (wg, longest)
        g.close()
    o.close()
    assert states["1"] != states["0"]


@pytest.mark.parametrize("nx,tp", [(8, 1), (30, 1), (45, 2), (48, 1), (64, 1), (100, 1), (150, 1)])
def test_engine_plan_is_the_host_plan(amd, nx, tp):
    """The engine's force-kernel launch is the plan the host-only planner
    gives for this device's CU count (whose every-plane-once invariants the
    CPU tests check over all sizes): same workgroup count, longest run and
    choice of the paired plan."""
    kc, paired, t = amd.force_plan_of(nx, tp, amd.device_cus())
    g = amd.Group(nx, tp)
    assert g.force_plan() == (len(t), int(t[:, 7].max()), paired)
    g.close()


def test_paired_runs_s100_against_even_split(amd, monkeypatch):
    """-s 100 (the headline size) on the paired plan and on the even split:
    every node and element field bit-identical after 60 cycles."""
    out = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("LULESH_PAIRED_RUNS", flag)
        g = amd.Group(100)
        assert g.force_plan()[2] == (flag == "1")
        g.set_max_cycles(60)
        assert g.run() == 0
        out[flag] = {f: g.field(f).copy() for f in NODE + ELEM}
        g.close()
    for f in NODE + ELEM:
        assert_bitwise(out["1"][f], out["0"][f], f"s100 paired vs even {f}")
