"""Unstructured-mesh (.lmesh) path on the GPU against the oracle: the
reference's own sedov15oct.lmesh to stoptime, its sedov21oct.lmesh (31 408
elements) after 100 cycles and at the stop time against committed oracle
digests, the block written as an unstructured mesh against the structured
GPU path, and a renumbered mesh (elements and nodes shuffled) -- all bit
for bit."""
import gzip
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, assert_bitwise
import umesh_util as U

pytestmark = pytest.mark.gpu

LMESH = os.path.join(GOLDEN, U.GOLDEN_LMESH)
U21 = json.load(open(os.path.join(GOLDEN, "umesh_sedov21oct.json")))
NODE = ["x", "y", "z", "xd", "yd", "zd", "nodalMass"]
ELEM = ["e", "p", "q", "v", "ss", "volo", "elemMass", "delv", "vdov", "arealg", "vnew",
        "delv_xi", "delv_eta", "delv_zeta", "delx_xi", "delx_eta", "delx_zeta"]


def compare(g, o, what, fields=NODE + ELEM):
    for f in fields:
        assert_bitwise(g.field(f), o.field(f), f"{what} {f}")


def test_sedov15oct_to_stoptime(amd, oracle_mod):
    print("case: sedov15oct.lmesh full run", flush=True)
    m = amd.Mesh.read(LMESH)
    g = amd.Group.from_mesh(m)
    m.close()
    o = oracle_mod.Oracle.from_mesh(oracle_mod.read_lmesh(LMESH), cbrt_mode=oracle_mod.CBRT_EXACT)
    assert g.run() == 0 and o.run() == 0
    rc, cyc, t, dt = g.sync()
    assert (cyc, t, dt) == (o.cycle, o.time, o.deltatime)
    compare(g, o, "sedov15oct")
    v = g.verify()
    assert v["origin_energy"] == o.field("e")[o.octant_corner]
    assert v["max_abs_diff"] == 0.0


def _check21(g, rec, what):
    rc, cyc, t, dt = g.sync()
    assert rc == 0 and cyc == rec["cycle"], what
    assert t == float.fromhex(rec["time"]) and dt == float.fromhex(rec["deltatime"]), what
    for f, h in rec["sha256"].items():
        got = hashlib.sha256(np.ascontiguousarray(g.field(f), dtype="<f8").tobytes()).hexdigest()
        assert got == h, f"{what} {f}"
    v = g.verify()
    for k, h in rec["verify"].items():
        assert v[k] == float.fromhex(h), f"{what} {k}"


def test_sedov21oct_to_stoptime(amd, tmp_path):
    # digests of the oracle (exact cube root) on the same mesh:
    # tests/golden/make_umesh_golden.py
    print("case: sedov21oct.lmesh 100 cycles, then to stoptime", flush=True)
    path = tmp_path / "sedov21oct.lmesh"
    path.write_bytes(gzip.decompress(open(os.path.join(GOLDEN, "sedov21oct.lmesh.gz"), "rb").read()))
    m = amd.Mesh.read(str(path))
    assert (m.num_elem, m.num_node) == (31408, 34446)
    g = amd.Group.from_mesh(m)
    m.close()
    g.set_max_cycles(100)
    assert g.run() == 0
    _check21(g, U21["its100_cbrt1"], "sedov21oct 100 cycles")
    g.set_max_cycles(-1)
    assert g.run() == 0
    _check21(g, U21["its-1_cbrt1"], "sedov21oct full run")
    # the reference's arithmetic (libm cube root): origin energy within 1e-10
    want = float.fromhex(U21["its-1_cbrt0"]["origin_energy"])
    assert abs(g.verify()["origin_energy"] - want) <= 1e-10 * abs(want)
    g.close()


def test_sedov15oct_regions_and_forces(amd, oracle_mod):
    print("case: sedov15oct.lmesh -r 5 -b 2 -c 2, keep forces, 60 cycles", flush=True)
    g = amd.Group.from_mesh(amd.Mesh.read(LMESH), nr=5, balance=2, cost=2)
    g.set_keep_forces(True)
    g.set_max_cycles(60)
    o = oracle_mod.Oracle.from_mesh(oracle_mod.read_lmesh(LMESH), nr=5, balance=2, cost=2,
                                    cbrt_mode=oracle_mod.CBRT_EXACT)
    assert np.array_equal(g.domains[0].elem_rep(), o.elem_rep())
    assert g.run() == 0 and o.run(60) == 0
    compare(g, o, "sedov15oct r5", NODE + ELEM + ["fx", "fy", "fz"])


@pytest.mark.parametrize("nx", [12, 30])
def test_block_as_umesh_matches_structured(amd, nx):
    # the block through the unstructured kernels (explicit connectivity,
    # corner lists) with the block's own elemBC and deposit gives the
    # structured kernels' result bit for bit
    print(f"case: block s={nx} as an unstructured mesh", flush=True)
    import pyoracle
    ref = amd.Group(nx)
    bc = pyoracle.Oracle(nx).elem_bc()
    m = amd.Mesh.from_arrays(U.structured_mesh(nx))
    m.set_elem_bc(bc)
    m.set_energy(U.einit(nx))
    g = amd.Group.from_mesh(m)
    cycles = -1 if nx <= 12 else 300
    for x in (ref, g):
        if cycles > 0:
            x.set_max_cycles(cycles)
        assert x.run() == 0
    assert g.sync()[1:] == ref.sync()[1:]
    ne = nx ** 3
    for f in NODE + ELEM:
        a, b = g.field(f), ref.field(f)
        if f.startswith("delv_"):
            a, b = a[:ne], b[:ne]
        assert_bitwise(a, b, f"s={nx} {f}")


def test_renumbered_mesh(amd, oracle_mod):
    # a genuinely unstructured numbering: the block with elements and nodes
    # shuffled (sum orders follow the new element order, on both sides)
    print("case: shuffled s=20 block, 150 cycles", flush=True)
    nx = 20
    rng = np.random.default_rng(11)
    m = U.structured_mesh(nx)
    m = U.renumber_elems(m, rng.permutation(m["num_elem"]))
    m = U.renumber_nodes(m, rng.permutation(m["num_node"]))
    g = amd.Group.from_mesh(amd.Mesh.from_arrays(m))
    g.set_max_cycles(150)
    o = oracle_mod.Oracle.from_mesh(m, cbrt_mode=oracle_mod.CBRT_EXACT)
    assert g.run() == 0 and o.run(150) == 0
    assert g.sync()[1] == o.cycle == 150
    compare(g, o, "shuffled s=20")


def test_umesh_group_rejects_comm(amd, tmp_path):
    print("case: comm calls on an unstructured group", flush=True)
    g = amd.Group.from_mesh(amd.Mesh.read(LMESH))
    d = g.domains[0]
    assert (d.nx, d.num_elem, d.num_node, d.all_elem) == (0, 875, 1126, 875)
    assert d.comm_recv(amd.MSG_COMM_SBN, 3, 1, 1, 1, True, False) == amd.EINVAL
    assert d.comm_send(amd.MSG_COMM_SBN, ["fx"], 1, 1, 1, True, False) == amd.EINVAL
    assert d.comm_sbn(["fx"]) == amd.EINVAL
    assert d.comm_sync_pos_vel() == amd.EINVAL
    assert d.comm_monoq() == amd.EINVAL
    assert g.exchange(amd.MSG_COMM_SBN, ["fx"]) == amd.EINVAL
    g.set_max_cycles(5)
    assert g.run() == 0
    path = tmp_path / "xyz.asc"
    g.write_solution(path)
    vals = open(path).read().split()
    assert len(vals) == 4 * 1126
    x = g.field("x")
    assert float(vals[4 * 7 + 1]) == pytest.approx(x[7], abs=1e-10)


def test_driver_unstructured(oracle_mod, tmp_path):
    import subprocess
    from conftest import REPO
    print("case: driver -u sedov15oct.lmesh", flush=True)
    driver = os.path.join(REPO, "cuda-ipc-benchmarks_amd", "lulesh_amd_run")
    out = subprocess.run([driver, "-u", LMESH], cwd=tmp_path, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    o = oracle_mod.Oracle.from_mesh(oracle_mod.read_lmesh(LMESH), cbrt_mode=oracle_mod.CBRT_EXACT)
    assert o.run() == 0
    lines = out.stdout.splitlines()
    assert lines[0].startswith("Running until t=0.010000, Problem size=875")
    assert f"   Iteration count     =  {o.cycle} " in lines
    assert f"FullPrecision OriginEnergy = {o.field('e')[o.octant_corner]:.17e}" in lines
    assert (tmp_path / "xyz.asc").stat().st_size > 0
    bad = subprocess.run([driver, "-u", str(tmp_path / "none.lmesh")], cwd=tmp_path,
                         capture_output=True, text=True, timeout=60)
    assert bad.returncode == 1 and "could not read" in bad.stderr
