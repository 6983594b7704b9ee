"""Unstructured-mesh (.lmesh) input on the CPU: the oracle's restatement of
the reference's reader and boundary derivation (LULESH/cuda/src/lulesh.cu:
855-1052), the library's host-side reader against it, and the properties
that pin the unstructured step (no GPU needed).

The reference's unstructured path runs only in its CUDA build, which cannot
run here; the step itself is the oracle's structured step, pinned bit for bit
against the compiled omp_4.0 reference, so the unstructured form is pinned by
(1) hand-derived boundary bits of the reference's own mesh file, (2) the
block written as an unstructured mesh reproducing the structured run bit for
bit, and (3) invariance under node renumbering.
"""
import gzip
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, REPO, assert_bitwise
import umesh_util as U

LMESH = os.path.join(GOLDEN, U.GOLDEN_LMESH)
REF_LMESH21 = "/root/reference/LULESH/cuda/src/sedov21oct.lmesh"
FIX21 = os.path.join(GOLDEN, "sedov21oct.lmesh.gz")
U21 = json.load(open(os.path.join(GOLDEN, "umesh_sedov21oct.json")))
FIELDS = ["x", "y", "z", "xd", "yd", "zd", "e", "p", "q", "v", "ss", "volo", "elemMass",
          "nodalMass", "delv", "vdov", "arealg", "delv_xi", "delv_eta", "delv_zeta"]


def test_reader_fixture(oracle_mod):
    m = oracle_mod.read_lmesh(LMESH)
    assert (m["num_elem"], m["num_node"]) == (875, 1126)
    assert [len(s) for s in m["symm"]] == [146, 146, 146] and len(m["free"]) == 91
    assert m["nodelist"][0].tolist() == [0, 5, 6, 7, 1, 2, 3, 4]
    assert m["nbr"][0].tolist() == [0, 5, 0, 7, 0, 1]
    assert (m["x"][1], m["y"][1], m["z"][1]) == (0.0, 0.0, 0.08)
    assert m["free"][-1] == 1125


def test_boundary_bits_hand_derived(oracle_mod):
    # element 0 = nodes (0,5,6,7,1,2,3,4); on a symmetry plane (x, y or z
    # = 0) are all but node 3 (0.071149, 0.071149, 0.071149), its local
    # corner 6.  Mask 0xbf: corners {0..3} (zeta-), {0,1,4,5} (eta-) and
    # {0,3,4,7} (xi-) complete; {4..7}, {2,3,6,7}, {1,2,5,6} miss corner 6.
    # None of its nodes is on the free surface.
    m = oracle_mod.read_lmesh(LMESH)
    assert not np.isin(m["nodelist"][0], m["free"]).any()
    o = oracle_mod.Oracle.from_mesh(m)
    bc = o.elem_bc()
    assert bc[0] == U.XI_M_SYMM | U.ETA_M_SYMM | U.ZETA_M_SYMM
    assert o.octant_corner == 0
    # an element with a zeta+ face on the free surface gets ZETA_P_SYMM
    onfree = np.isin(m["nodelist"], m["free"])
    zp = np.nonzero(onfree[:, 4:].all(axis=1))[0]
    assert zp.size > 0 and (bc[zp] & U.ZETA_P_SYMM).all()
    # no FREE or COMM bits are ever derived
    assert not (bc & (U.FREE_BITS | 0x24924)).any()


def test_library_reader_matches_oracle(oracle_mod):
    import lulesh_amd as la
    la.load()
    pm = oracle_mod.read_lmesh(LMESH)
    o = oracle_mod.Oracle.from_mesh(pm)
    m = la.Mesh.read(LMESH)
    assert (m.num_elem, m.num_node) == (875, 1126)
    assert m.octant_corner == o.octant_corner
    assert np.array_equal(m.elem_bc(), o.elem_bc())
    m2 = la.Mesh.from_arrays(pm)
    assert np.array_equal(m2.elem_bc(), o.elem_bc())


@pytest.mark.skipif(not os.path.exists(REF_LMESH21), reason="reference tree not present")
def test_library_reader_sedov21oct(oracle_mod):
    import lulesh_amd as la
    la.load()
    pm = oracle_mod.read_lmesh(REF_LMESH21)
    o = oracle_mod.Oracle.from_mesh(pm)
    m = la.Mesh.read(REF_LMESH21)
    assert (m.num_elem, m.num_node) == (31408, 34446)
    assert m.octant_corner == o.octant_corner
    assert np.array_equal(m.elem_bc(), o.elem_bc())


def test_library_reader_errors(tmp_path):
    import lulesh_amd as la
    L = la.load()
    import ctypes
    h = ctypes.c_void_p()
    assert L.lulesh_mesh_read(str(tmp_path / "missing.lmesh").encode(), ctypes.byref(h)) == la.EIO
    text = open(LMESH).read()
    short = tmp_path / "short.lmesh"
    short.write_text(text[: len(text) // 2])
    assert L.lulesh_mesh_read(str(short).encode(), ctypes.byref(h)) == la.EIO
    bad = tmp_path / "bad.lmesh"
    lines = text.split("\n")
    lines[1 + 1126] = "0 5 6 7 1 2 3 99999"          # node id out of range
    bad.write_text("\n".join(lines))
    assert L.lulesh_mesh_read(str(bad).encode(), ctypes.byref(h)) == la.EINVAL
    # element counts above 2^26 - 1 and node counts above 2^28 - 1 are
    # refused before any array is read
    one = np.zeros(8, dtype=np.float64)
    idx = np.zeros(8, dtype=np.int32)
    dp = one.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    ip = idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    for ne, nn in [(1, 1 << 28), (1 << 28, 8), (1 << 26, 8)]:
        assert L.lulesh_mesh_create(ne, nn, dp, dp, dp, ip, ip, ip, 0, ip, 0, ip, 0, ip, 0,
                                    ctypes.byref(h)) == la.EINVAL


def test_library_reader_huge_counts(tmp_path):
    """Counts in the header are checked against what the rest of the file
    can hold before any array is sized by them: a few dozen bytes claiming
    2^28 - 1 nodes, or a symmetry set of 10^9 entries, fail with EIO at
    once instead of committing gigabytes of host memory."""
    import ctypes
    import resource
    import lulesh_amd as la
    L = la.load()
    h = ctypes.c_void_p()
    cases = {
        "huge_nodes.lmesh": "1 268435455\n0 0 0\n",
        "huge_elems.lmesh": "67108863 8\n" + "0 0 0\n" * 8,
        "huge_set.lmesh": None,
        "mid_record.lmesh": None,
    }
    text = open(LMESH).read()
    tok = text.split()
    ne, nn = int(tok[0]), int(tok[1])
    head = 2 + 3 * nn + 14 * ne
    # a valid file up to the first symmetry set, which claims 10^9 entries
    cases["huge_set.lmesh"] = " ".join(tok[:head]) + " 1000000000 1 2 3\n"
    # a file cut in the middle of a node's coordinate triple
    cases["mid_record.lmesh"] = " ".join(tok[:2 + 3 * (nn // 2) + 1]) + "\n"
    before = resource.getrusage(resource.RUSAGE_SELF).ru_maxrss
    for name, body in cases.items():
        p = tmp_path / name
        p.write_text(body)
        assert L.lulesh_mesh_read(str(p).encode(), ctypes.byref(h)) == la.EIO, name
    grown_kb = resource.getrusage(resource.RUSAGE_SELF).ru_maxrss - before
    assert grown_kb < 256 * 1024   # far below the 6 GB a 2^28-node file would size


def test_region_weights_refused_at_the_boundary():
    """-r 11 -b 9 sums the region weights past 2^31 - 1 in the reference's
    int (lulesh-init.cc:384-388): group creation refuses it before touching
    the GPU, as it refuses the other bad arguments."""
    import ctypes
    import lulesh_amd as la
    L = la.load()
    h = ctypes.c_void_p()
    for nx, nr, b, c in [(8, 11, 9, 1), (8, 100, 5, 1), (8, 0, 1, 1), (8, 11, -1, 1), (8, 1 << 21, 0, 1)]:
        assert L.lulesh_group_create(nx, 1, nr, b, c, ctypes.byref(h)) == la.EINVAL
        assert L.lulesh_server_submit(None, nx, nr, b, c, 0.0) == la.EINVAL


def test_block_boundary_bits(oracle_mod):
    # the block as an unstructured mesh: the reader's bits are the block's
    # elemBC with every FREE face marked SYMM instead (the reference's
    # reader marks free-surface faces with the *_SYMM bits)
    nx = 6
    m = U.structured_mesh(nx)
    bc_block = oracle_mod.Oracle(nx).elem_bc()
    o = oracle_mod.Oracle.from_mesh(m)
    want = bc_block & ~U.FREE_BITS
    for free, symm in [(U.XI_P_FREE, U.XI_P_SYMM), (U.ETA_P_FREE, U.ETA_P_SYMM),
                       (U.ZETA_P_FREE, U.ZETA_P_SYMM)]:
        want |= np.where(bc_block & free, symm, 0)
    assert np.array_equal(o.elem_bc(), want)
    assert o.octant_corner == 0


@pytest.mark.parametrize("nx", [5, 8])
def test_block_as_umesh_matches_structured(oracle_mod, nx, tmp_path):
    # written to a file and read back, given the block's own elemBC and
    # deposit: the unstructured set-up and step reproduce the structured
    # run bit for bit (full run to stoptime)
    m = U.structured_mesh(nx)
    path = tmp_path / "block.lmesh"
    U.write_lmesh(path, m)
    m = oracle_mod.read_lmesh(str(path))
    ref = oracle_mod.Oracle(nx)
    o = oracle_mod.Oracle.from_mesh(m, elem_bc=ref.elem_bc(), e_deposit=U.einit(nx))
    assert o.deltatime == ref.deltatime
    assert ref.run() == 0 and o.run() == 0
    assert o.cycle == ref.cycle and o.time == ref.time
    ne = nx ** 3
    for f in FIELDS:   # (the block's gradient arrays carry ghost planes after ne)
        assert_bitwise(o.field(f)[:ne] if f.startswith("delv_") else o.field(f),
                       ref.field(f)[:ne] if f.startswith("delv_") else ref.field(f), f"s={nx} {f}")


def test_node_renumbering_invariance(oracle_mod):
    # renumbering nodes changes no sum order (per-node sums run over
    # elements, per-element work over the element's own corner order)
    m = oracle_mod.read_lmesh(LMESH)
    perm = np.random.default_rng(7).permutation(m["num_node"])
    a = oracle_mod.Oracle.from_mesh(m)
    b = oracle_mod.Oracle.from_mesh(U.renumber_nodes(m, perm))
    assert a.run(150) == 0 and b.run(150) == 0
    for f in ["x", "y", "z", "xd", "yd", "zd", "nodalMass"]:
        assert_bitwise(b.field(f)[perm], a.field(f), f"node field {f}")
    for f in ["e", "p", "q", "v", "ss"]:
        assert_bitwise(b.field(f), a.field(f), f"elem field {f}")


def test_regions_do_not_change_results(oracle_mod):
    # the EOS repetitions recompute the same values: -r/-b/-c change the
    # cost only (lulesh.cc:2677-2815)
    m = oracle_mod.read_lmesh(LMESH)
    a = oracle_mod.Oracle.from_mesh(m, nr=11, balance=1, cost=1)
    b = oracle_mod.Oracle.from_mesh(m, nr=5, balance=2, cost=3)
    assert not np.array_equal(a.elem_rep(), b.elem_rep())
    assert a.run(200) == 0 and b.run(200) == 0
    for f in ["e", "p", "q", "x", "xd"]:
        assert_bitwise(b.field(f), a.field(f), f)


def test_full_run_energy_is_finite(oracle_mod):
    # the deposit stays in the octant corner and the run reaches stoptime
    m = oracle_mod.read_lmesh(LMESH)
    o = oracle_mod.Oracle.from_mesh(m)
    assert o.field("e")[0] == U.E_BASE
    assert o.run() == 0
    assert abs(o.time - 1.0e-2) < 1e-15 and o.cycle > 100
    e = o.field("e")
    assert np.isfinite(e).all() and e.max() > 0
    assert o.verify()["origin_energy"] == e[0]


def test_corner_mesh_reproduces_block_corner(oracle_mod):
    # For its first cycles the blast stays near the origin: the m^3 corner
    # of the -s nx block, cut out as a mesh of its own (same coordinates and
    # deposit, free faces where the block continues), evolves bit for bit
    # as the corner of the full block, and its outer faces stay untouched.
    # A run too small for the oracle to repeat in full is checked this way.
    nx, m, cycles = 40, 12, 20
    ref = oracle_mod.Oracle(nx, cbrt_mode=oracle_mod.CBRT_EXACT)
    o = oracle_mod.Oracle.from_mesh(U.corner_mesh(nx, m), nr=1, balance=1, cost=0,
                                    cbrt_mode=oracle_mod.CBRT_EXACT,
                                    elem_bc=oracle_mod.Oracle(m).elem_bc(), e_deposit=U.einit(nx))
    assert o.deltatime == ref.deltatime
    assert ref.run(cycles) == 0 and o.run(cycles) == 0
    assert (o.cycle, o.time, o.deltatime) == (ref.cycle, ref.time, ref.deltatime)
    for f in U.CORNER_ELEM:
        assert_bitwise(o.field(f), U.corner_of(ref.field(f), nx, m), f"corner {f}")
    for f in U.CORNER_NODE:
        assert_bitwise(o.field(f), U.corner_of(ref.field(f), nx + 1, m + 1), f"corner {f}")
    e = o.field("e").reshape(m, m, m)
    assert e[0, 0, 0] > 0 and not e[-1].any() and not e[:, -1].any() and not e[:, :, -1].any()


def _unpack21(tmp_path):
    path = tmp_path / "sedov21oct.lmesh"
    path.write_bytes(gzip.decompress(open(FIX21, "rb").read()))
    return str(path)


@pytest.mark.skipif(not os.path.exists(REF_LMESH21), reason="reference tree not present")
def test_sedov21oct_fixture_is_the_reference_file(tmp_path):
    # tests/golden/sedov21oct.lmesh.gz holds the reference's mesh unchanged
    assert open(_unpack21(tmp_path), "rb").read() == open(REF_LMESH21, "rb").read()


def test_sedov21oct_oracle_100_cycles(oracle_mod, tmp_path):
    # the golden record the GPU test compares with (make_umesh_golden.py),
    # reproduced by the oracle from the committed fixture
    rec = U21["its100_cbrt1"]
    o = oracle_mod.Oracle.from_mesh(oracle_mod.read_lmesh(_unpack21(tmp_path)),
                                    cbrt_mode=oracle_mod.CBRT_EXACT)
    assert (o.num_elem, o.num_node) == (31408, 34446)
    assert o.run(100) == 0 and o.cycle == rec["cycle"] == 100
    assert o.time == float.fromhex(rec["time"]) and o.deltatime == float.fromhex(rec["deltatime"])
    for f, h in rec["sha256"].items():
        got = hashlib.sha256(np.ascontiguousarray(o.field(f), dtype="<f8").tobytes()).hexdigest()
        assert got == h, f
    # full runs: the same cycle count in both cube-root modes, and the
    # energies within 1e-10 of each other
    full, libm = U21["its-1_cbrt1"], U21["its-1_cbrt0"]
    assert full["status"] == libm["status"] == 0 and full["cycle"] == libm["cycle"] == 872
    a, b = float.fromhex(full["origin_energy"]), float.fromhex(libm["origin_energy"])
    assert abs(a - b) <= 1e-10 * abs(b)
