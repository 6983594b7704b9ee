"""The CPU oracle against the reference's own outputs.

Golden fixtures (tests/golden/ref_omp.json, ref_fields_s*.npz) come from the
unmodified LULESH/omp_4.0 sources compiled in the build container
(tests/golden/make_golden.py).  The oracle in libm mode must reproduce them
bit for bit: cycle count, final time and dt, every dumped field (sha256),
and the VerifyAndWriteFinalOutput symmetry figures.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, assert_bitwise

DB = json.load(open(os.path.join(GOLDEN, "ref_omp.json")))
FIELDS = ["e", "p", "q", "v", "ss", "x", "y", "z", "xd", "yd", "zd"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def _check(oracle_mod, key):
    rec = DB[key]
    o = oracle_mod.Oracle(rec["nx"], 1, rec["nr"], rec["balance"], rec["cost"],
                          oracle_mod.CBRT_LIBM)
    assert o.run(rec["its"]) == 0
    assert o.cycle == rec["cycle"]
    assert o.time == float.fromhex(rec["time"])
    assert o.deltatime == float.fromhex(rec["deltatime"])
    for f in FIELDS:
        assert sha(o.field(f)) == rec["sha256"][f], f"{key}: field {f}"
    v = o.verify()
    assert v["origin_energy"] == float.fromhex(rec["origin_energy"])
    assert v["total_abs_diff"] == float.fromhex(rec["total_abs_diff"])
    assert v["max_abs_diff"] == float.fromhex(rec["max_abs_diff"])
    mrd = float.fromhex(rec["max_rel_diff"])
    assert v["max_rel_diff"] == mrd or (np.isnan(mrd) and np.isnan(v["max_rel_diff"]))
    assert "%12.6e" % v["origin_energy"] == "%12.6e" % float(rec["printed_energy"])
    return o


@pytest.mark.parametrize("key", ["-s 4", "-s 8", "-s 8 -i 10", "-s 12 -r 5 -b 2 -c 2",
                                 "-s 16", "-s 16 -i 50", "-s 30 -i 100"])
def test_oracle_matches_reference(oracle_mod, key):
    _check(oracle_mod, key).close()


def test_oracle_matches_reference_s30_full(oracle_mod):
    # config 1 of the workload table: -s 30 to the default stoptime
    o = _check(oracle_mod, "-s 30")
    assert o.cycle == 932
    o.close()


@pytest.mark.parametrize("nx", [4, 8])
def test_oracle_fields_vs_fixture(oracle_mod, nx):
    ref = np.load(os.path.join(GOLDEN, f"ref_fields_s{nx}.npz"))
    o = oracle_mod.Oracle(nx)
    o.run()
    for f in FIELDS:
        assert_bitwise(o.field(f), ref[f], f"s={nx} {f}")


def test_oracle_vs_compiled_reference_live(oracle_mod, tmp_path):
    """Where oracle/_ref was built (the build container), run it now."""
    import struct
    import subprocess
    ref = os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref", "lulesh_ref")
    if not os.path.exists(ref):
        pytest.skip("oracle/_ref/lulesh_ref not built here")
    dump = str(tmp_path / "d.bin")
    env = dict(os.environ, OMP_NUM_THREADS="2", LULESH_REF_DUMP=dump)
    subprocess.run([ref, "-s", "6", "-r", "3", "-c", "2"], env=env, check=True,
                   capture_output=True)
    b = open(dump, "rb").read()
    nx, cyc = struct.unpack("ii", b[:8])
    a = np.frombuffer(b[24:], dtype=np.float64)
    o = oracle_mod.Oracle(6, 1, 3, 1, 2)
    o.run()
    assert o.cycle == cyc
    ne, nn = 6 ** 3, 7 ** 3
    assert_bitwise(o.field("e"), a[:ne], "e")
    assert_bitwise(o.field("x"), a[5 * ne:5 * ne + nn], "x")


def test_results_independent_of_regions(oracle_mod):
    # EvalEOSForElems repeats identical work rep times (lulesh.cc:2677-2815):
    # -r/-b/-c change only the cost, never a bit of the result
    base = oracle_mod.Oracle(8, 1, 11, 1, 1)
    base.run()
    for nr, b, c in [(1, 1, 1), (3, 4, 2), (16, 2, 1)]:
        o = oracle_mod.Oracle(8, 1, nr, b, c)
        o.run()
        for f in ["e", "x", "q"]:
            assert_bitwise(o.field(f), base.field(f), f"r={nr} b={b} c={c} {f}")


def test_region_layout_properties(oracle_mod):
    o = oracle_mod.Oracle(16, 1, 11, 1, 1)
    reg = o.region()
    rep = o.elem_rep()
    assert reg.min() >= 0 and reg.max() <= 10
    # rep rule of main, lulesh.cc:3178-3185 with nr=11, cost=1
    expect = np.where(reg < 5, 1, np.where(reg < 10, 2, 20))
    np.testing.assert_array_equal(rep, expect)
    # consecutive runs never repeat a region (lulesh-init.cc:421-428)
    runs = reg[np.r_[True, reg[1:] != reg[:-1]]]
    assert np.all(runs[1:] != runs[:-1])
