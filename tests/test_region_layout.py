"""Region layout and EOS repetition counts against the compiled reference.

tests/golden/ref_regions.json was written by tests/golden/make_region_golden.py
from the unmodified omp_4.0 build's own regNumList and elemRep
(CreateRegionIndexSets, lulesh-init.cc:377-484; main, lulesh.cc:3171-3192).
Results never depend on the layout (every EOS repetition recomputes the same
values), so only these tests show that the EOS work the kernels do -- and
the benchmark times -- is the reference's.  The library's layout is read
through its host-only query (no GPU needed); tests/test_gpu_eos_work.py
checks the copies a GPU group holds and its EOS work table.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

with open(os.path.join(GOLDEN, "ref_regions.json")) as f:
    REF = json.load(f)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()


@pytest.mark.parametrize("key", sorted(REF))
def test_library_region_layout(key):
    import lulesh_amd
    rec = REF[key]
    reg, rep = lulesh_amd.region_layout(rec["nx"], 1, 0, rec["nr"], rec["balance"], rec["cost"])
    assert sha(reg + 1) == rec["sha256_regNumList"], key
    assert sha(rep) == rec["sha256_elemRep"], key
    assert int(rep.sum()) == rec["elem_rep_sum"]
    assert np.bincount(reg, minlength=rec["nr"]).tolist() == rec["region_sizes"]


@pytest.mark.parametrize("key", [k for k in sorted(REF) if REF[k]["nx"] <= 40])
def test_oracle_region_layout(oracle_mod, key):
    rec = REF[key]
    o = oracle_mod.Oracle(rec["nx"], 1, rec["nr"], rec["balance"], rec["cost"])
    assert sha(o.region() + 1) == rec["sha256_regNumList"], key
    assert sha(o.elem_rep()) == rec["sha256_elemRep"], key


def test_s30_arrays():
    """element by element at -s 30 (the arrays themselves are committed)"""
    import lulesh_amd
    z = np.load(os.path.join(GOLDEN, "ref_regions_s30.npz"))
    reg, rep = lulesh_amd.region_layout(30, 1, 0, 11, 1, 1)
    np.testing.assert_array_equal(reg + 1, z["regNumList"].astype(np.int32))
    np.testing.assert_array_equal(rep, z["elemRep"].astype(np.int32))


def test_rep_rule():
    """rep = 1 for the lower half of the regions, 1 + cost up to the last
    (nr + 15) / 20, 10 (1 + cost) for those (lulesh.cc:3176-3185)"""
    import lulesh_amd
    for nr, cost in [(11, 1), (16, 2), (4, 3), (1, 1), (40, 0)]:
        reg, rep = lulesh_amd.region_layout(12, 1, 0, nr, 1, cost)
        want = np.where(reg < nr // 2, 1, np.where(reg < nr - (nr + 15) // 20, 1 + cost, 10 * (1 + cost)))
        np.testing.assert_array_equal(rep, want)


def test_region_weight_overflow_rejected():
    """The reference sums (i+1)^balance in an int: layouts whose sum passes
    2^31 - 1 are refused instead of wrapping (lulesh-init.cc:384-388)."""
    import lulesh_amd
    for nr, b in [(11, 9), (100, 5), (2, 31)]:
        with pytest.raises(lulesh_amd.LuleshError):
            lulesh_amd.region_layout(8, 1, 0, nr, b, 1)
    reg, _ = lulesh_amd.region_layout(8, 1, 0, 11, 8, 1)   # 1.5e9: still fits
    assert reg.max() < 11
