"""The EOS work a GPU group does is the reference's.

The region layout and elemRep a group holds equal the compiled reference's
(tests/golden/ref_regions.json, from the unmodified omp_4.0 build), and the
EOS work table the EOS kernel walks lists every element exactly elemRep
times: its length is the sum of elemRep, so the kernel evaluates the
equation of state as often as EvalEOSForElems does (lulesh.cc:2677-2815).
"""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

with open(os.path.join(GOLDEN, "ref_regions.json")) as f:
    REF = json.load(f)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()


def expected_table(rep):
    ne = rep.size
    ids = np.repeat(np.arange(ne, dtype=np.int64) & 0x7FFF, rep).astype(np.uint16)
    last = np.cumsum(rep) - 1
    ids[last] |= np.uint16(0x8000)
    return ids


@pytest.mark.parametrize("key", sorted(REF))
def test_group_regions_and_eos_table(amd, key):
    rec = REF[key]
    g = amd.Group(rec["nx"], 1, rec["nr"], rec["balance"], rec["cost"])
    d = g.domains[0]
    assert sha(d.region() + 1) == rec["sha256_regNumList"]
    rep = d.elem_rep()
    assert sha(rep) == rec["sha256_elemRep"]
    tab = d.eos_table()
    assert tab.size == rec["elem_rep_sum"] == int(rep.sum())
    np.testing.assert_array_equal(tab, expected_table(rep))
    g.close()


def test_decomposed_ranks_regions(amd, oracle_mod):
    """each subdomain seeds rand() with its rank (lulesh-init.cc:382): the
    group's ranks agree with the oracle's and their tables with their reps"""
    g = amd.Group(6, tp=2, nr=16, balance=2, cost=2)
    o = oracle_mod.Oracle(6, tp=2, nr=16, balance=2, cost=2)
    for r in range(8):
        d = g.domains[r]
        np.testing.assert_array_equal(d.region(), o.region(r))
        np.testing.assert_array_equal(d.elem_rep(), o.elem_rep(r))
        np.testing.assert_array_equal(d.eos_table(), expected_table(d.elem_rep()))
    g.close()
