"""The largest block one subdomain may hold: -s 645 -r 1 -c 0 (268 336 125
elements, 269 586 136 nodes, about 75 GB of HBM).  Element arrays end just
below 2^31 bytes and node arrays pass it, the top of the kernels' 32-bit
byte offsets.

The oracle cannot repeat a run of this size, so it is checked by locality:
for its first cycles the blast stays near the origin, and the 16^3 corner
of the block must hold, bit for bit, what the oracle leaves on that corner
cut out as a mesh of its own (the property is pinned on the CPU against the
full block, test_umesh_oracle.py); every node and element outside the
corner must still hold its initial state, which a wrong address anywhere
in the arrays would disturb.
"""
import numpy as np
import pytest

from conftest import assert_bitwise
import umesh_util as U

pytestmark = pytest.mark.gpu

NX, M, CYCLES = 645, 16, 20


def test_block_size_limits(amd):
    # one element array must stay below 2^31 bytes, and elements x the
    # largest EOS repetition count within 2^31 - 1 (the work table)
    with pytest.raises(amd.LuleshError):
        amd.Group(NX + 1, nr=1, balance=1, cost=0)
    with pytest.raises(amd.LuleshError):
        amd.Group(476)          # -r 11 -c 1: 20 repetitions at most
    amd.Group(4, nr=11, balance=1, cost=1).close()


def test_largest_block_corner_and_far_field(amd, oracle_mod):
    print(f"case: -s {NX} -r 1 -c 0, {CYCLES} cycles", flush=True)
    g = amd.Group(NX, nr=1, balance=1, cost=0)
    o = oracle_mod.Oracle.from_mesh(U.corner_mesh(NX, M), nr=1, balance=1, cost=0,
                                    cbrt_mode=oracle_mod.CBRT_EXACT,
                                    elem_bc=oracle_mod.Oracle(M).elem_bc(), e_deposit=U.einit(NX))
    assert g.sync()[3] == o.deltatime          # the same deposit (dt0 follows from it)
    g.set_max_cycles(CYCLES)
    assert g.run() == 0
    assert o.run(CYCLES) == 0
    assert g.sync()[1:] == (o.cycle, o.time, o.deltatime)
    print("stepped; comparing the corner", flush=True)
    e = o.field("e").reshape(M, M, M)
    assert e[0, 0, 0] > 0 and not e[-1].any() and not e[:, -1].any() and not e[:, :, -1].any()

    for f in U.CORNER_ELEM:
        a = g.field(f)
        assert_bitwise(U.corner_of(a, NX, M), o.field(f), f"corner {f}")
        if f in ("e", "p", "q", "v", "delv", "vdov"):
            a = a.reshape(NX, NX, NX)
            a[:M, :M, :M] = 1.0 if f == "v" else 0.0
            want = 1.0 if f == "v" else 0.0
            assert np.all(a == want), f"far field {f}"
        del a
    print("elements done; nodes", flush=True)
    pos = (1.125 * np.arange(NX + 1) / float(NX))
    for f in U.CORNER_NODE:
        a = g.field(f)
        assert_bitwise(U.corner_of(a, NX + 1, M + 1), o.field(f), f"corner {f}")
        a = a.reshape(NX + 1, NX + 1, NX + 1)
        if f in ("x", "y", "z"):
            # initial positions: 1.125 i / nx along the node's own axis
            shape = {"x": (1, 1, -1), "y": (1, -1, 1), "z": (-1, 1, 1)}[f]
            bad = a != pos.reshape(shape)
        else:
            bad = a != 0.0
        bad[:M + 1, :M + 1, :M + 1] = False
        assert not bad.any(), f"far field {f}: {np.argwhere(bad)[:4]}"
        del a, bad
    g.close()
    o.close()
