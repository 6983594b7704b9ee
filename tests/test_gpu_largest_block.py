"""The largest blocks one subdomain may hold: -s 645 -r 1 -c 0 (268 336 125
elements, 269 586 136 nodes, about 75 GB of HBM; element arrays end just
below 2^31 bytes and node arrays pass it, the top of the kernels' 32-bit
byte offsets) and, with the default regions, -s 475 -r 11 -c 1 (the
largest EOS work table those options allow).

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

M, CYCLES = 16, 20


def test_block_size_limits(amd):
    # one element array must stay below 2^31 bytes, and elements x the
    # largest EOS repetition count within 2^31 - 1 (the work table)
    with pytest.raises(amd.LuleshError):
        amd.Group(646, nr=1, balance=1, cost=0)
    with pytest.raises(amd.LuleshError):
        amd.Group(476)          # -r 11 -c 1: 20 repetitions at most
    amd.Group(4, nr=11, balance=1, cost=1).close()


# (the corner's results do not depend on the regions: the oracle's corner
# run uses one)
@pytest.mark.parametrize("nx,nr,cost", [(645, 1, 0), (475, 11, 1)])
def test_largest_block_corner_and_far_field(amd, oracle_mod, nx, nr, cost):
    print(f"case: -s {nx} -r {nr} -c {cost}, {CYCLES} cycles", flush=True)
    g = amd.Group(nx, nr=nr, balance=1, cost=cost)
    o = oracle_mod.Oracle.from_mesh(U.corner_mesh(nx, M), nr=1, balance=1, cost=0,
                                    cbrt_mode=oracle_mod.CBRT_EXACT,
                                    elem_bc=oracle_mod.Oracle(M).elem_bc(), e_deposit=U.einit(nx))
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
        assert_bitwise(U.corner_of(a, nx, M), o.field(f), f"corner {f}")
        if f in ("e", "p", "q", "v", "delv", "vdov"):
            a = a.reshape(nx, nx, nx)
            a[:M, :M, :M] = 1.0 if f == "v" else 0.0
            want = 1.0 if f == "v" else 0.0
            assert np.all(a == want), f"far field {f}"
        del a
    print("elements done; nodes", flush=True)
    pos = (1.125 * np.arange(nx + 1) / float(nx))
    for f in U.CORNER_NODE:
        a = g.field(f)
        assert_bitwise(U.corner_of(a, nx + 1, M + 1), o.field(f), f"corner {f}")
        a = a.reshape(nx + 1, nx + 1, nx + 1)
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
