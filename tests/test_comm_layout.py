"""Halo-exchange order and wire layout against the reference's own sources.

The expected values come from tests/comm_model.py, which replays an exchange
from tests/golden/comm_layout.json: message order, doRecv/doSend gates and
index maps read mechanically out of the reference (comm_backend.h's
shmRecvOffset, lulesh-comm.cc's CommRecv / CommSend / CommSBN /
CommSyncPosVel / CommMonoQ) by tests/golden/extract_comm_layout.py.  Neither
the library's plan tables (csrc/lulesh_comm.cpp) nor the oracle's
(oracle/lulesh_oracle.c) enter the model, so these tests pin both of them.

The field values are chosen so that the order of the adds at a shared node
changes the rounded sum, and every subdomain's positions carry its rank, so
a swapped message, a wrong final writer or a transposed index map changes
the result; test_model_is_order_sensitive shows that the data can tell
orders apart.
"""
import itertools

import numpy as np
import pytest

import comm_model as cm

DIRS = [d for d in itertools.product((-1, 0, 1), repeat=3) if d != (0, 0, 0)]


def sensitive_values(rng, n):
    """doubles whose sums round differently in different orders: full random
    mantissas within a few binades (2^-3 .. 2^4) and random signs, so that
    no term swamps the others and most reorderings of three or more terms
    change the last bits"""
    return (1.0 + rng.random(n)) * np.exp2(rng.integers(-3, 4, n)) * rng.choice([-1.0, 1.0], n)


def test_fixture_orders_agree():
    fx = cm.layout()
    assert len(fx["shm_order"]) == 26
    assert [(m["dir"], m["gate"], m["kind"]) for m in fx["shm_order"]] == \
        [(m["dir"], m["gate"], m["kind"]) for m in fx["recv_post"]]
    assert sorted(tuple(m["dir"]) for m in fx["recv_post"]) == sorted(DIRS)
    kinds = [m["kind"] for m in fx["recv_post"]]
    assert kinds == ["face"] * 6 + ["edge"] * 12 + ["corner"] * 8
    for nx in fx["sizes"]:
        sz = fx["sizes"][nx]
        for step in ("sbn", "posvel", "monoq"):
            assert [m["dir"] for m in sz[step]] == [m["dir"] for m in fx["recv_post"]][:len(sz[step])]


@pytest.mark.parametrize("tp,nx", [(3, 5), (2, 45), (3, 2)])
@pytest.mark.parametrize("xfer", [1, 3, 6])
@pytest.mark.parametrize("do_recv", [True, False])
@pytest.mark.parametrize("plane_only", [False, True])
def test_recv_offsets(oracle_mod, tp, nx, xfer, do_recv, plane_only):
    """shmRecvOffset of every (receiver, direction): the oracle's and the
    library's (host-only query) against the fixture's posting order"""
    import lulesh_amd
    o = oracle_mod.Oracle(nx, tp=tp)
    for r in range(tp ** 3):
        want = cm.recv_offsets(nx, tp, r, xfer, do_recv, plane_only)
        for d in DIRS:
            w = want.get(d, -1)
            assert o.recv_offset(r, *d, xfer, do_recv, plane_only) == w, (r, d)
            assert lulesh_amd.shm_recv_offset(nx, tp, r, *d, xfer, do_recv, plane_only) == w, (r, d)


@pytest.mark.parametrize("tp,nx", [(2, 45), (3, 5), (4, 3), (3, 100)])
def test_com_buf_size(oracle_mod, tp, nx):
    import lulesh_amd
    o = oracle_mod.Oracle(nx if nx <= 45 else 2, tp=tp) if nx <= 45 else None
    for r in range(tp ** 3):
        want = cm.com_buf_size(nx, tp, r)
        assert lulesh_amd.comm_buf_size_of(nx, tp, r) == want
        if o is not None:
            assert o.com_buf_size(r) == want
    # closed forms of lulesh-init.cc:334-356
    if (tp, nx) == (2, 45):
        assert cm.com_buf_size(45, 2, 0) == 39184    # corner: 3 faces, 3 edges, 1 corner
    if (tp, nx) == (3, 5):
        assert cm.com_buf_size(5, 3, 13) == 3008     # centre: 6 faces, 12 edges, 8 corners


def domain_fields(rng, nx, tp, step):
    n = (nx + 1) ** 3 if step != "monoq" else nx ** 3 + 6 * nx * nx
    nf = {"sbn": 3, "posvel": 6, "monoq": 3}[step]
    out = []
    for r in range(tp ** 3):
        fs = []
        for f in range(nf):
            if step == "posvel":
                # rank in the integer part, node and field below it
                v = r * 1.0e6 + np.arange(n) + f * 1.0e-3
            else:
                v = sensitive_values(rng, n)
                if step == "monoq":
                    v[nx ** 3:] = -1.0   # ghost planes start out stale
            fs.append(v)
        out.append(fs)
    return out


NAMES = {"sbn": ["fx", "fy", "fz"], "posvel": ["x", "y", "z", "xd", "yd", "zd"],
         "monoq": ["delv_xi", "delv_eta", "delv_zeta"]}
MSG = {"sbn": 1024, "posvel": 2048, "monoq": 3072}


@pytest.mark.parametrize("step", ["sbn", "posvel", "monoq"])
@pytest.mark.parametrize("tp,nx", [(2, 2), (2, 3), (2, 4), (3, 2), (3, 3)])
def test_oracle_exchange_matches_reference_order(oracle_mod, step, tp, nx):
    rng = np.random.default_rng(1000 * tp + nx)
    fields = domain_fields(rng, nx, tp, step)
    o = oracle_mod.Oracle(nx, tp=tp)
    for r in range(tp ** 3):
        for f, name in enumerate(NAMES[step]):
            o.set_field(name, fields[r][f], r)
    assert o.exchange(MSG[step], NAMES[step]) == 0
    bufs = cm.exchange(nx, tp, step, fields)
    for r in range(tp ** 3):
        for f, name in enumerate(NAMES[step]):
            got = o.field(name, r)
            assert np.array_equal(got.view(np.int64), fields[r][f].view(np.int64)), (r, name)
        rb = o.recv_buffer(r)
        buf, mask = bufs[r]
        assert np.array_equal(rb[mask].view(np.int64), buf[mask].view(np.int64)), f"receive buffer {r}"


@pytest.mark.parametrize("step", ["sbn", "posvel"])
def test_model_is_order_sensitive(step):
    """The same data unpacked in the reversed message order gives different
    fields: the exchange tests above would notice a wrong order."""
    nx, tp = 3, 3
    rng = np.random.default_rng(7)
    a = domain_fields(rng, nx, tp, step)
    b = [[v.copy() for v in fs] for fs in a]
    cm.exchange(nx, tp, step, a)
    rev = {r: list(range(cm.unpack_count(nx, tp, r, step)))[::-1] for r in range(tp ** 3)}
    cm.exchange(nx, tp, step, b, order=rev)
    changed = sum(int(np.sum(a[r][0].view(np.int64) != b[r][0].view(np.int64))) for r in range(tp ** 3))
    assert changed > 0
    if step != "sbn":
        return
    # and swapping two messages of the centre subdomain that share nodes
    # (planeMin and rowMin: the nodes of their common edge) is seen too
    c = [[v.copy() for v in fs] for fs in domain_fields(np.random.default_rng(7), nx, tp, step)]
    k = cm.unpack_count(nx, tp, 13, step)
    cm.exchange(nx, tp, step, c, order={13: [2, 1, 0] + list(range(3, k))})
    assert any(not np.array_equal(a[13][f].view(np.int64), c[13][f].view(np.int64))
               for f in range(len(a[13])))


def test_posvel_final_writer_is_the_last_sharer():
    """PosVel (doRecv = false): every copy of a shared node ends with the
    value of the sharer whose message the reference unpacks last -- on the
    max-side faces, edges and corners only, mixed-direction edges included."""
    nx, tp = 3, 3
    fields = domain_fields(None, nx, tp, "posvel")
    cm.exchange(nx, tp, "posvel", fields)
    n1 = nx + 1
    writers = {}
    for r in range(tp ** 3):
        src = np.floor(fields[r][0] / 1.0e6).astype(int)
        writers[r] = src
    # a node with no max-side neighbour keeps its own value
    for r in range(tp ** 3):
        col, row, plane = cm.position(r, tp)
        k, j, i = np.meshgrid(np.arange(n1), np.arange(n1), np.arange(n1), indexing="ij")
        k, j, i = k.ravel(), j.ravel(), i.ravel()
        upper = ((i == nx) & (col < tp - 1)) | ((j == nx) & (row < tp - 1)) | ((k == nx) & (plane < tp - 1))
        assert np.all(writers[r][~upper] == r)
        assert np.all(writers[r][upper] != r)
