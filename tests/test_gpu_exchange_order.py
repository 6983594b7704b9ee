"""The library's halo exchange on the GPU against the reference's message
order, wire layout and index maps (tests/comm_model.py replaying
tests/golden/comm_layout.json, which was read out of the reference's
sources).  Both forms are checked: one whole exchange step
(lulesh_group_exchange) and the reference's own call sequence on every
subdomain (CommRecv, then CommSend on all, then CommSBN / CommSyncPosVel /
CommMonoQ on all).  Fields and the words of the receive buffers the step
wrote must match bit for bit, with data for which a different add order or
final writer changes the result (tests/test_comm_layout.py)."""
import numpy as np
import pytest

import comm_model as cm
from test_comm_layout import MSG, NAMES, domain_fields

pytestmark = pytest.mark.gpu

CASES = [(2, 2), (2, 3), (2, 4), (3, 2), (3, 3)]


def seeded_group(amd, nx, tp, step, seed):
    rng = np.random.default_rng(seed)
    fields = domain_fields(rng, nx, tp, step)
    g = amd.Group(nx, tp=tp)
    for r in range(tp ** 3):
        for f, name in enumerate(NAMES[step]):
            g.domains[r].set_field(name, fields[r][f])
    return g, fields


def check(g, fields, bufs, tp, step, what):
    for r in range(tp ** 3):
        for f, name in enumerate(NAMES[step]):
            got = g.field(name, r)
            assert np.array_equal(got.view(np.int64), fields[r][f].view(np.int64)), \
                f"{what}: rank {r} {name}"
        buf, mask = bufs[r]
        rb = g.domains[r].recv_buffer()
        assert np.array_equal(rb[mask].view(np.int64), buf[mask].view(np.int64)), \
            f"{what}: receive buffer of rank {r}"


@pytest.mark.parametrize("step", ["sbn", "posvel", "monoq"])
@pytest.mark.parametrize("tp,nx", CASES)
def test_group_exchange_reference_order(amd, step, tp, nx):
    g, fields = seeded_group(amd, nx, tp, step, 100 * tp + nx)
    assert g.exchange(MSG[step], NAMES[step]) == 0
    bufs = cm.exchange(nx, tp, step, fields)
    check(g, fields, bufs, tp, step, f"group exchange {step} tp={tp} nx={nx}")
    g.close()


@pytest.mark.parametrize("step", ["sbn", "posvel", "monoq"])
@pytest.mark.parametrize("tp,nx", [(2, 3), (3, 3)])
def test_comm_calls_reference_order(amd, step, tp, nx):
    """The five comm entry points called as the reference's solver calls
    them (lulesh.cc:1372,1405-1408; 1574,1614-1617; 2496,2515-2519)."""
    g, fields = seeded_group(amd, nx, tp, step, 7 * tp + nx)
    msg = MSG[step]
    n = nx if step == "monoq" else nx + 1
    flag = step != "posvel"
    plane_only = step == "monoq"
    for d in g.domains:
        assert d.comm_recv(msg, len(NAMES[step]), n, n, n, flag, plane_only) == 0
    for d in g.domains:
        assert d.comm_send(msg, NAMES[step], n, n, n, flag, plane_only) == 0
    for d in g.domains:
        if step == "sbn":
            assert d.comm_sbn(NAMES[step]) == 0
        elif step == "posvel":
            assert d.comm_sync_pos_vel() == 0
        else:
            assert d.comm_monoq() == 0
    bufs = cm.exchange(nx, tp, step, fields)
    check(g, fields, bufs, tp, step, f"comm calls {step} tp={tp} nx={nx}")
    g.close()
