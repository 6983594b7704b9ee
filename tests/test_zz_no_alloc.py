"""lulesh_group_exchange_noalloc: the allocation-free exchange step against
lulesh_group_exchange, called through the C ABI.

Two calls in a row (the second on a smaller decomposition), on both of the
library's backends and all three exchange steps: fields and the receive
buffer words the step writes must equal those of lulesh_group_exchange on
an identically seeded group, bit for bit, and the reference's order
(tests/comm_model.py).  A decomposed group stepped after such an exchange
runs as after the allocating one.  A step larger than the scratch array is
refused with LULESH_ENOMEM.
"""
import ctypes

import numpy as np
import pytest

import comm_model as cm
from test_comm_layout import MSG, NAMES, domain_fields

pytestmark = pytest.mark.gpu


def noalloc(amd, g, step):
    L = amd.load()
    fn = L.lulesh_group_exchange_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_void_p]
    ids = np.array([amd.FIELDS[f] for f in NAMES[step]], dtype=np.int32)
    return fn(g.h, MSG[step], len(ids), ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), None)


def seeded(amd, nx, tp, step, seed, backend):
    fields = domain_fields(np.random.default_rng(seed), nx, tp, step)
    g = amd.Group(nx, tp=tp)
    # fresh receive buffers: creation's nodalMass exchange left words in them
    g.set_comm_backend(backend)
    for r in range(tp ** 3):
        for f, name in enumerate(NAMES[step]):
            g.domains[r].set_field(name, fields[r][f])
    return g, fields


@pytest.mark.parametrize("backend", ["device_local", "device_copy"])
@pytest.mark.parametrize("step", ["sbn", "posvel", "monoq"])
def test_noalloc_matches_exchange(amd, backend, step):
    for tp, nx in [(3, 4), (2, 3)]:       # the second call is the smaller step
        a, fields = seeded(amd, nx, tp, step, 5 * tp + nx, backend)
        b, _ = seeded(amd, nx, tp, step, 5 * tp + nx, backend)
        assert noalloc(amd, a, step) == 0
        assert b.exchange(MSG[step], NAMES[step]) == 0
        bufs = cm.exchange(nx, tp, step, fields)
        for r in range(tp ** 3):
            for f, name in enumerate(NAMES[step]):
                ga, gb = a.field(name, r), b.field(name, r)
                assert np.array_equal(ga.view(np.int64), gb.view(np.int64)), (tp, nx, r, name)
                assert np.array_equal(ga.view(np.int64), fields[r][f].view(np.int64)), (tp, nx, r, name)
            ra, rb = a.domains[r].recv_buffer(), b.domains[r].recv_buffer()
            assert np.array_equal(ra.view(np.int64), rb.view(np.int64)), (tp, nx, r, "receive buffer")
            buf, mask = bufs[r]
            assert np.array_equal(ra[mask].view(np.int64), buf[mask].view(np.int64))
        a.close()
        b.close()


def test_run_after_noalloc_exchange(amd):
    """the engine a decomposed group held before the exchange is rebuilt"""
    nx, tp = 4, 2
    runs = []
    for use_noalloc in (True, False):
        g = amd.Group(nx, tp=tp)
        g.set_max_cycles(10)
        assert g.run() == 0
        if use_noalloc:
            assert noalloc(amd, g, "posvel") == 0
        else:
            assert g.exchange(MSG["posvel"], NAMES["posvel"]) == 0
        g.set_max_cycles(30)
        assert g.run() == 0
        runs.append([g.field(f, r) for r in range(8) for f in ("e", "x", "xd")])
        g.close()
    for u, v in zip(*runs):
        assert np.array_equal(u.view(np.int64), v.view(np.int64))


def test_noalloc_refuses_steps_beyond_scratch(amd):
    g = amd.Group(45, tp=4)            # 64 subdomains: ~35 MB of send staging
    g.set_comm_backend("device_copy")
    assert noalloc(amd, g, "posvel") == amd.ENOMEM
    g.close()
