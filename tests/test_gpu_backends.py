"""Comm backends behind the reference's backend interface
(LULESH/cuda/src/comm/comm_backend.h:7-22).

* device_copy, the library's Mode A backend (pack into the sender's send
  buffer, one device-to-device copy per message into the receiver's slot,
  comm_ipc_common.h:192-225), installed with lulesh_group_set_comm_backend:
  every exchange step matches the reference's order and layout
  (tests/comm_model.py) bit for bit, receive buffers included, and a
  decomposed run to stoptime equals the oracle's bit for bit.
* tests/native/backend_shim_test: the reference-side shim of INTEGRATION.md
  compiled against a stand-in Domain with pointer-to-member fields, driving
  three groups -- device_local, device_copy and "host_staged", a backend
  defined in that program, outside the library, that stages every message
  through pinned host memory and uses every optional hook -- and requiring
  all three to agree bit for bit, exchange by exchange and over whole runs.
"""
import os
import subprocess

import numpy as np
import pytest

import comm_model as cm
from conftest import REPO, assert_bitwise
from test_comm_layout import MSG, NAMES, domain_fields

pytestmark = pytest.mark.gpu

SHIM = os.path.join(REPO, "tests", "native", "backend_shim_test")
FIELDS = ["e", "p", "q", "v", "ss", "x", "y", "z", "xd", "yd", "zd"]


@pytest.mark.parametrize("step", ["sbn", "posvel", "monoq"])
@pytest.mark.parametrize("tp,nx", [(2, 3), (3, 2), (3, 4)])
def test_device_copy_exchange_reference_order(amd, step, tp, nx):
    rng = np.random.default_rng(31 * tp + nx)
    fields = domain_fields(rng, nx, tp, step)
    g = amd.Group(nx, tp=tp)
    g.set_comm_backend("device_copy")
    for r in range(tp ** 3):
        for f, name in enumerate(NAMES[step]):
            g.domains[r].set_field(name, fields[r][f])
    assert g.exchange(MSG[step], NAMES[step]) == 0
    bufs = cm.exchange(nx, tp, step, fields)
    for r in range(tp ** 3):
        for f, name in enumerate(NAMES[step]):
            assert_bitwise(g.field(name, r), fields[r][f], f"device_copy {step} rank {r} {name}")
        buf, mask = bufs[r]
        rb = g.domains[r].recv_buffer()
        assert np.array_equal(rb[mask].view(np.int64), buf[mask].view(np.int64)), f"buffer {r}"
    g.close()


def test_device_copy_run_to_stoptime_vs_oracle(amd, oracle_mod):
    nx = 4
    o = oracle_mod.Oracle(nx, tp=2, cbrt_mode=oracle_mod.CBRT_EXACT)
    assert o.run() == 0
    g = amd.Group(nx, tp=2)
    g.set_comm_backend("device_copy")
    assert g.run() == 0
    assert g.sync()[1] == o.cycle
    for r in range(8):
        for f in FIELDS:
            assert_bitwise(g.field(f, r), o.field(f, r), f"device_copy run rank {r} {f}")
    g.close()


@pytest.mark.parametrize("args", [("4", "2", "-1"), ("3", "3", "30"), ("6", "2", "60")])
def test_shim_and_outside_backend(amd, args):
    if not os.path.exists(SHIM):
        pytest.fail(f"{SHIM} not built (make -C tests/native)")
    print(f"backend_shim_test {' '.join(args)}", flush=True)
    r = subprocess.run([SHIM, *args], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "BACKENDS_EQUAL" in r.stdout
