"""Failure handling, forced from the host: the reference exits with
VolumeError (-1) / QStopError (-2) (LULESH/omp_4.0/lulesh.h:41,
lulesh.cc:1263-1275, 2531-2546); the library returns those codes from the
run, at the same cycle as the oracle, and every kernel still exits
normally."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_qstop_error(amd, oracle_mod):
    o = oracle_mod.Oracle(8, cbrt_mode=oracle_mod.CBRT_EXACT)
    o.set_qstop(1e-30)
    assert o.run() == -2
    g = amd.Group(8)
    g.set_qstop(1e-30)
    assert g.run() == amd.QSTOP_ERROR
    assert g.sync()[1] == o.cycle


def test_volume_error(amd, oracle_mod):
    g = amd.Group(6)
    v = g.field("v")
    v[17] = -1.0
    g.domains[0].set_field("v", v)
    assert g.run() == amd.VOLUME_ERROR
    o = oracle_mod.Oracle(6)
    ov = o.field("v")
    ov[17] = -1.0
    o.set_field("v", ov)
    assert o.run() == -1
    assert g.sync()[1] == o.cycle == 1


def test_unexpected_direction(amd):
    g = amd.Group(4, tp=2)
    # rank 0 sits at the mesh origin: nothing arrives from the min sides
    assert g.domains[0].recv_offset(-1, 0, 0, 3, True, False) == -1
    assert g.domains[0].recv_offset(1, 0, 0, 3, True, False) >= 0


def test_stop_is_sticky_and_resumable_runs_are_noops(amd):
    g = amd.Group(4)
    assert g.run() == 0
    rc, c1, t1, _ = g.sync()
    g.enqueue(10)                 # cycles after the stop condition do nothing
    rc, c2, t2, _ = g.sync()
    assert (c1, t1) == (c2, t2)
    e = g.field("e")
    g.enqueue(3)
    assert np.array_equal(e, g.field("e"))
