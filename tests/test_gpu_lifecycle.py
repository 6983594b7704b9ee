"""Lifecycle accounting (the reference's PHASE_* lines,
LULESH/cuda/src/lulesh.cu:4918-4941) and the xyz.asc solution writer
(write_solution, lulesh.cu:4584-4603), through the C ABI and the driver."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(REPO, "cuda-ipc-benchmarks_amd", "lulesh_amd_run")


def test_write_solution_format(amd, tmp_path):
    g = amd.Group(4)
    assert g.run() == 0
    path = tmp_path / "xyz.asc"
    g.write_solution(path)
    lines = path.read_text().split("\n")
    assert lines[-1] == ""
    lines = lines[:-1]
    x, y, z = g.field("x"), g.field("y"), g.field("z")
    assert len(lines) == 4 * x.size
    for i in range(x.size):
        assert lines[4 * i] == "%10d" % i
        assert lines[4 * i + 1] == "%.10f" % x[i]
        assert lines[4 * i + 2] == "%.10f" % y[i]
        assert lines[4 * i + 3] == "%.10f" % z[i]


def test_phase_times_and_teardown(amd):
    g = amd.Group(6, tp=2)
    setup, win, peer, free = g.phase_times()
    assert setup > 0.0 and win == 0.0 and peer == 0.0 and free == 0.0
    g.set_max_cycles(3)
    assert g.run() == 0
    g.comm_teardown()
    assert g.phase_times()[3] >= 0.0
    # a decomposed group cannot step without its receive buffers
    g.set_max_cycles(6)
    with pytest.raises(amd.LuleshError):
        g.run()


def test_single_domain_steps_after_teardown(amd, oracle_mod):
    o = oracle_mod.Oracle(6, cbrt_mode=oracle_mod.CBRT_EXACT)
    o.run(10)
    g = amd.Group(6)
    g.comm_teardown()
    g.set_max_cycles(10)
    assert g.run() == 0
    assert np.array_equal(g.field("e").view(np.int64), o.field("e").view(np.int64))


def test_driver_phase_lines_and_solution(tmp_path):
    out = subprocess.run([DRIVER, "-s", "6", "-i", "20"], cwd=tmp_path, capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    names = [ln.split()[0] for ln in out.stdout.splitlines() if ln.startswith("PHASE_")]
    assert names == ["PHASE_COMM_SETUP_ms", "PHASE_WIN_ALLOCATE_ms", "PHASE_PEER_QUERY_ms",
                     "PHASE_COMM_FREE_ms", "PHASE_APP_TOTAL_ms"]
    assert "Final Origin Energy" in out.stdout
    assert (tmp_path / "xyz.asc").stat().st_size > 0


def test_driver_progress_and_iteration_cap(tmp_path, oracle_mod):
    # -p prints every cycle in the reference's format (lulesh.cc:3227-3230);
    # -z 5 stops after five cycles (lulesh.cc:3231-3234)
    out = subprocess.run([DRIVER, "-s", "6", "-p", "-z", "5", "-w", "0"], cwd=tmp_path,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    prog = [ln for ln in out.stdout.splitlines() if ln.startswith("cycle = ")]
    o = oracle_mod.Oracle(6, cbrt_mode=oracle_mod.CBRT_EXACT)
    want = []
    for c in range(1, 6):
        o.run(1)   # one more cycle
        assert o.cycle == c
        want.append("cycle = %d, time = %e, dt=%e" % (c, o.time, o.deltatime))
    assert prog == want
    assert "Iteration count     =  5 " in out.stdout
