"""The driver's command line against the reference's ParseCommandLineOptions
(LULESH/omp_4.0/lulesh-util.cc:60-173): usage, the refused -v, unknown
options and malformed values -- all decided before any GPU call, so these
run on the CPU.  Runs that step are GPU tests (tests/test_gpu_lifecycle.py,
tests/test_gpu_umesh.py)."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(REPO, "cuda-ipc-benchmarks_amd", "lulesh_amd_run")

pytestmark = pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")


def run(*args):
    return subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=60)


def test_usage():
    r = run("-h")
    assert r.returncode == 0
    for opt in ("-q", "-i <iterations>", "-s <size>", "-r <numregions>", "-b <balance>", "-c <cost>",
                "-f <numfiles>", "-p", "-v", "-z <cycles>", "-h", "-u <file>"):
        assert f" {opt} " in r.stdout, opt


def test_viz_refused():
    # a reference build without VIZ_MESH refuses -v the same way
    r = run("-s", "4", "-v")
    assert r.returncode != 0 and "-v requires" in r.stdout


def test_unknown_option():
    r = run("-x")
    assert r.returncode != 0
    assert "ERROR: Unknown command line argument: -x" in r.stdout and "Usage:" in r.stdout


@pytest.mark.parametrize("args", [["-s"], ["-s", "abc"], ["-z", "1x"], ["-f"], ["-u"]])
def test_malformed_values(args):
    r = run(*args)
    assert r.returncode != 0
