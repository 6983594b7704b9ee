"""The library's host-only code (the .lmesh reader, set-up, region layout,
exchange plans, the force kernel's column planner) and the oracle, built
with AddressSanitizer and UndefinedBehaviorSanitizer
(tests/native/host_sanitize_test, g++ and gcc with
-fsanitize=address,undefined; CPU only) and run on the reference's own mesh
plus malformed files: truncated mid-record, headers claiming huge counts,
node ids out of range, garbage.  Clean = exit status 0 and no sanitizer
report."""
import os
import subprocess

import pytest

from conftest import GOLDEN, REPO

NATIVE = os.path.join(REPO, "tests", "native")
EXE = os.path.join(NATIVE, "host_sanitize_test")
LMESH = os.path.join(GOLDEN, "sedov15oct.lmesh")


@pytest.fixture(scope="module")
def exe():
    r = subprocess.run(["make", "-C", NATIVE, "host_sanitize_test"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return EXE


def test_host_code_clean_under_asan_ubsan(exe, tmp_path):
    text = open(LMESH).read()
    tok = text.split()
    ne, nn = int(tok[0]), int(tok[1])
    files = {
        "ok": [LMESH],
        "bad": [],
    }

    def bad(name, body):
        p = tmp_path / name
        p.write_text(body)
        files["bad"].append(str(p))

    bad("empty.lmesh", "")
    bad("garbage.lmesh", "hello world\n")
    bad("half.lmesh", text[: len(text) // 2])
    bad("mid_coord.lmesh", " ".join(tok[: 2 + 3 * (nn // 3) + 2]))
    bad("mid_nodelist.lmesh", " ".join(tok[: 2 + 3 * nn + 8 * (ne // 2) + 5]))
    bad("huge_nodes.lmesh", "1 268435455\n0 0 0\n")
    bad("huge_elems.lmesh", "67108863 8\n" + "0 0 0\n" * 8)
    bad("negative.lmesh", "-5 8\n")
    bad("huge_set.lmesh", " ".join(tok[: 2 + 3 * nn + 14 * ne]) + " 1000000000 1 2 3\n")
    bad("set_overflow.lmesh", " ".join(tok[: 2 + 3 * nn + 14 * ne]) + " 99999999999 1\n")
    lines = text.split("\n")
    lines[1 + nn] = "0 5 6 7 1 2 3 99999"
    bad("node_out_of_range.lmesh", "\n".join(lines))
    args = [f"ok:{p}" for p in files["ok"]] + [f"bad:{p}" for p in files["bad"]]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, *args], capture_output=True, text=True, env=env, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out
    assert "host_sanitize_test: clean" in r.stdout
    for marker in ("ERROR: AddressSanitizer", "runtime error:", "ERROR: LeakSanitizer"):
        assert marker not in out, out
