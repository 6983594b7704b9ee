"""The C-ABI library loads and exports every entry point include/lulesh_amd.h
declares (no GPU calls)."""
import ctypes
import os
import re

from conftest import REPO

HEADER = os.path.join(REPO, "include", "lulesh_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"typedef struct lulesh_comm_backend_ops \{.*?\} lulesh_comm_backend_ops;", "",
                  text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s]*?[\s*]+(\w+)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while")))


def test_header_declares_comm_interface():
    names = declared_functions()
    for n in ["CommRecv", "CommSend", "CommSBN", "CommSyncPosVel", "CommMonoQ"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    import lulesh_amd
    lib = ctypes.CDLL(lulesh_amd.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    import lulesh_amd
    assert set(declared_functions()) == set(lulesh_amd.SIGNATURES)


def test_status_strings_without_gpu():
    import lulesh_amd
    lulesh_amd.load()
    assert lulesh_amd.status_string(-1) == "VolumeError"
    assert lulesh_amd.status_string(-2) == "QStopError"
