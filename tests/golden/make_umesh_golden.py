#!/usr/bin/env python3
"""Oracle values of the reference's larger unstructured mesh,
sedov21oct.lmesh (31 408 elements, 34 446 nodes; LULESH/cuda/src/), for the
GPU tests of the unstructured path.

    python tests/golden/make_umesh_golden.py --its 100 --cbrt 1    (~5 s)
    python tests/golden/make_umesh_golden.py --its -1 --cbrt 1     (~40 s)
    python tests/golden/make_umesh_golden.py --its -1 --cbrt 0     (~40 s)

The mesh is read from tests/golden/sedov21oct.lmesh.gz, a gzip copy of the
reference's file.  Writes (merges into) tests/golden/umesh_sedov21oct.json:
per run the status, cycle count, final time and deltatime, origin energy,
the plane-0 symmetry figures, and sha256 of every final node and element
field.  --cbrt 0 is the reference's libm cube root, --cbrt 1 the correctly
rounded one the GPU kernels use (bit-for-bit comparable).
"""
import argparse
import gzip
import hashlib
import json
import os
import shutil
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))
import pyoracle  # noqa: E402

FIXTURE = os.path.join(HERE, "sedov21oct.lmesh.gz")
OUT = os.path.join(HERE, "umesh_sedov21oct.json")
FIELDS = ("x", "y", "z", "xd", "yd", "zd", "nodalMass", "e", "p", "q", "v", "ss", "volo",
          "elemMass", "delv", "vdov", "arealg", "vnew", "delv_xi", "delv_eta", "delv_zeta",
          "delx_xi", "delx_eta", "delx_zeta")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def unpack(dst_dir):
    """The fixture decompressed into dst_dir; returns its path."""
    path = os.path.join(dst_dir, "sedov21oct.lmesh")
    with gzip.open(FIXTURE, "rb") as src, open(path, "wb") as dst:
        shutil.copyfileobj(src, dst)
    return path


def oracle_record(lmesh, its, cbrt):
    o = pyoracle.Oracle.from_mesh(pyoracle.read_lmesh(lmesh), cbrt_mode=cbrt)
    t0 = time.time()
    st = o.run(its)
    e0 = o.field("e")[o.octant_corner]
    return {"its": its, "cbrt_mode": cbrt, "status": st, "cycle": o.cycle, "time": o.time.hex(),
            "deltatime": o.deltatime.hex(), "origin_energy": e0.hex(), "printed_energy": "%12.6e" % e0,
            "elapsed_s": time.time() - t0, "verify": {k: v.hex() for k, v in o.verify().items()},
            "sha256": {f: sha(o.field(f)) for f in FIELDS}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--its", type=int, default=-1)
    ap.add_argument("--cbrt", type=int, default=1)
    ap.add_argument("--out", default=OUT)
    a = ap.parse_args()
    with tempfile.TemporaryDirectory() as d:
        rec = oracle_record(unpack(d), a.its, a.cbrt)
    db = json.load(open(a.out)) if os.path.exists(a.out) else {}
    db[f"its{a.its}_cbrt{a.cbrt}"] = rec
    json.dump(db, open(a.out, "w"), indent=1, sort_keys=True)
    print(rec["printed_energy"], rec["cycle"], rec["status"])


if __name__ == "__main__":
    main()
