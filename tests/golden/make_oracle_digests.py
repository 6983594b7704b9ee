#!/usr/bin/env python3
"""Writes tests/golden/oracle_digests.json: sha256 digests of the final
fields of the CPU oracle (oracle/, correctly rounded cube root) for the
headline mesh, so the GPU can be checked bit for bit at -s 100 without
shipping the 90 MB of fields.

The oracle in libm mode reproduces the compiled reference bit for bit at
these sizes (tests/golden/ref_omp.json holds the reference's own digests);
the correctly rounded mode differs from it only through cbrt's last bit.

    python tests/golden/make_oracle_digests.py              # -s 100 -i 200 (minutes)
    python tests/golden/make_oracle_digests.py --full       # adds the full run (about an hour)
    python tests/golden/make_oracle_digests.py --from-npz KEY NX ITS FILE.npz CYCLE TIMEHEX
        records fields saved earlier from the same oracle (numpy savez of
        e p q v ss x y z xd yd zd)
"""
import argparse
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))

FIELDS = ["e", "p", "q", "v", "ss", "x", "y", "z", "xd", "yd", "zd"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def record(nx, its, fields, cycle, time_hex):
    return {"nx": nx, "its": its, "cbrt": "exact", "cycle": cycle, "time": time_hex,
            "sha256": {f: sha(fields[f]) for f in FIELDS}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--full", action="store_true")
    ap.add_argument("--from-npz", nargs=6, metavar=("KEY", "NX", "ITS", "NPZ", "CYCLE", "TIMEHEX"))
    a = ap.parse_args()
    path = os.path.join(HERE, "oracle_digests.json")
    db = json.load(open(path)) if os.path.exists(path) else {}
    if a.from_npz:
        key, nx, its, npz, cyc, th = a.from_npz
        z = np.load(npz)
        db[key] = record(int(nx), int(its), {f: z[f] for f in FIELDS}, int(cyc), th)
    else:
        import pyoracle
        for nx, its in [(100, 200)] + ([(100, -1)] if a.full else []):
            o = pyoracle.Oracle(nx, cbrt_mode=pyoracle.CBRT_EXACT)
            assert o.run(its) == 0
            key = f"-s {nx}" + (f" -i {its}" if its >= 0 else " full")
            db[key] = record(nx, its, {f: o.field(f) for f in FIELDS}, o.cycle, o.time.hex())
            print(key, o.cycle, flush=True)
    json.dump(db, open(path, "w"), indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
