#!/usr/bin/env python3
"""Oracle values of the 8-subdomain problem (2x2x2 subdomains of -s 45, the
reference's 8-rank halo-exchange configuration), for comparison with the
reference's published 8-rank results and for the GPU tests.

    python tests/golden/make_multidomain_golden.py --its 200   (~3 min)
    python tests/golden/make_multidomain_golden.py --its -1    (~40 min)

Writes tests/golden/multidomain_s45.json: per run the cycle count, final
time, origin energy (rank 0), its %12.6e print, and sha256 of e, x, xd of
every subdomain.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))
import pyoracle  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--its", type=int, default=200)
    ap.add_argument("--nx", type=int, default=45)
    ap.add_argument("--cbrt", type=int, default=0)
    a = ap.parse_args()
    o = pyoracle.Oracle(a.nx, 2, 11, 1, 1, a.cbrt)
    t0 = time.time()
    st = o.run(a.its)
    rec = {"nx": a.nx, "tp": 2, "its": a.its, "cbrt_mode": a.cbrt, "status": st, "cycle": o.cycle,
           "time": o.time.hex(), "origin_energy": o.field("e")[0].hex(),
           "printed_energy": "%12.6e" % o.field("e")[0], "elapsed_s": time.time() - t0,
           "sha256": {f"{f}{r}": sha(o.field(f, r)) for r in range(8) for f in ("e", "x", "xd")}}
    path = os.path.join(HERE, "multidomain_s45.json")
    db = json.load(open(path)) if os.path.exists(path) else {}
    db[f"nx{a.nx}_its{a.its}_cbrt{a.cbrt}"] = rec
    json.dump(db, open(path, "w"), indent=1, sort_keys=True)
    print(rec["printed_energy"], rec["cycle"])


if __name__ == "__main__":
    main()
