#!/usr/bin/env python3
"""Oracle values of the 8-subdomain problem (2x2x2 subdomains of -s 45, the
reference's 8-rank halo-exchange configuration), for comparison with the
reference's published 8-rank results and for the GPU tests.

    python tests/golden/make_multidomain_golden.py --its 200            (~2 min)
    python tests/golden/make_multidomain_golden.py --its 200 --cbrt 1   (~2 min)
    python tests/golden/make_multidomain_golden.py --its -1             (~30 min)
    python tests/golden/make_multidomain_golden.py --its -1 --cbrt 1    (~30 min)

Writes (merges into) tests/golden/multidomain_s45.json, or the file named by
--out: per run the cycle count, final time, origin energy (rank 0), its
%12.6e print, the plane-0 symmetry figures of rank 0, and sha256 of the
final e, p, q, v, ss, x, y, z, xd, yd, zd of every subdomain.  --cbrt 0 is
the reference's libm cube root, --cbrt 1 the correctly rounded one the GPU
kernels use (bit-for-bit comparable).
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


DIGEST_FIELDS = ("e", "p", "q", "v", "ss", "x", "y", "z", "xd", "yd", "zd")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--its", type=int, default=200)
    ap.add_argument("--nx", type=int, default=45)
    ap.add_argument("--cbrt", type=int, default=0)
    ap.add_argument("--out", default=os.path.join(HERE, "multidomain_s45.json"))
    a = ap.parse_args()
    o = pyoracle.Oracle(a.nx, 2, 11, 1, 1, a.cbrt)
    t0 = time.time()
    st = o.run(a.its)
    rec = {"nx": a.nx, "tp": 2, "its": a.its, "cbrt_mode": a.cbrt, "status": st, "cycle": o.cycle,
           "time": o.time.hex(), "origin_energy": o.field("e")[0].hex(),
           "printed_energy": "%12.6e" % o.field("e")[0], "elapsed_s": time.time() - t0,
           "verify": {k: v.hex() for k, v in o.verify(0).items()},
           "sha256": {f"{f}{r}": sha(o.field(f, r)) for r in range(8) for f in DIGEST_FIELDS}}
    path = a.out
    db = json.load(open(path)) if os.path.exists(path) else {}
    db[f"nx{a.nx}_its{a.its}_cbrt{a.cbrt}"] = rec
    json.dump(db, open(path, "w"), indent=1, sort_keys=True)
    print(rec["printed_energy"], rec["cycle"])


if __name__ == "__main__":
    main()
