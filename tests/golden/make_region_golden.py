#!/usr/bin/env python3
"""Region layout and EOS repetition counts of the compiled reference.

Run in the build container after `make -C oracle ref`:

    python tests/golden/make_region_golden.py

Runs oracle/_ref/lulesh_ref (the unmodified omp_4.0 sources; one cycle is
enough, the layout is fixed at set-up) with LULESH_REF_REGIONS set, so its
final-output hook (oracle/ref_dump.cc) writes regNumList and elemRep as the
reference built them (CreateRegionIndexSets, lulesh-init.cc:377-484, and
main, lulesh.cc:3171-3192).  Writes tests/golden/ref_regions.json: per
configuration the region sizes, sum of elemRep and sha256 of both int32
arrays; and tests/golden/ref_regions_s30.npz with the -s 30 arrays
themselves (for a readable diff if a digest ever disagrees).
"""
import hashlib
import json
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(REPO, "oracle", "_ref", "lulesh_ref")

# (nx, nr, balance, cost): the reference's default -s 30 problem, the -s 100
# headline problem, and uneven layouts of the batched ensemble's ranges
CONFIGS = [(30, 11, 1, 1), (100, 11, 1, 1), (24, 16, 4, 2), (40, 16, 2, 1), (20, 1, 1, 1),
           (33, 7, 4, 2), (64, 5, 1, 2)]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()


def run(nx, nr, b, c):
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "regions.bin")
        env = dict(os.environ, LULESH_REF_REGIONS=path, OMP_NUM_THREADS="4")
        subprocess.run([REF, "-s", str(nx), "-r", str(nr), "-b", str(b), "-c", str(c), "-i", "1"],
                       env=env, check=True, capture_output=True)
        a = np.fromfile(path, dtype="<i4")
    ne, nreg = int(a[0]), int(a[1])
    reg = a[2:2 + ne]
    rep = a[2 + ne:2 + 2 * ne]
    sizes = a[2 + 2 * ne:2 + 2 * ne + nreg]
    assert ne == nx ** 3 and nreg == nr
    return reg, rep, sizes


def main():
    out = {}
    for nx, nr, b, c in CONFIGS:
        reg, rep, sizes = run(nx, nr, b, c)
        key = f"-s {nx} -r {nr} -b {b} -c {c}"
        out[key] = {"nx": nx, "nr": nr, "balance": b, "cost": c,
                    "region_sizes": [int(v) for v in sizes], "elem_rep_sum": int(rep.sum()),
                    "sha256_regNumList": sha(reg), "sha256_elemRep": sha(rep)}
        if (nx, nr, b, c) == (30, 11, 1, 1):
            np.savez_compressed(os.path.join(HERE, "ref_regions_s30.npz"),
                                regNumList=reg.astype(np.uint8), elemRep=rep.astype(np.uint8))
        print(key, out[key]["elem_rep_sum"])
    with open(os.path.join(HERE, "ref_regions.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
