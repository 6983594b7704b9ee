#!/usr/bin/env python3
"""Writes the golden fixtures of the reference CPU code (run in the build
container, where /root/reference exists).

The reference is the unmodified LULESH/omp_4.0 sources compiled by
`make -C oracle ref` into oracle/_ref/lulesh_ref, whose final-output hook
(oracle/ref_dump.cc) dumps the final state at full precision.  For each
configuration this records the cycle count, final time and dt, the origin
energy, the plane-0 symmetry figures (evaluated in the reference's loop
order from the dumped energies), the sha256 of every dumped field, and for
the two smallest meshes the fields themselves.

    python tests/golden/make_golden.py            # small configurations
    python tests/golden/make_golden.py --s100     # adds -s 100 -i 200 (minutes)

The full-length -s 100 record was made from a separate run of the
reference (LULESH_REF_DUMP=dump.bin oracle/_ref/lulesh_ref -s 100 > out.txt):
    python tests/golden/make_golden.py --from-dump "-s 100 full" "-s 100" dump.bin out.txt
"""
import argparse
import hashlib
import json
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(REPO, "oracle", "_ref", "lulesh_ref")

ELEM_FIELDS = ["e", "p", "q", "v", "ss"]
NODE_FIELDS = ["x", "y", "z", "xd", "yd", "zd"]

CONFIGS = [
    "-s 4", "-s 8", "-s 8 -i 10", "-s 12 -r 5 -b 2 -c 2", "-s 16", "-s 16 -i 50",
    "-s 30", "-s 30 -i 100",
]


def read_dump(path):
    b = open(path, "rb").read()
    nx, cyc = struct.unpack("ii", b[:8])
    t, dt = struct.unpack("dd", b[8:24])
    ne, nn = nx ** 3, (nx + 1) ** 3
    a = np.frombuffer(b[24:], dtype=np.float64)
    out, off = {}, 0
    for f in ELEM_FIELDS:
        out[f] = a[off:off + ne].copy()
        off += ne
    for f in NODE_FIELDS:
        out[f] = a[off:off + nn].copy()
        off += nn
    return nx, cyc, t, dt, out


def symmetry(e, nx):
    """VerifyAndWriteFinalOutput figures (lulesh-util.cc:192-210), same loop
    order, plain IEEE doubles."""
    mad, tad, mrd = np.float64(0.0), np.float64(0.0), np.float64(0.0)
    ev = np.asarray(e[: nx * nx], dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        for j in range(nx):
            for k in range(j + 1, nx):
                ad = abs(ev[j * nx + k] - ev[k * nx + j])
                tad += ad
                if mad < ad:
                    mad = ad
                rd = ad / ev[k * nx + j]   # IEEE: x/0 -> inf, 0/0 -> nan
                if mrd < rd:
                    mrd = rd
    return float(mad), float(tad), float(mrd)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def run(args, threads):
    with tempfile.TemporaryDirectory() as td:
        dump = os.path.join(td, "d.bin")
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), LULESH_REF_DUMP=dump)
        out = subprocess.run([REF] + args.split(), env=env, capture_output=True, text=True,
                             check=True).stdout
        return record(args, dump, out)


def record(args, dump, out):
    """Golden record of one reference run from its dump file and stdout."""
    nx, cyc, t, dt, fields = read_dump(dump)
    mad, tad, mrd = symmetry(fields["e"], nx)
    kw = dict(zip(args.split()[::2], args.split()[1::2]))
    rec = {
        "args": args, "nx": nx, "its": int(kw.get("-i", -1)), "nr": int(kw.get("-r", 11)),
        "balance": int(kw.get("-b", 1)), "cost": int(kw.get("-c", 1)),
        "cycle": cyc, "time": t.hex(), "deltatime": dt.hex(),
        "origin_energy": float(fields["e"][0]).hex(),
        "origin_energy_dec": "%.17e" % fields["e"][0],
        "max_abs_diff": mad.hex(), "total_abs_diff": tad.hex(), "max_rel_diff": mrd.hex(),
        "printed_energy": [ln.split("=")[1].strip() for ln in out.splitlines()
                           if "Final Origin Energy" in ln][0],
        "sha256": {f: sha(v) for f, v in fields.items()},
        # tolerance fingerprints for runs whose fields differ by roundoff
        "sum": {f: float(np.sum(v)).hex() for f, v in fields.items()},
        "l2": {f: float(np.sqrt(np.sum(v * v))).hex() for f, v in fields.items()},
    }
    return rec, fields


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--s100", action="store_true")
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--from-dump", nargs=4, metavar=("KEY", "ARGS", "DUMP", "STDOUT"),
                    help="record an earlier run of the reference (e.g. the full -s 100 run, "
                         "about an hour on 8 cores) from its dump file and saved stdout")
    a = ap.parse_args()
    path = os.path.join(HERE, "ref_omp.json")
    if a.from_dump:
        key, args, dump, out = a.from_dump
        db = json.load(open(path))
        rec, _ = record(args, dump, open(out).read())
        db[key] = rec
        json.dump(db, open(path, "w"), indent=1, sort_keys=True)
        print(key, rec["cycle"], rec["origin_energy_dec"])
        return
    if not os.path.exists(REF):
        subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    path = os.path.join(HERE, "ref_omp.json")
    db = json.load(open(path)) if os.path.exists(path) else {}
    cfgs = list(CONFIGS) + (["-s 100 -i 200"] if a.s100 else [])
    for c in cfgs:
        rec, fields = run(c, a.threads)
        db[c] = rec
        if rec["nx"] <= 8 and rec["its"] < 0:
            np.savez_compressed(os.path.join(HERE, f"ref_fields_s{rec['nx']}.npz"), **fields)
        print(c, rec["cycle"], rec["origin_energy_dec"], flush=True)
    json.dump(db, open(path, "w"), indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
