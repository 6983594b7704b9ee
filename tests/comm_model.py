"""The reference's halo exchange, modelled from the extracted layout fixture.

tests/golden/comm_layout.json holds what tests/golden/extract_comm_layout.py
read out of the reference's own sources: the receive posting order and its
doRecv gates (CommRecv, lulesh-comm.cc:59-353, and shmRecvOffset,
cuda comm/comm_backend.h:75-138), each CommSend message's sender indices in
wire order (lulesh-comm.cc:357-844), and each unpack's receiver indices in
wire order and in unpack order (CommSBN :848-1258, CommSyncPosVel
:1261-1680, CommMonoQ :1684-1837).  This module replays an exchange step
from those lists alone -- it shares no table with the library or the
oracle -- so comparing either of them with it checks their message order,
wire layout and index maps against the reference.
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LAYOUT = os.path.join(HERE, "golden", "comm_layout.json")

_fx = None


def layout():
    global _fx
    if _fx is None:
        with open(LAYOUT) as f:
            _fx = json.load(f)
    return _fx


def position(rank, tp):
    return rank % tp, (rank // tp) % tp, rank // (tp * tp)


def has_neighbour(rank, tp, d):
    c = position(rank, tp)
    return all(0 <= c[a] + d[a] < tp for a in range(3))


def neighbour(rank, tp, d):
    return rank + d[0] + d[1] * tp + d[2] * tp * tp


def present(msg, rank, tp, flag):
    """message posted / sent / unpacked by `rank` given the doRecv/doSend flag"""
    return has_neighbour(rank, tp, msg["dir"]) and (msg["gate"] is None or flag)


def align(n, pad):
    return (n + pad - 1) & ~(pad - 1)


def sizes(nx):
    """maxPlaneSize, maxEdgeSize of SetupCommBuffers (lulesh-init.cc:321-323)"""
    pad = layout()["constants"]["CACHE_COHERENCE_PAD_REAL"]
    return align((nx + 1) ** 2, pad), align(nx + 1, pad)


def recv_offsets(nx, tp, rank, xfer, do_recv, plane_only):
    """{direction: offset} of every message `rank` expects, counted along the
    fixture's posting order"""
    fx = layout()
    pad = fx["constants"]["CACHE_COHERENCE_PAD_REAL"]
    mp, me = sizes(nx)
    out, cnt = {}, {"face": 0, "edge": 0, "corner": 0}
    for m in fx["recv_post"]:
        if plane_only and m["kind"] != "face":
            break
        if not present(m, rank, tp, do_recv):
            continue
        out[tuple(m["dir"])] = cnt["face"] * xfer * mp + cnt["edge"] * xfer * me + cnt["corner"] * pad
        cnt[m["kind"]] += 1
    return out


def com_buf_size(nx, tp, rank):
    """comBufSize (lulesh-init.cc:334-356): 6-field face and edge slots plus a
    padded slot per corner neighbour"""
    fx = layout()
    pad = fx["constants"]["CACHE_COHERENCE_PAD_REAL"]
    nf = fx["constants"]["MAX_FIELDS_PER_MPI_COMM"]
    mp, me = sizes(nx)
    n = {"face": 0, "edge": 0, "corner": 0}
    for m in fx["recv_post"]:
        if has_neighbour(rank, tp, m["dir"]):
            n[m["kind"]] += 1
    return n["face"] * mp * nf + n["edge"] * me * nf + n["corner"] * pad


def exchange(nx, tp, step, fields, order=None):
    """Replays one exchange step on per-rank field arrays (a list over ranks
    of lists over fields of 1-D float64 arrays, changed in place).

    step: "sbn" (add, doRecv = true), "posvel" (copy, doRecv = false) or
    "monoq" (copy into the ghost planes, faces only).  All sends are packed
    from the fields before any unpack, as the reference's MPI ordering
    gives.  order, if given, maps rank -> a permutation of that rank's
    unpack list (to show that a different order gives different results).
    Returns {rank: (receive buffer, mask of the words this step wrote)},
    laid out as the reference lays it out; words outside the mask keep
    whatever an earlier exchange left there.
    """
    fx = layout()
    sz = fx["sizes"][str(nx)]
    ndom = tp ** 3
    flag = step != "posvel"
    plane_only = step == "monoq"
    send = sz["send_elems" if plane_only else "send_nodes"]
    unpack = sz[step]
    nf = len(fields[0])
    # pack: every sender's messages, keyed by (sender, direction)
    packed = {}
    for s in range(ndom):
        for m in send:
            if plane_only and m["kind"] != "face":
                continue
            if present(m, s, tp, flag):
                src = np.asarray(m["src"])
                packed[(s, tuple(m["dir"]))] = [fields[s][f][src].copy() for f in range(nf)]
    bufs = {}
    for r in range(ndom):
        offs = recv_offsets(nx, tp, r, nf, flag, plane_only)
        buf = np.zeros(com_buf_size(nx, tp, r))
        mask = np.zeros(buf.size, dtype=bool)
        msgs = [m for m in unpack if present(m, r, tp, flag)]
        if order is not None and r in order:
            msgs = [msgs[i] for i in order[r]]
        ghost = nx ** 3   # CommMonoQ's running ghost offset
        for m in msgs:
            d = tuple(m["dir"])
            payload = packed[(neighbour(r, tp, d), tuple(-v for v in d))]
            dst = np.asarray(m["dst"])
            if m.get("ghost_relative"):
                dst = dst + ghost
                ghost += m["advance"]
            off = offs[d]
            for f in range(nf):
                buf[off + f * len(dst): off + (f + 1) * len(dst)] = payload[f]
                mask[off + f * len(dst): off + (f + 1) * len(dst)] = True
                if m["op"] == "add":
                    fields[r][f][dst] += payload[f]
                else:
                    fields[r][f][dst] = payload[f]
        bufs[r] = (buf, mask)
    return bufs


def unpack_count(nx, tp, rank, step):
    flag = step != "posvel"
    return sum(present(m, rank, tp, flag) for m in layout()["sizes"][str(nx)][step])
