#!/usr/bin/env python3
"""Extracts the reference's halo-exchange layout into a data fixture.

Run in the build container, where the reference tree is present:

    python tests/golden/extract_comm_layout.py

Reads, mechanically and independently of this library's own tables:
  * the 26 receiver-offset entries of shmRecvOffset
    (LULESH/cuda/src/comm/comm_backend.h:75-138): kind, direction, doRecv gate;
  * CommRecv's receive posting (LULESH/omp_4.0/lulesh-comm.cc:59-353):
    direction (from the fromRank expression), gate, message length;
  * CommSend's packs (lulesh-comm.cc:357-844): direction (toRank), gate and,
    per message, which sender index lands at which wire position;
  * the unpacks of CommSBN (:848-1258, add), CommSyncPosVel (:1261-1680,
    copy) and CommMonoQ (:1684-1837, copy into the ghost planes): order,
    gate and which wire position lands on which receiver index;
  * MAX_FIELDS_PER_MPI_COMM and CACHE_COHERENCE_PAD_REAL (lulesh.h).

Every index map is evaluated here, for cubic subdomains of a few sizes, into
plain integer lists; the fixture (tests/golden/comm_layout.json) holds only
those lists, directions and flags.  The script cross-checks the sources
against each other (shmRecvOffset's order is CommRecv's; each unpack walks
the posting order; every send has the length its receive posts; the
condition flags name the direction the rank arithmetic names) and fails if
they disagree.
"""
import json
import os
import re
import sys

REF = "/root/reference/LULESH"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "comm_layout.json")

FLAG_DIR = {"colMin": (0, -1), "colMax": (0, 1), "rowMin": (1, -1), "rowMax": (1, 1),
            "planeMin": (2, -1), "planeMax": (2, 1)}
GATES = ("doRecv", "doSend")
SIZES = (2, 3, 4)     # elements per subdomain edge the index maps are evaluated at


def read(path):
    with open(path) as f:
        return f.read()


def strip_comments(s):
    """C/C++ source without its comments (string literals kept)"""
    out, i, n = [], 0, len(s)
    while i < n:
        if s.startswith("//", i):
            j = s.find("\n", i)
            i = n if j < 0 else j
        elif s.startswith("/*", i):
            j = s.find("*/", i + 2)
            i = n if j < 0 else j + 2
            out.append(" ")
        elif s[i] == '"':
            j = i + 1
            while j < n and s[j] != '"':
                j += 2 if s[j] == "\\" else 1
            out.append(s[i:j + 1])
            i = j + 1
        else:
            out.append(s[i])
            i += 1
    return "".join(out)


def match_close(s, i, op, cl):
    """index just past the bracket closing the one at s[i]"""
    assert s[i] == op
    depth = 0
    for k in range(i, len(s)):
        if s[k] == op:
            depth += 1
        elif s[k] == cl:
            depth -= 1
            if depth == 0:
                return k + 1
    raise ValueError("unbalanced")


def function_body(src, name):
    m = re.search(r"\n(?:void|static inline Index_t)\s+" + name + r"\s*\(", src)
    if not m:
        raise SystemExit(f"{name} not found")
    i = src.index("{", match_close(src, src.index("(", m.start()), "(", ")"))
    return src[i:match_close(src, i, "{", "}")]


def flag_blocks(body):
    """(condition, block text, start) of every `if (<flags>) { ... }` whose
    condition uses only the six neighbour flags and the doRecv/doSend gates"""
    out = []
    for m in re.finditer(r"\bif\s*\(", body):
        p0 = m.end() - 1
        p1 = match_close(body, p0, "(", ")")
        cond = body[p0 + 1:p1 - 1].strip()
        words = re.findall(r"[A-Za-z_]\w*", cond)
        if not words or any(w not in FLAG_DIR and w not in GATES for w in words):
            continue
        rest = body[p1:].lstrip()
        if not rest.startswith("{"):
            continue
        b0 = body.index("{", p1)
        out.append((cond, body[b0:match_close(body, b0, "{", "}")], m.start()))
    return out


def cond_info(cond):
    """direction named by the flags, the gate, and whether it is a group guard"""
    words = re.findall(r"[A-Za-z_]\w*", cond)
    d = [0, 0, 0]
    for w in words:
        if w in FLAG_DIR:
            ax, sg = FLAG_DIR[w]
            d[ax] = sg
    gate = [w for w in words if w in GATES]
    return tuple(d), (gate[0] if gate else None), ("|" in cond.replace("||", ""))


def rank_dir(expr):
    """(dc, dr, dp) of `myRank <expr>`: terms tp*tp (plane), tp (row), 1 (col)"""
    e = expr.replace(" ", "").replace("domain.tp()", "T")
    d = [0, 0, 0]
    for sign, term in re.findall(r"([+-])(T\*T|T|1)", e):
        s = 1 if sign == "+" else -1
        d[{"T*T": 2, "T": 1, "1": 0}[term]] += s
    rebuilt = "".join(f"{'+' if v > 0 else '-'}{t}" for v, t in ((d[2], "T*T"), (d[1], "T"), (d[0], "1")) if v)
    if sorted(rebuilt) != sorted(e):
        raise SystemExit(f"unparsed rank expression {expr!r}")
    return tuple(d)


def kind_of(off_expr):
    if "cmsg" in off_expr:
        return "corner"
    if "emsg" in off_expr:
        return "edge"
    return "face"


def defs_before(body, pos, name):
    """last `<name> = <expr> ;` before position pos"""
    ms = [m for m in re.finditer(r"\b" + name + r"\s*=\s*([^;]+);", body[:pos])]
    return ms[-1].group(1).strip() if ms else None


def loops(block):
    """(variable, bound) of the index loops, outermost first (not the field loop)"""
    out = []
    for m in re.finditer(r"for\s*\(\s*Index_t\s+(\w+)\s*=\s*0\s*;\s*\w+\s*<\s*([^;]+);", block):
        if m.group(1) != "fi":
            out.append((m.group(1), m.group(2).strip()))
    return out


def evaluate(lp, exprs, env):
    """index lists of each expression over the loop nest"""
    res = [[] for _ in exprs]

    def rec(k, e):
        if k == len(lp):
            for r, x in zip(res, exprs):
                r.append(int(eval(x, {"__builtins__": {}}, e)))
            return
        var, bound = lp[k]
        for v in range(int(eval(bound, {"__builtins__": {}}, e))):
            rec(k + 1, dict(e, **{var: v}))

    rec(0, dict(env))
    return res


def c_expr(x):
    x = x.strip()
    if not re.fullmatch(r"[\w\s\*\+\-\(\)]+", x):
        raise SystemExit(f"unexpected index expression {x!r}")
    return x


def parse_recv(body):
    msgs = []
    for cond, blk, _ in flag_blocks(body):
        d, gate, guard = cond_info(cond)
        if "MPI_Irecv" not in blk or guard:
            continue
        rd = rank_dir(re.search(r"fromRank\s*=\s*myRank([^;]*);", blk).group(1))
        if rd != d:
            raise SystemExit(f"CommRecv: flags {cond!r} name {d}, rank arithmetic {rd}")
        off = re.search(r"commDataRecv\[([^\]]*)\]", blk).group(1)
        call = blk[blk.index("MPI_Irecv"):]
        args = [a.strip() for a in call[call.index("(") + 1:].split(",")]
        count = args[1]
        if count.startswith("recvCount"):
            count = re.search(r"recvCount\s*=\s*([^;]+);", blk).group(1).strip()
        msgs.append({"dir": list(d), "gate": gate, "kind": kind_of(off),
                     "count": c_expr(count.replace("* xferFields", "").replace("xferFields", "1"))})
    return msgs


def parse_send(body, n):
    """per message: direction, gate, kind, and for field 0 the sender index
    at each wire position"""
    msgs = []
    for cond, blk, pos in flag_blocks(body):
        if "MPI_Isend" not in blk:
            continue
        d, gate, guard = cond_info(cond)
        if guard:
            continue
        m = re.search(r"toRank\s*=\s*myRank([^;]*);", blk)
        if m:
            rd = rank_dir(m.group(1))
        else:
            call = blk[blk.index("MPI_Isend"):]
            rd = rank_dir(re.search(r"myRank([^,]*),", call).group(1))
        if rd != d:
            raise SystemExit(f"CommSend: flags {cond!r} name {d}, rank arithmetic {rd}")
        off = re.search(r"commDataSend\[([^\]]*)\]", blk).group(1)
        env = {"dx": n, "dy": n, "dz": n}
        cnt = defs_before(body, pos, "sendCount")
        if cnt:
            env["sendCount"] = eval(cnt, {"__builtins__": {}}, env)
        kind = kind_of(off)
        if kind == "corner":
            im = re.search(r"Index_t\s+idx\s*=\s*([^;]+);", blk)
            idx = eval(c_expr(im.group(1)), {"__builtins__": {}}, env) if im else 0
            wire, src = [0], [int(idx)]
            assert re.search(r"comBuf\[fi\]\s*=\s*\(domain\.\*fieldData\[fi\]\)\((idx|0)\)", blk)
        else:
            a = re.search(r"destAddr\[([^\]]+)\]\s*=\s*\(domain\.\*src\)\((.*?)\)\s*;", blk)
            wire, src = evaluate(loops(blk), [c_expr(a.group(1)), c_expr(a.group(2))], env)
        order = sorted(range(len(wire)), key=lambda k: wire[k])
        if [wire[k] for k in order] != list(range(len(wire))):
            raise SystemExit("CommSend: wire positions are not a permutation")
        msgs.append({"dir": list(d), "gate": gate, "kind": kind, "src": [src[k] for k in order]})
    return msgs


def parse_unpack(body, n):
    """per message: direction, gate, kind, op, and for field 0 the receiver
    index of each wire position.  CommMonoQ writes at a running offset
    (fieldOffset, starting at numElem and advanced after each message it
    unpacks): its indices are relative to that offset, and `advance` is how
    far the message moves it."""
    msgs = []
    for cond, blk, pos in flag_blocks(body):
        d, gate, guard = cond_info(cond)
        if "commDataRecv" not in blk or guard:
            continue
        off = re.search(r"commDataRecv\[([^\]]*)\]", blk).group(1)
        kind = kind_of(off)
        env = {"dx": n, "dy": n, "dz": n}
        cnt = defs_before(body, pos, "opCount")
        if cnt and kind == "face":
            env["opCount"] = eval(cnt, {"__builtins__": {}}, env)
        if kind == "corner":
            im = re.search(r"Index_t\s+idx\s*=\s*([^;]+);", blk)
            idx = eval(c_expr(im.group(1)), {"__builtins__": {}}, env) if im else 0
            a = re.search(r"\(domain\.\*fieldData\[fi\]\)\((idx|0)\)\s*(\+?=)\s*comBuf\[fi\]", blk)
            op = a.group(2)
            wire, dst = [0], [int(idx)]
        else:
            a = re.search(r"\(domain\.\*dest\)\((.*?)\)\s*(\+?=)\s*srcAddr\[([^\]]+)\]\s*;", blk)
            op = a.group(2)
            dexpr = a.group(1)
            relative = "fieldOffset[fi]" in dexpr
            if relative:
                env["fo"] = 0
                dexpr = dexpr.replace("fieldOffset[fi]", "fo")
            dst, wire = evaluate(loops(blk), [c_expr(dexpr), c_expr(a.group(3))], env)
        order = sorted(range(len(wire)), key=lambda k: wire[k])
        if [wire[k] for k in order] != list(range(len(wire))):
            raise SystemExit("unpack: wire positions are not a permutation")
        m = {"dir": list(d), "gate": gate, "kind": kind,
             "op": "add" if op == "+=" else "copy", "dst": [dst[k] for k in order]}
        if kind != "corner" and relative:
            adv = re.search(r"fieldOffset\[fi\]\s*\+=\s*(\w+)\s*;", blk)
            m["ghost_relative"] = True
            m["advance"] = int(eval(adv.group(1), {"__builtins__": {}}, env)) if adv else 0
        msgs.append(m)
    return msgs


def parse_shm(body):
    out = []
    for m in re.finditer(r"SHM_(FACE|EDGE|CORNER)\(([^,]+),\s*(-?\d+),\s*(-?\d+),\s*(-?\d+)\)", body):
        kind, cond = m.group(1).lower(), m.group(2)
        d = (int(m.group(3)), int(m.group(4)), int(m.group(5)))
        fd, gate, _ = cond_info(cond)
        if fd != d:
            raise SystemExit(f"shmRecvOffset: flags {cond!r} name {fd}, entry says {d}")
        out.append({"dir": list(d), "gate": gate, "kind": kind})
    return out


def main():
    omp = strip_comments(read(os.path.join(REF, "omp_4.0", "lulesh-comm.cc")))
    hdr = strip_comments(read(os.path.join(REF, "omp_4.0", "lulesh.h")))
    shm_src = strip_comments(read(os.path.join(REF, "cuda", "src", "comm", "comm_backend.h")))
    consts = {
        "MAX_FIELDS_PER_MPI_COMM": int(re.search(r"#define\s+MAX_FIELDS_PER_MPI_COMM\s+(\d+)", hdr).group(1)),
        "CACHE_COHERENCE_PAD_REAL": eval(re.search(r"#define\s+CACHE_COHERENCE_PAD_REAL\s+\((\d+) / sizeof\(Real_t\)\)",
                                                   hdr).group(1)) // 8,
    }
    shm = parse_shm(function_body(shm_src, "shmRecvOffset"))
    recv = parse_recv(function_body(omp, "CommRecv"))
    if [(m["dir"], m["gate"] == "doRecv", m["kind"]) for m in recv] != \
            [(m["dir"], m["gate"] == "doRecv", m["kind"]) for m in shm]:
        raise SystemExit("shmRecvOffset and CommRecv post in different orders")
    counts = [m.pop("count") for m in recv]
    fx = {"source": "LULESH/cuda/src/comm/comm_backend.h:75-138, LULESH/omp_4.0/lulesh-comm.cc, "
                    "LULESH/omp_4.0/lulesh.h",
          "constants": consts, "shm_order": shm, "recv_post": recv, "sizes": {}}
    for nx in SIZES:
        e = {"send_nodes": parse_send(function_body(omp, "CommSend"), nx + 1),
             "send_elems": parse_send(function_body(omp, "CommSend"), nx),
             "sbn": parse_unpack(function_body(omp, "CommSBN"), nx + 1),
             "posvel": parse_unpack(function_body(omp, "CommSyncPosVel"), nx + 1),
             "monoq": parse_unpack(function_body(omp, "CommMonoQ"), nx)}
        # each unpack walks the posting order (faces only for MonoQ)
        for name in ("sbn", "posvel", "monoq"):
            got = [m["dir"] for m in e[name]]
            want = [m["dir"] for m in recv][:len(got)]
            if got != want:
                raise SystemExit(f"{name}: unpack order differs from CommRecv's posting order")
        if len(e["monoq"]) != 6 or len(e["sbn"]) != 26 or len(e["posvel"]) != 26:
            raise SystemExit("unexpected message counts")
        # per posted receive, the words per field it expects (node planes;
        # element planes for the faces of MonoQ), and the matching send's length
        e["recv_count_nodes"] = [int(eval(c, {"__builtins__": {}}, {"dx": nx + 1, "dy": nx + 1, "dz": nx + 1}))
                                 for c in counts]
        e["recv_count_elems"] = [int(eval(c, {"__builtins__": {}}, {"dx": nx, "dy": nx, "dz": nx}))
                                 for c in counts]
        for key, cnt in (("send_nodes", e["recv_count_nodes"]), ("send_elems", e["recv_count_elems"])):
            sent = {tuple(m["dir"]): len(m["src"]) for m in e[key]}
            for m, c in zip(recv, cnt):
                if sent[tuple(-v for v in m["dir"])] != c:
                    raise SystemExit(f"{key}: send toward {m['dir']} has a different length than its receive")
        for name in ("sbn", "posvel", "monoq"):
            for m, c in zip(e[name], e["recv_count_nodes" if name != "monoq" else "recv_count_elems"]):
                if len(m["dst"]) != c:
                    raise SystemExit(f"{name}: unpack of {m['dir']} has a different length than its receive")
        fx["sizes"][str(nx)] = e
    with open(OUT, "w") as f:
        json.dump(fx, f, separators=(",", ":"), sort_keys=True)
        f.write("\n")
    print(f"wrote {OUT}: {len(shm)} shm entries, {len(recv)} posted receives, sizes {SIZES}")


if __name__ == "__main__":
    sys.exit(main())
