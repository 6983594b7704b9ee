"""Unstructured-mesh helpers for the tests: the Sedov block written as an
unstructured mesh, renumberings, and a .lmesh writer.

The block's numbering, coordinates and boundary sets are those of the
structured set-up (LULESH/omp_4.0/lulesh-init.cc:189-238, 487-631); a face on
the mesh boundary names its own element as neighbour, as the .lmesh files of
the reference do (LULESH/cuda/src/sedov15oct.lmesh).
"""
import numpy as np

E_BASE = 3.948746e+7
GOLDEN_LMESH = "sedov15oct.lmesh"

# elemBC bits, LULESH/omp_4.0/lulesh.h:60-88
XI_M_SYMM, XI_M_FREE = 0x00001, 0x00002
XI_P_SYMM, XI_P_FREE = 0x00008, 0x00010
ETA_M_SYMM, ETA_M_FREE = 0x00040, 0x00080
ETA_P_SYMM, ETA_P_FREE = 0x00200, 0x00400
ZETA_M_SYMM, ZETA_M_FREE = 0x01000, 0x02000
ZETA_P_SYMM, ZETA_P_FREE = 0x08000, 0x10000
FREE_BITS = XI_M_FREE | XI_P_FREE | ETA_M_FREE | ETA_P_FREE | ZETA_M_FREE | ZETA_P_FREE


def einit(nx):
    """Structured deposit of -s nx, one domain (lulesh-init.cc:170-178)."""
    scale = (nx * 1) / 45.0
    return E_BASE * scale * scale * scale


def structured_mesh(nx):
    en = nx + 1
    idx = np.arange(en)
    k, j, i = np.meshgrid(idx, idx, idx, indexing="ij")
    x = (1.125 * i / float(nx)).ravel()
    y = (1.125 * j / float(nx)).ravel()
    z = (1.125 * k / float(nx)).ravel()
    e = np.arange(nx)
    ek, ej, ei = np.meshgrid(e, e, e, indexing="ij")
    n0 = (ek * en * en + ej * en + ei).ravel()
    off = np.array([0, 1, en + 1, en, en * en, en * en + 1, en * en + en + 1, en * en + en])
    nodelist = (n0[:, None] + off[None, :]).astype(np.int32)
    el = np.arange(nx ** 3).reshape(nx, nx, nx)   # [k][j][i]

    def shifted(axis, d):
        nb = np.roll(el, -d, axis=axis)
        edge = [slice(None)] * 3
        edge[axis] = -1 if d > 0 else 0
        nb[tuple(edge)] = el[tuple(edge)]
        return nb.ravel()

    nbr = np.stack([shifted(2, -1), shifted(2, 1), shifted(1, -1), shifted(1, 1),
                    shifted(0, -1), shifted(0, 1)], axis=1).astype(np.int32)
    ni, nj, nk = i.ravel(), j.ravel(), k.ravel()
    nid = np.arange(en ** 3, dtype=np.int32)
    symm = [nid[ni == 0], nid[nj == 0], nid[nk == 0]]
    free = nid[(ni == nx) | (nj == nx) | (nk == nx)]
    return dict(num_elem=nx ** 3, num_node=en ** 3, x=x, y=y, z=z, nodelist=nodelist, nbr=nbr,
                symm=symm, free=free)


def corner_mesh(nx, m):
    """The m^3 elements of the -s nx block nearest the origin, as a mesh of
    their own: the block's coordinates (spacing 1.125/nx), its symmetry
    planes, and free faces where the block continues.  Until the blast
    reaches those faces, a run on it leaves the corner in the state the
    full block's run leaves it."""
    c = structured_mesh(m)
    en = m + 1
    i = np.arange(en)
    k3, j3, i3 = np.meshgrid(i, i, i, indexing="ij")
    for name, idx in (("x", i3), ("y", j3), ("z", k3)):
        c[name] = (1.125 * idx / float(nx)).ravel()
    return c


CORNER_ELEM = ["e", "p", "q", "v", "ss", "volo", "elemMass", "delv", "vdov", "arealg", "vnew",
               "delx_xi", "delx_eta", "delx_zeta", "delv_xi", "delv_eta", "delv_zeta"]
CORNER_NODE = ["x", "y", "z", "xd", "yd", "zd"]


def corner_of(field, n, m):
    """The [k][j][i] < m corner of a field laid out n per edge."""
    return field[:n ** 3].reshape(n, n, n)[:m, :m, :m].ravel()


def renumber_nodes(m, perm):
    """Node n of m becomes node perm[n]."""
    perm = np.asarray(perm)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    out = dict(m)
    for c in ("x", "y", "z"):
        out[c] = np.ascontiguousarray(m[c][inv])
    out["nodelist"] = perm[m["nodelist"]].astype(np.int32)
    out["symm"] = [perm[s].astype(np.int32) for s in m["symm"]]
    out["free"] = perm[m["free"]].astype(np.int32)
    return out


def renumber_elems(m, perm):
    """Element e of m becomes element perm[e]."""
    perm = np.asarray(perm)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    out = dict(m)
    out["nodelist"] = np.ascontiguousarray(m["nodelist"][inv])
    out["nbr"] = perm[m["nbr"][inv]].astype(np.int32)
    return out


def write_lmesh(path, m):
    """The .lmesh text layout read by LULESH/cuda/src/lulesh.cu:855-976;
    coordinates with 17 significant digits, so they read back exactly."""
    with open(path, "w") as f:
        f.write(f"{m['num_elem']} {m['num_node']}\n")
        for a, b, c in zip(m["x"], m["y"], m["z"]):
            f.write(f"{a:.17g} {b:.17g} {c:.17g}\n")
        for row in m["nodelist"]:
            f.write(" ".join(str(int(v)) for v in row) + "\n")
        for row in m["nbr"]:
            f.write(" ".join(str(int(v)) for v in row) + "\n")
        for s in list(m["symm"]) + [m["free"]]:
            f.write(f"{len(s)}\n")
            for v in s:
                f.write(f"{int(v)}\n")
