"""ctypes binding of the CPU oracle (test infrastructure only).

Loads oracle/build/liblulesh_oracle.so (built by ``make -C oracle``).  Only
tests/, the smoke check and bench.py's CPU-baseline leg use this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liblulesh_oracle.so")
_lib = None

CBRT_LIBM = 0
CBRT_EXACT = 1

# field ids (same numbering as include/lulesh_amd.h)
FIELDS = {
    "x": 0, "y": 1, "z": 2, "xd": 3, "yd": 4, "zd": 5, "xdd": 6, "ydd": 7, "zdd": 8,
    "fx": 9, "fy": 10, "fz": 11, "nodalMass": 12,
    "e": 16, "p": 17, "q": 18, "ql": 19, "qq": 20, "v": 21, "volo": 22, "delv": 23,
    "vdov": 24, "arealg": 25, "ss": 26, "elemMass": 27, "delv_xi": 28, "delv_eta": 29,
    "delv_zeta": 30, "delx_xi": 31, "delx_eta": 32, "delx_zeta": 33, "vnew": 34,
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise FileNotFoundError(f"{_LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(_LIB_PATH)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        sigs = {
            "lo_create": (vp, [i, i, i, i, i, i]),
            "lo_destroy": (None, [vp]),
            "lo_run": (i, [vp, i]),
            "lo_cycle": (i, [vp]),
            "lo_time": (d, [vp]),
            "lo_deltatime": (d, [vp]),
            "lo_stoptime": (d, [vp]),
            "lo_num_elem": (i, [vp]),
            "lo_num_node": (i, [vp]),
            "lo_num_domains": (i, [vp]),
            "lo_get_field": (i, [vp, i, i, dp]),
            "lo_set_field": (i, [vp, i, i, dp]),
            "lo_get_elem_rep": (i, [vp, i, ip]),
            "lo_get_region": (i, [vp, i, ip]),
            "lo_set_qstop": (None, [vp, d]),
            "lo_set_stoptime": (None, [vp, d]),
            "lo_verify": (None, [vp, i, dp, dp, dp, dp]),
            "lo_exchange": (i, [vp, i, i, ip]),
            "lo_com_buf_size": (i, [vp, i]),
            "lo_recv_offset": (i, [vp, i, i, i, i, i, i, i]),
            "lo_get_recv_buffer": (i, [vp, i, dp]),
            "lo_cbrt_exact": (d, [d]),
            "lo_create_mesh": (vp, [i, i, dp, dp, dp, ip, ip, ip, i, ip, i, ip, i, ip, i, ip, d,
                                    i, i, i, i]),
            "lo_get_elem_bc": (i, [vp, i, ip]),
            "lo_octant_corner": (i, [vp]),
            "lo_all_elem": (i, [vp]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


E_DEPOSIT = 3.948746e+7   # energy the .lmesh reader deposits (lulesh.cu:1049-1050)


def read_lmesh(path):
    """The reference's .lmesh reader (LULESH/cuda/src/lulesh.cu:855-970),
    restated: whitespace-separated numbers in the order numElem numNode,
    numNode coordinate triples, numElem x 8 node ids, numElem x 6 face
    neighbours (xi-, xi+, eta-, eta+, zeta-, zeta+), then four counted node
    sets: symmetry X, Y, Z and the free surface."""
    with open(path) as f:
        tok = f.read().split()
    pos = 0

    def take(n, conv):
        nonlocal pos
        if pos + n > len(tok):
            raise ValueError(f"{path}: truncated")
        out = tok[pos:pos + n]
        pos += n
        return np.array([conv(t) for t in out])

    ne, nn = (int(v) for v in take(2, int))
    xyz = take(3 * nn, float).reshape(nn, 3)
    nodelist = take(8 * ne, int).astype(np.int32).reshape(ne, 8)
    nbr = take(6 * ne, int).astype(np.int32).reshape(ne, 6)
    sets = []
    for _ in range(4):
        cnt = int(take(1, int)[0])
        sets.append(take(cnt, int).astype(np.int32))
    return dict(num_elem=ne, num_node=nn, x=np.ascontiguousarray(xyz[:, 0]),
                y=np.ascontiguousarray(xyz[:, 1]), z=np.ascontiguousarray(xyz[:, 2]),
                nodelist=nodelist, nbr=nbr, symm=sets[:3], free=sets[3])


class Oracle:
    """tp^3 subdomains of nx^3 elements (tp=1: the plain problem), or one
    unstructured domain (Oracle.from_mesh)."""

    def __init__(self, nx, tp=1, nr=11, balance=1, cost=1, cbrt_mode=CBRT_LIBM, _handle=None):
        self._L = lib()
        self.h = _handle if _handle else self._L.lo_create(nx, tp, nr, balance, cost, cbrt_mode)
        if not self.h:
            raise ValueError("lo_create failed")
        self.nx, self.tp = nx, tp
        self.ndom = tp ** 3
        self.num_elem = self._L.lo_num_elem(self.h)
        self.num_node = self._L.lo_num_node(self.h)
        self.all_elem = self._L.lo_all_elem(self.h)

    @classmethod
    def from_mesh(cls, m, nr=11, balance=1, cost=1, cbrt_mode=CBRT_LIBM, elem_bc=None,
                  e_deposit=E_DEPOSIT):
        """One unstructured domain from a mesh dict (read_lmesh's layout)."""
        L = lib()
        keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a

        x, y, z = (arr(m[k], np.float64) for k in ("x", "y", "z"))
        nl = arr(m["nodelist"], np.int32)
        nb = arr(m["nbr"], np.int32)
        sx, sy, sz = (arr(s, np.int32) for s in m["symm"])
        fr = arr(m["free"], np.int32)
        bc = None if elem_bc is None else _ip(arr(elem_bc, np.int32))
        h = L.lo_create_mesh(int(m["num_elem"]), int(m["num_node"]), _dp(x), _dp(y), _dp(z), _ip(nl),
                             _ip(nb), _ip(sx), sx.size, _ip(sy), sy.size, _ip(sz), sz.size, _ip(fr),
                             fr.size, bc, float(e_deposit), nr, balance, cost, cbrt_mode)
        if not h:
            raise ValueError("lo_create_mesh rejected the mesh")
        return cls(0, 1, _handle=h)

    @property
    def octant_corner(self):
        return self._L.lo_octant_corner(self.h)

    def elem_bc(self, dom=0):
        out = np.empty(self.num_elem, dtype=np.int32)
        self._L.lo_get_elem_bc(self.h, dom, _ip(out))
        return out

    def close(self):
        if self.h:
            self._L.lo_destroy(self.h)
            self.h = None

    __del__ = close

    def run(self, max_cycles=-1):
        return self._L.lo_run(self.h, max_cycles)

    @property
    def cycle(self):
        return self._L.lo_cycle(self.h)

    @property
    def time(self):
        return self._L.lo_time(self.h)

    @property
    def deltatime(self):
        return self._L.lo_deltatime(self.h)

    def set_stoptime(self, t):
        self._L.lo_set_stoptime(self.h, t)

    def set_qstop(self, q):
        self._L.lo_set_qstop(self.h, q)

    def _flen(self, fid):
        if fid < 16:
            return self.num_node
        if fid in (28, 29, 30):
            return self.all_elem
        return self.num_elem

    def field(self, name, dom=0):
        fid = FIELDS[name]
        out = np.empty(self._flen(fid), dtype=np.float64)
        n = self._L.lo_get_field(self.h, dom, fid, _dp(out))
        if n < 0:
            raise KeyError(name)
        return out

    def set_field(self, name, values, dom=0):
        fid = FIELDS[name]
        a = np.ascontiguousarray(values, dtype=np.float64)
        assert a.size == self._flen(fid)
        self._L.lo_set_field(self.h, dom, fid, _dp(a))

    def elem_rep(self, dom=0):
        out = np.empty(self.num_elem, dtype=np.int32)
        self._L.lo_get_elem_rep(self.h, dom, _ip(out))
        return out

    def region(self, dom=0):
        out = np.empty(self.num_elem, dtype=np.int32)
        self._L.lo_get_region(self.h, dom, _ip(out))
        return out

    def verify(self, dom=0):
        v = [ctypes.c_double() for _ in range(4)]
        self._L.lo_verify(self.h, dom, *[ctypes.byref(x) for x in v])
        return dict(origin_energy=v[0].value, max_abs_diff=v[1].value,
                    total_abs_diff=v[2].value, max_rel_diff=v[3].value)

    def exchange(self, msg_type, fields):
        ids = np.array([FIELDS[f] for f in fields], dtype=np.int32)
        return self._L.lo_exchange(self.h, msg_type, len(ids), _ip(ids))

    def com_buf_size(self, dom):
        return self._L.lo_com_buf_size(self.h, dom)

    def recv_offset(self, dom, dc, dr, dp, xfer, do_recv, plane_only):
        return self._L.lo_recv_offset(self.h, dom, dc, dr, dp, xfer, int(do_recv), int(plane_only))

    def recv_buffer(self, dom):
        out = np.empty(self.com_buf_size(dom), dtype=np.float64)
        self._L.lo_get_recv_buffer(self.h, dom, _dp(out))
        return out


def cbrt_exact(x):
    return lib().lo_cbrt_exact(float(x))
