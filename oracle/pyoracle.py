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


class Oracle:
    """tp^3 subdomains of nx^3 elements (tp=1: the plain problem)."""

    def __init__(self, nx, tp=1, nr=11, balance=1, cost=1, cbrt_mode=CBRT_LIBM):
        self._L = lib()
        self.h = self._L.lo_create(nx, tp, nr, balance, cost, cbrt_mode)
        if not self.h:
            raise ValueError("lo_create failed")
        self.nx, self.tp = nx, tp
        self.ndom = tp ** 3
        self.num_elem = self._L.lo_num_elem(self.h)
        self.num_node = self._L.lo_num_node(self.h)

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
            return self.num_elem + 6 * self.nx * self.nx
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
