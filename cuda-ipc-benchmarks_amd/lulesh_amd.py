"""ctypes binding of liblulesh_amd.so (the C ABI of include/lulesh_amd.h).

Used by the tests, the benchmark and the smoke check.  There is no fallback:
if the HIP library is missing or no GPU is present, loading or initialising
raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblulesh_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "lulesh_amd.h")

OK, VOLUME_ERROR, QSTOP_ERROR, EINVAL, EHIP, ENOMEM, EDIRECTION, EBUSY = 0, -1, -2, -3, -4, -5, -6, -7
EIO = -8
MSG_COMM_SBN, MSG_SYNC_POS_VEL, MSG_MONOQ = 1024, 2048, 3072

FIELDS = {
    "x": 0, "y": 1, "z": 2, "xd": 3, "yd": 4, "zd": 5, "fx": 9, "fy": 10, "fz": 11,
    "nodalMass": 12, "e": 16, "p": 17, "q": 18, "v": 21, "volo": 22, "delv": 23, "vdov": 24,
    "arealg": 25, "ss": 26, "elemMass": 27, "delv_xi": 28, "delv_eta": 29, "delv_zeta": 30,
    "delx_xi": 31, "delx_eta": 32, "delx_zeta": 33, "vnew": 34,
}

_vp, _i, _d, _sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_size_t
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_pvp = ctypes.POINTER(ctypes.c_void_p)

# name -> (restype, argtypes); every extern "C" symbol of the header
SIGNATURES = {
    "lulesh_init": (_i, [_i]),
    "lulesh_device_cus": (_i, []),
    "lulesh_status_string": (ctypes.c_char_p, [_i]),
    "lulesh_version": (ctypes.c_char_p, []),
    "lulesh_group_create": (_i, [_i, _i, _i, _i, _i, _pvp]),
    "lulesh_mesh_read": (_i, [ctypes.c_char_p, _pvp]),
    "lulesh_mesh_create": (_i, [_i, _i, _dp, _dp, _dp, _ip, _ip, _ip, _i, _ip, _i, _ip, _i, _ip, _i, _pvp]),
    "lulesh_mesh_destroy": (_i, [_vp]),
    "lulesh_mesh_sizes": (_i, [_vp, _ip, _ip]),
    "lulesh_mesh_get_elem_bc": (_i, [_vp, _ip, _sz]),
    "lulesh_mesh_set_elem_bc": (_i, [_vp, _ip, _sz]),
    "lulesh_mesh_octant_corner": (_i, [_vp]),
    "lulesh_mesh_set_energy": (_i, [_vp, _d]),
    "lulesh_group_create_mesh": (_i, [_vp, _i, _i, _i, _pvp]),
    "lulesh_group_destroy": (_i, [_vp]),
    "lulesh_group_num_domains": (_i, [_vp]),
    "lulesh_group_domain": (_i, [_vp, _i, _pvp]),
    "lulesh_group_set_stoptime": (_i, [_vp, _d]),
    "lulesh_group_set_qstop": (_i, [_vp, _d]),
    "lulesh_group_set_max_cycles": (_i, [_vp, _i]),
    "lulesh_group_set_keep_forces": (_i, [_vp, _i]),
    "lulesh_group_comm_teardown": (_i, [_vp]),
    "lulesh_group_phase_times": (_i, [_vp, _dp]),
    "lulesh_domain_write_solution": (_i, [_vp, ctypes.c_char_p]),
    "lulesh_group_run": (_i, [_vp, _i, _vp]),
    "lulesh_group_enqueue": (_i, [_vp, _i, _vp]),
    "lulesh_group_sync": (_i, [_vp, _ip, _dp, _dp, _vp]),
    "lulesh_group_verify": (_i, [_vp, _i, _dp]),
    "lulesh_group_profile": (_i, [_vp, _i, _dp]),
    "lulesh_group_force_plan": (_i, [_vp, _ip, _ip, _ip]),
    "lulesh_batch_profile": (_i, [_vp, _i, _dp]),
    "lulesh_domain_rank": (_i, [_vp]),
    "lulesh_domain_sizes": (_i, [_vp, _ip, _ip, _ip, _ip]),
    "lulesh_domain_get_field": (_i, [_vp, _i, _dp, _sz]),
    "lulesh_domain_set_field": (_i, [_vp, _i, _dp, _sz]),
    "lulesh_domain_field_ptr": (_i, [_vp, _i, _pvp]),
    "lulesh_domain_get_elem_rep": (_i, [_vp, _ip, _sz]),
    "lulesh_domain_get_region": (_i, [_vp, _ip, _sz]),
    "CommRecv": (_i, [_vp, _i, _i, _i, _i, _i, _i, _i]),
    "CommSend": (_i, [_vp, _i, _i, _ip, _i, _i, _i, _i, _i, _vp]),
    "CommSBN": (_i, [_vp, _i, _ip, _vp]),
    "CommSyncPosVel": (_i, [_vp, _vp]),
    "CommMonoQ": (_i, [_vp, _vp]),
    "lulesh_comm_buf_size": (_i, [_vp]),
    "lulesh_comm_recv_offset": (_i, [_vp, _i, _i, _i, _i, _i, _i]),
    "lulesh_comm_get_recv_buffer": (_i, [_vp, _dp, _sz]),
    "lulesh_shm_recv_offset": (_i, [_i, _i, _i, _i, _i, _i, _i, _i, _i]),
    "lulesh_comm_buf_size_of": (_i, [_i, _i, _i]),
    "lulesh_region_layout": (_i, [_i, _i, _i, _i, _i, _i, _ip, _ip, _sz]),
    "lulesh_force_plan_of": (_i, [_i, _ip, _ip, _i, _i, _ip, _ip, _ip, _sz]),
    "lulesh_domain_get_eos_table": (_i, [_vp, ctypes.POINTER(ctypes.c_uint16), _sz]),
    "lulesh_device_local_backend": (_vp, []),
    "lulesh_device_copy_backend": (_vp, []),
    "lulesh_comm_apply_corner": (_i, [_vp, _i, _dp, _i, _ip, _i, _vp]),
    "lulesh_group_set_comm_backend": (_i, [_vp, _vp, _vp]),
    "lulesh_group_exchange": (_i, [_vp, _i, _i, _ip, _vp]),
    "lulesh_group_exchange_noalloc": (_i, [_vp, _i, _i, _ip, _vp]),
    "lulesh_batch_create": (_i, [_i, _ip, _ip, _ip, _ip, _pvp]),
    "lulesh_batch_destroy": (_i, [_vp]),
    "lulesh_batch_set_max_cycles": (_i, [_vp, _i]),
    "lulesh_batch_run": (_i, [_vp, _i, _vp]),
    "lulesh_batch_group": (_i, [_vp, _i, _pvp]),
    "lulesh_batch_results": (_i, [_vp, _ip, _ip]),
    "lulesh_batch_set_tiles": (_i, [_vp, _i, _i]),
    "lulesh_batch_enqueue": (_i, [_vp, _i]),
    "lulesh_batch_synchronize": (_i, [_vp]),
    "lulesh_batch_set_kchunk": (_i, [_vp, _i]),
    "lulesh_batch_autotune": (_i, [_vp, _i, _ip, _ip, _ip, _dp]),
    "lulesh_server_create": (_i, [_sz, _i, _pvp]),
    "lulesh_server_destroy": (_i, [_vp]),
    "lulesh_server_submit": (_i, [_vp, _i, _i, _i, _i, _d]),
    "lulesh_server_poll": (_i, [_vp, _i, _vp]),
    "lulesh_server_result": (_i, [_vp, _i, _ip, _ip, _dp, _dp]),
    "lulesh_server_stats": (_i, [_vp, _ip, _ip, _ip, ctypes.POINTER(_sz)]),
    "lulesh_server_now": (_d, [_vp]),
    "lulesh_server_times": (_i, [_vp, _i, _dp, _dp]),
    "lulesh_probe_cbrt": (_i, [_dp, _dp, _i]),
    "lulesh_probe_copy": (_d, [_sz, _i]),
    "lulesh_probe_sqrt_div": (_i, [_dp, _dp, _dp, _i]),
    "lulesh_probe_range_math": (_i, [_dp, _dp, _i]),
    "lulesh_probe_fp64": (_i, [_i, _i, _i, _dp, _dp]),
}

_lib = None
_inited = False


class LuleshError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        super().__init__(f"{what}: status {code} ({status_string(code)})")


def load():
    """Loads the library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} not built: run `make -C cuda-ipc-benchmarks_amd`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def status_string(code):
    return load().lulesh_status_string(code).decode()


def init(device=0):
    """Selects the GPU; raises if it is absent or not gfx950."""
    global _inited
    if not _inited:
        rc = load().lulesh_init(device)
        if rc != 0:
            raise LuleshError(rc, "lulesh_init")
        _inited = True


def _chk(rc, what):
    if rc < 0:
        raise LuleshError(rc, what)
    return rc


def _p(a, t=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(t))


class Domain:
    def __init__(self, group, handle):
        self.group = group
        self.h = handle
        L = load()
        v = [ctypes.c_int() for _ in range(4)]
        _chk(L.lulesh_domain_sizes(handle, *[ctypes.byref(x) for x in v]), "sizes")
        self.nx, self.num_elem, self.num_node, self.all_elem = [x.value for x in v]
        self.rank = L.lulesh_domain_rank(handle)

    def _len(self, fid):
        if fid < 16:
            return self.num_node
        if fid in (28, 29, 30):
            return self.all_elem
        return self.num_elem

    def field(self, name):
        fid = FIELDS[name]
        out = np.empty(self._len(fid), dtype=np.float64)
        _chk(load().lulesh_domain_get_field(self.h, fid, _p(out), out.size), f"get {name}")
        return out

    def set_field(self, name, values):
        fid = FIELDS[name]
        a = np.ascontiguousarray(values, dtype=np.float64)
        _chk(load().lulesh_domain_set_field(self.h, fid, _p(a), a.size), f"set {name}")

    def elem_rep(self):
        out = np.empty(self.num_elem, dtype=np.int32)
        _chk(load().lulesh_domain_get_elem_rep(self.h, _p(out, ctypes.c_int), out.size), "rep")
        return out

    def region(self):
        out = np.empty(self.num_elem, dtype=np.int32)
        _chk(load().lulesh_domain_get_region(self.h, _p(out, ctypes.c_int), out.size), "region")
        return out

    # the reference's comm interface
    def eos_table(self):
        """The EOS work table the EOS kernel walks (uint16 entries)."""
        L = load()
        out = np.empty(max(1, int(np.sum(self.elem_rep(), dtype=np.int64))), dtype=np.uint16)
        n = _chk(L.lulesh_domain_get_eos_table(self.h, _p(out, ctypes.c_uint16), out.size), "eos_table")
        return out[:n]

    def comm_recv(self, msg_type, xfer, dx, dy, dz, do_recv, plane_only):
        return load().CommRecv(self.h, msg_type, xfer, dx, dy, dz, int(do_recv), int(plane_only))

    def comm_send(self, msg_type, fields, dx, dy, dz, do_send, plane_only):
        ids = np.array([FIELDS[f] for f in fields], dtype=np.int32)
        return load().CommSend(self.h, msg_type, len(ids), _p(ids, ctypes.c_int), dx, dy, dz,
                               int(do_send), int(plane_only), None)

    def comm_sbn(self, fields):
        ids = np.array([FIELDS[f] for f in fields], dtype=np.int32)
        return load().CommSBN(self.h, len(ids), _p(ids, ctypes.c_int), None)

    def comm_sync_pos_vel(self):
        return load().CommSyncPosVel(self.h, None)

    def comm_monoq(self):
        return load().CommMonoQ(self.h, None)

    def com_buf_size(self):
        return load().lulesh_comm_buf_size(self.h)

    def recv_offset(self, dc, dr, dp, xfer, do_recv, plane_only):
        return load().lulesh_comm_recv_offset(self.h, dc, dr, dp, xfer, int(do_recv), int(plane_only))

    def recv_buffer(self):
        out = np.empty(self.com_buf_size(), dtype=np.float64)
        _chk(load().lulesh_comm_get_recv_buffer(self.h, _p(out), out.size), "recv buffer")
        return out


class Mesh:
    """An unstructured hexahedral mesh on the host (the reference's .lmesh
    input).  Needs no GPU: reading, checking and the boundary derivation
    run on the host."""

    def __init__(self, handle):
        self.h = handle
        ne, nn = ctypes.c_int(), ctypes.c_int()
        _chk(load().lulesh_mesh_sizes(handle, ctypes.byref(ne), ctypes.byref(nn)), "mesh_sizes")
        self.num_elem, self.num_node = ne.value, nn.value

    @classmethod
    def read(cls, path):
        h = ctypes.c_void_p()
        _chk(load().lulesh_mesh_read(str(path).encode(), ctypes.byref(h)), f"mesh_read {path}")
        return cls(h.value)

    @classmethod
    def from_arrays(cls, m):
        """m: dict with num_elem, num_node, x, y, z, nodelist (ne x 8), nbr
        (ne x 6), symm (three node lists) and free (node list)."""
        a = {k: np.ascontiguousarray(m[k], dtype=np.float64) for k in ("x", "y", "z")}
        nl = np.ascontiguousarray(m["nodelist"], dtype=np.int32)
        nb = np.ascontiguousarray(m["nbr"], dtype=np.int32)
        sx, sy, sz = (np.ascontiguousarray(v, dtype=np.int32) for v in m["symm"])
        fr = np.ascontiguousarray(m["free"], dtype=np.int32)
        ip = lambda v: _p(v, ctypes.c_int)
        h = ctypes.c_void_p()
        _chk(load().lulesh_mesh_create(int(m["num_elem"]), int(m["num_node"]), _p(a["x"]), _p(a["y"]),
                                       _p(a["z"]), ip(nl), ip(nb), ip(sx), sx.size, ip(sy), sy.size,
                                       ip(sz), sz.size, ip(fr), fr.size, ctypes.byref(h)), "mesh_create")
        return cls(h.value)

    def close(self):
        if self.h:
            load().lulesh_mesh_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def elem_bc(self):
        out = np.empty(self.num_elem, dtype=np.int32)
        _chk(load().lulesh_mesh_get_elem_bc(self.h, _p(out, ctypes.c_int), out.size), "elem_bc")
        return out

    def set_elem_bc(self, bc):
        a = np.ascontiguousarray(bc, dtype=np.int32)
        _chk(load().lulesh_mesh_set_elem_bc(self.h, _p(a, ctypes.c_int), a.size), "set_elem_bc")

    @property
    def octant_corner(self):
        return load().lulesh_mesh_octant_corner(self.h)

    def set_energy(self, e):
        _chk(load().lulesh_mesh_set_energy(self.h, float(e)), "set_energy")


def block_mesh_arrays(nx):
    """The Sedov block of -s nx (LULESH/omp_4.0/lulesh-init.cc:189-238) as
    unstructured-mesh arrays for Mesh.from_arrays: the block's numbering
    and coordinates, boundary faces naming their own element, symmetry sets
    on the x, y, z = 0 planes, the far faces as free surface."""
    en = nx + 1
    idx = np.arange(en)
    k, j, i = np.meshgrid(idx, idx, idx, indexing="ij")
    coord = lambda a: np.ascontiguousarray((1.125 * a / float(nx)).ravel())
    e = np.arange(nx)
    ek, ej, ei = np.meshgrid(e, e, e, indexing="ij")
    n0 = (ek * en * en + ej * en + ei).ravel()
    corner = np.array([0, 1, en + 1, en, en * en, en * en + 1, en * en + en + 1, en * en + en])
    el = np.arange(nx ** 3).reshape(nx, nx, nx)
    nbr = np.empty((nx ** 3, 6), dtype=np.int32)
    for f, (axis, d) in enumerate([(2, -1), (2, 1), (1, -1), (1, 1), (0, -1), (0, 1)]):
        nb = np.roll(el, -d, axis=axis)
        edge = [slice(None)] * 3
        edge[axis] = -1 if d > 0 else 0
        nb[tuple(edge)] = el[tuple(edge)]
        nbr[:, f] = nb.ravel()
    ni, nj, nk = i.ravel(), j.ravel(), k.ravel()
    nid = np.arange(en ** 3, dtype=np.int32)
    return dict(num_elem=nx ** 3, num_node=en ** 3, x=coord(i), y=coord(j), z=coord(k),
                nodelist=(n0[:, None] + corner[None, :]).astype(np.int32), nbr=nbr,
                symm=[nid[ni == 0], nid[nj == 0], nid[nk == 0]],
                free=nid[(ni == nx) | (nj == nx) | (nk == nx)])


class Group:
    """tp^3 subdomains of nx^3 elements on the GPU (tp=1: the plain problem),
    or one unstructured domain (Group.from_mesh)."""

    @classmethod
    def from_mesh(cls, mesh, nr=11, balance=1, cost=1):
        init()
        h = ctypes.c_void_p()
        _chk(load().lulesh_group_create_mesh(mesh.h, nr, balance, cost, ctypes.byref(h)), "group_create_mesh")
        return cls(0, 1, handle=h.value)

    def __init__(self, nx, tp=1, nr=11, balance=1, cost=1, handle=None, owned=True):
        init()
        L = load()
        if handle is None:
            h = ctypes.c_void_p()
            _chk(L.lulesh_group_create(nx, tp, nr, balance, cost, ctypes.byref(h)), "group_create")
            handle = h.value
        self.h = handle
        self.owned = owned
        self.nx, self.tp = nx, tp
        self.ndom = L.lulesh_group_num_domains(self.h)
        self.domains = []
        for r in range(self.ndom):
            d = ctypes.c_void_p()
            _chk(L.lulesh_group_domain(self.h, r, ctypes.byref(d)), "domain")
            self.domains.append(Domain(self, d.value))

    def close(self):
        if self.h and self.owned:
            load().lulesh_group_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stoptime(self, t):
        _chk(load().lulesh_group_set_stoptime(self.h, t), "stoptime")

    def set_qstop(self, q):
        _chk(load().lulesh_group_set_qstop(self.h, q), "qstop")

    def set_max_cycles(self, m):
        _chk(load().lulesh_group_set_max_cycles(self.h, m), "max_cycles")

    def comm_teardown(self):
        _chk(load().lulesh_group_comm_teardown(self.h), "comm_teardown")

    def phase_times(self):
        """(comm set-up, window alloc, peer query, comm teardown) in ms."""
        out = np.zeros(4, dtype=np.float64)
        _chk(load().lulesh_group_phase_times(self.h, _p(out)), "phase_times")
        return tuple(float(v) for v in out)

    def write_solution(self, path, rank=0):
        _chk(load().lulesh_domain_write_solution(self.domains[rank].h, str(path).encode()), "write_solution")

    def set_keep_forces(self, keep=True):
        """Store fx..fz every cycle (one-subdomain groups skip it otherwise)."""
        _chk(load().lulesh_group_set_keep_forces(self.h, 1 if keep else 0), "keep_forces")

    def run(self, cycles_per_sync=0):
        """Runs to the stop condition; returns the status (0 / -1 / -2)."""
        rc = load().lulesh_group_run(self.h, cycles_per_sync, None)
        if rc not in (0, VOLUME_ERROR, QSTOP_ERROR):
            raise LuleshError(rc, "run")
        return rc

    def enqueue(self, ncycles):
        _chk(load().lulesh_group_enqueue(self.h, ncycles, None), "enqueue")

    def sync(self):
        c, t, dt = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        rc = load().lulesh_group_sync(self.h, ctypes.byref(c), ctypes.byref(t), ctypes.byref(dt), None)
        if rc not in (0, VOLUME_ERROR, QSTOP_ERROR):
            raise LuleshError(rc, "sync")
        return rc, c.value, t.value, dt.value

    def verify(self, rank=0):
        out = np.empty(4, dtype=np.float64)
        _chk(load().lulesh_group_verify(self.h, rank, _p(out)), "verify")
        return dict(origin_energy=out[0], max_abs_diff=out[1], total_abs_diff=out[2],
                    max_rel_diff=out[3])

    def field(self, name, rank=0):
        return self.domains[rank].field(name)

    def profile(self, ncycles):
        """Per-kernel-class milliseconds over `ncycles` real cycles."""
        ms = np.zeros(7, dtype=np.float64)
        _chk(load().lulesh_group_profile(self.h, ncycles, _p(ms)), "profile")
        return ms

    def force_plan(self):
        """(workgroups, longest run in node planes, paired) of the force kernel."""
        v = [ctypes.c_int() for _ in range(3)]
        _chk(load().lulesh_group_force_plan(self.h, *[ctypes.byref(x) for x in v]), "force_plan")
        return v[0].value, v[1].value, bool(v[2].value)

    def exchange(self, msg_type, fields):
        ids = np.array([FIELDS[f] for f in fields], dtype=np.int32)
        return load().lulesh_group_exchange(self.h, msg_type, len(ids), _p(ids, ctypes.c_int), None)

    def exchange_noalloc(self, msg_type, fields):
        """The same exchange step without allocating (scratch array)."""
        ids = np.array([FIELDS[f] for f in fields], dtype=np.int32)
        return load().lulesh_group_exchange_noalloc(self.h, msg_type, len(ids), _p(ids, ctypes.c_int), None)

    def set_comm_backend(self, name):
        """Installs one of the library's comm backends: "device_local"
        (remote pack) or "device_copy" (pack locally, one copy per message)."""
        L = load()
        ops = {"device_local": L.lulesh_device_local_backend,
               "device_copy": L.lulesh_device_copy_backend}[name]()
        _chk(L.lulesh_group_set_comm_backend(self.h, ops, None), "set_comm_backend")


class Batch:
    def __init__(self, nx, nr, balance, cost):
        init()
        L = load()
        n = len(nx)
        arrs = [np.ascontiguousarray(a, dtype=np.int32) for a in (nx, nr, balance, cost)]
        h = ctypes.c_void_p()
        _chk(L.lulesh_batch_create(n, *[_p(a, ctypes.c_int) for a in arrs], ctypes.byref(h)),
             "batch_create")
        self.h = h.value
        self.n = n
        self.nx = arrs[0]

    def close(self):
        if self.h:
            load().lulesh_batch_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_max_cycles(self, m):
        _chk(load().lulesh_batch_set_max_cycles(self.h, m), "max_cycles")

    def set_tiles(self, te, tn):
        _chk(load().lulesh_batch_set_tiles(self.h, te, tn), "tiles")

    def set_kchunk(self, k):
        _chk(load().lulesh_batch_set_kchunk(self.h, k), "kchunk")

    def autotune(self, probe_cycles=5):
        a, n, k = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ms = ctypes.c_double()
        _chk(load().lulesh_batch_autotune(self.h, probe_cycles, ctypes.byref(a), ctypes.byref(n),
                                          ctypes.byref(k), ctypes.byref(ms)), "autotune")
        return dict(tile_elems=a.value, tile_nodes=n.value, kchunk=k.value, ms_per_cycle=ms.value)

    def run(self, cycles_per_sync=0):
        return load().lulesh_batch_run(self.h, cycles_per_sync, None)

    def enqueue(self, n):
        _chk(load().lulesh_batch_enqueue(self.h, n), "batch_enqueue")

    def synchronize(self):
        _chk(load().lulesh_batch_synchronize(self.h), "batch_sync")

    def profile(self, ncycles):
        ms = np.zeros(7, dtype=np.float64)
        _chk(load().lulesh_batch_profile(self.h, ncycles, _p(ms)), "profile")
        return ms

    def results(self):
        c = np.empty(self.n, dtype=np.int32)
        s = np.empty(self.n, dtype=np.int32)
        _chk(load().lulesh_batch_results(self.h, _p(c, ctypes.c_int), _p(s, ctypes.c_int)), "results")
        return c, s

    def group(self, i):
        g = ctypes.c_void_p()
        _chk(load().lulesh_batch_group(self.h, i, ctypes.byref(g)), "batch_group")
        return Group(int(self.nx[i]), handle=g.value, owned=False)


class Server:
    def __init__(self, hbm_budget=0, max_slots=64):
        init()
        h = ctypes.c_void_p()
        _chk(load().lulesh_server_create(hbm_budget, max_slots, ctypes.byref(h)), "server_create")
        self.h = h.value

    def close(self):
        if self.h:
            load().lulesh_server_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def now(self):
        return load().lulesh_server_now(self.h)

    def submit(self, nx, nr, balance, cost, arrival):
        return _chk(load().lulesh_server_submit(self.h, nx, nr, balance, cost, arrival), "submit")

    def poll(self, cycles=16):
        return _chk(load().lulesh_server_poll(self.h, cycles, None), "poll")

    def result(self, rid):
        st, cy = ctypes.c_int(), ctypes.c_int()
        e, t = ctypes.c_double(), ctypes.c_double()
        _chk(load().lulesh_server_result(self.h, rid, ctypes.byref(st), ctypes.byref(cy),
                                         ctypes.byref(e), ctypes.byref(t)), "result")
        ta, td = ctypes.c_double(), ctypes.c_double()
        _chk(load().lulesh_server_times(self.h, rid, ctypes.byref(ta), ctypes.byref(td)), "times")
        return dict(status=st.value, cycles=cy.value, origin_energy=e.value, total_abs_diff=t.value,
                    t_admit=ta.value, t_done=td.value)

    def stats(self):
        q, a, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        b = ctypes.c_size_t()
        _chk(load().lulesh_server_stats(self.h, ctypes.byref(q), ctypes.byref(a), ctypes.byref(d),
                                        ctypes.byref(b)), "stats")
        return dict(queued=q.value, active=a.value, done=d.value, hbm_in_use=b.value)


# host-only layout queries (no GPU needed)
def shm_recv_offset(nx, tp, recv_rank, dc, dr, dp, xfer, do_recv, plane_only):
    """shmRecvOffset of a tp^3 decomposition of nx^3-element subdomains."""
    return load().lulesh_shm_recv_offset(nx, tp, recv_rank, dc, dr, dp, xfer, int(do_recv), int(plane_only))


def comm_buf_size_of(nx, tp, rank):
    return _chk(load().lulesh_comm_buf_size_of(nx, tp, rank), "comm_buf_size_of")


def region_layout(nx, tp=1, rank=0, nr=11, balance=1, cost=1):
    """(region, elemRep) of rank `rank`: CreateRegionIndexSets' layout."""
    ne = nx ** 3
    reg = np.empty(ne, dtype=np.int32)
    rep = np.empty(ne, dtype=np.int32)
    _chk(load().lulesh_region_layout(nx, tp, rank, nr, balance, cost, _p(reg, ctypes.c_int),
                                     _p(rep, ctypes.c_int), ne), "region_layout")
    return reg, rep


def device_cus():
    """CUs of the current device (the engine's planning input)."""
    init()
    return _chk(load().lulesh_device_cus(), "device_cus")


def force_plan_of(nx, tp=1, cus=256, kchunk=0):
    """(kchunk_used, paired, tiles): the force kernel's column tiles of an
    engine over groups of tp^3 subdomains of nx^3 elements (nx, tp: ints for
    one group, or sequences, one entry per group) on a device of `cus` CUs,
    as the engine plans them (host only).  tiles is an (n, 8) int32 array of
    subdomain (numbered over the groups), i0, j0, k0, ex, ey, flags, kn in
    launch order."""
    L = load()
    nxs = np.atleast_1d(np.asarray(nx, dtype=np.int32))
    tps = np.broadcast_to(np.asarray(tp, dtype=np.int32), nxs.shape).copy()
    ng = int(nxs.size)
    a, b = _p(nxs, ctypes.c_int), _p(tps, ctypes.c_int)
    n = _chk(L.lulesh_force_plan_of(ng, a, b, kchunk, cus, None, None, None, 0), "force_plan_of")
    tiles = np.zeros((max(n, 1), 8), dtype=np.int32)
    kc, pr = ctypes.c_int(), ctypes.c_int()
    m = _chk(L.lulesh_force_plan_of(ng, a, b, kchunk, cus, ctypes.byref(kc), ctypes.byref(pr),
                                    _p(tiles, ctypes.c_int), n), "force_plan_of")
    assert m == n
    return kc.value, bool(pr.value), tiles[:n]


def probe_cbrt(x):
    init()
    a = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(a)
    _chk(load().lulesh_probe_cbrt(_p(a), _p(out), a.size), "probe_cbrt")
    return out


def probe_copy_ms(n_doubles, reps=20):
    """Average ms of a device-to-device streaming copy of n doubles."""
    init()
    ms = load().lulesh_probe_copy(n_doubles, reps)
    if ms < 0:
        raise LuleshError(ENOMEM, "probe_copy")
    return ms


def probe_fp64(waves_per_simd, iters=20000, one_chain=False):
    """(TFLOP/s, MHz): the FP64 rate 16 multiplies and adds per loop trip
    and lane reach, as eight independent chains or as one dependent chain,
    with `waves_per_simd` waves on every SIMD and no memory traffic, and the
    mean shader clock over their loops."""
    init()
    tf, mhz = ctypes.c_double(), ctypes.c_double()
    _chk(load().lulesh_probe_fp64(int(waves_per_simd), 1 if one_chain else 0, int(iters), ctypes.byref(tf),
                                  ctypes.byref(mhz)), "probe_fp64")
    return tf.value, mhz.value


def probe_sqrt_div(a, b):
    init()
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = np.empty(2 * a.size, dtype=np.float64)
    _chk(load().lulesh_probe_sqrt_div(_p(a), _p(b), _p(out), a.size), "probe_sqrt_div")
    return out[0::2], out[1::2]


def probe_range_math(a):
    """The kernels' range-limited sqrt and reciprocal of a, on the GPU."""
    init()
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.empty(2 * a.size, dtype=np.float64)
    _chk(load().lulesh_probe_range_math(_p(a), _p(out), a.size), "probe_range_math")
    return out[0::2], out[1::2]
