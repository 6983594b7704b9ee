// lulesh_kernels.h -- host-side launchers of the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "lulesh_dev.h"

namespace lulesh {

hipError_t launch_time_increment(GroupCtl *const *groups, int ngroups, DomDev *doms, const int *dom0,
                                 hipStream_t s);
hipError_t launch_elem_force(bool umesh, const DomDev *d, const Tile *t, int nt, int bs, hipStream_t s);
hipError_t launch_node_gather(bool umesh, bool fuse, const DomDev *d, const Tile *t, int nt, int bs,
                              hipStream_t s);
// fused force + ordered gather (column tiles of at most 256 threads, each
// over its own run of node planes)
// advance: also the node update (old -> new node-state buffer) as each
// node plane completes; store_f: write fx..fz; ti: the cycle's time
// increment is computed in the kernel (no time-increment launch before it)
hipError_t launch_force_march(DomDev *d, const ColTile *t, int nt, bool advance, bool store_f, bool ti,
                              hipStream_t s);
hipError_t launch_node_update(bool umesh, const DomDev *d, const Tile *t, int nt, int bs, hipStream_t s);
// CommSBN summed directly from the neighbours' partial forces + node update
hipError_t launch_node_update_direct(const DomDev *d, const Tile *t, int nt, int bs, hipStream_t s);
// the CommSyncPosVel final writers' snapshots of the current node state
hipError_t launch_pv_snapshot(const DomDev *d, const Tile *t, int nt, int bs, hipStream_t s);
hipError_t launch_kinematics(bool umesh, const DomDev *d, const Tile *t, int nt, int bs, hipStream_t s);
// umesh: the unstructured-mesh variants (DomDev::u_*)
// te elements per tile (<= bs threads per workgroup)
hipError_t launch_qeos(bool umesh, const DomDev *d, const Tile *t, int nt, int te, int bs, hipStream_t s);
// gate: skip the exchange once the group has stopped (NULL: always run)
hipError_t launch_pack(const PackEnt *e, int n, const double *const *fptr, int nfields,
                       const GroupCtl *gate, hipStream_t s);
hipError_t launch_unpack(const UnpackTgt *t, int n, const UnpackSrc *src, double *const *fptr,
                         int nfields, int mode, const GroupCtl *gate, hipStream_t s);
// one corner message of nf words at src applied to node idx (add or copy);
// skipped once the gate's group has stopped
hipError_t launch_apply_corner(double *const f[6], int nf, int idx, const double *src, bool add,
                               const GroupCtl *gate, hipStream_t s);
// Sedov set-up of one subdomain on the device: coordinates (block only),
// volo, elemMass, v = 1, nodal masses
hipError_t launch_setup(const DomDev &D, bool umesh, hipStream_t s);
hipError_t launch_hg_coef(double *coef, const double *ss, const double *mass, const double *volo,
                          const double *v, int n, hipStream_t s);
// the messages of m, each copied device to device (one launch)
hipError_t launch_copy_msgs(const CopyMsgs &m, hipStream_t s);
hipError_t launch_probe_copy(const double *a, double *b, size_t n, hipStream_t s);
hipError_t launch_probe_cbrt(const double *in, double *out, int n, hipStream_t s);
// each group's CtlSnap (groups: device array of GroupCtl pointers; out:
// pinned host memory, ngroups entries)
hipError_t launch_ctl_snapshot(GroupCtl *const *groups, int ngroups, CtlSnap *out, hipStream_t s);
hipError_t launch_probe_fp64(double *out, unsigned long long *clk, int grid, int iters, bool one_chain,
                             hipStream_t s);
hipError_t launch_probe_sqrt_div(const double *a, const double *b, double *out, int n, hipStream_t s);
hipError_t launch_probe_range_math(const double *a, double *out, int n, hipStream_t s);

}  // namespace lulesh
