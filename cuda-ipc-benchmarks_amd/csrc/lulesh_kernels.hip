// lulesh_kernels.hip -- HIP kernels of the LULESH Lagrange leapfrog step
// for gfx950 (MI355X).
//
// One cycle = TimeIncrement + LagrangeLeapFrog
// (LULESH/omp_4.0/lulesh.cc:3224-3225, 3040-3084), as the launches
//   k_time_increment      one wave per group: dt (MIN over its subdomains),
//                         time, cycle on the device
//   k_force_march         per column tile, marching up the node planes:
//                         stress + FB hourglass corner forces, the ordered
//                         node gather through LDS, and (subdomains without
//                         halo exchange) the node update (in
//                         lulesh_force_march.hip, built on its own)
//   k_node_update_direct  per node of a decomposed group: the neighbours'
//                         partial forces summed in receive-slot order
//                         (CommSBN) and each shared node's final writer
//                         pulled (CommSyncPosVel), no exchange launched
//   k_node_update         per node (after an explicit SBN exchange)
//   k_kinematics          per element: volume, arealg, vdov, MonoQ gradients
//                         (and the neighbours' MonoQ ghosts, CommMonoQ)
//   k_qeos                per element tile: MonoQ region, qstop, EOS x rep
//                         dealt over the workgroup, Courant/hydro dt minima
// plus the two-pass form of the force step (k_elem_force writes corner
// forces, k_node_gather<fuse> gathers them per node): the unstructured-mesh
// path, and a cross-check of the march on the block.  Every kernel walks a
// tile table (subdomain, first item)
// so one launch covers every subdomain of every group on the GPU.  Integer
// atomics are used only for the dt minima (MIN is exact) and the error
// code; all force sums run in the reference's element order, so results
// are bit-for-bit reproducible.
#include <hip/hip_runtime.h>

#include "lulesh_dev.h"
#include "lulesh_kernels.h"
#include "lulesh_math.h"
#include "lulesh_kernel_util.h"

namespace lulesh {

// ---------------------------------------------------------------------------
// TimeIncrement (lulesh.cc:207-263) and the loop test of main
// (lulesh.cc:3222), one thread per group.  gnewdt is formed per subdomain
// and reduced with MIN across the group (the MPI_Allreduce); every
// subdomain then applies the ratio clamps to its own deltatime.
// ---------------------------------------------------------------------------
// doms/dom0: the engine's subdomain descriptors and the index of each
// group's first one; the counted cycle n reads node-state buffer (n-1)&1
// and writes buffer n&1
__device__ __forceinline__ void ns_unpack(const NodeState *s, double *p[6])
{
    p[0] = *gp(&s->x); p[1] = *gp(&s->y); p[2] = *gp(&s->z);
    p[3] = *gp(&s->xd); p[4] = *gp(&s->yd); p[5] = *gp(&s->zd);
}

__global__ __launch_bounds__(256) void k_time_increment(GroupCtl *const *groups, int ngroups, DomDev *doms,
                                                         const int *dom0)
{
    // one wave per group, lane r = subdomain r (at most 4^3 = 64)
    const int gi = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int r = threadIdx.x & 63;
    if (gi >= ngroups) return;
    // Every load is issued ahead of the branches that use it, in as few
    // dependent rounds as the addresses allow: the group pointer and its
    // first subdomain's index; the group's loop state, subdomain 0's time
    // state and node-state buffers; then (groups of several subdomains
    // only) the other subdomains'.  The DomCtl array follows the GroupCtl
    // in its allocation.
    GroupCtl *const Gp = *gp(&groups[gi]);
    const int d0 = *gp(&dom0[gi]);
    DomCtl *const D = reinterpret_cast<DomCtl *>(Gp + 1);
    const double stoptime = *gp(&Gp->stoptime);
    const int max_cycles = *gp(&Gp->max_cycles), stop = *gp(&Gp->done) | *gp(&Gp->status);
    const int ndom = *gp(&Gp->ndom);
    double time = *gp(&D[0].time), olddt = *gp(&D[0].deltatime);
    unsigned long long dcb = *gp(&D[0].dtcourant_bits), dhb = *gp(&D[0].dthydro_bits);
    int cyc = *gp(&D[0].cycle);
    double *b0[6], *b1[6];
    ns_unpack(&doms[d0].ns[0], b0);
    ns_unpack(&doms[d0].ns[1], b1);
    // the values are needed here: the compiler may not sink their loads
    // behind the branches below (one wait instead of one per branch)
    __asm__ volatile("" ::"v"(stoptime), "v"(max_cycles), "v"(stop), "v"(ndom), "v"(time), "v"(olddt),
                     "v"(dcb), "v"(dhb), "v"(cyc), "v"(b0[0]), "v"(b0[1]), "v"(b0[2]), "v"(b0[3]),
                     "v"(b0[4]), "v"(b0[5]), "v"(b1[0]), "v"(b1[1]), "v"(b1[2]), "v"(b1[3]), "v"(b1[4]),
                     "v"(b1[5]));
    if (stop != 0) return;
    if (ti_stop(time, cyc, stoptime, max_cycles)) {
        if (r == 0) Gp->done = 1;
        return;
    }
    const bool mine = r < ndom;
    if (mine && r > 0) {
        time = D[r].time;
        olddt = D[r].deltatime;
        dcb = D[r].dtcourant_bits;
        dhb = D[r].dthydro_bits;
        cyc = D[r].cycle;
        ns_unpack(&doms[d0 + r].ns[0], b0);
        ns_unpack(&doms[d0 + r].ns[1], b1);
    }
    // each subdomain's candidate, then the MIN over the group (MIN is
    // exact, so any reduction order gives the sequential result)
    double newdt = mine ? ti_candidate(dcb, dhb) : __builtin_inf();
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(newdt, off, 64);
        newdt = o < newdt ? o : newdt;
    }
    if (!mine) return;
    const double deltatime = ti_deltatime(newdt, olddt, time, cyc, stoptime);
    const int cycle = cyc + 1;
    DomCtl &d = D[r];
    d.deltatime = deltatime;
    d.time = time + deltatime;
    d.cycle = cycle;
    // the running cycle's state, for the kernels that read the n-fields
    d.ndeltatime = deltatime;
    d.ntime = time + deltatime;
    d.ncycle = cycle;
    // CalcTimeConstraintsForElems restarts from 1e20 (lulesh.cc:3019-3020)
    d.dtcourant_bits = dbits(1.0e+20);
    d.dthydro_bits = dbits(1.0e+20);
    // the counted cycle n reads node-state buffer (n-1)&1, writes buffer n&1
    const bool odd = cycle & 1;
    double *cur[6], *old[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        cur[f] = odd ? b1[f] : b0[f];
        old[f] = odd ? b0[f] : b1[f];
    }
    DomDev &dd = doms[d0 + r];
    dd.cur = NodeState{cur[0], cur[1], cur[2], cur[3], cur[4], cur[5]};
    dd.old = NodeState{old[0], old[1], old[2], old[3], old[4], old[5]};
}

// node ids of element (i,j,k) in the structured block (BuildMesh,
// lulesh-init.cc:214-237): corners 0-3 on plane k, 4-7 on plane k+1
__device__ __forceinline__ void elem_nodes(int nx, int el, int n[8], int &i, int &j, int &k)
{
    const int s1 = nx + 1, s2 = s1 * s1;
    k = el / (nx * nx);
    const int rem = el - k * nx * nx;
    j = rem / nx;
    i = rem - j * nx;
    const int n0 = k * s2 + j * s1 + i;
    n[0] = n0;
    n[1] = n0 + 1;
    n[2] = n0 + s1 + 1;
    n[3] = n0 + s1;
    n[4] = n0 + s2;
    n[5] = n0 + s2 + 1;
    n[6] = n0 + s2 + s1 + 1;
    n[7] = n0 + s2 + s1;
}

// node ids of element el of an unstructured mesh (two 16-byte loads)
__device__ __forceinline__ void elem_nodes_u(const DomDev &D, int el, int n[8])
{
    const int4 *p = reinterpret_cast<const int4 *>(D.u_nodelist) + 2 * (size_t)el;
    const int4 a = p[0], b = p[1];
    n[0] = a.x; n[1] = a.y; n[2] = a.z; n[3] = a.w;
    n[4] = b.x; n[5] = b.y; n[6] = b.z; n[7] = b.w;
}

template <bool U>
__device__ __forceinline__ void elem_nodes_of(const DomDev &D, int el, int n[8])
{
    if (U) {
        elem_nodes_u(D, el, n);
    } else {
        int i, j, k;
        elem_nodes(D.nx, el, n, i, j, k);
    }
}

// ---------------------------------------------------------------------------
// IntegrateStressForElems + CalcHourglassControlForElems +
// CalcFBHourglassForceForElems (lulesh.cc:546-712, 837-1297): per element,
// the 8 stress corner forces -(sigma B) from the face normals and the 8
// Flanagan-Belytschko hourglass corner forces.  Both sets are stored
// separately because the reference sums them separately at each node.
// ---------------------------------------------------------------------------
template <bool U>
__global__ __launch_bounds__(256) void k_elem_force(const DomDev *doms, const Tile *tiles, int ntiles)
{
    const Tile t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
    // the tile's subdomain is uniform over the workgroup: one scalar copy of
    // its descriptor, which no store in the kernel can be taken to alias
    const DomDev D = doms[__builtin_amdgcn_readfirstlane(t.dom)];
    GroupCtl *G = D.grp;
    if (group_stopped(*G)) return;
    const int el = t.start + threadIdx.x;
    if (el >= D.numElem) return;
    int n[8];
    elem_nodes_of<U>(D, el, n);
    const size_t E = (size_t)D.numElem;
    const NodeState S = ns_old(D);

    double x[8], y[8], z[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) { x[c] = gp(S.x)[n[c]]; y[c] = gp(S.y)[n[c]]; z[c] = gp(S.z)[n[c]]; }

    // stress part
    {
        if (jac_volume(x, y, z) <= 0.0) raise_status(G, -1);      // lulesh.cc:1340-1349
        double bx[8], by[8], bz[8];
        node_normals16(x, y, z, bx, by, bz);
        // InitStressTermsForElems' sigma, over 16 against the normals x 16
        const double sig16 = (-gp(D.p)[el] - gp(D.q)[el]) * 0.0625;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            // one scalar base per (component, corner) row, one offset for all
            gst(D.corner + (0 * 8 + c) * E, el, -(sig16 * bx[c]));
            gst(D.corner + (1 * 8 + c) * E, el, -(sig16 * by[c]));
            gst(D.corner + (2 * 8 + c) * E, el, -(sig16 * bz[c]));
        }
    }

    // hourglass part
    const double vv = gp(D.v)[el];
    if (vv <= 0.0) raise_status(G, -1);                // lulesh.cc:1263-1265
    const double determ = gp(D.volo)[el] * vv;
    double dvdx[8], dvdy[8], dvdz[8];
    volume_derivs(x, y, z, dvdx, dvdy, dvdz);
    const double volinv = rcp_pos(determ);   // volo * v: positive unless the volume error is raised
    double hg[8][4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        double hx = x[0] * hg_gamma(m, 0), hy = y[0] * hg_gamma(m, 0), hz = z[0] * hg_gamma(m, 0);
#pragma unroll
        for (int c = 1; c < 8; ++c) {
            hx = hx + x[c] * hg_gamma(m, c);
            hy = hy + y[c] * hg_gamma(m, c);
            hz = hz + z[c] * hg_gamma(m, c);
        }
#pragma unroll
        for (int c = 0; c < 8; ++c)
            hg[c][m] = hg_gamma(m, c) - volinv * (dvdx[c] * hx + dvdy[c] * hy + dvdz[c] * hz);
    }
    const double coefficient = gp(D.hgcoef)[el];
    const __attribute__((address_space(1))) double *vel[3] = {gp(S.xd), gp(S.yd), gp(S.zd)};
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) {
        double vl[8], h[4];
#pragma unroll
        for (int c = 0; c < 8; ++c) vl[c] = vel[comp][n[c]];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            double acc = hg[0][m] * vl[0];
#pragma unroll
            for (int c = 1; c < 8; ++c) acc = acc + hg[c][m] * vl[c];
            h[m] = acc;
        }
#pragma unroll
        for (int c = 0; c < 8; ++c)
            gst(D.corner + ((3 + comp) * 8 + c) * E, el,
                coefficient * (hg[c][0] * h[0] + hg[c][1] * h[1] + hg[c][2] * h[2] + hg[c][3] * h[3]));
    }
}

// symmetry-plane membership of node nd (bit 0: x = 0 plane, 1: y, 2: z).
// On the block the planes exist on the subdomains at the mesh origin
// (SetupSymmetryPlanes, lulesh-init.cc:487-508); an unstructured mesh lists
// its symmetry node sets (lulesh.cu:939-965).
template <bool U>
__device__ __forceinline__ unsigned node_symm(const DomDev &D, int nd, int i, int j, int k)
{
    if (U) return gp(D.u_nsymm)[nd];
    return (D.col == 0 && i == 0 ? 1u : 0u) | (D.row == 0 && j == 0 ? 2u : 0u) |
           (D.plane == 0 && k == 0 ? 4u : 0u);
}

// node (i,j,k) of the block, from its index
__device__ __forceinline__ void node_ijk(int nx, int nd, int &i, int &j, int &k)
{
    const int s1 = nx + 1;
    k = nd / (s1 * s1);
    const int rem = nd - k * s1 * s1;
    j = rem / s1;
    i = rem - j * s1;
}

// CalcAccelerationForNodes + ApplyAccelerationBoundaryConditionsForNodes +
// CalcVelocityForNodes + CalcPositionForNodes (lulesh.cc:1415-1550) for
// node nd of subdomain D, symm from node_symm
struct NodeNew {
    double x, y, z, xd, yd, zd;
};

__device__ __forceinline__ NodeNew node_advance(const DomDev &D, const NodeState &src, const NodeState &dst,
                                                int nd, unsigned symm, double fx, double fy, double fz,
                                                double dt)
{
    const double one_over_nMass = rcp_pos(gp(D.nodalMass)[nd]);   // masses are positive
    double ax = fx * one_over_nMass, ay = fy * one_over_nMass, az = fz * one_over_nMass;
    if (symm & 1u) ax = 0.0;
    if (symm & 2u) ay = 0.0;
    if (symm & 4u) az = 0.0;
    double vx = gp(src.xd)[nd] + ax * dt;
    if (fabs(vx) < Phys::u_cut) vx = 0.0;
    double vy = gp(src.yd)[nd] + ay * dt;
    if (fabs(vy) < Phys::u_cut) vy = 0.0;
    double vz = gp(src.zd)[nd] + az * dt;
    if (fabs(vz) < Phys::u_cut) vz = 0.0;
    const double px = gp(src.x)[nd] + vx * dt;
    const double py = gp(src.y)[nd] + vy * dt;
    const double pz = gp(src.z)[nd] + vz * dt;
    gp(dst.xd)[nd] = vx;
    gp(dst.yd)[nd] = vy;
    gp(dst.zd)[nd] = vz;
    gp(dst.x)[nd] = px;
    gp(dst.y)[nd] = py;
    gp(dst.z)[nd] = pz;
    return NodeNew{px, py, pz, vx, vy, vz};
}

// ---------------------------------------------------------------------------
// Node gather (lulesh.cc:685-703 and 1134-1152): the node's corners in
// ascending element order -- elements (i-1,j-1,k-1) ... (i,j,k) -- summed
// from 0.0, stress first, then hourglass, f = stress + hourglass.
// ---------------------------------------------------------------------------
// An unstructured mesh lists each node's corners (8*element + corner) in
// ascending element order (the reference's nodeElemCornerList,
// lulesh-init.cc:243-313): the same two ordered sums over that list.
template <bool FUSE_UPDATE>
__device__ __forceinline__ void node_gather_u(const DomDev &D, int nd)
{
    const size_t E = (size_t)D.numElem;
    const int q0 = gp(D.u_nstart)[nd], cnt = gp(D.u_nstart)[nd + 1] - q0;
    double sx = 0.0, sy = 0.0, sz = 0.0, hx = 0.0, hy = 0.0, hz = 0.0;
    // corners in batches of 8 (a hexahedral mesh node has at most 8 unless
    // the mesh is irregular): every load of a batch is issued before the
    // first add, then the adds run in list order
    // an unstructured mesh has fewer than 2^26 elements, so a corner's
    // offset within one component's 8 rows fits 32 bits: scalar component
    // bases and one offset register per corner
    for (int b = 0; b < cnt; b += 8) {
        unsigned off[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int ent = b + t < cnt ? gp(D.u_ncorner)[q0 + b + t] : 0;
            off[t] = (unsigned)(ent & 7) * (unsigned)E + (unsigned)(ent >> 3);
        }
        // loaded on every path (a missing entry reads element 0's corner 0)
        double v[6][8];
#pragma unroll
        for (int t = 0; t < 8; ++t)
#pragma unroll
            for (int comp = 0; comp < 6; ++comp) v[comp][t] = gld(D.corner + comp * 8 * E, (unsigned)off[t]);
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            if (b + t >= cnt) break;
            sx += v[0][t];
            sy += v[1][t];
            sz += v[2][t];
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            if (b + t >= cnt) break;
            hx += v[3][t];
            hy += v[4][t];
            hz += v[5][t];
        }
    }
    const double fx = sx + hx, fy = sy + hy, fz = sz + hz;
    if (FUSE_UPDATE) {
        node_advance(D, ns_old(D), ns_new(D), nd, node_symm<true>(D, nd, 0, 0, 0), fx, fy, fz,
                     D.ctl->ndeltatime);
    } else {
        gp(D.fpx)[nd] = fx;
        gp(D.fpy)[nd] = fy;
        gp(D.fpz)[nd] = fz;
    }
}

template <bool FUSE_UPDATE, bool U>
__global__ __launch_bounds__(256) void k_node_gather(const DomDev *doms, const Tile *tiles, int ntiles)
{
    const Tile t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
    // the tile's subdomain is uniform over the workgroup: one scalar copy of
    // its descriptor, which no store in the kernel can be taken to alias
    const DomDev D = doms[__builtin_amdgcn_readfirstlane(t.dom)];
    if (group_stopped(*D.grp)) return;
    const int nd = t.start + threadIdx.x;
    if (nd >= D.numNode) return;
    if (U) {
        node_gather_u<FUSE_UPDATE>(D, nd);
        return;
    }
    const int nx = D.nx;
    int i, j, k;
    node_ijk(nx, nd, i, j, k);
    const size_t E = (size_t)D.numElem;

    // (di, dj, dk, corner): element (i+di, j+dj, k+dk) holds this node as
    // the listed corner; listed in ascending element index
    constexpr int L[8][4] = {{-1, -1, -1, 6}, {0, -1, -1, 7}, {-1, 0, -1, 5}, {0, 0, -1, 4},
                             {-1, -1, 0, 2}, {0, -1, 0, 3}, {-1, 0, 0, 1}, {0, 0, 0, 0}};
    int eidx[8];
    int ccorner[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int a = i + L[q][0], b = j + L[q][1], c = k + L[q][2];
        const bool ok = a >= 0 && a < nx && b >= 0 && b < nx && c >= 0 && c < nx;
        eidx[q] = ok ? (c * nx + b) * nx + a : -1;
        ccorner[q] = L[q][3];
    }
    double sx = 0.0, sy = 0.0, sz = 0.0, hx = 0.0, hy = 0.0, hz = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        if (eidx[q] < 0) continue;
        const size_t e = (size_t)eidx[q];
        const int c = ccorner[q];
        sx += gp(D.corner)[(0 * 8 + c) * E + e];
        sy += gp(D.corner)[(1 * 8 + c) * E + e];
        sz += gp(D.corner)[(2 * 8 + c) * E + e];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        if (eidx[q] < 0) continue;
        const size_t e = (size_t)eidx[q];
        const int c = ccorner[q];
        hx += gp(D.corner)[(3 * 8 + c) * E + e];
        hy += gp(D.corner)[(4 * 8 + c) * E + e];
        hz += gp(D.corner)[(5 * 8 + c) * E + e];
    }
    const double fx = sx + hx, fy = sy + hy, fz = sz + hz;
    if (FUSE_UPDATE) {
        node_advance(D, ns_old(D), ns_new(D), nd, node_symm<false>(D, nd, i, j, k), fx, fy, fz,
                     D.ctl->ndeltatime);
    } else {
        gp(D.fpx)[nd] = fx;
        gp(D.fpy)[nd] = fy;
        gp(D.fpz)[nd] = fz;
    }
}

template <bool U>
__global__ __launch_bounds__(256) void k_node_update(const DomDev *doms, const Tile *tiles, int ntiles)
{
    const Tile t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
    // the tile's subdomain is uniform over the workgroup: one scalar copy of
    // its descriptor, which no store in the kernel can be taken to alias
    const DomDev D = doms[__builtin_amdgcn_readfirstlane(t.dom)];
    if (group_stopped(*D.grp)) return;
    const int nd = t.start + threadIdx.x;
    if (nd >= D.numNode) return;
    int i = 0, j = 0, k = 0;
    if (!U) node_ijk(D.nx, nd, i, j, k);
    node_advance(D, ns_old(D), ns_new(D), nd, node_symm<U>(D, nd, i, j, k), gp(D.fx)[nd], gp(D.fy)[nd], gp(D.fz)[nd],
                 D.ctl->ndeltatime);
}

// CommSBN + the node update in one pass (device_local backend): a shared
// node adds the neighbours' own partial forces -- read where their force
// kernel left them, untouched during this launch -- in this subdomain's
// receive-slot order onto its own partial (lulesh-comm.cc:848-1257), as
// the reference's pack-then-unpack does, stores the sums as fx..fz and
// advances.
//
// With D.dir_pv, CommSyncPosVel too (lulesh-comm.cc:1261-1680): a copy
// that the exchange would overwrite takes its final writer's update, which
// it computes itself from the writer's force sum (the writer's partial and
// the writer's receive-slot order), the writer's nodal mass and the
// writer's pre-update state from the writer's snapshot of the previous
// cycle; a copy that is some other copy's final writer stores its new
// state in its own snapshot for the next cycle.
__device__ __forceinline__ void sbn_sum(const DomDev *doms, const DomDev &D, int nd, double &fx, double &fy,
                                        double &fz)
{
    fx = gp(D.fpx)[nd];
    fy = gp(D.fpy)[nd];
    fz = gp(D.fpz)[nd];
    const IdxSpan r = D.dir_node ? D.dir_node[nd] : IdxSpan{0, 0};
    for (int m = 0; m < r.count; ++m) {
        const DirRef s = D.dir_sbn[r.first + m];
        const DomDev &S = doms[D.dom_base + s.dom];
        fx += gp(S.fpx)[s.idx];
        fy += gp(S.fpy)[s.idx];
        fz += gp(S.fpz)[s.idx];
    }
}

__device__ __forceinline__ void node_update_pv(const DomDev *doms, const DomDev &D, int nd, int pi);

__global__ __launch_bounds__(256) void k_node_update_direct(const DomDev *doms, const Tile *tiles, int ntiles)
{
    const Tile t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
    if (t.pad) {
        // a tile of PosVel copies: (rank, node) entries of the group
        const DomDev &G0 = doms[t.dom];
        if (group_stopped(*G0.grp) || (int)threadIdx.x >= t.ecount) return;
        const DirRef c = G0.dir_pvlist[t.start + threadIdx.x];
        const DomDev &D = doms[G0.dom_base + c.dom];
        node_update_pv(doms, D, c.idx, gp(D.dir_pv)[c.idx]);
        return;
    }
    // the tile's subdomain is uniform over the workgroup: one scalar copy of
    // its descriptor, which no store in the kernel can be taken to alias
    const DomDev D = doms[__builtin_amdgcn_readfirstlane(t.dom)];
    if (group_stopped(*D.grp)) return;
    const int nd = t.start + threadIdx.x;
    if (nd >= D.numNode) return;
    if (D.dir_pv && gp(D.dir_pv)[nd] >= 0) return;   // in a PosVel tile
    int i, j, k;
    node_ijk(D.nx, nd, i, j, k);
    double fx, fy, fz;
    sbn_sum(doms, D, nd, fx, fy, fz);
    gp(D.fx)[nd] = fx;
    gp(D.fy)[nd] = fy;
    gp(D.fz)[nd] = fz;
    node_advance(D, ns_old(D), ns_new(D), nd, node_symm<false>(D, nd, i, j, k), fx, fy, fz,
                 D.ctl->ndeltatime);
}

// the update of a node copy that CommSyncPosVel touches
__device__ __forceinline__ void node_update_pv(const DomDev *doms, const DomDev &D, int nd, int pi)
{
    int i, j, k;
    node_ijk(D.nx, nd, i, j, k);
    double fx, fy, fz;
    sbn_sum(doms, D, nd, fx, fy, fz);
    gp(D.fx)[nd] = fx;
    gp(D.fy)[nd] = fy;
    gp(D.fz)[nd] = fz;
    const unsigned symm = node_symm<false>(D, nd, i, j, k);
    const double dt = D.ctl->ndeltatime;
    const PvNode P = D.dir_pvnode[pi];
    const int par = D.ctl->ncycle & 1;   // snapshots: read the other parity, write this one
    NodeNew v;
    if (P.fw_dom >= 0) {
        // the same physical node in the writer: same symmetry planes
        const DomDev &W = doms[D.dom_base + P.fw_dom];
        double gx, gy, gz;
        sbn_sum(doms, W, P.fw_idx, gx, gy, gz);
        const auto old = gp(W.pv_snap) + (size_t)(par ^ 1) * 6 * W.pv_nslot + P.fw_slot;
        const size_t ns = W.pv_nslot;
        const double one_over_nMass = rcp_pos(gp(W.nodalMass)[P.fw_idx]);   // masses are positive
        double ax = gx * one_over_nMass, ay = gy * one_over_nMass, az = gz * one_over_nMass;
        if (symm & 1u) ax = 0.0;
        if (symm & 2u) ay = 0.0;
        if (symm & 4u) az = 0.0;
        double vx = old[3 * ns] + ax * dt;
        if (fabs(vx) < Phys::u_cut) vx = 0.0;
        double vy = old[4 * ns] + ay * dt;
        if (fabs(vy) < Phys::u_cut) vy = 0.0;
        double vz = old[5 * ns] + az * dt;
        if (fabs(vz) < Phys::u_cut) vz = 0.0;
        v = NodeNew{old[0] + vx * dt, old[ns] + vy * dt, old[2 * ns] + vz * dt, vx, vy, vz};
        const NodeState N = ns_new(D);
        gp(N.x)[nd] = v.x;
        gp(N.y)[nd] = v.y;
        gp(N.z)[nd] = v.z;
        gp(N.xd)[nd] = v.xd;
        gp(N.yd)[nd] = v.yd;
        gp(N.zd)[nd] = v.zd;
    } else {
        v = node_advance(D, ns_old(D), ns_new(D), nd, symm, fx, fy, fz, dt);
    }
    if (P.own_slot >= 0) {
        const auto snap = gp(D.pv_snap) + (size_t)par * 6 * D.pv_nslot + P.own_slot;
        const size_t ns = D.pv_nslot;
        snap[0] = v.x;
        snap[ns] = v.y;
        snap[2 * ns] = v.z;
        snap[3 * ns] = v.xd;
        snap[4 * ns] = v.yd;
        snap[5 * ns] = v.zd;
    }
}

// the final writers' snapshots of the current node state, into the parity
// of the last counted cycle (the one the next cycle reads)
__global__ __launch_bounds__(256) void k_pv_snapshot(const DomDev *doms, const Tile *tiles, int ntiles)
{
    const Tile t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
    // the tile's subdomain is uniform over the workgroup: one scalar copy of
    // its descriptor, which no store in the kernel can be taken to alias
    const DomDev D = doms[__builtin_amdgcn_readfirstlane(t.dom)];
    const int nd = t.start + threadIdx.x;
    if (nd >= D.numNode || !D.dir_pv) return;
    const int pi = gp(D.dir_pv)[nd];
    if (pi < 0) return;
    const int slot = D.dir_pvnode[pi].own_slot;
    if (slot < 0) return;
    const NodeState S = D.cur;
    const size_t ns = D.pv_nslot;
    const auto snap = gp(D.pv_snap) + (size_t)(D.ctl->cycle & 1) * 6 * ns + slot;
    snap[0] = gp(S.x)[nd];
    snap[ns] = gp(S.y)[nd];
    snap[2 * ns] = gp(S.z)[nd];
    snap[3 * ns] = gp(S.xd)[nd];
    snap[4 * ns] = gp(S.yd)[nd];
    snap[5 * ns] = gp(S.zd)[nd];
}

// ---------------------------------------------------------------------------
// CalcKinematicsForElems + CalcLagrangeElements + CalcMonotonicQGradients-
// ForElems (lulesh.cc:1861-2054, 2059-2244): one gather of the element's
// positions and velocities feeds vnew, delv, arealg, vdov and the six
// MonoQ gradients.  The deviatoric strains of the reference are not
// stored: nothing reads them.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void kin_element(const DomDev *doms, const DomDev &D, GroupCtl *G, int el, double dt,
                                            const double x[8], const double y[8], const double z[8],
                                            const double xv[8], const double yv[8], const double zv[8])
{
    const double volo = gld(D.volo, el);
    const double volume = elem_volume(x, y, z);
    const double vnew = volume / volo;
    gst(D.vnew, el, vnew);
    gst(D.delv, el, vnew - gld(D.v, el));
    gst(D.arealg, el, char_length(x, y, z, volume));
    if (vnew <= 0.0) raise_status(G, -1);              // lulesh.cc:2037-2049

    {
        const double dt2 = 0.5 * dt;
        double xh[8], yh[8], zh[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            xh[c] = x[c] - dt2 * xv[c];
            yh[c] = y[c] - dt2 * yv[c];
            zh[c] = z[c] - dt2 * zv[c];
        }
        // B and the Jacobian determinant both 64x: inv_detJ * (B . v) is
        // the reference's quotient exactly
        const Jac64 J = elem_jacobian64(xh, yh, zh);
        double b[3][4];
        shape_b4(J, b);
        const double inv_detJ = rcp_pos(J.volume);   // positive for an element that is not inverted
        // CalcElemVelocityGradient diagonal, lulesh.cc:1809-1822
        const double dxx = inv_detJ * (b[0][0] * (xv[0] - xv[6]) + b[0][1] * (xv[1] - xv[7]) +
                                       b[0][2] * (xv[2] - xv[4]) + b[0][3] * (xv[3] - xv[5]));
        const double dyy = inv_detJ * (b[1][0] * (yv[0] - yv[6]) + b[1][1] * (yv[1] - yv[7]) +
                                       b[1][2] * (yv[2] - yv[4]) + b[1][3] * (yv[3] - yv[5]));
        const double dzz = inv_detJ * (b[2][0] * (zv[0] - zv[6]) + b[2][1] * (zv[1] - zv[7]) +
                                       b[2][2] * (zv[2] - zv[4]) + b[2][3] * (zv[3] - zv[5]));
        gst(D.vdov, el, dxx + dyy + dzz);
    }

    // MonoQ gradients at the end-of-step positions (each sqrt argument
    // below is a sum of squares plus ptiny: at least ptiny, in sqrt_pos's
    // range)
    const double ptiny = 1.e-36;
    const double vol = volo * vnew;
    const double norm = rcp_pos(vol + ptiny);   // >= ptiny unless the volume error is raised
    const double dxj = -0.25 * ((x[0] + x[1] + x[5] + x[4]) - (x[3] + x[2] + x[6] + x[7]));
    const double dyj = -0.25 * ((y[0] + y[1] + y[5] + y[4]) - (y[3] + y[2] + y[6] + y[7]));
    const double dzj = -0.25 * ((z[0] + z[1] + z[5] + z[4]) - (z[3] + z[2] + z[6] + z[7]));
    const double dxi = 0.25 * ((x[1] + x[2] + x[6] + x[5]) - (x[0] + x[3] + x[7] + x[4]));
    const double dyi = 0.25 * ((y[1] + y[2] + y[6] + y[5]) - (y[0] + y[3] + y[7] + y[4]));
    const double dzi = 0.25 * ((z[1] + z[2] + z[6] + z[5]) - (z[0] + z[3] + z[7] + z[4]));
    const double dxk = 0.25 * ((x[4] + x[5] + x[6] + x[7]) - (x[0] + x[1] + x[2] + x[3]));
    const double dyk = 0.25 * ((y[4] + y[5] + y[6] + y[7]) - (y[0] + y[1] + y[2] + y[3]));
    const double dzk = 0.25 * ((z[4] + z[5] + z[6] + z[7]) - (z[0] + z[1] + z[2] + z[3]));
    double ax, ay, az, dxv, dyv, dzv;

    ax = dyi * dzj - dzi * dyj; ay = dzi * dxj - dxi * dzj; az = dxi * dyj - dyi * dxj;
    gst(D.delx_zeta, el, vol / sqrt_pos(ax * ax + ay * ay + az * az + ptiny));
    ax *= norm; ay *= norm; az *= norm;
    dxv = 0.25 * ((xv[4] + xv[5] + xv[6] + xv[7]) - (xv[0] + xv[1] + xv[2] + xv[3]));
    dyv = 0.25 * ((yv[4] + yv[5] + yv[6] + yv[7]) - (yv[0] + yv[1] + yv[2] + yv[3]));
    dzv = 0.25 * ((zv[4] + zv[5] + zv[6] + zv[7]) - (zv[0] + zv[1] + zv[2] + zv[3]));
    const double dv_zeta = ax * dxv + ay * dyv + az * dzv;
    gst(D.delv_zeta, el, dv_zeta);

    ax = dyj * dzk - dzj * dyk; ay = dzj * dxk - dxj * dzk; az = dxj * dyk - dyj * dxk;
    gst(D.delx_xi, el, vol / sqrt_pos(ax * ax + ay * ay + az * az + ptiny));
    ax *= norm; ay *= norm; az *= norm;
    dxv = 0.25 * ((xv[1] + xv[2] + xv[6] + xv[5]) - (xv[0] + xv[3] + xv[7] + xv[4]));
    dyv = 0.25 * ((yv[1] + yv[2] + yv[6] + yv[5]) - (yv[0] + yv[3] + yv[7] + yv[4]));
    dzv = 0.25 * ((zv[1] + zv[2] + zv[6] + zv[5]) - (zv[0] + zv[3] + zv[7] + zv[4]));
    const double dv_xi = ax * dxv + ay * dyv + az * dzv;
    gst(D.delv_xi, el, dv_xi);

    ax = dyk * dzi - dzk * dyi; ay = dzk * dxi - dxk * dzi; az = dxk * dyi - dyk * dxi;
    gst(D.delx_eta, el, vol / sqrt_pos(ax * ax + ay * ay + az * az + ptiny));
    ax *= norm; ay *= norm; az *= norm;
    dxv = -0.25 * ((xv[0] + xv[1] + xv[5] + xv[4]) - (xv[3] + xv[2] + xv[6] + xv[7]));
    dyv = -0.25 * ((yv[0] + yv[1] + yv[5] + yv[4]) - (yv[3] + yv[2] + yv[6] + yv[7]));
    dzv = -0.25 * ((zv[0] + zv[1] + zv[5] + zv[4]) - (zv[3] + zv[2] + zv[6] + zv[7]));
    const double dv_eta = ax * dxv + ay * dyv + az * dzv;
    gst(D.delv_eta, el, dv_eta);

    // CommMonoQ done in place (device_local backend): an element on a
    // shared face fills its neighbours' ghost slots with its three
    // gradients, exactly what the face message would carry
    // (lulesh-comm.cc:1684-1837); qeos reads them after this launch
    if (D.dir_elem) {
        const IdxSpan r = D.dir_elem[el];
        if (r.count > 0) {
            for (int m = 0; m < r.count; ++m) {
                const DirRef g = D.dir_ghost[r.first + m];
                const DomDev &T = doms[D.dom_base + g.dom];
                gp(T.delv_xi)[g.idx] = dv_xi;
                gp(T.delv_eta)[g.idx] = dv_eta;
                gp(T.delv_zeta)[g.idx] = dv_zeta;
            }
        }
    }
}

template <bool U>
__global__ __launch_bounds__(256) void k_kinematics(const DomDev *doms, const Tile *tiles, int ntiles)
{
    const Tile t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
    // the tile's subdomain is uniform over the workgroup: one scalar copy of
    // its descriptor, which no store in the kernel can be taken to alias
    const DomDev D = doms[__builtin_amdgcn_readfirstlane(t.dom)];
    GroupCtl *G = D.grp;
    // the stop flag is tested once the node loads are in flight (reads
    // only: a stopped group's arrays are still allocated)
    const bool stopped = group_stopped(*G);
    if (t.pad == kFirstTile && threadIdx.x == 0) {
        // the subdomain's first tile makes the running cycle the completed
        // one (the other tiles read only the n-fields) and restarts the dt
        // minima for the EOS kernel (CalcTimeConstraintsForElems restarts
        // from 1e20, lulesh.cc:3019-3020) -- also after an error stopped
        // the cycle, which the reference has counted too
        DomCtl &C = *D.ctl;
        if (C.ncycle != C.cycle) {
            C.time = C.ntime;
            C.deltatime = C.ndeltatime;
            C.cycle = C.ncycle;
            C.dtcourant_bits = dbits(1.0e+20);
            C.dthydro_bits = dbits(1.0e+20);
        }
    }
    const int el = t.start + threadIdx.x;
    if (el >= D.numElem) return;
    int n[8];
    elem_nodes_of<U>(D, el, n);
    const double dt = D.ctl->ndeltatime;

    const NodeState S = ns_new(D);
    double x[8], y[8], z[8], xv[8], yv[8], zv[8];
    if (U) {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            x[c] = gld(S.x, n[c]); y[c] = gld(S.y, n[c]); z[c] = gld(S.z, n[c]);
            xv[c] = gld(S.xd, n[c]); yv[c] = gld(S.yd, n[c]); zv[c] = gld(S.zd, n[c]);
        }
    } else {
        // on the block, corners 1, 2, 5, 6 follow corners 0, 3, 4, 7 in
        // node order: one 16-byte load per pair and array
        const double *const A[6] = {S.x, S.y, S.z, S.xd, S.yd, S.zd};
        double *const V[6] = {x, y, z, xv, yv, zv};
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            gld_pair(A[f], n[0], V[f][0], V[f][1]);
            gld_pair(A[f], n[3], V[f][3], V[f][2]);
            gld_pair(A[f], n[4], V[f][4], V[f][5]);
            gld_pair(A[f], n[7], V[f][7], V[f][6]);
        }
    }
    if (stopped) return;
    kin_element(doms, D, G, el, dt, x, y, z, xv, yv, zv);
}

// one direction of CalcMonotonicQRegionForElems (lulesh.cc:2307-2340):
// the limiter phi from the element's own gradient and its two neighbours'
__device__ __forceinline__ double monoq_phi(double self, double dm, double dp)
{
    const double ptiny = 1.e-36;
    const double norm = 1. / (self + ptiny);
    double delvm = dm * norm, delvp = dp * norm;
    double phi = .5 * (delvm + delvp);
    delvm *= Phys::monoq_limiter_mult;
    delvp *= Phys::monoq_limiter_mult;
    // the four clamps as min / max instructions: for non-NaN values a tie
    // can only differ in the sign of a zero, and phi enters q only through
    // 1 - phi and phi * phi
    phi = fmin(phi, delvm);
    phi = fmin(phi, delvp);
    phi = fmax(phi, 0.);
    phi = fmin(phi, Phys::monoq_max_slope);
    return phi;
}

// Neighbour gradient across one face: interior neighbour, own value on a
// symmetry plane, 0 on a free surface, ghost plane on a subdomain face (the
// elemBC / lxim.. logic of SetupBoundaryConditions, lulesh-init.cc:576-630).
// Split into the index to load (the element's own where the face has no
// neighbour value) and whether the loaded value is kept, so that all of an
// element's loads can be issued before any of them is used.
__device__ __forceinline__ int face_idx(int el, bool at_edge, int nbr, int ghost_base, int ghost_off,
                                        bool symm_side, bool &keep)
{
    keep = !at_edge || ghost_base >= 0 || symm_side;
    return !at_edge ? nbr : (ghost_base >= 0 ? ghost_base + ghost_off : el);
}

// the same for an unstructured mesh, from the element's elemBC bits and
// face-neighbour list (CalcMonotonicQRegionForElems' switch,
// lulesh.cc:2307-2400): neighbour for COMM or no bit, own value for SYMM,
// 0 for FREE and any other combination
__device__ __forceinline__ int face_idx_u(int el, int nbr, int bits, int symm, int comm, bool &keep)
{
    const bool use_nbr = bits == 0 || bits == comm;
    keep = use_nbr || bits == symm;
    return use_nbr ? nbr : el;
}

__device__ __forceinline__ double ss_of(double pbvc, double e_new, double vv, double bvc,
                                        double p_new)
{
    double ssc = (pbvc * e_new + vv * vv * bvc * p_new) / Phys::refdens;
    if (ssc <= .1111111e-36) ssc = .3333333e-18;
    else ssc = sqrt_pos(ssc);                      // ssc > 1.1e-37 here
    return ssc;
}

// `if (x < c) x = c` as one maximum instruction: the same value for every
// non-NaN x when c is not zero, or when x cannot be -0.0 (in the EOS each
// pressure has just been snapped to +0.0 below its cut-off)
__device__ __forceinline__ double raise_to(double x, double c)
{
    return fmax(x, c);
}

struct EosOut {
    double p, e, q;
};

// the relative compression 1/v - 1 of EvalEOSForElems (lulesh.cc:2669-2676),
// zero at eosvmax; vnewc is already clamped to [eosvmin, eosvmax].  The
// last pressure evaluation's bvc is c1s * (compression + 1) and its pbvc
// c1s (lulesh.cc:2786-2789): the sound speed after the repetitions forms
// both again from vnewc instead of keeping them from the last repetition.
__device__ __forceinline__ double eos_compression(double vnewc)
{
    double compression = rcp_pos(vnewc) - 1.;
    if (vnewc >= Phys::eosvmax) compression = 0.;
    return compression;
}

// one rep of EvalEOSForElems (lulesh.cc:2677-2815); every rep recomputes
// the same values from the same inputs
__device__ __forceinline__ EosOut eos_rep(double vnewc, double e_old, double delvc, double p_old,
                                          double q_old, double qq_old, double ql_old)
{
    const double c1s = 2.0 / 3.0, sixth = 1.0 / 6.0;
    const double compression = eos_compression(vnewc);
    const double vchalf = vnewc - delvc * .5;
    // 1 / vchalf as the range-limited reciprocal: its value is used only
    // while eosvmin < vnewc < eosvmax, where vnewc is the kinematics' vnew
    // and delvc = vnew - v of the same cycle, so vchalf is (vnew + v) / 2
    // to rounding, with v (the last cycle's clamped volume) at least
    // eosvmin: far inside rcp_pos's range.  Outside that window the value
    // is replaced just below.
    double compHalfStep = rcp_pos(vchalf) - 1.;
    if (vnewc <= Phys::eosvmin) compHalfStep = compression;
    if (vnewc >= Phys::eosvmax) { p_old = 0.; compHalfStep = 0.; }
    const double work = 0.;
    double e_new = e_old - 0.5 * delvc * (p_old + q_old) + 0.5 * work;
    e_new = raise_to(e_new, Phys::emin);
    double bvc = c1s * (compHalfStep + 1.);
    double pbvc = c1s;
    double pHalfStep = bvc * e_new;
    if (fabs(pHalfStep) < Phys::p_cut) pHalfStep = 0.0;
    if (vnewc >= Phys::eosvmax) pHalfStep = 0.0;
    pHalfStep = raise_to(pHalfStep, Phys::pmin);
    // 1 + compHalfStep is about 1 / vchalf, about 1 / eosvmin, or 1
    const double vhalf = rcp_pos(1. + compHalfStep);
    double q_new;
    if (delvc > 0.) {
        q_new = 0.;
    } else {
        double ssc = (pbvc * e_new + vhalf * vhalf * bvc * pHalfStep) / Phys::refdens;
        if (ssc <= .1111111e-36) ssc = .3333333e-18;
        else ssc = sqrt_pos(ssc);                  // ssc > 1.1e-37 here
        q_new = (ssc * ql_old + qq_old);
    }
    e_new = e_new + 0.5 * delvc * (3.0 * (p_old + q_old) - 4.0 * (pHalfStep + q_new));
    e_new += 0.5 * work;
    if (fabs(e_new) < Phys::e_cut) e_new = 0.;
    e_new = raise_to(e_new, Phys::emin);
    bvc = c1s * (compression + 1.);
    pbvc = c1s;
    double p_new = bvc * e_new;
    if (fabs(p_new) < Phys::p_cut) p_new = 0.0;
    if (vnewc >= Phys::eosvmax) p_new = 0.0;
    p_new = raise_to(p_new, Phys::pmin);
    double q_tilde;
    if (delvc > 0.) q_tilde = 0.;
    else q_tilde = ss_of(pbvc, e_new, vnewc, bvc, p_new) * ql_old + qq_old;
    e_new = e_new - (7.0 * (p_old + q_old) - 8.0 * (pHalfStep + q_new) + (p_new + q_tilde)) * delvc * sixth;
    if (fabs(e_new) < Phys::e_cut) e_new = 0.;
    e_new = raise_to(e_new, Phys::emin);
    bvc = c1s * (compression + 1.);
    pbvc = c1s;
    p_new = bvc * e_new;
    if (fabs(p_new) < Phys::p_cut) p_new = 0.0;
    if (vnewc >= Phys::eosvmax) p_new = 0.0;
    p_new = raise_to(p_new, Phys::pmin);
    if (delvc <= 0.) {
        q_new = ss_of(pbvc, e_new, vnewc, bvc, p_new) * ql_old + qq_old;
        if (fabs(q_new) < Phys::q_cut) q_new = 0.;
    }
    EosOut o;
    o.p = p_new; o.e = e_new; o.q = q_new;
    return o;
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    return v;
}

// ---------------------------------------------------------------------------
// CalcMonotonicQRegionForElems + the qstop scan + EvalEOSForElems x rep +
// CalcCourantConstraintForElems + CalcHydroConstraintForElems
// (lulesh.cc:2249-2546, 2586-2836, 2874-3035), fused per element.
//
// The EOS is evaluated rep times per element, as in the reference, but the
// repetitions are spread over the workgroup: the tile's entries of the
// subdomain's evaluation table (element i listed elemRep[i] times, in
// element order) are dealt round-robin to the lanes, and the lane holding
// an element's last entry stores its result.  Regions of different cost
// sit next to each other in element order, so a per-lane loop would run
// every wave to its most expensive element, and a per-wave balance leaves
// the workgroup waiting on its most expensive wave.  Every repetition
// recomputes the same values from the same inputs, so the element's result
// is that of the reference's sequential loop.
// ---------------------------------------------------------------------------
// A tile holds te <= blockDim.x elements: the first te threads do the
// per-element work, all threads share the EOS repetitions (small meshes
// use te = 64 so a tile's repetitions spread over four times the lanes).
template <bool U>
__global__ __launch_bounds__(256) void k_qeos(const DomDev *doms, const Tile *tiles, int ntiles, int te)
{
    // 20 KB of LDS in all (eight workgroups per CU): the EOS inputs and
    // the three outputs; the workgroup minima reuse an input row
    __shared__ double s_in[7][256];
    __shared__ double s_out[3][256];
    const Tile t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
    // the tile's subdomain is uniform over the workgroup: one scalar copy of
    // its descriptor, which no store in the kernel can be taken to alias
    const DomDev D = doms[__builtin_amdgcn_readfirstlane(t.dom)];
    GroupCtl *G = D.grp;
    // the stop flag is tested after the element phase's loads (reads and
    // LDS writes only: a stopped group's arrays are still allocated)
    const bool stopped = group_stopped(*G);
    const double qstop = *gp(&G->qstop);
    // the tile's first EOS table entry, in flight during the element phase
    const auto tab = gp(D.evtab) + t.eoff;
    const int total = t.ecount;
    unsigned ent = tab[min((int)threadIdx.x, total - 1)];
    const int tid = threadIdx.x, lane = tid & 63;
    const int el = t.start + tid;
    const bool live = tid < te && el < D.numElem;
    bool q_err = false;   // raised after the element's work: no atomic on the load path
    // also read by the final phase: loaded once, with the element's other fields
    double vdov = 0.0, mass = 0.0, volo = 0.0, arealg = 0.0;
    bool clamped = false;   // vnew was moved into [eosvmin, eosvmax]
    if (live) {
        // every load of the element first (none behind a branch), then the
        // arithmetic
        const double gxi = gld(D.delv_xi, el), geta = gld(D.delv_eta, el), gzeta = gld(D.delv_zeta, el);
        vdov = gld(D.vdov, el);
        const double vnew_raw = gld(D.vnew, el);
        const double delx_xi = gld(D.delx_xi, el), delx_eta = gld(D.delx_eta, el);
        const double delx_zeta = gld(D.delx_zeta, el);
        mass = gld(D.elemMass, el);
        volo = gld(D.volo, el);
        arealg = gld(D.arealg, el);
        const double q_old = gld(D.q, el), e_old = gld(D.e, el), delv = gld(D.delv, el), p_old = gld(D.p, el);
        // --- MonoQ region ---
        int ix[6];
        bool keep[6];
        if (U) {
            const int4 *fp = reinterpret_cast<const int4 *>(D.u_face) + 2 * (size_t)el;
            const int4 a = fp[0], b = fp[1];
            const int bc = b.z;
            ix[0] = face_idx_u(el, a.x, bc & kXiM, kXiMSymm, kXiMComm, keep[0]);
            ix[1] = face_idx_u(el, a.y, bc & kXiP, kXiPSymm, kXiPComm, keep[1]);
            ix[2] = face_idx_u(el, a.z, bc & kEtaM, kEtaMSymm, kEtaMComm, keep[2]);
            ix[3] = face_idx_u(el, a.w, bc & kEtaP, kEtaPSymm, kEtaPComm, keep[3]);
            ix[4] = face_idx_u(el, b.x, bc & kZetaM, kZetaMSymm, kZetaMComm, keep[4]);
            ix[5] = face_idx_u(el, b.y, bc & kZetaP, kZetaPSymm, kZetaPComm, keep[5]);
        } else {
            const int nx = D.nx;
            const int k = el / (nx * nx);
            const int rem = el - k * nx * nx;
            const int j = rem / nx;
            const int i = rem - j * nx;
            ix[0] = face_idx(el, i == 0, el - 1, D.ghost[4], k * nx + j, D.col == 0, keep[0]);
            ix[1] = face_idx(el, i == nx - 1, el + 1, D.ghost[5], k * nx + j, false, keep[1]);
            ix[2] = face_idx(el, j == 0, el - nx, D.ghost[2], k * nx + i, D.row == 0, keep[2]);
            ix[3] = face_idx(el, j == nx - 1, el + nx, D.ghost[3], k * nx + i, false, keep[3]);
            ix[4] = face_idx(el, k == 0, el - nx * nx, D.ghost[0], j * nx + i, D.plane == 0, keep[4]);
            ix[5] = face_idx(el, k == nx - 1, el + nx * nx, D.ghost[1], j * nx + i, false, keep[5]);
        }
        const double r0 = gld(D.delv_xi, ix[0]), r1 = gld(D.delv_xi, ix[1]);
        const double r2 = gld(D.delv_eta, ix[2]), r3 = gld(D.delv_eta, ix[3]);
        const double r4 = gld(D.delv_zeta, ix[4]), r5 = gld(D.delv_zeta, ix[5]);
        // all loads issued; nothing above uses a loaded value
        __builtin_amdgcn_sched_barrier(0);
        const double xim = keep[0] ? r0 : 0.0, xip = keep[1] ? r1 : 0.0;
        const double etam = keep[2] ? r2 : 0.0, etap = keep[3] ? r3 : 0.0;
        const double zetam = keep[4] ? r4 : 0.0, zetap = keep[5] ? r5 : 0.0;
        const double phixi = monoq_phi(gxi, xim, xip);
        const double phieta = monoq_phi(geta, etam, etap);
        const double phizeta = monoq_phi(gzeta, zetam, zetap);
        double qlin, qquad;
        if (vdov > 0.) {
            qlin = 0.; qquad = 0.;
        } else {
            double dxi = gxi * delx_xi;
            double deta = geta * delx_eta;
            double dzeta = gzeta * delx_zeta;
            if (dxi > 0.) dxi = 0.;
            if (deta > 0.) deta = 0.;
            if (dzeta > 0.) dzeta = 0.;
            const double rho = mass / (volo * vnew_raw);
            qlin = -Phys::qlc_monoq * rho *
                   (dxi * (1. - phixi) + deta * (1. - phieta) + dzeta * (1. - phizeta));
            qquad = Phys::qqc_monoq * rho *
                    (dxi * dxi * (1. - phixi * phixi) + deta * deta * (1. - phieta * phieta) +
                     dzeta * dzeta * (1. - phizeta * phizeta));
        }

        // --- qstop on the q of the previous cycle (lulesh.cc:2531-2546) ---
        q_err = q_old > qstop;

        // --- EOS inputs (vnew clamped to [eosvmin, eosvmax], lulesh.cc:2642-2650) ---
        double vnewc = vnew_raw;
        vnewc = fmin(raise_to(vnewc, Phys::eosvmin), Phys::eosvmax);
        clamped = vnewc != vnew_raw;
        s_in[0][tid] = vnewc;
        s_in[1][tid] = e_old;
        s_in[2][tid] = delv;
        s_in[3][tid] = p_old;
        s_in[4][tid] = q_old;
        s_in[5][tid] = qquad;
        s_in[6][tid] = qlin;
    }
    // uniform over the workgroup unless another workgroup raises an error
    // status meanwhile: the first error code is kept and the group's state
    // past it is void, as in the reference, which exits there
    if (stopped) return;
    __syncthreads();

    // --- EOS repetitions, dealt over the workgroup ---
    {
        const int bs = blockDim.x;
        const unsigned base = (unsigned)t.start & kEvMask;
        for (int w = tid; w < total; w += bs) {
            // the next entry, in flight meanwhile; loaded unconditionally
            // (clamped index) so the wait for it can sit at the loop's end
            const unsigned next = tab[min(w + bs, total - 1)];
            const int o = (int)(((ent & kEvMask) - base) & kEvMask);
            const EosOut r = eos_rep(s_in[0][o], s_in[1][o], s_in[2][o], s_in[3][o], s_in[4][o],
                                     s_in[5][o], s_in[6][o]);
            if (ent & kEvLast) {
                s_out[0][o] = r.p;
                s_out[1][o] = r.e;
                s_out[2][o] = r.q;
            }
            ent = next;
        }
    }
    __syncthreads();

    unsigned long long cbits = dbits(1.0e+20), hbits = dbits(1.0e+20);
    if (live) {
        const double p_new = s_out[0][tid], e_new = s_out[1][tid];
        const double vnewc = s_in[0][tid];
        // the reference clamps the vnew array itself (lulesh.cc:2642-2650)
        if (clamped) gst(D.vnew, el, vnewc);
        gst(D.p, el, p_new);
        gst(D.e, el, e_new);
        gst(D.q, el, s_out[2][tid]);
        const double c1s = 2.0 / 3.0;
        const double ss = ss_of(c1s, e_new, vnewc, c1s * (eos_compression(vnewc) + 1.), p_new);
        gst(D.ss, el, ss);
        double vfin = vnewc;
        if (fabs(vfin - 1.0) < Phys::v_cut) vfin = 1.0;
        gst(D.v, el, vfin);
        // the next cycle's hourglass coefficient needs exactly these ss, v
        gst(D.hgcoef, el, hg_coefficient(ss, mass, volo, vfin));

        // --- time constraints ---
        if (vdov != 0.) {
            const double qqc2 = 64.0 * Phys::qqc * Phys::qqc;
            double dtf = ss * ss;
            if (vdov < 0.) dtf = dtf + qqc2 * arealg * arealg * vdov * vdov;
            dtf = sqrt_pos(dtf);                   // dtf >= ss^2 >= 1.1e-37
            dtf = arealg / dtf;
            const unsigned long long bb = dbits(dtf);
            cbits = bb < cbits ? bb : cbits;
            const double dtdvov = Phys::dvovmax / (fabs(vdov) + 1.e-20);
            const unsigned long long h = dbits(dtdvov);
            hbits = h < hbits ? h : hbits;
        }
    }
    // positive doubles order like their bit patterns: MIN via integer MIN;
    // reduce the workgroup first, then one atomic per workgroup and only
    // when it can lower the stored minimum (all workgroups of a subdomain
    // target the same two words)
    // (the last input row is free since the barrier after the repetitions)
    unsigned long long *s_min = reinterpret_cast<unsigned long long *>(&s_in[6][0]);
    cbits = wave_min_u64(cbits);
    hbits = wave_min_u64(hbits);
    if (lane == 0) {
        s_min[tid >> 6] = cbits;
        s_min[4 + (tid >> 6)] = hbits;
    }
    __syncthreads();
    if (tid == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int w = 1; w < nw; ++w) {
            cbits = s_min[w] < cbits ? s_min[w] : cbits;
            hbits = s_min[4 + w] < hbits ? s_min[4 + w] : hbits;
        }
        DomCtl &C = *D.ctl;
        if (cbits < *(volatile unsigned long long *)&C.dtcourant_bits) atomicMin(&C.dtcourant_bits, cbits);
        if (hbits < *(volatile unsigned long long *)&C.dthydro_bits) atomicMin(&C.dthydro_bits, hbits);
    }
    if (q_err) raise_status(G, -2);
}

// ---------------------------------------------------------------------------
// device_local halo exchange
// ---------------------------------------------------------------------------

// pack (CommSend, lulesh-comm.cc:357-844): gather boundary values of the
// sender into the receiver's buffer, fields back to back
__global__ __launch_bounds__(256) void k_pack(const PackEnt *ents, int n, const double *const *fptr,
                                              int nfields, const GroupCtl *gate)
{
    if (gate && group_stopped(*gate)) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const PackEnt e = ents[t];
    for (int f = 0; f < nfields; ++f)
        e.dst[(size_t)f * e.cnt] = fptr[e.src_dom * nfields + f][e.src_idx];
}

// unpack: mode 0 = add in slot order (CommSBN, lulesh-comm.cc:848-1257),
// 1 = overwrite, last slot wins (CommSyncPosVel, :1261-1680),
// 2 = ghost copy (CommMonoQ, :1684-1837)
__global__ __launch_bounds__(256) void k_unpack(const UnpackTgt *tg, int n, const UnpackSrc *src,
                                                double *const *fptr, int nfields, int mode,
                                                const GroupCtl *gate)
{
    if (gate && group_stopped(*gate)) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const UnpackTgt u = tg[t];
    for (int f = 0; f < nfields; ++f) {
        double *fp = fptr[u.dom * nfields + f];
        if (mode == 0) {
            double v = fp[u.idx];
            for (int m = 0; m < u.count; ++m) {
                const UnpackSrc s = src[u.first + m];
                v += s.src[(size_t)f * s.cnt];
            }
            fp[u.idx] = v;
        } else {
            const UnpackSrc s = src[u.first + u.count - 1];
            fp[u.idx] = s.src[(size_t)f * s.cnt];
        }
    }
}

// a corner message applied on its own (backends with corner hooks): one
// word per field onto one node, added (CommSBN) or copied (CommSyncPosVel)
struct CornerArgs {
    double *f[6];
    const double *src;
    int idx, nf, add;
};

__global__ __launch_bounds__(64) void k_apply_corner(CornerArgs a, const GroupCtl *gate)
{
    if (gate && group_stopped(*gate)) return;
    const int f = threadIdx.x;
    if (f >= a.nf) return;
    const double v = a.src[f];
    if (a.add) a.f[f][a.idx] += v;
    else a.f[f][a.idx] = v;
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static inline int grid_of(int n, int b) { return (n + b - 1) / b; }

hipError_t launch_time_increment(GroupCtl *const *groups, int ngroups, DomDev *doms, const int *dom0,
                                 hipStream_t s)
{
    // one 64-lane wave per group, four groups per workgroup
    hipLaunchKernelGGL(k_time_increment, dim3(grid_of(ngroups, 4)), dim3(256), 0, s, groups, ngroups, doms,
                       dom0);
    return hipGetLastError();
}

hipError_t launch_elem_force(bool umesh, const DomDev *d, const Tile *t, int nt, int bs, hipStream_t s)
{
    if (nt == 0) return hipSuccess;
    if (umesh) hipLaunchKernelGGL(k_elem_force<true>, dim3(nt), dim3(bs), 0, s, d, t, nt);
    else hipLaunchKernelGGL(k_elem_force<false>, dim3(nt), dim3(bs), 0, s, d, t, nt);
    return hipGetLastError();
}

hipError_t launch_node_gather(bool umesh, bool fuse, const DomDev *d, const Tile *t, int nt, int bs,
                              hipStream_t s)
{
    if (nt == 0) return hipSuccess;
    if (umesh) {
        if (fuse) hipLaunchKernelGGL((k_node_gather<true, true>), dim3(nt), dim3(bs), 0, s, d, t, nt);
        else hipLaunchKernelGGL((k_node_gather<false, true>), dim3(nt), dim3(bs), 0, s, d, t, nt);
    } else {
        if (fuse) hipLaunchKernelGGL((k_node_gather<true, false>), dim3(nt), dim3(bs), 0, s, d, t, nt);
        else hipLaunchKernelGGL((k_node_gather<false, false>), dim3(nt), dim3(bs), 0, s, d, t, nt);
    }
    return hipGetLastError();
}

hipError_t launch_node_update(bool umesh, const DomDev *d, const Tile *t, int nt, int bs, hipStream_t s)
{
    if (nt == 0) return hipSuccess;
    if (umesh) hipLaunchKernelGGL(k_node_update<true>, dim3(nt), dim3(bs), 0, s, d, t, nt);
    else hipLaunchKernelGGL(k_node_update<false>, dim3(nt), dim3(bs), 0, s, d, t, nt);
    return hipGetLastError();
}

hipError_t launch_node_update_direct(const DomDev *d, const Tile *t, int nt, int bs, hipStream_t s)
{
    if (nt == 0) return hipSuccess;
    hipLaunchKernelGGL(k_node_update_direct, dim3(nt), dim3(bs), 0, s, d, t, nt);
    return hipGetLastError();
}

hipError_t launch_pv_snapshot(const DomDev *d, const Tile *t, int nt, int bs, hipStream_t s)
{
    if (nt == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pv_snapshot, dim3(nt), dim3(bs), 0, s, d, t, nt);
    return hipGetLastError();
}

hipError_t launch_kinematics(bool umesh, const DomDev *d, const Tile *t, int nt, int bs, hipStream_t s)
{
    if (nt == 0) return hipSuccess;
    if (umesh) hipLaunchKernelGGL(k_kinematics<true>, dim3(nt), dim3(bs), 0, s, d, t, nt);
    else hipLaunchKernelGGL(k_kinematics<false>, dim3(nt), dim3(bs), 0, s, d, t, nt);
    return hipGetLastError();
}

hipError_t launch_qeos(bool umesh, const DomDev *d, const Tile *t, int nt, int te, int bs, hipStream_t s)
{
    if (nt == 0) return hipSuccess;
    if (te < 1 || te > bs || bs > 256) return hipErrorInvalidValue;
    if (umesh) hipLaunchKernelGGL(k_qeos<true>, dim3(nt), dim3(bs), 0, s, d, t, nt, te);
    else hipLaunchKernelGGL(k_qeos<false>, dim3(nt), dim3(bs), 0, s, d, t, nt, te);
    return hipGetLastError();
}

hipError_t launch_pack(const PackEnt *e, int n, const double *const *fptr, int nfields,
                       const GroupCtl *gate, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack, dim3(grid_of(n, 256)), dim3(256), 0, s, e, n, fptr, nfields, gate);
    return hipGetLastError();
}

hipError_t launch_unpack(const UnpackTgt *t, int n, const UnpackSrc *src, double *const *fptr,
                         int nfields, int mode, const GroupCtl *gate, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_unpack, dim3(grid_of(n, 256)), dim3(256), 0, s, t, n, src, fptr, nfields, mode,
                       gate);
    return hipGetLastError();
}

hipError_t launch_apply_corner(double *const f[6], int nf, int idx, const double *src, bool add,
                               const GroupCtl *gate, hipStream_t s)
{
    CornerArgs a;
    for (int i = 0; i < 6; ++i) a.f[i] = i < nf ? f[i] : nullptr;
    a.src = src;
    a.idx = idx;
    a.nf = nf;
    a.add = add ? 1 : 0;
    hipLaunchKernelGGL(k_apply_corner, dim3(1), dim3(64), 0, s, a, gate);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Sedov set-up on the device (Domain::Domain, lulesh-init.cc:16-184)
// ---------------------------------------------------------------------------

// BuildMesh coordinates, lulesh-init.cc:192-212: 1.125 * (loc*nx + i) / (tp*nx)
__global__ __launch_bounds__(256) void k_setup_coords(DomDev D)
{
    const int nd = blockIdx.x * blockDim.x + threadIdx.x;
    if (nd >= D.numNode) return;
    const int nx = D.nx, s1 = nx + 1;
    const int k = nd / (s1 * s1), rem = nd - k * s1 * s1, j = rem / s1, i = rem - j * s1;
    const double mesh = (double)(D.tp * nx);
    D.ns[0].x[nd] = 1.125 * (double)(D.col * nx + i) / mesh;
    D.ns[0].y[nd] = 1.125 * (double)(D.row * nx + j) / mesh;
    D.ns[0].z[nd] = 1.125 * (double)(D.plane * nx + k) / mesh;
}

// element volumes (lulesh-init.cc:151-162): volo = elemMass = volume, v = 1
template <bool U>
__global__ __launch_bounds__(256) void k_setup_volumes(DomDev D)
{
    const int el = blockIdx.x * blockDim.x + threadIdx.x;
    if (el >= D.numElem) return;
    int n[8];
    elem_nodes_of<U>(D, el, n);
    double x[8], y[8], z[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) { x[c] = D.ns[0].x[n[c]]; y[c] = D.ns[0].y[n[c]]; z[c] = D.ns[0].z[n[c]]; }
    const double vol = elem_volume(x, y, z);
    D.volo[el] = vol;
    D.elemMass[el] = vol;
    D.v[el] = 1.0;
}

// nodal masses (lulesh-init.cc:163-167): the reference adds volume/8 of
// each element, elements in ascending order, onto a node mass that starts
// at 0.0; per node that is the ordered sum over its elements
template <bool U>
__global__ __launch_bounds__(256) void k_setup_nodal_mass(DomDev D)
{
    const int nd = blockIdx.x * blockDim.x + threadIdx.x;
    if (nd >= D.numNode) return;
    if (U) {
        // an unstructured mesh: the node's elements, ascending, from its
        // corner list (a node twice in one element counts twice, as the
        // reference's per-corner loop adds it twice)
        double m = 0.0;
        for (int q = D.u_nstart[nd]; q < D.u_nstart[nd + 1]; ++q) m += D.volo[D.u_ncorner[q] >> 3] / 8.0;
        D.nodalMass[nd] = m;
        return;
    }
    const int nx = D.nx, s1 = nx + 1;
    const int k = nd / (s1 * s1), rem = nd - k * s1 * s1, j = rem / s1, i = rem - j * s1;
    double m = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int a = i - 1 + (q & 1), b = j - 1 + ((q >> 1) & 1), c = k - 1 + (q >> 2);
        if (a < 0 || a >= nx || b < 0 || b >= nx || c < 0 || c >= nx) continue;
        m += D.volo[(c * nx + b) * nx + a] / 8.0;
    }
    D.nodalMass[nd] = m;
}

hipError_t launch_setup(const DomDev &D, bool umesh, hipStream_t s)
{
    if (umesh) {
        // coordinates come from the mesh file
        hipLaunchKernelGGL(k_setup_volumes<true>, dim3(grid_of(D.numElem, 256)), dim3(256), 0, s, D);
        hipLaunchKernelGGL(k_setup_nodal_mass<true>, dim3(grid_of(D.numNode, 256)), dim3(256), 0, s, D);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_setup_coords, dim3(grid_of(D.numNode, 256)), dim3(256), 0, s, D);
    hipLaunchKernelGGL(k_setup_volumes<false>, dim3(grid_of(D.numElem, 256)), dim3(256), 0, s, D);
    hipLaunchKernelGGL(k_setup_nodal_mass<false>, dim3(grid_of(D.numNode, 256)), dim3(256), 0, s, D);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_hg_coef(double *coef, const double *ss, const double *mass,
                                                 const double *volo, const double *v, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) coef[t] = hg_coefficient(ss[t], mass[t], volo[t], v[t]);
}

hipError_t launch_hg_coef(double *coef, const double *ss, const double *mass, const double *volo,
                          const double *v, int n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_hg_coef, dim3(grid_of(n, 256)), dim3(256), 0, s, coef, ss, mass, volo, v, n);
    return hipGetLastError();
}

// element-wise probes of the device math, used by the parity tests
__global__ void k_probe_cbrt(const double *in, double *out, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = cbrt_cr(in[t]);
}

__global__ void k_probe_sqrt_div(const double *a, const double *b, double *out, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        out[2 * t] = sqrt(a[t]);
        out[2 * t + 1] = a[t] / b[t];
    }
}

__global__ void k_probe_range_math(const double *a, double *out, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        out[2 * t] = sqrt_pos(a[t]);
        out[2 * t + 1] = rcp_pos(a[t]);
    }
}

// streaming copy, the bandwidth yardstick of the profiling runs
__global__ __launch_bounds__(256) void k_probe_copy(const double2 *__restrict__ a, double2 *__restrict__ b,
                                                    size_t n2)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n2; t += stride) b[t] = a[t];
}

// the loop state of every group of an engine after its chunk of cycles,
// into pinned host memory: the host reads it after the stream completes
// instead of copying each group's control block
__global__ __launch_bounds__(64) void k_ctl_snapshot(GroupCtl *const *groups, int ngroups, CtlSnap *out)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    const GroupCtl &G = *groups[g];
    const DomCtl &C = G.dctl[0];
    CtlSnap v;
    v.done = G.done;
    v.status = G.status;
    v.cycle = C.cycle;
    v.pad = 0;
    v.time = C.time;
    v.deltatime = C.deltatime;
    out[g] = v;
}

// FP64 issue rate: 16 multiplies and adds per loop trip and lane, as eight
// independent chains or (ONE) as a single chain where each operation reads
// the one before (no memory traffic, no FMA); thread 0 of each workgroup
// records the shader cycles and the 100 MHz real-time ticks its loop took
template <bool ONE>
__global__ __launch_bounds__(256) void k_probe_fp64(double *out, int iters, unsigned long long *clk)
{
    double a[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) a[c] = 1.0 + 1e-9 * threadIdx.x + 0.125 * c;
    const double m = 0.9999999999, s = 1e-13;
    const unsigned long long t0 = __builtin_readcyclecounter(), r0 = wall_clock64();
    for (int i = 0; i < iters; ++i) {
        if (ONE) {
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                a[0] = a[0] * m;
                a[0] = a[0] + s;
            }
        } else {
#pragma unroll
            for (int c = 0; c < 8; ++c) a[c] = a[c] * m;
#pragma unroll
            for (int c = 0; c < 8; ++c) a[c] = a[c] + s;
        }
    }
    const unsigned long long t1 = __builtin_readcyclecounter(), r1 = wall_clock64();
    if (threadIdx.x == 0) {
        clk[2 * blockIdx.x] = t1 - t0;
        clk[2 * blockIdx.x + 1] = r1 - r0;
    }
    double t = 0.0;
#pragma unroll
    for (int c = 0; c < 8; ++c) t += a[c];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = t;
}

// message i is copied by the workgroups of grid row i
__global__ __launch_bounds__(256) void k_copy_msgs(CopyMsgs m)
{
    const int i = blockIdx.y;
    double *const dst = m.dst[i];
    const double *const src = m.src[i];
    const unsigned n = m.count[i];
    for (unsigned w = blockIdx.x * blockDim.x + threadIdx.x; w < n; w += gridDim.x * blockDim.x)
        dst[w] = src[w];
}

hipError_t launch_copy_msgs(const CopyMsgs &m, hipStream_t s)
{
    if (m.n <= 0) return hipSuccess;
    if (m.n > kMaxCopyMsgs) return hipErrorInvalidValue;
    unsigned longest = 0;
    for (int i = 0; i < m.n; ++i) longest = m.count[i] > longest ? m.count[i] : longest;
    const unsigned bx = longest == 0 ? 1u : (longest + 255u) / 256u < 64u ? (longest + 255u) / 256u : 64u;
    hipLaunchKernelGGL(k_copy_msgs, dim3(bx, m.n), dim3(256), 0, s, m);
    return hipGetLastError();
}

hipError_t launch_probe_copy(const double *a, double *b, size_t n, hipStream_t s)
{
    const size_t n2 = n / 2;
    hipLaunchKernelGGL(k_probe_copy, dim3(4096), dim3(256), 0, s, (const double2 *)a, (double2 *)b, n2);
    return hipGetLastError();
}

hipError_t launch_ctl_snapshot(GroupCtl *const *groups, int ngroups, CtlSnap *out, hipStream_t s)
{
    if (ngroups < 1) return hipSuccess;
    hipLaunchKernelGGL(k_ctl_snapshot, dim3((ngroups + 63) / 64), dim3(64), 0, s, groups, ngroups, out);
    return hipGetLastError();
}

// out: grid x 256 doubles, clk: 2 x grid counters
hipError_t launch_probe_fp64(double *out, unsigned long long *clk, int grid, int iters, bool one_chain,
                             hipStream_t s)
{
    if (grid < 1 || iters < 1) return hipErrorInvalidValue;
    if (one_chain) hipLaunchKernelGGL(k_probe_fp64<true>, dim3(grid), dim3(256), 0, s, out, iters, clk);
    else hipLaunchKernelGGL(k_probe_fp64<false>, dim3(grid), dim3(256), 0, s, out, iters, clk);
    return hipGetLastError();
}

hipError_t launch_probe_cbrt(const double *in, double *out, int n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_probe_cbrt, dim3(grid_of(n, 256)), dim3(256), 0, s, in, out, n);
    return hipGetLastError();
}

hipError_t launch_probe_sqrt_div(const double *a, const double *b, double *out, int n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_probe_sqrt_div, dim3(grid_of(n, 256)), dim3(256), 0, s, a, b, out, n);
    return hipGetLastError();
}

hipError_t launch_probe_range_math(const double *a, double *out, int n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_probe_range_math, dim3(grid_of(n, 256)), dim3(256), 0, s, a, out, n);
    return hipGetLastError();
}

}  // namespace lulesh
