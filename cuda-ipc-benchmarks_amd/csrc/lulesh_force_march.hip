// lulesh_force_march.hip -- the fused force kernel of the block path.
//
// Built on its own: the FP64-issue-bound march is scheduled for
// instruction-level parallelism (-amdgpu-sched-strategy=max-ilp in the
// Makefile), which the register-bound element kernels of lulesh_kernels.hip
// would pay for with occupancy.
#include <hip/hip_runtime.h>

#include "lulesh_dev.h"
#include "lulesh_kernel_util.h"
#include "lulesh_kernels.h"
#include "lulesh_math.h"

namespace lulesh {

// ---------------------------------------------------------------------------
// Fused force + ordered node gather, z-marching (the production path).
//
// A workgroup owns a column of (ex-1) x (ey-1) nodes and a chunk of node
// planes [k0, k1).  Its ex x ey threads (<= 256; 16 x 16 for large meshes,
// shapes fitted to the mesh for small ones) are the elements of the
// column's halo-extended element tile; the block walks element planes c = k0-1 .. k1-1.
// At each plane every thread computes its element's 8 stress corner forces
// and 8 hourglass corner forces (the same arithmetic as k_elem_force) into
// LDS, and each node thread adds its four neighbouring elements' corners:
// corners 2,3,1,0 of plane c complete node plane c, corners 6,7,5,4 start
// node plane c+1.  That is exactly the reference's per-node order (elements
// ascending: plane c-1 elements (i-1,j-1),(i,j-1),(i-1,j),(i,j), then the
// same four of plane c), summed from 0.0, stress and hourglass separately,
// f = stress + hourglass -- so the result is the reference's to the bit,
// without the 48-doubles-per-element corner array ever touching HBM.
// Elements outside the mesh contribute +0.0, which leaves a sum that
// started from +0.0 unchanged.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256, 2) void k_force_march(DomDev *doms, const ColTile *tiles, int ntiles,
                                                        int advance, int store_f, int ti)
{
    // corner forces of one phase, in pairs: lds[d][s][t] = component d of
    // element t's corners p and p+4 (the node-plane-c and node-plane-c+1
    // corners one node gathers from the same element), one 16-byte access,
    // for p = 0, 1, 3 in slots s = 0, 1, 2.  Corners 2 and 6 belong to the
    // node the element's own thread gathers: that pair stays in registers.
    __shared__ double2 lds[3][3][kMThreads];
    double2 own_c[3];
    __shared__ double sn[2][6][kPatch];          // node planes (x y z xd yd zd), by plane parity
    __shared__ double ti_dt;                     // this cycle's deltatime, or -1: the group stops
    const ColTile t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
    // the tile's subdomain is uniform over the workgroup: one scalar copy of
    // its descriptor, which no store in the kernel can be taken to alias
    const int di = __builtin_amdgcn_readfirstlane(t.dom);
    const DomDev D = doms[di];
    GroupCtl *G = D.grp;
    // The stop flag is tested once the first plane's loads are in flight
    // (reads only: a stopped group's arrays are still allocated), by wave 0
    // for the whole workgroup: the flag can change while the workgroup's
    // waves read it (the time increment below raises it in one workgroup
    // of the launch, and a volume error in any), and a wave that left
    // alone would leave its slots of the node patch unstaged for the rest.
    // Node-state buffers of the cycle about to run (cycle n = completed + 1
    // reads buffer (n-1)&1 and writes buffer n&1), from the completed
    // cycle count, which nothing in this launch writes.  With ti = 0 a
    // time-increment kernel has counted the cycle already.
    DomCtl *const Cg = reinterpret_cast<DomCtl *>(G + 1);   // the group's DomCtl array
    const int rank = (int)(D.ctl - Cg);
    const int done_cycles = ti ? *gp(&D.ctl->cycle) : *gp(&D.ctl->cycle) - 1;
    const NodeState S = (done_cycles & 1) ? D.ns[1] : D.ns[0];
    const NodeState Snew = (done_cycles & 1) ? D.ns[0] : D.ns[1];
    const int nx = D.nx, s1 = nx + 1, s2 = s1 * s1;
    const int tid = threadIdx.x;
    const int ex = t.ex, ey = t.ey, px = ex + 1;
    const int ta = tid % ex, tb = tid / ex;
    const bool slot_ok = tb < ey;
    const int a = t.i0 - 1 + ta, b = t.j0 - 1 + tb;
    const bool col_ok = slot_ok && a >= 0 && a < nx && b >= 0 && b < nx;
    const int ni = t.i0 + ta, nj = t.j0 + tb;
    const bool node_ok = ta < ex - 1 && tb < ey - 1 && ni <= nx && nj <= nx;
    const int k0 = t.k0;
    const int k1 = min(t.k0 + t.kn, nx + 1);
    // lower / upper corner slots seen from the node thread:
    // own element, +x, +y, +x+y neighbours
    const int t1 = tid + 1, t2 = tid + ex, t3 = tid + ex + 1;

    // Node patch: node (i0-1+pi, j0-1+pj) of a node plane, pi in [0, ex],
    // pj in [0, ey], sits in slot pj*ex + pi for pi < ex and in a last
    // column of its own, ex*(ey+1) + pj, for pi = ex.  Rows of ex slots
    // match the rows of ex threads, so a wave's lanes read consecutive
    // slots (no bank conflict where a group of lanes spans a row end), save
    // the one lane per row whose corner lies in the last column.  The
    // element's corners 0..3 are nodes (ta, tb), (ta+1, tb), (ta+1, tb+1),
    // (ta, tb+1) of its lower plane, 4..7 the same of the upper one.  A
    // thread loads the slot of the node it owns, (ta+1, tb+1) = corner 2
    // of its element; the first ex+ey+1 threads also load one slot of the
    // patch's row 0 / column 0.
    auto slot = [&](int pi, int pj) { return pi < ex ? pj * ex + pi : ex * (ey + 1) + pj; };
    const int cs[4] = {slot(ta, tb), slot(ta + 1, tb), slot(ta + 1, tb + 1), slot(ta, tb + 1)};
    const int own_sl = slot(ta + 1, tb + 1);
    const int own_i = ni, own_j = nj;
    const bool own_in = slot_ok && own_i <= nx && own_j <= nx;
    int xt_sl = -1, xt_i = 0, xt_j = 0;
    if (tid < ex + ey + 1) {
        const int pi = tid < px ? tid : 0, pj = tid < px ? 0 : tid - px + 1;
        xt_sl = slot(pi, pj);
        xt_i = t.i0 - 1 + pi;
        xt_j = t.j0 - 1 + pj;
        if (xt_i < 0 || xt_i > nx || xt_j < 0 || xt_j > nx) xt_sl = -1;
    }
    double pf_own[6], pf_xt[6];
    // The loads of a plane are issued on every path -- at a valid node
    // (node 0) where the slot or the plane does not exist, the values then
    // unused -- so the number of loads in flight is the same on all paths
    // and the waits for the older ones can be counted exactly instead of
    // draining the whole queue.
    auto fetch = [&](int kn, double own[6], double xt[6]) {   // issue the loads of node plane kn
        const int kc = kn < 0 ? 0 : (kn > nx ? nx : kn);
        const unsigned nd = opaque_idx(own_in ? kc * s2 + own_j * s1 + own_i : 0);
        own[0] = gld(S.x, nd); own[1] = gld(S.y, nd); own[2] = gld(S.z, nd);
        own[3] = gld(S.xd, nd); own[4] = gld(S.yd, nd); own[5] = gld(S.zd, nd);
        const unsigned nt = opaque_idx(xt_sl >= 0 ? kc * s2 + xt_j * s1 + xt_i : 0);
        xt[0] = gld(S.x, nt); xt[1] = gld(S.y, nt); xt[2] = gld(S.z, nt);
        xt[3] = gld(S.xd, nt); xt[4] = gld(S.yd, nt); xt[5] = gld(S.zd, nt);
    };
    auto stage = [&](int kn, const double own[6], const double xt[6]) {   // write plane kn to LDS
        if (kn < 0 || kn > nx) return;
        double (*P)[kPatch] = sn[kn & 1];
        if (own_in)
#pragma unroll
            for (int f = 0; f < 6; ++f) P[f][own_sl] = own[f];
        if (xt_sl >= 0)
#pragma unroll
            for (int f = 0; f < 6; ++f) P[f][xt_sl] = xt[f];
    };
    // element fields of plane c (p, q, v, volo, hgcoef), fetched a plane ahead
    double ef[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, ef_next[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    auto fetch_elem = [&](int c, double out[5]) {
        const unsigned el = opaque_idx((col_ok && c >= 0 && c < nx) ? (c * nx + b) * nx + a : 0);
        out[0] = gld(D.p, el); out[1] = gld(D.q, el); out[2] = gld(D.v, el); out[3] = gld(D.volo, el);
        out[4] = gld(D.hgcoef, el);
    };

    {
        // the first two planes' loads all in flight before either is staged
        double own1[6], xt1[6];
        fetch(k0 - 1, pf_own, pf_xt);
        fetch(k0, own1, xt1);
        fetch_elem(k0 - 1, ef);
        if (tid < 64) {
            // this cycle's deltatime, or -1 where the workgroup does nothing
            const bool stopped = __builtin_amdgcn_readfirstlane((int)group_stopped(*G)) != 0;
            if (stopped) {
                if (tid == 0) ti_dt = -1.0;
            } else if (!ti) {
                if (tid == 0) ti_dt = *gp(&D.ctl->ndeltatime);
            } else {
                // TimeIncrement (lulesh.cc:207-263) of this cycle, computed by
                // every workgroup of the group from the completed cycle's state:
                // lane l = subdomain l (at most 4^3 = 64), the group's MIN of
                // their candidates, then this subdomain's clamps
                const int ndom = *gp(&G->ndom);
                const double stoptime = *gp(&G->stoptime);
                const int max_cycles = *gp(&G->max_cycles);
                const bool mine = tid < ndom;
                const DomCtl &Cl = Cg[mine ? tid : 0];
                double newdt = mine ? ti_candidate(*gp(&Cl.dtcourant_bits), *gp(&Cl.dthydro_bits))
                                    : __builtin_inf();
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    const double o = __shfl_xor(newdt, off, 64);
                    newdt = o < newdt ? o : newdt;
                }
                if (tid == 0) {
                    const double time = *gp(&D.ctl->time), olddt = *gp(&D.ctl->deltatime);
                    const int cyc = *gp(&D.ctl->cycle);
                    const bool stop = ti_stop(*gp(&Cg[0].time), *gp(&Cg[0].cycle), stoptime, max_cycles);
                    const double deltatime = stop ? -1.0 : ti_deltatime(newdt, olddt, time, cyc, stoptime);
                    ti_dt = deltatime;
                    if (t.flags == kFirstTile) {
                        // the subdomain's first tile publishes the running
                        // cycle for the cycle's later kernels (no workgroup
                        // of this launch reads the n-fields or the
                        // descriptors' old/cur; the stop flag is read by
                        // each workgroup's wave 0 only, above)
                        if (stop) {
                            if (rank == 0) G->done = 1;
                        } else {
                            DomCtl &C = *D.ctl;
                            C.ndeltatime = deltatime;
                            C.ntime = time + deltatime;
                            C.ncycle = cyc + 1;
                            doms[di].old = S;
                            doms[di].cur = Snew;
                        }
                    }
                }
            }
        }
        stage(k0 - 1, pf_own, pf_xt);
        stage(k0, own1, xt1);
    }
    __syncthreads();
    const double dt = ti_dt;
    if (dt < 0.0) return;   // the group has stopped, or reaches its stop condition now

    // a non-positive volume (lulesh.cc:1263-1265, 1340-1349) is reported
    // once, after the planes: no atomic sits on the loop's path
    bool vol_err = false;
    double ps[3] = {0.0, 0.0, 0.0}, ph[3] = {0.0, 0.0, 0.0};   // node plane c+1 partials
    double fs[3] = {0.0, 0.0, 0.0};                              // node plane c stress sums
    for (int c = k0 - 1; c < k1; ++c) {
        const bool el_ok = col_ok && c >= 0 && c < nx;
        const double (*Lo)[kPatch] = sn[c & 1];
        const double (*Up)[kPatch] = sn[(c + 1) & 1];
        // this plane's nodal masses (first: the update waits for them
        // alone), then the next planes' loads, in flight meanwhile
        const double nmass = gld(D.nodalMass, opaque_idx((node_ok && c >= 0) ? c * s2 + nj * s1 + ni : 0));
        fetch(c + 2, pf_own, pf_xt);
        fetch_elem(c + 1, ef_next);

        // ---- stress corners ----
        if (el_ok) {
            double x[8], y[8], z[8];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                x[q] = Lo[0][cs[q]]; y[q] = Lo[1][cs[q]]; z[q] = Lo[2][cs[q]];
                x[q + 4] = Up[0][cs[q]]; y[q + 4] = Up[1][cs[q]]; z[q + 4] = Up[2][cs[q]];
            }
            // sigma / 16 against the normals x 16: the same products.  The
            // division is exact: p and q are 0 or at least their cut-offs
            const double sig16 = (-ef[0] - ef[1]) * 0.0625;
            vol_err |= jac_volume(x, y, z) <= 0.0;
            double bx[8], by[8], bz[8];
            node_normals16(x, y, z, bx, by, bz);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double2 vx = make_double2(-(sig16 * bx[q]), -(sig16 * bx[q + 4]));
                const double2 vy = make_double2(-(sig16 * by[q]), -(sig16 * by[q + 4]));
                const double2 vz = make_double2(-(sig16 * bz[q]), -(sig16 * bz[q + 4]));
                if (q == 2) {
                    own_c[0] = vx; own_c[1] = vy; own_c[2] = vz;
                } else {
                    const int sl = q == 3 ? 2 : q;
                    lds[0][sl][tid] = vx; lds[1][sl][tid] = vy; lds[2][sl][tid] = vz;
                }
            }
        } else {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                own_c[d] = make_double2(0.0, 0.0);
#pragma unroll
                for (int q = 0; q < 3; ++q) lds[d][q][tid] = make_double2(0.0, 0.0);
            }
        }
        __syncthreads();
        // the gathers between barriers are short and latency-bound: they
        // issue ahead of the other workgroup's element arithmetic on the
        // same SIMD, so this workgroup reaches its next barrier sooner
        __builtin_amdgcn_s_setprio(2);
        if (node_ok) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double2 p2 = own_c[d], p3 = lds[d][2][t1], p1 = lds[d][1][t2], p0 = lds[d][0][t3];
                if (c >= k0) {
                    double v = ps[d];
                    v += p2.x;
                    v += p3.x;
                    v += p1.x;
                    v += p0.x;
                    fs[d] = v;
                }
                if (c + 1 < k1) {
                    double v = 0.0;
                    v += p2.y;
                    v += p3.y;
                    v += p1.y;
                    v += p0.y;
                    ps[d] = v;
                }
            }
        }
        __syncthreads();
        __builtin_amdgcn_s_setprio(0);
        // ---- hourglass corners ----
        if (el_ok) {
            const double vv = ef[2];
            vol_err |= vv <= 0.0;
            HgModes H;
            {
                double x[8], y[8], z[8];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    x[q] = Lo[0][cs[q]]; y[q] = Lo[1][cs[q]]; z[q] = Lo[2][cs[q]];
                    x[q + 4] = Up[0][cs[q]]; y[q + 4] = Up[1][cs[q]]; z[q + 4] = Up[2][cs[q]];
                }
                hg_modes(x, y, z, ef[3] * vv, ef[4], H);
            }
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                double vl[8], h[8];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    vl[q] = Lo[3 + d][cs[q]];
                    vl[q + 4] = Up[3 + d][cs[q]];
                }
                hg_force(H, vl, h);
                own_c[d] = make_double2(h[2], h[6]);
                lds[d][0][tid] = make_double2(h[0], h[4]);
                lds[d][1][tid] = make_double2(h[1], h[5]);
                lds[d][2][tid] = make_double2(h[3], h[7]);
            }
        } else {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                own_c[d] = make_double2(0.0, 0.0);
#pragma unroll
                for (int q = 0; q < 3; ++q) lds[d][q][tid] = make_double2(0.0, 0.0);
            }
        }
        __syncthreads();
        __builtin_amdgcn_s_setprio(2);
        if (node_ok) {
            double f[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double2 p2 = own_c[d], p3 = lds[d][2][t1], p1 = lds[d][1][t2], p0 = lds[d][0][t3];
                if (c >= k0) {
                    double v = ph[d];
                    v += p2.x;
                    v += p3.x;
                    v += p1.x;
                    v += p0.x;
                    f[d] = fs[d] + v;
                }
                if (c + 1 < k1) {
                    double v = 0.0;
                    v += p2.y;
                    v += p3.y;
                    v += p1.y;
                    v += p0.y;
                    ph[d] = v;
                }
            }
            if (c >= k0) {
                const unsigned nd = opaque_idx(c * s2 + nj * s1 + ni);
                if (store_f) {
                    gst(D.fpx, nd, f[0]);
                    gst(D.fpy, nd, f[1]);
                    gst(D.fpz, nd, f[2]);
                }
                if (advance) {
                    // the node is complete: advance it from its old state
                    // (its own slot of the lower plane) into the new buffer
                    const double one_over_nMass = rcp_pos(nmass);   // masses are positive
                    double ax = f[0] * one_over_nMass, ay = f[1] * one_over_nMass,
                           az = f[2] * one_over_nMass;
                    if (D.col == 0 && ni == 0) ax = 0.0;
                    if (D.row == 0 && nj == 0) ay = 0.0;
                    if (D.plane == 0 && c == 0) az = 0.0;
                    double vx = Lo[3][own_sl] + ax * dt;
                    if (fabs(vx) < Phys::u_cut) vx = 0.0;
                    double vy = Lo[4][own_sl] + ay * dt;
                    if (fabs(vy) < Phys::u_cut) vy = 0.0;
                    double vz = Lo[5][own_sl] + az * dt;
                    if (fabs(vz) < Phys::u_cut) vz = 0.0;
                    const NodeState N = Snew;
                    gst(N.xd, nd, vx);
                    gst(N.yd, nd, vy);
                    gst(N.zd, nd, vz);
                    gst(N.x, nd, Lo[0][own_sl] + vx * dt);
                    gst(N.y, nd, Lo[1][own_sl] + vy * dt);
                    gst(N.z, nd, Lo[2][own_sl] + vz * dt);
                }
            }
        }
        // the lower plane is done with (this thread's own slot was read
        // just above): it takes node plane c+2
        if (c + 2 <= k1) stage(c + 2, pf_own, pf_xt);
#pragma unroll
        for (int q = 0; q < 5; ++q) ef[q] = ef_next[q];
        __syncthreads();
        __builtin_amdgcn_s_setprio(0);
    }
    if (vol_err) raise_status(G, -1);
}


hipError_t launch_force_march(DomDev *d, const ColTile *t, int nt, bool advance, bool store_f, bool ti,
                              hipStream_t s)
{
    if (nt == 0) return hipSuccess;
    hipLaunchKernelGGL(k_force_march, dim3(nt), dim3(kMThreads), 0, s, d, t, nt, advance ? 1 : 0,
                       store_f ? 1 : 0, ti ? 1 : 0);
    return hipGetLastError();
}

}  // namespace lulesh
