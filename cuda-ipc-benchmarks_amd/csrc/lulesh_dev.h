// lulesh_dev.h -- device-visible data layout of the MI355X LULESH step.
//
// Layout in HBM (per subdomain, structure of arrays, FP64):
//   node state    x y z xd yd zd, two buffers for a one-subdomain group
//                 (cycle n reads buffer (n-1)&1 and writes buffer n&1, so a
//                 kernel can advance nodes that other workgroups still read)
//   node fields   fx fy fz nodalMass                       [(nx+1)^3]
//   element       e p q v volo delv vdov arealg ss elemMass vnew
//                 delx_xi delx_eta delx_zeta                [nx^3]
//   gradients     delv_xi delv_eta delv_zeta   [nx^3 + 6 nx^2 ghost planes]
//   corner forces corner[48][nx^3]  (stress x,y,z then hourglass x,y,z,
//                 8 corners each; element index fastest -> coalesced)
//   elemRep       int32 [nx^3]
// The mesh is the reference's structured Sedov block (BuildMesh,
// LULESH/omp_4.0/lulesh-init.cc:189-238), so element->node connectivity,
// face neighbours and boundary masks are computed from (i,j,k) instead of
// being read from nodelist / lxim.. / elemBC arrays.
#pragma once
#include <stdint.h>

namespace lulesh {

constexpr int kCornerComps = 48;     // 6 components x 8 corners
// node slots per plane of the force kernel's LDS patch: (ex+1)*(ey+1) must fit
constexpr int kForcePatch = 320;

struct IdxSpan {          // [first, first + count) of a list
    int first;
    int count;
};

struct DirRef {           // an index of another subdomain of the group
    int dom;
    int idx;
};

// a node copy touched by CommSyncPosVel: its final writer (the last
// message's sender, -1 if none) with the writer's node index and snapshot
// slot, and this copy's own snapshot slot (-1 unless it is some other
// copy's final writer)
struct PvNode {
    int fw_dom;
    int fw_idx;
    int fw_slot;
    int own_slot;
};

// node positions and velocities
struct NodeState {
    double *x, *y, *z, *xd, *yd, *zd;
};

// per-subdomain descriptor, read by every kernel through the tile table
struct DomDev {
    // ns[b]: buffer b of the node state; ns[1] aliases ns[0] when the
    // subdomain updates its nodes in place (groups with halo exchanges)
    NodeState ns[2];
    // the buffers the running cycle reads (old: the state it starts from)
    // and writes (cur); k_time_increment sets them in the stepping
    // engine's copy of this descriptor when it counts the cycle
    NodeState old, cur;
    double *fx, *fy, *fz, *nodalMass;
    // where the force kernels store a cycle's forces: fx..fz, or, when the
    // stepping engine sums shared nodes directly (k_node_update_direct),
    // separate arrays holding this subdomain's own partial forces
    double *fpx, *fpy, *fpz;
    // direct halo steps (device_local backend): per node {first, count} of
    // its CommSBN sources in receive-slot order, per element {first,
    // count} of the neighbours' MonoQ ghost slots it fills; null otherwise.
    // dom of a DirRef is a rank of the group, whose subdomain 0 is at
    // index dom_base of the engine's descriptor array.
    const IdxSpan *dir_node;
    const IdxSpan *dir_elem;
    const DirRef *dir_sbn;
    const DirRef *dir_ghost;
    // CommSyncPosVel done in place: per node an index into dir_pvnode, or
    // -1 for a node no PosVel message touches; null when the PosVel
    // exchange runs.  pv_snap holds, by cycle parity, the state after the
    // cycle of this subdomain's nodes that are another copy's final writer
    // ([2][6][pv_nslot]).
    const int *dir_pv;
    const struct PvNode *dir_pvnode;
    double *pv_snap;
    int pv_nslot;
    // the group's node copies with dir_pv >= 0, as (rank, node): the node
    // update runs them in tiles of their own (Tile::pad = 1)
    const DirRef *dir_pvlist;
    int dom_base;
    double *e, *p, *q, *v, *volo, *delv, *vdov, *arealg, *ss, *elemMass, *vnew;
    double *delv_xi, *delv_eta, *delv_zeta, *delx_xi, *delx_eta, *delx_zeta;
    double *corner;
    // FB hourglass coefficient -hgcoef*0.01*ss*elemMass/cbrt(volo*v) of
    // the coming force computation (CalcFBHourglassForceForElems,
    // lulesh.cc:1012-1052), produced where ss and v are produced
    double *hgcoef;
    // unstructured mesh (the reference's .lmesh input): explicit
    // connectivity in place of the (i,j,k) formulas; null for the block.
    //   u_nodelist  [numElem][8] node ids
    //   u_face      [numElem][8] lxim lxip letam letap lzetam lzetap elemBC 0
    //   u_nstart    [numNode+1]  start of each node's corners in u_ncorner
    //   u_ncorner   8*element + corner, elements ascending per node
    //   u_nsymm     per node: 1 symmX, 2 symmY, 4 symmZ
    const int *u_nodelist;
    const int *u_face;
    const int *u_nstart;
    const int *u_ncorner;
    const unsigned char *u_nsymm;
    const int *elemRep;
    // one entry per EOS evaluation, elements in ascending order, each
    // repeated elemRep times (entry format: kEvLast / kEvMask)
    const unsigned short *evtab;
    int nx, numElem, numNode, allElem;
    int col, row, plane, tp;
    // ghost element offsets of the velocity-gradient arrays for the six
    // faces (planeMin, planeMax, rowMin, rowMax, colMin, colMax); -1 when
    // the face is a physical boundary (SetupBoundaryConditions,
    // lulesh-init.cc:533-631)
    int ghost[6];
    struct DomCtl *ctl;      // this subdomain's time state
    struct GroupCtl *grp;    // its group's loop state
};

// per-subdomain time state (every rank of the reference keeps its own).
// time, deltatime and cycle are the state after the last completed cycle;
// ntime, ndeltatime and ncycle the running cycle's, published by the
// cycle's first kernel (its TimeIncrement) and copied into the former by
// the kinematics kernel, so the first kernel of a cycle can compute the
// increment in every workgroup from state nothing in its launch writes.
// Kernels between the two read the n-fields.
struct DomCtl {
    double time, deltatime;
    unsigned long long dtcourant_bits, dthydro_bits;   // atomicMin slots
    int cycle;
    int pad;
    double ntime, ndeltatime;
    int ncycle;
    int npad;
};

// one allocation per group: GroupCtl followed by its ndom DomCtl
struct GroupCtl {
    double stoptime, qstop;
    DomCtl *dctl;     // the group's subdomain time states (contiguous)
    int max_cycles;   // < 0: none
    int done;         // stop condition reached
    int status;       // 0, -1 VolumeError, -2 QStopError
    int ndom;
};

// what a host poll needs of a group after an engine's chunk of cycles: its
// loop state and the first subdomain's completed-cycle time state, copied
// by the chunk's last kernel into pinned host memory (one entry per group)
struct CtlSnap {
    int done, status, cycle, pad;
    double time, deltatime;
};

// element or node tile: [start, start + blockDim) of subdomain dom; for
// element tiles eoff/ecount locate the tile's EOS evaluations in the
// subdomain's evaluation table.  A node-update tile with pad = 1 covers
// entries [start, start + ecount) of the group's PosVel copy list instead,
// dom being the group's first subdomain.
struct Tile {
    int dom;
    int start;
    long long eoff;
    int ecount;
    int pad;
};

// Tile::pad of an element tile that is its subdomain's first (the one
// that completes the running cycle's time state in the kinematics kernel),
// and ColTile::flags of a subdomain's first column tile (the one that
// publishes the cycle's time increment in the force kernel)
constexpr int kFirstTile = 2;

// EOS evaluation table entry: element index mod 2^15 in the low 15 bits,
// bit 15 set on the element's last repetition
constexpr unsigned kEvLast = 0x8000u, kEvMask = 0x7fffu;

// column tile of the fused force kernel: (ex-1) x (ey-1) nodes from
// (i0, j0), node planes [k0, k0 + kn); its ex x ey threads are the
// elements of the tile plus one halo row and column on the low side
struct ColTile {
    int dom;
    int i0, j0, k0;
    int ex, ey;
    int flags;   // kFirstTile or 0
    int kn;      // node planes
};

// elemBC bits (LULESH/omp_4.0/lulesh.h:60-88)
constexpr int kXiM = 0x00007, kXiMSymm = 0x00001, kXiMFree = 0x00002, kXiMComm = 0x00004;
constexpr int kXiP = 0x00038, kXiPSymm = 0x00008, kXiPFree = 0x00010, kXiPComm = 0x00020;
constexpr int kEtaM = 0x001c0, kEtaMSymm = 0x00040, kEtaMFree = 0x00080, kEtaMComm = 0x00100;
constexpr int kEtaP = 0x00e00, kEtaPSymm = 0x00200, kEtaPFree = 0x00400, kEtaPComm = 0x00800;
constexpr int kZetaM = 0x07000, kZetaMSymm = 0x01000, kZetaMFree = 0x02000, kZetaMComm = 0x04000;
constexpr int kZetaP = 0x38000, kZetaPSymm = 0x08000, kZetaPFree = 0x10000, kZetaPComm = 0x20000;

// physics constants of Domain::Domain (LULESH/omp_4.0/lulesh-init.cc:20-38)
struct Phys {
    static constexpr double e_cut = 1.0e-7, p_cut = 1.0e-7, q_cut = 1.0e-7;
    static constexpr double v_cut = 1.0e-10, u_cut = 1.0e-7;
    static constexpr double hgcoef = 3.0;
    static constexpr double monoq_max_slope = 1.0, monoq_limiter_mult = 2.0;
    static constexpr double qlc_monoq = 0.5, qqc_monoq = 2.0 / 3.0, qqc = 2.0;
    static constexpr double eosvmax = 1.0e+9, eosvmin = 1.0e-9;
    static constexpr double pmin = 0., emin = -1.0e+15;
    static constexpr double dvovmax = 0.1, refdens = 1.0;
    static constexpr double dtfixed = -1.0e-6, dtmax = 1.0e-2;
    static constexpr double deltatimemultlb = 1.1, deltatimemultub = 1.2;
};

// halo-exchange tables (device_local backend)
struct PackEnt {          // one element of one message
    double *dst;          // receiver buffer address of field 0
    int src_dom;
    int src_idx;
    int cnt;              // message length per field (field stride)
    int pad;
};

struct UnpackTgt {        // one receiving index and its messages
    int dom;
    int idx;              // node index, or ghost element index
    int first;            // into the UnpackSrc list
    int count;
};

struct UnpackSrc {
    const double *src;    // receive-buffer address of field 0
    int cnt;              // field stride
    int pad;
};

// device-to-device copies of messages (the device_copy backend's sends),
// passed to the copy kernel by value
constexpr int kMaxCopyMsgs = 64;
struct CopyMsgs {
    double *dst[kMaxCopyMsgs];
    const double *src[kMaxCopyMsgs];
    unsigned count[kMaxCopyMsgs];   // doubles
    int n;
};

}  // namespace lulesh
