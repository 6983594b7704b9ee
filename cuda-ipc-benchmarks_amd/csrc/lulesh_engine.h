// lulesh_engine.h -- host-side objects behind the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "lulesh_amd.h"
#include "lulesh_comm.h"
#include "lulesh_dev.h"
#include "lulesh_setup.h"

namespace lulesh {
struct Engine;

// a halo-exchange step of one group, resident on the device
// device tables of the direct halo steps of one group (device_local
// backend): see DomDev::dir_*
struct DirectTables {
    IdxSpan *node = nullptr;       // sum of numNode over the group's subdomains
    IdxSpan *elem = nullptr;       // sum of numElem
    DirRef *sbn = nullptr, *ghost = nullptr;
    // CommSyncPosVel: per node (sum of numNode) an index into pvnode or
    // -1; the snapshots of the final-writer nodes, per rank at snap_off
    int *pv = nullptr;
    PvNode *pvnode = nullptr;
    double *snap = nullptr;
    std::vector<size_t> snap_off;   // per rank, into snap ([2][6][nslot] each)
    std::vector<int> nslot;         // per rank
    DirRef *pvlist = nullptr;       // (rank, node) of every copy with pv >= 0
    int npv = 0;
    std::vector<size_t> node_off, elem_off;   // per rank
    ~DirectTables();
};

struct DevPlan {
    int msgType = 0, nfields = 0, mode = 0;
    PackEnt *pack = nullptr;
    UnpackTgt *tgt = nullptr;
    UnpackSrc *src = nullptr;
    double **fptr = nullptr;
    int npack = 0, ntgt = 0;
    std::vector<int> pack_start, tgt_start;
    // the step's messages (host side), for the backend's send / staging /
    // corner / MonoQ hooks
    std::vector<MsgRef> sends, recvs;
    std::vector<int> send_start, recv_start;
    std::vector<int> fields;
    std::vector<double *> host_fptr;   // [rank * nfields + f]
    std::vector<double *> send_base;   // per rank: where its local packs are staged
    bool corner_hook = false, monoq_hook = false;
    bool owned = true;                 // pack/tgt/src/fptr allocated for this plan
    ~DevPlan();
};
}  // namespace lulesh

struct lulesh_domain {
    lulesh_group *g = nullptr;
    int rank = 0;
    lulesh::HostDomain host;
    lulesh::DomDev dev{};
    void *arena = nullptr;
    size_t arena_bytes = 0;
    double *recv = nullptr;
    double *send = nullptr;     // send buffer (comBufSize), for backends that pack locally
    double *fpart[3] = {nullptr, nullptr, nullptr};   // own partial forces (groups with halos)
};

struct lulesh_group {
    int nx = 0, tp = 1, ndom = 1;
    bool umesh = false;         // one unstructured domain (nx = 0)
    std::vector<lulesh_domain *> doms;
    void *ctl_mem = nullptr;
    lulesh::GroupCtl *gctl = nullptr;
    lulesh::DomCtl *dctl = nullptr;
    double stoptime = 1.0e-2, qstop = 1.0e+12;
    int max_cycles = -1;
    bool keep_forces = false;   // store fx..fz even when no exchange needs them
    // lifecycle phases (max over subdomains), milliseconds
    double phase_setup_ms = 0.0, phase_free_ms = 0.0;
    bool comm_down = false;     // receive buffers released
    const lulesh_comm_backend_ops *backend = nullptr;
    void *backend_ctx = nullptr;
    std::map<std::string, std::unique_ptr<lulesh::DevPlan>> plans;
    std::unique_ptr<lulesh::DirectTables> direct;   // built on first use
    lulesh::Engine *engine = nullptr;   // standalone stepping
    lulesh::CopyMsgs dc_pending{};      // device_copy backend: sends not yet copied
    // the node state changed behind the engine's back (PosVel snapshots
    // stale) where it could not be freed at once: rebuilt at the next step
    bool engine_stale = false;
    hipStream_t stream = nullptr;
};

namespace lulesh {

// Steps a set of groups together: one tile table per kernel class covers
// every subdomain of every member group, and a cycle of the whole set is
// captured once into a HIP graph and replayed.
struct Engine {
    std::vector<lulesh_group *> groups;
    bool fused = true;          // every member has one subdomain
    DomDev *d_doms = nullptr;
    int ndoms = 0;
    GroupCtl **d_groups = nullptr;
    int *d_dom0 = nullptr;      // index in d_doms of each group's first subdomain
    Tile *d_etiles = nullptr, *d_ntiles = nullptr, *d_qtiles = nullptr;
    int n_etiles = 0, n_ntiles = 0, n_qtiles = 0;
    Tile *d_utiles = nullptr;   // node tiles + PosVel copy tiles of the direct update
    int n_utiles = 0;
    int qeos_te = 256;          // elements per EOS tile (256 threads per workgroup)
    int tile_e = 256, tile_n = 256;
    // fused force + gather (z-marching column tiles) or the two-pass
    // corner-array path
    bool march = true;
    // node planes per column tile (0 = auto): a column's planes are split
    // into max(1, planes / kchunk) runs of kchunk or kchunk + 1 planes
    int kchunk = 0;
    int kchunk_used = 8;
    bool paired_runs = false;   // force-march runs balanced for the CU pairs (one round)
    ColTile *d_ctiles = nullptr;
    int n_ctiles = 0;
    // node update done by the force march as each node plane completes
    // (every member double-buffers its node state, no halo exchange sits
    // between force and update); LULESH_NO_FUSED_UPDATE=1 keeps them apart
    bool fuse_update = true;
    // the force march computes each cycle's time increment in every
    // workgroup (no time-increment launch); LULESH_NO_FUSED_TI=1 launches
    // k_time_increment first instead
    bool fused_ti = true;
    bool store_forces = true;   // fx..fz written (exchange or keep_forces)
    // halo groups on the device_local backend: CommSBN summed inside the
    // node update, CommSyncPosVel written by each shared node's final
    // writer, CommMonoQ written by the kinematics kernel, instead of
    // pack/unpack launches (LULESH_NO_DIRECT=1 keeps the exchanges,
    // LULESH_NO_DIRECT_POSVEL=1 keeps the PosVel one)
    bool direct = false;
    bool direct_pv = false;
    bool umesh = false;         // members are unstructured meshes
    // captured graphs of n cycles, one per (n, stream), most recent last;
    // a caller that alternates a chunk and its remainder replays both
    // without capturing inside its loop
    struct CachedGraph {
        int ncycles = 0;
        hipStream_t stream = nullptr;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<CachedGraph> graphs;
    bool use_graph = true;
    // pinned host memory, one entry per member group (engine order): each
    // engine_enqueue ends with a kernel that writes the members' loop
    // state here, valid once that enqueue's stream has completed
    CtlSnap *h_snap = nullptr;
    hipStream_t snap_stream = nullptr;   // of the last enqueue (drained before h_snap is freed)
    bool snap_pending = false;
};

// cycles, when given: each member's completed cycle count (engine order),
// as a snapshot of a finished stream shows it, instead of reading every
// member's control block
int engine_create(const std::vector<lulesh_group *> &groups, Engine **out,
                  const std::vector<int> *cycles = nullptr);
void engine_destroy(Engine *e);
int engine_set_tiles(Engine *e, int tile_e, int tile_n);
int engine_set_kchunk(Engine *e, int kchunk);
int engine_enqueue(Engine *e, int ncycles, hipStream_t s);
int engine_enqueue_cycle(Engine *e, hipStream_t s);
int engine_profile(Engine *e, int ncycles, hipStream_t s, double ms[7]);

int group_create(int nx, int tp, int nr, int balance, int cost, lulesh_group **out);
int group_create_mesh(const lulesh_mesh *m, int nr, int balance, int cost, lulesh_group **out);
// tables of the direct halo steps (device_local backend only)
int group_direct(lulesh_group *g, DirectTables **out);
bool group_device_local(const lulesh_group *g);
void group_free(lulesh_group *g);
void group_state_changed(lulesh_group *g);
int group_reset_ctl(lulesh_group *g);
int group_set_param(lulesh_group *g);
int group_read_ctl(lulesh_group *g, GroupCtl *G, DomCtl *D0, hipStream_t s);
// gated: skip when the group has stopped (the exchanges inside the time loop)
int group_exchange(lulesh_group *g, int msgType, int nf, const int *fields, hipStream_t s,
                   bool gated = false);
int group_plan(lulesh_group *g, int msgType, int nf, const int *fields, bool doFlag, bool planeOnly,
               DevPlan **out);
double *field_dev_ptr(lulesh_domain *d, int f, size_t *len);
bool corner_path_enabled();

}  // namespace lulesh
