// lulesh_engine.cpp -- the stepping engine, the batched ensemble and the
// continuous-batching server.
//
// An Engine steps any set of groups together.  One cycle is
//   k_time_increment  (all groups)
//   k_force_march     (stress + hourglass forces and the ordered node gather,
//                      all subdomains; for one-subdomain groups also the
//                      node update, old node-state buffer -> new one;
//                      LULESH_FORCE_PATH=twopass selects k_elem_force +
//                      k_node_gather instead)
//   [SBN exchange]    multi-subdomain groups
//   [k_node_update]   when not done by k_force_march
//   [PosVel exchange]
//   k_kinematics
//   [MonoQ exchange]
//   k_qeos            (MonoQ region, qstop, EOS x rep, dt constraints)
// and `n` cycles are captured once into a HIP graph and replayed, so the
// host launches one graph per n cycles and never waits inside a cycle: dt,
// time, the stop test and error codes all live in device control blocks.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "lulesh_amd.h"
#include "lulesh_engine.h"
#include "lulesh_kernels.h"
#include "lulesh_plan.h"

#define HIPCHK(call)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            fprintf(stderr, "lulesh_amd: %s:%d: %s: %s\n", __FILE__, __LINE__, #call,    \
                    hipGetErrorString(e_));                                              \
            return LULESH_EHIP;                                                          \
        }                                                                                \
    } while (0)

namespace lulesh {

static const int kSbn[3] = {LULESH_FX, LULESH_FY, LULESH_FZ};
static const int kPosVel[6] = {LULESH_X, LULESH_Y, LULESH_Z, LULESH_XD, LULESH_YD, LULESH_ZD};
static const int kMonoQ[3] = {LULESH_DELV_XI, LULESH_DELV_ETA, LULESH_DELV_ZETA};

// Graph execs may still be queued or running on the stream they were last
// launched on: that stream is drained before an exec is destroyed.
static void drop_cached(Engine::CachedGraph &c)
{
    if (c.exec) {
        (void)hipStreamSynchronize(c.stream);
        (void)hipGraphExecDestroy(c.exec);
    }
    if (c.graph) (void)hipGraphDestroy(c.graph);
    c = Engine::CachedGraph{};
}

static void engine_drop_graph(Engine *e)
{
    for (auto &c : e->graphs) drop_cached(c);
    e->graphs.clear();
}

// at most this many captured shapes are kept per engine
static const size_t kGraphCache = 4;

static int device_cus()
{
    int cus = 256, dev = 0;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    return std::max(cus, 1);
}

static int build_tiles(Engine *e)
{
    std::vector<Tile> et, nt;
    // column tiles of the force march (the block's (i,j,k) structure; an
    // unstructured engine uses the element/node tiles only)
    ForcePlan fp;
    if (!e->umesh) {
        ForcePlanInput pin;
        for (auto *g : e->groups) {
            pin.group_nx.emplace_back();
            for (auto *d : g->doms) pin.group_nx.back().push_back(d->host.nx);
        }
        pin.kchunk = e->kchunk;
        pin.cus = device_cus();
        pin.paired_allowed = paired_runs_wanted() && e->groups.size() == 1 && e->kchunk == 0;
        const int rc = plan_force_columns(pin, fp);
        if (rc) return rc;
    }
    const std::vector<ColTile> &ct = fp.tiles;
    e->paired_runs = fp.paired;
    e->kchunk_used = fp.kchunk_used;
    if (e->d_ctiles) (void)hipFree(e->d_ctiles);
    e->d_ctiles = nullptr;
    e->n_ctiles = (int)ct.size();
    HIPCHK(hipMalloc((void **)&e->d_ctiles, std::max<size_t>(1, ct.size()) * sizeof(ColTile)));
    if (!ct.empty()) HIPCHK(hipMemcpy(e->d_ctiles, ct.data(), ct.size() * sizeof(ColTile), hipMemcpyHostToDevice));
    for (int di = 0, gi = 0; gi < (int)e->groups.size(); ++gi)
        for (auto *d : e->groups[gi]->doms) {
            const HostDomain &h = d->host;
            for (int s = 0; s < h.numElem; s += e->tile_e) {
                const int end = std::min(s + e->tile_e, h.numElem);
                const long long a = h.evprefix64[s / 64], b = h.evprefix64[(end + 63) / 64];
                et.push_back(Tile{di, s, a, (int)(b - a), s == 0 ? kFirstTile : 0});
            }
            for (int s = 0; s < h.numNode; s += e->tile_n) nt.push_back(Tile{di, s, 0, 0, 0});
            ++di;
        }
    // EOS tiles: 256 elements when the engine has enough elements to give
    // every CU two such workgroups, else 128 or 64, so that the repetitions
    // of a small mesh spread over more lanes than it has elements
    long long etot = 0;
    for (auto *g : e->groups)
        for (auto *d : g->doms) etot += d->host.numElem;
    const int cus = device_cus();
    e->qeos_te = etot >= 2LL * 256 * cus ? 256 : etot >= 2LL * 128 * cus ? 128 : 64;
    if (const char *qe = getenv("LULESH_QEOS_TILE")) {
        const int v = atoi(qe);
        if (v == 64 || v == 128 || v == 256) e->qeos_te = v;
    }
    std::vector<Tile> qt;
    for (int di = 0, gi = 0; gi < (int)e->groups.size(); ++gi)
        for (auto *d : e->groups[gi]->doms) {
            const HostDomain &h = d->host;
            for (int s = 0; s < h.numElem; s += e->qeos_te) {
                const int end = std::min(s + e->qeos_te, h.numElem);
                const long long a = h.evprefix64[s / 64], b = h.evprefix64[(end + 63) / 64];
                qt.push_back(Tile{di, s, a, (int)(b - a), 0});
            }
            ++di;
        }
    // EOS launch order: the kernels map workgroup b to tile chunk b % 8 (one
    // chunk of consecutive tiles per XCD), slot b / 8; inside each chunk the
    // tiles with the most EOS evaluations go first, so the long workgroups
    // do not start last and run alone at the end of the launch
    {
        const size_t n = qt.size(), q = n / 8, r = n % 8;
        size_t a = 0;
        for (size_t x = 0; x < 8; ++x) {
            const size_t len = x < r ? q + 1 : q;
            std::stable_sort(qt.begin() + a, qt.begin() + a + len,
                             [](const Tile &u, const Tile &v) { return u.ecount > v.ecount; });
            a += len;
        }
    }
    // the direct update: the node tiles, then tiles of each decomposed
    // group's PosVel copies (those nodes leave the node tiles at once)
    std::vector<Tile> pvt;
    if (e->direct_pv)
        for (int di = 0, gi = 0; gi < (int)e->groups.size(); ++gi) {
            lulesh_group *g = e->groups[gi];
            if (g->ndom > 1 && g->direct)
                for (int s = 0; s < g->direct->npv; s += e->tile_n)
                    pvt.push_back(Tile{di, s, 0, std::min(e->tile_n, g->direct->npv - s), 1});
            di += g->ndom;
        }
    // the PosVel tiles (long chains of dependent loads) lead each XCD's
    // chunk, so they start first and overlap the plain node tiles
    std::vector<Tile> ut;
    {
        const size_t n = nt.size() + pvt.size(), q = n / 8, r = n % 8;
        size_t ri = 0, pi = 0;
        for (size_t x = 0; x < 8; ++x) {
            const size_t len = x < r ? q + 1 : q;
            const size_t pend = pvt.size() * (x + 1) / 8;
            size_t k = 0;
            for (; pi < pend && k < len; ++pi, ++k) ut.push_back(pvt[pi]);
            for (; k < len && ri < nt.size(); ++k) ut.push_back(nt[ri++]);
            for (; k < len && pi < pvt.size(); ++k) ut.push_back(pvt[pi++]);
        }
    }
    if (e->d_etiles) (void)hipFree(e->d_etiles);
    if (e->d_ntiles) (void)hipFree(e->d_ntiles);
    if (e->d_qtiles) (void)hipFree(e->d_qtiles);
    if (e->d_utiles) (void)hipFree(e->d_utiles);
    e->d_etiles = e->d_ntiles = e->d_qtiles = e->d_utiles = nullptr;
    e->n_utiles = (int)ut.size();
    HIPCHK(hipMalloc((void **)&e->d_utiles, std::max<size_t>(1, ut.size()) * sizeof(Tile)));
    if (!ut.empty()) HIPCHK(hipMemcpy(e->d_utiles, ut.data(), ut.size() * sizeof(Tile), hipMemcpyHostToDevice));
    e->n_etiles = (int)et.size();
    e->n_ntiles = (int)nt.size();
    e->n_qtiles = (int)qt.size();
    HIPCHK(hipMalloc((void **)&e->d_etiles, std::max<size_t>(1, et.size()) * sizeof(Tile)));
    HIPCHK(hipMalloc((void **)&e->d_ntiles, std::max<size_t>(1, nt.size()) * sizeof(Tile)));
    HIPCHK(hipMalloc((void **)&e->d_qtiles, std::max<size_t>(1, qt.size()) * sizeof(Tile)));
    if (!et.empty()) HIPCHK(hipMemcpy(e->d_etiles, et.data(), et.size() * sizeof(Tile), hipMemcpyHostToDevice));
    if (!nt.empty()) HIPCHK(hipMemcpy(e->d_ntiles, nt.data(), nt.size() * sizeof(Tile), hipMemcpyHostToDevice));
    if (!qt.empty()) HIPCHK(hipMemcpy(e->d_qtiles, qt.data(), qt.size() * sizeof(Tile), hipMemcpyHostToDevice));
    return LULESH_OK;
}

int engine_create(const std::vector<lulesh_group *> &groups, Engine **out, const std::vector<int> *cycles)
{
    if (cycles && cycles->size() != groups.size()) return LULESH_EINVAL;
    Engine *e = new Engine();
    e->groups = groups;
    const char *nti = getenv("LULESH_NO_FUSED_TI");
    e->fused_ti = !(nti && nti[0] == '1');
    const char *ng = getenv("LULESH_NO_GRAPH");
    e->use_graph = !(ng && ng[0] == '1');
    const char *fp = getenv("LULESH_FORCE_PATH");
    e->march = !(fp && !strcmp(fp, "twopass"));
    const char *kc = getenv("LULESH_KCHUNK");
    if (kc) e->kchunk = atoi(kc);
    std::vector<DomDev> doms;
    std::vector<GroupCtl *> gptr;
    std::vector<int> dom0;
    // an engine steps block groups or unstructured groups, not both: the
    // kernels' connectivity is chosen per launch
    e->umesh = !groups.empty() && groups[0]->umesh;
    for (auto *g : groups)
        if (g->umesh != e->umesh) {
            engine_destroy(e);
            return LULESH_EINVAL;
        }
    if (e->umesh) e->march = false;
    for (size_t gi = 0; gi < groups.size(); ++gi) {
        lulesh_group *g = groups[gi];
        if (g->ndom > 1) e->fused = false;
        // node-state buffers as of the group's last counted cycle
        int cycle;
        if (cycles) {
            cycle = (*cycles)[gi];
        } else {
            GroupCtl G;
            DomCtl D0;
            const int rc0 = group_read_ctl(g, &G, &D0, g->stream);
            if (rc0) { engine_destroy(e); return rc0; }
            cycle = D0.cycle;
        }
        dom0.push_back((int)doms.size());
        for (auto *d : g->doms) {
            DomDev D = d->dev;
            D.cur = D.ns[cycle & 1];
            D.old = D.ns[(cycle & 1) ^ 1];
            doms.push_back(D);
        }
        gptr.push_back(g->gctl);
        if (g->ndom > 1) {
            // build the exchange plans now: no allocation may happen while a
            // graph is being captured
            DevPlan *p;
            int rc = group_plan(g, LULESH_MSG_COMM_SBN, 3, kSbn, true, false, &p);
            if (!rc) rc = group_plan(g, LULESH_MSG_SYNC_POS_VEL, 6, kPosVel, false, false, &p);
            if (!rc) rc = group_plan(g, LULESH_MSG_MONOQ, 3, kMonoQ, true, true, &p);
            if (rc) { engine_destroy(e); return rc; }
        }
    }
    e->ndoms = (int)doms.size();
    {
        const char *nd = getenv("LULESH_NO_DIRECT");
        e->direct = !e->fused && !(nd && nd[0] == '1');
        for (auto *g : groups)
            if (g->ndom > 1 && !group_device_local(g)) e->direct = false;
        const char *npv = getenv("LULESH_NO_DIRECT_POSVEL");
        e->direct_pv = e->direct && !(npv && npv[0] == '1');
        if (e->direct) {
            for (size_t gi = 0; gi < groups.size(); ++gi) {
                lulesh_group *g = groups[gi];
                DirectTables *t = nullptr;
                if (g->ndom > 1) {
                    const int rc0 = group_direct(g, &t);
                    if (rc0) { engine_destroy(e); return rc0; }
                }
                for (int r = 0; r < g->ndom; ++r) {
                    DomDev &D = doms[dom0[gi] + r];
                    D.dom_base = dom0[gi];
                    if (!t) continue;
                    lulesh_domain *d = g->doms[r];
                    D.fpx = d->fpart[0];
                    D.fpy = d->fpart[1];
                    D.fpz = d->fpart[2];
                    D.dir_node = t->node + t->node_off[r];
                    D.dir_elem = t->elem + t->elem_off[r];
                    D.dir_sbn = t->sbn;
                    D.dir_ghost = t->ghost;
                }
            }
            for (size_t gi = 0; gi < groups.size() && e->direct_pv; ++gi) {
                lulesh_group *g = groups[gi];
                if (g->ndom < 2) continue;
                DirectTables *t = g->direct.get();
                for (int r = 0; r < g->ndom; ++r) {
                    DomDev &D = doms[dom0[gi] + r];
                    D.dir_pv = t->pv + t->node_off[r];
                    D.dir_pvnode = t->pvnode;
                    D.pv_snap = t->snap + t->snap_off[r];
                    D.pv_nslot = t->nslot[r];
                    D.dir_pvlist = t->pvlist;
                }
            }
        }
    }
    {
        const char *fu = getenv("LULESH_NO_FUSED_UPDATE");
        e->fuse_update = e->fused && e->march && !(fu && fu[0] == '1');
        for (const DomDev &D : doms)
            if (D.ns[0].x == D.ns[1].x) e->fuse_update = false;
        // forces go to memory when an exchange or a separate update reads
        // them, or a member group asked to keep them
        e->store_forces = !e->fused || (e->march && !e->fuse_update);
        for (auto *g : groups)
            if (g->keep_forces) e->store_forces = true;
    }
    if (hipMalloc((void **)&e->d_doms, std::max<size_t>(1, doms.size()) * sizeof(DomDev)) != hipSuccess ||
        hipMalloc((void **)&e->d_groups, std::max<size_t>(1, gptr.size()) * sizeof(GroupCtl *)) != hipSuccess ||
        hipMalloc((void **)&e->d_dom0, std::max<size_t>(1, dom0.size()) * sizeof(int)) != hipSuccess ||
        hipHostMalloc((void **)&e->h_snap, std::max<size_t>(1, gptr.size()) * sizeof(CtlSnap),
                      hipHostMallocCoherent) != hipSuccess) {
        engine_destroy(e);
        return LULESH_ENOMEM;
    }
    if (!doms.empty())
        (void)hipMemcpy(e->d_doms, doms.data(), doms.size() * sizeof(DomDev), hipMemcpyHostToDevice);
    if (!gptr.empty())
        (void)hipMemcpy(e->d_groups, gptr.data(), gptr.size() * sizeof(GroupCtl *), hipMemcpyHostToDevice);
    if (!dom0.empty())
        (void)hipMemcpy(e->d_dom0, dom0.data(), dom0.size() * sizeof(int), hipMemcpyHostToDevice);
    int rc = build_tiles(e);
    if (rc) { engine_destroy(e); return rc; }
    if (e->direct_pv) {
        // the final writers' snapshots start from the current node state
        bool ok = true;
        for (auto *g : groups) ok = ok && hipStreamSynchronize(g->stream) == hipSuccess;
        if (!ok || launch_pv_snapshot(e->d_doms, e->d_ntiles, e->n_ntiles, e->tile_n, nullptr) != hipSuccess ||
            hipStreamSynchronize(nullptr) != hipSuccess) {
            engine_destroy(e);
            return LULESH_EHIP;
        }
    }
    *out = e;
    return LULESH_OK;
}

void engine_destroy(Engine *e)
{
    if (!e) return;
    engine_drop_graph(e);
    if (e->d_doms) (void)hipFree(e->d_doms);
    if (e->d_groups) (void)hipFree(e->d_groups);
    if (e->d_dom0) (void)hipFree(e->d_dom0);
    if (e->d_etiles) (void)hipFree(e->d_etiles);
    if (e->d_qtiles) (void)hipFree(e->d_qtiles);
    if (e->d_ntiles) (void)hipFree(e->d_ntiles);
    if (e->d_ctiles) (void)hipFree(e->d_ctiles);
    if (e->d_utiles) (void)hipFree(e->d_utiles);
    if (e->snap_pending) (void)hipStreamSynchronize(e->snap_stream);
    if (e->h_snap) (void)hipHostFree(e->h_snap);
    delete e;
}

int engine_set_tiles(Engine *e, int tile_e, int tile_n)
{
    if (tile_e > 0) {
        if (tile_e > 256 || tile_e % 64) return LULESH_EINVAL;
        e->tile_e = tile_e;
    }
    if (tile_n > 0) {
        if (tile_n > 256 || tile_n % 64) return LULESH_EINVAL;
        e->tile_n = tile_n;
    }
    engine_drop_graph(e);
    return build_tiles(e);
}

// Optional per-kernel-class timing of engine_enqueue_cycle: an event pair
// around every step on the engine's stream (only outside graph capture).
struct Prof {
    hipEvent_t a = nullptr, b = nullptr;
    double ms[7] = {0, 0, 0, 0, 0, 0, 0};
};

static int prof_begin(Prof *p, hipStream_t s)
{
    if (p) HIPCHK(hipEventRecord(p->a, s));
    return LULESH_OK;
}

static int prof_end(Prof *p, int k, hipStream_t s)
{
    if (!p) return LULESH_OK;
    HIPCHK(hipEventRecord(p->b, s));
    HIPCHK(hipEventSynchronize(p->b));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, p->a, p->b));
    p->ms[k] += t;
    return LULESH_OK;
}

#define STEP(k, body)                                    \
    do {                                                 \
        int rc_ = prof_begin(prof, s);                   \
        if (rc_) return rc_;                             \
        body;                                            \
        rc_ = prof_end(prof, k, s);                      \
        if (rc_) return rc_;                             \
    } while (0)

static int exchange_all(Engine *e, int msg, int nf, const int *f, hipStream_t s)
{
    for (auto *g : e->groups) {
        const int rc = group_exchange(g, msg, nf, f, s, true);
        if (rc) return rc;
    }
    return LULESH_OK;
}

// kernel classes: 0 time_increment, 1 elem_force, 2 node_gather,
// 3 node_update, 4 kinematics, 5 qeos, 6 halo exchanges
static int enqueue_cycle(Engine *e, hipStream_t s, Prof *prof)
{
    const int ng = (int)e->groups.size();
    int rc = LULESH_OK;
    // the force march computes the cycle's time increment itself; the
    // two-pass (unstructured) path has a launch of its own for it
    if (!e->march || !e->fused_ti)
        STEP(0, HIPCHK(launch_time_increment(e->d_groups, ng, e->d_doms, e->d_dom0, s)));
    bool updated;   // node update done inside the force step
    if (e->march) {
        STEP(1, HIPCHK(launch_force_march(e->d_doms, e->d_ctiles, e->n_ctiles, e->fuse_update,
                                          e->store_forces, e->fused_ti, s)));
        updated = e->fuse_update;
    } else {
        STEP(1, HIPCHK(launch_elem_force(e->umesh, e->d_doms, e->d_etiles, e->n_etiles, e->tile_e, s)));
        const bool gather_update = e->fused && !e->store_forces;
        STEP(2, HIPCHK(launch_node_gather(e->umesh, gather_update, e->d_doms, e->d_ntiles, e->n_ntiles, e->tile_n, s)));
        updated = gather_update;
    }
    if (e->direct) {
        // CommSBN summed straight from the neighbours' partial forces
        STEP(3, HIPCHK(launch_node_update_direct(e->d_doms, e->d_utiles, e->n_utiles, e->tile_n, s)));
    } else {
        if (!e->fused) {
            STEP(6, rc = exchange_all(e, LULESH_MSG_COMM_SBN, 3, kSbn, s));
            if (rc) return rc;
        }
        if (!updated)
            STEP(3, HIPCHK(launch_node_update(e->umesh, e->d_doms, e->d_ntiles, e->n_ntiles, e->tile_n, s)));
    }
    if (!e->fused && !e->direct_pv) {   // direct_pv: the node update wrote every copy
        STEP(6, rc = exchange_all(e, LULESH_MSG_SYNC_POS_VEL, 6, kPosVel, s));
        if (rc) return rc;
    }
    STEP(4, HIPCHK(launch_kinematics(e->umesh, e->d_doms, e->d_etiles, e->n_etiles, e->tile_e, s)));
    if (!e->fused && !e->direct) {   // direct: the kinematics kernel filled the ghosts
        STEP(6, rc = exchange_all(e, LULESH_MSG_MONOQ, 3, kMonoQ, s));
        if (rc) return rc;
    }
    STEP(5, HIPCHK(launch_qeos(e->umesh, e->d_doms, e->d_qtiles, e->n_qtiles, e->qeos_te, 256, s)));
    return LULESH_OK;
}

int engine_enqueue_cycle(Engine *e, hipStream_t s) { return enqueue_cycle(e, s, nullptr); }

int engine_set_kchunk(Engine *e, int kchunk)
{
    if (kchunk < 0) return LULESH_EINVAL;
    e->kchunk = kchunk;
    engine_drop_graph(e);
    return build_tiles(e);
}

int engine_profile(Engine *e, int ncycles, hipStream_t s, double ms[7])
{
    Prof p;
    HIPCHK(hipEventCreate(&p.a));
    HIPCHK(hipEventCreate(&p.b));
    int rc = LULESH_OK;
    for (int c = 0; c < ncycles && rc == LULESH_OK; ++c) rc = enqueue_cycle(e, s, &p);
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
    for (int k = 0; k < 7; ++k) ms[k] = p.ms[k];
    return rc;
}

int engine_enqueue(Engine *e, int ncycles, hipStream_t s)
{
    if (ncycles <= 0) return LULESH_OK;
    if (!e->use_graph) {
        for (int c = 0; c < ncycles; ++c) {
            const int rc = engine_enqueue_cycle(e, s);
            if (rc) return rc;
        }
        e->snap_stream = s;
        e->snap_pending = true;
        HIPCHK(launch_ctl_snapshot(e->d_groups, (int)e->groups.size(), e->h_snap, s));
        return LULESH_OK;
    }
    for (size_t i = 0; i < e->graphs.size(); ++i) {
        if (e->graphs[i].ncycles == ncycles && e->graphs[i].stream == s) {
            Engine::CachedGraph c = e->graphs[i];
            e->graphs.erase(e->graphs.begin() + (long)i);
            e->graphs.push_back(c);
            e->snap_stream = s;
            e->snap_pending = true;
            HIPCHK(hipGraphLaunch(c.exec, s));
            return LULESH_OK;
        }
    }
    if (e->graphs.size() >= kGraphCache) {   // least recently used
        drop_cached(e->graphs.front());
        e->graphs.erase(e->graphs.begin());
    }
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = LULESH_OK;
    for (int c = 0; c < ncycles && rc == LULESH_OK; ++c) rc = engine_enqueue_cycle(e, s);
    if (rc == LULESH_OK && launch_ctl_snapshot(e->d_groups, (int)e->groups.size(), e->h_snap, s) != hipSuccess)
        rc = LULESH_EHIP;
    Engine::CachedGraph c;
    c.ncycles = ncycles;
    c.stream = s;
    const hipError_t ce = hipStreamEndCapture(s, &c.graph);
    if (rc) {
        if (c.graph) (void)hipGraphDestroy(c.graph);
        return rc;
    }
    if (ce != hipSuccess) {
        if (c.graph) (void)hipGraphDestroy(c.graph);
        HIPCHK(ce);
    }
    const hipError_t ie = hipGraphInstantiate(&c.exec, c.graph, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        (void)hipGraphDestroy(c.graph);
        HIPCHK(ie);
    }
    e->graphs.push_back(c);
    e->snap_stream = s;
    e->snap_pending = true;
    HIPCHK(hipGraphLaunch(c.exec, s));
    return LULESH_OK;
}

}  // namespace lulesh

using namespace lulesh;

// ===========================================================================
// batched ensemble
// ===========================================================================
struct lulesh_batch {
    std::vector<lulesh_group *> groups;
    Engine *engine = nullptr;           // steps engine_groups
    std::vector<lulesh_group *> engine_groups;
    long long engine_zones = 0;         // elements of engine_groups
    int tile_e = 0, tile_n = 0;         // the caller's tile shapes (0: default)
    hipStream_t stream = nullptr;
};

static long long zones_of(const lulesh_group *g)
{
    long long z = 0;
    for (auto *d : g->doms) z += d->host.numElem;
    return z;
}

extern "C" {

static int batch_create(int nprob, const int *nx, const int *nr, const int *balance, const int *cost,
                        lulesh_batch **out)
{
    if (nprob < 1 || !nx || !nr || !balance || !cost || !out) return LULESH_EINVAL;
    lulesh_batch *b = new lulesh_batch();
    for (int i = 0; i < nprob; ++i) {
        lulesh_group *g = nullptr;
        const int rc = group_create(nx[i], 1, nr[i], balance[i], cost[i], &g);
        if (rc) {
            group_free(g);
            lulesh_batch_destroy(b);
            return rc;
        }
        b->groups.push_back(g);
    }
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        lulesh_batch_destroy(b);
        return LULESH_EHIP;
    }
    const int rc = engine_create(b->groups, &b->engine);
    if (rc) { lulesh_batch_destroy(b); return rc; }
    b->engine_groups = b->groups;
    for (auto *g : b->groups) b->engine_zones += zones_of(g);
    *out = b;
    return LULESH_OK;
}

int lulesh_batch_create(int nprob, const int *nx, const int *nr, const int *balance, const int *cost,
                        lulesh_batch **out)
{
    try {
        return batch_create(nprob, nx, nr, balance, cost, out);
    } catch (const std::bad_alloc &) {
        return LULESH_ENOMEM;
    }
}

int lulesh_batch_destroy(lulesh_batch *b)
{
    if (!b) return LULESH_OK;
    if (b->engine) engine_destroy(b->engine);
    for (auto *g : b->groups) group_free(g);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return LULESH_OK;
}

int lulesh_batch_set_max_cycles(lulesh_batch *b, int m)
{
    if (!b) return LULESH_EINVAL;
    HIPCHK(hipStreamSynchronize(b->stream));
    for (auto *g : b->groups) {
        g->max_cycles = m;
        const int rc = group_set_param(g);
        if (rc) return rc;
    }
    return LULESH_OK;
}

int lulesh_batch_set_tiles(lulesh_batch *b, int tile_e, int tile_n)
{
    if (!b) return LULESH_EINVAL;
    const int rc = engine_set_tiles(b->engine, tile_e, tile_n);
    if (rc == LULESH_OK) {
        if (tile_e > 0) b->tile_e = tile_e;
        if (tile_n > 0) b->tile_n = tile_n;
    }
    return rc;
}

int lulesh_batch_group(lulesh_batch *b, int i, lulesh_group **out)
{
    if (!b || i < 0 || i >= (int)b->groups.size() || !out) return LULESH_EINVAL;
    *out = b->groups[i];
    return LULESH_OK;
}

// a member's loop state as the last poll saw it
struct MemberState {
    int done = 0, status = 0, cycle = 0;
};

// What a lulesh_batch_run call knows of the members: their states in
// b->groups order, and where each member sits in that order.
struct BatchView {
    std::vector<MemberState> st;
    std::unordered_map<const lulesh_group *, size_t> idx;
    const MemberState &of(const lulesh_group *g) const { return st[idx.at(g)]; }
};

// The members still running, and the first error any member raised.
// With `snap`, the engine's members are read from the snapshot its last
// chunk wrote (the batch's stream has completed it) and the others keep the
// state last seen: during lulesh_batch_run nothing steps a member the
// engine no longer holds.  Without it (the start of a run: host calls may
// have changed any member since) every member's control block is read.
static int batch_poll(lulesh_batch *b, BatchView &v, bool snap, std::vector<lulesh_group *> &running,
                      int *first_err)
{
    running.clear();
    *first_err = 0;
    HIPCHK(hipStreamSynchronize(b->stream));
    if (snap) {
        const std::vector<lulesh_group *> &eg = b->engine->groups;
        for (size_t k = 0; k < eg.size(); ++k) {
            const CtlSnap &c = b->engine->h_snap[k];
            v.st[v.idx.at(eg[k])] = MemberState{c.done, c.status, c.cycle};
        }
    } else {
        for (size_t i = 0; i < b->groups.size(); ++i) {
            GroupCtl G;
            DomCtl D0;
            const int rc = group_read_ctl(b->groups[i], &G, &D0, b->stream);
            if (rc) return rc;
            v.st[i] = MemberState{G.done, G.status, D0.cycle};
        }
    }
    for (size_t i = 0; i < b->groups.size(); ++i) {
        const MemberState &m = v.st[i];
        if (m.status && !*first_err) *first_err = m.status;
        if (!m.done && !m.status) running.push_back(b->groups[i]);
    }
    return LULESH_OK;
}

// an engine over `members` (z elements) in place of the current one; the
// members' cycle counts are the poll's
static int batch_rebuild(lulesh_batch *b, const std::vector<lulesh_group *> &members, long long z,
                         const BatchView &v)
{
    std::vector<int> cycles;
    for (auto *g : members) cycles.push_back(v.of(g).cycle);
    Engine *e = nullptr;
    int rc = engine_create(members, &e, &cycles);
    if (rc == LULESH_OK && (b->tile_e || b->tile_n)) rc = engine_set_tiles(e, b->tile_e, b->tile_n);
    if (rc) {
        if (e) engine_destroy(e);
        return rc;
    }
    engine_destroy(b->engine);
    b->engine = e;
    b->engine_groups = members;
    b->engine_zones = z;
    return LULESH_OK;
}

// Keeps the chip full as members finish: once the running members hold
// under 7/8 of the elements the engine steps, the engine is rebuilt over
// them alone, so its column tiles are re-chunked for the smaller set (more,
// shorter tiles) and finished members' tiles are no longer launched
// (measured on the 64-problem ensemble: 5.22e9 zone-cycles/s at 5/8, 5.33
// at 6/8, 5.41 at 7/8, 5.29 rebuilding whenever a member finishes).
// Results do not depend on the tiling.
static int batch_shrink(lulesh_batch *b, const std::vector<lulesh_group *> &running, const BatchView &v)
{
    long long z = 0;
    for (auto *g : running) z += zones_of(g);
    if (running.empty() || z * 8 >= b->engine_zones * 7) return LULESH_OK;
    return batch_rebuild(b, running, z, v);
}

// a member re-armed after the engine dropped it (a raised cycle cap or stop
// time): the engine covers every running member again
static int batch_cover(lulesh_batch *b, const std::vector<lulesh_group *> &running, const BatchView &v)
{
    for (auto *g : running)
        if (std::find(b->engine_groups.begin(), b->engine_groups.end(), g) == b->engine_groups.end()) {
            long long z = 0;
            for (auto *h : running) z += zones_of(h);
            return batch_rebuild(b, running, z, v);
        }
    return LULESH_OK;
}

static int batch_run(lulesh_batch *b, int cycles_per_sync)
{
    if (cycles_per_sync <= 0) cycles_per_sync = 32;
    BatchView v;
    v.st.resize(b->groups.size());
    for (size_t i = 0; i < b->groups.size(); ++i) v.idx[b->groups[i]] = i;
    std::vector<lulesh_group *> running;
    int err0;
    int rc0 = batch_poll(b, v, false, running, &err0);
    if (rc0 == LULESH_OK) rc0 = batch_cover(b, running, v);
    if (rc0) return rc0;
    if (running.empty()) return err0;
    for (;;) {
        int rc = engine_enqueue(b->engine, cycles_per_sync, b->stream);
        if (rc) return rc;
        int err;
        rc = batch_poll(b, v, true, running, &err);
        if (rc) return rc;
        if (running.empty()) return err;
        rc = batch_shrink(b, running, v);
        if (rc) return rc;
    }
}

int lulesh_batch_run(lulesh_batch *b, int cycles_per_sync, lulesh_stream_t stream)
{
    if (!b) return LULESH_EINVAL;
    (void)stream;
    try {
        return batch_run(b, cycles_per_sync);
    } catch (const std::bad_alloc &) {
        return LULESH_ENOMEM;
    }
}

int lulesh_batch_results(lulesh_batch *b, int *cycles, int *status)
{
    if (!b) return LULESH_EINVAL;
    HIPCHK(hipStreamSynchronize(b->stream));
    for (size_t i = 0; i < b->groups.size(); ++i) {
        GroupCtl G;
        DomCtl D0;
        const int rc = group_read_ctl(b->groups[i], &G, &D0, b->stream);
        if (rc) return rc;
        if (cycles) cycles[i] = D0.cycle;
        if (status) status[i] = G.status;
    }
    return LULESH_OK;
}

// Enqueues `ncycles` cycles of the whole ensemble without waiting (used by
// the tile autotuner and the benchmark's timed region).
int lulesh_batch_enqueue(lulesh_batch *b, int ncycles)
{
    if (!b) return LULESH_EINVAL;
    return engine_enqueue(b->engine, ncycles, b->stream);
}

int lulesh_batch_profile(lulesh_batch *b, int ncycles, double ms[7])
{
    if (!b || ncycles < 0 || !ms) return LULESH_EINVAL;
    return engine_profile(b->engine, ncycles, b->stream, ms);
}

// Tile-shape autotuner: times `probe_cycles` cycles of the ensemble (one
// captured graph each) for every candidate (element tile, node tile, column
// chunk) and installs the fastest.  The probe cycles are real cycles of the
// problems; results do not depend on the tile shapes.
int lulesh_batch_autotune(lulesh_batch *b, int probe_cycles, int *best_tile_e, int *best_tile_n,
                          int *best_kchunk, double *best_ms)
{
    if (!b || probe_cycles < 1) return LULESH_EINVAL;
    const int te[] = {64, 128, 256};
    const int tn[] = {128, 256};
    const int kc[] = {0, 4, 8};
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    double best = 1e30;
    int be = 256, bn = 256, bk = 0;
    for (int a : te)
        for (int n : tn)
            for (int k : kc) {
                int rc = engine_set_tiles(b->engine, a, n);
                if (!rc) rc = engine_set_kchunk(b->engine, k);
                if (!rc) rc = engine_enqueue(b->engine, probe_cycles, b->stream);   // capture + warm
                if (rc) return rc;
                HIPCHK(hipEventRecord(e0, b->stream));
                rc = engine_enqueue(b->engine, probe_cycles, b->stream);
                if (rc) return rc;
                HIPCHK(hipEventRecord(e1, b->stream));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) { best = ms; be = a; bn = n; bk = k; }
            }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    int rc = engine_set_tiles(b->engine, be, bn);
    if (!rc) rc = engine_set_kchunk(b->engine, bk);
    b->tile_e = be;
    b->tile_n = bn;
    if (best_tile_e) *best_tile_e = be;
    if (best_tile_n) *best_tile_n = bn;
    if (best_kchunk) *best_kchunk = bk ? bk : b->engine->kchunk_used;
    if (best_ms) *best_ms = best / probe_cycles;
    return rc;
}

int lulesh_batch_set_kchunk(lulesh_batch *b, int kchunk)
{
    if (!b) return LULESH_EINVAL;
    return engine_set_kchunk(b->engine, kchunk);
}

int lulesh_batch_synchronize(lulesh_batch *b)
{
    if (!b) return LULESH_EINVAL;
    HIPCHK(hipStreamSynchronize(b->stream));
    return LULESH_OK;
}

}  // extern "C"

// ===========================================================================
// continuous-batching server
// ===========================================================================
namespace {

struct Request {
    int id;
    int nx, nr, balance, cost;
    double arrival;
    size_t bytes;
    int cycle = 0;   // completed cycles as last seen (a new group has run none)
};

struct Result {
    int status = 1;   // 1 = not finished
    int cycles = 0;
    double energy = 0, tad = 0;
    double t_admit = -1, t_done = -1;
};

// device footprint of one admitted single-domain problem: its field arena
// plus control block and tile-table slack
size_t request_bytes(int nx, int nr, int cost)
{
    const long long ne = (long long)nx * nx * nx;
    return domain_device_bytes(nx, corner_path_enabled(), ne * max_elem_rep(nr, cost)) + 4096;
}

}  // namespace

struct lulesh_server {
    size_t budget = 0, in_use = 0;
    int max_slots = 0;
    std::deque<Request> queue;          // arrival order
    std::vector<std::pair<Request, lulesh_group *>> active;
    std::vector<Result> results;
    Engine *engine = nullptr;
    bool dirty = true;
    hipStream_t stream = nullptr;
    std::chrono::steady_clock::time_point t0;
    int done = 0;
    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

extern "C" {

int lulesh_server_create(size_t budget, int max_slots, lulesh_server **out)
{
    if (!out || max_slots < 1) return LULESH_EINVAL;
    lulesh_server *s = new lulesh_server();
    if (budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) { delete s; return LULESH_EHIP; }
        budget = fr - fr / 16;   // leave headroom for the runtime
    }
    s->budget = budget;
    s->max_slots = max_slots;
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        return LULESH_EHIP;
    }
    s->t0 = std::chrono::steady_clock::now();
    *out = s;
    return LULESH_OK;
}

int lulesh_server_destroy(lulesh_server *s)
{
    if (!s) return LULESH_OK;
    if (s->engine) engine_destroy(s->engine);
    for (auto &a : s->active) group_free(a.second);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return LULESH_OK;
}

double lulesh_server_now(lulesh_server *s) { return s ? s->now() : 0.0; }

static int server_submit(lulesh_server *s, int nx, int nr, int balance, int cost, double arrival)
{
    if (!s || !problem_ok(nx, 1, nr, balance, cost)) return LULESH_EINVAL;
    Request r;
    r.id = (int)s->results.size();
    r.nx = nx; r.nr = nr; r.balance = balance; r.cost = cost;
    r.arrival = arrival;
    r.bytes = request_bytes(nx, nr, cost);
    if (r.bytes > s->budget) return LULESH_ENOMEM;
    s->queue.push_back(r);
    s->results.push_back(Result());
    return r.id;
}

int lulesh_server_submit(lulesh_server *s, int nx, int nr, int balance, int cost, double arrival)
{
    try {
        return server_submit(s, nx, nr, balance, cost, arrival);
    } catch (const std::bad_alloc &) {
        return LULESH_ENOMEM;
    }
}

// admit arrived requests in arrival order while a slot and HBM are free
static int server_admit(lulesh_server *s)
{
    const double t = s->now();
    while (!s->queue.empty() && (int)s->active.size() < s->max_slots) {
        const Request &r = s->queue.front();
        if (r.arrival > t) break;
        if (s->in_use + r.bytes > s->budget) break;
        lulesh_group *g = nullptr;
        const int rc = group_create(r.nx, 1, r.nr, r.balance, r.cost, &g);
        if (rc) {
            group_free(g);
            return rc;
        }
        s->results[r.id].t_admit = s->now();
        s->in_use += r.bytes;
        s->active.push_back({r, g});
        s->queue.pop_front();
        s->dirty = true;
    }
    return LULESH_OK;
}

static int server_poll(lulesh_server *s, int cycles, lulesh_stream_t stream)
{
    if (!s) return LULESH_EINVAL;
    (void)stream;
    if (cycles <= 0) cycles = 16;
    int rc;
    if (s->active.empty()) {
        // nothing on the device to overlap with: admit now
        rc = server_admit(s);
        if (rc) return rc;
        if (s->active.empty()) {
            if (!s->queue.empty()) {
                const double wait = s->queue.front().arrival - s->now();
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.05)));
            }
            return 0;
        }
    }
    if (s->dirty) {
        // the members' completed cycles are known: from the last chunk's
        // snapshot, or 0 for a group admitted since
        std::vector<lulesh_group *> gs;
        std::vector<int> cyc;
        for (auto &a : s->active) {
            gs.push_back(a.second);
            cyc.push_back(a.first.cycle);
        }
        if (s->engine) engine_destroy(s->engine);
        s->engine = nullptr;
        rc = engine_create(gs, &s->engine, &cyc);
        if (rc) return rc;
        s->dirty = false;
    }
    rc = engine_enqueue(s->engine, cycles, s->stream);
    if (rc) return rc;
    // the engine's members are the first entries of the active list, in
    // order (every change of the list rebuilds the engine before the next
    // chunk); requests admitted while the chunk runs (their set-up is host
    // work and kernels on their own streams) join at the next chunk
    const size_t nmem = s->engine->groups.size();
    // an admission error is reported once this chunk's results are taken:
    // the next rebuild needs every member's cycle count from its snapshot
    const int admit_rc = server_admit(s);
    HIPCHK(hipStreamSynchronize(s->stream));
    // the chunk's snapshot: one entry per engine member
    const CtlSnap *snap = s->engine->h_snap;
    int completed = 0;
    std::vector<std::pair<Request, lulesh_group *>> keep;
    keep.reserve(s->active.size());
    for (size_t i = 0; i < s->active.size(); ++i) {
        auto &a = s->active[i];
        if (i >= nmem) {
            keep.push_back(a);
            continue;
        }
        const CtlSnap &cs = snap[i];
        a.first.cycle = cs.cycle;
        if (cs.done || cs.status) {
            Result &res = s->results[a.first.id];
            res.status = cs.status;
            res.cycles = cs.cycle;
            double v[4] = {0, 0, 0, 0};
            lulesh_group_verify(a.second, 0, v);
            res.energy = v[0];
            res.tad = v[2];
            res.t_done = s->now();
            s->in_use -= a.first.bytes;
            group_free(a.second);
            s->dirty = true;
            ++completed;
            ++s->done;
        } else {
            keep.push_back(a);
        }
    }
    s->active.swap(keep);
    return admit_rc ? admit_rc : completed;
}

int lulesh_server_poll(lulesh_server *s, int cycles, lulesh_stream_t stream)
{
    try {
        return server_poll(s, cycles, stream);
    } catch (const std::bad_alloc &) {
        return LULESH_ENOMEM;
    }
}

int lulesh_server_result(lulesh_server *s, int id, int *status, int *cycles, double *energy, double *tad)
{
    if (!s || id < 0 || id >= (int)s->results.size()) return LULESH_EINVAL;
    const Result &r = s->results[id];
    if (status) *status = r.status;
    if (cycles) *cycles = r.cycles;
    if (energy) *energy = r.energy;
    if (tad) *tad = r.tad;
    return LULESH_OK;
}

int lulesh_server_times(lulesh_server *s, int id, double *t_admit, double *t_done)
{
    if (!s || id < 0 || id >= (int)s->results.size()) return LULESH_EINVAL;
    if (t_admit) *t_admit = s->results[id].t_admit;
    if (t_done) *t_done = s->results[id].t_done;
    return LULESH_OK;
}

int lulesh_server_stats(lulesh_server *s, int *queued, int *active, int *done, size_t *hbm)
{
    if (!s) return LULESH_EINVAL;
    if (queued) *queued = (int)s->queue.size();
    if (active) *active = (int)s->active.size();
    if (done) *done = s->done;
    if (hbm) *hbm = s->in_use;
    return LULESH_OK;
}

}  // extern "C"
