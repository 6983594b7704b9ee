// lulesh_api.cpp -- groups, subdomains, the device_local comm backend, the
// stepping engine (HIP graphs) and the C ABI of include/lulesh_amd.h.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <limits.h>
#include <stddef.h>

#include <algorithm>
#include <chrono>
#include <new>
#include <map>
#include <unordered_map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "lulesh_amd.h"
#include "lulesh_comm.h"
#include "lulesh_dev.h"
#include "lulesh_engine.h"
#include "lulesh_kernels.h"
#include "lulesh_mesh.h"
#include "lulesh_setup.h"

using namespace lulesh;

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

// ---------------------------------------------------------------------------
// subdomain storage
// ---------------------------------------------------------------------------
static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// The per-element corner-force array (48 doubles per element) is needed
// only by the two-pass force path, kept for cross-checking the fused one.
bool corner_path_enabled()
{
    const char *fp = getenv("LULESH_FORCE_PATH");
    return fp && !strcmp(fp, "twopass");
}

int domain_alloc(lulesh_domain *d)
{
    const HostDomain &h = d->host;
    const size_t nn = h.numNode, ne = h.numElem, all = h.allElem;
    struct Slot { double **p; size_t n; };
    DomDev &D = d->dev;
    // a one-subdomain group double-buffers its node state (no halo
    // exchange touches it between the node update and its readers)
    const bool two = d->g->ndom == 1;
    NodeState &a = D.ns[0], &b = D.ns[1];
    const size_t nn2 = two ? nn : 0;
    std::vector<Slot> slots = {
        {&a.x, nn}, {&a.y, nn}, {&a.z, nn}, {&a.xd, nn}, {&a.yd, nn}, {&a.zd, nn},
        {&b.x, nn2}, {&b.y, nn2}, {&b.z, nn2}, {&b.xd, nn2}, {&b.yd, nn2}, {&b.zd, nn2},
        {&D.fx, nn}, {&D.fy, nn}, {&D.fz, nn}, {&D.nodalMass, nn},
        {&d->fpart[0], two ? 0 : nn}, {&d->fpart[1], two ? 0 : nn}, {&d->fpart[2], two ? 0 : nn},
        {&D.e, ne}, {&D.p, ne}, {&D.q, ne}, {&D.v, ne}, {&D.volo, ne}, {&D.delv, ne},
        {&D.vdov, ne}, {&D.arealg, ne}, {&D.ss, ne}, {&D.elemMass, ne}, {&D.vnew, ne},
        {&D.delv_xi, all}, {&D.delv_eta, all}, {&D.delv_zeta, all},
        {&D.delx_xi, ne}, {&D.delx_eta, ne}, {&D.delx_zeta, ne}, {&D.hgcoef, ne},
        {&D.corner, corner_path_enabled() || h.umesh ? (size_t)kCornerComps * ne : 0}};
    // integer tables: elemRep, then the unstructured connectivity
    struct ISlot { const void **p; size_t bytes; };
    const size_t um = h.umesh ? 1 : 0;
    std::vector<ISlot> islots = {
        {(const void **)&D.elemRep, ne * sizeof(int)},
        {(const void **)&D.u_nodelist, um * 8 * ne * sizeof(int)},
        {(const void **)&D.u_face, um * 8 * ne * sizeof(int)},
        {(const void **)&D.u_nstart, um * (nn + 1) * sizeof(int)},
        {(const void **)&D.u_ncorner, um * 8 * ne * sizeof(int)},
        {(const void **)&D.u_nsymm, um * nn},
        {(const void **)&D.evtab, (size_t)h.evtotal * sizeof(unsigned short)}};
    size_t bytes = 0;
    for (auto &s : slots) bytes += align256(s.n * sizeof(double));
    for (auto &s : islots) bytes += align256(s.bytes);
    void *arena = nullptr;
    if (hipMalloc(&arena, bytes) != hipSuccess) return LULESH_ENOMEM;
    d->arena = arena;
    d->arena_bytes = bytes;
    char *p = (char *)arena;
    for (auto &s : slots) {
        *s.p = (double *)p;
        p += align256(s.n * sizeof(double));
    }
    for (auto &s : islots) {
        *s.p = s.bytes ? (const void *)p : nullptr;
        p += align256(s.bytes);
    }
    if (!two) b = a;
    D.cur = a;   // cycle 0: the set-up state is in buffer 0
    D.old = b;
    D.fpx = D.fx;
    D.fpy = D.fy;
    D.fpz = D.fz;
    D.dir_node = D.dir_elem = nullptr;
    D.dir_sbn = D.dir_ghost = nullptr;
    D.dir_pv = nullptr;
    D.dir_pvnode = nullptr;
    D.pv_snap = nullptr;
    D.pv_nslot = 0;
    D.dir_pvlist = nullptr;
    D.dom_base = 0;
    HIPCHK(hipMemset(arena, 0, bytes));
    return LULESH_OK;
}

int domain_upload(lulesh_domain *d)
{
    const HostDomain &h = d->host;
    DomDev &D = d->dev;
    D.nx = h.nx; D.numElem = h.numElem; D.numNode = h.numNode; D.allElem = h.allElem;
    D.col = h.col; D.row = h.row; D.plane = h.plane; D.tp = h.tp;
    for (int f = 0; f < 6; ++f) D.ghost[f] = h.ghost[f];
    if (h.umesh) {
        // coordinates and connectivity tables before the set-up kernels
        // that read them (blocking copies on the null stream)
        const size_t nn = h.numNode, ne = h.numElem;
        HIPCHK(hipMemcpy(D.ns[0].x, h.ux.data(), nn * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(D.ns[0].y, h.uy.data(), nn * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(D.ns[0].z, h.uz.data(), nn * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy((void *)D.u_nodelist, h.u_nodelist.data(), 8 * ne * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy((void *)D.u_face, h.u_face.data(), 8 * ne * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy((void *)D.u_nstart, h.u_nstart.data(), (nn + 1) * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy((void *)D.u_ncorner, h.u_ncorner.data(), 8 * ne * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy((void *)D.u_nsymm, h.u_nsymm.data(), nn, hipMemcpyHostToDevice));
    }
    // the arena was cleared on the null stream; the set-up kernels follow
    // it there, and the blocking copies below wait for them
    HIPCHK(launch_setup(D, h.umesh, nullptr));
    HIPCHK(hipMemcpy(D.e + h.e0_index, &h.e0, 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy((void *)D.elemRep, h.elemRep.data(), (size_t)h.numElem * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy((void *)D.evtab, h.evtab.data(), (size_t)h.evtotal * 2, hipMemcpyHostToDevice));
    HIPCHK(hipStreamSynchronize(nullptr));
    // the device copies are the ones used
    HostDomain &hm = d->host;
    hm.evtab = std::vector<unsigned short>();
    hm.ux = hm.uy = hm.uz = std::vector<double>();
    hm.u_nodelist = hm.u_face = hm.u_nstart = hm.u_ncorner = std::vector<int>();
    hm.u_nsymm = std::vector<unsigned char>();
    return LULESH_OK;
}

// recompute the hourglass coefficients after set-up or a host write of one
// of their inputs
int domain_refresh_hgcoef(lulesh_domain *d, hipStream_t s)
{
    HIPCHK(launch_hg_coef(d->dev.hgcoef, d->dev.ss, d->dev.elemMass, d->dev.volo, d->dev.v,
                          d->host.numElem, s));
    return LULESH_OK;
}

// the node-state buffer holding the current positions and velocities:
// buffer (cycles done) & 1 when the subdomain double-buffers them; -1 when
// the control block cannot be read
static int node_buf(lulesh_domain *d)
{
    const DomDev &D = d->dev;
    if (D.ns[0].x == D.ns[1].x || !D.ctl) return 0;
    DomCtl c;
    if (hipStreamSynchronize(d->g->stream) != hipSuccess) return -1;
    if (hipMemcpy(&c, D.ctl, sizeof(DomCtl), hipMemcpyDeviceToHost) != hipSuccess) return -1;
    return c.cycle & 1;
}

double *field_dev_ptr(lulesh_domain *d, int f, size_t *len)
{
    DomDev &D = d->dev;
    const size_t nn = d->host.numNode, ne = d->host.numElem, all = d->host.allElem;
    if (f >= LULESH_X && f <= LULESH_ZD) {
        const int b = node_buf(d);
        if (b < 0) {
            *len = 0;
            return nullptr;
        }
        const NodeState &S = D.ns[b];
        double *const p[6] = {S.x, S.y, S.z, S.xd, S.yd, S.zd};
        *len = nn;
        return p[f - LULESH_X];
    }
    switch (f) {
    case LULESH_FX: *len = nn; return D.fx;
    case LULESH_FY: *len = nn; return D.fy;
    case LULESH_FZ: *len = nn; return D.fz;
    case LULESH_NODAL_MASS: *len = nn; return D.nodalMass;
    case LULESH_E: *len = ne; return D.e;
    case LULESH_P: *len = ne; return D.p;
    case LULESH_Q: *len = ne; return D.q;
    case LULESH_V: *len = ne; return D.v;
    case LULESH_VOLO: *len = ne; return D.volo;
    case LULESH_DELV: *len = ne; return D.delv;
    case LULESH_VDOV: *len = ne; return D.vdov;
    case LULESH_AREALG: *len = ne; return D.arealg;
    case LULESH_SS: *len = ne; return D.ss;
    case LULESH_ELEM_MASS: *len = ne; return D.elemMass;
    case LULESH_DELV_XI: *len = all; return D.delv_xi;
    case LULESH_DELV_ETA: *len = all; return D.delv_eta;
    case LULESH_DELV_ZETA: *len = all; return D.delv_zeta;
    case LULESH_DELX_XI: *len = ne; return D.delx_xi;
    case LULESH_DELX_ETA: *len = ne; return D.delx_eta;
    case LULESH_DELX_ZETA: *len = ne; return D.delx_zeta;
    case LULESH_VNEW: *len = ne; return D.vnew;
    default: *len = 0; return nullptr;
    }
}

// ---------------------------------------------------------------------------
// device_local backend: receive buffers in HBM; senders write straight into
// them (remote pack), so there is nothing to send and nothing to wait on
// beyond stream order.
// ---------------------------------------------------------------------------
static int dl_alloc_recv(lulesh_domain *, size_t n, double **recv, void *)
{
    void *p = nullptr;
    if (hipMalloc(&p, (n ? n : 1) * sizeof(double)) != hipSuccess) return LULESH_ENOMEM;
    if (hipMemset(p, 0, (n ? n : 1) * sizeof(double)) != hipSuccess) return LULESH_EHIP;
    *recv = (double *)p;
    return LULESH_OK;
}

static int dl_teardown(lulesh_domain *, double *recv, void *)
{
    if (recv) (void)hipFree(recv);
    return LULESH_OK;
}

static double *dl_pack_dest(lulesh_domain *, lulesh_domain *to, size_t offset, void *)
{
    return to->recv + offset;
}

static int dl_send_finish(lulesh_domain *, lulesh_stream_t, void *) { return LULESH_OK; }

static double *dl_recv_base(lulesh_domain *d, void *) { return d->recv; }

static const lulesh_comm_backend_ops kDeviceLocal = {
    "device_local", dl_alloc_recv, dl_teardown, dl_pack_dest, dl_send_finish, dl_recv_base,
    nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};

// ---------------------------------------------------------------------------
// device_copy backend (the reference's Mode A, comm_ipc_common.h:192-225):
// every message is packed into the sender's own send buffer, then copied
// device to device into the receiver's slot, as COMM_SEND_MSG /
// COMM_SEND_CORNER do over IPC.  The copies are queued as the messages are
// handed over (their packs are already on the stream) and issued by the
// next COMM_SEND_FINISH as one copy kernel for everything queued (one
// hipMemcpyAsync per message cost 0.29 ms of a 2x2x2 -s 45 cycle: 56 copies
// per exchange step).
// ---------------------------------------------------------------------------
static double *dc_pack_dest(lulesh_domain *, lulesh_domain *, size_t, void *) { return nullptr; }

static int dc_flush(lulesh_domain *from, lulesh_stream_t stream)
{
    CopyMsgs &q = from->g->dc_pending;
    const hipError_t e = launch_copy_msgs(q, (hipStream_t)stream);
    q.n = 0;
    return e == hipSuccess ? LULESH_OK : LULESH_EHIP;
}

static int dc_send(lulesh_domain *from, lulesh_domain *to, size_t offset, const double *src, size_t count,
                   lulesh_stream_t stream, void *)
{
    if (count > 0xffffffffu) return LULESH_EINVAL;
    CopyMsgs &q = from->g->dc_pending;
    if (q.n == kMaxCopyMsgs) {
        const int rc = dc_flush(from, stream);
        if (rc) return rc;
    }
    q.dst[q.n] = to->recv + offset;
    q.src[q.n] = src;
    q.count[q.n] = (unsigned)count;
    ++q.n;
    return LULESH_OK;
}

static int dc_send_finish(lulesh_domain *from, lulesh_stream_t stream, void *) { return dc_flush(from, stream); }

static const lulesh_comm_backend_ops kDeviceCopy = {
    "device_copy", dl_alloc_recv, dl_teardown, dc_pack_dest, dc_send_finish, dl_recv_base,
    nullptr, nullptr, nullptr, nullptr, dc_send, dc_send, nullptr, nullptr, nullptr, nullptr};

// ---------------------------------------------------------------------------
// comm plans
// ---------------------------------------------------------------------------
static std::string plan_key(int msgType, int nf, const int *fields, bool doFlag, bool planeOnly)
{
    std::string k = std::to_string(msgType) + ":" + std::to_string(doFlag) + std::to_string(planeOnly);
    for (int i = 0; i < nf; ++i) k += "," + std::to_string(fields[i]);
    return k;
}

DevPlan::~DevPlan()
{
    if (!owned) return;
    if (pack) (void)hipFree(pack);
    if (tgt) (void)hipFree(tgt);
    if (src) (void)hipFree(src);
    if (fptr) (void)hipFree(fptr);
}

int group_plan(lulesh_group *g, int msgType, int nf, const int *fields, bool doFlag, bool planeOnly,
               DevPlan **out)
{
    if (nf < 1 || nf > 6) return LULESH_EINVAL;
    if (g->comm_down && g->ndom > 1) return LULESH_EINVAL;   // receive buffers released
    const std::string key = plan_key(msgType, nf, fields, doFlag, planeOnly);
    auto it = g->plans.find(key);
    if (it != g->plans.end()) { *out = it->second.get(); return LULESH_OK; }

    const lulesh_comm_backend_ops *ops = g->backend;
    // a backend may pack locally: every subdomain gets its send buffer
    if (ops != &kDeviceLocal)
        for (auto *d : g->doms)
            if (!d->send) {
                HIPCHK(hipMalloc((void **)&d->send, (size_t)std::max(1, d->host.comBufSize) * sizeof(double)));
                HIPCHK(hipMemset(d->send, 0, (size_t)std::max(1, d->host.comBufSize) * sizeof(double)));
            }
    const int mode = msgType == LULESH_MSG_COMM_SBN ? 0 : msgType == LULESH_MSG_SYNC_POS_VEL ? 1 : 2;
    PlanInput in;
    in.tp = g->tp;
    for (auto *d : g->doms) in.doms.push_back(&d->host);
    in.nfields = nf;
    in.doFlag = doFlag;
    in.planeOnly = planeOnly;
    in.pack_dest = [g](int from, int to, size_t off) {
        return g->backend->pack_dest(g->doms[from], g->doms[to], off, g->backend_ctx);
    };
    in.send_base = [g](int from) { return g->doms[from]->send; };
    in.recv_base = [g](int dom) -> const double * {
        return g->backend->recv_base(g->doms[dom], g->backend_ctx);
    };
    in.skip_corners = (mode == 0 && ops->add_corner) || (mode == 1 && ops->copy_corner);
    in.skip_all = mode == 2 && ops->monoq_copy;
    HostPlan hp;
    const int rc = build_plan(in, hp);
    if (rc != 0) return LULESH_EDIRECTION;
    for (const MsgRef &m : hp.sends)
        if (m.local && !g->doms[m.from]->send) return LULESH_EINVAL;

    std::unique_ptr<DevPlan> p(new DevPlan());
    p->msgType = msgType;
    p->nfields = nf;
    p->mode = mode;
    p->sends = hp.sends;
    p->recvs = hp.recvs;
    p->send_start = hp.send_start;
    p->recv_start = hp.recv_start;
    p->fields.assign(fields, fields + nf);
    p->corner_hook = in.skip_corners;
    p->monoq_hook = in.skip_all;
    p->npack = (int)hp.pack.size();
    p->ntgt = (int)hp.tgt.size();
    p->pack_start = hp.pack_start;
    p->tgt_start = hp.tgt_start;
    HIPCHK(hipMalloc((void **)&p->pack, std::max<size_t>(1, hp.pack.size()) * sizeof(PackEnt)));
    HIPCHK(hipMalloc((void **)&p->tgt, std::max<size_t>(1, hp.tgt.size()) * sizeof(UnpackTgt)));
    HIPCHK(hipMalloc((void **)&p->src, std::max<size_t>(1, hp.src.size()) * sizeof(UnpackSrc)));
    if (!hp.pack.empty())
        HIPCHK(hipMemcpy(p->pack, hp.pack.data(), hp.pack.size() * sizeof(PackEnt), hipMemcpyHostToDevice));
    if (!hp.tgt.empty())
        HIPCHK(hipMemcpy(p->tgt, hp.tgt.data(), hp.tgt.size() * sizeof(UnpackTgt), hipMemcpyHostToDevice));
    if (!hp.src.empty())
        HIPCHK(hipMemcpy(p->src, hp.src.data(), hp.src.size() * sizeof(UnpackSrc), hipMemcpyHostToDevice));
    std::vector<double *> fp((size_t)g->ndom * nf);
    for (int r = 0; r < g->ndom; ++r)
        for (int f = 0; f < nf; ++f) {
            size_t len;
            fp[(size_t)r * nf + f] = field_dev_ptr(g->doms[r], fields[f], &len);
            if (!fp[(size_t)r * nf + f]) return LULESH_EINVAL;
        }
    HIPCHK(hipMalloc((void **)&p->fptr, fp.size() * sizeof(double *)));
    HIPCHK(hipMemcpy(p->fptr, fp.data(), fp.size() * sizeof(double *), hipMemcpyHostToDevice));
    p->host_fptr = fp;
    for (auto *d : g->doms) p->send_base.push_back(d->send);
    *out = p.get();
    g->plans[key] = std::move(p);
    return LULESH_OK;
}

// where messages to d land: COMM_RECV_POST_BUF, else the unpack buffer
static double *post_buf(lulesh_group *g, lulesh_domain *d)
{
    const lulesh_comm_backend_ops *ops = g->backend;
    return ops->recv_post_buf ? ops->recv_post_buf(d, g->backend_ctx) : ops->recv_base(d, g->backend_ctx);
}

// words of one message over all its fields (a corner message carries one
// word per field)
static size_t msg_words(const DevPlan *p, const MsgRef &m) { return (size_t)p->nfields * m.cnt; }

// CommSend of ranks [r0, r1): the pack kernel, then each locally packed
// message handed to the backend (COMM_SEND_MSG / COMM_SEND_CORNER)
int plan_pack(lulesh_group *g, DevPlan *p, int rank, hipStream_t s, bool gated = false)
{
    const int a = rank < 0 ? 0 : p->pack_start[rank];
    const int b = rank < 0 ? p->npack : p->pack_start[rank + 1];
    // a step begins: copies the device_copy backend queued for a step that
    // failed before its COMM_SEND_FINISH are dropped
    g->dc_pending.n = 0;
    HIPCHK(launch_pack(p->pack + a, b - a, p->fptr, p->nfields, gated ? g->gctl : nullptr, s));
    const lulesh_comm_backend_ops *ops = g->backend;
    const int m0 = rank < 0 ? 0 : p->send_start[rank];
    const int m1 = rank < 0 ? (int)p->sends.size() : p->send_start[rank + 1];
    for (int i = m0; i < m1; ++i) {
        const MsgRef &m = p->sends[i];
        if (!m.local) continue;
        lulesh_domain *from = g->doms[m.from], *to = g->doms[m.to];
        const double *src = p->send_base[m.from] + m.send_off;
        const size_t words = msg_words(p, m);
        auto hook = m.kind == 3 ? ops->send_corner : ops->send_msg;
        if (hook) {
            const int rc = hook(from, to, m.recv_off, src, words, (lulesh_stream_t)s, g->backend_ctx);
            if (rc) return rc;
        } else {
            HIPCHK(hipMemcpyAsync(post_buf(g, to) + m.recv_off, src, words * sizeof(double),
                                  hipMemcpyDefault, s));
        }
    }
    return LULESH_OK;
}

// CommSBN / CommSyncPosVel / CommMonoQ of one rank (or all): staging of the
// delivered messages into the unpack buffer (COMM_UNPACK_H2D), the unpack
// kernel, then the messages a backend hook applies itself (corners,
// COMM_ADD_CORNER / COMM_COPY_CORNER; MonoQ planes, COMM_MONOQ_COPY)
int plan_unpack(lulesh_group *g, DevPlan *p, int rank, hipStream_t s, bool gated = false)
{
    const lulesh_comm_backend_ops *ops = g->backend;
    const int m0 = rank < 0 ? 0 : p->recv_start[rank];
    const int m1 = rank < 0 ? (int)p->recvs.size() : p->recv_start[rank + 1];
    for (int i = m0; i < m1; ++i) {
        const MsgRef &m = p->recvs[i];
        lulesh_domain *d = g->doms[m.to];
        double *base = ops->recv_base(d, g->backend_ctx);
        const double *post = post_buf(g, d);
        if (post == base) continue;
        const size_t bytes = msg_words(p, m) * sizeof(double);
        if (ops->unpack_h2d) {
            const int rc = ops->unpack_h2d(d, base + m.recv_off, post + m.recv_off, bytes, (lulesh_stream_t)s,
                                           g->backend_ctx);
            if (rc) return rc;
        } else {
            HIPCHK(hipMemcpyAsync(base + m.recv_off, post + m.recv_off, bytes, hipMemcpyDefault, s));
        }
    }
    const int a = rank < 0 ? 0 : p->tgt_start[rank];
    const int b = rank < 0 ? p->ntgt : p->tgt_start[rank + 1];
    HIPCHK(launch_unpack(p->tgt + a, b - a, p->src, p->fptr, p->nfields, p->mode,
                         gated ? g->gctl : nullptr, s));
    for (int i = m0; i < m1 && (p->corner_hook || p->monoq_hook); ++i) {
        const MsgRef &m = p->recvs[i];
        lulesh_domain *d = g->doms[m.to];
        const double *base = ops->recv_base(d, g->backend_ctx);
        if (p->corner_hook && m.kind == 3) {
            if (gated && p->mode == 0) {
                // a cycle after the group's stop condition must leave its
                // forces alone: in the stepping engine the corner adds run
                // as the library's own launch, which tests the stop flag
                // (the copies of the other hooks rewrite what is already
                // there, so they may run)
                HIPCHK(launch_apply_corner(&p->host_fptr[(size_t)m.to * p->nfields], p->nfields, m.first,
                                           base + m.recv_off, true, g->gctl, s));
                continue;
            }
            auto hook = p->mode == 0 ? ops->add_corner : ops->copy_corner;
            const int rc = hook(d, m.first, base + m.recv_off, p->nfields, p->fields.data(), (lulesh_stream_t)s,
                                g->backend_ctx);
            if (rc) return rc;
        } else if (p->monoq_hook) {
            for (int f = 0; f < p->nfields; ++f) {
                double *dst = p->host_fptr[(size_t)m.to * p->nfields + f] + m.first;
                const int rc = ops->monoq_copy(d, dst, base + m.recv_off + (size_t)f * m.cnt,
                                               (size_t)m.cnt * sizeof(double), (lulesh_stream_t)s, g->backend_ctx);
                if (rc) return rc;
            }
        }
    }
    return LULESH_OK;
}

// the three exchange steps of a cycle and their field lists
// (lulesh.cc:1400-1408, 1607-1617, 2511-2519)
static const int kPosVelFields[6] = {LULESH_X, LULESH_Y, LULESH_Z, LULESH_XD, LULESH_YD, LULESH_ZD};
static const int kMonoQFields[3] = {LULESH_DELV_XI, LULESH_DELV_ETA, LULESH_DELV_ZETA};

DirectTables::~DirectTables()
{
    if (node) (void)hipFree(node);
    if (elem) (void)hipFree(elem);
    if (sbn) (void)hipFree(sbn);
    if (ghost) (void)hipFree(ghost);
    if (pv) (void)hipFree(pv);
    if (pvnode) (void)hipFree(pvnode);
    if (snap) (void)hipFree(snap);
    if (pvlist) (void)hipFree(pvlist);
}

// The direct halo steps are the device_local exchange plans read the other
// way round: the plan says which sender index lands at which receive-buffer
// word (pack) and which words a receiving index takes in slot order
// (unpack); joining the two on the buffer address gives, per receiving
// index, its senders in slot order.
bool group_device_local(const lulesh_group *g) { return g->backend == &kDeviceLocal; }

int group_direct(lulesh_group *g, DirectTables **out)
{
    if (g->direct) { *out = g->direct.get(); return LULESH_OK; }
    if (g->backend != &kDeviceLocal || g->comm_down || g->ndom < 2) return LULESH_EINVAL;
    auto plan = [g](int nf, bool planeOnly, HostPlan &hp, bool doFlag = true) {
        PlanInput in;
        in.tp = g->tp;
        for (auto *d : g->doms) in.doms.push_back(&d->host);
        in.nfields = nf;
        in.doFlag = doFlag;
        in.planeOnly = planeOnly;
        in.pack_dest = [g](int from, int to, size_t off) {
            return g->backend->pack_dest(g->doms[from], g->doms[to], off, g->backend_ctx);
        };
        in.recv_base = [g](int dom) -> const double * {
            return g->backend->recv_base(g->doms[dom], g->backend_ctx);
        };
        return build_plan(in, hp);
    };
    HostPlan sbn, mq, pvp;
    if (plan(3, false, sbn) || plan(3, true, mq) || plan(6, false, pvp, false)) return LULESH_EDIRECTION;
    auto senders = [](const HostPlan &hp) {
        std::unordered_map<const double *, DirRef> m;
        for (const PackEnt &e : hp.pack) m[e.dst] = DirRef{e.src_dom, e.src_idx};
        return m;
    };
    const int nd = g->ndom;
    std::unique_ptr<DirectTables> t(new DirectTables());
    t->node_off.assign(nd + 1, 0);
    t->elem_off.assign(nd + 1, 0);
    for (int r = 0; r < nd; ++r) {
        t->node_off[r + 1] = t->node_off[r] + (size_t)g->doms[r]->host.numNode;
        t->elem_off[r + 1] = t->elem_off[r] + (size_t)g->doms[r]->host.numElem;
    }
    std::vector<IdxSpan> node(t->node_off[nd], IdxSpan{0, 0}), elem(t->elem_off[nd], IdxSpan{0, 0});
    std::vector<DirRef> sbn_list, ghost_list;
    {
        const auto from = senders(sbn);
        for (const UnpackTgt &u : sbn.tgt) {
            node[t->node_off[u.dom] + u.idx] = IdxSpan{(int)sbn_list.size(), u.count};
            for (int m = 0; m < u.count; ++m) {
                const auto it = from.find(sbn.src[u.first + m].src);
                if (it == from.end()) return LULESH_EDIRECTION;
                sbn_list.push_back(it->second);
            }
        }
    }
    {
        const auto from = senders(mq);
        std::map<std::pair<int, int>, std::vector<DirRef>> fills;   // sender element -> ghost slots
        for (const UnpackTgt &u : mq.tgt) {
            if (u.count != 1) return LULESH_EDIRECTION;
            const auto it = from.find(mq.src[u.first].src);
            if (it == from.end()) return LULESH_EDIRECTION;
            fills[{it->second.dom, it->second.idx}].push_back(DirRef{u.dom, u.idx});
        }
        for (auto &kv : fills) {
            elem[t->elem_off[kv.first.first] + kv.first.second] = IdxSpan{(int)ghost_list.size(), (int)kv.second.size()};
            ghost_list.insert(ghost_list.end(), kv.second.begin(), kv.second.end());
        }
    }
    // CommSyncPosVel (lulesh-comm.cc:1261-1680): a receiving copy keeps the
    // last message's value, the final writer's own update as packed before
    // any unpack.  A final writer's own copy can itself be overwritten (a
    // mixed-direction edge), so each copy computes its final writer's
    // update from that writer's own force sum and its pre-update state,
    // which the writer keeps in a snapshot by cycle parity.
    std::vector<int> pv(t->node_off[nd], -1);
    std::vector<PvNode> pvnode;
    {
        const auto from = senders(pvp);
        std::map<std::pair<int, int>, int> slot_of;   // final-writer node -> its snapshot slot
        t->nslot.assign(nd, 0);
        std::vector<std::pair<DirRef, DirRef>> fw;    // (copy, its final writer)
        for (const UnpackTgt &u : pvp.tgt) {
            if (u.count < 1) continue;
            const auto it = from.find(pvp.src[u.first + u.count - 1].src);
            if (it == from.end()) return LULESH_EDIRECTION;
            const DirRef w = it->second;
            if (!slot_of.count({w.dom, w.idx})) slot_of[{w.dom, w.idx}] = t->nslot[w.dom]++;
            fw.push_back({DirRef{u.dom, u.idx}, w});
        }
        auto entry = [&](int dom, int idx) -> PvNode & {
            int &k = pv[t->node_off[dom] + idx];
            if (k < 0) {
                k = (int)pvnode.size();
                pvnode.push_back(PvNode{-1, 0, 0, -1});
            }
            return pvnode[k];
        };
        for (auto &c : fw) {
            PvNode &e = entry(c.first.dom, c.first.idx);
            e.fw_dom = c.second.dom;
            e.fw_idx = c.second.idx;
            e.fw_slot = slot_of[{c.second.dom, c.second.idx}];
        }
        for (auto &kv : slot_of) entry(kv.first.first, kv.first.second).own_slot = kv.second;
        t->snap_off.assign(nd + 1, 0);
        for (int r = 0; r < nd; ++r) t->snap_off[r + 1] = t->snap_off[r] + 12 * (size_t)t->nslot[r];
    }
    std::vector<DirRef> pvl;
    for (int r = 0; r < nd; ++r)
        for (size_t n = 0; n < t->node_off[r + 1] - t->node_off[r]; ++n)
            if (pv[t->node_off[r] + n] >= 0) pvl.push_back(DirRef{r, (int)n});
    t->npv = (int)pvl.size();
    HIPCHK(hipMalloc((void **)&t->pvlist, std::max<size_t>(1, pvl.size()) * sizeof(DirRef)));
    if (!pvl.empty())
        HIPCHK(hipMemcpy(t->pvlist, pvl.data(), pvl.size() * sizeof(DirRef), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc((void **)&t->pv, pv.size() * sizeof(int)));
    HIPCHK(hipMalloc((void **)&t->pvnode, std::max<size_t>(1, pvnode.size()) * sizeof(PvNode)));
    HIPCHK(hipMalloc((void **)&t->snap, std::max<size_t>(1, t->snap_off[nd]) * sizeof(double)));
    HIPCHK(hipMemcpy(t->pv, pv.data(), pv.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!pvnode.empty())
        HIPCHK(hipMemcpy(t->pvnode, pvnode.data(), pvnode.size() * sizeof(PvNode), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc((void **)&t->node, node.size() * sizeof(IdxSpan)));
    HIPCHK(hipMalloc((void **)&t->elem, elem.size() * sizeof(IdxSpan)));
    HIPCHK(hipMalloc((void **)&t->sbn, std::max<size_t>(1, sbn_list.size()) * sizeof(DirRef)));
    HIPCHK(hipMalloc((void **)&t->ghost, std::max<size_t>(1, ghost_list.size()) * sizeof(DirRef)));
    HIPCHK(hipMemcpy(t->node, node.data(), node.size() * sizeof(IdxSpan), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(t->elem, elem.data(), elem.size() * sizeof(IdxSpan), hipMemcpyHostToDevice));
    if (!sbn_list.empty())
        HIPCHK(hipMemcpy(t->sbn, sbn_list.data(), sbn_list.size() * sizeof(DirRef), hipMemcpyHostToDevice));
    if (!ghost_list.empty())
        HIPCHK(hipMemcpy(t->ghost, ghost_list.data(), ghost_list.size() * sizeof(DirRef),
                         hipMemcpyHostToDevice));
    g->direct = std::move(t);
    *out = g->direct.get();
    return LULESH_OK;
}

int group_exchange(lulesh_group *g, int msgType, int nf, const int *fields, hipStream_t s, bool gated)
{
    if (g->ndom == 1) return LULESH_OK;
    const bool doFlag = msgType != LULESH_MSG_SYNC_POS_VEL;
    const bool planeOnly = msgType == LULESH_MSG_MONOQ;
    DevPlan *p = nullptr;
    int rc = group_plan(g, msgType, nf, fields, doFlag, planeOnly, &p);
    if (rc) return rc;
    rc = plan_pack(g, p, -1, s, gated);
    if (rc) return rc;
    for (auto *d : g->doms) {
        rc = g->backend->send_finish(d, s, g->backend_ctx);
        if (rc) return rc;
    }
    return plan_unpack(g, p, -1, s, gated);
}

// ---------------------------------------------------------------------------
// groups
// ---------------------------------------------------------------------------
// a group under construction: released with everything it holds if set-up
// fails part way
struct GroupDeleter {
    void operator()(lulesh_group *g) const { group_free(g); }
};
using GroupPtr = std::unique_ptr<lulesh_group, GroupDeleter>;

static int group_finish(GroupPtr &g, lulesh_group **out);

int group_create(int nx, int tp, int nr, int balance, int cost, lulesh_group **out)
{
    if (!out || !problem_ok(nx, tp, nr, balance, cost)) return LULESH_EINVAL;
    GroupPtr g(new lulesh_group());
    g->nx = nx;
    g->tp = tp;
    g->ndom = tp * tp * tp;
    g->backend = &kDeviceLocal;
    HIPCHK(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    for (int r = 0; r < g->ndom; ++r) {
        lulesh_domain *d = new lulesh_domain();
        d->g = g.get();
        d->rank = r;
        g->doms.push_back(d);
        build_host_domain(d->host, nx, tp, r, nr, balance, cost);
        int rc = domain_alloc(d);
        if (rc) return rc;
        rc = domain_upload(d);
        if (rc) return rc;
        // the comm-setup phase: the backend's whole receive-side set-up,
        // per subdomain (lulesh.cu:515-527 times commAllocRecv per rank)
        const auto t0 = std::chrono::steady_clock::now();
        rc = g->backend->alloc_recv(d, (size_t)d->host.comBufSize, &d->recv, g->backend_ctx);
        if (rc) return rc;
        // a backend's set-up is done on return or ordered on the null
        // stream (hipMalloc + hipMemset); the library's streams do not
        // block on it, so groups stepping meanwhile are not waited for
        HIPCHK(hipStreamSynchronize(nullptr));
        g->phase_setup_ms = std::max(
            g->phase_setup_ms, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    return group_finish(g, out);
}

// control blocks, hourglass coefficients and the initial nodalMass
// exchange of a group whose subdomains are set up
static int group_finish(GroupPtr &g, lulesh_group **out)
{
    // control blocks: GroupCtl then ndom DomCtl in one allocation
    const size_t cbytes = sizeof(GroupCtl) + (size_t)g->ndom * sizeof(DomCtl);
    HIPCHK(hipMalloc(&g->ctl_mem, cbytes));
    g->gctl = (GroupCtl *)g->ctl_mem;
    g->dctl = (DomCtl *)((char *)g->ctl_mem + sizeof(GroupCtl));
    g->stoptime = 1.0e-2;   // lulesh-init.cc:139
    g->qstop = 1.0e+12;
    g->max_cycles = -1;
    int rc = group_reset_ctl(g.get());
    if (rc) return rc;
    for (int r = 0; r < g->ndom; ++r) {
        lulesh_domain *d = g->doms[r];
        DomDev &D = d->dev;
        D.ctl = g->dctl + r;
        D.grp = g->gctl;
    }
    for (auto *d : g->doms) {
        rc = domain_refresh_hgcoef(d, g->stream);
        if (rc) return rc;
    }
    // initial nodalMass exchange (lulesh.cc:3147-3161)
    const int fm = LULESH_NODAL_MASS;
    rc = group_exchange(g.get(), LULESH_MSG_COMM_SBN, 1, &fm, g->stream);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(g->stream));
    *out = g.release();
    return LULESH_OK;
}

int group_create_mesh(const lulesh_mesh *m, int nr, int balance, int cost, lulesh_group **out)
{
    if (!out || !m || !mesh_problem_ok(m->numElem, nr, balance, cost)) return LULESH_EINVAL;
    if ((int)m->elemBC.size() != m->numElem) return LULESH_EINVAL;
    GroupPtr g(new lulesh_group());
    g->nx = 0;
    g->tp = 1;
    g->ndom = 1;
    g->umesh = true;
    g->backend = &kDeviceLocal;
    HIPCHK(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    lulesh_domain *d = new lulesh_domain();
    d->g = g.get();
    d->rank = 0;
    g->doms.push_back(d);
    build_host_domain_mesh(d->host, *m, nr, balance, cost);
    int rc = domain_alloc(d);
    if (rc) return rc;
    rc = domain_upload(d);
    if (rc) return rc;
    // no comm buffers: the unstructured path is single-domain
    return group_finish(g, out);
}

int group_reset_ctl(lulesh_group *g)
{
    std::vector<char> buf(sizeof(GroupCtl) + (size_t)g->ndom * sizeof(DomCtl));
    GroupCtl *G = (GroupCtl *)buf.data();
    DomCtl *D = (DomCtl *)(buf.data() + sizeof(GroupCtl));
    memset(buf.data(), 0, buf.size());
    G->stoptime = g->stoptime;
    G->qstop = g->qstop;
    G->dctl = g->dctl;
    G->max_cycles = g->max_cycles;
    G->ndom = g->ndom;
    double big = 1.0e+20;
    unsigned long long bb;
    memcpy(&bb, &big, 8);
    for (int r = 0; r < g->ndom; ++r) {
        D[r].time = 0.0;
        D[r].deltatime = g->doms.empty() ? 0.0 : g->doms[r]->host.deltatime0;
        D[r].dtcourant_bits = bb;
        D[r].dthydro_bits = bb;
        D[r].cycle = 0;
        D[r].ntime = D[r].time;
        D[r].ndeltatime = D[r].deltatime;
        D[r].ncycle = 0;
    }
    HIPCHK(hipMemcpy(g->ctl_mem, buf.data(), buf.size(), hipMemcpyHostToDevice));
    return LULESH_OK;
}

int group_set_param(lulesh_group *g)
{
    // stoptime / qstop / max_cycles live in the device GroupCtl.  The group
    // may have cycles queued on its own stream, a caller's stream or a
    // batch's stream: the whole device is drained first, and only the
    // parameter words and the done flag are written, so a status a kernel
    // raised is never overwritten.
    HIPCHK(hipDeviceSynchronize());
    static_assert(offsetof(GroupCtl, qstop) == offsetof(GroupCtl, stoptime) + sizeof(double),
                  "stoptime and qstop are adjacent");
    static_assert(offsetof(GroupCtl, done) == offsetof(GroupCtl, max_cycles) + sizeof(int),
                  "max_cycles and done are adjacent");
    const double sq[2] = {g->stoptime, g->qstop};
    // the stop test is re-evaluated at the next TimeIncrement, as the
    // reference's loop condition is on every iteration: a raised cap or
    // stop time resumes a group that had reached the old one
    const int md[2] = {g->max_cycles, 0};
    HIPCHK(hipMemcpy(&g->gctl->stoptime, sq, sizeof(sq), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(&g->gctl->max_cycles, md, sizeof(md), hipMemcpyHostToDevice));
    return LULESH_OK;
}

void group_free(lulesh_group *g)
{
    if (!g) return;
    if (g->engine) engine_destroy(g->engine);
    g->plans.clear();
    for (auto *d : g->doms) {
        if (d->recv) g->backend->teardown(d, d->recv, g->backend_ctx);
        if (d->send) (void)hipFree(d->send);
        if (d->arena) (void)hipFree(d->arena);
        delete d;
    }
    if (g->backend && g->backend->runtime_shutdown) (void)g->backend->runtime_shutdown(g->backend_ctx);
    if (g->ctl_mem) (void)hipFree(g->ctl_mem);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
}

// a host-side write or exchange changed the node state of a decomposed
// group: its engine's PosVel snapshots are stale, so the next run builds a
// new engine (which takes them afresh)
void group_state_changed(lulesh_group *g)
{
    if (g->ndom > 1 && g->engine) {
        (void)hipStreamSynchronize(g->stream);
        engine_destroy(g->engine);
        g->engine = nullptr;
    }
}

int group_read_ctl(lulesh_group *g, GroupCtl *G, DomCtl *D0, hipStream_t s)
{
    std::vector<char> buf(sizeof(GroupCtl) + sizeof(DomCtl));
    HIPCHK(hipMemcpyAsync(buf.data(), g->ctl_mem, buf.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    memcpy(G, buf.data(), sizeof(GroupCtl));
    memcpy(D0, buf.data() + sizeof(GroupCtl), sizeof(DomCtl));
    return LULESH_OK;
}

}  // namespace lulesh

// ===========================================================================
// scratch of lulesh_group_exchange_noalloc
// ===========================================================================
// One file-scope device array: the plan tables and send staging of an
// allocation-free exchange step.  Each GPU the library runs on holds its
// own copy until the process ends.  Host code never uses the array's name
// as an address: its device address is queried once per device (at
// lulesh_init, else at the first such call on the device) and kept in
// g_scratch.
static __device__ __align__(256) unsigned char scratch[LULESH_EXCHANGE_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch[kMaxDevices] = {};
}  // namespace

// the current device's scratch address (queried on first use)
static int scratch_base(unsigned char **out)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return LULESH_EINVAL;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = (unsigned char *)p;
    }
    *out = g_scratch[dev];
    return LULESH_OK;
}

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

static hipStream_t stream_of(lulesh_group *g, lulesh_stream_t s)
{
    return s ? (hipStream_t)s : g->stream;
}

int lulesh_init(int device)
{
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return LULESH_EINVAL;
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        fprintf(stderr, "lulesh_amd: device %d is %s, built for gfx950\n", device, prop.gcnArchName);
        return LULESH_EINVAL;
    }
    unsigned char *sc = nullptr;   // the scratch address of this device
    return scratch_base(&sc);
}

int lulesh_device_cus(void)
{
    int dev = 0, cus = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    return cus > 0 ? cus : LULESH_EHIP;
}

const char *lulesh_status_string(int status)
{
    switch (status) {
    case LULESH_OK: return "ok";
    case LULESH_VOLUME_ERROR: return "VolumeError";
    case LULESH_QSTOP_ERROR: return "QStopError";
    case LULESH_EINVAL: return "invalid argument";
    case LULESH_EHIP: return "HIP runtime error";
    case LULESH_ENOMEM: return "out of device memory";
    case LULESH_EDIRECTION: return "no message expected from that direction";
    case LULESH_EBUSY: return "server queue full";
    default: return "unknown";
    }
}

const char *lulesh_version(void) { return "lulesh_amd 0.1 (gfx950)"; }

// Host allocation failures inside the library (std::bad_alloc) end as
// LULESH_ENOMEM at the C boundary; no C++ exception crosses it.
int lulesh_group_create(int nx, int tp, int nr, int balance, int cost, lulesh_group **out)
{
    if (!out) return LULESH_EINVAL;
    try {
        lulesh_group *g = nullptr;
        const int rc = group_create(nx, tp, nr, balance, cost, &g);
        if (rc) { group_free(g); return rc; }
        *out = g;
        return LULESH_OK;
    } catch (const std::bad_alloc &) {
        return LULESH_ENOMEM;
    }
}

int lulesh_group_create_mesh(const lulesh_mesh *m, int nr, int balance, int cost, lulesh_group **out)
{
    try {
        return group_create_mesh(m, nr, balance, cost, out);
    } catch (const std::bad_alloc &) {
        return LULESH_ENOMEM;
    }
}

int lulesh_group_destroy(lulesh_group *g)
{
    group_free(g);
    return LULESH_OK;
}

int lulesh_group_num_domains(const lulesh_group *g) { return g ? g->ndom : LULESH_EINVAL; }

int lulesh_group_domain(lulesh_group *g, int rank, lulesh_domain **out)
{
    if (!g || rank < 0 || rank >= g->ndom || !out) return LULESH_EINVAL;
    *out = g->doms[rank];
    return LULESH_OK;
}

int lulesh_group_set_stoptime(lulesh_group *g, double t)
{
    if (!g) return LULESH_EINVAL;
    g->stoptime = t;
    return group_set_param(g);
}

int lulesh_group_set_qstop(lulesh_group *g, double q)
{
    if (!g) return LULESH_EINVAL;
    g->qstop = q;
    return group_set_param(g);
}

int lulesh_group_set_max_cycles(lulesh_group *g, int m)
{
    if (!g) return LULESH_EINVAL;
    g->max_cycles = m;
    return group_set_param(g);
}

int lulesh_group_set_keep_forces(lulesh_group *g, int keep)
{
    if (!g) return LULESH_EINVAL;
    g->keep_forces = keep != 0;
    // the captured cycle depends on it
    if (g->engine) { engine_destroy(g->engine); g->engine = nullptr; }
    return LULESH_OK;
}

// drops an engine whose PosVel snapshots an allocation-free exchange made
// stale (the exchange itself could not free it)
static void drop_stale_engine(lulesh_group *g)
{
    if (g->engine_stale && g->engine) {
        (void)hipDeviceSynchronize();
        engine_destroy(g->engine);
        g->engine = nullptr;
    }
    g->engine_stale = false;
}

int lulesh_group_enqueue(lulesh_group *g, int ncycles, lulesh_stream_t stream)
{
    if (!g || ncycles < 0) return LULESH_EINVAL;
    drop_stale_engine(g);
    if (!g->engine) {
        std::vector<lulesh_group *> one{g};
        const int rc = engine_create(one, &g->engine);
        if (rc) return rc;
    }
    return engine_enqueue(g->engine, ncycles, stream_of(g, stream));
}

int lulesh_group_sync(lulesh_group *g, int *cycle, double *time, double *dt, lulesh_stream_t stream)
{
    if (!g) return LULESH_EINVAL;
    GroupCtl G;
    DomCtl D0;
    const int rc = group_read_ctl(g, &G, &D0, stream_of(g, stream));
    if (rc) return rc;
    if (cycle) *cycle = D0.cycle;
    if (time) *time = D0.time;
    if (dt) *dt = D0.deltatime;
    return G.status;
}

int lulesh_group_run(lulesh_group *g, int cycles_per_sync, lulesh_stream_t stream)
{
    if (!g) return LULESH_EINVAL;
    if (cycles_per_sync <= 0) cycles_per_sync = 32;
    hipStream_t s = stream_of(g, stream);
    for (;;) {
        int rc = lulesh_group_enqueue(g, cycles_per_sync, s);
        if (rc) return rc;
        GroupCtl G;
        DomCtl D0;
        rc = group_read_ctl(g, &G, &D0, s);
        if (rc) return rc;
        if (G.status) return G.status;
        if (G.done) return LULESH_OK;
    }
}

int lulesh_group_profile(lulesh_group *g, int ncycles, double ms[7])
{
    if (!g || ncycles < 0 || !ms) return LULESH_EINVAL;
    drop_stale_engine(g);
    if (!g->engine) {
        std::vector<lulesh_group *> one{g};
        const int rc = engine_create(one, &g->engine);
        if (rc) return rc;
    }
    return engine_profile(g->engine, ncycles, g->stream, ms);
}

int lulesh_group_force_plan(lulesh_group *g, int *workgroups, int *node_planes_max, int *paired)
{
    if (!g) return LULESH_EINVAL;
    drop_stale_engine(g);
    if (!g->engine) {
        std::vector<lulesh_group *> one{g};
        const int rc = engine_create(one, &g->engine);
        if (rc) return rc;
    }
    int longest = 0;
    if (g->engine->n_ctiles > 0) {
        std::vector<ColTile> t(g->engine->n_ctiles);
        if (hipMemcpy(t.data(), g->engine->d_ctiles, t.size() * sizeof(ColTile), hipMemcpyDeviceToHost) !=
            hipSuccess)
            return LULESH_EHIP;
        for (const ColTile &c : t) longest = std::max(longest, c.kn);
    }
    if (workgroups) *workgroups = g->engine->n_ctiles;
    if (node_planes_max) *node_planes_max = longest;
    if (paired) *paired = g->engine->paired_runs ? 1 : 0;
    return LULESH_OK;
}

int lulesh_group_verify(lulesh_group *g, int rank, double out[4])
{
    if (!g || rank < 0 || rank >= g->ndom || !out) return LULESH_EINVAL;
    lulesh_domain *d = g->doms[rank];
    if (g->umesh) {
        // the unstructured path has no plane-0 symmetry test in the
        // reference (VerifyAndWriteFinalOutput's structured branch only,
        // LULESH/cuda/src/lulesh.cu:4700-4737): the octant corner's energy
        HIPCHK(hipStreamSynchronize(g->stream));
        HIPCHK(hipMemcpy(&out[0], d->dev.e + d->host.e0_index, 8, hipMemcpyDeviceToHost));
        out[1] = out[2] = out[3] = 0.0;
        return LULESH_OK;
    }
    const int nx = g->nx;
    std::vector<double> e((size_t)nx * nx);
    HIPCHK(hipStreamSynchronize(g->stream));
    HIPCHK(hipMemcpy(e.data(), d->dev.e, e.size() * 8, hipMemcpyDeviceToHost));
    // VerifyAndWriteFinalOutput, lulesh-util.cc:192-210
    double MaxAbsDiff = 0.0, TotalAbsDiff = 0.0, MaxRelDiff = 0.0;
    for (int j = 0; j < nx; ++j)
        for (int k = j + 1; k < nx; ++k) {
            const double AbsDiff = fabs(e[(size_t)j * nx + k] - e[(size_t)k * nx + j]);
            TotalAbsDiff += AbsDiff;
            if (MaxAbsDiff < AbsDiff) MaxAbsDiff = AbsDiff;
            const double RelDiff = AbsDiff / e[(size_t)k * nx + j];
            if (MaxRelDiff < RelDiff) MaxRelDiff = RelDiff;
        }
    out[0] = e[0];
    out[1] = MaxAbsDiff;
    out[2] = TotalAbsDiff;
    out[3] = MaxRelDiff;
    return LULESH_OK;
}

int lulesh_domain_rank(const lulesh_domain *d) { return d ? d->rank : LULESH_EINVAL; }

int lulesh_domain_sizes(const lulesh_domain *d, int *nx, int *ne, int *nn, int *all)
{
    if (!d) return LULESH_EINVAL;
    if (nx) *nx = d->host.nx;
    if (ne) *ne = d->host.numElem;
    if (nn) *nn = d->host.numNode;
    if (all) *all = d->host.allElem;
    return LULESH_OK;
}

int lulesh_domain_get_field(lulesh_domain *d, int f, double *host, size_t n)
{
    if (!d || !host) return LULESH_EINVAL;
    // a one-subdomain group stores its forces only when asked to
    if (f >= LULESH_FX && f <= LULESH_FZ && d->g->ndom == 1 && !d->g->keep_forces) return LULESH_EINVAL;
    size_t len;
    double *p = field_dev_ptr(d, f, &len);
    if (!p || n < len) return LULESH_EINVAL;
    HIPCHK(hipStreamSynchronize(d->g->stream));
    HIPCHK(hipMemcpy(host, p, len * 8, hipMemcpyDeviceToHost));
    return (int)len;
}

int lulesh_domain_set_field(lulesh_domain *d, int f, const double *host, size_t n)
{
    if (!d || !host) return LULESH_EINVAL;
    size_t len;
    double *p = field_dev_ptr(d, f, &len);
    if (!p || n < len) return LULESH_EINVAL;
    HIPCHK(hipStreamSynchronize(d->g->stream));
    HIPCHK(hipMemcpy(p, host, len * 8, hipMemcpyHostToDevice));
    if (f >= LULESH_X && f <= LULESH_ZD) group_state_changed(d->g);
    if (f == LULESH_SS || f == LULESH_V || f == LULESH_VOLO || f == LULESH_ELEM_MASS) {
        const int rc = domain_refresh_hgcoef(d, d->g->stream);
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(d->g->stream));
    }
    return (int)len;
}

int lulesh_domain_field_ptr(lulesh_domain *d, int f, double **dev)
{
    if (!d || !dev) return LULESH_EINVAL;
    size_t len;
    *dev = field_dev_ptr(d, f, &len);
    return *dev ? LULESH_OK : LULESH_EINVAL;
}

int lulesh_domain_get_elem_rep(lulesh_domain *d, int *host, size_t n)
{
    if (!d || !host || n < d->host.elemRep.size()) return LULESH_EINVAL;
    memcpy(host, d->host.elemRep.data(), d->host.elemRep.size() * sizeof(int));
    return (int)d->host.elemRep.size();
}

int lulesh_domain_get_region(lulesh_domain *d, int *host, size_t n)
{
    if (!d || !host || n < d->host.region.size()) return LULESH_EINVAL;
    memcpy(host, d->host.region.data(), d->host.region.size() * sizeof(int));
    return (int)d->host.region.size();
}

// ---------------------------------------------------------------------------
// the reference's comm interface
// ---------------------------------------------------------------------------
static int check_dims(const lulesh_domain *d, int msgType, int dx, int dy, int dz)
{
    if (d->g->umesh) return LULESH_EINVAL;   // no comm on the unstructured path
    const int n = d->host.nx + (msgType == LULESH_MSG_MONOQ ? 0 : 1);
    if (dx != n || dy != n || dz != n) return LULESH_EINVAL;
    if (msgType != LULESH_MSG_COMM_SBN && msgType != LULESH_MSG_SYNC_POS_VEL &&
        msgType != LULESH_MSG_MONOQ)
        return LULESH_EINVAL;
    return LULESH_OK;
}

// CommRecv (lulesh-comm.cc:59-353).  device_local is one-sided: senders
// write into this subdomain's buffer directly, so posting a receive only
// validates the request (COMM_RECV_SKIP of the one-sided backends,
// cuda comm/comm_backend.h:15).
int CommRecv(lulesh_domain *d, int msgType, int xferFields, int dx, int dy, int dz, int doRecv,
             int planeOnly)
{
    if (!d || xferFields < 1 || xferFields > 6) return LULESH_EINVAL;
    const int rc = check_dims(d, msgType, dx, dy, dz);
    if (rc) return rc;
    if ((msgType == LULESH_MSG_MONOQ) != (planeOnly != 0)) return LULESH_EINVAL;
    lulesh_group *g = d->g;
    if (g->ndom == 1) return LULESH_OK;   // numRanks == 1: nothing to post
    if (g->backend->recv_skip) {
        const int r = g->backend->recv_skip(d, msgType, xferFields, doRecv, planeOnly, g->backend_ctx);
        if (r < 0) return r;
    }
    return LULESH_OK;
}

// CommSend (lulesh-comm.cc:357-844 / CommSendGpu, lulesh-comms-gpu.cu:348)
int CommSend(lulesh_domain *d, int msgType, int xferFields, const int *fields, int dx, int dy, int dz,
             int doSend, int planeOnly, lulesh_stream_t stream)
{
    if (!d || !fields || xferFields < 1 || xferFields > 6) return LULESH_EINVAL;
    int rc = check_dims(d, msgType, dx, dy, dz);
    if (rc) return rc;
    lulesh_group *g = d->g;
    if (g->ndom == 1) return LULESH_OK;   // numRanks == 1: nothing to send
    DevPlan *p = nullptr;
    rc = group_plan(g, msgType, xferFields, fields, doSend != 0, planeOnly != 0, &p);
    if (rc) return rc;
    hipStream_t s = stream_of(g, stream);
    rc = plan_pack(g, p, d->rank, s);
    if (rc) return rc;
    return g->backend->send_finish(d, s, g->backend_ctx);
}

// CommSBN (lulesh-comm.cc:848-1257 / CommSBNGpu, lulesh-comms-gpu.cu:805)
int CommSBN(lulesh_domain *d, int xferFields, const int *fields, lulesh_stream_t stream)
{
    if (!d || !fields || xferFields < 1 || xferFields > 6) return LULESH_EINVAL;
    if (d->g->umesh) return LULESH_EINVAL;
    group_state_changed(d->g);
    lulesh_group *g = d->g;
    if (g->ndom == 1) return LULESH_OK;
    DevPlan *p = nullptr;
    int rc = group_plan(g, LULESH_MSG_COMM_SBN, xferFields, fields, true, false, &p);
    if (rc) return rc;
    return plan_unpack(g, p, d->rank, stream_of(g, stream));
}

// CommSyncPosVel (lulesh-comm.cc:1261-1680 / CommSyncPosVelGpu,
// lulesh-comms-gpu.cu:1231)
int CommSyncPosVel(lulesh_domain *d, lulesh_stream_t stream)
{
    if (!d) return LULESH_EINVAL;
    if (d->g->umesh) return LULESH_EINVAL;
    group_state_changed(d->g);
    lulesh_group *g = d->g;
    if (g->ndom == 1) return LULESH_OK;
    DevPlan *p = nullptr;
    int rc = group_plan(g, LULESH_MSG_SYNC_POS_VEL, 6, kPosVelFields, false, false, &p);
    if (rc) return rc;
    return plan_unpack(g, p, d->rank, stream_of(g, stream));
}

// CommMonoQ (lulesh-comm.cc:1684-1837 / CommMonoQGpu, lulesh-comms-gpu.cu:1665)
int CommMonoQ(lulesh_domain *d, lulesh_stream_t stream)
{
    if (!d) return LULESH_EINVAL;
    if (d->g->umesh) return LULESH_EINVAL;
    lulesh_group *g = d->g;
    if (g->ndom == 1) return LULESH_OK;
    DevPlan *p = nullptr;
    int rc = group_plan(g, LULESH_MSG_MONOQ, 3, kMonoQFields, true, true, &p);
    if (rc) return rc;
    return plan_unpack(g, p, d->rank, stream_of(g, stream));
}

int lulesh_comm_buf_size(const lulesh_domain *d) { return d ? d->host.comBufSize : LULESH_EINVAL; }

int lulesh_comm_recv_offset(const lulesh_domain *d, int dc, int dr, int dp, int xfer, int doRecv,
                            int planeOnly)
{
    if (!d || d->g->umesh) return LULESH_EINVAL;
    return recv_offset(d->host, dc, dr, dp, xfer, doRecv != 0, planeOnly != 0);
}

int lulesh_comm_get_recv_buffer(lulesh_domain *d, double *host, size_t n)
{
    if (!d || d->g->umesh || !host || n < (size_t)d->host.comBufSize) return LULESH_EINVAL;
    lulesh_group *g = d->g;
    const double *base = g->backend->recv_base(d, g->backend_ctx);
    HIPCHK(hipStreamSynchronize(g->stream));
    HIPCHK(hipMemcpy(host, base, (size_t)d->host.comBufSize * 8, hipMemcpyDeviceToHost));
    return d->host.comBufSize;
}

int lulesh_domain_get_eos_table(lulesh_domain *d, uint16_t *host, size_t n)
{
    if (!d || !host) return LULESH_EINVAL;
    const long long total = d->host.evtotal;
    if (n < (size_t)total) return LULESH_EINVAL;
    HIPCHK(hipStreamSynchronize(d->g->stream));
    if (total) HIPCHK(hipMemcpy(host, d->dev.evtab, (size_t)total * sizeof(uint16_t), hipMemcpyDeviceToHost));
    return (int)total;
}

const lulesh_comm_backend_ops *lulesh_device_local_backend(void) { return &kDeviceLocal; }

const lulesh_comm_backend_ops *lulesh_device_copy_backend(void) { return &kDeviceCopy; }

int lulesh_group_set_comm_backend(lulesh_group *g, const lulesh_comm_backend_ops *ops, void *ctx)
{
    if (!g || g->umesh || !ops || !ops->alloc_recv || !ops->teardown || !ops->pack_dest ||
        !ops->send_finish || !ops->recv_base)
        return LULESH_EINVAL;
    HIPCHK(hipDeviceSynchronize());
    if (ops->runtime_init) {   // COMM_RUNTIME_INIT
        const int rc = ops->runtime_init(ctx);
        if (rc) return rc;
    }
    const lulesh_comm_backend_ops *old_ops = g->backend;
    void *old_ctx = g->backend_ctx;
    // receive buffers belong to the backend: hand them over
    g->phase_setup_ms = 0.0;
    for (auto *d : g->doms) {
        double *old = d->recv;
        double *fresh = nullptr;
        const auto t0 = std::chrono::steady_clock::now();
        int rc = ops->alloc_recv(d, (size_t)d->host.comBufSize, &fresh, ctx);
        if (rc) return rc;
        HIPCHK(hipDeviceSynchronize());
        g->phase_setup_ms = std::max(
            g->phase_setup_ms, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        if (old) old_ops->teardown(d, old, old_ctx);
        d->recv = fresh;
    }
    if (old_ops && old_ops->runtime_shutdown) (void)old_ops->runtime_shutdown(old_ctx);
    g->comm_down = false;
    g->backend = ops;
    g->backend_ctx = ctx;
    g->plans.clear();
    g->direct.reset();
    if (g->engine) { engine_destroy(g->engine); g->engine = nullptr; }
    return LULESH_OK;
}

int lulesh_comm_apply_corner(lulesh_domain *d, int node, const double *src, int xferFields,
                             const int *fields, int add, lulesh_stream_t stream)
{
    if (!d || !src || !fields || xferFields < 1 || xferFields > 6 || node < 0 ||
        node >= d->host.numNode)
        return LULESH_EINVAL;
    double *f[6];
    for (int i = 0; i < xferFields; ++i) {
        size_t len = 0;
        f[i] = field_dev_ptr(d, fields[i], &len);
        if (!f[i] || fields[i] >= 16) return LULESH_EINVAL;   // node fields only
    }
    HIPCHK(launch_apply_corner(f, xferFields, node, src, add != 0, nullptr, stream_of(d->g, stream)));
    return LULESH_OK;
}

int lulesh_group_comm_teardown(lulesh_group *g)
{
    if (!g) return LULESH_EINVAL;
    if (g->comm_down) return LULESH_OK;
    HIPCHK(hipStreamSynchronize(g->stream));
    if (g->engine) { engine_destroy(g->engine); g->engine = nullptr; }
    g->plans.clear();
    g->direct.reset();
    g->phase_free_ms = 0.0;
    for (auto *d : g->doms) {
        // the comm-free phase, per subdomain (lulesh.cu:4909-4913)
        const auto t0 = std::chrono::steady_clock::now();
        if (d->recv) {
            const int rc = g->backend->teardown(d, d->recv, g->backend_ctx);
            if (rc) return rc;
        }
        HIPCHK(hipDeviceSynchronize());
        d->recv = nullptr;
        g->phase_free_ms = std::max(
            g->phase_free_ms, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    g->comm_down = true;
    return LULESH_OK;
}

int lulesh_group_phase_times(const lulesh_group *g, double out[4])
{
    if (!g || !out) return LULESH_EINVAL;
    out[0] = g->phase_setup_ms;
    out[1] = 0.0;   // window construction: no window in the device-local backend
    out[2] = 0.0;   // peer query / mapping: none either
    out[3] = g->phase_free_ms;
    return LULESH_OK;
}

int lulesh_domain_write_solution(lulesh_domain *d, const char *path)
{
    if (!d || !path) return LULESH_EINVAL;
    const size_t nn = (size_t)d->host.numNode;
    std::vector<double> x(nn), y(nn), z(nn);
    int rc = lulesh_domain_get_field(d, LULESH_X, x.data(), nn);
    if (rc < 0) return rc;
    rc = lulesh_domain_get_field(d, LULESH_Y, y.data(), nn);
    if (rc < 0) return rc;
    rc = lulesh_domain_get_field(d, LULESH_Z, z.data(), nn);
    if (rc < 0) return rc;
    FILE *f = fopen(path, "wb");
    if (!f) return LULESH_EINVAL;
    // write_solution's record per node (lulesh.cu:4584-4603)
    for (size_t i = 0; i < nn; ++i)
        fprintf(f, "%10d\n%.10f\n%.10f\n%.10f\n", (int)i, x[i], y[i], z[i]);
    return fclose(f) == 0 ? LULESH_OK : LULESH_EINVAL;
}

int lulesh_group_exchange(lulesh_group *g, int msgType, int nf, const int *fields, lulesh_stream_t stream)
{
    if (!g || g->umesh || !fields) return LULESH_EINVAL;
    group_state_changed(g);
    return group_exchange(g, msgType, nf, fields, stream_of(g, stream));
}

// lulesh_group_exchange with its plan tables and send staging in the
// device scratch array instead of allocations: the same plan (built on the
// host), the same pack / send / unpack path, the same results.  Scratch
// layout: the send staging of every subdomain (comBufSize doubles each,
// 256-byte aligned), then the pack entries, unpack targets, unpack sources
// and field pointers.  Every scratch word a step reads, the same step
// writes first, so nothing needs clearing.  Returns after the step
// completes (the next call may overwrite the scratch).
int lulesh_group_exchange_noalloc(lulesh_group *g, int msgType, int nf, const int *fields,
                                  lulesh_stream_t stream)
{
    if (!g || g->umesh || !fields || nf < 1 || nf > 6) return LULESH_EINVAL;
    if (msgType != LULESH_MSG_COMM_SBN && msgType != LULESH_MSG_SYNC_POS_VEL && msgType != LULESH_MSG_MONOQ)
        return LULESH_EINVAL;
    if (g->ndom == 1) return LULESH_OK;
    if (g->comm_down) return LULESH_EINVAL;
    unsigned char *base = nullptr;
    int rc = scratch_base(&base);
    if (rc) return rc;
    try {
        const int ndom = g->ndom;
        const bool doFlag = msgType != LULESH_MSG_SYNC_POS_VEL;
        const bool planeOnly = msgType == LULESH_MSG_MONOQ;
        const int mode = msgType == LULESH_MSG_COMM_SBN ? 0 : msgType == LULESH_MSG_SYNC_POS_VEL ? 1 : 2;
        const lulesh_comm_backend_ops *ops = g->backend;
        std::vector<double *> fp((size_t)ndom * nf);
        for (int r = 0; r < ndom; ++r)
            for (int f = 0; f < nf; ++f) {
                size_t len;
                fp[(size_t)r * nf + f] = field_dev_ptr(g->doms[r], fields[f], &len);
                if (!fp[(size_t)r * nf + f]) return LULESH_EINVAL;
            }
        std::vector<double *> stage(ndom);
        size_t off = 0;
        for (int r = 0; r < ndom; ++r) {
            stage[r] = (double *)(base + off);
            off += align256((size_t)std::max(1, g->doms[r]->host.comBufSize) * sizeof(double));
        }
        PlanInput in;
        in.tp = g->tp;
        for (auto *d : g->doms) in.doms.push_back(&d->host);
        in.nfields = nf;
        in.doFlag = doFlag;
        in.planeOnly = planeOnly;
        in.pack_dest = [g](int from, int to, size_t o) {
            return g->backend->pack_dest(g->doms[from], g->doms[to], o, g->backend_ctx);
        };
        in.send_base = [&stage](int from) { return stage[from]; };
        in.recv_base = [g](int dom) -> const double * {
            return g->backend->recv_base(g->doms[dom], g->backend_ctx);
        };
        in.skip_corners = (mode == 0 && ops->add_corner) || (mode == 1 && ops->copy_corner);
        in.skip_all = mode == 2 && ops->monoq_copy;
        HostPlan hp;
        if (build_plan(in, hp) != 0) return LULESH_EDIRECTION;
        const size_t o_pack = align256(off);
        const size_t o_tgt = align256(o_pack + hp.pack.size() * sizeof(PackEnt));
        const size_t o_src = align256(o_tgt + hp.tgt.size() * sizeof(UnpackTgt));
        const size_t o_fp = align256(o_src + hp.src.size() * sizeof(UnpackSrc));
        const size_t end = o_fp + fp.size() * sizeof(double *);
        if (end > (size_t)LULESH_EXCHANGE_SCRATCH_BYTES) {
            fprintf(stderr, "lulesh_amd: lulesh_group_exchange_noalloc: this step needs %zu bytes of "
                            "scratch, the library has %zu\n", end, (size_t)LULESH_EXCHANGE_SCRATCH_BYTES);
            return LULESH_ENOMEM;
        }
        hipStream_t s = stream_of(g, stream);
        if (!hp.pack.empty())
            HIPCHK(hipMemcpyAsync(base + o_pack, hp.pack.data(), hp.pack.size() * sizeof(PackEnt),
                                  hipMemcpyHostToDevice, s));
        if (!hp.tgt.empty())
            HIPCHK(hipMemcpyAsync(base + o_tgt, hp.tgt.data(), hp.tgt.size() * sizeof(UnpackTgt),
                                  hipMemcpyHostToDevice, s));
        if (!hp.src.empty())
            HIPCHK(hipMemcpyAsync(base + o_src, hp.src.data(), hp.src.size() * sizeof(UnpackSrc),
                                  hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(base + o_fp, fp.data(), fp.size() * sizeof(double *), hipMemcpyHostToDevice, s));
        DevPlan p;
        p.owned = false;
        p.msgType = msgType;
        p.nfields = nf;
        p.mode = mode;
        p.pack = (PackEnt *)(base + o_pack);
        p.tgt = (UnpackTgt *)(base + o_tgt);
        p.src = (UnpackSrc *)(base + o_src);
        p.fptr = (double **)(base + o_fp);
        p.npack = (int)hp.pack.size();
        p.ntgt = (int)hp.tgt.size();
        p.pack_start = hp.pack_start;
        p.tgt_start = hp.tgt_start;
        p.sends = hp.sends;
        p.recvs = hp.recvs;
        p.send_start = hp.send_start;
        p.recv_start = hp.recv_start;
        p.fields.assign(fields, fields + nf);
        p.host_fptr = fp;
        p.send_base = stage;
        p.corner_hook = in.skip_corners;
        p.monoq_hook = in.skip_all;
        rc = plan_pack(g, &p, -1, s);
        for (int r = 0; r < ndom && rc == LULESH_OK; ++r) rc = ops->send_finish(g->doms[r], s, g->backend_ctx);
        if (rc == LULESH_OK) rc = plan_unpack(g, &p, -1, s);
        // drained before the host tables go and before the next call may
        // overwrite the scratch
        HIPCHK(hipStreamSynchronize(s));
        g->engine_stale = true;   // node state changed behind a decomposed group's engine
        return rc;
    } catch (const std::bad_alloc &) {
        return LULESH_ENOMEM;
    }
}

// element-wise probes of the device arithmetic (parity tests)
int lulesh_probe_cbrt(const double *host_in, double *host_out, int n)
{
    if (n < 0 || (n > 0 && (!host_in || !host_out))) return LULESH_EINVAL;
    if (n == 0) return LULESH_OK;
    double *din = nullptr, *dout = nullptr;
    hipError_t e = hipMalloc((void **)&din, (size_t)n * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&dout, (size_t)n * 8);
    if (e == hipSuccess) e = hipMemcpy(din, host_in, (size_t)n * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_probe_cbrt(din, dout, n, 0);
    if (e == hipSuccess) e = hipMemcpy(host_out, dout, (size_t)n * 8, hipMemcpyDeviceToHost);
    if (din) (void)hipFree(din);
    if (dout) (void)hipFree(dout);
    return e == hipSuccess ? LULESH_OK : LULESH_EHIP;
}

double lulesh_probe_copy(size_t n, int reps)
{
    if (n < 2 || reps < 1) return -1.0;
    double *a = nullptr, *b = nullptr;
    if (hipMalloc((void **)&a, n * 8) != hipSuccess) return -1.0;
    if (hipMalloc((void **)&b, n * 8) != hipSuccess) { (void)hipFree(a); return -1.0; }
    (void)hipMemset(a, 0, n * 8);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)launch_probe_copy(a, b, n, 0);
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r) (void)launch_probe_copy(a, b, n, 0);
    (void)hipEventRecord(e1, 0);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return ms / reps;
}

int lulesh_probe_fp64(int waves_per_simd, int one_chain, int iters, double *tflops, double *clock_mhz)
{
    if (waves_per_simd < 1 || waves_per_simd > 8 || (one_chain != 0 && one_chain != 1) || iters < 1 ||
        iters > 1000000 || !tflops || !clock_mhz)
        return LULESH_EINVAL;
    int dev = 0, cus = 0, wall_khz = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    HIPCHK(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, dev));
    if (cus < 1 || wall_khz < 1) return LULESH_EHIP;
    // w workgroups of four waves per CU: w waves on each of its four SIMDs
    const int grid = cus * waves_per_simd;
    double *out = nullptr;
    unsigned long long *clk = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<unsigned long long> h((size_t)2 * grid);
    float ms = 0.f;
    hipError_t e = hipMalloc((void **)&out, (size_t)grid * 256 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&clk, h.size() * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = launch_probe_fp64(out, clk, grid, iters, one_chain != 0, 0);   // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, 0);
    if (e == hipSuccess) e = launch_probe_fp64(out, clk, grid, iters, one_chain != 0, 0);
    if (e == hipSuccess) e = hipEventRecord(e1, 0);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess) e = hipMemcpy(h.data(), clk, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (out) (void)hipFree(out);
    if (clk) (void)hipFree(clk);
    if (e != hipSuccess) return LULESH_EHIP;
    double cycles = 0.0, ticks = 0.0;
    for (int b = 0; b < grid; ++b) {
        cycles += (double)h[2 * b];
        ticks += (double)h[2 * b + 1];
    }
    if (ms <= 0.f || ticks <= 0.0) return LULESH_EHIP;
    *tflops = (double)grid * 256 * iters * 16 / (ms * 1e-3) / 1e12;
    *clock_mhz = cycles / ticks * wall_khz / 1000.0;
    return LULESH_OK;
}

int lulesh_probe_range_math(const double *a, double *out2, int n)
{
    if (n < 0 || (n > 0 && (!a || !out2))) return LULESH_EINVAL;
    if (n == 0) return LULESH_OK;
    double *da = nullptr, *dout = nullptr;
    hipError_t e = hipMalloc((void **)&da, (size_t)n * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&dout, (size_t)n * 16);
    if (e == hipSuccess) e = hipMemcpy(da, a, (size_t)n * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_probe_range_math(da, dout, n, 0);
    if (e == hipSuccess) e = hipMemcpy(out2, dout, (size_t)n * 16, hipMemcpyDeviceToHost);
    if (da) (void)hipFree(da);
    if (dout) (void)hipFree(dout);
    return e == hipSuccess ? LULESH_OK : LULESH_EHIP;
}

int lulesh_probe_sqrt_div(const double *a, const double *b, double *out2, int n)
{
    if (n < 0 || (n > 0 && (!a || !b || !out2))) return LULESH_EINVAL;
    if (n == 0) return LULESH_OK;
    double *da = nullptr, *db = nullptr, *dout = nullptr;
    hipError_t e = hipMalloc((void **)&da, (size_t)n * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&db, (size_t)n * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&dout, (size_t)n * 16);
    if (e == hipSuccess) e = hipMemcpy(da, a, (size_t)n * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db, b, (size_t)n * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_probe_sqrt_div(da, db, dout, n, 0);
    if (e == hipSuccess) e = hipMemcpy(out2, dout, (size_t)n * 16, hipMemcpyDeviceToHost);
    if (da) (void)hipFree(da);
    if (db) (void)hipFree(db);
    if (dout) (void)hipFree(dout);
    return e == hipSuccess ? LULESH_OK : LULESH_EHIP;
}

}  // extern "C"
