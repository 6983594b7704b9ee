// backend_shim_test.cpp -- the reference-side binding of INTEGRATION.md,
// compiled against a stand-in Domain whose fields are reached through
// pointer-to-member accessors as the reference's are
// (LULESH/cuda/src/lulesh.h:300: typedef Real_t &(Domain::*Domain_member)(Index_t)),
// plus a comm backend written outside the library: "host_staged", which
// delivers every message through pinned host memory the way the
// reference's staged MPI backend does (COMM_RECV_POST_BUF in host memory,
// COMM_SEND_MSG as a device-to-host copy, COMM_UNPACK_H2D, COMM_MONOQ_COPY,
// COMM_ADD_CORNER / COMM_COPY_CORNER through the library's corner launch).
//
// The program steps the same decomposed Sedov problem on three groups --
// the library's device_local and device_copy backends and host_staged --
// driving each exchange step through the shim's CommRecv / CommSendGpu /
// CommSBNGpu / CommSyncPosVelGpu / CommMonoQGpu in the reference's call
// order, and requires every field and every receive buffer to agree bit
// for bit between the three, after single exchanges and after whole runs.
// Exit status 0 = all equal.  Test infrastructure: built by the package
// build, run by tests/test_gpu_backends.py.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <random>
#include <vector>

#include "lulesh_amd.h"

typedef double Real_t;
typedef int Index_t;
typedef int Int_t;

// ---------------------------------------------------------------------------
// stand-in for the reference's Domain: only what the shim touches
// ---------------------------------------------------------------------------
struct Domain {
    lulesh_domain *amd = nullptr;
    std::vector<Real_t> m_store[16];
    Real_t &x(Index_t i) { return m_store[0][i]; }
    Real_t &y(Index_t i) { return m_store[1][i]; }
    Real_t &z(Index_t i) { return m_store[2][i]; }
    Real_t &xd(Index_t i) { return m_store[3][i]; }
    Real_t &yd(Index_t i) { return m_store[4][i]; }
    Real_t &zd(Index_t i) { return m_store[5][i]; }
    Real_t &fx(Index_t i) { return m_store[6][i]; }
    Real_t &fy(Index_t i) { return m_store[7][i]; }
    Real_t &fz(Index_t i) { return m_store[8][i]; }
    Real_t &nodalMass(Index_t i) { return m_store[9][i]; }
    Real_t &delv_xi(Index_t i) { return m_store[10][i]; }
    Real_t &delv_eta(Index_t i) { return m_store[11][i]; }
    Real_t &delv_zeta(Index_t i) { return m_store[12][i]; }
};
typedef Real_t &(Domain::*Domain_member)(Index_t);
typedef hipStream_t cudaStream_t;

// ---------------------------------------------------------------------------
// the shim ("Reference-side binding")
// ---------------------------------------------------------------------------
static int to_field(Domain_member m)
{
    if (m == &Domain::fx) return LULESH_FX;
    if (m == &Domain::fy) return LULESH_FY;
    if (m == &Domain::fz) return LULESH_FZ;
    if (m == &Domain::x) return LULESH_X;
    if (m == &Domain::y) return LULESH_Y;
    if (m == &Domain::z) return LULESH_Z;
    if (m == &Domain::xd) return LULESH_XD;
    if (m == &Domain::yd) return LULESH_YD;
    if (m == &Domain::zd) return LULESH_ZD;
    if (m == &Domain::delv_xi) return LULESH_DELV_XI;
    if (m == &Domain::delv_eta) return LULESH_DELV_ETA;
    if (m == &Domain::delv_zeta) return LULESH_DELV_ZETA;
    if (m == &Domain::nodalMass) return LULESH_NODAL_MASS;
    return -1;
}

// entry points return a status < 0, or 0 / a count on success
static void check(int rc, const char *what)
{
    if (rc < 0) {
        fprintf(stderr, "backend_shim_test: %s failed: %d (%s)\n", what, rc, lulesh_status_string(rc));
        exit(2);
    }
}

void CommRecv(Domain &domain, Int_t msgType, Index_t xferFields, Index_t dx, Index_t dy, Index_t dz,
              bool doRecv, bool planeOnly)
{
    check(CommRecv(domain.amd, msgType, xferFields, dx, dy, dz, doRecv, planeOnly), "CommRecv");
}

void CommSendGpu(Domain &domain, Int_t msgType, Index_t xferFields, Domain_member *fieldData, Index_t dx,
                 Index_t dy, Index_t dz, bool doSend, bool planeOnly, cudaStream_t stream)
{
    int f[6];
    for (int i = 0; i < xferFields; ++i) f[i] = to_field(fieldData[i]);
    check(CommSend(domain.amd, msgType, xferFields, f, dx, dy, dz, doSend, planeOnly, (void *)stream),
          "CommSend");
}

void CommSBNGpu(Domain &domain, Int_t xferFields, Domain_member *fieldData, cudaStream_t *streams)
{
    int f[6];
    for (int i = 0; i < xferFields; ++i) f[i] = to_field(fieldData[i]);
    check(CommSBN(domain.amd, xferFields, f, (void *)streams[0]), "CommSBN");
}

void CommSyncPosVelGpu(Domain &domain, cudaStream_t *streams)
{
    check(CommSyncPosVel(domain.amd, (void *)streams[0]), "CommSyncPosVel");
}

void CommMonoQGpu(Domain &domain, cudaStream_t stream) { check(CommMonoQ(domain.amd, (void *)stream), "CommMonoQ"); }

// ---------------------------------------------------------------------------
// host_staged: a backend defined outside the library
// ---------------------------------------------------------------------------
struct Staged {
    struct Slot {
        lulesh_domain *d;
        double *dev;    // unpack buffer (COMM_RECV_BASE)
        double *post;   // pinned host buffer messages land in (COMM_RECV_POST_BUF)
    };
    std::vector<Slot> slots;
    int inits = 0, shutdowns = 0, skips = 0, sends = 0, corners_sent = 0, h2d = 0;
    int corner_adds = 0, corner_copies = 0, monoq = 0;
    Slot *of(lulesh_domain *d)
    {
        for (auto &s : slots)
            if (s.d == d) return &s;
        return nullptr;
    }
};

static int st_alloc_recv(lulesh_domain *d, size_t n, double **recv, void *ctx)
{
    Staged *s = (Staged *)ctx;
    void *dev = nullptr, *host = nullptr;
    const size_t bytes = (n ? n : 1) * sizeof(double);
    if (hipMalloc(&dev, bytes) != hipSuccess) return LULESH_ENOMEM;
    if (hipHostMalloc(&host, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipFree(dev);
        return LULESH_ENOMEM;
    }
    if (hipMemset(dev, 0, bytes) != hipSuccess) return LULESH_EHIP;
    memset(host, 0, bytes);
    s->slots.push_back({d, (double *)dev, (double *)host});
    *recv = (double *)dev;
    return LULESH_OK;
}

static int st_teardown(lulesh_domain *d, double *recv, void *ctx)
{
    Staged::Slot *sl = ((Staged *)ctx)->of(d);
    if (sl) {
        (void)hipHostFree(sl->post);
        sl->d = nullptr;
    }
    (void)hipFree(recv);
    return LULESH_OK;
}

static double *st_pack_dest(lulesh_domain *, lulesh_domain *, size_t, void *) { return nullptr; }
static int st_send_finish(lulesh_domain *, lulesh_stream_t, void *) { return LULESH_OK; }
static double *st_recv_base(lulesh_domain *d, void *ctx)
{
    Staged::Slot *sl = ((Staged *)ctx)->of(d);
    return sl ? sl->dev : nullptr;
}
static double *st_post(lulesh_domain *d, void *ctx)
{
    Staged::Slot *sl = ((Staged *)ctx)->of(d);
    return sl ? sl->post : nullptr;
}

static int st_runtime_init(void *ctx) { ((Staged *)ctx)->inits++; return LULESH_OK; }
static int st_runtime_shutdown(void *ctx) { ((Staged *)ctx)->shutdowns++; return LULESH_OK; }
static int st_recv_skip(lulesh_domain *, int, int, int, int, void *ctx) { ((Staged *)ctx)->skips++; return LULESH_OK; }
static int st_send_msg(lulesh_domain *, lulesh_domain *to, size_t off, const double *src, size_t count,
                       lulesh_stream_t stream, void *ctx)
{
    Staged *s = (Staged *)ctx;
    s->sends++;
    return hipMemcpyAsync(st_post(to, ctx) + off, src, count * sizeof(double), hipMemcpyDeviceToHost,
                          (hipStream_t)stream) == hipSuccess ? LULESH_OK : LULESH_EHIP;
}
static int st_send_corner(lulesh_domain *from, lulesh_domain *to, size_t off, const double *src, size_t count,
                          lulesh_stream_t stream, void *ctx)
{
    ((Staged *)ctx)->corners_sent++;
    return st_send_msg(from, to, off, src, count, stream, ctx);
}
static int st_unpack_h2d(lulesh_domain *, double *dst, const double *src, size_t bytes, lulesh_stream_t stream,
                         void *ctx)
{
    ((Staged *)ctx)->h2d++;
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream) == hipSuccess ? LULESH_OK
                                                                                                      : LULESH_EHIP;
}
static int st_add_corner(lulesh_domain *d, int node, const double *src, int nf, const int *fields,
                         lulesh_stream_t stream, void *ctx)
{
    ((Staged *)ctx)->corner_adds++;
    return lulesh_comm_apply_corner(d, node, src, nf, fields, 1, stream);
}
static int st_copy_corner(lulesh_domain *d, int node, const double *src, int nf, const int *fields,
                          lulesh_stream_t stream, void *ctx)
{
    ((Staged *)ctx)->corner_copies++;
    return lulesh_comm_apply_corner(d, node, src, nf, fields, 0, stream);
}
static int st_monoq_copy(lulesh_domain *, double *dst, const double *src, size_t bytes, lulesh_stream_t stream,
                         void *ctx)
{
    ((Staged *)ctx)->monoq++;
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream) == hipSuccess
               ? LULESH_OK
               : LULESH_EHIP;
}

static const lulesh_comm_backend_ops kHostStaged = {
    "host_staged", st_alloc_recv, st_teardown, st_pack_dest, st_send_finish, st_recv_base,
    st_runtime_init, st_runtime_shutdown, st_recv_skip, st_post, st_send_msg, st_send_corner,
    st_unpack_h2d, st_add_corner, st_copy_corner, st_monoq_copy};

// ---------------------------------------------------------------------------
// the comparison
// ---------------------------------------------------------------------------
static const int kNodeFields[] = {LULESH_X, LULESH_Y, LULESH_Z, LULESH_XD, LULESH_YD, LULESH_ZD,
                                  LULESH_FX, LULESH_FY, LULESH_FZ};
static const int kElemFields[] = {LULESH_E, LULESH_P, LULESH_Q, LULESH_V, LULESH_SS, LULESH_DELV_XI,
                                  LULESH_DELV_ETA, LULESH_DELV_ZETA};

static std::vector<double> get(lulesh_domain *d, int f)
{
    int nx, ne, nn, all;
    lulesh_domain_sizes(d, &nx, &ne, &nn, &all);
    const size_t n = f < 16 ? nn : (f >= 28 && f <= 30) ? all : ne;
    std::vector<double> v(n);
    check(lulesh_domain_get_field(d, f, v.data(), n), "get_field");
    return v;
}

// fields (and, with bufs, receive buffers) of every subdomain, bit for bit
static int compare(lulesh_group *a, lulesh_group *b, const char *what, bool bufs = true)
{
    const int nd = lulesh_group_num_domains(a);
    int bad = 0;
    for (int r = 0; r < nd; ++r) {
        lulesh_domain *da, *db;
        lulesh_group_domain(a, r, &da);
        lulesh_group_domain(b, r, &db);
        for (int f : kNodeFields) bad += get(da, f) != get(db, f);
        for (int f : kElemFields) bad += get(da, f) != get(db, f);
        if (!bufs) continue;
        const int n = lulesh_comm_buf_size(da);
        std::vector<double> ra(n), rb(n);
        check(lulesh_comm_get_recv_buffer(da, ra.data(), n), "recv buffer");
        check(lulesh_comm_get_recv_buffer(db, rb.data(), n), "recv buffer");
        bad += memcmp(ra.data(), rb.data(), n * sizeof(double)) != 0;
    }
    if (bad) fprintf(stderr, "backend_shim_test: %s: %d arrays differ\n", what, bad);
    return bad;
}

static void seed(lulesh_group *g, int f, unsigned s)
{
    std::mt19937_64 rng(s);
    std::uniform_real_distribution<double> u(-8.0, 8.0);
    const int nd = lulesh_group_num_domains(g);
    for (int r = 0; r < nd; ++r) {
        lulesh_domain *d;
        lulesh_group_domain(g, r, &d);
        std::vector<double> v = get(d, f);
        for (auto &x : v) x = u(rng);
        check(lulesh_domain_set_field(d, f, v.data(), v.size()), "set_field");
    }
}

// one exchange step through the shim, in the reference's call order
static void shim_step(std::vector<Domain> &doms, int nx, int step, hipStream_t s)
{
    Domain_member sbn[3] = {&Domain::fx, &Domain::fy, &Domain::fz};
    Domain_member pv[6] = {&Domain::x, &Domain::y, &Domain::z, &Domain::xd, &Domain::yd, &Domain::zd};
    Domain_member mq[3] = {&Domain::delv_xi, &Domain::delv_eta, &Domain::delv_zeta};
    cudaStream_t streams[1] = {s};
    const int n = nx + 1;
    if (step == 0) {   // lulesh.cc:1372, 1405-1408
        for (auto &d : doms) CommRecv(d, LULESH_MSG_COMM_SBN, 3, n, n, n, true, false);
        for (auto &d : doms) CommSendGpu(d, LULESH_MSG_COMM_SBN, 3, sbn, n, n, n, true, false, s);
        for (auto &d : doms) CommSBNGpu(d, 3, sbn, streams);
    } else if (step == 1) {   // lulesh.cc:1574, 1614-1617
        for (auto &d : doms) CommRecv(d, LULESH_MSG_SYNC_POS_VEL, 6, n, n, n, false, false);
        for (auto &d : doms) CommSendGpu(d, LULESH_MSG_SYNC_POS_VEL, 6, pv, n, n, n, false, false, s);
        for (auto &d : doms) CommSyncPosVelGpu(d, streams);
    } else {   // lulesh.cc:2496, 2515-2519
        for (auto &d : doms) CommRecv(d, LULESH_MSG_MONOQ, 3, nx, nx, nx, true, true);
        for (auto &d : doms) CommSendGpu(d, LULESH_MSG_MONOQ, 3, mq, nx, nx, nx, true, true, s);
        for (auto &d : doms) CommMonoQGpu(d, s);
    }
}

int main(int argc, char **argv)
{
    const int nx = argc > 1 ? atoi(argv[1]) : 4;
    const int tp = argc > 2 ? atoi(argv[2]) : 2;
    const int cycles = argc > 3 ? atoi(argv[3]) : -1;
    check(lulesh_init(0), "lulesh_init");
    Staged staged;
    lulesh_group *g[3];
    for (int i = 0; i < 3; ++i) check(lulesh_group_create(nx, tp, 11, 1, 1, &g[i]), "group_create");
    // fresh receive buffers everywhere (creation's nodalMass exchange left
    // words in device_local's ones)
    check(lulesh_group_set_comm_backend(g[0], lulesh_device_local_backend(), nullptr), "device_local");
    check(lulesh_group_set_comm_backend(g[1], lulesh_device_copy_backend(), nullptr), "device_copy");
    check(lulesh_group_set_comm_backend(g[2], &kHostStaged, &staged), "host_staged");
    if (staged.inits != 1) { fprintf(stderr, "runtime_init not called once\n"); return 1; }
    const int nd = lulesh_group_num_domains(g[0]);
    std::vector<Domain> doms[3];
    for (int i = 0; i < 3; ++i) {
        doms[i].resize(nd);
        for (int r = 0; r < nd; ++r) lulesh_group_domain(g[i], r, &doms[i][r].amd);
    }
    hipStream_t s;
    if (hipStreamCreate(&s) != hipSuccess) return 2;
    int bad = 0;
    const int seeded[3][6] = {{LULESH_FX, LULESH_FY, LULESH_FZ, -1, -1, -1},
                              {LULESH_X, LULESH_Y, LULESH_Z, LULESH_XD, LULESH_YD, LULESH_ZD},
                              {LULESH_DELV_XI, LULESH_DELV_ETA, LULESH_DELV_ZETA, -1, -1, -1}};
    const char *names[3] = {"CommSBN", "CommSyncPosVel", "CommMonoQ"};
    for (int step = 0; step < 3; ++step) {
        for (int i = 0; i < 3; ++i) {
            for (int k = 0; k < 6 && seeded[step][k] >= 0; ++k) seed(g[i], seeded[step][k], 77 + 13 * step + k);
            shim_step(doms[i], nx, step, s);
        }
        if (hipStreamSynchronize(s) != hipSuccess) return 2;
        bad += compare(g[0], g[1], names[step]);
        bad += compare(g[0], g[2], names[step]);
        printf("exchange %s: device_local == device_copy == host_staged: %s\n", names[step],
               bad ? "NO" : "yes");
    }
    // whole runs from a fresh start: the stepping engine exchanges through
    // each backend every cycle
    for (int i = 0; i < 3; ++i) check(lulesh_group_destroy(g[i]), "destroy");
    for (int i = 0; i < 3; ++i) check(lulesh_group_create(nx, tp, 11, 1, 1, &g[i]), "group_create");
    Staged staged2;
    check(lulesh_group_set_comm_backend(g[1], lulesh_device_copy_backend(), nullptr), "device_copy");
    check(lulesh_group_set_comm_backend(g[2], &kHostStaged, &staged2), "host_staged");
    int cyc[3];
    for (int i = 0; i < 3; ++i) {
        if (cycles > 0) check(lulesh_group_set_max_cycles(g[i], cycles), "max_cycles");
        check(lulesh_group_run(g[i], 0, nullptr), "run");
        double t, dt;
        check(lulesh_group_sync(g[i], &cyc[i], &t, &dt, nullptr), "sync");
    }
    // device_local steps decomposed groups without its receive buffers (the
    // direct halo steps), the other two exchange through them
    const int rb = compare(g[0], g[1], "run", false) + compare(g[1], g[2], "run");
    bad += rb + (cyc[0] != cyc[1]) + (cyc[0] != cyc[2]);
    double v[3][4];
    for (int i = 0; i < 3; ++i) lulesh_group_verify(g[i], 0, v[i]);
    printf("run to %s: %d / %d / %d cycles, origin energy %.17e / %.17e / %.17e: %s\n",
           cycles > 0 ? "cycle cap" : "stoptime", cyc[0], cyc[1], cyc[2], v[0][0], v[1][0], v[2][0],
           rb ? "NO" : "yes");
    for (int i = 0; i < 3; ++i) check(lulesh_group_destroy(g[i]), "destroy");
    printf("host_staged hooks: init %d shutdown %d recv_skip %d send_msg %d send_corner %d unpack_h2d %d "
           "add_corner %d copy_corner %d monoq_copy %d\n",
           staged.inits, staged.shutdowns, staged.skips, staged.sends, staged.corners_sent, staged.h2d,
           staged.corner_adds, staged.corner_copies, staged.monoq);
    const bool hooks_ok = staged.inits == 1 && staged.shutdowns == 1 && staged.skips > 0 && staged.sends > 0 &&
                          staged.corners_sent > 0 && staged.h2d > 0 && staged.corner_adds > 0 &&
                          staged.corner_copies > 0 && staged.monoq > 0 && staged2.sends > 0;
    if (!hooks_ok) fprintf(stderr, "backend_shim_test: a hook was not exercised\n");
    printf("%s\n", bad == 0 && hooks_ok ? "BACKENDS_EQUAL" : "BACKENDS_DIFFER");
    return bad == 0 && hooks_ok ? 0 : 1;
}
