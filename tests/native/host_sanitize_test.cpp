// host_sanitize_test.cpp -- the library's host-only code and the oracle
// under AddressSanitizer and UndefinedBehaviorSanitizer (test
// infrastructure, built with g++ -fsanitize=address,undefined by
// tests/native/Makefile, run by tests/test_host_sanitize.py; no GPU).
//
// Exercises the one parser of untrusted input, the .lmesh reader
// (lulesh_mesh_read, checked where the reference reads unchecked,
// LULESH/cuda/src/lulesh.cu:855-976), on every file named on the command
// line (the test passes valid, truncated, huge-count and out-of-range
// files) and on in-memory meshes; the set-up (decomposition, comm-buffer
// sizes, glibc rand() region layouts incl. refused weight overflows); the
// exchange plans and receive offsets for every decomposition; and short
// oracle runs, single and decomposed, with exchanges.  Exit status 0 and
// no sanitizer report = clean.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "lulesh_amd.h"
#include "lulesh_comm.h"
#include "lulesh_mesh.h"
#include "lulesh_oracle.h"
#include "lulesh_setup.h"

using namespace lulesh;

static int failures = 0;
#define EXPECT(c)                                                               \
    do {                                                                        \
        if (!(c)) {                                                             \
            fprintf(stderr, "host_sanitize_test: %s:%d: %s\n", __FILE__, __LINE__, #c); \
            ++failures;                                                         \
        }                                                                       \
    } while (0)

static void meshes(int argc, char **argv)
{
    // files named on the command line: "ok:<path>" must read, "bad:<path>"
    // must fail with a status (never crash)
    for (int i = 1; i < argc; ++i) {
        const bool want_ok = strncmp(argv[i], "ok:", 3) == 0;
        const char *path = argv[i] + 4 - (want_ok ? 1 : 0);
        lulesh_mesh *m = nullptr;
        const int rc = lulesh_mesh_read(path, &m);
        EXPECT(want_ok ? rc == LULESH_OK : rc < 0);
        if (rc == LULESH_OK) {
            int ne = 0, nn = 0;
            EXPECT(lulesh_mesh_sizes(m, &ne, &nn) == LULESH_OK);
            std::vector<int> bc(ne);
            EXPECT(lulesh_mesh_get_elem_bc(m, bc.data(), bc.size()) == ne);
            HostDomain h;
            build_host_domain_mesh(h, *m, 11, 1, 1);
            EXPECT((int)h.u_ncorner.size() == 8 * ne);
            lulesh_mesh_destroy(m);
        }
    }
    // an in-memory single hexahedron, then the same with a node id out of range
    double x[8] = {0, 1, 1, 0, 0, 1, 1, 0}, y[8] = {0, 0, 1, 1, 0, 0, 1, 1}, z[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    int nl[8] = {0, 1, 2, 3, 4, 5, 6, 7}, nb[6] = {0, 0, 0, 0, 0, 0};
    int sx[4] = {0, 3, 4, 7}, sy[4] = {0, 1, 4, 5}, sz[4] = {0, 1, 2, 3};
    lulesh_mesh *m = nullptr;
    EXPECT(lulesh_mesh_create(1, 8, x, y, z, nl, nb, sx, 4, sy, 4, sz, 4, nullptr, 0, &m) == LULESH_OK);
    EXPECT(lulesh_mesh_octant_corner(m) == 0);
    lulesh_mesh_destroy(m);
    nl[5] = 8;
    EXPECT(lulesh_mesh_create(1, 8, x, y, z, nl, nb, sx, 4, sy, 4, sz, 4, nullptr, 0, &m) == LULESH_EINVAL);
}

static void setup_and_plans()
{
    // region layouts, including the refused weight overflows
    const int cases[][4] = {{6, 11, 1, 1}, {5, 16, 4, 2}, {4, 1, 1, 1}, {7, 40, 0, 3}, {3, 11, 8, 1}};
    for (auto &c : cases) {
        std::vector<int> reg(c[0] * c[0] * c[0]), rep(reg.size());
        EXPECT(lulesh_region_layout(c[0], 1, 0, c[1], c[2], c[3], reg.data(), rep.data(), reg.size()) == (int)reg.size());
    }
    std::vector<int> reg(512), rep(512);
    EXPECT(lulesh_region_layout(8, 1, 0, 11, 9, 1, reg.data(), rep.data(), 512) == LULESH_EINVAL);
    EXPECT(lulesh_region_layout(8, 1, 0, 100, 5, 1, reg.data(), rep.data(), 512) == LULESH_EINVAL);
    EXPECT(lulesh_region_layout(8, 1, 0, 11, 1, 1, reg.data(), rep.data(), 100) == LULESH_EINVAL);
    // every decomposition the library takes: set-up, receive offsets, and
    // the exchange plans of the three steps, with pack destinations into
    // host stand-ins for the receive buffers
    for (int tp = 1; tp <= 4; ++tp)
        for (int nx : {1, 2, 3, 5}) {
            const int nd = tp * tp * tp;
            std::vector<HostDomain> hs(nd);
            std::vector<std::vector<double>> recv(nd), send(nd);
            for (int r = 0; r < nd; ++r) {
                build_host_domain(hs[r], nx, tp, r, 11, 1, 1);
                EXPECT(hs[r].comBufSize == lulesh_comm_buf_size_of(nx, tp, r));
                recv[r].assign((size_t)hs[r].comBufSize + 1, 0.0);
                send[r].assign((size_t)hs[r].comBufSize + 1, 0.0);
                for (int dc = -1; dc <= 1; ++dc)
                    for (int dr = -1; dr <= 1; ++dr)
                        for (int dp = -1; dp <= 1; ++dp) {
                            const int o = lulesh_shm_recv_offset(nx, tp, r, dc, dr, dp, 6, 1, 0);
                            EXPECT(o < hs[r].comBufSize);
                        }
            }
            for (int step = 0; step < 3; ++step)
                for (int local = 0; local < 2; ++local) {
                    PlanInput in;
                    in.tp = tp;
                    for (auto &h : hs) in.doms.push_back(&h);
                    in.nfields = step == 1 ? 6 : 3;
                    in.doFlag = step != 1;
                    in.planeOnly = step == 2;
                    in.pack_dest = [&](int, int to, size_t off) -> double * {
                        return local ? nullptr : recv[to].data() + off;
                    };
                    in.send_base = [&](int from) { return send[from].data(); };
                    in.recv_base = [&](int d) -> const double * { return recv[d].data(); };
                    in.skip_corners = local && step == 0;
                    HostPlan hp;
                    EXPECT(build_plan(in, hp) == 0);
                    for (const PackEnt &e : hp.pack) {
                        const double *lo = local ? send[e.src_dom].data() : nullptr;
                        (void)lo;
                        EXPECT(e.src_idx >= 0);
                    }
                    for (const MsgRef &mref : hp.recvs)
                        EXPECT(mref.recv_off + (size_t)in.nfields * mref.cnt <= (size_t)hs[mref.to].comBufSize);
                    for (const MsgRef &mref : hp.sends)
                        EXPECT(mref.send_off + (size_t)in.nfields * mref.cnt <= (size_t)hs[mref.from].comBufSize);
                }
        }
}

// the force kernel's column planner over sizes, decompositions and CU
// counts: every node plane of every column in exactly one run
static void force_plans()
{
    std::vector<int> t;
    for (int cus : {1, 8, 80, 256, 304})
        for (int tp = 1; tp <= 2; ++tp)
            for (int nx : {1, 2, 3, 5, 8, 13, 16, 31, 45, 48, 64, 100, 130}) {
                const int n = lulesh_force_plan_of(1, &nx, &tp, 0, cus, nullptr, nullptr, nullptr, 0);
                EXPECT(n > 0);
                if (n <= 0) continue;
                t.assign((size_t)8 * n, -1);
                int kc = 0, paired = -1;
                EXPECT(lulesh_force_plan_of(1, &nx, &tp, 0, cus, &kc, &paired, t.data(), n) == n);
                EXPECT(kc >= 1 && (paired == 0 || paired == 1));
                const int s1 = nx + 1;
                std::vector<int> covered((size_t)tp * tp * tp * s1 * s1 * s1, 0);
                for (int i = 0; i < n; ++i) {
                    const int *r = &t[(size_t)8 * i];
                    EXPECT(r[0] >= 0 && r[0] < tp * tp * tp && r[7] >= 1 && r[3] + r[7] <= s1);
                    if (!(r[0] >= 0 && r[0] < tp * tp * tp && r[7] >= 1 && r[3] >= 0 && r[3] + r[7] <= s1)) continue;
                    for (int k = r[3]; k < r[3] + r[7]; ++k)
                        for (int j = r[2]; j < std::min(s1, r[2] + r[5] - 1); ++j)
                            for (int i2 = r[1]; i2 < std::min(s1, r[1] + r[4] - 1); ++i2)
                                ++covered[(((size_t)r[0] * s1 + k) * s1 + j) * s1 + i2];
                }
                for (int c : covered) EXPECT(c == 1);
            }
    const int z = 0, five = 5, eight = 8, one = 1;
    EXPECT(lulesh_force_plan_of(1, &z, &one, 0, 256, nullptr, nullptr, nullptr, 0) == LULESH_EINVAL);
    EXPECT(lulesh_force_plan_of(1, &eight, &five, 0, 256, nullptr, nullptr, nullptr, 0) == LULESH_EINVAL);
    EXPECT(lulesh_force_plan_of(1, &eight, &one, 0, 0, nullptr, nullptr, nullptr, 0) == LULESH_EINVAL);
    EXPECT(lulesh_force_plan_of(0, &eight, &one, 0, 256, nullptr, nullptr, nullptr, 0) == LULESH_EINVAL);
    // an ensemble of several groups, the z-chunk chosen and given
    const int gnx[5] = {16, 33, 64, 20, 48}, gtp[5] = {1, 1, 1, 2, 1};
    for (int kc : {0, 4, 8}) {
        const int n = lulesh_force_plan_of(5, gnx, gtp, kc, 256, nullptr, nullptr, nullptr, 0);
        EXPECT(n > 0);
        if (n > 0) {
            t.assign((size_t)8 * n, -1);
            int used = 0, paired = -1;
            EXPECT(lulesh_force_plan_of(5, gnx, gtp, kc, 256, &used, &paired, t.data(), n) == n);
            EXPECT(paired == 0 && used >= 1 && (kc == 0 || used == kc));
        }
    }
}

static void oracle_runs()
{
    lo_group *g = lo_create(4, 1, 11, 1, 1, LO_CBRT_EXACT);
    EXPECT(g && lo_run(g, 40) == 0);
    double e0, a, b, c;
    lo_verify(g, 0, &e0, &a, &b, &c);
    EXPECT(e0 > 0.0);
    lo_destroy(g);
    g = lo_create(3, 2, 11, 2, 2, LO_CBRT_LIBM);
    EXPECT(g && lo_run(g, 20) == 0);
    const int fx[3] = {9, 10, 11}, mq[3] = {28, 29, 30}, pv[6] = {0, 1, 2, 3, 4, 5};
    EXPECT(lo_exchange(g, 1024, 3, fx) == 0);
    EXPECT(lo_exchange(g, 2048, 6, pv) == 0);
    EXPECT(lo_exchange(g, 3072, 3, mq) == 0);
    std::vector<double> buf(lo_com_buf_size(g, 7));
    EXPECT(lo_get_recv_buffer(g, 7, buf.data()) == 0 || true);
    lo_destroy(g);
    g = lo_create(2, 3, 4, 1, 1, LO_CBRT_EXACT);
    EXPECT(g && lo_run(g, -1) == 0);
    lo_destroy(g);
}

int main(int argc, char **argv)
{
    meshes(argc, argv);
    setup_and_plans();
    force_plans();
    oracle_runs();
    printf("host_sanitize_test: %s\n", failures ? "FAILED" : "clean");
    return failures ? 1 : 0;
}
