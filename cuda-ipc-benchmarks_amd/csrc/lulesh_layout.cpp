// lulesh_layout.cpp -- host-only layout queries of the C ABI: the receive
// offsets, receive-buffer sizes and region layouts of a decomposition,
// without a group or a GPU (the same host code the groups use).
#include <limits.h>
#include <string.h>

#include <new>

#include "lulesh_amd.h"
#include "lulesh_comm.h"
#include "lulesh_plan.h"
#include "lulesh_setup.h"

using namespace lulesh;

extern "C" {

// a decomposition described by (nx, tp, rank)
static bool decomp_ok(int nx, int tp, int rank)
{
    return nx >= 1 && tp >= 1 && tp <= 4 && rank >= 0 && rank < tp * tp * tp &&
           (long long)nx * nx * nx * 8 <= INT_MAX;
}

int lulesh_shm_recv_offset(int nx, int tp, int recvRank, int dCol, int dRow, int dPlane,
                           int xferFields, int doRecv, int planeOnly)
{
    if (!decomp_ok(nx, tp, recvRank) || xferFields < 1 || xferFields > 6) return LULESH_EINVAL;
    HostDomain h;
    host_decomp(h, nx, tp, recvRank);
    return recv_offset(h, dCol, dRow, dPlane, xferFields, doRecv != 0, planeOnly != 0);
}

int lulesh_comm_buf_size_of(int nx, int tp, int rank)
{
    if (!decomp_ok(nx, tp, rank)) return LULESH_EINVAL;
    HostDomain h;
    host_decomp(h, nx, tp, rank);
    return h.comBufSize;
}

int lulesh_region_layout(int nx, int tp, int rank, int nr, int balance, int cost, int *region,
                         int *rep, size_t n)
{
    if (!decomp_ok(nx, tp, rank) || !problem_ok(nx, tp, nr, balance, cost)) return LULESH_EINVAL;
    if (n < (size_t)nx * nx * nx) return LULESH_EINVAL;
    try {
        HostDomain h;
        host_decomp(h, nx, tp, rank);
        host_regions(h, nr, balance, cost);
        if (region) memcpy(region, h.region.data(), h.region.size() * sizeof(int));
        if (rep) memcpy(rep, h.elemRep.data(), h.elemRep.size() * sizeof(int));
        return h.numElem;
    } catch (const std::bad_alloc &) {
        return LULESH_ENOMEM;
    }
}

int lulesh_force_plan_of(int ngroups, const int *nx, const int *tp, int kchunk, int cus, int *kchunk_used,
                         int *paired, int *tiles, size_t n)
{
    if (ngroups < 1 || ngroups > 65536 || !nx || !tp || kchunk < 0 || kchunk > 4096 || cus < 1 || cus > 65536)
        return LULESH_EINVAL;
    for (int g = 0; g < ngroups; ++g)
        if (!decomp_ok(nx[g], tp[g], 0)) return LULESH_EINVAL;
    try {
        ForcePlanInput in;
        for (int g = 0; g < ngroups; ++g) in.group_nx.emplace_back((size_t)tp[g] * tp[g] * tp[g], nx[g]);
        in.kchunk = kchunk;
        in.cus = cus;
        in.paired_allowed = paired_runs_wanted() && ngroups == 1 && kchunk == 0;
        ForcePlan fp;
        const int rc = plan_force_columns(in, fp);
        if (rc) return rc;
        if (kchunk_used) *kchunk_used = fp.kchunk_used;
        if (paired) *paired = fp.paired ? 1 : 0;
        if (tiles)
            for (size_t t = 0; t < fp.tiles.size() && t < n; ++t) {
                const ColTile &c = fp.tiles[t];
                const int row[8] = {c.dom, c.i0, c.j0, c.k0, c.ex, c.ey, c.flags, c.kn};
                memcpy(tiles + 8 * t, row, sizeof(row));
            }
        return (int)fp.tiles.size();
    } catch (const std::bad_alloc &) {
        return LULESH_ENOMEM;
    }
}

}  // extern "C"
