// lulesh_comm.cpp -- halo-exchange plans of the device_local backend.
//
// A plan turns one exchange step of the reference (CommSend on every
// subdomain, then CommSBN / CommSyncPosVel / CommMonoQ on every subdomain,
// LULESH/omp_4.0/lulesh-comm.cc) into two flat tables:
//   pack entries   one per (message, element): sender index -> receiver
//                  buffer address, at the reference's wire offsets
//                  (shmRecvOffset, LULESH/cuda/src/comm/comm_backend.h:75-138)
//   unpack targets one per receiving index, with its messages in the
//                  receiver's slot order (faces, edges, corners as posted by
//                  CommRecv, lulesh-comm.cc:59-353)
// so the pack and unpack of a whole group are one kernel launch each, and a
// node shared by several messages sees them applied in the reference's
// order -- the sums are the reference's sums, bit for bit.
#include "lulesh_comm.h"

#include <algorithm>
#include <map>

namespace lulesh {

namespace {

struct Dir {
    int dc, dr, dp;
    int gated;   // present only when doRecv (receive table) / doSend (send table)
};

// CommRecv posting order; direction = sender relative to receiver
const Dir kRecvOrder[26] = {
    {0, 0, -1, 1}, {0, 0, 1, 0}, {0, -1, 0, 1}, {0, 1, 0, 0}, {-1, 0, 0, 1}, {1, 0, 0, 0},
    {-1, -1, 0, 1}, {0, -1, -1, 1}, {-1, 0, -1, 1}, {1, 1, 0, 0}, {0, 1, 1, 0}, {1, 0, 1, 0},
    {-1, 1, 0, 0}, {0, -1, 1, 0}, {-1, 0, 1, 0}, {1, -1, 0, 1}, {0, 1, -1, 1}, {1, 0, -1, 1},
    {-1, -1, -1, 1}, {-1, -1, 1, 0}, {1, -1, -1, 1}, {1, -1, 1, 0},
    {-1, 1, -1, 1}, {-1, 1, 1, 0}, {1, 1, -1, 1}, {1, 1, 1, 0}};

// CommSend order (lulesh-comm.cc:357-844); direction = receiver relative to
// sender
const Dir kSendOrder[26] = {
    {0, 0, -1, 0}, {0, 0, 1, 1}, {0, -1, 0, 0}, {0, 1, 0, 1}, {-1, 0, 0, 0}, {1, 0, 0, 1},
    {-1, -1, 0, 0}, {0, -1, -1, 0}, {-1, 0, -1, 0}, {1, 1, 0, 1}, {0, 1, 1, 1}, {1, 0, 1, 1},
    {-1, 1, 0, 1}, {0, -1, 1, 1}, {-1, 0, 1, 1}, {1, -1, 0, 0}, {0, 1, -1, 0}, {1, 0, -1, 0},
    {-1, -1, -1, 0}, {-1, -1, 1, 1}, {1, -1, -1, 0}, {1, -1, 1, 1},
    {-1, 1, -1, 0}, {-1, 1, 1, 1}, {1, 1, -1, 0}, {1, 1, 1, 1}};

int ndims(const Dir &d) { return (d.dc != 0) + (d.dr != 0) + (d.dp != 0); }

bool has_nbr(const HostDomain &h, int dc, int dr, int dp)
{
    if (dc < 0 && !h.colMin) return false;
    if (dc > 0 && !h.colMax) return false;
    if (dr < 0 && !h.rowMin) return false;
    if (dr > 0 && !h.rowMax) return false;
    if (dp < 0 && !h.planeMin) return false;
    if (dp > 0 && !h.planeMax) return false;
    return true;
}

// indices (ascending) of the boundary of an n^3 index box that faces
// direction (dc,dr,dp): the index maps of the reference's pack/unpack loops
void boundary_set(int n, int dc, int dr, int dp, std::vector<int> &out)
{
    out.clear();
    const int c0 = dc > 0 ? n - 1 : 0, c1 = dc ? c0 + 1 : n;
    const int r0 = dr > 0 ? n - 1 : 0, r1 = dr ? r0 + 1 : n;
    const int p0 = dp > 0 ? n - 1 : 0, p1 = dp ? p0 + 1 : n;
    for (int p = p0; p < p1; ++p)
        for (int r = r0; r < r1; ++r)
            for (int c = c0; c < c1; ++c) out.push_back((p * n + r) * n + c);
}

}  // namespace

int recv_offset(const HostDomain &h, int dc, int dr, int dp, int xfer, bool doRecv, bool planeOnly)
{
    const int maxPlaneComm = xfer * h.maxPlaneSize, maxEdgeComm = xfer * h.maxEdgeSize;
    int pmsg = 0, emsg = 0, cmsg = 0;
    for (const Dir &o : kRecvOrder) {
        const int nd = ndims(o);
        if (nd > 1 && planeOnly) break;
        if (!has_nbr(h, o.dc, o.dr, o.dp) || (o.gated && !doRecv)) continue;
        const int off = pmsg * maxPlaneComm + emsg * maxEdgeComm + cmsg * 16;
        if (o.dc == dc && o.dr == dr && o.dp == dp) return off;
        if (nd == 1) ++pmsg;
        else if (nd == 2) ++emsg;
        else ++cmsg;
    }
    return -1;
}

int build_plan(const PlanInput &in, HostPlan &out)
{
    const int tp = in.tp;
    const int ndom = tp * tp * tp;
    const bool planeOnly = in.planeOnly;
    const bool doFlag = in.doFlag;
    const int n = planeOnly ? in.doms[0]->nx : in.doms[0]->nx + 1;
    const int nf = in.nfields;
    out.pack.clear();
    out.tgt.clear();
    out.src.clear();
    out.sends.clear();
    out.recvs.clear();
    out.pack_start.assign(ndom + 1, 0);
    out.tgt_start.assign(ndom + 1, 0);
    out.send_start.assign(ndom + 1, 0);
    out.recv_start.assign(ndom + 1, 0);
    std::vector<int> idx;

    for (int s = 0; s < ndom; ++s) {
        out.pack_start[s] = (int)out.pack.size();
        out.send_start[s] = (int)out.sends.size();
        const HostDomain &hs = *in.doms[s];
        // the sender's own buffer slots (commDataSend, lulesh-comm.cc:357-844)
        int spmsg = 0, semsg = 0, scmsg = 0;
        for (const Dir &o : kSendOrder) {
            const int nd = ndims(o);
            if (planeOnly && nd > 1) break;
            if (!has_nbr(hs, o.dc, o.dr, o.dp) || (o.gated && !doFlag)) continue;
            const int r = s + o.dc + o.dr * tp + o.dp * tp * tp;
            const int off = recv_offset(*in.doms[r], -o.dc, -o.dr, -o.dp, nf, doFlag, planeOnly);
            if (off < 0) return -6;
            const size_t soff = (size_t)spmsg * nf * hs.maxPlaneSize + (size_t)semsg * nf * hs.maxEdgeSize +
                                (size_t)scmsg * 16;
            if (nd == 1) ++spmsg;
            else if (nd == 2) ++semsg;
            else ++scmsg;
            boundary_set(n, o.dc, o.dr, o.dp, idx);
            double *dst = in.pack_dest(s, r, (size_t)off);
            const bool local = dst == nullptr;
            if (local) dst = in.send_base(s) + soff;
            const int cnt = (int)idx.size();
            out.sends.push_back(MsgRef{s, r, nd, (size_t)off, soff, cnt, local, idx[0]});
            for (int i = 0; i < cnt; ++i) {
                PackEnt e;
                e.dst = dst + i;
                e.src_dom = s;
                e.src_idx = idx[i];
                e.cnt = cnt;
                e.pad = 0;
                out.pack.push_back(e);
            }
        }
    }
    out.pack_start[ndom] = (int)out.pack.size();
    out.send_start[ndom] = (int)out.sends.size();

    for (int r = 0; r < ndom; ++r) {
        out.tgt_start[r] = (int)out.tgt.size();
        out.recv_start[r] = (int)out.recvs.size();
        const HostDomain &hr = *in.doms[r];
        // receiving index -> its messages in slot order
        std::map<int, std::vector<UnpackSrc>> lists;
        const double *base = in.recv_base(r);
        int pmsg = 0, emsg = 0, cmsg = 0;
        int ghost = hr.numElem;
        for (const Dir &o : kRecvOrder) {
            const int nd = ndims(o);
            if (planeOnly && nd > 1) break;
            if (!has_nbr(hr, o.dc, o.dr, o.dp) || (o.gated && !doFlag)) continue;
            const int off = pmsg * nf * hr.maxPlaneSize + emsg * nf * hr.maxEdgeSize + cmsg * 16;
            boundary_set(n, o.dc, o.dr, o.dp, idx);
            const int cnt = (int)idx.size();
            const int s = r + o.dc + o.dr * tp + o.dp * tp * tp;
            out.recvs.push_back(MsgRef{s, r, nd, (size_t)off, 0, cnt, false, planeOnly ? ghost : idx[0]});
            const bool skip = in.skip_all || (in.skip_corners && nd == 3);
            for (int i = 0; i < cnt && !skip; ++i) {
                UnpackSrc u;
                u.src = base + off + i;
                u.cnt = cnt;
                u.pad = 0;
                const int target = planeOnly ? ghost + i : idx[i];
                lists[target].push_back(u);
            }
            if (planeOnly) ghost += cnt;
            if (nd == 1) ++pmsg;
            else if (nd == 2) ++emsg;
            else ++cmsg;
        }
        for (auto &kv : lists) {
            UnpackTgt t;
            t.dom = r;
            t.idx = kv.first;
            t.first = (int)out.src.size();
            t.count = (int)kv.second.size();
            out.tgt.push_back(t);
            out.src.insert(out.src.end(), kv.second.begin(), kv.second.end());
        }
    }
    out.tgt_start[ndom] = (int)out.tgt.size();
    out.recv_start[ndom] = (int)out.recvs.size();
    return 0;
}

}  // namespace lulesh
