// lulesh_comm.h -- halo-exchange plan construction (host side).
#pragma once
#include <stddef.h>

#include <functional>
#include <vector>

#include "lulesh_dev.h"
#include "lulesh_setup.h"

namespace lulesh {

struct PlanInput {
    int tp = 1;
    std::vector<const HostDomain *> doms;    // rank order
    int nfields = 0;
    bool doFlag = true;      // doSend == doRecv of the exchange step
    bool planeOnly = false;  // MonoQ: faces only, element planes
    // where the pack kernel writes the message from -> to that lands at
    // `offset` of the receiver's buffer; null: into the sender's own send
    // buffer (send_base), at the sender's CommSend slot, for a send hook to
    // deliver
    std::function<double *(int from, int to, size_t offset)> pack_dest;
    std::function<double *(int from)> send_base;
    std::function<const double *(int dom)> recv_base;
    // leave these messages out of the unpack tables (a backend hook applies
    // them): corner messages, or every message (MonoQ copies)
    bool skip_corners = false;
    bool skip_all = false;
};

// one message of an exchange step, in the sender's send order (sends) or
// the receiver's slot order (recvs)
struct MsgRef {
    int from, to;
    int kind;             // 1 face, 2 edge, 3 corner
    size_t recv_off;      // offset in the receiver's buffer
    size_t send_off;      // offset in the sender's send buffer (local packs)
    int cnt;              // words per field
    bool local;           // packed into the sender's send buffer
    int first;            // receiver index of the message's first word: the
                          // node of a corner message, the ghost element of
                          // a MonoQ message
};

struct HostPlan {
    std::vector<PackEnt> pack;
    std::vector<UnpackTgt> tgt;
    std::vector<UnpackSrc> src;
    std::vector<int> pack_start;   // per sender rank, into pack
    std::vector<int> tgt_start;    // per receiver rank, into tgt
    std::vector<MsgRef> sends, recvs;
    std::vector<int> send_start;   // per sender rank, into sends
    std::vector<int> recv_start;   // per receiver rank, into recvs
};

// shmRecvOffset (LULESH/cuda/src/comm/comm_backend.h:75-138); -1 if the
// receiver expects no message from that direction
int recv_offset(const HostDomain &h, int dc, int dr, int dp, int xfer, bool doRecv, bool planeOnly);

// returns 0, or -6 when a send has no matching receive slot
int build_plan(const PlanInput &in, HostPlan &out);

}  // namespace lulesh
