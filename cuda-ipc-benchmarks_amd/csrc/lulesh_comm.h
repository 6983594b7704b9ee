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
    std::function<double *(int from, int to, size_t offset)> pack_dest;
    std::function<const double *(int dom)> recv_base;
};

struct HostPlan {
    std::vector<PackEnt> pack;
    std::vector<UnpackTgt> tgt;
    std::vector<UnpackSrc> src;
    std::vector<int> pack_start;   // per sender rank, into pack
    std::vector<int> tgt_start;    // per receiver rank, into tgt
};

// shmRecvOffset (LULESH/cuda/src/comm/comm_backend.h:75-138); -1 if the
// receiver expects no message from that direction
int recv_offset(const HostDomain &h, int dc, int dr, int dp, int xfer, bool doRecv, bool planeOnly);

// returns 0, or -6 when a send has no matching receive slot
int build_plan(const PlanInput &in, HostPlan &out);

}  // namespace lulesh
