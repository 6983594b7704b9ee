// lulesh_plan.h -- host-only planning of the force march's column tiles.
#pragma once
#include <vector>

#include "lulesh_dev.h"

namespace lulesh {

// What the planner needs of an engine: the nx of each subdomain, per member
// group in engine order; the z-chunk asked for (0: chosen by the model);
// the device's CU count; whether a paired plan may replace the even split
// (one group, no z-chunk asked for).
struct ForcePlanInput {
    std::vector<std::vector<int>> group_nx;
    int kchunk = 0;
    int cus = 256;
    bool paired_allowed = false;
};

struct ForcePlan {
    std::vector<ColTile> tiles;   // in tile-index (launch) order
    int kchunk_used = 1;
    bool paired = false;
};

// Whether a paired plan may be used at all (LULESH_PAIRED_RUNS=0 turns it
// off: every engine then takes the even split).
bool paired_runs_wanted();

// Column-tile shape for a mesh of nx elements per edge (ex x ey slots).
void tile_shape(int nx, int &ex, int &ey);

// The force march's column tiles for the engine described by `in`:
// LULESH_OK, or LULESH_EINVAL when a mesh has no tile shape or cus < 1.
int plan_force_columns(const ForcePlanInput &in, ForcePlan &out);

}  // namespace lulesh
