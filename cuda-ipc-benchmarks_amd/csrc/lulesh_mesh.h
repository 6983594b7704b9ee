// lulesh_mesh.h -- unstructured hexahedral meshes (the reference's .lmesh
// input, LULESH/cuda/src/lulesh.cu:855-1052).
#pragma once
#include <vector>

#include "lulesh_amd.h"

struct lulesh_mesh {
    int numElem = 0, numNode = 0;
    std::vector<double> x, y, z;     // [numNode]
    std::vector<int> nodelist;       // [numElem][8]
    std::vector<int> nbr;            // [numElem][6]: xi-, xi+, eta-, eta+, zeta-, zeta+
    std::vector<int> symm[3];        // symmetry node sets X, Y, Z
    std::vector<int> freeSurf;       // free-surface node set
    std::vector<int> elemBC;         // derived by mesh_finish, or set by the caller
    int octantCorner = 0;
    double e_deposit = 3.948746e+7;  // lulesh.cu:1049-1050
};

namespace lulesh {

// element limit of an unstructured mesh: the ordered node gather addresses
// one component's eight corner rows (8 x numElem doubles) with 32-bit
// byte offsets
constexpr int kMaxMeshElems = 1 << 26;

// validates every index and derives elemBC and the octant corner
int mesh_finish(lulesh_mesh *m);

}  // namespace lulesh
