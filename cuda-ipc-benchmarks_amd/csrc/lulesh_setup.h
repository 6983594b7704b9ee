// lulesh_setup.h -- host-side Sedov set-up of one subdomain.
#pragma once
#include <stddef.h>

#include <vector>

struct lulesh_mesh;

namespace lulesh {

struct HostDomain {
    int nx = 0, tp = 1, rank = 0, col = 0, row = 0, plane = 0;
    int numElem = 0, numNode = 0, allElem = 0;
    bool rowMin = false, rowMax = false, colMin = false, colMax = false;
    bool planeMin = false, planeMax = false;
    int maxPlaneSize = 0, maxEdgeSize = 0, comBufSize = 0;
    int ghost[6] = {-1, -1, -1, -1, -1, -1};
    std::vector<int> elemRep, region;
    // EOS evaluation table (uploaded, then released) and the number of
    // evaluations before element 64*i, i = 0 .. ceil(numElem/64)
    std::vector<unsigned short> evtab;
    std::vector<long long> evprefix64;
    long long evtotal = 0;
    double e0 = 0.0, deltatime0 = 0.0, volo0 = 0.0;
    int e0_index = 0;   // the element that starts with energy e0
    // unstructured mesh: coordinates and connectivity tables in the layout
    // of DomDev::u_* (uploaded, then released)
    bool umesh = false;
    std::vector<double> ux, uy, uz;
    std::vector<int> u_nodelist, u_face, u_nstart, u_ncorner;
    std::vector<unsigned char> u_nsymm;
};

// decomposition of rank `rank` of tp^3 subdomains: position, neighbour
// flags, comm-buffer sizes (SetupCommBuffers) and ghost offsets
void host_decomp(HostDomain &h, int nx, int tp, int rank);
// region layout (glibc rand()) and elemRep of a decomposed subdomain
void host_regions(HostDomain &h, int nr, int balance, int cost);
// decomposition, comm-buffer sizes, region layout (glibc rand()), e0 and
// dt0; the mesh, volumes and nodal masses are built on the device
void build_host_domain(HostDomain &h, int nx, int tp, int rank, int nr, int balance, int cost);
// one unstructured domain: regions, EOS table, connectivity tables, e0 at
// the octant corner and dt0
void build_host_domain_mesh(HostDomain &h, const lulesh_mesh &m, int nr, int balance, int cost);
size_t domain_device_bytes(int nx, bool corner_array, long long evtotal);
// largest elemRep the region layout of (nr, cost) can assign
long long max_elem_rep(int nr, int cost);
// most regions a problem may have
constexpr int kMaxRegions = 1 << 20;
// problem parameters the library accepts: 32-bit element and node
// indexing, at most 4x4x4 subdomains, an EOS table under 2^31 entries,
// at most kMaxRegions regions whose weights sum_i (i+1)^balance stay
// below 2^31 - 1 (the reference sums them in an int)
bool problem_ok(int nx, int tp, int nr, int balance, int cost);
bool mesh_problem_ok(int numElem, int nr, int balance, int cost);

}  // namespace lulesh
