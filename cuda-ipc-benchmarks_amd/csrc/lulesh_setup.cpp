// lulesh_setup.cpp -- host-side Sedov set-up of one subdomain.
//
// Reproduces the host half of Domain::Domain and its helpers
// (LULESH/omp_4.0/lulesh-init.cc): decomposition, comm-buffer sizes, ghost
// offsets, glibc rand() for the region layout, e0 and dt0.  Coordinates,
// volumes and nodal masses are produced on the device (k_setup_*), with
// the same formulas and the same summation order.
#include "lulesh_setup.h"

#include <math.h>
#include <limits.h>
#include <stdlib.h>

#include "lulesh_mesh.h"

namespace lulesh {

static const unsigned kEvLastHost = 0x8000u, kEvMaskHost = 0x7fffu;

// CalcElemVolume (LULESH/omp_4.0/lulesh.cc:1627-1709), host copy for set-up
static double host_elem_volume(const double x[8], const double y[8], const double z[8])
{
    const double twelveth = 1.0 / 12.0;
    const double dx61 = x[6] - x[1], dy61 = y[6] - y[1], dz61 = z[6] - z[1];
    const double dx70 = x[7] - x[0], dy70 = y[7] - y[0], dz70 = z[7] - z[0];
    const double dx63 = x[6] - x[3], dy63 = y[6] - y[3], dz63 = z[6] - z[3];
    const double dx20 = x[2] - x[0], dy20 = y[2] - y[0], dz20 = z[2] - z[0];
    const double dx50 = x[5] - x[0], dy50 = y[5] - y[0], dz50 = z[5] - z[0];
    const double dx64 = x[6] - x[4], dy64 = y[6] - y[4], dz64 = z[6] - z[4];
    const double dx31 = x[3] - x[1], dy31 = y[3] - y[1], dz31 = z[3] - z[1];
    const double dx72 = x[7] - x[2], dy72 = y[7] - y[2], dz72 = z[7] - z[2];
    const double dx43 = x[4] - x[3], dy43 = y[4] - y[3], dz43 = z[4] - z[3];
    const double dx57 = x[5] - x[7], dy57 = y[5] - y[7], dz57 = z[5] - z[7];
    const double dx14 = x[1] - x[4], dy14 = y[1] - y[4], dz14 = z[1] - z[4];
    const double dx25 = x[2] - x[5], dy25 = y[2] - y[5], dz25 = z[2] - z[5];
    auto det3 = [](double a, double b, double c, double d, double e, double f,
                   double g, double h, double i) {
        return a * (e * i - f * h) + d * (c * h - b * i) + g * (b * f - c * e);
    };
    double vol = det3(dx31 + dx72, dx63, dx20, dy31 + dy72, dy63, dy20, dz31 + dz72, dz63, dz20) +
                 det3(dx43 + dx57, dx64, dx70, dy43 + dy57, dy64, dy70, dz43 + dz57, dz64, dz70) +
                 det3(dx14 + dx25, dx61, dx50, dy14 + dy25, dy61, dy50, dz14 + dz25, dz61, dz50);
    return vol * twelveth;
}

// Domain::CreateRegionIndexSets (lulesh-init.cc:377-484) and the elemRep
// assignment of main (lulesh.cc:3171-3192).  Results never depend on the
// regions -- every EOS rep recomputes the same values -- only the cost does.
static void create_regions(HostDomain &h, int nr, int balance, int cost)
{
    const int ne = h.numElem;
    srand(h.rank);
    h.region.assign(ne, 0);
    std::vector<int> regNum(ne, 0);
    int nextIndex = 0;
    if (nr == 1) {
        while (nextIndex < ne) regNum[nextIndex++] = 1;
    } else {
        std::vector<int> binEnd(nr);
        int costDenominator = 0;
        for (int i = 0; i < nr; ++i) {
            costDenominator += pow((i + 1), balance);
            binEnd[i] = costDenominator;
        }
        int lastReg = -1;
        while (nextIndex < ne) {
            int regionVar = rand() % costDenominator;
            int i = 0;
            while (regionVar >= binEnd[i]) i++;
            int regionNum = ((i + h.rank) % nr) + 1;
            while (regionNum == lastReg) {
                regionVar = rand() % costDenominator;
                i = 0;
                while (regionVar >= binEnd[i]) i++;
                regionNum = ((i + h.rank) % nr) + 1;
            }
            const int binSize = rand() % 1000;
            int elements;
            if (binSize < 773) elements = rand() % 15 + 1;
            else if (binSize < 937) elements = rand() % 16 + 16;
            else if (binSize < 970) elements = rand() % 32 + 32;
            else if (binSize < 974) elements = rand() % 64 + 64;
            else if (binSize < 978) elements = rand() % 128 + 128;
            else if (binSize < 981) elements = rand() % 256 + 256;
            else elements = rand() % 1537 + 512;
            const int runto = elements + nextIndex;
            while (nextIndex < runto && nextIndex < ne) regNum[nextIndex++] = regionNum;
            lastReg = regionNum;
        }
    }
    h.elemRep.assign(ne, 1);
    for (int i = 0; i < ne; ++i) {
        const int r = regNum[i] - 1;
        int rep;
        if (r < nr / 2) rep = 1;
        else if (r < (nr - (nr + 15) / 20)) rep = 1 + cost;
        else rep = 10 * (1 + cost);
        h.region[i] = r;
        h.elemRep[i] = rep;
    }
}

long long max_elem_rep(int nr, int cost)
{
    return (nr + 15) / 20 >= 1 ? 10 * (1 + (long long)cost) : 1 + (long long)cost;
}

// CreateRegionIndexSets accumulates the region weights (i+1)^balance in
// an int (lulesh-init.cc:384-388): every partial sum must stay below
// INT_MAX, or the bins wrap and the region search runs off their end.  The
// region count is also capped (one int per region on the host).
static bool region_weights_ok(int nr, int balance)
{
    if (nr < 1 || nr > kMaxRegions || balance < 0) return false;
    if (nr == 1) return true;
    double sum = 0.0;
    for (int i = 0; i < nr; ++i) {
        sum = floor(sum + pow((double)(i + 1), (double)balance));
        if (!(sum <= (double)INT_MAX)) return false;
    }
    return true;
}

bool problem_ok(int nx, int tp, int nr, int balance, int cost)
{
    if (nx < 1 || tp < 1 || tp > 4 || nr < 1 || balance < 0 || cost < 0) return false;
    if (!region_weights_ok(nr, balance)) return false;
    const long long ne = (long long)nx * nx * nx;
    if (ne * 8 > 2147483647LL) return false;
    return ne * max_elem_rep(nr, cost) <= 2147483647LL;
}

bool mesh_problem_ok(int numElem, int nr, int balance, int cost)
{
    if (numElem < 1 || numElem >= kMaxMeshElems || nr < 1 || balance < 0 || cost < 0) return false;
    if (!region_weights_ok(nr, balance)) return false;
    return (long long)numElem * max_elem_rep(nr, cost) <= INT_MAX;
}

// The EOS work list: element i appears elemRep[i] times, in element order,
// its last appearance flagged.  A workgroup of the EOS kernel deals the
// entries of its tile round-robin over its lanes, so every lane does the
// same number of evaluations whatever the region layout.
static void build_eval_table(HostDomain &h)
{
    const int ne = h.numElem;
    long long total = 0;
    h.evprefix64.assign((size_t)(ne + 63) / 64 + 1, 0);
    for (int i = 0; i < ne; ++i) {
        if (i % 64 == 0) h.evprefix64[i / 64] = total;
        total += h.elemRep[i];
    }
    h.evprefix64.back() = total;
    h.evtotal = total;
    h.evtab.resize((size_t)total);
    size_t w = 0;
    for (int i = 0; i < ne; ++i) {
        const unsigned short id = (unsigned short)(i & kEvMaskHost);
        for (int r = 1; r < h.elemRep[i]; ++r) h.evtab[w++] = id;
        h.evtab[w++] = (unsigned short)(id | kEvLastHost);
    }
}

void host_decomp(HostDomain &h, int nx, int tp, int rank)
{
    h.nx = nx;
    h.tp = tp;
    h.rank = rank;
    // InitMeshDecomp, lulesh-init.cc:635-712 (one domain per rank)
    h.col = rank % tp;
    h.row = (rank / tp) % tp;
    h.plane = rank / (tp * tp);
    const int en = nx + 1;
    h.numElem = nx * nx * nx;
    h.numNode = en * en * en;
    h.allElem = h.numElem + 6 * nx * nx;

    // SetupCommBuffers, lulesh-init.cc:317-372
    h.rowMin = h.row != 0;
    h.rowMax = h.row != tp - 1;
    h.colMin = h.col != 0;
    h.colMax = h.col != tp - 1;
    h.planeMin = h.plane != 0;
    h.planeMax = h.plane != tp - 1;
    const int pad = 16;   // 128-byte coherence pad in doubles
    h.maxPlaneSize = ((en * en) + pad - 1) & ~(pad - 1);
    h.maxEdgeSize = (en + pad - 1) & ~(pad - 1);
    const int faces = h.rowMin + h.rowMax + h.colMin + h.colMax + h.planeMin + h.planeMax;
    const int edges = (h.rowMin & h.colMin) + (h.rowMin & h.planeMin) + (h.colMin & h.planeMin) +
                      (h.rowMax & h.colMax) + (h.rowMax & h.planeMax) + (h.colMax & h.planeMax) +
                      (h.rowMax & h.colMin) + (h.rowMin & h.planeMax) + (h.colMin & h.planeMax) +
                      (h.rowMin & h.colMax) + (h.rowMax & h.planeMin) + (h.colMax & h.planeMin);
    int corners = 0;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c)
                corners += (a ? h.rowMax : h.rowMin) & (b ? h.colMax : h.colMin) &
                           (c ? h.planeMax : h.planeMin);
    h.comBufSize = faces * h.maxPlaneSize * 6 + edges * h.maxEdgeSize * 6 + corners * pad;

    // ghost offsets, SetupBoundaryConditions, lulesh-init.cc:545-574
    int pidx = h.numElem;
    const bool present[6] = {h.planeMin, h.planeMax, h.rowMin, h.rowMax, h.colMin, h.colMax};
    for (int f = 0; f < 6; ++f) {
        if (present[f]) { h.ghost[f] = pidx; pidx += nx * nx; }
        else h.ghost[f] = -1;
    }
}

void host_regions(HostDomain &h, int nr, int balance, int cost) { create_regions(h, nr, balance, cost); }

void build_host_domain(HostDomain &h, int nx, int tp, int rank, int nr, int balance, int cost)
{
    host_decomp(h, nx, tp, rank);
    create_regions(h, nr, balance, cost);
    build_eval_table(h);

    // volume of element 0 for dt0: its corner coordinates by the BuildMesh
    // formula (lulesh-init.cc:192-212)
    {
        const double mesh = (double)(tp * nx);
        double xl[8], yl[8], zl[8];
        for (int c = 0; c < 8; ++c) {
            const int di = ((c + 1) >> 1) & 1, dj = (c >> 1) & 1, dk = c >> 2;
            xl[c] = 1.125 * (double)(h.col * nx + di) / mesh;
            yl[c] = 1.125 * (double)(h.row * nx + dj) / mesh;
            zl[c] = 1.125 * (double)(h.plane * nx + dk) / mesh;
        }
        h.volo0 = host_elem_volume(xl, yl, zl);
    }

    // initial energy and dt0, lulesh-init.cc:170-180
    const double ebase = 3.948746e+7;
    const double scale = (nx * tp) / 45.0;
    const double einit = ebase * scale * scale * scale;
    h.e0 = (h.row + h.col + h.plane == 0) ? einit : 0.0;
    h.deltatime0 = (.5 * cbrt(h.volo0)) / sqrt(2.0 * einit);
}

void build_host_domain_mesh(HostDomain &h, const lulesh_mesh &m, int nr, int balance, int cost)
{
    h.umesh = true;
    h.nx = 0;
    h.tp = 1;
    h.rank = h.col = h.row = h.plane = 0;
    h.numElem = m.numElem;
    h.numNode = m.numNode;
    h.allElem = m.numElem;   // one domain: no ghost planes, no comm buffers
    create_regions(h, nr, balance, cost);
    build_eval_table(h);

    const int ne = m.numElem, nn = m.numNode;
    h.ux = m.x;
    h.uy = m.y;
    h.uz = m.z;
    h.u_nodelist = m.nodelist;
    h.u_face.assign(8 * (size_t)ne, 0);
    for (int i = 0; i < ne; ++i) {
        for (int f = 0; f < 6; ++f) h.u_face[8 * (size_t)i + f] = m.nbr[6 * (size_t)i + f];
        h.u_face[8 * (size_t)i + 6] = m.elemBC[i];
    }
    // node -> corner lists, elements ascending (the order the reference's
    // corner list has, lulesh-init.cc:243-313)
    h.u_nstart.assign((size_t)nn + 1, 0);
    for (int v : m.nodelist) ++h.u_nstart[(size_t)v + 1];
    for (int n = 0; n < nn; ++n) h.u_nstart[n + 1] += h.u_nstart[n];
    h.u_ncorner.assign(8 * (size_t)ne, 0);
    std::vector<int> fill(h.u_nstart.begin(), h.u_nstart.end() - 1);
    for (int i = 0; i < ne; ++i)
        for (int j = 0; j < 8; ++j) {
            const int n = m.nodelist[8 * (size_t)i + j];
            h.u_ncorner[fill[n]++] = 8 * i + j;
        }
    h.u_nsymm.assign((size_t)nn, 0);
    for (int a = 0; a < 3; ++a)
        for (int n : m.symm[a]) h.u_nsymm[n] |= (unsigned char)(1u << a);

    // element 0's volume, the deposit at the octant corner, and dt0 by the
    // analytic CFL estimate of lulesh.cu:1205 with the deposited energy
    double xl[8], yl[8], zl[8];
    for (int c = 0; c < 8; ++c) {
        const int n = m.nodelist[c];
        xl[c] = m.x[n];
        yl[c] = m.y[n];
        zl[c] = m.z[n];
    }
    h.volo0 = host_elem_volume(xl, yl, zl);
    h.e0 = m.e_deposit;
    h.e0_index = m.octantCorner;
    h.deltatime0 = (.5 * cbrt(h.volo0)) / sqrt(2.0 * m.e_deposit);
}

size_t domain_device_bytes(int nx, bool corner_array, long long evtotal)
{
    const size_t ne = (size_t)nx * nx * nx, nn = (size_t)(nx + 1) * (nx + 1) * (nx + 1);
    const size_t all = ne + 6 * (size_t)nx * nx;
    // a one-subdomain problem: 16 node arrays (two node-state buffers, f,
    // nodalMass), 12 element fields, 3 ghosted gradients, 3 plain
    // gradients, [48 corner components], elemRep, the EOS work table;
    // 256-byte alignment slack
    return 8 * (16 * nn + 12 * ne + 3 * all + 3 * ne + (corner_array ? 48 * ne : 0)) + 4 * ne +
           2 * (size_t)evtotal + 64 * 256;
}

}  // namespace lulesh
