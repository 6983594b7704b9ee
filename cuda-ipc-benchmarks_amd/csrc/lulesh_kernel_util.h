// lulesh_kernel_util.h -- device helpers shared by the kernel translation
// units (lulesh_kernels.hip and lulesh_force_march.hip, which is built with
// its own instruction-scheduling strategy).
#pragma once
#include <hip/hip_runtime.h>

#include "lulesh_dev.h"
#include "lulesh_kernels.h"
#include "lulesh_math.h"

namespace lulesh {

// Consecutive workgroups are dealt round-robin to the 8 XCDs; give each XCD
// a contiguous run of tiles so neighbouring elements share an L2.
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    const int q = nb >> 3, r = nb & 7;
    const int xcd = b & 7, slot = b >> 3;
    return xcd < r ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
}

__device__ __forceinline__ bool group_stopped(const GroupCtl &G)
{
    return (G.done | G.status) != 0;
}

// node state at the start of the running cycle and the buffer the cycle's
// node update writes (the same buffer when the subdomain does not
// double-buffer its node state)
__device__ __forceinline__ NodeState ns_old(const DomDev &D) { return D.old; }
__device__ __forceinline__ NodeState ns_new(const DomDev &D) { return D.cur; }

__device__ __forceinline__ void raise_status(GroupCtl *G, int code)
{
    atomicCAS(&G->status, 0, code);
}

__device__ __forceinline__ unsigned long long dbits(double v)
{
    return (unsigned long long)__double_as_longlong(v);
}

// TimeIncrement, lulesh.cc:207-263.  A subdomain's candidate from its
// Courant and hydro minima of the last cycle (before the group's MIN, the
// reference's MPI_Allreduce, lulesh.cc:227-236)...
__device__ __forceinline__ double ti_candidate(unsigned long long dcb, unsigned long long dhb)
{
    double gnewdt = 1.0e+20;
    const double dc = __longlong_as_double((long long)dcb);
    const double dh = __longlong_as_double((long long)dhb);
    if (dc < gnewdt) gnewdt = dc / 2.0;
    if (dh < gnewdt) gnewdt = dh * 2.0 / 3.0;
    return gnewdt;
}

// ... and, given the group's MIN newdt, a subdomain's new deltatime (ratio
// clamps against its own olddt, dtmax, the approach to the stop time)
__device__ __forceinline__ double ti_deltatime(double newdt, double olddt, double time, int cyc,
                                               double stoptime)
{
    double deltatime = olddt;
    double targetdt = stoptime - time;
    if ((Phys::dtfixed <= 0.0) && (cyc != 0)) {
        double dt = newdt;
        const double ratio = dt / olddt;
        if (ratio >= 1.0) {
            if (ratio < Phys::deltatimemultlb) dt = olddt;
            else if (ratio > Phys::deltatimemultub) dt = olddt * Phys::deltatimemultub;
        }
        if (dt > Phys::dtmax) dt = Phys::dtmax;
        deltatime = dt;
    }
    if ((targetdt > deltatime) && (targetdt < (4.0 * deltatime / 3.0)))
        targetdt = 2.0 * deltatime / 3.0;
    if (targetdt < deltatime) deltatime = targetdt;
    return deltatime;
}

// main's loop test (lulesh.cc:3222) on the group's subdomain 0
__device__ __forceinline__ bool ti_stop(double time, int cyc, double stoptime, int max_cycles)
{
    return !(time < stoptime) || (max_cycles >= 0 && cyc >= max_cycles);
}

constexpr int kMThreads = 256;   // threads (element slots) per column tile
constexpr int kPatch = kForcePatch;   // node slots per plane: (ex+1)*(ey+1)


// FB hourglass corner forces (CalcFBHourglassForceForElems,
// lulesh.cc:906-1056) of one velocity component at a time
struct HgModes {
    double g[8][4];
    double coefficient;
};

__device__ __forceinline__ void hg_modes(const double x[8], const double y[8], const double z[8],
                                         double determ, double coefficient, HgModes &H)
{
    // the four modes' coordinate sums first, so x..z can die before the
    // volume derivatives and the mode vectors are live together
    // the twelve sums advance together (corner by corner), so that no add
    // waits on the one just issued; each is still summed left to right
    double hmx[4], hmy[4], hmz[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        hmx[m] = x[0] * hg_gamma(m, 0);
        hmy[m] = y[0] * hg_gamma(m, 0);
        hmz[m] = z[0] * hg_gamma(m, 0);
    }
#pragma unroll
    for (int c = 1; c < 8; ++c)
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            hmx[m] = hmx[m] + x[c] * hg_gamma(m, c);
            hmy[m] = hmy[m] + y[c] * hg_gamma(m, c);
            hmz[m] = hmz[m] + z[c] * hg_gamma(m, c);
        }
    double dvdx[8], dvdy[8], dvdz[8];
    volume_derivs(x, y, z, dvdx, dvdy, dvdz);
    const double volinv = rcp_pos(determ);   // volo * v: positive unless the volume error is raised
#pragma unroll
    for (int c = 0; c < 8; ++c)
#pragma unroll
        for (int m = 0; m < 4; ++m)
            H.g[c][m] = hg_gamma(m, c) - volinv * (dvdx[c] * hmx[m] + dvdy[c] * hmy[m] + dvdz[c] * hmz[m]);
    H.coefficient = coefficient;
}

// coefficient = - hourg * 0.01 * ss * mass / cbrt(volo * v), lulesh.cc:1012-1052
__device__ __forceinline__ double hg_coefficient(double ss, double mass, double volo, double v)
{
    const double volume13 = cbrt_cr(volo * v);
    return -Phys::hgcoef * 0.01 * ss * mass / volume13;
}

__device__ __forceinline__ void hg_force(const HgModes &H, const double vl[8], double out[8])
{
    // the four mode sums, then the eight corner sums, each advanced in
    // lockstep with its siblings (the same left-to-right sums)
    double h[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) h[m] = H.g[0][m] * vl[0];
#pragma unroll
    for (int c = 1; c < 8; ++c)
#pragma unroll
        for (int m = 0; m < 4; ++m) h[m] = h[m] + H.g[c][m] * vl[c];
    double acc[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[c] = H.g[c][0] * h[0];
#pragma unroll
    for (int m = 1; m < 4; ++m)
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[c] = acc[c] + H.g[c][m] * h[m];
#pragma unroll
    for (int c = 0; c < 8; ++c) out[c] = H.coefficient * acc[c];
}

}  // namespace lulesh
