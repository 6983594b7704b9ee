// lulesh_plan.cpp -- host-only planning of the force march's column tiles:
// the tile shape per mesh, the z-runs of each column (even split or the
// paired plan) and their launch order.  No HIP call: the engine passes the
// device's CU count, and lulesh_force_plan_of answers the same question
// without a GPU.
#include <math.h>
#include <stdlib.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "lulesh_amd.h"
#include "lulesh_plan.h"

namespace lulesh {

// Column-tile shape for a mesh of nx elements per edge: ex x ey element
// slots (<= 256) covering (ex-1) x (ey-1) nodes each, with an
// (ex+1) x (ey+1) node patch per plane in LDS.  The kernel's LDS allows two
// workgroups per CU whatever their wave count, and a workgroup's plane
// step takes about as long with three waves as with four, so the cost is
// the number of tiles: pick the fewest tiles per node plane, then the
// fewest busy waves, then the fewest element slots, then the wider x
// extent (longer contiguous rows in each plane's global loads).
void tile_shape(int nx, int &ex, int &ey)
{
    const int s1 = nx + 1;
    long best_t = -1, best_w = -1, best_s = -1;
    ex = ey = 16;
    for (int a = 4; a <= 64; ++a)
        for (int b = 4; b <= 64 && a * b <= 256; ++b) {
            if ((a + 1) * (b + 1) > kForcePatch) continue;   // node patch held in LDS
            const long tiles = (long)((s1 + a - 2) / (a - 1)) * ((s1 + b - 2) / (b - 1));
            const long waves = tiles * ((a * b + 63) / 64);   // busy waves per plane
            const long slots = tiles * a * b;
            if (best_t < 0 || tiles < best_t || (tiles == best_t && waves < best_w) ||
                (tiles == best_t && waves == best_w && slots < best_s) ||
                (tiles == best_t && waves == best_w && slots == best_s && a > ex)) {
                best_t = tiles;
                best_w = waves;
                best_s = slots;
                ex = a;
                ey = b;
            }
        }
}

// A column's s1 node planes are split into runs of about kc planes, either
// evenly (max(1, s1 / kc) runs of kc or kc + 1 planes, the longer first:
// no short remainder run) or as runs of kc with the remainder last.
struct ZRuns {
    int n, len, extra;   // runs; planes per run; see planes()
    bool even;
    int planes(int r) const
    {
        if (even) return len + (r < extra ? 1 : 0);
        return r == n - 1 ? extra : len;
    }
};
static ZRuns z_runs(int s1, int kc, bool even)
{
    kc = std::max(kc, 1);
    if (even) {
        const int n = std::max(1, s1 / kc);
        return ZRuns{n, s1 / n, s1 % n, true};
    }
    const int n = (s1 + kc - 1) / kc;
    return ZRuns{n, kc, s1 - kc * (n - 1), false};
}

// Model of the force kernel's time for one z-chunking, in plane steps.  A
// workgroup over a run of k node planes takes about k + 2 plane steps (k + 1
// element planes and its start-up); the kernel runs two workgroups per CU,
// so with T workgroups it takes about ceil(T / (2 CUs)) rounds of the
// longest run, and a workgroup alone on its CU (T <= CUs) runs its steps
// about 0.72 times as long.  Ties go to more workgroups (the dispatcher
// balances a longer queue better), then to the even split.  Fitted to the
// force kernel's times on MI355X over every split tried at -s 32..200 (for
// each size the model's pick is within 0.4 % of the fastest measured;
// -s 48 runs 29.0 us with 204 workgroups of 3 planes against 32.5 us with
// 300 of 2, -s 100 100 us with 480 of 10-11 against 110 us with 1008 of 5).
struct ZCost {
    double time = 0.0;
    long tiles = 0;
};
// a subdomain's node planes, tile shape and node columns, computed once per
// plan (the chunk candidates below all reuse them)
struct DomShape {
    int s1, ex, ey;
    long cols;
};
static std::vector<DomShape> dom_shapes(const ForcePlanInput &in)
{
    std::vector<DomShape> v;
    for (const auto &g : in.group_nx)
        for (const int nx : g) {
            DomShape d;
            d.s1 = nx + 1;
            tile_shape(nx, d.ex, d.ey);
            d.cols = (long)((d.s1 + d.ex - 2) / (d.ex - 1)) * ((d.s1 + d.ey - 2) / (d.ey - 1));
            v.push_back(d);
        }
    return v;
}
static ZCost z_cost(const ForcePlanInput &in, const std::vector<DomShape> &shapes, int kc, bool even)
{
    long tiles = 0;
    int longest = 0;
    for (const DomShape &d : shapes) {
        const ZRuns z = z_runs(d.s1, kc, even);
        tiles += d.cols * z.n;
        longest = std::max(longest, std::max(z.planes(0), z.planes(z.n - 1)));
    }
    const long slots = 2L * in.cus;
    ZCost c;
    c.time = (double)((tiles + slots - 1) / slots) * (longest + 2) * (tiles <= in.cus ? 0.72 : 1.0);
    c.tiles = tiles;
    return c;
}

// The z-chunking of the engine's column tiles: the kchunk given (0 = every
// kc in [1, 32]) with the better split by the model above.  An engine of
// several groups (a batch, a server) mixes run lengths, which the model's
// rounds do not describe: there kc minimises rounds x (kc + 1) over runs of
// kc planes, and each group's columns are split evenly into runs of kc or
// kc + 1 (measured on MI355X: served streams 5.15 with the single-group
// model, 5.25 with runs of kc and the remainder last, 5.28e9 zone-cycles/s
// with even runs; the 64-problem ensemble 4.95 -> 5.04e9 at kc = 8); the
// batch autotuner times its kc candidates on top.
static void choose_zruns(const ForcePlanInput &in, const std::vector<DomShape> &shapes, int &kc_out,
                         bool &even_out)
{
    const int lo = in.kchunk > 0 ? in.kchunk : 1, hi = in.kchunk > 0 ? in.kchunk : 32;
    if (in.group_nx.size() > 1) {
        const long slots = 2L * in.cus;
        long best_cost = -1;
        kc_out = lo;
        even_out = true;
        for (int kc = lo; kc <= hi; ++kc) {
            long tiles = 0;
            for (const DomShape &d : shapes) tiles += d.cols * ((d.s1 + kc - 1) / kc);
            const long cost = ((tiles + slots - 1) / slots) * (kc + 1);
            if (best_cost < 0 || cost < best_cost) {
                best_cost = cost;
                kc_out = kc;
            }
        }
        return;
    }
    bool have = false;
    ZCost best;
    for (int kc = lo; kc <= hi; ++kc)
        for (int ev = 1; ev >= 0; --ev) {
            const ZCost c = z_cost(in, shapes, kc, ev != 0);
            const bool better = !have || c.time < best.time - 1e-9 ||
                                (c.time < best.time + 1e-9 && c.tiles > best.tiles);
            if (better) {
                have = true;
                best = c;
                kc_out = kc;
                even_out = ev != 0;
            }
        }
}

// ---------------------------------------------------------------------------
// Paired z-runs: one launch round balanced for the two workgroups of a CU.
//
// With at most two workgroups per CU, the dispatcher gives blocks 0..CUs-1
// one CU each and block b + CUs the CU of block b.  The two do not progress
// alike: the older workgroup's waves win the issue arbitration whenever
// both are ready, so it runs almost as fast as alone while the younger
// takes the cycles it leaves.  Measured per workgroup on MI355X at -s 100
// (wall-clock stamps at each workgroup's start and end, runs of 10-11
// planes): older 87.5 us, younger 107 us, alone 80 us -- relative rates
// 1.0, 0.756 and 1.094, and the launch lasts as long as the younger ones.
// So a round is balanced by giving the older slots longer runs: for a
// target time T, a slot of rate v takes a run of at most floor(T v) - 2
// node planes (a run of k planes costs about k + 2 plane steps: k + 1
// element planes and its start-up); the columns are filled bottom-up over
// the slots in launch order, each column's planes are then spread over its
// slots in proportion to their rates, and T is the least for which every
// column fits.  The number of workgroups is chosen by the pair model below
// (a CU's finish time with its two runs), against the even split.
// ---------------------------------------------------------------------------
static const double kRateOlder = 1.0, kRateYounger = 0.756, kRateAlone = 1.094;
static const double kRunStartup = 2.0;

// host mirror of the kernels' blockIdx -> tile index map (xcd_tile)
static int xcd_tile_host(int b, int nb)
{
    const int q = nb >> 3, r = nb & 7;
    const int xcd = b & 7, slot = b >> 3;
    return xcd < r ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
}

static double slot_rate(int b, int nb, int cus)
{
    if (b >= cus) return kRateYounger;
    return b + cus < nb ? kRateOlder : kRateAlone;
}

// finish time (plane steps) of a CU whose older workgroup has a run of cost
// ca and whose younger one cost cb (< 0: none): both at their paired rates
// until one ends, the other alone after that
static double pair_time(double ca, double cb)
{
    if (cb < 0.0) return ca / kRateAlone;
    const double ta = ca / kRateOlder, tb = cb / kRateYounger;
    if (ta <= tb) return ta + (cb - kRateYounger * ta) / kRateAlone;
    return tb + (ca - kRateOlder * tb) / kRateAlone;
}

struct ColRef {
    int dom, i0, j0, ex, ey, planes;
};
struct PairedPlan {
    double time = -1.0;
    std::vector<ColTile> tiles;   // in tile-index order
};

// Slots handed to the columns for target T: the slots in launch order
// (tile-index order, each XCD's chunk in turn), the columns in round-robin
// order, so that each XCD works on a z-slab of every column at once (the
// column tiles' halo rows and columns then come from its own L2), each
// slot taking up to floor(T v) - 2 planes of its column.  Records each
// slot's column when col_of is given.  Returns the slots used, or -1 when
// the columns do not fit in nb.
static int paired_assign(const std::vector<ColRef> &cols, int nb, int cus, double T, const std::vector<int> &blk,
                         std::vector<int> *col_of)
{
    const int nc = (int)cols.size();
    // the three capacities (slot classes: older, younger, alone)
    const int cap_o = (int)floor(T * kRateOlder - kRunStartup);
    const int cap_y = (int)floor(T * kRateYounger - kRunStartup);
    const int cap_a = (int)floor(T * kRateAlone - kRunStartup);
    thread_local std::vector<int> rem;
    rem.resize(nc);
    for (int c = 0; c < nc; ++c) rem[c] = cols[c].planes;
    if (col_of) col_of->assign(nb, -1);
    int t = 0, active = nc, ci = 0;
    while (active > 0) {
        while (rem[ci] == 0) ci = ci + 1 == nc ? 0 : ci + 1;
        if (t >= nb) return -1;
        const int b = blk[t];
        const int cap = b >= cus ? cap_y : (b + cus < nb ? cap_o : cap_a);
        if (cap < 1) return -1;
        const int take = std::min(cap, rem[ci]);
        rem[ci] -= take;
        if (rem[ci] == 0) --active;
        if (col_of) (*col_of)[t] = ci;
        ++t;
        ci = ci + 1 == nc ? 0 : ci + 1;
    }
    return t;
}

// The pair-model time of the runs for nb workgroups and target T (-1 when
// the columns do not fit or leave a slot unused), and the tiles when asked:
// each column's planes are spread over its slots in proportion to their
// rates (equal finish times), at least one plane each, bottom-up in launch
// order.
static double paired_fill(const std::vector<ColRef> &cols, int nb, int cus, double T, const std::vector<int> &blk,
                          std::vector<ColTile> *tiles)
{
    thread_local std::vector<int> col_of, start, order, len;
    thread_local std::vector<double> frac, cost;
    if (paired_assign(cols, nb, cus, T, blk, &col_of) != nb) return -1.0;
    const int nc = (int)cols.size();
    // each column's slots, in launch order (counting sort by column)
    start.assign(nc + 1, 0);
    for (int t = 0; t < nb; ++t) ++start[col_of[t] + 1];
    for (int c = 0; c < nc; ++c) start[c + 1] += start[c];
    order.resize(nb);
    {
        thread_local std::vector<int> pos;
        pos.assign(start.begin(), start.end() - 1);
        for (int t = 0; t < nb; ++t) order[pos[col_of[t]]++] = t;
    }
    len.assign(nb, 0);
    frac.assign(nb, 0.0);
    for (int c = 0; c < nc; ++c) {
        const int f = start[c], m = start[c + 1] - f, planes = cols[c].planes;
        double vsum = 0.0;
        for (int s = f; s < f + m; ++s) vsum += slot_rate(blk[order[s]], nb, cus);
        const double tau = (planes + m * kRunStartup) / vsum;
        int tot = 0;
        for (int s = f; s < f + m; ++s) {
            const double want = std::max(1.0, tau * slot_rate(blk[order[s]], nb, cus) - kRunStartup);
            len[s] = std::max(1, (int)floor(want));
            frac[s] = want - floor(want);
            tot += len[s];
        }
        while (tot < planes) {   // largest remainders first
            int bi = f;
            for (int s = f + 1; s < f + m; ++s)
                if (frac[s] > frac[bi]) bi = s;
            ++len[bi];
            frac[bi] = -1.0;
            ++tot;
        }
        while (tot > planes) {   // trim the longest runs
            int bi = f;
            for (int s = f + 1; s < f + m; ++s)
                if (len[s] > len[bi]) bi = s;
            if (len[bi] <= 1) return -1.0;
            --len[bi];
            --tot;
        }
    }
    cost.assign(nb, -1.0);
    for (int s = 0; s < nb; ++s) cost[blk[order[s]]] = len[s] + kRunStartup;
    double worst = 0.0;
    for (int b = 0; b < std::min(nb, cus); ++b)
        worst = std::max(worst, pair_time(cost[b], b + cus < nb ? cost[b + cus] : -1.0));
    if (tiles) {
        tiles->assign(nb, ColTile{});
        for (int c = 0; c < nc; ++c) {
            const ColRef &cr = cols[c];
            for (int s = start[c], k0 = 0; s < start[c + 1]; k0 += len[s], ++s)
                (*tiles)[order[s]] = ColTile{cr.dom, cr.i0, cr.j0, k0, cr.ex, cr.ey,
                                             (k0 | cr.j0 | cr.i0) == 0 ? kFirstTile : 0, len[s]};
        }
    }
    return worst;
}

// The least T for which the columns fit in nb slots (bisection; a column
// alone on the slowest slot bounds it).
static double paired_target(const std::vector<ColRef> &cols, int nb, int cus, int longest,
                            const std::vector<int> &blk)
{
    double lo = 0.0, hi = (longest + kRunStartup) / kRateYounger + 1.0;
    for (int it = 0; it < 16; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (paired_assign(cols, nb, cus, mid, blk, nullptr) >= 0) hi = mid;
        else lo = mid;
    }
    return hi;
}

// The best paired plan over the workgroup counts [columns, 2 CUs], or an
// empty plan when the columns do not fit one round.  Plans are kept per
// column list: a process that builds engines for the same problem shape
// searches once.
static PairedPlan paired_plan(const std::vector<ColRef> &cols, int cus)
{
    PairedPlan best;
    const int nc = (int)cols.size();
    int longest = 0;
    for (const ColRef &c : cols) longest = std::max(longest, c.planes);
    if (nc == 0 || nc > 2 * cus) return best;
    std::vector<int> key{cus};
    for (const ColRef &c : cols) key.insert(key.end(), {c.dom, c.i0, c.j0, c.ex, c.ey, c.planes});
    static std::mutex mu;
    static std::map<std::vector<int>, PairedPlan> cache;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    std::vector<int> blk;
    int best_nb = -1;
    double best_T = 0.0;
    for (int nb = nc; nb <= 2 * cus; ++nb) {
        blk.resize(nb);
        for (int b = 0; b < nb; ++b) blk[xcd_tile_host(b, nb)] = b;
        const double T = paired_target(cols, nb, cus, longest, blk);
        const double t = paired_fill(cols, nb, cus, T, blk, nullptr);
        if (t > 0.0 && (best.time < 0.0 || t < best.time - 1e-9)) {
            best.time = t;
            best_nb = nb;
            best_T = T;
        }
    }
    if (best_nb > 0) {
        blk.resize(best_nb);
        for (int b = 0; b < best_nb; ++b) blk[xcd_tile_host(b, best_nb)] = b;
        paired_fill(cols, best_nb, cus, best_T, blk, &best.tiles);
    }
    std::lock_guard<std::mutex> lk(mu);
    if (cache.size() > 256) cache.clear();
    cache[key] = best;
    return best;
}

// pair-model time of a plan given as tiles in tile-index order (one round)
static double plan_pair_time(const std::vector<ColTile> &tiles, int cus)
{
    const int nb = (int)tiles.size();
    if (nb > 2 * cus) return -1.0;
    std::vector<double> cost(nb, -1.0);
    for (int b = 0; b < nb; ++b) cost[b] = tiles[xcd_tile_host(b, nb)].kn + kRunStartup;
    double worst = 0.0;
    for (int b = 0; b < std::min(nb, cus); ++b)
        worst = std::max(worst, pair_time(cost[b], b + cus < nb ? cost[b + cus] : -1.0));
    return worst;
}

bool paired_runs_wanted()
{
    const char *pr = getenv("LULESH_PAIRED_RUNS");
    return !(pr && atoi(pr) == 0);
}

int plan_force_columns(const ForcePlanInput &in, ForcePlan &out)
{
    out = ForcePlan{};
    if (in.cus < 1) return LULESH_EINVAL;
    const std::vector<DomShape> shapes = dom_shapes(in);
    for (const DomShape &d : shapes)
        if (d.ex * d.ey > 256 || (d.ex + 1) * (d.ey + 1) > kForcePatch) return LULESH_EINVAL;
    int kc = 1;
    bool even = true;
    choose_zruns(in, shapes, kc, even);
    std::vector<ColTile> &ct = out.tiles;
    std::vector<ColRef> cols;
    for (int di = 0; di < (int)shapes.size(); ++di) {
        const int s1 = shapes[di].s1, ex = shapes[di].ex, ey = shapes[di].ey;
        const ZRuns z = z_runs(s1, kc, even);
        for (int r = 0, k0 = 0; r < z.n; k0 += z.planes(r), ++r)
            for (int j0 = 0; j0 < s1; j0 += ey - 1)
                for (int i0 = 0; i0 < s1; i0 += ex - 1)
                    ct.push_back(ColTile{di, i0, j0, k0, ex, ey, (k0 | j0 | i0) == 0 ? kFirstTile : 0,
                                         z.planes(r)});
        for (int j0 = 0; j0 < s1; j0 += ey - 1)
            for (int i0 = 0; i0 < s1; i0 += ex - 1) cols.push_back(ColRef{di, i0, j0, ex, ey, s1});
    }
    out.kchunk_used = kc;
    if (in.paired_allowed && !ct.empty()) {
        const PairedPlan p = paired_plan(cols, in.cus);
        const double t_even = plan_pair_time(ct, in.cus);
        if (p.time > 0.0 && t_even > 0.0 && p.time < t_even - 1e-9) {
            ct = p.tiles;
            out.paired = true;
        }
    }
    return LULESH_OK;
}

}  // namespace lulesh
