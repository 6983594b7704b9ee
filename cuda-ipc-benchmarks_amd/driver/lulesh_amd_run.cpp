// lulesh_amd_run -- command-line driver with the reference's options and
// final output (LULESH/omp_4.0/lulesh.cc:3089-3275, lulesh-util.cc:60-224),
// running the Lagrange step on the GPU through the C ABI.
//
//   -s <size>   elements per edge of each subdomain   (default 30)
//   -u <file>   an unstructured mesh in the .lmesh format instead of the
//               block (the CUDA reference's -u, LULESH/cuda/src/lulesh.cu:
//               4754-4770; one domain)
//   -i <its>    cycle cap                              (default none)
//   -r <nr>     regions                                (default 11)
//   -b <bal>    region balance                         (default 1)
//   -c <cost>   region cost                            (default 1)
//   -t <tp>     subdomains per edge; tp^3 subdomains share the GPU (default 1)
//   -p          print progress every cycle (one cycle per launch, as the
//               reference's loop prints, lulesh.cc:3227-3230)
//   -q          quiet
//   -z <n>      stop after n cycles of this run (the reference's iteration
//               cap, lulesh.cc:3231-3234; 0 = none)
//   -f <n>      accepted for the reference's viz-file count; no viz output
//   -v          refused, as in a reference build without VIZ_MESH: the
//               Silo output needs a library this build does not have
//   -h          usage
//   -w <0|1>    write subdomain 0's final coordinates to xyz.asc, as the
//               CUDA reference's write_solution does (default 1)
//
// After the run it prints the reference's PHASE_* lifecycle lines
// (LULESH/cuda/src/lulesh.cu:4918-4941).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>

#include "lulesh_amd.h"

static int parse_int(const char *s, int *out)
{
    char *end = nullptr;
    const long v = strtol(s, &end, 10);
    if (end == s || (*end != '\0' && *end != ' ')) return 0;
    *out = (int)v;
    return 1;
}

static void usage(const char *exe)
{
    // PrintCommandLineOptions, lulesh-util.cc:30-48, with this driver's own
    // options after the reference's
    printf("Usage: %s [opts]\n", exe);
    printf(" where [opts] is one or more of:\n");
    printf(" -q              : quiet mode - suppress all stdout\n");
    printf(" -i <iterations> : number of cycles to run\n");
    printf(" -s <size>       : length of cube mesh along side\n");
    printf(" -r <numregions> : Number of distinct regions (def: 11)\n");
    printf(" -b <balance>    : Load balance between regions of a domain (def: 1)\n");
    printf(" -c <cost>       : Extra cost of more expensive regions (def: 1)\n");
    printf(" -f <numfiles>   : Number of files to split viz dump into (no viz output here)\n");
    printf(" -p              : Print out progress\n");
    printf(" -v              : Output viz file (not available: needs the Silo library)\n");
    printf(" -z <cycles>     : Stop after this many cycles of the run\n");
    printf(" -h              : This message\n");
    printf(" -u <file>       : Unstructured mesh (.lmesh) instead of the cube\n");
    printf(" -t <tp>         : tp^3 subdomains on the one GPU (def: 1)\n");
    printf(" -w <0|1>        : Write xyz.asc (def: 1)\n");
    printf("\n\n");
}

int main(int argc, char **argv)
{
    const auto t_app = std::chrono::steady_clock::now();
    int its = -1, nx = 30, nr = 11, bal = 1, cost = 1, tp = 1, prog = 0, quiet = 0, wsol = 1;
    int zcap = 0, nfiles = 0;
    const char *mesh_path = nullptr;
    for (int i = 1; i < argc; ++i) {
        int *dst = nullptr;
        if (!strcmp(argv[i], "-h")) {
            usage(argv[0]);
            return 0;
        }
        if (!strcmp(argv[i], "-v")) {
            printf("Use of -v requires the Silo library, which this build does not have\n");
            return 1;
        }
        if (!strcmp(argv[i], "-u")) {
            if (i + 1 >= argc) {
                fprintf(stderr, "file name required after -u\n");
                return 1;
            }
            mesh_path = argv[++i];
            continue;
        }
        if (!strcmp(argv[i], "-s")) dst = &nx;
        else if (!strcmp(argv[i], "-i")) dst = &its;
        else if (!strcmp(argv[i], "-r")) dst = &nr;
        else if (!strcmp(argv[i], "-b")) dst = &bal;
        else if (!strcmp(argv[i], "-c")) dst = &cost;
        else if (!strcmp(argv[i], "-t")) dst = &tp;
        else if (!strcmp(argv[i], "-w")) dst = &wsol;
        else if (!strcmp(argv[i], "-z")) dst = &zcap;
        else if (!strcmp(argv[i], "-f")) dst = &nfiles;
        else if (!strcmp(argv[i], "-p")) { prog = 1; continue; }
        else if (!strcmp(argv[i], "-q")) { quiet = 1; continue; }
        else {
            usage(argv[0]);
            printf("ERROR: Unknown command line argument: %s\n", argv[i]);
            return 1;
        }
        if (i + 1 >= argc || !parse_int(argv[i + 1], dst)) {
            fprintf(stderr, "integer value required after %s\n", argv[i]);
            return 1;
        }
        ++i;
    }
    (void)nfiles;   // viz files: no viz output is written
    int rc = lulesh_init(0);
    if (rc) { fprintf(stderr, "lulesh_init: %s\n", lulesh_status_string(rc)); return 1; }
    lulesh_group *g = nullptr;
    long long zones = (long long)nx * nx * nx;   // per subdomain
    if (mesh_path) {
        lulesh_mesh *m = nullptr;
        rc = lulesh_mesh_read(mesh_path, &m);
        if (rc) { fprintf(stderr, "could not read %s: %s\n", mesh_path, lulesh_status_string(rc)); return 1; }
        int ne = 0;
        lulesh_mesh_sizes(m, &ne, nullptr);
        zones = ne;
        tp = 1;
        rc = lulesh_group_create_mesh(m, nr, bal, cost, &g);
        lulesh_mesh_destroy(m);
    } else {
        rc = lulesh_group_create(nx, tp, nr, bal, cost, &g);
    }
    if (rc) { fprintf(stderr, "create: %s\n", lulesh_status_string(rc)); return 1; }
    // the run starts at cycle 0, so an iteration cap of z cycles is a cycle
    // cap of z (the reference's loop breaks after its z-th iteration)
    if (zcap > 0) its = its < 0 ? zcap : (its < zcap ? its : zcap);
    if (its >= 0) lulesh_group_set_max_cycles(g, its);
    const int ndom = tp * tp * tp;
    if (!quiet) {
        if (mesh_path) {
            // lulesh.cu:4842-4846
            printf("Running until t=%f, Problem size=%lld \n", 1.0e-2, zones);
        } else {
            printf("Running problem size %d^3 per domain until completion\n", nx);
            printf("Num subdomains: %d (on one GPU)\n", ndom);
            printf("Total number of elements: %lld\n\n", (long long)ndom * zones);
        }
    }
    const auto t0 = std::chrono::steady_clock::now();
    int cycle = 0;
    double time = 0, dt = 0;
    for (;;) {
        rc = lulesh_group_enqueue(g, prog ? 1 : 64, nullptr);
        if (rc) break;
        const int prev = cycle;
        rc = lulesh_group_sync(g, &cycle, &time, &dt, nullptr);
        if (rc) break;
        if (cycle == prev) break;   // stop condition reached on the device
        if (prog && !quiet) printf("cycle = %d, time = %e, dt=%e\n", cycle, time, dt);
    }
    const double elapsed =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc) {
        fprintf(stderr, "run stopped: %s\n", lulesh_status_string(rc));
        lulesh_group_destroy(g);
        return rc == LULESH_VOLUME_ERROR || rc == LULESH_QSTOP_ERROR ? -rc : 1;
    }
    double v[4];
    lulesh_group_verify(g, 0, v);
    if (!quiet) {
        // VerifyAndWriteFinalOutput, lulesh-util.cc:177-224 (grind time per
        // element of the mesh on the unstructured path, lulesh.cu:4689-4697)
        const double grind1 = ((elapsed * 1e6) / cycle) / (double)zones;
        const double grind2 = grind1 / ndom;
        if (mesh_path) {
            // the reference prints no verification block here; the cycle
            // count and the octant corner's energy are this driver's
            printf("Run completed (unstructured mesh):  \n");
            printf("   Problem size        =  %lld elements\n", zones);
            printf("   Iteration count     =  %i \n", cycle);
            printf("   Final Origin Energy = %12.6e   (octant corner)\n", v[0]);
        } else {
            printf("Run completed:  \n");
            printf("   Problem size        =  %i \n", nx);
            printf("   MPI tasks           =  %i \n", ndom);
            printf("   Iteration count     =  %i \n", cycle);
            printf("   Final Origin Energy = %12.6e \n", v[0]);
            printf("   Testing Plane 0 of Energy Array on rank 0:\n");
            printf("        MaxAbsDiff   = %12.6e\n", v[1]);
            printf("        TotalAbsDiff = %12.6e\n", v[2]);
            printf("        MaxRelDiff   = %12.6e\n\n", v[3]);
        }
        printf("\nElapsed time         = %10.2f (s)\n", elapsed);
        printf("Grind time (us/z/c)  = %10.8g (per dom)  (%10.8g overall)\n", grind1, grind2);
        printf("FOM                  = %10.8g (z/s)\n\n", 1000.0 / grind2);
        printf("FullPrecision OriginEnergy = %.17e\n", v[0]);
    }
    if (wsol) {
        lulesh_domain *d0 = nullptr;
        if (lulesh_group_domain(g, 0, &d0) == 0) {
            rc = lulesh_domain_write_solution(d0, "xyz.asc");
            if (rc) fprintf(stderr, "write_solution: %s\n", lulesh_status_string(rc));
        }
    }
    rc = lulesh_group_comm_teardown(g);
    if (rc) fprintf(stderr, "comm_teardown: %s\n", lulesh_status_string(rc));
    double ph[4] = {0, 0, 0, 0};
    lulesh_group_phase_times(g, ph);
    lulesh_group_destroy(g);
    const double app_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_app).count();
    printf("PHASE_COMM_SETUP_ms   %.4f\n", ph[0]);
    printf("PHASE_WIN_ALLOCATE_ms %.4f\n", ph[1]);
    printf("PHASE_PEER_QUERY_ms   %.4f\n", ph[2]);
    printf("PHASE_COMM_FREE_ms    %.4f\n", ph[3]);
    printf("PHASE_APP_TOTAL_ms    %.4f\n", app_ms);
    fflush(stdout);
    return 0;
}
