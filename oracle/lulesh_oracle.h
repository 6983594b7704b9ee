/*
 * lulesh_oracle.h -- CPU oracle for the LULESH Lagrange leapfrog step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a sequential, deterministic
 * restatement of the OpenMP-4.0 LULESH 2.0 arithmetic
 * (LULESH/omp_4.0/lulesh.cc, lulesh-init.cc, lulesh-comm.cc) used as the
 * checker for the HIP implementation.  Only tests/, the smoke check and the
 * CPU-baseline leg of bench.py may load it; the product library never links
 * or calls it.
 *
 * One "group" holds tp^3 subdomains of nx^3 elements each (tp = 1 is the
 * plain single-domain problem).  Subdomain exchanges follow the reference's
 * MPI message order (lulesh-comm.cc) through an in-process loopback, so an
 * 8-subdomain group reproduces an 8-rank run without MPI.
 */
#ifndef LULESH_ORACLE_H
#define LULESH_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lo_group lo_group;

/* cbrt used by the per-cycle hourglass coefficient
 * (LULESH/omp_4.0/lulesh.cc:1014).  LO_CBRT_LIBM calls the C library
 * exactly as the reference does; LO_CBRT_EXACT uses a correctly rounded
 * cube root (the value the HIP kernels compute), so the GPU path can be
 * compared bit for bit.  Set-up (dt0, lulesh-init.cc:180) always uses libm,
 * as the HIP library's host-side set-up does. */
enum { LO_CBRT_LIBM = 0, LO_CBRT_EXACT = 1 };

/* Status codes: 0 ok, -1 VolumeError, -2 QStopError (lulesh.h:22-26 of
 * omp_4.0, enum VolumeError/QStopError).  The reference exits; the oracle
 * returns the code and stops advancing. */

lo_group *lo_create(int nx, int tp, int nr, int balance, int cost, int cbrt_mode);
void      lo_destroy(lo_group *g);

/* One unstructured domain: the reference's -u <file.lmesh> input
 * (LULESH/cuda/src/lulesh.cu:855-1052, single domain only).  Arrays are the
 * file's contents: coordinates [numNode], nodelist [numElem][8], face
 * neighbours [numElem][6] (lxim lxip letam letap lzetam lzetap), the X/Y/Z
 * symmetry node sets and the free-surface node set.  elemBC is derived as
 * the reference's reader derives it unless `elemBC` is given; the element
 * found as the octant corner receives e_deposit (3.948746e+7 in the
 * reference).  Returns NULL on an index out of range. */
lo_group *lo_create_mesh(int numElem, int numNode, const double *x, const double *y, const double *z,
                         const int *nodelist, const int *nbr, const int *symmX, int nSymmX,
                         const int *symmY, int nSymmY, const int *symmZ, int nSymmZ,
                         const int *freeSurf, int nFree, const int *elemBC, double e_deposit,
                         int nr, int balance, int cost, int cbrt_mode);
int       lo_get_elem_bc(const lo_group *g, int dom, int *out);
int       lo_octant_corner(const lo_group *g);
int       lo_all_elem(const lo_group *g);     /* per subdomain, with ghost planes */

/* Advance until time >= stoptime or `max_cycles` more cycles were taken
 * (max_cycles < 0: no cap).  Returns the status code. */
int    lo_run(lo_group *g, int max_cycles);

int    lo_cycle(const lo_group *g);
double lo_time(const lo_group *g);
double lo_deltatime(const lo_group *g);
double lo_stoptime(const lo_group *g);
int    lo_num_elem(const lo_group *g);   /* per subdomain */
int    lo_num_node(const lo_group *g);   /* per subdomain */
int    lo_num_domains(const lo_group *g);

/* Field ids shared with include/lulesh_amd.h (enum lulesh_field). */
int    lo_get_field(const lo_group *g, int dom, int field, double *out);
int    lo_set_field(lo_group *g, int dom, int field, const double *in);
int    lo_get_elem_rep(const lo_group *g, int dom, int *out);
int    lo_get_region(const lo_group *g, int dom, int *out); /* region number-1 per element */
void   lo_set_qstop(lo_group *g, double qstop);
void   lo_set_stoptime(lo_group *g, double t);

/* Final-output figures of VerifyAndWriteFinalOutput
 * (LULESH/omp_4.0/lulesh-util.cc:177-224) for subdomain `dom`.  For an
 * unstructured mesh: the energy of the octant-corner element, and zeros for
 * the plane-0 symmetry figures, which the reference does not compute there. */
void   lo_verify(const lo_group *g, int dom, double *origin_energy,
                 double *max_abs_diff, double *total_abs_diff,
                 double *max_rel_diff);

/* Halo-exchange building blocks, exposed for bit-exact halo tests:
 * lo_exchange runs pack (CommSend) for every subdomain, then the matching
 * unpack (CommSBN / CommSyncPosVel / CommMonoQ) for every subdomain, on the
 * fields named.  msg_type is 1024 (SBN, add), 2048 (PosVel, copy) or
 * 3072 (MonoQ, ghost copy). */
int    lo_exchange(lo_group *g, int msg_type, int nfields, const int *fields);
/* Receive-buffer layout (lulesh-init.cc:317-372; comm_backend.h:75-138). */
int    lo_com_buf_size(const lo_group *g, int dom);
int    lo_recv_offset(const lo_group *g, int recv_dom, int dcol, int drow,
                      int dplane, int xfer_fields, int do_recv, int plane_only);
/* Copy of the receive buffer after the last exchange's pack phase. */
int    lo_get_recv_buffer(const lo_group *g, int dom, double *out);

/* Correctly rounded cube root (oracle's own implementation, checked against
 * exact integer arithmetic in tests). */
double lo_cbrt_exact(double x);

#ifdef __cplusplus
}
#endif
#endif
