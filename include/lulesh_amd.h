/*
 * lulesh_amd.h -- C ABI of the MI355X LULESH Lagrange step and its halo
 * exchange.
 *
 * The library advances LULESH 2.0 Sedov problems (FP64) on one AMD Instinct
 * MI355X (gfx950) with hand-written HIP kernels.  A *group* is one
 * decomposed problem: tp^3 subdomains of nx^3 elements, all resident in the
 * same GPU's HBM; tp = 1 is the plain single-domain problem.  Subdomains
 * exchange ghost planes through the reference's comm-backend interface
 * (CommRecv / CommSend / CommSBN / CommSyncPosVel / CommMonoQ), served by
 * an on-device "device_local" backend: pack kernels write straight into the
 * receiving subdomain's receive buffer at the reference's wire offsets, and
 * unpack kernels apply the messages in the reference's slot order, so the
 * result is bit-reproducible.
 *
 * All entry points return an int status (enum lulesh_status); nothing in
 * the library calls exit().  Plain pointers and sizes only.
 */
#ifndef LULESH_AMD_H
#define LULESH_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Opaque HIP stream (hipStream_t); NULL = the library's own stream. */
typedef void *lulesh_stream_t;

typedef struct lulesh_group  lulesh_group;   /* one decomposed problem        */
typedef struct lulesh_domain lulesh_domain;  /* one subdomain of a group      */
typedef struct lulesh_batch  lulesh_batch;   /* ensemble of independent groups */
typedef struct lulesh_server lulesh_server;  /* continuous-batching server     */
typedef struct lulesh_mesh   lulesh_mesh;    /* host-side unstructured mesh   */

/* Status codes.  VOLUME / QSTOP keep the reference's exit values
 * (LULESH/omp_4.0/lulesh.h:41, enum { VolumeError = -1, QStopError = -2 }). */
enum lulesh_status {
    LULESH_OK           = 0,
    LULESH_VOLUME_ERROR = -1,   /* negative element volume            */
    LULESH_QSTOP_ERROR  = -2,   /* q above qstop                      */
    LULESH_EINVAL       = -3,   /* bad argument                       */
    LULESH_EHIP         = -4,   /* HIP runtime error                  */
    LULESH_ENOMEM       = -5,   /* allocation / HBM budget            */
    LULESH_EDIRECTION   = -6,   /* no message expected from a direction
                                   (shmRecvOffset abort, cuda
                                   comm/comm_backend.h:134-137)      */
    LULESH_EBUSY        = -7,   /* server queue full                  */
    LULESH_EIO          = -8    /* mesh file unreadable or truncated  */
};

/* Field selector; replaces the reference's Domain_member pointer-to-member
 * (LULESH/cuda/src/lulesh.h:300).  Node fields are < 16. */
enum lulesh_field {
    LULESH_X = 0, LULESH_Y = 1, LULESH_Z = 2,
    LULESH_XD = 3, LULESH_YD = 4, LULESH_ZD = 5,
    LULESH_FX = 9, LULESH_FY = 10, LULESH_FZ = 11,
    LULESH_NODAL_MASS = 12,
    LULESH_E = 16, LULESH_P = 17, LULESH_Q = 18,
    LULESH_V = 21, LULESH_VOLO = 22, LULESH_DELV = 23, LULESH_VDOV = 24,
    LULESH_AREALG = 25, LULESH_SS = 26, LULESH_ELEM_MASS = 27,
    LULESH_DELV_XI = 28, LULESH_DELV_ETA = 29, LULESH_DELV_ZETA = 30,
    LULESH_DELX_XI = 31, LULESH_DELX_ETA = 32, LULESH_DELX_ZETA = 33,
    LULESH_VNEW = 34
};

/* Message tags, LULESH/omp_4.0/lulesh.h:92-94. */
#define LULESH_MSG_COMM_SBN     1024
#define LULESH_MSG_SYNC_POS_VEL 2048
#define LULESH_MSG_MONOQ        3072

/* ------------------------------------------------------------------ */
/* library                                                             */
/* ------------------------------------------------------------------ */

/* Selects the device (0 on a one-GPU machine) and checks it is gfx950. */
int lulesh_init(int device);
/* CUs of the current device: what the engine plans its launches for (no
 * counterpart in the reference; see lulesh_force_plan_of). */
int lulesh_device_cus(void);
const char *lulesh_status_string(int status);
const char *lulesh_version(void);

/* ------------------------------------------------------------------ */
/* problems                                                            */
/* ------------------------------------------------------------------ */

/* Sedov set-up of tp^3 subdomains with nx^3 elements each, nr regions,
 * balance and cost as the reference's -s/-r/-b/-c options
 * (LULESH/omp_4.0/lulesh-init.cc:16-184, lulesh-util.cc:60-173).  Replaces
 * Domain::Domain + InitMeshDecomp + the initial nodalMass exchange
 * (lulesh.cc:3140-3161).  LULESH_EINVAL when nx^3 * 8 passes 2^31 - 1
 * (32-bit byte offsets into an element array: nx <= 645) or nx^3 times the
 * largest EOS repetition count does (10 (1 + cost) with nr >= 5, else
 * 1 + cost: nx <= 475 at nr 11, cost 1), and when the region weights
 * overflow the reference's int sum (sum over regions of (i+1)^balance above
 * 2^31 - 1, lulesh-init.cc:384-388: e.g. -r 11 -b 9) or nr passes 2^20.
 * LULESH_ENOMEM when host or device memory runs out. */
int lulesh_group_create(int nx, int tp, int nr, int balance, int cost,
                        lulesh_group **out);
int lulesh_group_destroy(lulesh_group *g);

/* Unstructured hexahedral meshes: the reference's -u <file.lmesh> input
 * (LULESH/cuda/src/lulesh.cu:855-1052), one domain only, as there.
 *
 * lulesh_mesh_read parses the text format of lulesh.cu:855-976: numElem
 * numNode; numNode coordinate triples; numElem x 8 node ids (corners in
 * the reference's hexahedron order); numElem x 6 face neighbours (xi-,
 * xi+, eta-, eta+, zeta-, zeta+; a boundary face names the element
 * itself); then four counted node lists: symmetry X, symmetry Y, symmetry
 * Z, free surface.  elemBC and the octant corner are derived as
 * lulesh.cu:978-1045 derives them: a face whose four nodes all lie in the
 * union of the symmetry sets gets its *_SYMM bit, the first element whose
 * bits are then exactly XI_M|ETA_M|ZETA_M symmetric is the octant corner,
 * and a face whose four nodes are all free-surface nodes also gets its
 * *_SYMM bit (the reference's choice, kept).  LULESH_EIO: unreadable or
 * truncated file; LULESH_EINVAL: an index out of range, numElem above
 * 2^26 - 1 or numNode above 2^28 - 1 (the kernels address the arrays with
 * 32-bit byte offsets).  lulesh_mesh_create takes the same content from arrays
 * (nodelist [numElem][8], nbr [numElem][6]). */
int lulesh_mesh_read(const char *path, lulesh_mesh **out);
int lulesh_mesh_create(int numElem, int numNode, const double *x, const double *y, const double *z,
                       const int *nodelist, const int *nbr, const int *symmX, int nSymmX,
                       const int *symmY, int nSymmY, const int *symmZ, int nSymmZ,
                       const int *freeSurf, int nFree, lulesh_mesh **out);
int lulesh_mesh_destroy(lulesh_mesh *m);
int lulesh_mesh_sizes(const lulesh_mesh *m, int *numElem, int *numNode);
/* elemBC per element (bits of LULESH/omp_4.0/lulesh.h:60-88); the setter
 * replaces the derived bits. */
int lulesh_mesh_get_elem_bc(const lulesh_mesh *m, int *out, size_t n);
int lulesh_mesh_set_elem_bc(lulesh_mesh *m, const int *bc, size_t n);
int lulesh_mesh_octant_corner(const lulesh_mesh *m);
/* Energy deposited in the octant-corner element: 3.948746e+7 as read
 * (lulesh.cu:1049-1050). */
int lulesh_mesh_set_energy(lulesh_mesh *m, double e);
/* A one-domain group on a mesh (nr, balance, cost as lulesh_group_create;
 * the mesh may be destroyed afterwards).  The deposit stays in the octant
 * corner and dt0 = 0.5 cbrt(volo[0]) / sqrt(2 e_deposit), the analytic CFL
 * estimate of lulesh.cu:1205.  (The reference's own start-up then
 * overwrites e[0] with an energy scaled by the structured edge length,
 * which is atoi(<file name>) = 0 on this path: no energy and an infinite
 * dt0.)  On such a group the comm calls and exchanges fail with
 * LULESH_EINVAL (no subdomains: SetupCommBuffers has no unstructured form,
 * LULESH/cuda/README.md:111-113), lulesh_domain_sizes reports nx = 0 and
 * allElem = numElem, and lulesh_group_verify reports the octant corner's
 * energy with zero symmetry figures (the reference prints none). */
int lulesh_group_create_mesh(const lulesh_mesh *m, int nr, int balance, int cost, lulesh_group **out);

int lulesh_group_num_domains(const lulesh_group *g);
int lulesh_group_domain(lulesh_group *g, int rank, lulesh_domain **out);

/* Overrides for tests and drivers. */
int lulesh_group_set_stoptime(lulesh_group *g, double stoptime);
int lulesh_group_set_qstop(lulesh_group *g, double qstop);
/* Cap on the cycle count, the reference's -i option (default: none). */
int lulesh_group_set_max_cycles(lulesh_group *g, int max_cycles);
/* Lifecycle accounting (lulesh.cu:4918-4941, comm_backend.h:33-48).
 * lulesh_group_comm_teardown releases the backend's receive buffers (after
 * the last exchange; a multi-subdomain group cannot step afterwards) and
 * times it.  lulesh_group_phase_times fills, in milliseconds and as the
 * maximum over subdomains: [0] comm set-up (the backend's alloc_recv at
 * group creation or backend change), [1] window allocation and [2] peer
 * query (breakdowns a window-based backend would give; 0 here), [3] comm
 * teardown. */
int lulesh_group_comm_teardown(lulesh_group *g);
int lulesh_group_phase_times(const lulesh_group *g, double out[4]);
/* The final node coordinates of one subdomain as the reference's
 * write_solution writes them (xyz.asc, lulesh.cu:4584-4603): per node its
 * index ("%10d") and x, y, z ("%.10f"), one value per line. */
int lulesh_domain_write_solution(lulesh_domain *d, const char *path);

/* Forces (LULESH_FX..FZ) of the last cycle.  A group with one subdomain
 * advances each node as soon as its force is complete and does not store
 * the force unless asked to with keep = 1 (costs one extra node-array
 * write per component per cycle); lulesh_domain_get_field of FX..FZ on
 * such a group fails with LULESH_EINVAL otherwise.  Groups with several
 * subdomains always store them: the CommSBN exchange sums them. */
int lulesh_group_set_keep_forces(lulesh_group *g, int keep);

/* Runs the time loop of main (lulesh.cc:3222-3235) on the device until
 * time >= stoptime, the cycle cap, or an error.  cycles_per_sync cycles are
 * enqueued (as one captured HIP graph) between host checks of the device
 * control block; 0 picks a default.  Returns the status. */
int lulesh_group_run(lulesh_group *g, int cycles_per_sync, lulesh_stream_t stream);

/* Enqueues exactly `ncycles` leapfrog cycles (each one a TimeIncrement +
 * LagrangeLeapFrog, lulesh.cc:3224-3225) without waiting.  Cycles after the
 * stop condition are no-ops on the device. */
int lulesh_group_enqueue(lulesh_group *g, int ncycles, lulesh_stream_t stream);
/* Waits for the stream and reports the status, cycle count and time of
 * rank 0. */
int lulesh_group_sync(lulesh_group *g, int *cycle, double *time,
                      double *deltatime, lulesh_stream_t stream);

/* Runs `ncycles` cycles launch by launch on the group's stream with a HIP
 * event pair around each kernel class and returns the summed milliseconds:
 * ms[0] time increment (0 where the force kernel computes it), [1] element
 * forces (+ gather, + update, + time increment on the fused path), [2] node
 * gather (+ update) of the two-pass path, [3] node update, [4] kinematics +
 * gradients, [5] MonoQ + EOS + dt, [6] halo exchanges.  These are real
 * cycles of the run. */
int lulesh_group_profile(lulesh_group *g, int ncycles, double ms[7]);

/* Launch shape of the group's fused force kernel (a diagnostic; no
 * counterpart in the reference, whose force kernels are per element and
 * per node): the number of column-run workgroups, the longest run in node
 * planes, and whether the runs are the paired plan (longer runs for the
 * first workgroup of each CU, shorter for the second; see the engine) or
 * the even split.  An unstructured group reports 0 workgroups. */
int lulesh_group_force_plan(lulesh_group *g, int *workgroups, int *node_planes_max, int *paired);

/* Final-output figures of VerifyAndWriteFinalOutput
 * (LULESH/omp_4.0/lulesh-util.cc:177-224) for subdomain `rank`:
 * out[0] origin energy e(0), out[1] MaxAbsDiff, out[2] TotalAbsDiff,
 * out[3] MaxRelDiff of the plane-0 energy symmetry. */
int lulesh_group_verify(lulesh_group *g, int rank, double out[4]);

/* ------------------------------------------------------------------ */
/* subdomains                                                          */
/* ------------------------------------------------------------------ */

int lulesh_domain_rank(const lulesh_domain *d);
/* elements per edge, element count, node count, element count incl. the
 * six ghost planes of the velocity gradients */
int lulesh_domain_sizes(const lulesh_domain *d, int *nx, int *num_elem,
                        int *num_node, int *all_elem);
/* Host copies of a field (n = number of doubles the buffer holds). */
int lulesh_domain_get_field(lulesh_domain *d, int field, double *host, size_t n);
int lulesh_domain_set_field(lulesh_domain *d, int field, const double *host, size_t n);
/* Device pointer of a field (the allocation stays valid for the life of
 * the group).  A group with one subdomain keeps its node positions and
 * velocities (X..ZD) in two buffers that alternate every cycle: the pointer
 * returned for those fields holds the current state until the next cycle
 * runs.  Ask again after stepping. */
int lulesh_domain_field_ptr(lulesh_domain *d, int field, double **dev_ptr);
/* Region bookkeeping of CreateRegionIndexSets (lulesh-init.cc:377-484) and
 * the per-element repetition count of main (lulesh.cc:3171-3192). */
int lulesh_domain_get_elem_rep(lulesh_domain *d, int *host, size_t n);
int lulesh_domain_get_region(lulesh_domain *d, int *host, size_t n);
/* Host copy of the EOS work table the EOS kernel walks (the repetitions of
 * EvalEOSForElems, lulesh.cc:2677-2815, made explicit): one uint16 per EOS
 * evaluation, elements ascending, element i listed elemRep[i] times; low
 * 15 bits = element index mod 2^15, bit 15 set on the element's last
 * repetition.  Returns the number of entries (the sum of elemRep), or the
 * status when n is smaller. */
int lulesh_domain_get_eos_table(lulesh_domain *d, uint16_t *host, size_t n);

/* ------------------------------------------------------------------ */
/* halo exchange: the reference's comm interface                       */
/* ------------------------------------------------------------------ */

/* Same names, argument meaning and call order as lulesh-comm.cc
 * (LULESH/omp_4.0/lulesh-comm.cc:59, 357, 848, 1261, 1684) and the GPU
 * forms of LULESH/cuda/src/lulesh.h:321-340.  `fields` holds lulesh_field
 * values in place of Domain_member pointers.  Within one exchange step the
 * caller issues CommSend for every subdomain of the group before any
 * unpack (CommSBN / CommSyncPosVel / CommMonoQ), which is the ordering MPI
 * gives the reference.  All work is enqueued on `stream`. */
int CommRecv(lulesh_domain *d, int msgType, int xferFields,
             int dx, int dy, int dz, int doRecv, int planeOnly);
int CommSend(lulesh_domain *d, int msgType, int xferFields, const int *fields,
             int dx, int dy, int dz, int doSend, int planeOnly,
             lulesh_stream_t stream);
int CommSBN(lulesh_domain *d, int xferFields, const int *fields,
            lulesh_stream_t stream);
int CommSyncPosVel(lulesh_domain *d, lulesh_stream_t stream);
int CommMonoQ(lulesh_domain *d, lulesh_stream_t stream);

/* Receive-buffer wire layout: comBufSize of SetupCommBuffers
 * (lulesh-init.cc:317-372) and the receiver offset of a message from
 * direction (dcol,drow,dplane) (shmRecvOffset, cuda
 * comm/comm_backend.h:75-138); -1 = no message expected. */
int lulesh_comm_buf_size(const lulesh_domain *d);
int lulesh_comm_recv_offset(const lulesh_domain *d, int dcol, int drow, int dplane,
                            int xferFields, int doRecv, int planeOnly);
/* Host copy of a subdomain's receive buffer (comBufSize doubles). */
int lulesh_comm_get_recv_buffer(lulesh_domain *d, double *host, size_t n);

/* The same layout without a group (host only; usable without a GPU):
 * shmRecvOffset(domain, recvRank, dCol, dRow, dPlane, xferFields, doRecv,
 * planeOnly) of cuda comm/comm_backend.h:75-138 for subdomains of nx^3
 * elements in a tp^3 decomposition (-1: no message expected from that
 * direction), and comBufSize of rank `rank` (lulesh-init.cc:317-372). */
int lulesh_shm_recv_offset(int nx, int tp, int recvRank, int dCol, int dRow, int dPlane,
                           int xferFields, int doRecv, int planeOnly);
int lulesh_comm_buf_size_of(int nx, int tp, int rank);
/* Region layout of CreateRegionIndexSets (lulesh-init.cc:377-484, glibc
 * rand() seeded with the rank) and elemRep of main (lulesh.cc:3171-3192)
 * for rank `rank` of a tp^3 decomposition, host only: region[i] is the
 * 0-based region of element i, rep[i] its EOS repetition count.  n = the
 * capacity of each array (nx^3 needed).  Returns nx^3. */
int lulesh_region_layout(int nx, int tp, int rank, int nr, int balance, int cost,
                         int *region, int *rep, size_t n);
/* The force kernel's column tiles that an engine stepping `ngroups`
 * groups (group g: tp[g]^3 subdomains of nx[g]^3 elements) gets on a device
 * of `cus` CUs with z-chunk `kchunk` (0: chosen), host only (no counterpart
 * in the reference; for tests and tuning without a GPU): the planner the
 * engine runs.  Writes up to n tiles of 8 ints each, in launch order --
 * subdomain (numbered over the groups in order), i0, j0, k0 (first node
 * column, row, plane), ex, ey (element slots), flags (2: the subdomain's
 * first tile) and kn (node planes of the run) -- and the z-chunk used and
 * whether the paired plan was taken.  Returns the number of tiles (write
 * none with tiles = NULL). */
int lulesh_force_plan_of(int ngroups, const int *nx, const int *tp, int kchunk, int cus, int *kchunk_used,
                         int *paired, int *tiles, size_t n);

/* Backend hook table: the run-time form of the compile-time backend API of
 * LULESH/cuda/src/comm/comm_backend.h:7-22 (dispatch :140-173).  The first
 * five hooks are required; every later one is optional -- zero-initialise
 * the struct and leave a hook NULL for the device_local behaviour named in
 * its comment.  A backend decides HOW message words move; which words, in
 * which order and to which offsets stays the library's (the reference's
 * wire layout).  Hooks are called on the host while the library enqueues
 * an exchange on `stream`; they must only enqueue stream-ordered work (the
 * stepping engine captures exchanges into HIP graphs).
 *
 * The library ships two: "device_local" (remote pack, Mode C of
 * LULESH/cuda/src/comm/comm_ipc_common.h:177-190: the pack kernel writes
 * into the receiver's buffer) and "device_copy" (Mode A, :192-225: pack
 * into the sender's own send buffer, then one device-to-device copy per
 * message into the receiver's slot). */
typedef struct lulesh_comm_backend_ops {
    const char *name;
    /* commAllocRecv: receive buffer of comBufSize doubles for subdomain d,
     * ready when the call returns or once the null stream completes */
    int (*alloc_recv)(lulesh_domain *d, size_t comBufSize, double **recv, void *ctx);
    /* commTeardown */
    int (*teardown)(lulesh_domain *d, double *recv, void *ctx);
    /* COMM_PACK_DEST: where the pack kernel of sender `from` writes the
     * message to receiver `to` that lands at `offset` doubles into the
     * receiver's buffer; NULL = into the sender's own send buffer, at its
     * CommSend slot, for send_msg / send_corner to deliver */
    double *(*pack_dest)(lulesh_domain *from, lulesh_domain *to, size_t offset, void *ctx);
    /* COMM_SEND_FINISH: after all sends of `from` */
    int (*send_finish)(lulesh_domain *from, lulesh_stream_t stream, void *ctx);
    /* COMM_RECV_BASE: the device buffer the unpack kernels read */
    double *(*recv_base)(lulesh_domain *d, void *ctx);

    /* ---- optional ---- */
    /* COMM_RUNTIME_INIT / COMM_RUNTIME_SHUTDOWN: when installed on a group
     * (before alloc_recv) / when the group lets the backend go (replaced or
     * destroyed).  NULL: nothing. */
    int (*runtime_init)(void *ctx);
    int (*runtime_shutdown)(void *ctx);
    /* COMM_RECV_SKIP: called by CommRecv; a status < 0 fails it.  NULL:
     * nothing to post (one-sided). */
    int (*recv_skip)(lulesh_domain *d, int msgType, int xferFields, int doRecv, int planeOnly,
                     void *ctx);
    /* COMM_RECV_POST_BUF: where delivered messages land (device or pinned
     * host memory, comBufSize doubles).  NULL: recv_base. */
    double *(*recv_post_buf)(lulesh_domain *d, void *ctx);
    /* COMM_SEND_MSG / COMM_SEND_CORNER: deliver one locally packed face or
     * edge / corner message: `count` doubles at device address `src` go to
     * `offset` of the receiver's post buffer.  NULL: one stream-ordered
     * copy into recv_post_buf(to) + offset. */
    int (*send_msg)(lulesh_domain *from, lulesh_domain *to, size_t offset, const double *src,
                    size_t count, lulesh_stream_t stream, void *ctx);
    int (*send_corner)(lulesh_domain *from, lulesh_domain *to, size_t offset, const double *src,
                       size_t count, lulesh_stream_t stream, void *ctx);
    /* COMM_UNPACK_H2D: before a subdomain's unpack, per message, when the
     * post buffer is not recv_base: copy `bytes` from `src` (post buffer)
     * to `dst` (recv_base).  NULL: one stream-ordered copy. */
    int (*unpack_h2d)(lulesh_domain *d, double *dst, const double *src, size_t bytes,
                      lulesh_stream_t stream, void *ctx);
    /* COMM_ADD_CORNER / COMM_COPY_CORNER: apply one corner message (one
     * word per field at device address `src`) to node `node` of d, adding
     * (CommSBN) or overwriting (CommSyncPosVel); lulesh_comm_apply_corner
     * is the library's own launch.  Corner messages are the last a corner
     * node receives, so they are applied after the unpack of the faces and
     * edges.  NULL: in the unpack kernel. */
    int (*add_corner)(lulesh_domain *d, int node, const double *src, int xferFields,
                      const int *fields, lulesh_stream_t stream, void *ctx);
    int (*copy_corner)(lulesh_domain *d, int node, const double *src, int xferFields,
                       const int *fields, lulesh_stream_t stream, void *ctx);
    /* COMM_MONOQ_COPY: one field of one MonoQ message, `bytes` from `src`
     * (recv_base) into its ghost plane at device address `dst`.  NULL: in
     * the unpack kernel. */
    int (*monoq_copy)(lulesh_domain *d, double *dst, const double *src, size_t bytes,
                      lulesh_stream_t stream, void *ctx);
} lulesh_comm_backend_ops;

const lulesh_comm_backend_ops *lulesh_device_local_backend(void);
const lulesh_comm_backend_ops *lulesh_device_copy_backend(void);
/* Installs a backend on a group (the set-up exchange of lulesh_group_create
 * used device_local): runtime_init, then a fresh receive buffer per
 * subdomain from alloc_recv; the previous backend's buffers are torn down
 * and its runtime_shutdown called. */
int lulesh_group_set_comm_backend(lulesh_group *g, const lulesh_comm_backend_ops *ops,
                                  void *ctx);
/* The library's corner unpack (the reference's AddCorner / CopyCorner
 * kernels, lulesh-comms-gpu.cu:324-339), for backends that override the
 * corner hooks: node `node` of d gets (add = 1) or takes (add = 0) the
 * xferFields words at device address src, field by field. */
int lulesh_comm_apply_corner(lulesh_domain *d, int node, const double *src, int xferFields,
                             const int *fields, int add, lulesh_stream_t stream);
/* One whole exchange step on every subdomain of the group: CommSend on
 * all, then the matching unpack on all, through the installed backend. */
int lulesh_group_exchange(lulesh_group *g, int msgType, int xferFields,
                          const int *fields, lulesh_stream_t stream);

/* Bytes of the device scratch array lulesh_group_exchange_noalloc uses,
 * one per GPU, held until the process ends. */
#define LULESH_EXCHANGE_SCRATCH_BYTES (32u << 20)
/* lulesh_group_exchange without allocating: no counterpart in the
 * reference.  For callers that cannot allocate device memory (tight
 * serving loops, drivers that forbid hipMalloc).  Same step, same backend
 * hooks, same results as lulesh_group_exchange, with the step's plan
 * tables and send staging in the library's scratch array instead of
 * allocations (lulesh_group_exchange allocates them on a step's first
 * use): calls no hipMalloc or hipFree.  Returns after the step has
 * completed on the GPU.  LULESH_ENOMEM when a step needs more than
 * LULESH_EXCHANGE_SCRATCH_BYTES (8 subdomains of -s 45 need about 4 MB
 * for the six PosVel fields).  Every call on one GPU shares that GPU's
 * scratch: calls that overlap on one GPU (from several threads) race. */
int lulesh_group_exchange_noalloc(lulesh_group *g, int msgType, int xferFields,
                                  const int *fields, lulesh_stream_t stream);

/* ------------------------------------------------------------------ */
/* batched ensemble                                                    */
/* ------------------------------------------------------------------ */

/* Many independent single-domain Sedov problems advanced together: every
 * kernel of a cycle covers all problems through one tile table, each
 * problem keeps its own dt / time / cycle, and one cycle of the whole
 * ensemble is replayed as one captured HIP graph. */
int lulesh_batch_create(int nprob, const int *nx, const int *nr,
                        const int *balance, const int *cost, lulesh_batch **out);
int lulesh_batch_destroy(lulesh_batch *b);
int lulesh_batch_set_max_cycles(lulesh_batch *b, int max_cycles);
/* Runs every problem to its stop condition; returns the first error. */
int lulesh_batch_run(lulesh_batch *b, int cycles_per_sync, lulesh_stream_t stream);
int lulesh_batch_group(lulesh_batch *b, int i, lulesh_group **out);
/* per-problem cycles and status after a run */
int lulesh_batch_results(lulesh_batch *b, int *cycles, int *status);
/* Element/block tile shapes of the batched kernels (the autotuner's
 * output): tile_elems = elements per workgroup for element kernels,
 * tile_nodes for node kernels (64, 128, 192 or 256); 0 keeps the current
 * value. */
int lulesh_batch_set_tiles(lulesh_batch *b, int tile_elems, int tile_nodes);
/* Node planes per column tile of the fused force kernel (0 = automatic): a
 * column of P node planes is split into max(1, P / kchunk) runs of kchunk
 * or kchunk + 1 planes. */
int lulesh_batch_set_kchunk(lulesh_batch *b, int kchunk);
/* Autotuner: times `probe_cycles` real cycles of the ensemble for each
 * candidate element tile {64,128,256}, node tile {128,256} and column chunk
 * {auto,4,8}, installs the fastest and reports it (ms per cycle). */
int lulesh_batch_autotune(lulesh_batch *b, int probe_cycles, int *tile_elems, int *tile_nodes,
                          int *kchunk, double *ms_per_cycle);
/* Enqueue `ncycles` cycles of the whole ensemble without waiting / wait. */
int lulesh_batch_enqueue(lulesh_batch *b, int ncycles);
int lulesh_batch_synchronize(lulesh_batch *b);
/* lulesh_group_profile for the whole ensemble */
int lulesh_batch_profile(lulesh_batch *b, int ncycles, double ms[7]);

/* ------------------------------------------------------------------ */
/* serving                                                             */
/* ------------------------------------------------------------------ */

/* Continuous-batching server: requests are admitted into slots while the
 * free-HBM budget allows, their state stays resident, all admitted
 * problems are stepped together, and finished ones free their slot. */
int lulesh_server_create(size_t hbm_budget_bytes, int max_slots, lulesh_server **out);
int lulesh_server_destroy(lulesh_server *s);
/* Queues a request; returns its id (>= 0) or a status < 0. */
int lulesh_server_submit(lulesh_server *s, int nx, int nr, int balance, int cost,
                         double arrival_time);
/* Admits what fits, advances all admitted problems by `cycles` cycles and
 * retires finished ones; returns the number of requests completed in this
 * call. */
int lulesh_server_poll(lulesh_server *s, int cycles, lulesh_stream_t stream);
/* Result of a completed request: status, cycle count, origin energy, and
 * the TotalAbsDiff symmetry figure. */
int lulesh_server_result(lulesh_server *s, int id, int *status, int *cycles,
                         double *origin_energy, double *total_abs_diff);
int lulesh_server_stats(lulesh_server *s, int *queued, int *active, int *done,
                        size_t *hbm_in_use);
/* Server clock (seconds since lulesh_server_create); arrival times of
 * lulesh_server_submit are on this clock and a request is admitted only
 * once it has arrived.  lulesh_server_times gives admission and completion
 * times of a request on the same clock. */
double lulesh_server_now(lulesh_server *s);
int lulesh_server_times(lulesh_server *s, int id, double *t_admit, double *t_done);

/* ------------------------------------------------------------------ */
/* arithmetic probes (parity tests of the device math)                 */
/* ------------------------------------------------------------------ */
/* out[i] = correctly rounded cube root of in[i], computed on the GPU */
int lulesh_probe_cbrt(const double *in, double *out, int n);
/* Average milliseconds of a streaming device copy of n doubles (reps
 * launches): the measured-bandwidth yardstick of the profiling runs. */
double lulesh_probe_copy(size_t n, int reps);
/* out2[2i] = sqrt(a[i]), out2[2i+1] = a[i] / b[i], computed on the GPU */
int lulesh_probe_sqrt_div(const double *a, const double *b, double *out2, int n);
/* out2[2i] = sqrt(a[i]) and out2[2i+1] = 1 / a[i] as the EOS and
 * kinematics kernels form them where the physics bounds the argument
 * (sqrt: 2^-767 <= a[i] < inf; 1/x: 2^-767 <= |a[i]| < 2^1021) */
int lulesh_probe_range_math(const double *a, double *out2, int n);
/* FP64 issue ceiling (a diagnostic with no counterpart in the reference):
 * a kernel of 16 FP64 multiplies and adds per loop trip and lane, as eight
 * independent chains (one_chain 0) or one dependent chain (one_chain 1),
 * with waves_per_simd (1-8) waves on every SIMD and no memory traffic, runs
 * `iters` (1-10^6) loop trips; reports the FP64 rate it reached in TFLOP/s
 * and the mean shader clock over its loops in MHz.  The measured-FP64
 * yardstick of the profiling runs, beside lulesh_probe_copy. */
int lulesh_probe_fp64(int waves_per_simd, int one_chain, int iters, double *tflops, double *clock_mhz);

#ifdef __cplusplus
}
#endif
#endif /* LULESH_AMD_H */
