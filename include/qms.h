/*
 * qms.h -- C ABI of libqms.so, the MI355X time-stepping core of the AlpineQMS
 * quasi-magnetostatic A-V solver (DanRich123/AlpineQMS, 3D/QMS_3D.f90).
 *
 * The reference has no library interface: one Fortran program assembles a
 * CSR matrix every step and calls MKL PARDISO.  Each entry point below
 * replaces one stage of that program; the comment names the reference lines
 * it stands in for.  Conventions:
 *   - every function returns QMS_OK (0) or a negative QMS_ERR_* code, never
 *     exits; qms_last_error() gives the message;
 *   - host buffers are borrowed for the duration of the call and copied;
 *   - device memory is owned by the handle; one handle is used by one thread
 *     at a time; distinct handles may run concurrently (each has its own HIP
 *     stream);
 *   - node-sized field arrays have extents (nx, ny, nz) = cells + 1, x
 *     fastest (the reference's node-sized storage, QMS_3D.f90:288-299);
 *     cell arrays have extents cells, x fastest (Fortran order).
 */
#ifndef QMS_H
#define QMS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QMS_OK 0
#define QMS_ERR_ARG (-1)      /* invalid argument or size */
#define QMS_ERR_HIP (-2)      /* HIP runtime error (message has the detail) */
#define QMS_ERR_NOCONV (-3)   /* Krylov solve did not reach the tolerance */
#define QMS_ERR_NOMEM (-4)    /* device allocation failed */
#define QMS_ERR_STATE (-5)    /* call out of order (e.g. step before steady) */
#define QMS_BUSY 1            /* qms_poll: an asynchronous run is still going */

#define QMS_FP64 0            /* Krylov and multigrid in FP64 */
#define QMS_MIXED 1           /* FP64 refinement around an FP32 inner solve (Krylov basis, operator
                                 input/output, A-block multigrid); V-block multigrid in FP64 */

#define QMS_FIELD_AX 0
#define QMS_FIELD_AY 1
#define QMS_FIELD_AZ 2
#define QMS_FIELD_V 3

typedef struct qms_handle qms_handle;

typedef struct {
    int cells[3];        /* x_size, y_size, z_size (inputs.txt line 1, QMS_3D.f90:123) */
    double h[3];         /* del_x, del_y, del_z (line 2, :124) */
    double dt;           /* del_t (line 3, :125) */
    double bc[12];       /* Bx,By,Bz then gxx..gzz (lines 8-11, :129-132) */
    int precision;       /* QMS_FP64 or QMS_MIXED */
    double tol;          /* ||D^-1 (b - M x)|| <= tol ||D^-1 b||, D = diag(M) */
    int restart;         /* FGMRES restart length */
    int max_iters;       /* Krylov iteration cap per solve */
    int mg_nu;           /* Jacobi sweeps before and after the coarse correction on level 0 of
                            the A block */
    double mg_omega;     /* Jacobi damping of the A block */
    double mg_alpha;     /* coarse-correction scaling of the A block */
    int use_graphs;      /* 1: each FGMRES iteration replays a captured hipGraph (default);
                            0: its kernels are launched one by one (same results bit for bit) */
    double inner_tol;    /* QMS_MIXED: each FP32 inner FGMRES cycle stops once it has reduced the
                            FP64 residual it started from by this factor (<= 0: 1e-6) */
    int mg_nu_coarse;    /* Jacobi sweeps before and after the correction on the coarse levels
                            (<= 0: mg_nu) */
    double floor_tol;    /* a solve whose FGMRES stagnates (the residual estimate improves by less
                            than 10% over two iterations, or a cycle does not halve the residual)
                            is accepted when its FP64 scaled residual is at most this, and counted
                            in qms_stats.stalled_solves (<= 0: 100 x tol) */
    int mg_fused_tail;   /* 1: the multigrid level above the coarsest (at most 8 nodes per axis)
                            and the coarsest solve run as one single-workgroup launch per
                            V-cycle; 0: one launch per sweep.  Same results bit for bit. */
    double mg_omega_v;   /* Jacobi damping of the V block (<= 0: mg_omega) */
    int mg_nu_v;         /* Jacobi sweeps before and after the coarse correction on level 0 of
                            the V block (<= 0: mg_nu) */
    double mg_alpha_v;   /* coarse-correction scaling of the V block (<= 0: mg_alpha) */
    double mg_cheb_lmax; /* > 0: Chebyshev smoothing of the A block: the nu sweeps of a level are
                            Richardson sweeps x += w_k D^-1 (b - A x) with w_k = 1 / theta_k,
                            theta_k the roots of the degree-nu Chebyshev polynomial on
                            [mg_cheb_lmax / mg_cheb_ratio, mg_cheb_lmax] (mg_cheb_lmax: a bound on
                            the spectrum of D^-1 A), the largest root first before the coarse
                            correction and in reverse order after it; <= 0: damped Jacobi with
                            mg_omega */
    double mg_cheb_ratio;   /* ratio of that interval (<= 1: 4) */
    double mg_cheb_lmax_v;  /* the same for the V block (<= 0: damped Jacobi with mg_omega_v) */
    double mg_cheb_ratio_v; /* ratio of the V block's interval (<= 1: 4) */
    int mg_fused_restrict;  /* 1: level 0's residual, its restriction to level 1 and level 1's
                               first sweep run as one launch (where level 1 is not the tail's);
                               0: two launches.  Same results bit for bit. */
} qms_config;

typedef struct {
    int last_iters;          /* Krylov iterations of the last solve */
    int last_restarts;       /* FP64 restarts (residual recomputations) */
    double last_relres;      /* final scaled relative residual */
    double last_relres0;     /* initial scaled relative residual */
    int64_t total_iters;     /* since creation */
    int64_t solves;
    int64_t device_bytes;    /* device memory held by the handle */
    int n_levels_a, n_levels_v;
    int coef_classes;        /* distinct coefficient tuples (0: operator streams the arrays) */
    int64_t queued_iters;    /* iterations queued since creation (>= total_iters: the solver
                                queues iterations ahead and reads convergence between chunks) */
    int mg_classes_a, mg_classes_v;   /* row classes of the first coarse multigrid level of the A and
                                         V blocks (0: that level stores explicit stencil arrays) */
    int64_t stalled_solves;  /* solves accepted on stagnation within floor_tol (two stagnant cycles
                                above it fail the solve with QMS_ERR_NOCONV) */
    int fused_restrict_blocks;  /* multigrid blocks (0..2) whose last V-cycle ran level 0's residual,
                                   its restriction and level 1's first sweep as one launch
                                   (mg_fused_restrict) */
} qms_stats;

/* Fills defaults: FP64, tol 1e-12, restart 30, 3000 iterations; multigrid:
 * on level 0 nu 4 (A block) and 6 (V block), 3 on the coarse levels,
 * Chebyshev smoothing on [0.65, 2.6] (A block) and [0.25, 2.0] (V block)
 * (the damped-Jacobi omegas 0.7 and 0.85 apply when it is switched off),
 * alpha 1.5 (A block) and 1.9 (V block), fused tail and fused level-0
 * restriction on; graphs on,
 * inner_tol 1e-6. */
int qms_config_default(qms_config *cfg);

/* Setup and material coefficients: replaces QMS_3D.f90:172-269 (sizes,
 * geometry read, material -> Yee means) and :289-396 (allocations, Dirichlet
 * masks).  cell_slot holds cells_x*cells_y*cells_z slot numbers (x fastest);
 * slot_props holds nslots rows {mur_x, mur_y, mur_z, sigma_x, sigma_y,
 * sigma_z}; slot 0 must be vacuum {1,1,1,0,0,0} (material ID 0 and IDs absent
 * from the table, :234-243). */
int qms_create(const qms_config *cfg, const uint16_t *cell_slot, const double *slot_props, int nslots,
               qms_handle **out);

/* Steady-state initialisation at frame velocity v: replaces the steady
 * assembly and PARDISO phase 13, QMS_3D.f90:430-1317, including the unpack
 * :1320-1343.  Initial guess is the Dirichlet lift (zero interior). */
int qms_steady(qms_handle *h, const double v[3]);

/* One implicit time step at frame velocity v (the velocity of the step, from
 * trajectory_1): replaces the transient assembly, PARDISO phase 23/33 and
 * unpack, QMS_3D.f90:1452-2234.  Warm-started from the current state. */
int qms_step(qms_handle *h, const double v[3]);

/* nsteps time steps with velocities v[3*s..3*s+2], fields kept on the device
 * between steps (the device-resident time loop). */
int qms_run(qms_handle *h, int nsteps, const double *v);

/* Asynchronous form of qms_run for serving many designs at once: the steps
 * run on a library thread (each handle has its own HIP stream, so the work of
 * different handles overlaps on the GPU) and the call returns at once.  The
 * velocities are copied.  qms_poll returns QMS_BUSY while the run is going,
 * then its status; qms_wait blocks until it ends and returns its status.  No
 * other call may use the handle while a run is going (QMS_ERR_STATE).  The
 * reference steps one design per process (QMS_3D.f90:1424-2302); this has no
 * counterpart there. */
int qms_run_async(qms_handle *h, int nsteps, const double *v);
int qms_poll(qms_handle *h);
int qms_wait(qms_handle *h);

/* Device bytes a handle with this configuration holds once it has solved a
 * steady state and a transient step (state, Krylov basis, both multigrid
 * hierarchies, coefficients; the class table at its maximum size), for
 * admission by memory footprint.  Host-only: needs no GPU.  Returns 0 on an
 * invalid configuration. */
int64_t qms_estimate_bytes(const qms_config *cfg);

/* Copy the state (4 node-sized arrays Ax, Ay, Az, V) to / from the host. */
int qms_get_fields(qms_handle *h, double *out);
int qms_set_fields(qms_handle *h, const double *in);

/* Cell-centred H (unit_mu = 0, QMS_3D.f90:1346-1376) or B (unit_mu = 1) as
 * 3 arrays of cells_x*cells_y*cells_z (Hx, Hy, Hz). */
int qms_get_h(qms_handle *h, int unit_mu, double *out);

/* The Yee material coefficients the design's rows use, as 6 node-sized arrays
 * (Mu_x, Mu_y, Mu_z, Sigma_x, Sigma_y, Sigma_z), QMS_3D.f90:211-269; Mu is
 * returned as 1/(1/Mu) of the stored reciprocal. */
int qms_get_coeffs(qms_handle *h, double *out);

/* y = M(v, dt) x on 4 node-sized arrays, where M is the reference's CSR
 * matrix (dt <= 0: steady S(v), :436-1296; dt > 0: transient T(v, dt),
 * :1452-2191).  Fixed rows are identity rows; entries with no unknown are 0. */
int qms_apply(qms_handle *h, const double v[3], double dt, const double *x, double *y);

/* b(A^n) of the transient rows (dt > 0) or the steady RHS (dt <= 0) on 4
 * node-sized arrays, QMS_3D.f90:459-470, :1475, :1624, :1772, :1906-1928. */
int qms_rhs(qms_handle *h, const double v[3], double dt, const double *aprev, double *b);

/* r = D^-1 (b - M(v, dt) x) on free rows, b - x on fixed rows, 0 where no
 * unknown exists (D = diag(M)): the FP64 residual the Krylov solver restarts
 * from, i.e. the row-scaled residual of the reference's system
 * (QMS_3D.f90:1314-1317, :2193-2203), on 4 node-sized arrays. */
int qms_residual(qms_handle *h, const double v[3], double dt, const double *x, const double *b, double *r);

int qms_get_stats(const qms_handle *h, qms_stats *st);

/* Planes per operator workgroup (1..64): each 256-thread workgroup of the
 * operator kernel owns a column of nodes (qms_set_colwidth) and marches it
 * through this many z-planes, staging the fields through LDS.  qms_create
 * picks the deepest of 1, 2, 4, 8, 16 planes that still gives at least 768
 * columns (three per CU), the fastest choice measured at 32^3..160^3 nodes.
 * A tuning knob with no counterpart in the reference (whose assembly loop is
 * serial, QMS_3D.f90:448-1293). */
int qms_set_zchunk(qms_handle *h, int planes);

/* Operator coefficients from the class table (on != 0, the default when
 * the design has at most 65535 distinct (1/Mu, Sigma) node tuples) or
 * streamed from the six node arrays (on = 0); the rows are bit-identical
 * either way.  qms_create builds the table: a 2-byte class per node and the
 * distinct tuples, instead of 48 bytes per node read on every apply.
 * Performance knob, no reference counterpart. */
int qms_use_coef_classes(qms_handle *h, int on);

/* Operator column width in nodes (32, the default, or 64): a workgroup owns
 * 32 x 8 or 64 x 4 nodes of each plane it marches through.  Tuning knob, no
 * reference counterpart. */
int qms_set_colwidth(qms_handle *h, int nodes);

/* Times `reps` launches of one kernel on the handle's stream with HIP events
 * and reports the mean duration and the algorithmic HBM bytes per launch.
 * v: frame velocity of the rows (NULL: 0); rows are transient when dt > 0.
 * which: 0 fused operator apply (FP64), 1 level-0 A-block smoother sweep,
 * 2 level-0 V-block smoother sweep (precision of the handle's multigrid),
 * 3 plain copy of the 4-field state (calibration of byte counters),
 * 4 step right-hand side, 5 cell-centred H, 6 FP64 residual b - M x,
 * 7 Gram-Schmidt step of FGMRES against restart + 1 basis vectors (the
 * largest of a restart cycle: multi-dot, reduction, update, norm),
 * 8 one A-block V-cycle, 9 one V-block V-cycle, 10 one FGMRES iteration of
 * the handle's precision (preconditioner, operator, Gram-Schmidt against 11
 * basis vectors, Givens update); 8-10 are composite and report 0 bytes;
 * 11 the copy kernel of 3 on two 2 GiB arrays allocated for the call (4 GiB
 * moved per launch: the sustained HBM rate of a plain stream).
 * which | QMS_TIME_COLD: the same kernel with cold caches, as inside a solve
 * where the Krylov basis streams through the caches between two launches:
 * the launches rotate through copies of the kernel's arrays, enough that at
 * least 640 MB of other arrays pass between two uses of one copy (more than
 * the L2s and the 256 MB Infinity Cache hold), back to back and timed as a
 * whole (allocates the copies for the call; up to 64 of them, so a design
 * under about 10 MB per copy stays cached).  Without it, launches reuse the
 * handle's arrays and a working set below 256 MB can stay in the Infinity
 * Cache between them.  The Gram-Schmidt step (7) streams its 4 GB basis
 * either way; the smoother sweeps (1, 2) are timed as without the flag.
 * Operator bytes are the workload's: x, y (and b) plus the six FP64
 * coefficient arrays, also when the kernel reads the class table instead.
 * No reference counterpart (measurement only). */
#define QMS_TIME_COLD 256
int qms_time_kernel(qms_handle *h, int which, const double v[3], int reps, double *ms_per_launch,
                    double *bytes_per_launch);

/* One classical Gram-Schmidt step of the solver's FGMRES, as it runs inside
 * every Krylov iteration (the library's replacement of PARDISO's solve phase,
 * QMS_3D.f90:1314-1317, :2193-2203; PARDISO itself is closed): with the nv
 * vectors of `basis` (nv x 4 node-sized arrays, 1 <= nv <= restart + 1) and
 * w, computes h_i = V_i . w, then w <- w - sum_i h_i V_i and h_nv = w . w,
 * with the fused multi-dot, two-stage reduction and update kernels.
 * Writes h[0..nv] and the updated w.  Exposed for parity tests. */
int qms_gram_schmidt(qms_handle *h, int nv, const double *basis, const double *w, double *hcoef, double *wout);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the step of qms_gram_schmidt on buffers the caller
 * already holds on the current device, with no allocation of any kind.
 * basis_dev: nv vectors of n doubles, vector i at basis_dev + i*n;
 * w_dev: n doubles, orthogonalised in place; hcoef_dev: nv + 1 doubles out
 * (h_i = V_i . w, then |w|^2 after the update).  Queued on `stream` (a
 * hipStream_t; NULL: the null stream); returns once queued.  The kernels'
 * scratch (basis pointer table, partial sums) is a fixed static device
 * buffer of QMS_NOALLOC_SCRATCH_BYTES per GPU, kept until the process ends;
 * every call on one GPU uses that GPU's buffer, so calls that overlap on one
 * GPU (other streams, other threads) race: order them.  QMS_ERR_ARG for
 * n < 1, nv outside [1, QMS_NOALLOC_MAX_VECTORS] or a null buffer.  For
 * n = 4 n1 the results are bit for bit those of qms_gram_schmidt. */
#define QMS_NOALLOC_MAX_VECTORS 64
#define QMS_NOALLOC_SCRATCH_BYTES (1u << 20)
int qms_gram_schmidt_noalloc(int64_t n, int nv, const double *basis_dev, double *w_dev, double *hcoef_dev,
                             void *stream);

/* Copy the handle's work vector (4 node-sized arrays) to the host: the last
 * operator output D^-1 M x written by qms_apply, qms_time_kernel or
 * qms_batch_apply, where D = diag(M) and fixed rows are identity rows. */
int qms_get_work(qms_handle *h, double *out);

/* ---- Batches: many independent designs in one persistent operator launch.
 * No counterpart in the reference (one design per process, QMS_3D.f90:117);
 * the building block of stepping many designs together.  A batch borrows n
 * existing handles (each must stay alive and idle while the batch is used)
 * and owns one HIP stream. */
typedef struct qms_batch qms_batch;
int qms_batch_create(qms_handle *const *handles, int n, qms_batch **out);
/* Grid size of the batched launch.  0: one workgroup per natural piece (a
 * design's column over one z-chunk).  blocks > 0: that many persistent
 * workgroups (rounded up to a multiple of 8), each running a list of column
 * z-spans of about equal cost (qms_batch_plan). */
int qms_batch_set_blocks(qms_batch *b, int blocks);
/* w_d = D_d^-1 M_d(v_d) x_d for every design d (x_d: the handle's state,
 * v: 3n velocities; transient != 0: transient rows with each design's dt).
 * Results land in each handle's work vector (qms_get_work), identical bit
 * for bit to a single-design qms_apply of the same rows. */
int qms_batch_apply(qms_batch *b, const double *v, int transient);
/* Autotuner: times each design's operator at column widths {64, 32} x
 * z-chunks {1,2,3,4,6,8} and keeps the fastest (qms_set_colwidth,
 * qms_set_zchunk); then times the batched launch with those per-design
 * shapes and with one shape for all designs ({64, 32} x z-chunks
 * {2, 4, 8, 16, 32, 64}), each at grid sizes {one workgroup per natural piece,
 * R, 2R} where R is the number of workgroups the device holds at once, and
 * keeps the fastest.  zchunks[n], colwidths[n] and blocks (any may be NULL)
 * receive the choices. */
int qms_batch_tune(qms_batch *b, const double *v, int transient, int reps, int *zchunks, int *colwidths,
                   int *blocks);
/* Mean time of one batched launch (mode 0) or of the n single-design
 * launches it replaces (mode 1), and the algorithmic bytes they move. */
int qms_batch_time(qms_batch *b, int mode, const double *v, int transient, int reps, double *ms_per_launch,
                   double *bytes_per_launch);
/* The schedule of a batched launch, computed on the host with no GPU call
 * (for inspection and tests).  Design d has ncols[d] workgroup columns per
 * plane, nz[d] planes and natural z-chunk band[d]; blocks as for
 * qms_batch_set_blocks.  Writes *nslots (the grid size) and *nsegs.  When
 * segs and slot_off are non-null and seg_cap >= *nsegs, off_cap >= *nslots + 1,
 * also writes the segments (4 values each: design, column, first plane, end
 * plane) and the slot offsets: the workgroup of slot s runs segments
 * slot_off[s] .. slot_off[s+1]-1 in order, and workgroup b has slot
 * (b % 8) * (*nslots / 8) + b / 8.  Returns QMS_ERR_ARG on invalid input or
 * short buffers. */
int qms_batch_plan(int n, const int *ncols, const int *nz, const int *band, int blocks, unsigned *segs,
                   int seg_cap, unsigned *slot_off, int off_cap, int *nsegs, int *nslots);
const char *qms_batch_last_error(const qms_batch *b);
void qms_batch_destroy(qms_batch *b);

void qms_destroy(qms_handle *h);

/* Message of the last failing call on h (h = NULL: last failing
 * qms_create). */
const char *qms_last_error(const qms_handle *h);

#ifdef __cplusplus
}
#endif
#endif
