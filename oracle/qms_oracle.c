/*
 * qms_oracle.c -- CPU restatement of the per-step operator of QMS_3D.f90.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP library is
 * compared against; it is loaded by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg, never by the product path.
 *
 * It restates, in plain C, what the reference program computes before each
 * PARDISO call: the Yee-grid material coefficients, the Dirichlet shell, the
 * CSR matrix of the coupled A-V system (steady or transient) with its
 * right-hand side, and the cell-centred H field that is written after each
 * solve.  Row order, column order and the floating-point evaluation order of
 * every coefficient follow the reference so that the assembled values agree
 * with it to the last bit wherever the reference's compiler keeps source
 * order.  The linear solve itself (MKL PARDISO in the reference) is done by
 * the Python side of the oracle with an FP64 sparse direct solver.
 *
 * Indexing: all (i,j,k) below are 1-based, as in the reference.  Node-sized
 * arrays have extents (Nx,Ny,Nz) = cells+1 and are stored x-fastest
 * (Fortran order).  CSR output is 0-based.
 *
 * Parity pins (tests/test_oracle_*.py): closed-form uniform/gradient field,
 * an independent incidence-matrix construction of every block (curl, grad,
 * penalty, mass, advection), symmetry/antisymmetry of the blocks, structural
 * counts (1/17/37 entries per row, strictly increasing columns), brute-force
 * dense solves.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Default-real (float32) literals of the reference, promoted to double.
 * QMS_3D.f90:465 (1/3.0), :1237-1281 (1.0e-12), :1432-1439 (1E-8). */
#define QO_THIRD ((double)(1.0f / 3.0f))
#define QO_SIGFLOOR ((double)1.0e-12f)
#define QO_NUDGE ((double)1.0e-8f)

/* mu_0 literal, QMS_3D.f90:22 (1.25663706E-6_dp, a double literal). */
#define QO_MU0 1.25663706E-6

typedef struct {
    int cx, cy, cz;       /* cells per axis (inputs.txt line 1) */
    int Nx, Ny, Nz;       /* nodes per axis */
    double hx, hy, hz;    /* del_x, del_y, del_z */
    int64_t n1;           /* Nx*Ny*Nz */
    int64_t offy, offz, offv, n;  /* CSR block offsets and n_dims */
} qo_grid;

static void qo_grid_init(qo_grid *g, int cx, int cy, int cz, double hx,
                         double hy, double hz) {
    g->cx = cx; g->cy = cy; g->cz = cz;
    g->Nx = cx + 1; g->Ny = cy + 1; g->Nz = cz + 1;
    g->hx = hx; g->hy = hy; g->hz = hz;
    g->n1 = (int64_t)g->Nx * g->Ny * g->Nz;
    /* QMS_3D.f90:175 and the block offsets used at :455-457, :628-631,
     * :802-805, :976-979. */
    g->offy = (int64_t)cx * g->Ny * g->Nz;
    g->offz = g->offy + (int64_t)g->Nx * cy * g->Nz;
    g->offv = g->offz + (int64_t)g->Nx * g->Ny * cz;
    g->n = g->offv + g->n1;
}

/* Node-sized array element, 1-based. */
#define NODE(g, a, i, j, k) \
    (a)[((i) - 1) + (int64_t)((j) - 1) * (g)->Nx + (int64_t)((k) - 1) * (g)->Nx * (g)->Ny]

/* 0-based global unknown indices, QMS_3D.f90:455-457, :628-631, :802-805,
 * :976-979. */
#define CAX(g, i, j, k) (((i) - 1) + (int64_t)((j) - 1) * (g)->cx + (int64_t)((k) - 1) * (g)->cx * (g)->Ny)
#define CAY(g, i, j, k) ((g)->offy + ((i) - 1) + (int64_t)((j) - 1) * (g)->Nx + (int64_t)((k) - 1) * (g)->Nx * (g)->cy)
#define CAZ(g, i, j, k) ((g)->offz + ((i) - 1) + (int64_t)((j) - 1) * (g)->Nx + (int64_t)((k) - 1) * (g)->Nx * (g)->Ny)
#define CVV(g, i, j, k) ((g)->offv + ((i) - 1) + (int64_t)((j) - 1) * (g)->Nx + (int64_t)((k) - 1) * (g)->Nx * (g)->Ny)

/* ------------------------------------------------------------------------ */
/* Sizes                                                                    */
/* ------------------------------------------------------------------------ */

/* n_dims, QMS_3D.f90:175. */
int64_t qo_ndims(int cx, int cy, int cz) {
    qo_grid g;
    qo_grid_init(&g, cx, cy, cz, 1, 1, 1);
    return g.n;
}

/* Exact CSR entry count: fixed rows hold 1 entry, interior A rows 17 and
 * interior V rows 37 (QMS_3D.f90:448-1293). */
int64_t qo_nnz(int cx, int cy, int cz) {
    int Nx = cx + 1, Ny = cy + 1, Nz = cz + 1;
    int64_t n = qo_ndims(cx, cy, cz);
    int64_t ax = (int64_t)((cx - 2) > 0 ? cx - 2 : 0) * ((Ny - 4) > 0 ? Ny - 4 : 0) * ((Nz - 4) > 0 ? Nz - 4 : 0);
    int64_t ay = (int64_t)((Nx - 4) > 0 ? Nx - 4 : 0) * ((cy - 2) > 0 ? cy - 2 : 0) * ((Nz - 4) > 0 ? Nz - 4 : 0);
    int64_t az = (int64_t)((Nx - 4) > 0 ? Nx - 4 : 0) * ((Ny - 4) > 0 ? Ny - 4 : 0) * ((cz - 2) > 0 ? cz - 2 : 0);
    int64_t vv = (int64_t)((Nx - 4) > 0 ? Nx - 4 : 0) * ((Ny - 4) > 0 ? Ny - 4 : 0) * ((Nz - 4) > 0 ? Nz - 4 : 0);
    return n + 16 * (ax + ay + az) + 36 * vv;
}

/* The reference's CSR capacity bound exact_nn, QMS_3D.f90:179-181. */
int64_t qo_exact_nn(int cx, int cy, int cz) {
    int64_t Nx = cx + 1, Ny = cy + 1, Nz = cz + 1;
    int64_t nb = (2 * Nx + 2 * Ny - 4) * Nz;
    int64_t ni = ((Nx - 2) * (Ny - 2)) * Nz;
    return nb * 1 + ni * (15 * 3 + 37);
}

/* ------------------------------------------------------------------------ */
/* Material -> Yee coefficients, QMS_3D.f90:211-269                          */
/* ------------------------------------------------------------------------ */

/*
 * geom: cx*cy*cz material IDs as f64, Fortran order (geom.bin).
 * mat:  nmat rows of 7 doubles {id, mur_x, mur_y, mur_z, sig_x, sig_y, sig_z}
 *       (inputs.txt material block, QMS_3D.f90:153-161).
 * coef: 6 node-sized arrays, in order Mu_x, Mu_y, Mu_z, Sigma_x, Sigma_y,
 *       Sigma_z.
 * Returns 0, or -1 if an ID in geom is negative or above the largest table
 * ID (the reference indexes temp_mat_id out of bounds there, :234-243).
 */
int qo_coeffs(int cx, int cy, int cz, const double *geom, int nmat,
              const double *mat, double *coef) {
    qo_grid G, *g = &G;
    qo_grid_init(g, cx, cy, cz, 1, 1, 1);
    const int64_t n1 = g->n1, nc = (int64_t)cx * cy * cz;
    double *mux = coef, *muy = coef + n1, *muz = coef + 2 * n1;
    double *sx = coef + 3 * n1, *sy = coef + 4 * n1, *sz = coef + 5 * n1;
    for (int64_t p = 0; p < n1; ++p) {
        mux[p] = muy[p] = muz[p] = QO_MU0;
        sx[p] = sy[p] = sz[p] = 0.0;
    }
    /* temp_mat_id(0:maxID), :234-238 */
    int maxid = 0;
    for (int r = 0; r < nmat; ++r) {
        int id = (int)mat[7 * r];
        if (id > maxid) maxid = id;
    }
    int *slot = (int *)calloc((size_t)maxid + 1, sizeof(int));
    for (int r = 0; r < nmat; ++r) slot[(int)mat[7 * r]] = r + 1;
    /* cell properties, :240-254 */
    double *cm = (double *)malloc(sizeof(double) * 6 * (size_t)nc);
    int bad = 0;
    for (int64_t p = 0; p < nc; ++p) {
        double m0 = QO_MU0, m1 = QO_MU0, m2 = QO_MU0, s0 = 0, s1 = 0, s2 = 0;
        int id = (int)geom[p];
        if (id < 0 || id > maxid) {
            bad = 1;
        } else if (slot[id]) {
            const double *row = mat + 7 * (slot[id] - 1);
            m0 = row[1] * QO_MU0; m1 = row[2] * QO_MU0; m2 = row[3] * QO_MU0;
            s0 = row[4]; s1 = row[5]; s2 = row[6];
        }
        cm[p] = m0; cm[nc + p] = m1; cm[2 * nc + p] = m2;
        cm[3 * nc + p] = s0; cm[4 * nc + p] = s1; cm[5 * nc + p] = s2;
    }
    free(slot);
    if (bad) { free(cm); return -1; }
#define CELL(c, i, j, k) cm[(int64_t)(c) * nc + ((i) - 1) + (int64_t)((j) - 1) * cx + (int64_t)((k) - 1) * cx * cy]
    /* Yee means for indices 2..cells only, :258-269 */
    for (int k = 2; k <= cz; ++k)
        for (int j = 2; j <= cy; ++j)
            for (int i = 2; i <= cx; ++i) {
                double a, b;
                a = CELL(0, i, j, k); b = CELL(0, i - 1, j, k);
                NODE(g, mux, i, j, k) = 2.0 * a * b / (a + b);
                a = CELL(1, i, j, k); b = CELL(1, i, j - 1, k);
                NODE(g, muy, i, j, k) = 2.0 * a * b / (a + b);
                a = CELL(2, i, j, k); b = CELL(2, i, j, k - 1);
                NODE(g, muz, i, j, k) = 2.0 * a * b / (a + b);
                NODE(g, sx, i, j, k) = (CELL(3, i, j, k) + CELL(3, i, j, k - 1) + CELL(3, i, j - 1, k) + CELL(3, i, j - 1, k - 1)) / 4.0;
                NODE(g, sy, i, j, k) = (CELL(4, i, j, k) + CELL(4, i - 1, j, k) + CELL(4, i, j, k - 1) + CELL(4, i - 1, j, k - 1)) / 4.0;
                NODE(g, sz, i, j, k) = (CELL(5, i, j, k) + CELL(5, i - 1, j, k) + CELL(5, i, j - 1, k) + CELL(5, i - 1, j - 1, k)) / 4.0;
            }
#undef CELL
    free(cm);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Dirichlet shell and boundary values, QMS_3D.f90:361-396, :459-470,       */
/* :633-644, :808-818, :982-996                                              */
/* ------------------------------------------------------------------------ */

static int fixed_ax(const qo_grid *g, int i, int j, int k) {
    return j <= 2 || j >= g->Ny - 1 || k <= 2 || k >= g->Nz - 1 || i == 1 || i == g->cx;
}
static int fixed_ay(const qo_grid *g, int i, int j, int k) {
    return i <= 2 || i >= g->Nx - 1 || k <= 2 || k >= g->Nz - 1 || j == 1 || j == g->cy;
}
static int fixed_az(const qo_grid *g, int i, int j, int k) {
    return i <= 2 || i >= g->Nx - 1 || j <= 2 || j >= g->Ny - 1 || k == 1 || k == g->cz;
}
static int fixed_v(const qo_grid *g, int i, int j, int k) {
    return i <= 2 || i >= g->Nx - 1 || j <= 2 || j >= g->Ny - 1 || k <= 2 || k >= g->Nz - 1;
}

/* bc: {Bx, By, Bz, gxx, gxy, gxz, gyx, gyy, gyz, gzx, gzy, gzz}.
 * Positions are the unstaggered node positions (i-1)*dx etc. */
static double bc_ax(const qo_grid *g, const double *bc, int i, int j, int k) {
    double x = (i - 1) * g->hx, y = (j - 1) * g->hy, z = (k - 1) * g->hz;
    double gy = bc[6] * x + bc[7] * y + bc[8] * z;
    double gz = bc[9] * x + bc[10] * y + bc[11] * z;
    return 0.5 * bc[1] * z - 0.5 * bc[2] * y + QO_THIRD * (gy * z - gz * y);
}
static double bc_ay(const qo_grid *g, const double *bc, int i, int j, int k) {
    double x = (i - 1) * g->hx, y = (j - 1) * g->hy, z = (k - 1) * g->hz;
    double gz = bc[9] * x + bc[10] * y + bc[11] * z;
    double gx = bc[3] * x + bc[4] * y + bc[5] * z;
    return 0.5 * bc[2] * x - 0.5 * bc[0] * z + QO_THIRD * (gz * x - gx * z);
}
static double bc_az(const qo_grid *g, const double *bc, int i, int j, int k) {
    double x = (i - 1) * g->hx, y = (j - 1) * g->hy, z = (k - 1) * g->hz;
    double gx = bc[3] * x + bc[4] * y + bc[5] * z;
    double gy = bc[6] * x + bc[7] * y + bc[8] * z;
    return 0.5 * bc[0] * y - 0.5 * bc[1] * x + QO_THIRD * (gx * y - gy * x);
}
static double bc_v(const qo_grid *g, const double *bc, const double *v, int i,
                   int j, int k) {
    double x = (i - 1) * g->hx, y = (j - 1) * g->hy, z = (k - 1) * g->hz;
    double vx = v[0], vy = v[1], vz = v[2];
    double Bx = bc[0], By = bc[1], Bz = bc[2];
    double gx = bc[3] * x + bc[4] * y + bc[5] * z;
    double gy = bc[6] * x + bc[7] * y + bc[8] * z;
    double gz = bc[9] * x + bc[10] * y + bc[11] * z;
    return -(1.0 * (vy * Bz - vz * By) * x) - 1.0 * (vz * Bx - vx * Bz) * y
           - 1.0 * (vx * By - vy * Bx) * z - 0.5 * (vy * gz - vz * gy) * x
           - 0.5 * (vz * gx - vx * gz) * y - 0.5 * (vx * gy - vy * gx) * z;
}

/* Writes the Dirichlet values into node-sized field arrays (Ax, Ay, Az, V),
 * leaving interior entries untouched.  Used to form initial guesses and by
 * the fixed-row tests. */
void qo_bc_fields(int cx, int cy, int cz, double hx, double hy, double hz,
                  const double *bc, const double *v, double *fields) {
    qo_grid G, *g = &G;
    qo_grid_init(g, cx, cy, cz, hx, hy, hz);
    double *ax = fields, *ay = fields + g->n1, *az = fields + 2 * g->n1, *vv = fields + 3 * g->n1;
    for (int k = 1; k <= g->Nz; ++k)
        for (int j = 1; j <= g->Ny; ++j)
            for (int i = 1; i <= g->Nx; ++i) {
                if (i <= g->cx && fixed_ax(g, i, j, k)) NODE(g, ax, i, j, k) = bc_ax(g, bc, i, j, k);
                if (j <= g->cy && fixed_ay(g, i, j, k)) NODE(g, ay, i, j, k) = bc_ay(g, bc, i, j, k);
                if (k <= g->cz && fixed_az(g, i, j, k)) NODE(g, az, i, j, k) = bc_az(g, bc, i, j, k);
                if (fixed_v(g, i, j, k)) NODE(g, vv, i, j, k) = bc_v(g, bc, v, i, j, k);
            }
}

/* ------------------------------------------------------------------------ */
/* CSR assembly, QMS_3D.f90:436-1296 (steady) and :1452-2191 (transient)     */
/* ------------------------------------------------------------------------ */

typedef struct {
    int64_t *ptr;
    int32_t *col;
    double *val;
    int64_t nz;
} qo_csr;

static inline void put(qo_csr *m, int64_t c, double w) {
    m->col[m->nz] = (int32_t)c;
    m->val[m->nz] = w;
    m->nz++;
}

/*
 * Assemble the system matrix and right-hand side.
 *   coef   6 node arrays from qo_coeffs
 *   bc     {B0[3], G[3][3]}
 *   v      frame velocity (solve_vx, solve_vy, solve_vz)
 *   dt     > 0: transient rows T(v,dt) with RHS from aprev; <= 0: steady S(v)
 *   aprev  3 node arrays (Ax, Ay, Az of the previous step); unused if steady
 *   ptr    n+1, col/val qo_nnz entries, b n entries
 * Returns the entry count written.
 */
int64_t qo_assemble(int cx, int cy, int cz, double hx, double hy, double hz,
                    const double *coef, const double *bc, const double *v,
                    double dt, const double *aprev, int64_t *ptr,
                    int32_t *col, double *val, double *b) {
    qo_grid G, *g = &G;
    qo_grid_init(g, cx, cy, cz, hx, hy, hz);
    const int64_t n1 = g->n1;
    const double *MX = coef, *MY = coef + n1, *MZ = coef + 2 * n1;
    const double *SX = coef + 3 * n1, *SY = coef + 4 * n1, *SZ = coef + 5 * n1;
    const double vx = v[0], vy = v[1], vz = v[2];
    const int tr = dt > 0;
    const double *pAx = aprev, *pAy = aprev ? aprev + n1 : 0, *pAz = aprev ? aprev + 2 * n1 : 0;
    qo_csr M = {ptr, col, val, 0}, *m = &M;
    int64_t row = 0;
#define mux(i, j, k) NODE(g, MX, i, j, k)
#define muy(i, j, k) NODE(g, MY, i, j, k)
#define muz(i, j, k) NODE(g, MZ, i, j, k)
#define sgx(i, j, k) NODE(g, SX, i, j, k)
#define sgy(i, j, k) NODE(g, SY, i, j, k)
#define sgz(i, j, k) NODE(g, SZ, i, j, k)
#define AX(i, j, k) CAX(g, i, j, k)
#define AY(i, j, k) CAY(g, i, j, k)
#define AZ(i, j, k) CAZ(g, i, j, k)
#define VV(i, j, k) CVV(g, i, j, k)

    /* ---- Ax rows, :448-615 / :1457-1597 ---- */
    for (int k = 1; k <= g->Nz; ++k)
        for (int j = 1; j <= g->Ny; ++j)
            for (int i = 1; i <= cx; ++i, ++row) {
                ptr[row] = m->nz;
                if (fixed_ax(g, i, j, k)) {
                    b[row] = bc_ax(g, bc, i, j, k);
                    put(m, row, 1.0);
                    continue;
                }
                const double s = sgx(i, j, k);
                b[row] = tr ? (s / dt) * NODE(g, pAx, i, j, k) : 0.0;
                put(m, AX(i, j, k - 1), -1.0 / muy(i, j, k - 1) / (hz * hz) + (s * vz) / (2.0 * hz));
                put(m, AX(i, j - 1, k), -1.0 / muz(i, j - 1, k) / (hy * hy) + (s * vy) / (2.0 * hy));
                put(m, AX(i - 1, j, k), -1.0 / mux(i - 1, j, k) / (hx * hx));
                {
                    double d = (1.0 / muz(i, j, k) + 1.0 / muz(i, j - 1, k)) / (hy * hy)
                             + (1.0 / muy(i, j, k) + 1.0 / muy(i, j, k - 1)) / (hz * hz)
                             + (1.0 / mux(i, j, k) + 1.0 / mux(i - 1, j, k)) / (hx * hx);
                    if (tr) d = d + s / dt;
                    put(m, AX(i, j, k), d);
                }
                put(m, AX(i + 1, j, k), -1.0 / mux(i, j, k) / (hx * hx));
                put(m, AX(i, j + 1, k), -1.0 / muz(i, j, k) / (hy * hy) - (s * vy) / (2.0 * hy));
                put(m, AX(i, j, k + 1), -1.0 / muy(i, j, k) / (hz * hz) - (s * vz) / (2.0 * hz));
                put(m, AY(i, j - 1, k), 1.0 / muz(i, j - 1, k) / (hx * hy) - (s * vy) / (2.0 * hx));
                put(m, AY(i + 1, j - 1, k), -1.0 / muz(i, j - 1, k) / (hx * hy) + (s * vy) / (2.0 * hx));
                put(m, AY(i, j, k), -1.0 / muz(i, j, k) / (hx * hy) - (s * vy) / (2.0 * hx));
                put(m, AY(i + 1, j, k), 1.0 / muz(i, j, k) / (hx * hy) + (s * vy) / (2.0 * hx));
                put(m, AZ(i, j, k - 1), 1.0 / muy(i, j, k - 1) / (hx * hz) - (s * vz) / (2.0 * hx));
                put(m, AZ(i + 1, j, k - 1), -1.0 / muy(i, j, k - 1) / (hx * hz) + (s * vz) / (2.0 * hx));
                put(m, AZ(i, j, k), -1.0 / muy(i, j, k) / (hx * hz) - (s * vz) / (2.0 * hx));
                put(m, AZ(i + 1, j, k), 1.0 / muy(i, j, k) / (hx * hz) + (s * vz) / (2.0 * hx));
                put(m, VV(i, j, k), -s / hx);
                put(m, VV(i + 1, j, k), s / hx);
            }

    /* ---- Ay rows, :620-789 / :1603-1745 ---- */
    for (int k = 1; k <= g->Nz; ++k)
        for (int j = 1; j <= cy; ++j)
            for (int i = 1; i <= g->Nx; ++i, ++row) {
                ptr[row] = m->nz;
                if (fixed_ay(g, i, j, k)) {
                    b[row] = bc_ay(g, bc, i, j, k);
                    put(m, row, 1.0);
                    continue;
                }
                const double s = sgy(i, j, k);
                b[row] = tr ? (s / dt) * NODE(g, pAy, i, j, k) : 0.0;
                put(m, AX(i - 1, j, k), 1.0 / muz(i - 1, j, k) / (hx * hy) - (s * vx) / (2.0 * hy));
                put(m, AX(i, j, k), -1.0 / muz(i, j, k) / (hx * hy) - (s * vx) / (2.0 * hy));
                put(m, AX(i - 1, j + 1, k), -1.0 / muz(i - 1, j, k) / (hx * hy) + (s * vx) / (2.0 * hy));
                put(m, AX(i, j + 1, k), 1.0 / muz(i, j, k) / (hx * hy) + (s * vx) / (2.0 * hy));
                put(m, AY(i, j, k - 1), -1.0 / mux(i, j, k - 1) / (hz * hz) + (s * vz) / (2.0 * hz));
                put(m, AY(i, j - 1, k), -1.0 / muy(i, j - 1, k) / (hy * hy));
                put(m, AY(i - 1, j, k), -1.0 / muz(i - 1, j, k) / (hx * hx) + (s * vx) / (2.0 * hx));
                {
                    double d = (1.0 / muz(i, j, k) + 1.0 / muz(i - 1, j, k)) / (hx * hx)
                             + (1.0 / mux(i, j, k) + 1.0 / mux(i, j, k - 1)) / (hz * hz)
                             + (1.0 / muy(i, j, k) + 1.0 / muy(i, j - 1, k)) / (hy * hy);
                    if (tr) d = d + s / dt;
                    put(m, AY(i, j, k), d);
                }
                put(m, AY(i + 1, j, k), -1.0 / muz(i, j, k) / (hx * hx) - (s * vx) / (2.0 * hx));
                put(m, AY(i, j + 1, k), -1.0 / muy(i, j, k) / (hy * hy));
                put(m, AY(i, j, k + 1), -1.0 / mux(i, j, k) / (hz * hz) - (s * vz) / (2.0 * hz));
                put(m, AZ(i, j, k - 1), 1.0 / mux(i, j, k - 1) / (hy * hz) - (s * vz) / (2.0 * hy));
                put(m, AZ(i, j + 1, k - 1), -1.0 / mux(i, j, k - 1) / (hy * hz) + (s * vz) / (2.0 * hy));
                put(m, AZ(i, j, k), -1.0 / mux(i, j, k) / (hy * hz) - (s * vz) / (2.0 * hy));
                put(m, AZ(i, j + 1, k), 1.0 / mux(i, j, k) / (hy * hz) + (s * vz) / (2.0 * hy));
                put(m, VV(i, j, k), -s / hy);
                put(m, VV(i, j + 1, k), s / hy);
            }

    /* ---- Az rows, :794-963 / :1751-1893 ---- */
    for (int k = 1; k <= cz; ++k)
        for (int j = 1; j <= g->Ny; ++j)
            for (int i = 1; i <= g->Nx; ++i, ++row) {
                ptr[row] = m->nz;
                if (fixed_az(g, i, j, k)) {
                    b[row] = bc_az(g, bc, i, j, k);
                    put(m, row, 1.0);
                    continue;
                }
                const double s = sgz(i, j, k);
                b[row] = tr ? (s / dt) * NODE(g, pAz, i, j, k) : 0.0;
                put(m, AX(i - 1, j, k), 1.0 / muy(i - 1, j, k) / (hx * hz) - (s * vx) / (2.0 * hz));
                put(m, AX(i, j, k), -1.0 / muy(i, j, k) / (hx * hz) - (s * vx) / (2.0 * hz));
                put(m, AX(i - 1, j, k + 1), -1.0 / muy(i - 1, j, k) / (hx * hz) + (s * vx) / (2.0 * hz));
                put(m, AX(i, j, k + 1), 1.0 / muy(i, j, k) / (hx * hz) + (s * vx) / (2.0 * hz));
                put(m, AY(i, j - 1, k), 1.0 / mux(i, j - 1, k) / (hy * hz) - (s * vy) / (2.0 * hz));
                put(m, AY(i, j, k), -1.0 / mux(i, j, k) / (hy * hz) - (s * vy) / (2.0 * hz));
                put(m, AY(i, j - 1, k + 1), -1.0 / mux(i, j - 1, k) / (hy * hz) + (s * vy) / (2.0 * hz));
                put(m, AY(i, j, k + 1), 1.0 / mux(i, j, k) / (hy * hz) + (s * vy) / (2.0 * hz));
                put(m, AZ(i, j, k - 1), -1.0 / muz(i, j, k - 1) / (hz * hz));
                put(m, AZ(i, j - 1, k), -1.0 / mux(i, j - 1, k) / (hy * hy) + (s * vy) / (2.0 * hy));
                put(m, AZ(i - 1, j, k), -1.0 / muy(i - 1, j, k) / (hx * hx) + (s * vx) / (2.0 * hx));
                {
                    double d = (1.0 / muy(i, j, k) + 1.0 / muy(i - 1, j, k)) / (hx * hx)
                             + (1.0 / mux(i, j, k) + 1.0 / mux(i, j - 1, k)) / (hy * hy)
                             + (1.0 / muz(i, j, k) + 1.0 / muz(i, j, k - 1)) / (hz * hz);
                    if (tr) d = d + s / dt;
                    put(m, AZ(i, j, k), d);
                }
                put(m, AZ(i + 1, j, k), -1.0 / muy(i, j, k) / (hx * hx) - (s * vx) / (2.0 * hx));
                put(m, AZ(i, j + 1, k), -1.0 / mux(i, j, k) / (hy * hy) - (s * vy) / (2.0 * hy));
                put(m, AZ(i, j, k + 1), -1.0 / muz(i, j, k) / (hz * hz));
                put(m, VV(i, j, k), -s / hz);
                put(m, VV(i, j, k + 1), s / hz);
            }

    /* ---- V rows, :968-1293 / :1899-2189 ---- */
    const double hx2 = hx * hx, hy2 = hy * hy, hz2 = hz * hz;
    for (int k = 1; k <= g->Nz; ++k)
        for (int j = 1; j <= g->Ny; ++j)
            for (int i = 1; i <= g->Nx; ++i, ++row) {
                ptr[row] = m->nz;
                if (fixed_v(g, i, j, k)) {
                    b[row] = bc_v(g, bc, v, i, j, k);
                    put(m, row, 1.0);
                    continue;
                }
                const double x0 = sgx(i, j, k), x1 = sgx(i - 1, j, k);
                const double y0 = sgy(i, j, k), y1 = sgy(i, j - 1, k);
                const double z0 = sgz(i, j, k), z1 = sgz(i, j, k - 1);
                if (tr) {
                    /* divergence of sigma*A^n/dt, :1925-1928 */
                    b[row] = (x0 * NODE(g, pAx, i, j, k) - x1 * NODE(g, pAx, i - 1, j, k)) / (dt * hx)
                           + (y0 * NODE(g, pAy, i, j, k) - y1 * NODE(g, pAy, i, j - 1, k)) / (dt * hy)
                           + (z0 * NODE(g, pAz, i, j, k) - z1 * NODE(g, pAz, i, j, k - 1)) / (dt * hz);
                } else {
                    b[row] = 0.0;
                }
                double w;
                /* Ax columns */
                put(m, AX(i - 1, j, k - 1), x1 * vz / (2.0 * hz * hx) - z1 * vx / (2.0 * hz2));
                put(m, AX(i, j, k - 1), -x0 * vz / (2.0 * hz * hx) - z1 * vx / (2.0 * hz2));
                put(m, AX(i - 1, j - 1, k), x1 * vy / (2.0 * hy * hx) - y1 * vx / (2.0 * hy2));
                put(m, AX(i, j - 1, k), -x0 * vy / (2.0 * hy * hx) - y1 * vx / (2.0 * hy2));
                w = (y0 + y1) * vx / (2.0 * hy2) + (z0 + z1) * vx / (2.0 * hz2);
                if (tr) w = w - x1 / (hx * dt);
                put(m, AX(i - 1, j, k), w);
                w = (y0 + y1) * vx / (2.0 * hy2) + (z0 + z1) * vx / (2.0 * hz2);
                if (tr) w = w + x0 / (hx * dt);
                put(m, AX(i, j, k), w);
                put(m, AX(i - 1, j + 1, k), -x1 * vy / (2.0 * hy * hx) - y0 * vx / (2.0 * hy2));
                put(m, AX(i, j + 1, k), x0 * vy / (2.0 * hy * hx) - y0 * vx / (2.0 * hy2));
                put(m, AX(i - 1, j, k + 1), -x1 * vz / (2.0 * hz * hx) - z0 * vx / (2.0 * hz2));
                put(m, AX(i, j, k + 1), x0 * vz / (2.0 * hz * hx) - z0 * vx / (2.0 * hz2));
                /* Ay columns */
                put(m, AY(i, j - 1, k - 1), y1 * vz / (2.0 * hz * hy) - z1 * vy / (2.0 * hz2));
                put(m, AY(i, j, k - 1), -y0 * vz / (2.0 * hz * hy) - z1 * vy / (2.0 * hz2));
                put(m, AY(i - 1, j - 1, k), -x1 * vy / (2.0 * hx2) + y1 * vx / (2.0 * hx * hy));
                w = (x0 + x1) * vy / (2.0 * hx2) + (z0 + z1) * vy / (2.0 * hz2);
                if (tr) w = w - y1 / (hy * dt);
                put(m, AY(i, j - 1, k), w);
                put(m, AY(i + 1, j - 1, k), -x0 * vy / (2.0 * hx2) - y1 * vx / (2.0 * hx * hy));
                put(m, AY(i - 1, j, k), -x1 * vy / (2.0 * hx2) - y0 * vx / (2.0 * hx * hy));
                w = (x0 + x1) * vy / (2.0 * hx2) + (z0 + z1) * vy / (2.0 * hz2);
                if (tr) w = w + y0 / (hy * dt);
                put(m, AY(i, j, k), w);
                put(m, AY(i + 1, j, k), -x0 * vy / (2.0 * hx2) + y0 * vx / (2.0 * hx * hy));
                put(m, AY(i, j - 1, k + 1), -y1 * vz / (2.0 * hz * hy) - z0 * vy / (2.0 * hz2));
                put(m, AY(i, j, k + 1), y0 * vz / (2.0 * hz * hy) - z0 * vy / (2.0 * hz2));
                /* Az columns */
                put(m, AZ(i, j - 1, k - 1), -y1 * vz / (2.0 * hy2) + z1 * vy / (2.0 * hy * hz));
                put(m, AZ(i - 1, j, k - 1), -x1 * vz / (2.0 * hx2) + z1 * vx / (2.0 * hx * hz));
                w = (x0 + x1) * vz / (2.0 * hx2) + (y0 + y1) * vz / (2.0 * hy2);
                if (tr) w = w - z1 / (hz * dt);
                put(m, AZ(i, j, k - 1), w);
                put(m, AZ(i + 1, j, k - 1), -x0 * vz / (2.0 * hx2) - z1 * vx / (2.0 * hx * hz));
                put(m, AZ(i, j + 1, k - 1), -y0 * vz / (2.0 * hy2) - z1 * vy / (2.0 * hy * hz));
                put(m, AZ(i, j - 1, k), -y1 * vz / (2.0 * hy2) - z0 * vy / (2.0 * hy * hz));
                put(m, AZ(i - 1, j, k), -x1 * vz / (2.0 * hx2) - z0 * vx / (2.0 * hx * hz));
                w = (x0 + x1) * vz / (2.0 * hx2) + (y0 + y1) * vz / (2.0 * hy2);
                if (tr) w = w + z0 / (hz * dt);
                put(m, AZ(i, j, k), w);
                put(m, AZ(i + 1, j, k), -x0 * vz / (2.0 * hx2) + z0 * vx / (2.0 * hx * hz));
                put(m, AZ(i, j + 1, k), -y0 * vz / (2.0 * hy2) + z0 * vy / (2.0 * hy * hz));
                /* V columns: sigma-floored 7-point Laplacian, :1237-1281 */
                const double fx0 = fmax(x0, QO_SIGFLOOR), fx1 = fmax(x1, QO_SIGFLOOR);
                const double fy0 = fmax(y0, QO_SIGFLOOR), fy1 = fmax(y1, QO_SIGFLOOR);
                const double fz0 = fmax(z0, QO_SIGFLOOR), fz1 = fmax(z1, QO_SIGFLOOR);
                put(m, VV(i, j, k - 1), -fz1 / hz2);
                put(m, VV(i, j - 1, k), -fy1 / hy2);
                put(m, VV(i - 1, j, k), -fx1 / hx2);
                put(m, VV(i, j, k), (fx0 + fx1) / hx2 + (fy0 + fy1) / hy2 + (fz0 + fz1) / hz2);
                put(m, VV(i + 1, j, k), -fx0 / hx2);
                put(m, VV(i, j + 1, k), -fy0 / hy2);
                put(m, VV(i, j, k + 1), -fz0 / hz2);
            }
    ptr[row] = m->nz;
#undef mux
#undef muy
#undef muz
#undef sgx
#undef sgy
#undef sgz
#undef AX
#undef AY
#undef AZ
#undef VV
    return m->nz;
}

/* ------------------------------------------------------------------------ */
/* Unknown vector <-> node-sized fields, QMS_3D.f90:1320-1343               */
/* ------------------------------------------------------------------------ */

/* x (n_dims) -> fields (4 node arrays Ax, Ay, Az, V); entries outside each
 * component's sub-box are set to 0 (node-sized storage, :288-299). */
void qo_unpack(int cx, int cy, int cz, const double *x, double *fields) {
    qo_grid G, *g = &G;
    qo_grid_init(g, cx, cy, cz, 1, 1, 1);
    memset(fields, 0, sizeof(double) * 4 * (size_t)g->n1);
    double *ax = fields, *ay = fields + g->n1, *az = fields + 2 * g->n1, *vv = fields + 3 * g->n1;
    for (int k = 1; k <= g->Nz; ++k)
        for (int j = 1; j <= g->Ny; ++j)
            for (int i = 1; i <= g->Nx; ++i) {
                if (i <= cx) NODE(g, ax, i, j, k) = x[CAX(g, i, j, k)];
                if (j <= cy) NODE(g, ay, i, j, k) = x[CAY(g, i, j, k)];
                if (k <= cz) NODE(g, az, i, j, k) = x[CAZ(g, i, j, k)];
                NODE(g, vv, i, j, k) = x[CVV(g, i, j, k)];
            }
}

void qo_pack(int cx, int cy, int cz, const double *fields, double *x) {
    qo_grid G, *g = &G;
    qo_grid_init(g, cx, cy, cz, 1, 1, 1);
    const double *ax = fields, *ay = fields + g->n1, *az = fields + 2 * g->n1, *vv = fields + 3 * g->n1;
    for (int k = 1; k <= g->Nz; ++k)
        for (int j = 1; j <= g->Ny; ++j)
            for (int i = 1; i <= g->Nx; ++i) {
                if (i <= cx) x[CAX(g, i, j, k)] = NODE(g, ax, i, j, k);
                if (j <= cy) x[CAY(g, i, j, k)] = NODE(g, ay, i, j, k);
                if (k <= cz) x[CAZ(g, i, j, k)] = NODE(g, az, i, j, k);
                x[CVV(g, i, j, k)] = NODE(g, vv, i, j, k);
            }
}

/* ------------------------------------------------------------------------ */
/* Cell-centred H, QMS_3D.f90:1346-1376                                      */
/* ------------------------------------------------------------------------ */

/* fields: node arrays Ax, Ay, Az (V unused); H: 3 arrays of cx*cy*cz in
 * Fortran order (Hx, Hy, Hz). */
void qo_hfield(int cx, int cy, int cz, double hx, double hy, double hz,
               const double *coef, const double *fields, double *H) {
    qo_grid G, *g = &G;
    qo_grid_init(g, cx, cy, cz, hx, hy, hz);
    const int64_t n1 = g->n1, nc = (int64_t)cx * cy * cz;
    const double *MX = coef, *MY = coef + n1, *MZ = coef + 2 * n1;
    const double *Ax = fields, *Ay = fields + n1, *Az = fields + 2 * n1;
#define F(a, i, j, k) NODE(g, a, i, j, k)
    for (int k = 1; k <= cz; ++k)
        for (int j = 1; j <= cy; ++j)
            for (int i = 1; i <= cx; ++i) {
                int64_t p = (i - 1) + (int64_t)(j - 1) * cx + (int64_t)(k - 1) * cx * cy;
                H[p] = 0.5 * (((F(Az, i, j + 1, k) - F(Az, i, j, k)) / hy - (F(Ay, i, j, k + 1) - F(Ay, i, j, k)) / hz) / F(MX, i, j, k)
                            + ((F(Az, i + 1, j + 1, k) - F(Az, i + 1, j, k)) / hy - (F(Ay, i + 1, j, k + 1) - F(Ay, i + 1, j, k)) / hz) / F(MX, i + 1, j, k));
                H[nc + p] = 0.5 * (((F(Ax, i, j, k + 1) - F(Ax, i, j, k)) / hz - (F(Az, i + 1, j, k) - F(Az, i, j, k)) / hx) / F(MY, i, j, k)
                                 + ((F(Ax, i, j + 1, k + 1) - F(Ax, i, j + 1, k)) / hz - (F(Az, i + 1, j + 1, k) - F(Az, i, j + 1, k)) / hx) / F(MY, i, j + 1, k));
                H[2 * nc + p] = 0.5 * (((F(Ay, i + 1, j, k) - F(Ay, i, j, k)) / hx - (F(Ax, i, j + 1, k) - F(Ax, i, j, k)) / hy) / F(MZ, i, j, k)
                                     + ((F(Ay, i + 1, j, k + 1) - F(Ay, i, j, k + 1)) / hx - (F(Ax, i, j + 1, k + 1) - F(Ax, i, j, k + 1)) / hy) / F(MZ, i, j, k + 1));
            }
#undef F
}

/* ------------------------------------------------------------------------ */
/* Velocity schedule, QMS_3D.f90:2331-2350 and the first-step nudge          */
/* :1431-1449                                                                */
/* ------------------------------------------------------------------------ */

double qo_trajectory(double v0, double a, int n_a, int n, double dt) {
    if (n <= n_a) return v0 + a * n * dt;
    return v0 + a * n_a * dt;
}

/* Velocity used by transient step `counter` (1-based) for one axis.
 * v_prev is the velocity of the steady solve (trajectory at n = 0). */
double qo_step_velocity(double v0, double a, int n_a, int counter, double dt) {
    double nudge = 0.0;
    if (counter == 1) {
        double vs = qo_trajectory(v0, a, n_a, 0, dt);
        if (a * dt < v0 * QO_NUDGE) nudge = vs * QO_NUDGE / dt;
    }
    return qo_trajectory(v0, a + nudge, n_a, counter, dt);
}
