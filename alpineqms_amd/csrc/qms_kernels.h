// Internal interfaces between the HIP kernel files and the solver driver.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <functional>
#include "qms_grid.h"

namespace qms {

// Yee material coefficients, node-sized arrays.
struct Coeffs {
    double *im[3];   // 1/Mu_x, 1/Mu_y, 1/Mu_z
    double *sg[3];   // Sigma_x, Sigma_y, Sigma_z
    float *im32[3];
    float *sg32[3];
};
struct CoeffsD {
    const double *im[3];
    const double *sg[3];
    // Optional compressed form of the same six arrays: a 2-byte class per
    // node and a table of the distinct (1/Mu_x, 1/Mu_y, 1/Mu_z, Sigma_x,
    // Sigma_y, Sigma_z) tuples, 6 doubles per class (null when not built).
    const uint16_t *cls;
    const double *ctab;
};
struct CoeffsF {
    const float *im[3];
    const float *sg[3];
};

struct OpK;
struct OpParams {
    double v[3];   // frame velocity (solve_vx, solve_vy, solve_vz)
    double dt;     // del_t (transient rows only)
    int tr;        // 1: transient rows T(v,dt), 0: steady rows S(v)
    int adv;       // 1 if any velocity component is non-zero
    int zchunk;    // planes per operator workgroup (z-march depth)
    int colw;      // operator column width in nodes (64 or 32)
    // Device copy of make_opk(g, *this): when set, the operator, RHS and
    // Dirichlet kernels read their row constants from it instead of their
    // launch arguments (captured graphs then follow the per-step velocity).
    const OpK *kdev = nullptr;
};

// Per-launch constants of the operator rows, computed once on the host so
// that the kernels multiply instead of dividing (an FP64 division costs
// about ten dependent VALU instructions).
struct OpK {
    int sy, sz;                  // node strides along y and z (nx, nx*ny)
    int tr, adv;
    double ih2[3];               // 1/h_a^2
    double ixy, ixz, iyz;        // 1/(h_a h_b)
    double ih[3];                // 1/h_a
    double idt;                  // 1/dt (transient rows)
    double iht[3];               // 1/(h_a dt)
    double v1[3][3];             // v_a / (2 h_b)
    double w[12];                // v_a / (2 h_b h_c) for the (a,b,c) of the V row's advection terms
    double v[3];                 // frame velocity (Dirichlet values of V)
    double hv[3];                // v_a / 2 (edge average of the advection terms)
};
OpK make_opk(const Grid &g, const OpParams &op);

// Workgroup columns of the z-marching operator: tx x (256/tx) nodes
// (tx = 64 or 32), kz planes.
struct ZTiling {
    int ntx, nty, kz, tx;
};
ZTiling ztiling(const Grid &g, int kz, int tx, unsigned &ncol);

// Hardware workgroup index -> tile index such that each of the 8 XCDs (which
// receive workgroups round-robin) works on one contiguous range of tiles, and
// neighbouring tiles share that XCD's L2.
__device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned nb) {
    const unsigned q = nb >> 3, r = nb & 7u, x = b & 7u;
    return x * q + min(x, r) + (b >> 3);
}

// One design of a batched operator launch (device resident).
struct BatchItem {
    Grid g;
    ZTiling t;
    CoeffsD c;
    OpK k;
    const double *x;
    double *y;
};
// One segment of a batched launch: planes [kb, ke) of workgroup column `col`
// (x fastest, then y, as in ZTiling) of design d.  16 bytes, one scalar load.
struct BatchSeg {
    unsigned d, col;
    int kb, ke;
};
// slots workgroups (a multiple of 8); the workgroup of slot s runs segments
// [slot_off[s], slot_off[s+1]) in order, where workgroup b has slot
// (b % 8) * (slots / 8) + b / 8, so each XCD runs one contiguous slot range.
// slot_off_dev null: slots = the segment count, one workgroup per segment,
// each XCD running one contiguous range of segments.
// tx, cls: column width and coefficient form shared by every design, or tx = 0.
void launch_operator_batch(const BatchItem *items_dev, const BatchSeg *segs_dev, const unsigned *slot_off_dev,
                           unsigned slots, int tx, int cls, hipStream_t s);
// Workgroups of that kernel variant resident on the device at once (a
// multiple of 8).
int batch_resident_blocks(int tx, int cls);

enum { OP_RESID = 0, OP_APPLY = 1, OP_VRHS = 2 };

void launch_coeffs(const Grid &g, const uint16_t *slot, const double *mat, const Coeffs &c, hipStream_t s);
void launch_diag(const Grid &g, const CoeffsD &c, const OpParams &op, double *D, hipStream_t s);
void launch_rhs(const Grid &g, const CoeffsD &c, const BC &bc, const OpParams &op, const double *aprev, double *b,
                hipStream_t s);
void launch_setbc(const Grid &g, const BC &bc, const OpParams &op, double *x, hipStream_t s);
void launch_resid(const Grid &g, const CoeffsD &c, const OpParams &op, const double *x, const double *b, double *r,
                  hipStream_t s);
void launch_apply(const Grid &g, const CoeffsD &c, const OpParams &op, const double *x, double *y, hipStream_t s);
// The V-block right-hand side of the block-triangular preconditioner, in
// FP64: out[p] = D_V v_V - (M_VA z_A) on free V rows, 0 elsewhere (n1 entries;
// z in the A block's precision, v the FP64 Krylov vector).
void launch_vrhs(const Grid &g, const CoeffsD &c, const OpParams &op, const float *z, const double *v, double *out,
                 hipStream_t s);
void launch_vrhs(const Grid &g, const CoeffsD &c, const OpParams &op, const double *z, const double *v, double *out,
                 hipStream_t s);
// The same with v's V field given as its own FP64 array vv (n1 values): the
// V segment of a mixed-precision Krylov vector.
void launch_vrhs_seg(const Grid &g, const CoeffsD &c, const OpParams &op, const float *z, const double *vv,
                     double *out, hipStream_t s);
// y = D^-1 M x for an FP32-stored x (FP64 arithmetic and output)
void launch_apply_f32in(const Grid &g, const CoeffsD &c, const OpParams &op, const float *x, double *y, hipStream_t s);
// y = D^-1 M x with x stored in FP32 and y a mixed-precision Krylov vector
// (A rows in FP32, the V row in FP64 from byte offset offv); FP64 arithmetic
void launch_apply_f32(const Grid &g, const CoeffsD &c, const OpParams &op, const float *x, float *y, int64_t offv,
                      hipStream_t s);
void launch_hfield(const Grid &g, const CoeffsD &c, const double *x, double *H, int unit_mu, hipStream_t s);

// ---------------------------------------------------------------------------
// Multigrid hierarchy of one diagonal block of the preconditioner (the
// coupled A block or the V block).  Level 0 is evaluated on the fly from the
// coefficient arrays; levels >= 1 hold explicit Galerkin stencils in
// difference form:
//   (A x)_p = e_p x_p + sum_d c_{p,d} (x_p - x_{p+d}) + sum_s a_{p,s} y_{p+o_s}
// where d runs over the 6 axis neighbours of the same component and s over
// the 4+4 cross-component couplings (A block only).
// ---------------------------------------------------------------------------
constexpr int MG_MAX_LEVELS = 16;

struct LevelDims {
    int nx, ny, nz;
    int64_t n1;
};

// Grid constants of the level-0 rows, in the working precision.
template <typename T>
struct FineConst {
    T ih2[3];        // 1/h_a^2
    T ixy, ixz, iyz; // 1/(h_a h_b)
    T idt;           // 1/dt (0: steady, no mass term)
    T sigfloor;      // sigma floor of the V Laplacian
};

// Device coefficient arrays of one precision: 1/Mu_{x,y,z}, Sigma_{x,y,z}.
template <typename T>
struct CoeffsT {
    const T *im[3];
    const T *sg[3];
    // the same values as a 2-byte class per node and a table of FP64 tuples
    // (1/Mu_x, 1/Mu_y, 1/Mu_z, Sigma_x, Sigma_y, Sigma_z); null if not built
    const uint16_t *cls;
    const double *ctab;
};
// Host copies of the FP64 coefficient arrays (node-sized).
struct HostCoef {
    const double *im[3];
    const double *sg[3];
};

template <typename T>
struct MgLevel {
    LevelDims d{0, 0, 0, 0};
    int ncomp = 0;             // 3 (A block) or 1 (V block)
    T *cond = nullptr;         // [ncomp][6][n1] conductances (-x,+x,-y,+y,-z,+z)
    T *exc = nullptr;          // [ncomp][n1] excess diagonal
    T *cross = nullptr;        // [ncomp][8][n1] cross-component coefficients (A block)
    T *dinv = nullptr;         // [ncomp][n1] 1/diag, 0 on inactive rows
    // The same rows as a 2-byte class per node and a table of the distinct
    // row tuples (MG_CLS_RS values per component: cond[6], exc, dinv,
    // cross[8]); when set, cond/exc/cross/dinv are not allocated.
    uint16_t *cls = nullptr;
    T *tab = nullptr;
    int ncls = 0;
    T *x = nullptr, *b = nullptr, *r = nullptr, *t = nullptr;   // [ncomp][n1] (x, b unused on level 0)
    T *dg = nullptr;           // level 0: [ncomp][n1] row diagonals, 0 on inactive rows
};
// Table row layout of a class-compressed level: per component, values
// [0..5] conductances, [6] excess diagonal, [7] 1/diag (0: inactive row),
// [8..15] cross-component coefficients (A block only).
constexpr int MG_CLS_RS_A = 16, MG_CLS_RS_V = 8;

template <typename T>
struct Hierarchy {
    int block = 0;             // 0: A block (3 components), 1: V block
    int nlev = 0;
    Grid g{};
    const T *im[3] = {nullptr, nullptr, nullptr};
    const T *sg[3] = {nullptr, nullptr, nullptr};
    FineConst<T> fk{};
    const uint16_t *cls = nullptr;   // coefficient classes (level-0 A-block z-march)
    const double *ctab = nullptr;
    ZTiling zt{};                    // level-0 z-march columns (32 x 8 nodes, zt.kz planes)
    unsigned zcols = 0;
    MgLevel<T> lev[MG_MAX_LEVELS];
    int ncoarse = 0;           // active unknowns of the coarsest level
    int *coarse_idx = nullptr; // [ncoarse] -> comp*n1 + p on the coarsest level
    double *coarse_inv = nullptr;   // [ncoarse*ncoarse] row-major FP64 inverse
    bool fused_restrict = false;    // the last V-cycle restricted level 0's residual in the z-march
};

struct MGParams {
    int nu;          // pre/post smoothing sweeps on level 0
    int nu_coarse;   // pre/post smoothing sweeps on the coarse levels
    double omega;    // Jacobi damping
    double alpha;    // coarse-correction scaling
    int tail;        // run the level above the coarsest and the coarsest solve as one launch
    double cheb_lmax;   // > 0: Chebyshev-Richardson weights on [cheb_lmax / cheb_ratio, cheb_lmax]
    double cheb_ratio;  //   instead of omega
    int fuse_restrict;  // level 0's residual, restriction and level 1's first sweep in one launch
};

// Weight of sweep it (0-based) of a level with nu sweeps, before (post =
// false) or after the coarse correction: damped Jacobi's omega, or 1 /
// theta_k with theta_k = (hi + lo) / 2 + (hi - lo) / 2 cos((2k + 1) pi / 2nu),
// the roots of the degree-nu Chebyshev polynomial on [lo, hi], k = it before
// the correction (largest root first) and nu - 1 - it after it.  nu sweeps
// then apply the Chebyshev residual polynomial of D^-1 A, the one of degree
// nu that is smallest on [lo, hi] among those equal to 1 at 0.
double sweep_weight(const MGParams &mp, int nu, int it, bool post);

// Device allocations of the hierarchy go through the owner (which frees them).
struct MgAlloc {
    std::function<void *(size_t)> alloc;
    hipStream_t stream;
};

FineConst<double> fine_const(const Grid &g, double dt);
// Builds the coarse levels on the host in FP64 from the host coefficients and
// uploads them in precision T; dt = 0 builds the steady operator.
template <typename T>
void mg_setup(const Grid &g, const HostCoef &hc, const CoeffsT<T> &dc, double dt, int block, const MgAlloc &al,
              Hierarchy<T> &h);
// x0 <- one V-cycle applied to b0 (x0, b0: [ncomp][n1] of level 0).
// first_done: level 0's first sweep from zero is already in
// mg_first_target(h, mp, x0) (written by launch_d2t_first / launch_first_sweep).
template <typename T>
void mg_vcycle(Hierarchy<T> &h, const MGParams &mp, T *x0, const T *b0, hipStream_t s, bool first_done = false);
// Where the V-cycle reads level 0's first sweep from (with first_done).
template <typename T>
T *mg_first_target(Hierarchy<T> &h, const MGParams &mp, T *x0);
// Level 0's first sweep from zero as a stream: xf = omega b / diag (0 on
// inactive rows), b = the level-0 right-hand side.
template <typename T>
void launch_first_sweep(const Hierarchy<T> &h, const MGParams &mp, const T *b, T *xf, hipStream_t s);
// dst = src * scale (the A block's right-hand side D_A v_A) together with
// level 0's first sweep from zero on it, xf = omega dst / diag.
template <typename TS, typename T>
void launch_d2t_first(const Hierarchy<T> &h, const MGParams &mp, const TS *src, const double *scale, T *dst, T *xf,
                      hipStream_t s);
// One level-0 smoother sweep (for kernel timing).
template <typename T>
void mg_fine_sweep(Hierarchy<T> &h, const MGParams &mp, const T *x, const T *b, T *xo, hipStream_t s);
template <typename T>
void launch_fill(T *x, int64_t n, T v, hipStream_t s);

// conversions / vector kernels at the Krylov / multigrid boundary
void launch_cvt(const double *src, float *dst, int64_t n, hipStream_t s);

// Krylov helpers.  A Krylov vector has two segments: the A components (3 n1
// values of type TE from byte 0) and the V component (n1 FP64 values from
// byte offset offv).  TE = double with offv = 24 n1 is a flat FP64 vector;
// TE = float is the mixed-precision basis.  Dots, norms and coefficients are
// FP64.
template <typename TE>
void launch_multidot(const TE *const *V, int nv, const TE *w, int64_t n1, int64_t offv, double *partial, int nblk,
                     hipStream_t s);
void launch_reduce_partials(const double *partial, int nv, int nblk, double *out, hipStream_t s);
template <typename TE>
void launch_orth_update(TE *const *V, int nv, const double *h, TE *w, int64_t n1, int64_t offv, double *partial,
                        int nblk, hipStream_t s);
// dst (two-segment) = src (flat FP64, 4 n1 values) * a
template <typename TE>
void launch_scale(TE *dst, const double *src, double a, int64_t n1, int64_t offv, hipStream_t s);
void launch_axpy_list(double *x, const double *const *Z, int nz, const double *y_dev, int64_t n, hipStream_t s);
void launch_axpy_list(double *x, const float *const *Z, int nz, const double *y_dev, int64_t n, hipStream_t s);
template <typename TE>
void launch_scale_dev(TE *dst, const TE *src, const double *a_dev, int64_t n1, int64_t offv, hipStream_t s);
// FGMRES cycle scalars on the device (st[KS_*])
// KS_FLOOR: below this estimate a cycle that stops improving ends
// (KS_STALL = 1); KS_HIST..KS_HIST+3: the estimates of the last four
// iterations (ring by iteration index).
enum {
    KS_BNORM = 0, KS_TOL = 1, KS_INVH = 2, KS_EST = 3, KS_K = 4, KS_ITERS = 5, KS_FLOOR = 6, KS_STALL = 7,
    KS_HIST = 8, KS_COUNT = 12
};
void launch_cycle_start(int m, double beta, double bnorm, double tol, double floor, double *g, double *st,
                        hipStream_t s);
void launch_givens(int j, int m, const double *hcol, double *H, double *cs, double *sn, double *g, double *st,
                   hipStream_t s);
void launch_backsub(int k, int m, const double *H, const double *g, double *y, hipStream_t s);
void launch_sqrt_to(const double *src, double *dst, hipStream_t s);
void launch_extrapolate(double *x, double *h1, double *h2, int nh, int64_t n, hipStream_t s);
void launch_norm2(const double *w, int64_t n, double *partial, int nblk, hipStream_t s);
int reduce_blocks(int64_t n);
// Sets the message qms_last_error(NULL) returns (for C-ABI calls without a handle).
void set_last_error(const char *msg);
void launch_copy(double *dst, const double *src, int64_t n, hipStream_t s);

}  // namespace qms
