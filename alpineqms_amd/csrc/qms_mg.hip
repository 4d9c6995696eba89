// Multigrid for the two diagonal blocks of the solver's block-triangular
// preconditioner (the linear solve replaces PARDISO, QMS_3D.f90:1314-1317,
// :2193-2203):
//   A block: K + Sigma/dt on the free Ax, Ay, Az unknowns -- curl-curl plus
//            the component-wise gauge penalty plus the mass term of the
//            reference's A rows (QMS_3D.f90:476-607, :650-781, :824-955,
//            :1504, :1681, :1857), without the frame-velocity terms; the three
//            components stay coupled through the curl-curl cross terms.
//   V block: the sigma-floored Laplacian of the V rows (:1237-1281).
// Couplings to fixed (Dirichlet) unknowns are eliminated: the preconditioner
// acts on corrections, which vanish there.
//
// Coarse spaces: 2x2x2 aggregation of each component's active unknowns
// (coarse index = fine index >> 1), Galerkin coarse operators.  This
// aggregation keeps the fine pattern (7-point self stencil plus the 4+4
// cross-component offsets), so every coarse level is one structured stencil
// in difference form:
//   (A x)_p = e_p x_p + sum_d c_{p,d} (x_p - x_{p+d}) + sum_s a_{p,s} y_{p+o_s}
// which keeps the relative accuracy of 1e-3 S/m background links next to
// 6e7 S/m conductors in FP32.
//
// The coarse operators are built once per design on the host in FP64 (from the
// FP64 Yee coefficients) and uploaded in the working precision; the device
// runs only element-wise stencil kernels (smoother, residual, transfers) and a
// one-workgroup dense solve on the coarsest level.  Level 0 is evaluated on
// the fly from the coefficient arrays.
//
// Every stencil table below is a constexpr function of compile-time indices
// inside fully unrolled loops, so no kernel indexes a private array at run
// time (none needs scratch memory).
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>
#include <algorithm>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>
#include "qms_kernels.h"

// Multiply-adds are fused only within one source expression, never across
// statements: the same row code then rounds identically in every kernel that
// inlines it (the one-workgroup tail is bit for bit its launch-per-step form).
#pragma clang fp contract(on)

namespace qms {

// ---------------------------------------------------------------------------
// Stencil tables
// ---------------------------------------------------------------------------
// Self neighbour d of a node: -x, +x, -y, +y, -z, +z.
QMS_HD constexpr int so_i(int d) { return d == 0 ? -1 : d == 1 ? 1 : 0; }
QMS_HD constexpr int so_j(int d) { return d == 2 ? -1 : d == 3 ? 1 : 0; }
QMS_HD constexpr int so_k(int d) { return d == 4 ? -1 : d == 5 ? 1 : 0; }

// Cross-component coupling slot s of component c: slots 0-3 couple to the
// first other component, 4-7 to the second, at the offsets where the
// reference's rows read them (Ax row :546-592, Ay row :727-770, Az row
// :844-888).  Encoded as tc*64 + (di+1)*16 + (dj+1)*4 + (dk+1).
QMS_HD constexpr int xo_code(int c, int s) {
    constexpr int T[3][8] = {
        // Ax row: Ay at (0,-1,0) (1,-1,0) (0,0,0) (1,0,0); Az at (0,0,-1) (1,0,-1) (0,0,0) (1,0,0)
        {1 * 64 + 1 * 16 + 0 * 4 + 1, 1 * 64 + 2 * 16 + 0 * 4 + 1, 1 * 64 + 1 * 16 + 1 * 4 + 1, 1 * 64 + 2 * 16 + 1 * 4 + 1,
         2 * 64 + 1 * 16 + 1 * 4 + 0, 2 * 64 + 2 * 16 + 1 * 4 + 0, 2 * 64 + 1 * 16 + 1 * 4 + 1, 2 * 64 + 2 * 16 + 1 * 4 + 1},
        // Ay row: Ax at (-1,0,0) (0,0,0) (-1,1,0) (0,1,0); Az at (0,0,-1) (0,1,-1) (0,0,0) (0,1,0)
        {0 * 64 + 0 * 16 + 1 * 4 + 1, 0 * 64 + 1 * 16 + 1 * 4 + 1, 0 * 64 + 0 * 16 + 2 * 4 + 1, 0 * 64 + 1 * 16 + 2 * 4 + 1,
         2 * 64 + 1 * 16 + 1 * 4 + 0, 2 * 64 + 1 * 16 + 2 * 4 + 0, 2 * 64 + 1 * 16 + 1 * 4 + 1, 2 * 64 + 1 * 16 + 2 * 4 + 1},
        // Az row: Ax at (-1,0,0) (0,0,0) (-1,0,1) (0,0,1); Ay at (0,-1,0) (0,0,0) (0,-1,1) (0,0,1)
        {0 * 64 + 0 * 16 + 1 * 4 + 1, 0 * 64 + 1 * 16 + 1 * 4 + 1, 0 * 64 + 0 * 16 + 1 * 4 + 2, 0 * 64 + 1 * 16 + 1 * 4 + 2,
         1 * 64 + 1 * 16 + 0 * 4 + 1, 1 * 64 + 1 * 16 + 1 * 4 + 1, 1 * 64 + 1 * 16 + 0 * 4 + 2, 1 * 64 + 1 * 16 + 1 * 4 + 2}};
    return T[c][s];
}
QMS_HD constexpr int xo_c(int c, int s) { return xo_code(c, s) >> 6; }
QMS_HD constexpr int xo_i(int c, int s) { return ((xo_code(c, s) >> 4) & 3) - 1; }
QMS_HD constexpr int xo_j(int c, int s) { return ((xo_code(c, s) >> 2) & 3) - 1; }
QMS_HD constexpr int xo_k(int c, int s) { return (xo_code(c, s) & 3) - 1; }

QMS_HD int64_t lidx(const LevelDims &d, int i, int j, int k) {
    return (int64_t)i + (int64_t)d.nx * ((int64_t)j + (int64_t)d.ny * k);
}
QMS_HD bool inside(const LevelDims &d, int i, int j, int k) {
    return i >= 0 && j >= 0 && k >= 0 && i < d.nx && j < d.ny && k < d.nz;
}

// ---------------------------------------------------------------------------
// Level-0 rows from the Yee coefficients (shared by the device kernels and
// the host Galerkin build).  IM(a, di, dj, dk) is 1/Mu_a at the node offset,
// SG(a) is Sigma_a at the node.
// ---------------------------------------------------------------------------
template <typename T>
struct Row {
    T cond[6];
    T exc;
    T cross[8];
};

// Coefficient access at node (i,j,k) of a node-sized array.
template <typename T>
struct CoefAt {
    const T *im[3];
    const T *sg[3];
    int64_t p;
    int sy, sz;
    QMS_HD T IM(int a, int di, int dj, int dk) const { return im[a][p + di + (int64_t)dj * sy + (int64_t)dk * sz]; }
    QMS_HD T SG(int a, int di, int dj, int dk) const { return sg[a][p + di + (int64_t)dj * sy + (int64_t)dk * sz]; }
};

// A-block row of component C at a free node (see the file comment), before
// the Dirichlet elimination.  The values are the reference's coefficients:
// Ax row :476-607 (+ :1504), Ay row :650-781 (+ :1681), Az row :824-955
// (+ :1857).
template <int C, typename T, typename CA>
QMS_HD void row_a_raw(const CA &ca, const FineConst<T> &fk, Row<T> &r) {
    if (C == 0) {
        const T a = ca.IM(2, 0, -1, 0) * fk.ixy, b = ca.IM(2, 0, 0, 0) * fk.ixy;
        const T e = ca.IM(1, 0, 0, -1) * fk.ixz, f = ca.IM(1, 0, 0, 0) * fk.ixz;
        r.cond[0] = ca.IM(0, -1, 0, 0) * fk.ih2[0]; r.cond[1] = ca.IM(0, 0, 0, 0) * fk.ih2[0];
        r.cond[2] = ca.IM(2, 0, -1, 0) * fk.ih2[1]; r.cond[3] = ca.IM(2, 0, 0, 0) * fk.ih2[1];
        r.cond[4] = ca.IM(1, 0, 0, -1) * fk.ih2[2]; r.cond[5] = ca.IM(1, 0, 0, 0) * fk.ih2[2];
        r.cross[0] = a; r.cross[1] = -a; r.cross[2] = -b; r.cross[3] = b;
        r.cross[4] = e; r.cross[5] = -e; r.cross[6] = -f; r.cross[7] = f;
        r.exc = ca.SG(0, 0, 0, 0) * fk.idt;
    } else if (C == 1) {
        const T a = ca.IM(2, -1, 0, 0) * fk.ixy, b = ca.IM(2, 0, 0, 0) * fk.ixy;
        const T e = ca.IM(0, 0, 0, -1) * fk.iyz, f = ca.IM(0, 0, 0, 0) * fk.iyz;
        r.cond[0] = ca.IM(2, -1, 0, 0) * fk.ih2[0]; r.cond[1] = ca.IM(2, 0, 0, 0) * fk.ih2[0];
        r.cond[2] = ca.IM(1, 0, -1, 0) * fk.ih2[1]; r.cond[3] = ca.IM(1, 0, 0, 0) * fk.ih2[1];
        r.cond[4] = ca.IM(0, 0, 0, -1) * fk.ih2[2]; r.cond[5] = ca.IM(0, 0, 0, 0) * fk.ih2[2];
        r.cross[0] = a; r.cross[1] = -b; r.cross[2] = -a; r.cross[3] = b;
        r.cross[4] = e; r.cross[5] = -e; r.cross[6] = -f; r.cross[7] = f;
        r.exc = ca.SG(1, 0, 0, 0) * fk.idt;
    } else {
        const T a = ca.IM(1, -1, 0, 0) * fk.ixz, b = ca.IM(1, 0, 0, 0) * fk.ixz;
        const T e = ca.IM(0, 0, -1, 0) * fk.iyz, f = ca.IM(0, 0, 0, 0) * fk.iyz;
        r.cond[0] = ca.IM(1, -1, 0, 0) * fk.ih2[0]; r.cond[1] = ca.IM(1, 0, 0, 0) * fk.ih2[0];
        r.cond[2] = ca.IM(0, 0, -1, 0) * fk.ih2[1]; r.cond[3] = ca.IM(0, 0, 0, 0) * fk.ih2[1];
        r.cond[4] = ca.IM(2, 0, 0, -1) * fk.ih2[2]; r.cond[5] = ca.IM(2, 0, 0, 0) * fk.ih2[2];
        r.cross[0] = a; r.cross[1] = -b; r.cross[2] = -a; r.cross[3] = b;
        r.cross[4] = e; r.cross[5] = -f; r.cross[6] = -e; r.cross[7] = f;
        r.exc = ca.SG(2, 0, 0, 0) * fk.idt;
    }
}

// V-block row at a free node: floored-sigma Laplacian (:1237-1281).
template <typename T, typename CA>
QMS_HD void row_v_raw(const CA &ca, const FineConst<T> &fk, Row<T> &r) {
    const T fl = fk.sigfloor;
    auto S = [&](int a, int di, int dj, int dk) {
        const T s = ca.SG(a, di, dj, dk);
        return s > fl ? s : fl;
    };
    r.cond[0] = S(0, -1, 0, 0) * fk.ih2[0]; r.cond[1] = S(0, 0, 0, 0) * fk.ih2[0];
    r.cond[2] = S(1, 0, -1, 0) * fk.ih2[1]; r.cond[3] = S(1, 0, 0, 0) * fk.ih2[1];
    r.cond[4] = S(2, 0, 0, -1) * fk.ih2[2]; r.cond[5] = S(2, 0, 0, 0) * fk.ih2[2];
    r.exc = (T)0;
#pragma unroll
    for (int s = 0; s < 8; ++s) r.cross[s] = (T)0;
}

// Dirichlet elimination: a coupling to a fixed neighbour moves into the
// excess diagonal (self) or is dropped (cross).
template <int C, int NC, typename T>
QMS_HD void eliminate(const Grid &g, int i, int j, int k, Row<T> &r) {
    constexpr int F = NC == 1 ? FV : C;
#pragma unroll
    for (int d = 0; d < 6; ++d)
        if (!free_f(g, F, i + so_i(d), j + so_j(d), k + so_k(d))) { r.exc += r.cond[d]; r.cond[d] = (T)0; }
    if (NC == 3) {
#pragma unroll
        for (int s = 0; s < 8; ++s)
            if (!free_f(g, xo_c(C, s), i + xo_i(C, s), j + xo_j(C, s), k + xo_k(C, s))) r.cross[s] = (T)0;
    }
}

// Coefficient access at node (i,j,k) through the class table: the class of
// the node at the offset selects a row (1/Mu_x, 1/Mu_y, 1/Mu_z, Sigma_x,
// Sigma_y, Sigma_z) of the FP64 table.  The values are the arrays' values
// (the FP32 arrays hold the same doubles rounded to float).
template <typename T>
struct CoefCls {
    const uint16_t *cls;
    const double *ctab;
    int64_t p;
    int sy, sz;
    QMS_HD const double *tuple(int di, int dj, int dk) const {
        return ctab + 6 * (int)cls[p + di + (int64_t)dj * sy + (int64_t)dk * sz];
    }
    QMS_HD T IM(int a, int di, int dj, int dk) const { return (T)tuple(di, dj, dk)[a]; }
    QMS_HD T SG(int a, int di, int dj, int dk) const { return (T)tuple(di, dj, dk)[3 + a]; }
};

// ---------------------------------------------------------------------------
// Row providers for the device kernels
// ---------------------------------------------------------------------------
// Level 0: rows evaluated from the coefficient arrays, or (CLS) from the
// class table.
template <typename T, int NC, bool CLS>
struct FineProv {
    Grid g;
    const T *im[3];
    const T *sg[3];
    const uint16_t *cls;
    const double *ctab;
    FineConst<T> fk;
    __device__ __forceinline__ LevelDims dims() const { return LevelDims{g.nx, g.ny, g.nz, g.n1}; }
    template <int C>
    __device__ __forceinline__ bool active(int i, int j, int k) const { return free_f(g, NC == 1 ? FV : C, i, j, k); }
    template <int C>
    __device__ __forceinline__ bool row(int i, int j, int k, Row<T> &r) const {
        if (!active<C>(i, j, k)) return false;
        if constexpr (CLS) {
            const CoefCls<T> ca{cls, ctab, gidx(g, i, j, k), g.nx, g.nx * g.ny};
            if (NC == 3) row_a_raw<C, T>(ca, fk, r);
            else row_v_raw<T>(ca, fk, r);
        } else {
            const CoefAt<T> ca{{im[0], im[1], im[2]}, {sg[0], sg[1], sg[2]}, gidx(g, i, j, k), g.nx, g.nx * g.ny};
            if (NC == 3) row_a_raw<C, T>(ca, fk, r);
            else row_v_raw<T>(ca, fk, r);
        }
        eliminate<C, NC, T>(g, i, j, k, r);
        return true;
    }
};

// Levels >= 1 compressed: a 2-byte class per node selects a row tuple
// (layout MG_CLS_RS_*) of the level's table.
template <typename T, int NC>
struct ClsProv {
    static constexpr int RS = NC == 3 ? MG_CLS_RS_A : MG_CLS_RS_V;
    LevelDims d;
    const uint16_t *cls;
    const T *tab;
    __device__ __forceinline__ LevelDims dims() const { return d; }
    __device__ __forceinline__ const T *tuple(int C, int i, int j, int k) const {
        return tab + (int64_t)cls[lidx(d, i, j, k)] * (NC * RS) + C * RS;
    }
    template <int C>
    __device__ __forceinline__ bool active(int i, int j, int k) const { return tuple(C, i, j, k)[7] != (T)0; }
    template <int C>
    __device__ __forceinline__ bool row(int i, int j, int k, Row<T> &r) const {
        const T *t = tuple(C, i, j, k);
        if (t[7] == (T)0) return false;
#pragma unroll
        for (int q = 0; q < 6; ++q) r.cond[q] = t[q];
        r.exc = t[6];
        if (NC == 3) {
#pragma unroll
            for (int s = 0; s < 8; ++s) r.cross[s] = t[8 + s];
        }
        return true;
    }
    template <int C>
    __device__ __forceinline__ bool row_all(int i, int j, int k, Row<T> &r) const {
        const T *t = tuple(C, i, j, k);
#pragma unroll
        for (int q = 0; q < 6; ++q) r.cond[q] = t[q];
        r.exc = t[6];
#pragma unroll
        for (int s = 0; s < 8; ++s) r.cross[s] = NC == 3 ? t[8 + s] : (T)0;
        return t[7] != (T)0;
    }
};

// Levels >= 1: explicit arrays.
template <typename T, int NC>
struct ExplProv {
    LevelDims d;
    const T *cond, *exc, *cross, *dinv;
    __device__ __forceinline__ LevelDims dims() const { return d; }
    template <int C>
    __device__ __forceinline__ bool active(int i, int j, int k) const {
        return dinv[(int64_t)C * d.n1 + lidx(d, i, j, k)] != (T)0;
    }
    template <int C>
    __device__ __forceinline__ bool row(int i, int j, int k, Row<T> &r) const {
        const int64_t p = lidx(d, i, j, k);
        if (dinv[(int64_t)C * d.n1 + p] == (T)0) return false;
#pragma unroll
        for (int q = 0; q < 6; ++q) r.cond[q] = cond[((int64_t)C * 6 + q) * d.n1 + p];
        r.exc = exc[(int64_t)C * d.n1 + p];
        if (NC == 3) {
#pragma unroll
            for (int s = 0; s < 8; ++s) r.cross[s] = cross[((int64_t)C * 8 + s) * d.n1 + p];
        }
        return true;
    }
    // The row loaded whatever its state; returns whether it is active.
    template <int C>
    __device__ __forceinline__ bool row_all(int i, int j, int k, Row<T> &r) const {
        const int64_t p = lidx(d, i, j, k);
#pragma unroll
        for (int q = 0; q < 6; ++q) r.cond[q] = cond[((int64_t)C * 6 + q) * d.n1 + p];
        r.exc = exc[(int64_t)C * d.n1 + p];
#pragma unroll
        for (int s = 0; s < 8; ++s) r.cross[s] = NC == 3 ? cross[((int64_t)C * 8 + s) * d.n1 + p] : (T)0;
        return dinv[(int64_t)C * d.n1 + p] != (T)0;
    }
};

// (A x) of component C at (i,j,k).  Only couplings with a non-zero
// coefficient read their neighbour, and every read is checked against the
// level's box.
template <int C, int NC, typename T>
__device__ __forceinline__ T apply_row(const LevelDims &d, const T *__restrict__ x, int i, int j, int k, const Row<T> &r,
                                       T &diag) {
    const T x0 = x[(int64_t)C * d.n1 + lidx(d, i, j, k)];
    T y = r.exc * x0;
    diag = r.exc;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        diag += r.cond[q];
        const int ni = i + so_i(q), nj = j + so_j(q), nk = k + so_k(q);
        if (r.cond[q] != (T)0 && inside(d, ni, nj, nk))
            y += r.cond[q] * (x0 - x[(int64_t)C * d.n1 + lidx(d, ni, nj, nk)]);
    }
    if (NC == 3) {
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int ni = i + xo_i(C, s), nj = j + xo_j(C, s), nk = k + xo_k(C, s);
            if (r.cross[s] != (T)0 && inside(d, ni, nj, nk))
                y += r.cross[s] * x[(int64_t)xo_c(C, s) * d.n1 + lidx(d, ni, nj, nk)];
        }
    }
    return y;
}

template <typename T>
__device__ __forceinline__ T row_diag(const Row<T> &r) {
    T d = r.exc;
#pragma unroll
    for (int q = 0; q < 6; ++q) d += r.cond[q];
    return d;
}

__device__ __forceinline__ bool node_of(const LevelDims &d, int64_t p, int &i, int &j, int &k) {
    if (p >= d.n1) return false;
    i = (int)(p % d.nx);
    j = (int)((p / d.nx) % d.ny);
    k = (int)(p / ((int64_t)d.nx * d.ny));
    return true;
}

// ---------------------------------------------------------------------------
// Kernels (one thread per node, all components of the block)
// ---------------------------------------------------------------------------
// Weighted Jacobi (Richardson) sweep: xo = x + w (b - A x) / diag, w the sweep's weight
// (fixed damping or a Chebyshev weight); FIRST: x = 0, xo = w b / diag.
// Inactive rows get 0.
template <int C, int NC, bool FIRST, typename T, typename P>
__device__ __forceinline__ void smooth_comp(const P &pv, const LevelDims &d, int64_t p, int i, int j, int k,
                                            const T *__restrict__ x, const T *__restrict__ b, T *__restrict__ xo,
                                            T omega) {
    Row<T> r;
    const int64_t q = (int64_t)C * d.n1 + p;
    if (!pv.template row<C>(i, j, k, r)) { xo[q] = (T)0; return; }
    if (FIRST) {
        xo[q] = omega * b[q] / row_diag(r);
    } else {
        T diag;
        const T ax = apply_row<C, NC, T>(d, x, i, j, k, r, diag);
        xo[q] = x[q] + omega * (b[q] - ax) / diag;
    }
}

template <typename T, typename P, int NC, bool FIRST>
__global__ void __launch_bounds__(256) k_mg_smooth(P pv, const T *__restrict__ x, const T *__restrict__ b,
                                                   T *__restrict__ xo, T omega) {
    const LevelDims d = pv.dims();
    int i, j, k;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (!node_of(d, p, i, j, k)) return;
    smooth_comp<0, NC, FIRST, T>(pv, d, p, i, j, k, x, b, xo, omega);
    if (NC == 3) {
        smooth_comp<1, NC, FIRST, T>(pv, d, p, i, j, k, x, b, xo, omega);
        smooth_comp<2, NC, FIRST, T>(pv, d, p, i, j, k, x, b, xo, omega);
    }
}

// Row diagonals of level 0 (0 on inactive rows), computed once per
// hierarchy: the divisor of the first sweep from zero, which then reads only
// b and this array.
template <int C, typename T, typename P>
__device__ __forceinline__ T diag_comp(const P &pv, int i, int j, int k) {
    Row<T> r;
    return pv.template row<C>(i, j, k, r) ? row_diag(r) : (T)0;
}

template <typename T, typename P, int NC>
__global__ void __launch_bounds__(256) k_mg_diag(P pv, T *__restrict__ dg) {
    const LevelDims d = pv.dims();
    int i, j, k;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (!node_of(d, p, i, j, k)) return;
    dg[p] = diag_comp<0, T>(pv, i, j, k);
    if constexpr (NC == 3) {
        dg[d.n1 + p] = diag_comp<1, T>(pv, i, j, k);
        dg[2 * d.n1 + p] = diag_comp<2, T>(pv, i, j, k);
    }
}

// The first sweep from zero, xf = omega b / diag (0 on inactive rows): the
// expression of k_mg_smooth's FIRST form, as a stream.
template <typename T>
__global__ void __launch_bounds__(256) k_first_sweep(const T *__restrict__ b, const T *__restrict__ dg, T omega,
                                                     T *__restrict__ xf, int64_t n) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const T d = dg[p];
    xf[p] = d != (T)0 ? omega * b[p] / d : (T)0;
}

// dst = src * scale (the FP64 product, rounded to T) and the first sweep from
// zero on it.
template <typename TS, typename T>
__global__ void __launch_bounds__(256) k_d2t_first(const TS *__restrict__ src, const double *__restrict__ scale,
                                                   const T *__restrict__ dg, T omega, T *__restrict__ dst,
                                                   T *__restrict__ xf, int64_t n) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const T bv = (T)((double)src[p] * scale[p]);
    const T d = dg[p];
    dst[p] = bv;
    xf[p] = d != (T)0 ? omega * bv / d : (T)0;
}

template <int C, int NC, typename T, typename P>
__device__ __forceinline__ void resid_comp(const P &pv, const LevelDims &d, int64_t p, int i, int j, int k,
                                           const T *__restrict__ x, const T *__restrict__ b, T *__restrict__ r) {
    Row<T> row;
    const int64_t q = (int64_t)C * d.n1 + p;
    if (!pv.template row<C>(i, j, k, row)) { r[q] = (T)0; return; }
    T diag;
    r[q] = b[q] - apply_row<C, NC, T>(d, x, i, j, k, row, diag);
}

template <typename T, typename P, int NC>
__global__ void __launch_bounds__(256) k_mg_resid(P pv, const T *__restrict__ x, const T *__restrict__ b,
                                                  T *__restrict__ r) {
    const LevelDims d = pv.dims();
    int i, j, k;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (!node_of(d, p, i, j, k)) return;
    resid_comp<0, NC, T>(pv, d, p, i, j, k, x, b, r);
    if (NC == 3) {
        resid_comp<1, NC, T>(pv, d, p, i, j, k, x, b, r);
        resid_comp<2, NC, T>(pv, d, p, i, j, k, x, b, r);
    }
}

// ---------------------------------------------------------------------------
// Coarse levels (explicit or class-compressed rows), branch-free: these
// levels are small, so a launch costs the latency of one thread's chain of
// loads rather than bandwidth.  Every neighbour index is clamped into the
// level's box -- a coupling to a node outside the box is zero after the
// Dirichlet elimination (the Galerkin build checks this), so the clamped
// value is multiplied by 0 -- every load is unconditional, and an inactive
// row is selected to 0 at the end.  The loads of all components can then be
// in flight together.  The arithmetic is that of k_mg_smooth / k_mg_resid:
// a skipped zero term there adds a zero product here.
// MODE 0: weighted Jacobi sweep, 1: first sweep from x = 0, 2: residual.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int64_t lidx_clamped(const LevelDims &d, int i, int j, int k) {
    return lidx(d, max(0, min(i, d.nx - 1)), max(0, min(j, d.ny - 1)), max(0, min(k, d.nz - 1)));
}

// The coarse-grid correction of the next level, applied where a value is
// read: x + alpha * xc[parent].  It is added whatever the node's state: an
// inactive node's value only meets zero couplings (a coupling to a node of
// no active row is zero after the elimination), and an inactive row's own
// output is 0.  For active nodes the value is the prolongation's own.
template <typename T>
struct Prolong {
    const T *xc;       // null: no correction
    LevelDims cd;
    T alpha;
    // PRO: the kernel variant that applies the correction (compile time, so
    // the plain sweeps carry no test or registers for it)
    template <bool PRO>
    __device__ __forceinline__ T at(const T *__restrict__ x, int64_t q, int c, int i, int j, int k) const {
        if constexpr (PRO) return x[q] + alpha * xc[(int64_t)c * cd.n1 + lidx(cd, i >> 1, j >> 1, k >> 1)];
        else return x[q];
    }
};

// The output of component C's row from the loaded row r (act: the row is
// active), the right-hand side entry bq and X(c, i, j, k), the value of
// component c at a node, which clamps the node into the box.  Shared by the
// one-launch-per-sweep kernels and the one-workgroup tail of the V-cycle.
template <int C, int NC, int MODE, typename T, typename XF>
__device__ __forceinline__ T cop_eval(const Row<T> &r, bool act, const XF &X, int i, int j, int k, T bq, T omega) {
    T v;
    if (MODE == 1) {
        v = omega * bq / row_diag(r);
    } else {
        const T x0 = X(C, i, j, k);
        T y = r.exc * x0;
        T diag = r.exc;
#pragma unroll
        for (int e = 0; e < 6; ++e) {
            diag += r.cond[e];
            y += r.cond[e] * (x0 - X(C, i + so_i(e), j + so_j(e), k + so_k(e)));
        }
        if (NC == 3) {
#pragma unroll
            for (int s = 0; s < 8; ++s) y += r.cross[s] * X(xo_c(C, s), i + xo_i(C, s), j + xo_j(C, s), k + xo_k(C, s));
        }
        v = MODE == 0 ? x0 + omega * (bq - y) / diag : bq - y;
    }
    return act ? v : (T)0;
}

// Component C's output at a node of a coarse level (stored by the caller, so
// that every component's loads are issued before the first store, which the
// compiler cannot move loads of the non-restrict class table across).
template <int C, int NC, int MODE, bool PRO, typename T, typename P>
__device__ __forceinline__ T cop_comp(const P &pv, const LevelDims &d, int64_t p, int i, int j, int k,
                                      const T *__restrict__ x, const T *__restrict__ b, T omega,
                                      const Prolong<T> &pr) {
    Row<T> r;
    const bool act = pv.template row_all<C>(i, j, k, r);
    auto X = [&](int c, int ni, int nj, int nk) {
        ni = max(0, min(ni, d.nx - 1)); nj = max(0, min(nj, d.ny - 1)); nk = max(0, min(nk, d.nz - 1));
        return pr.template at<PRO>(x, (int64_t)c * d.n1 + lidx(d, ni, nj, nk), c, ni, nj, nk);
    };
    return cop_eval<C, NC, MODE, T>(r, act, X, i, j, k, b[(int64_t)C * d.n1 + p], omega);
}

template <typename T, typename P, int NC, int MODE, bool PRO>
__global__ void __launch_bounds__(256) k_mg_cop(P pv, const T *__restrict__ x, const T *__restrict__ b,
                                                T *__restrict__ out, T omega, Prolong<T> pr) {
    const LevelDims d = pv.dims();
    int i, j, k;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (!node_of(d, p, i, j, k)) return;
    const T v0 = cop_comp<0, NC, MODE, PRO, T>(pv, d, p, i, j, k, x, b, omega, pr);
    if constexpr (NC == 3) {
        const T v1 = cop_comp<1, NC, MODE, PRO, T>(pv, d, p, i, j, k, x, b, omega, pr);
        const T v2 = cop_comp<2, NC, MODE, PRO, T>(pv, d, p, i, j, k, x, b, omega, pr);
        out[d.n1 + p] = v1;
        out[2 * d.n1 + p] = v2;
    }
    out[p] = v0;
}

// ---------------------------------------------------------------------------
// Level 0 of either block on a z-march: a workgroup owns a 32 x 8 column of
// nodes and walks it through zt.kz planes; the components of planes k-1, k,
// k+1 (with a one-node x/y halo) sit in an LDS ring, so every x value
// is read from memory about once per column instead of once per stencil use
// (the per-node kernel issues about 70 loads per node).  Coefficients come
// from the class table when there is one (2 bytes per node instead of six
// arrays) and the plane-(k-1) tuple is carried in registers.  The arithmetic
// is that of k_mg_smooth / k_mg_resid with FineProv, term for term.
// ---------------------------------------------------------------------------
template <typename T>
struct A0Z {
    Grid g;
    ZTiling zt;
    const T *im[3];
    const T *sg[3];
    const uint16_t *cls;
    const double *ctab;
    FineConst<T> fk;
    const uint16_t *rcls;      // V block: row class per node (null: rows from the coefficient tuples)
    const T *rtab;             // and the row table (MG_CLS_RS_V values per class)
};

// Coefficient tuples of a node and of its -x, -y and -z neighbours.
template <typename T>
struct ZmTuple {
    T im0[3], imx[3], imy[3], imz[3], sg0[3], sgx[3], sgy[3], sgz[3];
    __device__ __forceinline__ T IM(int a, int di, int dj, int dk) const {
        return dk ? imz[a] : di ? imx[a] : dj ? imy[a] : im0[a];
    }
    __device__ __forceinline__ T SG(int a, int di, int dj, int dk) const {
        return dk ? sgz[a] : di ? sgx[a] : dj ? sgy[a] : sg0[a];
    }
};

template <typename T, bool CLS>
__device__ __forceinline__ void load_tuple(const A0Z<T> &z, int64_t q, T im[3], T sg[3]) {
    if constexpr (CLS) {
        const double *t = z.ctab + 6 * (int)z.cls[q];
#pragma unroll
        for (int a = 0; a < 3; ++a) { im[a] = (T)t[a]; sg[a] = (T)t[3 + a]; }
    } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) { im[a] = z.im[a][q]; sg[a] = z.sg[a][q]; }
    }
}

constexpr int A0_TX = 32, A0_TY = 8, A0_SY = A0_TX + 2, A0_PL = A0_SY * (A0_TY + 2);

// Workgroup barrier ordering LDS only: it waits for this wave's LDS
// operations, not for its global loads, so loads issued before it (the next
// plane's prefetch) stay in flight across it.  Used where no thread reads
// global memory another thread of the workgroup wrote before the barrier.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <typename T>
struct A0View {
    const T *lds;      // [comp][slot][A0_PL]
    int b[3];          // slot offsets of planes k-1, k, k+1
    int ctr;
    __device__ __forceinline__ T get(int c, int di, int dj, int dk) const {
        return lds[c * 3 * A0_PL + b[dk + 1] + ctr + di + dj * A0_SY];
    }
};

template <int C, int NC, typename T>
__device__ __forceinline__ T apply_lds(const A0View<T> &xv, const Row<T> &r, T &diag, T &x0) {
    x0 = xv.get(C, 0, 0, 0);
    T y = r.exc * x0;
    diag = r.exc;
    // every LDS value is a finite field value, so a zero coupling adds a zero
    // product: no branches
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        diag += r.cond[q];
        y += r.cond[q] * (x0 - xv.get(C, so_i(q), so_j(q), so_k(q)));
    }
    if (NC == 3) {
#pragma unroll
        for (int s = 0; s < 8; ++s) y += r.cross[s] * xv.get(xo_c(C, s), xo_i(C, s), xo_j(C, s), xo_k(C, s));
    }
    return y;
}

// Row of component C of an NC-component block (NC = 3: the A block's Ax, Ay,
// Az rows; NC = 1: the V block's row) at a node of the z-march.
template <int C, int NC, int MODE, typename T>
__device__ __forceinline__ T a0_value(const Grid &g, const FineConst<T> &fk, const ZmTuple<T> &cf,
                                      const A0View<T> &xv, int i, int j, int k, const T *bq, T omega) {
    if (!free_f(g, NC == 1 ? FV : C, i, j, k)) return (T)0;
    Row<T> r;
    if (NC == 3) row_a_raw<C, T>(cf, fk, r);
    else row_v_raw<T>(cf, fk, r);
    eliminate<C, NC, T>(g, i, j, k, r);
    T diag, x0;
    const T y = apply_lds<C, NC, T>(xv, r, diag, x0);
    return MODE == 0 ? x0 + omega * (bq[C] - y) / diag : bq[C] - y;
}
template <int C, int NC, int MODE, typename T>
__device__ __forceinline__ void a0_row(const Grid &g, const FineConst<T> &fk, const ZmTuple<T> &cf,
                                       const A0View<T> &xv, int i, int j, int k, int64_t p, const T *bq,
                                       T *__restrict__ out, T omega) {
    out[(int64_t)C * g.n1 + p] = a0_value<C, NC, MODE, T>(g, fk, cf, xv, i, j, k, bq, omega);
}

// MODE 3 of the z-march below: the residual r = b - A x of level 0
// restricted to level 1, bc = P^T r, together with level 1's first sweep
// from zero, xf = w1 bc / diag (0 on inactive rows), in place of the
// residual launch and k_mg_restrict_first: the residuals of two planes stay
// in LDS and each coarse node sums its (up to) 2 x 2 x 2 children there, in
// the order and with the expression of k_mg_restrict_first.  Needs an even
// z-chunk, so that both child planes of a coarse node are in one column.
struct NoRestrict {};
template <typename T, typename P>
struct A0Restrict {
    P pv;          // level 1's rows
    T *bc, *xf;    // level 1's right-hand side and first sweep
    T w1;          // level 1's first sweep weight
};

// Level 0 of either block (NC = 3: A block, NC = 1: V block) on the z-march.
// MODE 0: weighted Jacobi sweep out = x + w (b - A x)/diag; MODE 1: residual
// out = b - A x; MODE 3: the residual restricted to level 1 (rr; out unused).
// Inactive rows get 0.
template <typename T, int NC, int MODE, bool CLS, bool PRO, bool RC = false, typename RR = NoRestrict>
__global__ void __launch_bounds__(256) k_a0_zm(A0Z<T> z, const T *__restrict__ x, const T *__restrict__ b,
                                               T *__restrict__ out, T omega, Prolong<T> pr, RR rr) {
    static_assert(!RC || NC == 1, "row classes: V block only");
    static_assert((MODE == 3) != std::is_same<RR, NoRestrict>::value, "MODE 3 takes the restriction's outputs");
    constexpr int RM = MODE == 3 ? 1 : MODE;     // the row arithmetic: MODE 3 forms residuals
    __shared__ T lds[NC * 3 * A0_PL];
    // MODE 3: the residuals of the last two planes, [plane & 1][comp][node of the column]
    __shared__ T rs[MODE == 3 ? 2 * NC * A0_TX * A0_TY : 1];
    Grid g = z.g;
    g.nx = __builtin_amdgcn_readfirstlane(z.g.nx);
    g.ny = __builtin_amdgcn_readfirstlane(z.g.ny);
    g.nz = __builtin_amdgcn_readfirstlane(z.g.nz);
    g.n1 = (int64_t)g.nx * g.ny * g.nz;
    const unsigned tid = xcd_tile(blockIdx.x, gridDim.x);
    const unsigned tx = tid % (unsigned)z.zt.ntx, r1 = tid / (unsigned)z.zt.ntx;
    const unsigned ty = r1 % (unsigned)z.zt.nty, tz = r1 / (unsigned)z.zt.nty;
    const int kb = (int)tz * z.zt.kz, ke = min(kb + z.zt.kz, g.nz);
    const int i0 = (int)tx * A0_TX, j0 = (int)ty * A0_TY;
    const int li = (int)(threadIdx.x % A0_TX), lj = (int)(threadIdx.x / A0_TX);
    const int i = i0 + li, j = j0 + lj;
    const int ic = min(i, g.nx - 1), jc = min(j, g.ny - 1);
    const bool live = i < g.nx && j < g.ny;
    const int ctr = (lj + 1) * A0_SY + (li + 1);
    const int64_t n1 = g.n1, plane = (int64_t)g.nx * g.ny;
    // staging of a plane (with its one-node x/y halo): each thread owns the
    // ring positions threadIdx.x and threadIdx.x + 256 (< A0_PL), loaded into
    // registers (xr) one plane ahead and stored to LDS the next iteration
    static_assert(A0_PL <= 512, "two staging positions per thread");
    T xr[2][NC];
    auto stage_load = [&](int kp) {
        kp = max(0, min(kp, g.nz - 1));
#pragma unroll
        for (int h2 = 0; h2 < 2; ++h2) {
            const int e = (int)threadIdx.x + 256 * h2;
            if (e < A0_PL) {
                const int ei = e % A0_SY, ej = e / A0_SY;
                const int gi = max(0, min(i0 - 1 + ei, g.nx - 1)), gj = max(0, min(j0 - 1 + ej, g.ny - 1));
                const int64_t q = gi + (int64_t)g.nx * gj + plane * kp;
#pragma unroll
                for (int c = 0; c < NC; ++c) xr[h2][c] = pr.template at<PRO>(x, c * n1 + q, c, gi, gj, kp);
            }
        }
    };
    auto stage_store = [&](int slot) {
#pragma unroll
        for (int h2 = 0; h2 < 2; ++h2) {
            const int e = (int)threadIdx.x + 256 * h2;
            if (e < A0_PL) {
#pragma unroll
                for (int c = 0; c < NC; ++c) lds[(c * 3 + slot) * A0_PL + e] = xr[h2][c];
            }
        }
    };
    stage_load(kb - 1);
    stage_store(0);
    stage_load(kb);
    stage_store(1);
    stage_load(kb + 1);
    const int64_t pc = ic + (int64_t)g.nx * jc;
    const int dxm = ic > 0 ? -1 : 0;
    const int64_t dym = jc > 0 ? -(int64_t)g.nx : 0;
    ZmTuple<T> cf;
    if constexpr (!RC) load_tuple<T, CLS>(z, pc + plane * max(kb - 1, 0), cf.imz, cf.sgz);
    for (int kk = kb; kk < ke; ++kk) {
        const int r = kk - kb;
        const int sm = r % 3, s0 = (r + 1) % 3, sp = (r + 2) % 3;
        const int64_t p = pc + plane * kk;
        const T *rt = nullptr;
        if constexpr (RC) {
            rt = z.rtab + MG_CLS_RS_V * (int)z.rcls[p];
        } else {
            load_tuple<T, CLS>(z, p, cf.im0, cf.sg0);
            load_tuple<T, CLS>(z, p + dxm, cf.imx, cf.sgx);
            load_tuple<T, CLS>(z, p + dym, cf.imy, cf.sgy);
        }
        T bq[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) bq[c] = live ? b[c * n1 + p] : (T)0;
        stage_store(sp);                    // plane kk + 1, loaded the iteration before
        if (kk + 1 < ke) stage_load(kk + 2);   // in flight across the barriers below
        lds_barrier();
        if (live) {
            const A0View<T> xv{lds, {sm * A0_PL, s0 * A0_PL, sp * A0_PL}, ctr};
            if constexpr (RC) {
                // the node's row from the table: conductances, excess
                // diagonal, 1/diag (0: inactive row)
                Row<T> rw;
#pragma unroll
                for (int q = 0; q < 6; ++q) rw.cond[q] = rt[q];
                rw.exc = rt[6];
                T diag, x0;
                const T y = apply_lds<0, 1, T>(xv, rw, diag, x0);
                const T v = RM == 0 ? x0 + omega * (bq[0] - y) / diag : bq[0] - y;
                const T o = rt[7] != (T)0 ? v : (T)0;
                if constexpr (MODE == 3) rs[(kk & 1) * A0_TX * A0_TY + threadIdx.x] = o;
                else out[p] = o;
            } else if constexpr (MODE == 3) {
                const int sl = (kk & 1) * NC * A0_TX * A0_TY + (int)threadIdx.x;
                rs[sl] = a0_value<0, NC, 1, T>(g, z.fk, cf, xv, i, j, kk, bq, omega);
                if constexpr (NC == 3) {
                    rs[sl + A0_TX * A0_TY] = a0_value<1, NC, 1, T>(g, z.fk, cf, xv, i, j, kk, bq, omega);
                    rs[sl + 2 * A0_TX * A0_TY] = a0_value<2, NC, 1, T>(g, z.fk, cf, xv, i, j, kk, bq, omega);
                }
            } else {
                a0_row<0, NC, MODE, T>(g, z.fk, cf, xv, i, j, kk, p, bq, out, omega);
                if constexpr (NC == 3) {
                    a0_row<1, NC, MODE, T>(g, z.fk, cf, xv, i, j, kk, p, bq, out, omega);
                    a0_row<2, NC, MODE, T>(g, z.fk, cf, xv, i, j, kk, p, bq, out, omega);
                }
            }
        } else if constexpr (MODE == 3) {
#pragma unroll
            for (int c = 0; c < NC; ++c) rs[((kk & 1) * NC + c) * A0_TX * A0_TY + threadIdx.x] = (T)0;
        }
        if constexpr (MODE == 3) {
            if ((kk & 1) || kk == g.nz - 1) {
                lds_barrier();
                // coarse plane kk / 2: the 16 x 4 coarse nodes of this column
                const LevelDims fd{g.nx, g.ny, g.nz, g.n1};
                const LevelDims cd = rr.pv.dims();
                const int I = i0 / 2 + (int)(threadIdx.x % (A0_TX / 2)), J = j0 / 2 + (int)(threadIdx.x / (A0_TX / 2));
                const int K = kk >> 1;
                if (threadIdx.x < (A0_TX / 2) * (A0_TY / 2) && I < cd.nx && J < cd.ny) {
                    const int64_t pcn = lidx(cd, I, J, K);
                    auto comp = [&](auto cc) {
                        constexpr int C = decltype(cc)::value;
                        T sum = (T)0;
#pragma unroll
                        for (int ch = 0; ch < 8; ++ch) {
                            const int fi = 2 * I + (ch & 1), fj = 2 * J + ((ch >> 1) & 1), fk2 = 2 * K + (ch >> 2);
                            if (inside(fd, fi, fj, fk2))
                                sum += rs[((fk2 & 1) * NC + C) * A0_TX * A0_TY + (fj - j0) * A0_TX + (fi - i0)];
                        }
                        Row<T> row;
                        const bool act = rr.pv.template row_all<C>(I, J, K, row);
                        const T v = rr.w1 * sum / row_diag(row);
                        rr.bc[(int64_t)C * cd.n1 + pcn] = sum;
                        rr.xf[(int64_t)C * cd.n1 + pcn] = act ? v : (T)0;
                    };
                    comp(std::integral_constant<int, 0>());
                    if constexpr (NC == 3) {
                        comp(std::integral_constant<int, 1>());
                        comp(std::integral_constant<int, 2>());
                    }
                }
            }
        }
        if constexpr (!RC) {
#pragma unroll
            for (int a = 0; a < 3; ++a) { cf.imz[a] = cf.im0[a]; cf.sgz[a] = cf.sg0[a]; }
        }
        lds_barrier();
    }
}

// bc = P^T r: the sum over the (up to) 2x2x2 children of each coarse node.
template <typename T>
__global__ void __launch_bounds__(256) k_mg_restrict(LevelDims fd, LevelDims cd, int ncomp, const T *__restrict__ r,
                                                     T *__restrict__ bc) {
    int I, J, K;
    const int64_t pc = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (!node_of(cd, pc, I, J, K)) return;
    for (int c = 0; c < ncomp; ++c) {
        T s = (T)0;
#pragma unroll
        for (int ch = 0; ch < 8; ++ch) {
            const int i = 2 * I + (ch & 1), j = 2 * J + ((ch >> 1) & 1), k = 2 * K + (ch >> 2);
            if (inside(fd, i, j, k)) s += r[(int64_t)c * fd.n1 + lidx(fd, i, j, k)];
        }
        bc[(int64_t)c * cd.n1 + pc] = s;
    }
}

// The restriction bc = P^T r together with the coarse level's first sweep
// from zero, xo = omega bc / diag (0 on inactive rows): the same sums and
// the same expression as k_mg_restrict followed by k_mg_cop's first sweep,
// in one launch.
template <int C, typename T, typename P>
__device__ __forceinline__ void restrict_first_comp(const P &pv, const LevelDims &fd, int I, int J, int K,
                                                    const T *__restrict__ r, T omega, T &sum, T &first) {
    T s = (T)0;
#pragma unroll
    for (int ch = 0; ch < 8; ++ch) {
        const int i = 2 * I + (ch & 1), j = 2 * J + ((ch >> 1) & 1), k = 2 * K + (ch >> 2);
        if (inside(fd, i, j, k)) s += r[(int64_t)C * fd.n1 + lidx(fd, i, j, k)];
    }
    Row<T> row;
    const bool act = pv.template row_all<C>(I, J, K, row);
    const T v = omega * s / row_diag(row);
    sum = s;
    first = act ? v : (T)0;
}

// Every component's loads are issued before the first store (see k_mg_cop).
template <typename T, typename P, int NC>
__global__ void __launch_bounds__(256) k_mg_restrict_first(P pv, LevelDims fd, const T *__restrict__ r,
                                                           T *__restrict__ bc, T *__restrict__ xo, T omega) {
    const LevelDims cd = pv.dims();
    int I, J, K;
    const int64_t pc = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (!node_of(cd, pc, I, J, K)) return;
    T s[NC], f[NC];
    restrict_first_comp<0, T>(pv, fd, I, J, K, r, omega, s[0], f[0]);
    if constexpr (NC == 3) {
        restrict_first_comp<1, T>(pv, fd, I, J, K, r, omega, s[1], f[1]);
        restrict_first_comp<2, T>(pv, fd, I, J, K, r, omega, s[2], f[2]);
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        bc[(int64_t)c * cd.n1 + pc] = s[c];
        xo[(int64_t)c * cd.n1 + pc] = f[c];
    }
}

// Coarsest level: x = Ainv b over its active unknowns, every other entry 0
// (one workgroup of MG_COARSE_THREADS).  b is gathered into LDS once.  Row r
// is summed by G threads (G = min(8, threads / n)), thread g over columns
// [g n / G, (g+1) n / G) in ascending order; the inverse is stored
// column-major, so the threads of a wave read consecutive entries of a
// column, and the G partial sums are added in order.
constexpr int MG_COARSE_MAX = 1024, MG_COARSE_THREADS = 1024;

// The solve proper, on bs[0..n) (gathered b, FP64) with partial sums in
// ps[0..MG_COARSE_THREADS); put(r, v) stores row r's result.  Every thread of
// the workgroup calls it; the partition is that of MG_COARSE_THREADS threads
// whatever the workgroup size (a smaller workgroup takes several of those
// slots per thread), so the sums are the same bit for bit.
template <typename F>
__device__ __forceinline__ void coarse_core(int n, const double *__restrict__ invT, const double *bs, double *ps,
                                            F put) {
    constexpr int VT = MG_COARSE_THREADS;
    const int G = n > 0 ? max(1, min(8, VT / n)) : 1;
    const int RP = VT / G;
    for (int r0 = 0; r0 < n; r0 += RP) {
        for (int t = (int)threadIdx.x; t < VT; t += (int)blockDim.x) {
            const int r = r0 + t % RP, g = t / RP;
            double s = 0.0;
            if (r < n && g < G) {
                const int c0 = g * n / G, c1 = (g + 1) * n / G;
                double a[4] = {0.0, 0.0, 0.0, 0.0};
                int c = c0;
                // 16 loads in flight, then the products in the order of the
                // 4-wide loop below
                for (; c + 16 <= c1; c += 16) {
                    double v[16];
#pragma unroll
                    for (int u = 0; u < 16; ++u) v[u] = invT[(int64_t)(c + u) * n + r];
#pragma unroll
                    for (int u = 0; u < 16; ++u) a[u & 3] += v[u] * bs[c + u];
                }
                for (; c + 4 <= c1; c += 4) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) a[u] += invT[(int64_t)(c + u) * n + r] * bs[c + u];
                }
                for (; c < c1; ++c) a[0] += invT[(int64_t)c * n + r] * bs[c];
                s = (a[0] + a[1]) + (a[2] + a[3]);
            }
            ps[t] = s;
        }
        __syncthreads();
        for (int t = (int)threadIdx.x; t < VT; t += (int)blockDim.x) {
            const int r = r0 + t % RP, g = t / RP;
            if (r < n && g == 0) {
                double tot = 0.0;
                for (int u = 0; u < G; ++u) tot += ps[t + u * RP];
                put(r, tot);
            }
        }
        __syncthreads();
    }
}

template <typename T>
__global__ void __launch_bounds__(MG_COARSE_THREADS) k_mg_coarse(int n, int64_t nv, const int *__restrict__ idx,
                                                                 const double *__restrict__ invT,
                                                                 const T *__restrict__ b, T *__restrict__ x) {
    __shared__ double bs[MG_COARSE_MAX];
    __shared__ double ps[MG_COARSE_THREADS];
    for (int c = threadIdx.x; c < n; c += blockDim.x) bs[c] = (double)b[idx[c]];
    for (int64_t q = threadIdx.x; q < nv; q += blockDim.x) x[q] = (T)0;
    __syncthreads();
    coarse_core(n, invT, bs, ps, [&](int r, double v) { x[idx[r]] = (T)v; });
}

// ---------------------------------------------------------------------------
// The tail of the V-cycle in one workgroup: the level above the coarsest
// (at most MG_TAIL_THREADS nodes, one thread per node, its rows held in
// registers and its vectors in LDS) and the coarsest solve, in one launch
// instead of 2 nu + 4.  Input: the residual rf of the level below (fd);
// output: xout, the tail level's x after its post-smoothing.  The steps and
// their arithmetic are those of the one-launch-per-step path (the restriction
// fused with the first sweep, nu - 1 sweeps, the residual, the restriction to
// the coarsest level, its dense solve, the prolongation fused into the first
// post-sweep, nu - 1 sweeps), with barriers between them.
// ---------------------------------------------------------------------------
constexpr int MG_TAIL_LDS = 49152;
template <typename T>
constexpr int mg_tail_rows() { return MG_TAIL_LDS / (4 * (int)sizeof(T)); }
// Threads of the tail, one per node: the level above the coarsest has at
// most 8 nodes per axis (the coarsest has at most 4), so at most 512 nodes.
// 512 threads leave 256 registers per lane for the rows held in registers
// (about 90 for the three FP64 rows of an A-block node).
constexpr int MG_TAIL_THREADS = 512;
// Sweeps per side of the tail level that the tail takes weights for.
constexpr int MG_TAIL_MAX_NU = 8;
template <typename T>
struct TailWeights {
    T pre[MG_TAIL_MAX_NU], post[MG_TAIL_MAX_NU];   // sweep weights before and after the coarse solve
};

template <typename T, typename P, int NC>
__global__ void __launch_bounds__(MG_TAIL_THREADS) k_mg_tail(P pv, LevelDims fd, const T *__restrict__ rf,
                                                             LevelDims cd, int n, const int *__restrict__ idx,
                                                             const double *__restrict__ invT, T *__restrict__ xout,
                                                             TailWeights<T> wt, T alpha, int nu) {
    constexpr int R = mg_tail_rows<T>();
    static_assert(R * sizeof(T) >= sizeof(double) * MG_COARSE_THREADS && R * sizeof(T) >= sizeof(double) * MG_COARSE_MAX,
                  "the coarse solve's gathered b and partial sums live in two of the tail level's vectors");
    // four vectors of the tail level (x ping-pong, b, r), then the coarsest
    // level's b and x in node layout; the gathered coarse b and the partial
    // sums of its solve reuse the r vector and the free x buffer
    __shared__ double raw[MG_TAIL_LDS / 8 + 2 * MG_COARSE_MAX * sizeof(T) / 8];
    T *const v0 = (T *)raw, *const v1 = v0 + R, *const bl = v1 + R, *const rl = bl + R;
    T *const bcs = rl + R, *const xcs = bcs + MG_COARSE_MAX;
    const LevelDims d = pv.dims();
    const int64_t nvc = (int64_t)NC * cd.n1;
    int i = 0, j = 0, k = 0;
    const int p = (int)threadIdx.x;
    const bool live = node_of(d, p, i, j, k);
    Row<T> rw[NC];
    bool act[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) act[c] = false;
    if (live) {
        act[0] = pv.template row_all<0>(i, j, k, rw[0]);
        if constexpr (NC == 3) {
            act[1] = pv.template row_all<1>(i, j, k, rw[1]);
            act[2] = pv.template row_all<2>(i, j, k, rw[2]);
        }
    }
    // LDS index of the node's 27 neighbours (offsets -1..1 per axis), each
    // clamped into the box, computed once: the rows only ever read these
    int nb[27];
#pragma unroll
    for (int o = 0; o < 27; ++o) {
        const int ni = max(0, min(i + o % 3 - 1, d.nx - 1)), nj = max(0, min(j + (o / 3) % 3 - 1, d.ny - 1)),
                  nk = max(0, min(k + o / 9 - 1, d.nz - 1));
        nb[o] = (int)lidx(d, ni, nj, nk);
    }
    // X(c, ni, nj, nk) of the rows: (ni - i, nj - j, nk - k) are compile-time
    // offsets once the row code is inlined
    auto Xof = [&](const T *xs) {
        return [=, &nb, &d](int c, int ni, int nj, int nk) {
            return xs[c * d.n1 + nb[(ni - i + 1) + 3 * (nj - j + 1) + 9 * (nk - k + 1)]];
        };
    };
    // one sweep with weight omega (MODE 0) or the residual (MODE 2) of every
    // row of the node
    auto pass = [&](auto mode, const auto &X, T *dst, T omega) {
        constexpr int MODE = decltype(mode)::value;
        if (live) {
            T o[NC];
            o[0] = cop_eval<0, NC, MODE, T>(rw[0], act[0], X, i, j, k, bl[p], omega);
            if constexpr (NC == 3) {
                o[1] = cop_eval<1, NC, MODE, T>(rw[1], act[1], X, i, j, k, bl[d.n1 + p], omega);
                o[2] = cop_eval<2, NC, MODE, T>(rw[2], act[2], X, i, j, k, bl[2 * d.n1 + p], omega);
            }
#pragma unroll
            for (int c = 0; c < NC; ++c) dst[c * d.n1 + p] = o[c];
        }
        __syncthreads();
    };
    using M0 = std::integral_constant<int, 0>;
    using M2 = std::integral_constant<int, 2>;
    // restriction of rf fused with the first sweep from zero
    T *cur = v0, *oth = v1;
    if (live) {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            T s = (T)0;
#pragma unroll
            for (int ch = 0; ch < 8; ++ch) {
                const int fi = 2 * i + (ch & 1), fj = 2 * j + ((ch >> 1) & 1), fk = 2 * k + (ch >> 2);
                if (inside(fd, fi, fj, fk)) s += rf[(int64_t)c * fd.n1 + lidx(fd, fi, fj, fk)];
            }
            const T v = wt.pre[0] * s / row_diag(rw[c]);
            bl[c * d.n1 + p] = s;
            cur[c * d.n1 + p] = act[c] ? v : (T)0;
        }
    }
    __syncthreads();
    for (int it = 1; it < nu; ++it) {
        pass(M0(), Xof(cur), oth, wt.pre[it]);
        T *t = cur; cur = oth; oth = t;
    }
    pass(M2(), Xof(cur), rl, (T)0);
    // restriction to the coarsest level
    for (int pc = p; pc < cd.n1; pc += (int)blockDim.x) {
        int ci, cj, ck;
        node_of(cd, pc, ci, cj, ck);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            T s = (T)0;
#pragma unroll
            for (int ch = 0; ch < 8; ++ch) {
                const int fi = 2 * ci + (ch & 1), fj = 2 * cj + ((ch >> 1) & 1), fk = 2 * ck + (ch >> 2);
                if (inside(d, fi, fj, fk)) s += rl[c * d.n1 + lidx(d, fi, fj, fk)];
            }
            bcs[c * cd.n1 + pc] = s;
        }
    }
    __syncthreads();
    // coarsest solve (gathered b in the r vector, the MG_COARSE_THREADS
    // partial sums in the free x buffer)
    double *bs = (double *)rl, *ps = (double *)oth;
    for (int c = p; c < n; c += blockDim.x) bs[c] = (double)bcs[idx[c]];
    for (int64_t q = p; q < nvc; q += blockDim.x) xcs[q] = (T)0;
    __syncthreads();
    coarse_core(n, invT, bs, ps, [&](int r, double v) { xcs[idx[r]] = (T)v; });
    // post-smoothing, the first sweep reading x + alpha * (the coarse x of
    // the parent node)
    {
        const T *xs = cur;
        auto XP = [=, &d, &cd](int c, int ni, int nj, int nk) {
            ni = max(0, min(ni, d.nx - 1)); nj = max(0, min(nj, d.ny - 1)); nk = max(0, min(nk, d.nz - 1));
            return xs[c * d.n1 + lidx(d, ni, nj, nk)] + alpha * xcs[c * cd.n1 + lidx(cd, ni >> 1, nj >> 1, nk >> 1)];
        };
        pass(M0(), XP, oth, wt.post[0]);
        T *t = cur; cur = oth; oth = t;
    }
    for (int it = 1; it < nu; ++it) {
        pass(M0(), Xof(cur), oth, wt.post[it]);
        T *t = cur; cur = oth; oth = t;
    }
    if (live) {
#pragma unroll
        for (int c = 0; c < NC; ++c) xout[(int64_t)c * d.n1 + p] = cur[c * d.n1 + p];
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_fill(T *__restrict__ x, int64_t n, T v) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) x[p] = v;
}

template <typename T>
__global__ void __launch_bounds__(256) k_copy_t(T *__restrict__ dst, const T *__restrict__ src, int64_t n) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) dst[p] = src[p];
}

// ---------------------------------------------------------------------------
// Host: hierarchy construction in FP64
// ---------------------------------------------------------------------------
namespace {

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// One explicit level in FP64 on the host.
struct HostLevel {
    LevelDims d;
    int nc;
    std::vector<double> cond, exc, cross, dinv;   // [nc][6][n1], [nc][n1], [nc][8][n1], [nc][n1]
};

template <int C, int NC>
bool host_fine_row(const Grid &g, const HostCoef &hc, const FineConst<double> &fk, int i, int j, int k,
                   Row<double> &r) {
    if (!free_f(g, NC == 1 ? FV : C, i, j, k)) return false;
    const CoefAt<double> ca{{hc.im[0], hc.im[1], hc.im[2]}, {hc.sg[0], hc.sg[1], hc.sg[2]}, gidx(g, i, j, k), g.nx,
                            g.nx * g.ny};
    if (NC == 3) row_a_raw<C, double>(ca, fk, r);
    else row_v_raw<double>(ca, fk, r);
    eliminate<C, NC, double>(g, i, j, k, r);
    return true;
}

bool host_expl_row(const HostLevel &L, int c, int i, int j, int k, Row<double> &r) {
    const int64_t p = lidx(L.d, i, j, k), n1 = L.d.n1;
    if (L.dinv[(size_t)(c * n1 + p)] == 0.0) return false;
    for (int q = 0; q < 6; ++q) r.cond[q] = L.cond[(size_t)((c * 6 + q) * n1 + p)];
    r.exc = L.exc[(size_t)(c * n1 + p)];
    for (int s = 0; s < 8; ++s) r.cross[s] = L.nc == 3 ? L.cross[(size_t)((c * 8 + s) * n1 + p)] : 0.0;
    return true;
}

// Galerkin product of one level with 2x2x2 aggregation.  `row(c,i,j,k,r)`
// gives the fine row of component c at (i,j,k) (false if inactive).
template <typename RowFn>
HostLevel galerkin(const LevelDims &fd, int nc, RowFn row) {
    HostLevel C;
    C.nc = nc;
    C.d.nx = (fd.nx + 1) / 2;
    C.d.ny = (fd.ny + 1) / 2;
    C.d.nz = (fd.nz + 1) / 2;
    C.d.n1 = (int64_t)C.d.nx * C.d.ny * C.d.nz;
    const int64_t n1 = C.d.n1;
    C.cond.assign((size_t)(nc * 6 * n1), 0.0);
    C.exc.assign((size_t)(nc * n1), 0.0);
    C.dinv.assign((size_t)(nc * n1), 0.0);
    if (nc == 3) C.cross.assign((size_t)(nc * 8 * n1), 0.0);
    for (int K = 0; K < C.d.nz; ++K)
        for (int J = 0; J < C.d.ny; ++J)
            for (int I = 0; I < C.d.nx; ++I) {
                const int64_t P = lidx(C.d, I, J, K);
                for (int c = 0; c < nc; ++c) {
                    double cc[6] = {0, 0, 0, 0, 0, 0}, ce = 0.0, cx[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                    bool any = false;
                    for (int ch = 0; ch < 8; ++ch) {
                        const int i = 2 * I + (ch & 1), j = 2 * J + ((ch >> 1) & 1), k = 2 * K + (ch >> 2);
                        if (!inside(fd, i, j, k)) continue;
                        Row<double> r;
                        if (!row(c, i, j, k, r)) continue;
                        any = true;
                        ce += r.exc;
                        for (int q = 0; q < 6; ++q) {
                            const int ni = i + so_i(q), nj = j + so_j(q), nk = k + so_k(q);
                            // a link inside the aggregate cancels in P^T A P
                            if ((ni >> 1) != I || (nj >> 1) != J || (nk >> 1) != K) cc[q] += r.cond[q];
                        }
                        if (nc == 3) {
                            for (int s = 0; s < 8; ++s) {
                                if (r.cross[s] == 0.0) continue;
                                const int oi = ((i + xo_i(c, s)) >> 1) - I, oj = ((j + xo_j(c, s)) >> 1) - J,
                                          ok = ((k + xo_k(c, s)) >> 1) - K;
                                const int base = s & 4;
                                int hit = -1;
                                for (int s2 = 0; s2 < 4; ++s2)
                                    if (xo_i(c, base + s2) == oi && xo_j(c, base + s2) == oj &&
                                        xo_k(c, base + s2) == ok)
                                        hit = base + s2;
                                if (hit < 0) throw std::logic_error("galerkin: coarse cross offset outside the pattern");
                                cx[hit] += r.cross[s];
                            }
                        }
                    }
                    if (!any) continue;
                    double diag = ce;
                    for (int q = 0; q < 6; ++q) {
                        C.cond[(size_t)((c * 6 + q) * n1 + P)] = cc[q];
                        diag += cc[q];
                    }
                    C.exc[(size_t)(c * n1 + P)] = ce;
                    if (nc == 3)
                        for (int s = 0; s < 8; ++s) C.cross[(size_t)((c * 8 + s) * n1 + P)] = cx[s];
                    if (!(diag > 0.0)) throw std::runtime_error("galerkin: coarse row without a positive diagonal");
                    C.dinv[(size_t)(c * n1 + P)] = 1.0 / diag;
                    // the coarse kernels clamp neighbour reads into the box and rely on every
                    // coupling that leaves it being zero
                    for (int q = 0; q < 6; ++q)
                        if (cc[q] != 0.0 && !inside(C.d, I + so_i(q), J + so_j(q), K + so_k(q)))
                            throw std::logic_error("galerkin: coarse coupling outside the level's box");
                    if (nc == 3)
                        for (int s = 0; s < 8; ++s)
                            if (cx[s] != 0.0 && !inside(C.d, I + xo_i(c, s), J + xo_j(c, s), K + xo_k(c, s)))
                                throw std::logic_error("galerkin: coarse coupling outside the level's box");
                }
            }
    return C;
}

// Dense FP64 inverse by Gauss-Jordan with partial pivoting.
void dense_inverse(std::vector<double> &A, int n, std::vector<double> &inv) {
    inv.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
    for (int col = 0; col < n; ++col) {
        int piv = col;
        double best = fabs(A[(size_t)col * n + col]);
        for (int r = col + 1; r < n; ++r) {
            const double v = fabs(A[(size_t)r * n + col]);
            if (v > best) { best = v; piv = r; }
        }
        if (best == 0.0) throw std::runtime_error("singular coarsest-level operator");
        if (piv != col)
            for (int c = 0; c < n; ++c) {
                std::swap(A[(size_t)col * n + c], A[(size_t)piv * n + c]);
                std::swap(inv[(size_t)col * n + c], inv[(size_t)piv * n + c]);
            }
        const double dd = 1.0 / A[(size_t)col * n + col];
        for (int c = 0; c < n; ++c) { A[(size_t)col * n + c] *= dd; inv[(size_t)col * n + c] *= dd; }
        for (int r = 0; r < n; ++r) {
            if (r == col) continue;
            const double m = A[(size_t)r * n + col];
            if (m == 0.0) continue;
            for (int c = 0; c < n; ++c) {
                A[(size_t)r * n + c] -= m * A[(size_t)col * n + c];
                inv[(size_t)r * n + c] -= m * inv[(size_t)col * n + c];
            }
        }
    }
}

// Distinct row tuples of an explicit level, keyed by their bit patterns
// (layout MG_CLS_RS_*).  Coarse rows are sums of fine rows in a fixed order,
// so two aggregates with the same local configuration of materials and
// Dirichlet shell give bit-identical tuples: a design built from a few boxes
// has a few thousand.  Returns false (the level keeps its explicit arrays)
// above 65535 classes.
bool class_compress(const HostLevel &H, std::vector<uint16_t> &cls, std::vector<double> &tab) {
    const int nc = H.nc, RS = nc == 3 ? MG_CLS_RS_A : MG_CLS_RS_V, W = nc * RS;
    const int64_t n1 = H.d.n1;
    cls.assign((size_t)n1, 0);
    tab.clear();
    std::vector<double> t((size_t)W);
    std::unordered_multimap<uint64_t, uint32_t> seen;
    seen.reserve(4096);
    uint32_t ncls = 0;
    for (int64_t p = 0; p < n1; ++p) {
        for (int c = 0; c < nc; ++c) {
            double *r = t.data() + c * RS;
            for (int q = 0; q < 6; ++q) r[q] = H.cond[(size_t)((c * 6 + q) * n1 + p)];
            r[6] = H.exc[(size_t)(c * n1 + p)];
            r[7] = H.dinv[(size_t)(c * n1 + p)];
            if (nc == 3)
                for (int s = 0; s < 8; ++s) r[8 + s] = H.cross[(size_t)((c * 8 + s) * n1 + p)];
        }
        uint64_t hsh = 0xcbf29ce484222325ull;
        for (int e = 0; e < W; ++e) {
            uint64_t u;
            memcpy(&u, &t[(size_t)e], 8);
            hsh = (hsh ^ u) * 0x100000001b3ull;
            hsh ^= hsh >> 31;
        }
        uint32_t id = UINT32_MAX;
        auto range = seen.equal_range(hsh);
        for (auto it = range.first; it != range.second; ++it)
            if (memcmp(tab.data() + (size_t)it->second * W, t.data(), sizeof(double) * W) == 0) {
                id = it->second;
                break;
            }
        if (id == UINT32_MAX) {
            if (ncls >= 65535) return false;
            id = ncls++;
            seen.emplace(hsh, id);
            tab.insert(tab.end(), t.begin(), t.end());
        }
        cls[(size_t)p] = (uint16_t)id;
    }
    return true;
}

template <typename T>
T *upload(const std::vector<double> &src, std::vector<T> &stage, const MgAlloc &al) {
    stage.resize(src.size());
    for (size_t e = 0; e < src.size(); ++e) stage[e] = (T)src[e];
    T *p = (T *)al.alloc(sizeof(T) * (stage.size() ? stage.size() : 1));
    if (!stage.empty()) hip_check(hipMemcpyAsync(p, stage.data(), sizeof(T) * stage.size(), hipMemcpyHostToDevice,
                                                 al.stream), "h2d");
    hip_check(hipStreamSynchronize(al.stream), "sync");
    return p;
}

}  // namespace

FineConst<double> fine_const(const Grid &g, double dt) {
    FineConst<double> k;
    const double h[3] = {g.hx, g.hy, g.hz};
    for (int a = 0; a < 3; ++a) k.ih2[a] = 1.0 / (h[a] * h[a]);
    k.ixy = 1.0 / (h[0] * h[1]);
    k.ixz = 1.0 / (h[0] * h[2]);
    k.iyz = 1.0 / (h[1] * h[2]);
    k.idt = dt > 0 ? 1.0 / dt : 0.0;
    k.sigfloor = QMS_SIGFLOOR;
    return k;
}

static inline dim3 g1(int64_t n);
template <typename T, typename F>
static void on_level(const Hierarchy<T> &h, int l, F f);

template <typename T>
void mg_setup(const Grid &g, const HostCoef &hc, const CoeffsT<T> &dc, double dt, int block, const MgAlloc &al,
              Hierarchy<T> &h) {
    h = Hierarchy<T>();
    h.block = block;
    const int nc = block == 0 ? 3 : 1;
    const FineConst<double> fk = fine_const(g, dt);
    const FineConst<double> fkd = fk;
    h.fk.ih2[0] = (T)fkd.ih2[0]; h.fk.ih2[1] = (T)fkd.ih2[1]; h.fk.ih2[2] = (T)fkd.ih2[2];
    h.fk.ixy = (T)fkd.ixy; h.fk.ixz = (T)fkd.ixz; h.fk.iyz = (T)fkd.iyz;
    h.fk.idt = (T)fkd.idt; h.fk.sigfloor = (T)fkd.sigfloor;
    h.g = g;
    for (int a = 0; a < 3; ++a) { h.im[a] = dc.im[a]; h.sg[a] = dc.sg[a]; }
    // Level 0 reads its coefficients through the class table (2 bytes per
    // node) in FP64 and in the V block.  The FP32 A block reads the six FP32
    // arrays instead: measured faster at 128^3 (26.2 against 28.6 us per
    // sweep), where the dependent loads class -> tuple cost more time than
    // the 24 bytes per node they save.
    const bool direct = std::is_same<T, float>::value && block == 0;
    h.cls = direct ? nullptr : dc.cls;
    h.ctab = direct ? nullptr : dc.ctab;
    {
        // level-0 z-march: 32 x 8 columns, the deepest z-chunk of 2..16 planes
        // that leaves at least 768 workgroups (three per CU) for the A block
        // and 1536 for the one-component V block, whose workgroups are
        // lighter (measured fastest at 128^3: 8 and 4 planes)
        const int64_t cols2d = (int64_t)((g.nx + A0_TX - 1) / A0_TX) * ((g.ny + A0_TY - 1) / A0_TY);
        const int64_t want = block == 0 ? 768 : 1536;
        // (at least 2, so that every coarse plane's two child planes lie in
        // one column: the fused residual restriction needs it)
        int kz = 2;
        for (int k : {4, 8, 16})
            if (cols2d * ((g.nz + k - 1) / k) >= want) kz = k;
        h.zt = ztiling(g, kz, A0_TX, h.zcols);
    }
    // level 0 work vectors
    LevelDims d0{g.nx, g.ny, g.nz, g.n1};
    h.lev[0].d = d0;
    h.lev[0].ncomp = nc;
    h.lev[0].t = (T *)al.alloc(sizeof(T) * (size_t)(nc * g.n1));
    h.lev[0].r = (T *)al.alloc(sizeof(T) * (size_t)(nc * g.n1));
    h.lev[0].dg = (T *)al.alloc(sizeof(T) * (size_t)(nc * g.n1));
    if (block == 1) {
        // V block: level 0's rows as a 2-byte class per node and a table of
        // the distinct rows (built in FP64 from the same row formulas), so
        // that a sweep reads one class and one row per node instead of four
        // coefficient tuples and forming the row
        HostLevel H0;
        H0.d = d0;
        H0.nc = 1;
        H0.cond.assign((size_t)(6 * g.n1), 0.0);
        H0.exc.assign((size_t)g.n1, 0.0);
        H0.dinv.assign((size_t)g.n1, 0.0);
        Row<double> rw;
        for (int k = 0; k < g.nz; ++k)
            for (int j = 0; j < g.ny; ++j)
                for (int i = 0; i < g.nx; ++i) {
                    if (!host_fine_row<0, 1>(g, hc, fk, i, j, k, rw)) continue;
                    const int64_t p = gidx(g, i, j, k);
                    double dg = rw.exc;
                    for (int q = 0; q < 6; ++q) { H0.cond[(size_t)(q * g.n1 + p)] = rw.cond[q]; dg += rw.cond[q]; }
                    H0.exc[(size_t)p] = rw.exc;
                    H0.dinv[(size_t)p] = 1.0 / dg;
                }
        std::vector<uint16_t> cls0;
        std::vector<double> tab0;
        if (class_compress(H0, cls0, tab0)) {
            std::vector<T> stage0;
            h.lev[0].ncls = (int)(tab0.size() / MG_CLS_RS_V);
            h.lev[0].tab = upload<T>(tab0, stage0, al);
            h.lev[0].cls = (uint16_t *)al.alloc(sizeof(uint16_t) * cls0.size());
            hip_check(hipMemcpyAsync(h.lev[0].cls, cls0.data(), sizeof(uint16_t) * cls0.size(), hipMemcpyHostToDevice,
                                     al.stream), "h2d");
            hip_check(hipStreamSynchronize(al.stream), "sync");
        }
    }
    on_level(h, 0, [&](auto pv, auto ncc) {
        using P = decltype(pv);
        constexpr int NC = decltype(ncc)::value;
        hipLaunchKernelGGL((k_mg_diag<T, P, NC>), g1(g.n1), dim3(256), 0, al.stream, pv, h.lev[0].dg);
    });
    // coarse levels
    std::vector<HostLevel> hl;
    auto fine_row = [&](int c, int i, int j, int k, Row<double> &r) {
        if (nc == 1) return host_fine_row<0, 1>(g, hc, fk, i, j, k, r);
        if (c == 0) return host_fine_row<0, 3>(g, hc, fk, i, j, k, r);
        if (c == 1) return host_fine_row<1, 3>(g, hc, fk, i, j, k, r);
        return host_fine_row<2, 3>(g, hc, fk, i, j, k, r);
    };
    LevelDims cur = d0;
    int nlev = 1;
    while (nlev < MG_MAX_LEVELS) {
        const int mx = std::max(cur.nx, std::max(cur.ny, cur.nz));
        if (mx <= 4 && nlev >= 2) break;
        HostLevel C = nlev == 1 ? galerkin(cur, nc, fine_row)
                                : galerkin(cur, nc, [&](int c, int i, int j, int k, Row<double> &r) {
                                      return host_expl_row(hl.back(), c, i, j, k, r);
                                  });
        cur = C.d;
        hl.push_back(std::move(C));
        ++nlev;
    }
    h.nlev = nlev;
    std::vector<T> stage;
    for (int l = 1; l < nlev; ++l) {
        const HostLevel &H = hl[(size_t)(l - 1)];
        MgLevel<T> &L = h.lev[l];
        L.d = H.d;
        L.ncomp = nc;
        std::vector<uint16_t> cls;
        std::vector<double> tab;
        if (class_compress(H, cls, tab)) {
            L.ncls = (int)(tab.size() / (size_t)(nc * (nc == 3 ? MG_CLS_RS_A : MG_CLS_RS_V)));
            L.tab = upload<T>(tab, stage, al);
            L.cls = (uint16_t *)al.alloc(sizeof(uint16_t) * cls.size());
            hip_check(hipMemcpyAsync(L.cls, cls.data(), sizeof(uint16_t) * cls.size(), hipMemcpyHostToDevice,
                                     al.stream), "h2d");
            hip_check(hipStreamSynchronize(al.stream), "sync");
        } else {
            L.cond = upload<T>(H.cond, stage, al);
            L.exc = upload<T>(H.exc, stage, al);
            L.dinv = upload<T>(H.dinv, stage, al);
            L.cross = nc == 3 ? upload<T>(H.cross, stage, al) : nullptr;
        }
        L.x = (T *)al.alloc(sizeof(T) * (size_t)(nc * H.d.n1));
        L.b = (T *)al.alloc(sizeof(T) * (size_t)(nc * H.d.n1));
        L.r = (T *)al.alloc(sizeof(T) * (size_t)(nc * H.d.n1));
        L.t = (T *)al.alloc(sizeof(T) * (size_t)(nc * H.d.n1));
    }
    // coarsest level: dense FP64 inverse over its active unknowns
    const HostLevel &H = hl.back();
    const int64_t n1 = H.d.n1;
    std::vector<int> map((size_t)(nc * n1), -1), idx;
    for (int64_t q = 0; q < nc * n1; ++q)
        if (H.dinv[(size_t)q] != 0.0) { map[(size_t)q] = (int)idx.size(); idx.push_back((int)q); }
    const int n = (int)idx.size();
    std::vector<double> A((size_t)n * n, 0.0), inv;
    for (int r = 0; r < n; ++r) {
        const int q = idx[(size_t)r];
        const int c = (int)(q / n1);
        const int64_t p = q % n1;
        const int i = (int)(p % H.d.nx), j = (int)((p / H.d.nx) % H.d.ny), k = (int)(p / ((int64_t)H.d.nx * H.d.ny));
        Row<double> rw;
        host_expl_row(H, c, i, j, k, rw);
        double diag = rw.exc;
        for (int q2 = 0; q2 < 6; ++q2) {
            if (rw.cond[q2] == 0.0) continue;
            diag += rw.cond[q2];
            const int ni = i + so_i(q2), nj = j + so_j(q2), nk = k + so_k(q2);
            if (!inside(H.d, ni, nj, nk)) continue;
            const int m = map[(size_t)(c * n1 + lidx(H.d, ni, nj, nk))];
            if (m >= 0) A[(size_t)r * n + m] -= rw.cond[q2];
        }
        A[(size_t)r * n + r] += diag;
        if (nc == 3)
            for (int s = 0; s < 8; ++s) {
                if (rw.cross[s] == 0.0) continue;
                const int ni = i + xo_i(c, s), nj = j + xo_j(c, s), nk = k + xo_k(c, s);
                if (!inside(H.d, ni, nj, nk)) continue;
                const int m = map[(size_t)(xo_c(c, s) * n1 + lidx(H.d, ni, nj, nk))];
                if (m >= 0) A[(size_t)r * n + m] += rw.cross[s];
            }
    }
    if (n > MG_COARSE_MAX) throw std::runtime_error("mg_setup: coarsest level too large for the dense solve");
    if (n > 0) dense_inverse(A, n, inv);
    // column-major for the device solve
    std::vector<double> invT((size_t)n * n);
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) invT[(size_t)c * n + r] = inv[(size_t)r * n + c];
    inv.swap(invT);
    h.ncoarse = n;
    h.coarse_idx = (int *)al.alloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    h.coarse_inv = (double *)al.alloc(sizeof(double) * (size_t)(n > 0 ? (size_t)n * n : 1));
    if (n > 0) {
        hip_check(hipMemcpyAsync(h.coarse_idx, idx.data(), sizeof(int) * n, hipMemcpyHostToDevice, al.stream), "h2d");
        hip_check(hipMemcpyAsync(h.coarse_inv, inv.data(), sizeof(double) * (size_t)n * n, hipMemcpyHostToDevice,
                                 al.stream), "h2d");
        hip_check(hipStreamSynchronize(al.stream), "sync");
    }
}

// ---------------------------------------------------------------------------
// Host: V-cycle launches
// ---------------------------------------------------------------------------
static inline dim3 g1(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

template <typename T, int NC, bool CLS>
static FineProv<T, NC, CLS> fine_prov(const Hierarchy<T> &h) {
    FineProv<T, NC, CLS> p;
    p.g = h.g;
    for (int a = 0; a < 3; ++a) { p.im[a] = h.im[a]; p.sg[a] = h.sg[a]; }
    p.cls = h.cls;
    p.ctab = h.ctab;
    p.fk = h.fk;
    return p;
}
template <typename T, int NC>
static ExplProv<T, NC> expl_prov(const MgLevel<T> &L) {
    return ExplProv<T, NC>{L.d, L.cond, L.exc, L.cross, L.dinv};
}
template <typename T, int NC>
static ClsProv<T, NC> cls_prov(const MgLevel<T> &L) {
    return ClsProv<T, NC>{L.d, L.cls, L.tab};
}

template <typename P>
struct is_coarse_prov : std::false_type {};
template <typename T, int NC>
struct is_coarse_prov<ClsProv<T, NC>> : std::true_type {};
template <typename T, int NC>
struct is_coarse_prov<ExplProv<T, NC>> : std::true_type {};

// Dispatch a kernel template on the level's row provider (level 0 from the
// coefficient arrays or the class table; coarse levels class-compressed or
// explicit) and the block.
template <typename T, typename F>
static void on_level(const Hierarchy<T> &h, int l, F f) {
    using I3 = std::integral_constant<int, 3>;
    using I1 = std::integral_constant<int, 1>;
    if (l == 0 && h.block == 1 && h.lev[0].cls) {
        f(cls_prov<T, 1>(h.lev[0]), I1());
    } else if (l == 0) {
        if (h.cls) {
            if (h.block == 0) f(fine_prov<T, 3, true>(h), I3());
            else f(fine_prov<T, 1, true>(h), I1());
        } else {
            if (h.block == 0) f(fine_prov<T, 3, false>(h), I3());
            else f(fine_prov<T, 1, false>(h), I1());
        }
    } else if (h.lev[l].cls) {
        if (h.block == 0) f(cls_prov<T, 3>(h.lev[l]), I3());
        else f(cls_prov<T, 1>(h.lev[l]), I1());
    } else {
        if (h.block == 0) f(expl_prov<T, 3>(h.lev[l]), I3());
        else f(expl_prov<T, 1>(h.lev[l]), I1());
    }
}

template <typename T>
static A0Z<T> a0z(const Hierarchy<T> &h) {
    A0Z<T> z;
    z.g = h.g;
    z.zt = h.zt;
    for (int a = 0; a < 3; ++a) { z.im[a] = h.im[a]; z.sg[a] = h.sg[a]; }
    z.cls = h.cls;
    z.ctab = h.ctab;
    z.fk = h.fk;
    z.rcls = h.block == 1 ? h.lev[0].cls : nullptr;
    z.rtab = h.block == 1 ? h.lev[0].tab : nullptr;
    return z;
}

// Level 0 of either block, z-march form (MODE 0: sweep, 1: residual).
template <typename T, int MODE>
static void a0_launch(const Hierarchy<T> &h, const T *x, const T *b, T *out, T w, const Prolong<T> &pr,
                      hipStream_t s) {
    const dim3 grid(h.zcols), block(256);
    auto go = [&](auto nc, auto cls, auto pro) {
        constexpr int NC = decltype(nc)::value;
        constexpr bool CLS = decltype(cls)::value, PRO = decltype(pro)::value;
        hipLaunchKernelGGL((k_a0_zm<T, NC, MODE, CLS, PRO>), grid, block, 0, s, a0z(h), x, b, out, w, pr,
                           NoRestrict{});
    };
    using I3 = std::integral_constant<int, 3>;
    using I1 = std::integral_constant<int, 1>;
    using Y = std::true_type;
    using N = std::false_type;
    const bool pro = MODE == 0 && pr.xc != nullptr;
    if (h.block == 1 && h.lev[0].cls) {
        if (pro)
            hipLaunchKernelGGL((k_a0_zm<T, 1, MODE, false, true, true>), grid, block, 0, s, a0z(h), x, b, out, w, pr,
                               NoRestrict{});
        else
            hipLaunchKernelGGL((k_a0_zm<T, 1, MODE, false, false, true>), grid, block, 0, s, a0z(h), x, b, out, w, pr,
                               NoRestrict{});
        return;
    }
    if (h.block == 0) {
        if (h.cls) { if (pro) go(I3(), Y(), Y()); else go(I3(), Y(), N()); }
        else { if (pro) go(I3(), N(), Y()); else go(I3(), N(), N()); }
    } else {
        if (h.cls) { if (pro) go(I1(), Y(), Y()); else go(I1(), Y(), N()); }
        else { if (pro) go(I1(), N(), Y()); else go(I1(), N(), N()); }
    }
}

// Level 0's residual restricted to level 1, with level 1's first sweep from
// zero (MODE 3 of the z-march), when level 1 is class-compressed and the
// z-chunk even; false: nothing launched.
template <typename T>
static bool a0_restrict_launch(const Hierarchy<T> &h, const T *x, const T *b, T *bc, T *xf, T w1, hipStream_t s) {
    if (h.zt.kz % 2) return false;
    bool done = false;
    const dim3 grid(h.zcols), block(256);
    const Prolong<T> none{nullptr, LevelDims{0, 0, 0, 0}, (T)0};
    on_level(h, 1, [&](auto pv, auto ncc) {
        using P = decltype(pv);
        constexpr int NC = decltype(ncc)::value;
        if constexpr (is_coarse_prov<P>::value) {
            using RR = A0Restrict<T, P>;
            const RR rr{pv, bc, xf, w1};
            auto go = [&](auto cls, auto rc) {
                constexpr bool CLS = decltype(cls)::value, RC = decltype(rc)::value;
                hipLaunchKernelGGL((k_a0_zm<T, NC, 3, CLS, false, RC, RR>), grid, block, 0, s, a0z(h), x, b,
                                   (T *)nullptr, (T)0, none, rr);
            };
            using Y = std::true_type;
            using N = std::false_type;
            if constexpr (NC == 1) {
                if (h.lev[0].cls) go(N(), Y());
                else if (h.cls) go(Y(), N());
                else go(N(), N());
            } else {
                if (h.cls) go(Y(), N());
                else go(N(), N());
            }
            done = true;
        }
    });
    return done;
}

// One smoother sweep of level l: from x = 0 (first), or from x plus the
// next level's correction xc scaled by alpha (pr.xc non-null: the
// prolongation fused into the first post-smoothing sweep).
template <typename T>
static void smooth(const Hierarchy<T> &h, int l, const T *x, const T *b, T *xo, T w, bool first, hipStream_t s,
                   const Prolong<T> &pr = Prolong<T>{nullptr, LevelDims{0, 0, 0, 0}, (T)0}) {
    const int64_t n1 = h.lev[l].d.n1;
    if (l == 0 && !first) { a0_launch<T, 0>(h, x, b, xo, w, pr, s); return; }
    on_level(h, l, [&](auto pv, auto nc) {
        using P = decltype(pv);
        constexpr int NC = decltype(nc)::value;
        if constexpr (is_coarse_prov<P>::value) {
            if (first) hipLaunchKernelGGL((k_mg_cop<T, P, NC, 1, false>), g1(n1), dim3(256), 0, s, pv, x, b, xo, w, pr);
            else if (pr.xc) hipLaunchKernelGGL((k_mg_cop<T, P, NC, 0, true>), g1(n1), dim3(256), 0, s, pv, x, b, xo, w, pr);
            else hipLaunchKernelGGL((k_mg_cop<T, P, NC, 0, false>), g1(n1), dim3(256), 0, s, pv, x, b, xo, w, pr);
        } else {
            if (first) hipLaunchKernelGGL((k_mg_smooth<T, P, NC, true>), g1(n1), dim3(256), 0, s, pv, x, b, xo, w);
            else hipLaunchKernelGGL((k_mg_smooth<T, P, NC, false>), g1(n1), dim3(256), 0, s, pv, x, b, xo, w);
        }
    });
}
template <typename T>
static void resid(const Hierarchy<T> &h, int l, const T *x, const T *b, T *r, hipStream_t s) {
    const int64_t n1 = h.lev[l].d.n1;
    const Prolong<T> none{nullptr, LevelDims{0, 0, 0, 0}, (T)0};
    if (l == 0) { a0_launch<T, 1>(h, x, b, r, (T)0, none, s); return; }
    on_level(h, l, [&](auto pv, auto nc) {
        using P = decltype(pv);
        constexpr int NC = decltype(nc)::value;
        if constexpr (is_coarse_prov<P>::value)
            hipLaunchKernelGGL((k_mg_cop<T, P, NC, 2, false>), g1(n1), dim3(256), 0, s, pv, x, b, r, (T)0, none);
        else
            hipLaunchKernelGGL((k_mg_resid<T, P, NC>), g1(n1), dim3(256), 0, s, pv, x, b, r);
    });
}
double sweep_weight(const MGParams &mp, int nu, int it, bool post) {
    if (!(mp.cheb_lmax > 0.0)) return mp.omega;
    const double hi = mp.cheb_lmax, lo = hi / mp.cheb_ratio;
    const int k = post ? nu - 1 - it : it;
    return 1.0 / (0.5 * (hi + lo) + 0.5 * (hi - lo) * cos((2.0 * k + 1.0) * M_PI / (2.0 * nu)));
}

// Smoothing sweeps before and after the coarse correction on level l.
static inline int level_nu(const MGParams &mp, int l) { return l == 0 ? mp.nu : mp.nu_coarse; }

// Whether level l (above the coarsest) and the coarsest solve run as one
// k_mg_tail launch: l is a coarse level whose nodes, vectors and coarse
// level fit one workgroup.
template <typename T>
static bool tail_fits(const Hierarchy<T> &h, const MGParams &mp, int l) {
    if (!mp.tail || l < 1 || l != h.nlev - 2 || level_nu(mp, l) > MG_TAIL_MAX_NU) return false;
    const MgLevel<T> &L = h.lev[l], &K = h.lev[l + 1];
    return L.d.n1 <= MG_TAIL_THREADS && (int64_t)L.ncomp * L.d.n1 <= mg_tail_rows<T>() &&
           (int64_t)K.ncomp * K.d.n1 <= MG_COARSE_MAX && h.ncoarse <= MG_COARSE_MAX;
}

// Where level l's first sweep from zero writes: the level's scratch vector
// L.t.  The sweeps then ping-pong between L.t and x, so pre-smoothing ends
// in L.t for odd nu and in x for even nu, and the nu post-smoothing sweeps
// that start there end in x either way, with no copy.
template <typename T>
static T *first_target(Hierarchy<T> &h, const MGParams &, int l) {
    return h.lev[l].t;
}

// first_done: the caller's restriction already wrote the first sweep from
// zero into first_target (coarse levels).
template <typename T>
static void vcycle_rec(Hierarchy<T> &h, const MGParams &mp, int l, T *x, const T *b, hipStream_t s,
                       bool first_done = false) {
    MgLevel<T> &L = h.lev[l];
    const int nc = L.ncomp;
    const int64_t nv = (int64_t)nc * L.d.n1;
    if (l == h.nlev - 1) {
        hipLaunchKernelGGL(k_mg_coarse<T>, dim3(1), dim3(MG_COARSE_THREADS), 0, s, h.ncoarse, nv, (const int *)h.coarse_idx,
                           (const double *)h.coarse_inv, b, x);
        return;
    }
    const int nu = level_nu(mp, l);
    // pre-smoothing from zero, ping-pong between L.t and x
    T *cur = L.t, *oth = x;
    if (!first_done) smooth<T>(h, l, nullptr, b, cur, (T)sweep_weight(mp, nu, 0, false), true, s);
    for (int it = 1; it < nu; ++it) {
        smooth<T>(h, l, cur, b, oth, (T)sweep_weight(mp, nu, it, false), false, s);
        std::swap(cur, oth);
    }
    MgLevel<T> &C = h.lev[l + 1];
    // level 0: the residual, its restriction and level 1's first sweep in one
    // launch where that applies (level 1 neither the tail nor the coarsest)
    const bool restricted = l == 0 && mp.fuse_restrict && l + 1 < h.nlev - 1 && !tail_fits(h, mp, l + 1) &&
                            a0_restrict_launch(h, (const T *)cur, b, C.b, first_target(h, mp, l + 1),
                                               (T)sweep_weight(mp, level_nu(mp, l + 1), 0, false), s);
    if (l == 0) h.fused_restrict = restricted;
    if (!restricted) resid<T>(h, l, cur, b, L.r, s);
    if (tail_fits(h, mp, l + 1)) {
        // levels l+1 .. nlev-1 in one workgroup, leaving level l+1's x
        const MgLevel<T> &K = h.lev[h.nlev - 1];
        const int nt = level_nu(mp, l + 1);
        TailWeights<T> tw;
        for (int it = 0; it < MG_TAIL_MAX_NU; ++it) {
            tw.pre[it] = it < nt ? (T)sweep_weight(mp, nt, it, false) : (T)0;
            tw.post[it] = it < nt ? (T)sweep_weight(mp, nt, it, true) : (T)0;
        }
        on_level(h, l + 1, [&](auto pv, auto ncc) {
            using P = decltype(pv);
            constexpr int NC = decltype(ncc)::value;
            if constexpr (is_coarse_prov<P>::value)
                hipLaunchKernelGGL((k_mg_tail<T, P, NC>), dim3(1), dim3(MG_TAIL_THREADS), 0, s, pv, L.d,
                                   (const T *)L.r, K.d, h.ncoarse, (const int *)h.coarse_idx,
                                   (const double *)h.coarse_inv, C.x, tw, (T)mp.alpha, nt);
        });
    } else if (l + 1 < h.nlev - 1) {
        // restriction fused with the next level's first sweep
        T *xf = first_target(h, mp, l + 1);
        if (!restricted)
            on_level(h, l + 1, [&](auto pv, auto ncc) {
                using P = decltype(pv);
                constexpr int NC = decltype(ncc)::value;
                if constexpr (is_coarse_prov<P>::value)
                    hipLaunchKernelGGL((k_mg_restrict_first<T, P, NC>), g1(C.d.n1), dim3(256), 0, s, pv, L.d,
                                       (const T *)L.r, C.b, xf, (T)sweep_weight(mp, level_nu(mp, l + 1), 0, false));
            });
        vcycle_rec<T>(h, mp, l + 1, C.x, C.b, s, true);
    } else {
        hipLaunchKernelGGL(k_mg_restrict<T>, g1(C.d.n1), dim3(256), 0, s, L.d, C.d, nc, (const T *)L.r, C.b);
        vcycle_rec<T>(h, mp, l + 1, C.x, C.b, s);
    }
    // post-smoothing from where the pre-smoothing ended (ending in x); the
    // first sweep reads x + alpha * P xc (the prolongation)
    oth = cur == x ? L.t : x;
    for (int it = 0; it < nu; ++it) {
        const T w = (T)sweep_weight(mp, nu, it, true);
        if (it == 0) smooth<T>(h, l, cur, b, oth, w, false, s, Prolong<T>{C.x, C.d, (T)mp.alpha});
        else smooth<T>(h, l, cur, b, oth, w, false, s);
        std::swap(cur, oth);
    }
    if (cur != x) hipLaunchKernelGGL(k_copy_t<T>, g1(nv), dim3(256), 0, s, x, (const T *)cur, nv);
}

template <typename T>
void mg_vcycle(Hierarchy<T> &h, const MGParams &mp, T *x0, const T *b0, hipStream_t s, bool first_done) {
    vcycle_rec<T>(h, mp, 0, x0, b0, s, first_done);
}

template <typename T>
T *mg_first_target(Hierarchy<T> &h, const MGParams &mp, T *) {
    return first_target(h, mp, 0);
}

template <typename T>
void launch_first_sweep(const Hierarchy<T> &h, const MGParams &mp, const T *b, T *xf, hipStream_t s) {
    const int64_t n = (int64_t)h.lev[0].ncomp * h.lev[0].d.n1;
    hipLaunchKernelGGL(k_first_sweep<T>, g1(n), dim3(256), 0, s, b, (const T *)h.lev[0].dg,
                       (T)sweep_weight(mp, level_nu(mp, 0), 0, false), xf, n);
}

template <typename TS, typename T>
void launch_d2t_first(const Hierarchy<T> &h, const MGParams &mp, const TS *src, const double *scale, T *dst, T *xf,
                      hipStream_t s) {
    const int64_t n = (int64_t)h.lev[0].ncomp * h.lev[0].d.n1;
    hipLaunchKernelGGL((k_d2t_first<TS, T>), g1(n), dim3(256), 0, s, src, scale, (const T *)h.lev[0].dg,
                       (T)sweep_weight(mp, level_nu(mp, 0), 0, false), dst, xf, n);
}

template <typename T>
void mg_fine_sweep(Hierarchy<T> &h, const MGParams &mp, const T *x, const T *b, T *xo, hipStream_t s) {
    smooth<T>(h, 0, x, b, xo, (T)mp.omega, false, s);
}

template <typename T>
void launch_fill(T *x, int64_t n, T v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill<T>, g1(n), dim3(256), 0, s, x, n, v);
}

template void mg_setup<float>(const Grid &, const HostCoef &, const CoeffsT<float> &, double, int, const MgAlloc &,
                              Hierarchy<float> &);
template void mg_setup<double>(const Grid &, const HostCoef &, const CoeffsT<double> &, double, int, const MgAlloc &,
                               Hierarchy<double> &);
template void mg_vcycle<float>(Hierarchy<float> &, const MGParams &, float *, const float *, hipStream_t, bool);
template void mg_vcycle<double>(Hierarchy<double> &, const MGParams &, double *, const double *, hipStream_t, bool);
template float *mg_first_target<float>(Hierarchy<float> &, const MGParams &, float *);
template double *mg_first_target<double>(Hierarchy<double> &, const MGParams &, double *);
template void launch_first_sweep<float>(const Hierarchy<float> &, const MGParams &, const float *, float *,
                                        hipStream_t);
template void launch_first_sweep<double>(const Hierarchy<double> &, const MGParams &, const double *, double *,
                                         hipStream_t);
template void launch_d2t_first<float, float>(const Hierarchy<float> &, const MGParams &, const float *,
                                             const double *, float *, float *, hipStream_t);
template void launch_d2t_first<double, double>(const Hierarchy<double> &, const MGParams &, const double *,
                                               const double *, double *, double *, hipStream_t);
template void mg_fine_sweep<float>(Hierarchy<float> &, const MGParams &, const float *, const float *, float *,
                                   hipStream_t);
template void mg_fine_sweep<double>(Hierarchy<double> &, const MGParams &, const double *, const double *, double *,
                                    hipStream_t);
template void launch_fill<float>(float *, int64_t, float, hipStream_t);
template void launch_fill<double>(double *, int64_t, double, hipStream_t);

// ---------------------------------------------------------------------------
// Vector conversions at the Krylov / multigrid boundary
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_cvt(const double *__restrict__ src, float *__restrict__ dst, int64_t n) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) dst[p] = (float)src[p];
}
void launch_cvt(const double *src, float *dst, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_cvt, g1(n), dim3(256), 0, s, src, dst, n);
}

}  // namespace qms
