// Fused matrix-free operator of the coupled A-V system on the Yee grid, and
// the per-step kernels around it (coefficients, diagonal, RHS, Dirichlet
// values, cell-centred H).
//
// In the operator one thread owns one node p = (i,j,k) and evaluates the
// four rows (Ax, Ay, Az, V) that live there, exactly the rows the reference
// assembles into CSR every step (steady QMS_3D.f90:448-1293, transient
// :1457-2189).  Arithmetic is FP64 whatever the storage type of the input
// vector: the V rows mix conductivities 1e-3..6e7 S/m, and only FP64
// evaluation keeps the weak background couplings that pin each conductor's
// floating potential.  Self couplings are evaluated in difference form
// c*(x_p - x_q) so that a constant on a conductor produces its true (tiny)
// residual.
#include <hip/hip_runtime.h>
#include <type_traits>
#include "qms_kernels.h"

// Multiply-adds are fused only within one source expression, never across
// statements: the same row code then rounds identically in every kernel that
// inlines it (single-design and batched launches give bit-identical rows).
#pragma clang fp contract(on)

namespace qms {

// ---------------------------------------------------------------------------
// Material IDs -> Yee coefficients, QMS_3D.f90:211-269.
// slot: cells (cx*cy*cz, x fastest) holding 0 (vacuum) or 1..nmat.
// mat:  per slot {mur_x, mur_y, mur_z, sig_x, sig_y, sig_z}, slot 0 = vacuum.
// Outputs node arrays 1/Mu_x, 1/Mu_y, 1/Mu_z, Sigma_x, Sigma_y, Sigma_z
// (FP64) and FP32 copies for the multigrid smoother.
// ---------------------------------------------------------------------------
__global__ void k_coeffs(Grid g, const uint16_t *__restrict__ slot, const double *__restrict__ mat,
                         Coeffs c) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.n1) return;
    const int i = (int)(p % g.nx), j = (int)((p / g.nx) % g.ny), k = (int)(p / ((int64_t)g.nx * g.ny));
    const int cx = g.nx - 1, cy = g.ny - 1;
    double mux = QMS_MU0, muy = QMS_MU0, muz = QMS_MU0, sx = 0.0, sy = 0.0, sz = 0.0;
    // 1-based Yee means for indices 2..cells  <=>  0-based nodes 1..cells-1
    if (i >= 1 && i <= g.nx - 2 && j >= 1 && j <= g.ny - 2 && k >= 1 && k <= g.nz - 2) {
        auto cell = [&](int a, int b, int d) -> const double * {
            return mat + 6 * slot[(int64_t)a + (int64_t)cx * ((int64_t)b + (int64_t)cy * d)];
        };
        const double *c0 = cell(i, j, k);
        const double *cxm = cell(i - 1, j, k), *cym = cell(i, j - 1, k), *czm = cell(i, j, k - 1);
        const double *cyzm = cell(i, j - 1, k - 1), *cxzm = cell(i - 1, j, k - 1), *cxym = cell(i - 1, j - 1, k);
        double a, b;
        a = c0[0] * QMS_MU0; b = cxm[0] * QMS_MU0; mux = 2.0 * a * b / (a + b);
        a = c0[1] * QMS_MU0; b = cym[1] * QMS_MU0; muy = 2.0 * a * b / (a + b);
        a = c0[2] * QMS_MU0; b = czm[2] * QMS_MU0; muz = 2.0 * a * b / (a + b);
        sx = (c0[3] + czm[3] + cym[3] + cyzm[3]) / 4.0;
        sy = (c0[4] + cxm[4] + czm[4] + cxzm[4]) / 4.0;
        sz = (c0[5] + cxm[5] + cym[5] + cxym[5]) / 4.0;
    }
    const double imx = 1.0 / mux, imy = 1.0 / muy, imz = 1.0 / muz;
    c.im[0][p] = imx; c.im[1][p] = imy; c.im[2][p] = imz;
    c.sg[0][p] = sx; c.sg[1][p] = sy; c.sg[2][p] = sz;
    c.im32[0][p] = (float)imx; c.im32[1][p] = (float)imy; c.im32[2][p] = (float)imz;
    c.sg32[0][p] = (float)sx; c.sg32[1][p] = (float)sy; c.sg32[2][p] = (float)sz;
}

// ---------------------------------------------------------------------------
// Row evaluation.  Each function returns (M x) for its row and the row's
// diagonal (the scaling D used by the Krylov solver).  `k.tr` enables the
// transient sigma/dt terms, `k.adv` the frame-velocity terms.
//
// Accessors take a node offset o = di + dj*sy + dk*sz relative to the row's
// node; the field and coefficient pointers are already advanced to it.
// ---------------------------------------------------------------------------
template <typename T>
struct View {
    const T *ax, *ay, *az, *v;
    int sy, sz;
    __device__ __forceinline__ int off(int di, int dj, int dk) const { return di + dj * sy + dk * sz; }
    __device__ __forceinline__ double AX(int di, int dj, int dk) const { return (double)ax[off(di, dj, dk)]; }
    __device__ __forceinline__ double AY(int di, int dj, int dk) const { return (double)ay[off(di, dj, dk)]; }
    __device__ __forceinline__ double AZ(int di, int dj, int dk) const { return (double)az[off(di, dj, dk)]; }
    __device__ __forceinline__ double VV(int di, int dj, int dk) const { return (double)v[off(di, dj, dk)]; }
};

struct CView {
    const double *imx, *imy, *imz, *sx, *sy_, *sz_;
    int sy, sz;
    __device__ __forceinline__ int off(int di, int dj, int dk) const { return di + dj * sy + dk * sz; }
    __device__ __forceinline__ double IMX(int di, int dj, int dk) const { return imx[off(di, dj, dk)]; }
    __device__ __forceinline__ double IMY(int di, int dj, int dk) const { return imy[off(di, dj, dk)]; }
    __device__ __forceinline__ double IMZ(int di, int dj, int dk) const { return imz[off(di, dj, dk)]; }
    __device__ __forceinline__ double SX(int di, int dj, int dk) const { return sx[off(di, dj, dk)]; }
    __device__ __forceinline__ double SY(int di, int dj, int dk) const { return sy_[off(di, dj, dk)]; }
    __device__ __forceinline__ double SZ(int di, int dj, int dk) const { return sz_[off(di, dj, dk)]; }
};

// Row-uniform views: one wave covers 64 consecutive nodes of one grid row.
// The field pointers are wave-uniform; each lane adds its own byte offset.
template <typename T>
__device__ __forceinline__ View<T> view_at(const T *x, int64_t n1, int64_t p, const OpK &k) {
    return View<T>{x + p, x + n1 + p, x + 2 * n1 + p, x + 3 * n1 + p, k.sy, k.sz};
}
__device__ __forceinline__ CView cview_at(const CoeffsD &c, int64_t p, const OpK &k) {
    return CView{c.im[0] + p, c.im[1] + p, c.im[2] + p, c.sg[0] + p, c.sg[1] + p, c.sg[2] + p, k.sy, k.sz};
}

// Which blocks of the row to include.
enum : int { B_AA = 1, B_AV = 2, B_VA = 4, B_VV = 8, B_ALL = 15 };

// Ax row, QMS_3D.f90:476-607 (+ :1504, :1475)
template <typename XV, typename CV>
__device__ __forceinline__ double row_ax(const XV &x, const CV &c, const OpK &k, int blocks, double &diag) {
    const double s = c.SX(0, 0, 0);
    const double imy_km = c.IMY(0, 0, -1), imy_k = c.IMY(0, 0, 0);
    const double imz_jm = c.IMZ(0, -1, 0), imz_j = c.IMZ(0, 0, 0);
    const double imx_im = c.IMX(-1, 0, 0), imx_i = c.IMX(0, 0, 0);
    const double ihx2 = k.ih2[0], ihy2 = k.ih2[1], ihz2 = k.ih2[2];
    double d = (imz_j + imz_jm) * ihy2 + (imy_k + imy_km) * ihz2 + (imx_i + imx_im) * ihx2;
    if (k.tr) d = d + s * k.idt;
    diag = d;
    double y = 0.0;
    if (blocks & B_AA) {
        const double x0 = x.AX(0, 0, 0);
        const double xzm = x.AX(0, 0, -1), xym = x.AX(0, -1, 0), xxm = x.AX(-1, 0, 0), xxp = x.AX(1, 0, 0), xyp = x.AX(0, 1, 0), xzp = x.AX(0, 0, 1);
        y += imy_km * ihz2 * (x0 - xzm) + imz_jm * ihy2 * (x0 - xym) + imx_im * ihx2 * (x0 - xxm)
           + imx_i * ihx2 * (x0 - xxp) + imz_j * ihy2 * (x0 - xyp) + imy_k * ihz2 * (x0 - xzp);
        if (k.tr) y += (s * k.idt) * x0;
        const double ay_m0 = x.AY(0, -1, 0), ay_p0 = x.AY(1, -1, 0), ay_00 = x.AY(0, 0, 0), ay_10 = x.AY(1, 0, 0);
        const double az_m0 = x.AZ(0, 0, -1), az_p0 = x.AZ(1, 0, -1), az_00 = x.AZ(0, 0, 0), az_10 = x.AZ(1, 0, 0);
        const double ixy = k.ixy, ixz = k.ixz;
        y += imz_jm * ixy * (ay_m0 - ay_p0) + imz_j * ixy * (ay_10 - ay_00);
        y += imy_km * ixz * (az_m0 - az_p0) + imy_k * ixz * (az_10 - az_00);
        if (k.adv) {
            y += s * k.v1[2][2] * (xzm - xzp) + s * k.v1[1][1] * (xym - xyp);
            y += s * k.v1[1][0] * (-ay_m0 + ay_p0 - ay_00 + ay_10);
            y += s * k.v1[2][0] * (-az_m0 + az_p0 - az_00 + az_10);
        }
    }
    if (blocks & B_AV) y += s * k.ih[0] * (x.VV(1, 0, 0) - x.VV(0, 0, 0));
    return y;
}

// Ay row, QMS_3D.f90:650-781 (+ :1681, :1624)
template <typename XV, typename CV>
__device__ __forceinline__ double row_ay(const XV &x, const CV &c, const OpK &k, int blocks, double &diag) {
    const double s = c.SY(0, 0, 0);
    const double imx_km = c.IMX(0, 0, -1), imx_k = c.IMX(0, 0, 0);
    const double imy_jm = c.IMY(0, -1, 0), imy_j = c.IMY(0, 0, 0);
    const double imz_im = c.IMZ(-1, 0, 0), imz_i = c.IMZ(0, 0, 0);
    const double ihx2 = k.ih2[0], ihy2 = k.ih2[1], ihz2 = k.ih2[2];
    double d = (imz_i + imz_im) * ihx2 + (imx_k + imx_km) * ihz2 + (imy_j + imy_jm) * ihy2;
    if (k.tr) d = d + s * k.idt;
    diag = d;
    double y = 0.0;
    if (blocks & B_AA) {
        const double x0 = x.AY(0, 0, 0);
        const double xzm = x.AY(0, 0, -1), xym = x.AY(0, -1, 0), xxm = x.AY(-1, 0, 0), xxp = x.AY(1, 0, 0), xyp = x.AY(0, 1, 0), xzp = x.AY(0, 0, 1);
        y += imx_km * ihz2 * (x0 - xzm) + imy_jm * ihy2 * (x0 - xym) + imz_im * ihx2 * (x0 - xxm)
           + imz_i * ihx2 * (x0 - xxp) + imy_j * ihy2 * (x0 - xyp) + imx_k * ihz2 * (x0 - xzp);
        if (k.tr) y += (s * k.idt) * x0;
        const double ax_m0 = x.AX(-1, 0, 0), ax_00 = x.AX(0, 0, 0), ax_m1 = x.AX(-1, 1, 0), ax_01 = x.AX(0, 1, 0);
        const double az_m0 = x.AZ(0, 0, -1), az_m1 = x.AZ(0, 1, -1), az_00 = x.AZ(0, 0, 0), az_01 = x.AZ(0, 1, 0);
        const double ixy = k.ixy, iyz = k.iyz;
        y += imz_im * ixy * (ax_m0 - ax_m1) + imz_i * ixy * (ax_01 - ax_00);
        y += imx_km * iyz * (az_m0 - az_m1) + imx_k * iyz * (az_01 - az_00);
        if (k.adv) {
            y += s * k.v1[2][2] * (xzm - xzp) + s * k.v1[0][0] * (xxm - xxp);
            y += s * k.v1[0][1] * (-ax_m0 - ax_00 + ax_m1 + ax_01);
            y += s * k.v1[2][1] * (-az_m0 + az_m1 - az_00 + az_01);
        }
    }
    if (blocks & B_AV) y += s * k.ih[1] * (x.VV(0, 1, 0) - x.VV(0, 0, 0));
    return y;
}

// Az row, QMS_3D.f90:824-955 (+ :1857, :1772)
template <typename XV, typename CV>
__device__ __forceinline__ double row_az(const XV &x, const CV &c, const OpK &k, int blocks, double &diag) {
    const double s = c.SZ(0, 0, 0);
    const double imz_km = c.IMZ(0, 0, -1), imz_k = c.IMZ(0, 0, 0);
    const double imx_jm = c.IMX(0, -1, 0), imx_j = c.IMX(0, 0, 0);
    const double imy_im = c.IMY(-1, 0, 0), imy_i = c.IMY(0, 0, 0);
    const double ihx2 = k.ih2[0], ihy2 = k.ih2[1], ihz2 = k.ih2[2];
    double d = (imy_i + imy_im) * ihx2 + (imx_j + imx_jm) * ihy2 + (imz_k + imz_km) * ihz2;
    if (k.tr) d = d + s * k.idt;
    diag = d;
    double y = 0.0;
    if (blocks & B_AA) {
        const double x0 = x.AZ(0, 0, 0);
        const double xzm = x.AZ(0, 0, -1), xym = x.AZ(0, -1, 0), xxm = x.AZ(-1, 0, 0), xxp = x.AZ(1, 0, 0), xyp = x.AZ(0, 1, 0), xzp = x.AZ(0, 0, 1);
        y += imz_km * ihz2 * (x0 - xzm) + imx_jm * ihy2 * (x0 - xym) + imy_im * ihx2 * (x0 - xxm)
           + imy_i * ihx2 * (x0 - xxp) + imx_j * ihy2 * (x0 - xyp) + imz_k * ihz2 * (x0 - xzp);
        if (k.tr) y += (s * k.idt) * x0;
        const double ax_m0 = x.AX(-1, 0, 0), ax_00 = x.AX(0, 0, 0), ax_m1 = x.AX(-1, 0, 1), ax_01 = x.AX(0, 0, 1);
        const double ay_m0 = x.AY(0, -1, 0), ay_00 = x.AY(0, 0, 0), ay_m1 = x.AY(0, -1, 1), ay_01 = x.AY(0, 0, 1);
        const double ixz = k.ixz, iyz = k.iyz;
        y += imy_im * ixz * (ax_m0 - ax_m1) + imy_i * ixz * (ax_01 - ax_00);
        y += imx_jm * iyz * (ay_m0 - ay_m1) + imx_j * iyz * (ay_01 - ay_00);
        if (k.adv) {
            y += s * k.v1[1][1] * (xym - xyp) + s * k.v1[0][0] * (xxm - xxp);
            y += s * k.v1[0][2] * (-ax_m0 - ax_00 + ax_m1 + ax_01);
            y += s * k.v1[1][2] * (-ay_m0 - ay_00 + ay_m1 + ay_01);
        }
    }
    if (blocks & B_AV) y += s * k.ih[2] * (x.VV(0, 0, 1) - x.VV(0, 0, 0));
    return y;
}

// V row, QMS_3D.f90:1003-1285 (+ :1975-2126 transient terms)
template <typename XV, typename CV>
__device__ __forceinline__ double row_v(const XV &x, const CV &c, const OpK &k, int blocks, double &diag) {
    const double ihx2 = k.ih2[0], ihy2 = k.ih2[1], ihz2 = k.ih2[2];
    const double x0 = c.SX(0, 0, 0), x1 = c.SX(-1, 0, 0);
    const double y0 = c.SY(0, 0, 0), y1 = c.SY(0, -1, 0);
    const double z0 = c.SZ(0, 0, 0), z1 = c.SZ(0, 0, -1);
    const double fx0 = fmax(x0, QMS_SIGFLOOR), fx1 = fmax(x1, QMS_SIGFLOOR);
    const double fy0 = fmax(y0, QMS_SIGFLOOR), fy1 = fmax(y1, QMS_SIGFLOOR);
    const double fz0 = fmax(z0, QMS_SIGFLOOR), fz1 = fmax(z1, QMS_SIGFLOOR);
    diag = (fx0 + fx1) * ihx2 + (fy0 + fy1) * ihy2 + (fz0 + fz1) * ihz2;
    double y = 0.0;
    if (blocks & B_VV) {
        const double v0 = x.VV(0, 0, 0);
        y += fz1 * ihz2 * (v0 - x.VV(0, 0, -1)) + fy1 * ihy2 * (v0 - x.VV(0, -1, 0)) + fx1 * ihx2 * (v0 - x.VV(-1, 0, 0))
           + fx0 * ihx2 * (v0 - x.VV(1, 0, 0)) + fy0 * ihy2 * (v0 - x.VV(0, 1, 0)) + fz0 * ihz2 * (v0 - x.VV(0, 0, 1));
    }
    if (blocks & B_VA) {
        if (k.tr) {
            y += x0 * k.iht[0] * x.AX(0, 0, 0) - x1 * k.iht[0] * x.AX(-1, 0, 0)
               + y0 * k.iht[1] * x.AY(0, 0, 0) - y1 * k.iht[1] * x.AY(0, -1, 0)
               + z0 * k.iht[2] * x.AZ(0, 0, 0) - z1 * k.iht[2] * x.AZ(0, 0, -1);
        }
        if (k.adv) {
            // k.w holds the twelve v_a / (2 h_b h_c) of this row (axes x=0, y=1, z=2)
            double w = 0.0;
            // Ax columns
            w += (x1 * k.w[0] - z1 * k.w[1]) * x.AX(-1, 0, -1);
            w += (-x0 * k.w[0] - z1 * k.w[1]) * x.AX(0, 0, -1);
            w += (x1 * k.w[2] - y1 * k.w[3]) * x.AX(-1, -1, 0);
            w += (-x0 * k.w[2] - y1 * k.w[3]) * x.AX(0, -1, 0);
            const double axc = (y0 + y1) * k.w[3] + (z0 + z1) * k.w[1];
            w += axc * (x.AX(-1, 0, 0) + x.AX(0, 0, 0));
            w += (-x1 * k.w[2] - y0 * k.w[3]) * x.AX(-1, 1, 0);
            w += (x0 * k.w[2] - y0 * k.w[3]) * x.AX(0, 1, 0);
            w += (-x1 * k.w[0] - z0 * k.w[1]) * x.AX(-1, 0, 1);
            w += (x0 * k.w[0] - z0 * k.w[1]) * x.AX(0, 0, 1);
            // Ay columns
            w += (y1 * k.w[4] - z1 * k.w[5]) * x.AY(0, -1, -1);
            w += (-y0 * k.w[4] - z1 * k.w[5]) * x.AY(0, 0, -1);
            w += (-x1 * k.w[6] + y1 * k.w[7]) * x.AY(-1, -1, 0);
            const double ayc = (x0 + x1) * k.w[6] + (z0 + z1) * k.w[5];
            w += ayc * (x.AY(0, -1, 0) + x.AY(0, 0, 0));
            w += (-x0 * k.w[6] - y1 * k.w[7]) * x.AY(1, -1, 0);
            w += (-x1 * k.w[6] - y0 * k.w[7]) * x.AY(-1, 0, 0);
            w += (-x0 * k.w[6] + y0 * k.w[7]) * x.AY(1, 0, 0);
            w += (-y1 * k.w[4] - z0 * k.w[5]) * x.AY(0, -1, 1);
            w += (y0 * k.w[4] - z0 * k.w[5]) * x.AY(0, 0, 1);
            // Az columns
            w += (-y1 * k.w[8] + z1 * k.w[9]) * x.AZ(0, -1, -1);
            w += (-x1 * k.w[10] + z1 * k.w[11]) * x.AZ(-1, 0, -1);
            const double azc = (x0 + x1) * k.w[10] + (y0 + y1) * k.w[8];
            w += azc * (x.AZ(0, 0, -1) + x.AZ(0, 0, 0));
            w += (-x0 * k.w[10] - z1 * k.w[11]) * x.AZ(1, 0, -1);
            w += (-y0 * k.w[8] - z1 * k.w[9]) * x.AZ(0, 1, -1);
            w += (-y1 * k.w[8] - z0 * k.w[9]) * x.AZ(0, -1, 0);
            w += (-x1 * k.w[10] - z0 * k.w[11]) * x.AZ(-1, 0, 0);
            w += (-x0 * k.w[10] + z0 * k.w[11]) * x.AZ(1, 0, 0);
            w += (-y0 * k.w[8] + z0 * k.w[9]) * x.AZ(0, 1, 0);
            y += w;
        }
    }
    return y;
}

// ---------------------------------------------------------------------------
// All four rows of one node in flux form: the same operator as row_ax..row_v,
// regrouped around the Yee grid's face and edge quantities so that each is
// formed once and shared between rows:
//   B  = curl A on the 12 faces around the node (z-faces at (0,0),(0,-1),
//        (-1,0),(-1,-1) in (i,j); y-faces and x-faces likewise),
//   F  = B / Mu on those faces      -> curl-curl part of the A rows,
//   P  = (1/Mu_a) dA_a/da           -> gauge-penalty part of the A rows,
//   E  = sigma (v x avg B) on edges -> advection of the A rows and, through
//        G^T, of the V row (QMS_3D.f90:546-592, :1003-1285),
//   Phi = max(sigma, floor) dV/da   -> Laplacian of the V row.
// A row:  curl(F) - div(P)_a + sigma A/dt + sigma dV/da + E     (+ transient)
// V row:  sum_a [ (sigma A/dt)_a - Phi_a - E_a ]_(-a)^(0) / h_a
// Every term is a difference of neighbouring values, as in the row form.
// ---------------------------------------------------------------------------
// 1/d for a positive normal d (a row diagonal): the hardware reciprocal
// estimate refined by two Newton steps, r <- r + r (1 - d r), which leaves
// it within about one unit in the last place -- cheaper than a correctly
// rounded division, whose result it may differ from in the last bit.
__device__ __forceinline__ double recip(double d) {
    double r = __builtin_amdgcn_rcp(d);
    r = fma(fma(-d, r, 1.0), r, r);
    r = fma(fma(-d, r, 1.0), r, r);
    return r;
}

// Faces of plane k-1 that the rows of plane k reuse: y-faces at (i,j) and
// (i-1,j), x-faces at (i,j) and (i,j-1).  A z-marching thread carries them
// from one plane to the next (`first`: compute them from plane k-1 instead).
struct FaceCarry {
    double by0, bym, bx0, bxm;
    double ez;             // sigma (v x avg B)_z on the Az edge of plane k-1
};

template <typename XV, typename CV>
__device__ __forceinline__ void rows_flux(const XV &x, const CV &c, const OpK &k, double r[4], double d[4],
                                          FaceCarry &fc, bool first) {
    const double ihx = k.ih[0], ihy = k.ih[1], ihz = k.ih[2];
    // face curls; Bz(di,dj) at plane k, By(di,dk) at row j, Bx(dj,dk) at column i
    auto Bz = [&](int di, int dj) {
        return (x.AY(di + 1, dj, 0) - x.AY(di, dj, 0)) * ihx - (x.AX(di, dj + 1, 0) - x.AX(di, dj, 0)) * ihy;
    };
    auto By = [&](int di, int dk) {
        return (x.AX(di, 0, dk + 1) - x.AX(di, 0, dk)) * ihz - (x.AZ(di + 1, 0, dk) - x.AZ(di, 0, dk)) * ihx;
    };
    auto Bx = [&](int dj, int dk) {
        return (x.AZ(0, dj + 1, dk) - x.AZ(0, dj, dk)) * ihy - (x.AY(0, dj, dk + 1) - x.AY(0, dj, dk)) * ihz;
    };
    if (first) {
        fc.by0 = By(0, -1);
        fc.bym = By(-1, -1);
        fc.bx0 = Bx(0, -1);
        fc.bxm = Bx(-1, -1);
    }
    const double bz00 = Bz(0, 0), bz0m = Bz(0, -1), bzm0 = Bz(-1, 0);
    const double by00 = By(0, 0), by0m = fc.by0, bym0 = By(-1, 0);
    const double bx00 = Bx(0, 0), bx0m = fc.bx0, bxm0 = Bx(-1, 0);
    const double imx0 = c.IMX(0, 0, 0), imy0 = c.IMY(0, 0, 0), imz0 = c.IMZ(0, 0, 0);
    const double imx_y = c.IMX(0, -1, 0), imx_z = c.IMX(0, 0, -1), imx_x = c.IMX(-1, 0, 0);
    const double imy_x = c.IMY(-1, 0, 0), imy_z = c.IMY(0, 0, -1), imy_y = c.IMY(0, -1, 0);
    const double imz_x = c.IMZ(-1, 0, 0), imz_y = c.IMZ(0, -1, 0), imz_z = c.IMZ(0, 0, -1);
    const double sx0 = c.SX(0, 0, 0), sy0 = c.SY(0, 0, 0), sz0 = c.SZ(0, 0, 0);
    const double sx1 = c.SX(-1, 0, 0), sy1 = c.SY(0, -1, 0), sz1 = c.SZ(0, 0, -1);
    // curl-curl
    const double fz0 = imz0 * bz00, fzy = imz_y * bz0m, fzx = imz_x * bzm0;
    const double fy0 = imy0 * by00, fyz = imy_z * by0m, fyx = imy_x * bym0;
    const double fx0 = imx0 * bx00, fxz = imx_z * bx0m, fxy = imx_y * bxm0;
    double rax = (fz0 - fzy) * ihy - (fy0 - fyz) * ihz;
    double ray = (fx0 - fxz) * ihz - (fz0 - fzx) * ihx;
    double raz = (fy0 - fyx) * ihx - (fx0 - fxy) * ihy;
    // gauge penalty
    const double ax0 = x.AX(0, 0, 0), ay0 = x.AY(0, 0, 0), az0 = x.AZ(0, 0, 0), v0 = x.VV(0, 0, 0);
    const double axm = x.AX(-1, 0, 0), aym = x.AY(0, -1, 0), azm = x.AZ(0, 0, -1);
    rax -= (imx0 * (x.AX(1, 0, 0) - ax0) - imx_x * (ax0 - axm)) * k.ih2[0];
    ray -= (imy0 * (x.AY(0, 1, 0) - ay0) - imy_y * (ay0 - aym)) * k.ih2[1];
    raz -= (imz0 * (x.AZ(0, 0, 1) - az0) - imz_z * (az0 - azm)) * k.ih2[2];
    // sigma grad V on the A rows
    const double vxp = x.VV(1, 0, 0), vyp = x.VV(0, 1, 0), vzp = x.VV(0, 0, 1);
    const double vxm = x.VV(-1, 0, 0), vym = x.VV(0, -1, 0), vzm = x.VV(0, 0, -1);
    rax += sx0 * (vxp - v0) * ihx;
    ray += sy0 * (vyp - v0) * ihy;
    raz += sz0 * (vzp - v0) * ihz;
    // V row: floored-sigma Laplacian
    const double fx0f = fmax(sx0, QMS_SIGFLOOR), fx1f = fmax(sx1, QMS_SIGFLOOR);
    const double fy0f = fmax(sy0, QMS_SIGFLOOR), fy1f = fmax(sy1, QMS_SIGFLOOR);
    const double fz0f = fmax(sz0, QMS_SIGFLOOR), fz1f = fmax(sz1, QMS_SIGFLOOR);
    double rv = -((fx0f * (vxp - v0) - fx1f * (v0 - vxm)) * k.ih2[0]
                + (fy0f * (vyp - v0) - fy1f * (v0 - vym)) * k.ih2[1]
                + (fz0f * (vzp - v0) - fz1f * (v0 - vzm)) * k.ih2[2]);
    double dax = (imz0 + imz_y) * k.ih2[1] + (imy0 + imy_z) * k.ih2[2] + (imx0 + imx_x) * k.ih2[0];
    double day = (imz0 + imz_x) * k.ih2[0] + (imx0 + imx_z) * k.ih2[2] + (imy0 + imy_y) * k.ih2[1];
    double daz = (imy0 + imy_x) * k.ih2[0] + (imx0 + imx_y) * k.ih2[1] + (imz0 + imz_z) * k.ih2[2];
    d[3] = (fx0f + fx1f) * k.ih2[0] + (fy0f + fy1f) * k.ih2[1] + (fz0f + fz1f) * k.ih2[2];
    if (k.tr) {
        dax = dax + sx0 * k.idt;
        day = day + sy0 * k.idt;
        daz = daz + sz0 * k.idt;
        rax += (sx0 * k.idt) * ax0;
        ray += (sy0 * k.idt) * ay0;
        raz += (sz0 * k.idt) * az0;
        rv += (sx0 * ax0 - sx1 * axm) * k.iht[0] + (sy0 * ay0 - sy1 * aym) * k.iht[1]
            + (sz0 * az0 - sz1 * azm) * k.iht[2];
    }
    if (k.adv) {
        const double hvx = k.hv[0], hvy = k.hv[1], hvz = k.hv[2];
        const double bzmm = Bz(-1, -1), bymm = fc.bym, bxmm = fc.bxm;
        // sigma (v x avg B) on the node's own edges and on the -x, -y, -z edges
        const double ex0 = sx0 * (hvy * (bz00 + bz0m) - hvz * (by00 + by0m));
        const double ey0 = sy0 * (hvz * (bx00 + bx0m) - hvx * (bz00 + bzm0));
        const double ez0 = sz0 * (hvx * (by00 + bym0) - hvy * (bx00 + bxm0));
        const double exm = sx1 * (hvy * (bzm0 + bzmm) - hvz * (bym0 + bymm));
        const double eym = sy1 * (hvz * (bxm0 + bxmm) - hvx * (bz0m + bzmm));
        // the Az edge of plane k-1 was the previous plane's own edge
        const double ezm = first ? sz1 * (hvx * (by0m + bymm) - hvy * (bx0m + bxmm)) : fc.ez;
        fc.ez = ez0;
        rax += ex0;
        ray += ey0;
        raz += ez0;
        rv -= (ex0 - exm) * ihx + (ey0 - eym) * ihy + (ez0 - ezm) * ihz;
    }
    r[0] = rax; r[1] = ray; r[2] = raz; r[3] = rv;
    d[0] = dax; d[1] = day; d[2] = daz;
    fc.by0 = by00;
    fc.bym = bym0;
    fc.bx0 = bx00;
    fc.bxm = bxm0;
}

// ---------------------------------------------------------------------------
// Tiled launch geometry.  Tiles are numbered x-fastest, then y, then z, and
// the hardware workgroup index is remapped so that each of the 8 XCDs (which
// receive workgroups round-robin) works on one contiguous range of tiles:
// neighbouring rows and planes then share that XCD's L2.  Plane tiles
// (Tiling): a workgroup of 256 threads owns tx x (256/tx) nodes of one
// z-plane.
// ---------------------------------------------------------------------------
struct Tiling {
    int txs;            // log2 of the tile width
    int ntx, nty;       // tiles per row / per plane column
};


__device__ __forceinline__ bool tile_node(const Grid &g, const Tiling &t, int &i, int &j, int &k) {
    const unsigned tid = xcd_tile(blockIdx.x, gridDim.x);
    const unsigned tx = tid % (unsigned)t.ntx, rest = tid / (unsigned)t.ntx;
    const unsigned ty = rest % (unsigned)t.nty;
    k = (int)(rest / (unsigned)t.nty);
    i = (int)((tx << t.txs) + (threadIdx.x & ((1u << t.txs) - 1u)));
    j = (int)(ty * (256u >> t.txs) + (threadIdx.x >> t.txs));
    return i < g.nx && j < g.ny && k < g.nz;
}

// ---------------------------------------------------------------------------
// z-marching operator.  A workgroup owns a TX x (256/TX) column of nodes
// (TX = 64 or 32) and walks it through kz consecutive planes.  The four
// fields of planes k-1, k, k+1 (with a one-node halo in x and y) sit in an
// LDS ring of three slots, so each field value is read from HBM once per
// column instead of once per stencil use.  The material coefficients of
// plane k are read per node, from the class table or the six arrays (their
// -x and -y neighbours hit L1), and the plane-(k-1) values they need are
// carried in registers from the previous plane.
// ---------------------------------------------------------------------------
template <int TX>
struct ZShape {
    static constexpr int TY = 256 / TX;              // rows per column
    static constexpr int SY = TX + 2;                // LDS row stride (one-node halo each side)
    static constexpr int PL = SY * (TY + 2);         // LDS plane size
};
constexpr int ZPL_MAX = ZShape<64>::PL > ZShape<32>::PL ? ZShape<64>::PL : ZShape<32>::PL;

template <int TX>
struct LdsView {
    const double *lds;     // [field][slot][PL]
    int b[3];              // slot offsets (elements) of planes k-1, k, k+1
    int ctr;               // this node's index within a plane slot
    __device__ __forceinline__ double get(int f, int di, int dj, int dk) const {
        return lds[f * 3 * ZShape<TX>::PL + b[dk + 1] + ctr + di + dj * ZShape<TX>::SY];
    }
    __device__ __forceinline__ double AX(int di, int dj, int dk) const { return get(0, di, dj, dk); }
    __device__ __forceinline__ double AY(int di, int dj, int dk) const { return get(1, di, dj, dk); }
    __device__ __forceinline__ double AZ(int di, int dj, int dk) const { return get(2, di, dj, dk); }
    __device__ __forceinline__ double VV(int di, int dj, int dk) const { return get(3, di, dj, dk); }
};

struct ZmCView {
    double imx0, imy0, imz0, sx0, sy0, sz0;    // at (0,0,0)
    double imxz, imyz, imzz, szz;              // at (0,0,-1)
    const double *imx, *imy, *imz, *sx, *sy_;  // pointers at the node, for the -x / -y neighbours
    int dxm, dym;                              // offsets of those neighbours (0 at the domain edge)
    __device__ __forceinline__ double pick(const double *a, double v0, double vz, int di, int dj, int dk) const {
        return dk ? vz : di ? a[dxm] : dj ? a[dym] : v0;
    }
    __device__ __forceinline__ double IMX(int di, int dj, int dk) const { return pick(imx, imx0, imxz, di, dj, dk); }
    __device__ __forceinline__ double IMY(int di, int dj, int dk) const { return pick(imy, imy0, imyz, di, dj, dk); }
    __device__ __forceinline__ double IMZ(int di, int dj, int dk) const { return pick(imz, imz0, imzz, di, dj, dk); }
    __device__ __forceinline__ double SX(int di, int, int) const { return di ? sx[dxm] : sx0; }
    __device__ __forceinline__ double SY(int, int dj, int) const { return dj ? sy_[dym] : sy0; }
    __device__ __forceinline__ double SZ(int, int, int dk) const { return dk ? szz : sz0; }
};

// Coefficients of plane k from the class table: the node's class and those of
// its -x and -y neighbours select rows of the (small, cache-resident) table,
// so the six FP64 arrays are not streamed from HBM.  Values are the same
// doubles as the arrays hold.
struct ZmClassView {
    double imx0, imy0, imz0, sx0, sy0, sz0;    // at (0,0,0)
    double imxz, imyz, imzz, szz;              // at (0,0,-1)
    double imxx, imyx, imzx, sxx;              // at (-1,0,0)
    double imxy, imyy, imzy, syy;              // at (0,-1,0)
    __device__ __forceinline__ double pick(double v0, double vx, double vy, double vz, int di, int dj, int dk) const {
        return dk ? vz : di ? vx : dj ? vy : v0;
    }
    __device__ __forceinline__ double IMX(int di, int dj, int dk) const { return pick(imx0, imxx, imxy, imxz, di, dj, dk); }
    __device__ __forceinline__ double IMY(int di, int dj, int dk) const { return pick(imy0, imyx, imyy, imyz, di, dj, dk); }
    __device__ __forceinline__ double IMZ(int di, int dj, int dk) const { return pick(imz0, imzx, imzy, imzz, di, dj, dk); }
    __device__ __forceinline__ double SX(int di, int, int) const { return di ? sxx : sx0; }
    __device__ __forceinline__ double SY(int, int dj, int) const { return dj ? syy : sy0; }
    __device__ __forceinline__ double SZ(int, int, int dk) const { return dk ? szz : sz0; }
};

// Planes [kb, ke) of the workgroup column (tx, ty) of a design's tiling.
struct ColSpan {
    unsigned tx, ty;
    int kb, ke;
};
// Workgroup column `tid` of a single-design launch: x fastest, then y, then
// the z-chunk.
__device__ __forceinline__ ColSpan zspan(const Grid &g, const ZTiling &t, unsigned tid) {
    const unsigned tx = tid % (unsigned)t.ntx, r1 = tid / (unsigned)t.ntx;
    const unsigned ty = r1 % (unsigned)t.nty, tz = r1 / (unsigned)t.nty;
    const int kb = (int)tz * t.kz;
    return ColSpan{tx, ty, kb, min(kb + t.kz, g.nz)};
}

// One workgroup column span of one design.  Every thread of the workgroup
// must call this (it synchronises the workgroup); `lds` holds 4*3*ZPL_MAX
// doubles.  The rows of a plane do not depend on where the span starts (the
// carried plane-(kb-1) values are recomputed by the same expressions), so any
// split of a column into spans gives the same bits.  STG selects the staging
// form (see `stage`); the results do not depend on it.
//
// Output: with TOUT = float the output is a two-segment Krylov vector (the A
// rows in FP32 from byte 0, the V row in FP64 from byte offset `offv`).
// OP_VRHS reads the V field of b at element offset `bv` (3 n1 in a flat
// vector, 0 when b points at a two-segment vector's V segment).
template <typename TIN, typename TOUT, int MODE, int TX, bool CLS, bool STG = false, typename TB = double>
__device__ __forceinline__ void op_column(const Grid &gin, const ColSpan cs, const CoeffsD &c, const OpK &k,
                                          const TIN *__restrict__ x, const TB *__restrict__ b,
                                          TOUT *__restrict__ out, double *lds, int64_t offv = 0,
                                          int64_t bv = -1) {
    // The node counts, held in scalar registers for the whole z-march: where
    // the grid comes from memory (the batched kernel) the compiler would
    // otherwise reload them in every plane, and each reload's wait also
    // waits for the outstanding LDS reads.
    Grid g = gin;
    g.nx = __builtin_amdgcn_readfirstlane(gin.nx);
    g.ny = __builtin_amdgcn_readfirstlane(gin.ny);
    g.nz = __builtin_amdgcn_readfirstlane(gin.nz);
    g.n1 = (int64_t)g.nx * g.ny * g.nz;
    using S = ZShape<TX>;
    const int i0 = (int)cs.tx * TX, j0 = (int)cs.ty * S::TY;
    const int kb = cs.kb, ke = cs.ke;
    const int li = (int)(threadIdx.x % TX), lj = (int)(threadIdx.x / TX);
    const int i = i0 + li, j = j0 + lj;
    const int ic = min(i, g.nx - 1), jc = min(j, g.ny - 1);
    const bool live = i < g.nx && j < g.ny;
    const int ctr = (lj + 1) * S::SY + (li + 1);
    const int64_t n1 = g.n1;
    const int64_t plane = (int64_t)g.nx * g.ny;

    // Stage plane kp into LDS slot `slot`.  STG: all eight loads of a thread
    // are issued before its first LDS write (one exposed memory latency per
    // plane, at the cost of the registers that hold them); otherwise the two
    // rounds of four loads run one after the other.
    static_assert(S::PL > 256 && S::PL <= 512, "two staging rounds per plane");
    auto stage = [&](int kp, int slot) {
        kp = max(0, min(kp, g.nz - 1));
        if constexpr (STG) {
            TIN v[2][4];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int e = (int)threadIdx.x + 256 * u;
                if (u == 0 || e < S::PL) {
                    const int ei = e % S::SY, ej = e / S::SY;
                    const int gi = max(0, min(i0 - 1 + ei, g.nx - 1)), gj = max(0, min(j0 - 1 + ej, g.ny - 1));
                    const int64_t q = gi + (int64_t)g.nx * gj + plane * kp;
#pragma unroll
                    for (int f = 0; f < 4; ++f) v[u][f] = x[f * n1 + q];
                }
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int e = (int)threadIdx.x + 256 * u;
                if (u == 0 || e < S::PL) {
#pragma unroll
                    for (int f = 0; f < 4; ++f) lds[(f * 3 + slot) * S::PL + e] = (double)v[u][f];
                }
            }
        } else {
            for (int e = (int)threadIdx.x; e < S::PL; e += 256) {
                const int ei = e % S::SY, ej = e / S::SY;
                const int gi = max(0, min(i0 - 1 + ei, g.nx - 1)), gj = max(0, min(j0 - 1 + ej, g.ny - 1));
                const int64_t q = gi + (int64_t)g.nx * gj + plane * kp;
#pragma unroll
                for (int f = 0; f < 4; ++f) lds[(f * 3 + slot) * S::PL + e] = (double)x[f * n1 + q];
            }
        }
    };

    stage(kb - 1, 0);
    stage(kb, 1);
    const int64_t pc = ic + (int64_t)g.nx * jc;        // in-plane index of the node
    double imxz, imyz, imzz, szz;
    {
        const int64_t q = pc + plane * max(kb - 1, 0);
        if (CLS) {
            const double *t0 = c.ctab + 6 * (int)c.cls[q];
            imxz = t0[0]; imyz = t0[1]; imzz = t0[2]; szz = t0[5];
        } else {
            imxz = c.im[0][q]; imyz = c.im[1][q]; imzz = c.im[2][q]; szz = c.sg[2][q];
        }
    }
    const int dxm = ic > 0 ? -1 : 0, dym = jc > 0 ? -g.nx : 0;
    // Class indices of the node and its -x / -y neighbours, one plane ahead.
    unsigned c0 = 0, cx = 0, cy = 0;
    if constexpr (CLS) {
        const int64_t p = pc + plane * kb;
        c0 = c.cls[p]; cx = c.cls[p + dxm]; cy = c.cls[p + dym];
    }
    FaceCarry fc{0.0, 0.0, 0.0, 0.0, 0.0};
    for (int kk = kb; kk < ke; ++kk) {
        const int r = kk - kb;
        const int sm = r % 3, s0 = (r + 1) % 3, sp = (r + 2) % 3;
        const int64_t p = pc + plane * kk;
        // The coefficients of plane kk (and the class indices of plane kk+1)
        // are requested before the plane-(kk+1) staging loads, so their
        // latencies overlap.
        // Residual with the class table: b of the node's four rows is
        // requested with the coefficients, so its latency overlaps the
        // staging (19% faster at 128^3).  With streamed coefficients the
        // registers that takes would cost a wave per SIMD, so b is read when
        // the rows are written.
        double bq[4] = {0.0, 0.0, 0.0, 0.0};
        if constexpr (MODE == OP_RESID && CLS) {
            if (live) {
#pragma unroll
                for (int f = 0; f < 4; ++f) bq[f] = (double)b[f * n1 + p];
            }
        }
        typename std::conditional<CLS, ZmClassView, ZmCView>::type cv;
        if constexpr (CLS) {
            const double *t0 = c.ctab + 6 * c0;
            const double *tx_ = c.ctab + 6 * cx;
            const double *ty_ = c.ctab + 6 * cy;
            cv = ZmClassView{t0[0], t0[1], t0[2], t0[3], t0[4], t0[5], imxz, imyz, imzz, szz,
                             tx_[0], tx_[1], tx_[2], tx_[3], ty_[0], ty_[1], ty_[2], ty_[4]};
            if (kk + 1 < ke) {
                const int64_t pn = p + plane;
                c0 = c.cls[pn]; cx = c.cls[pn + dxm]; cy = c.cls[pn + dym];
            }
        } else {
            cv = ZmCView{c.im[0][p], c.im[1][p], c.im[2][p], c.sg[0][p], c.sg[1][p], c.sg[2][p],
                         imxz, imyz, imzz, szz,
                         c.im[0] + p, c.im[1] + p, c.im[2] + p, c.sg[0] + p, c.sg[1] + p, dxm, dym};
        }
        stage(kk + 1, sp);
        __syncthreads();
        const LdsView<TX> xv{lds, {sm * S::PL, s0 * S::PL, sp * S::PL}, ctr};
        if (MODE == OP_VRHS) {
            // V-block right-hand side of the preconditioner, in FP64:
            // D_V v_V - M_VA z_A (n1 entries, 0 on inactive V rows)
            if (live) {
                double rv = 0.0;
                if (free_v(g, i, j, kk)) {
                    double d;
                    const double va = row_v(xv, cv, k, B_VA, d);
                    rv = d * (double)b[(bv < 0 ? 3 * n1 : bv) + p] - va;
                }
                out[p] = (TOUT)rv;
            }
        } else {
            double rr4[4], dd4[4];
            rows_flux(xv, cv, k, rr4, dd4, fc, r == 0);
            const double rax = rr4[0], ray = rr4[1], raz = rr4[2], rv = rr4[3];
            const double dax = dd4[0], day = dd4[1], daz = dd4[2], dv = dd4[3];
            if (live) {
                auto emit = [&](int f, double mx, double d) {
                    const int64_t q = f * n1 + p;
                    double rr;
                    if (free_f(g, f, i, j, kk)) {
                        rr = ((MODE == OP_RESID) ? (CLS ? bq[f] : (double)b[q]) - mx : mx) * recip(d);
                    } else if (exists_f(g, f, i, j, kk)) {
                        const double xq = xv.get(f, 0, 0, 0);
                        rr = (MODE == OP_RESID) ? (CLS ? bq[f] : (double)b[q]) - xq : xq;
                    } else {
                        rr = 0.0;
                    }
                    if (std::is_same<TOUT, float>::value && f == 3)
                        ((double *)((char *)out + offv))[p] = rr;
                    else
                        out[q] = (TOUT)rr;
                };
                emit(0, rax, dax);
                emit(1, ray, day);
                emit(2, raz, daz);
                emit(3, rv, dv);
            }
        }
        imxz = cv.imx0; imyz = cv.imy0; imzz = cv.imz0; szz = cv.sz0;
        __syncthreads();
    }
}

// The row constants as the kernel sees them: the launch argument, or (KP) a
// device-resident copy, so that a captured launch stays valid when the frame
// velocity changes between steps.  The copy is read through a read-only,
// non-aliased kernel-argument pointer: uniform loads the compiler issues as
// scalar loads.
template <bool KP>
__device__ __forceinline__ OpK op_consts(const OpK &k, const OpK *__restrict__ kp) {
    if constexpr (KP) return *kp;
    else return k;
}

template <typename TIN, typename TOUT, int MODE, int TX, bool CLS, bool KP, typename TB>
__global__ void __launch_bounds__(256) k_operator(Grid g, ZTiling t, CoeffsD c, OpK k0, const OpK *__restrict__ kp,
                                                  const TIN *__restrict__ x, const TB *__restrict__ b,
                                                  TOUT *__restrict__ out, int64_t offv, int64_t bv) {
    __shared__ double lds[4 * 3 * ZShape<TX>::PL];
    const OpK k = op_consts<KP>(k0, kp);
    // The unrolled staging is used where the kernel still fits in 128 VGPRs
    // (four waves per SIMD) with it: the class-table apply on 32-wide
    // columns.  Elsewhere it costs a wave per SIMD and runs slower.
    constexpr bool STG = MODE == OP_APPLY && TX == 32 && CLS;
    op_column<TIN, TOUT, MODE, TX, CLS, STG, TB>(g, zspan(g, t, xcd_tile(blockIdx.x, gridDim.x)), c, k, x, b, out,
                                                 lds, offv, bv);
}

// ---------------------------------------------------------------------------
// Batched operator: y_d = D_d^-1 M_d x_d for many independent designs in one
// persistent launch.  The host splits every design's workgroup columns into
// z-spans (segments) and deals them to the workgroups so that each runs about
// the same number of planes, and so that the workgroups of one XCD run
// neighbouring columns of a design at the same time and share its L2 (the
// x/y halo of a column is its neighbours' data).  Each workgroup runs its own
// list of segments; every loop bound is workgroup-uniform, so the grid
// always drains.
// ---------------------------------------------------------------------------
// TX = 0: the column width and coefficient form are read per design (any mix
// in one launch).  TX = 32 / 64: every design of the launch uses that width
// and, per CLS, the class table or the streamed arrays; only that variant is
// compiled in, so the kernel needs the registers and LDS of one variant.
// Every variant is held to 128 VGPRs (four waves per SIMD); the mixed one
// would take 133 and run three.
// The design's pointers, read through the constant address space: a generic
// pointer loaded from constant memory is taken to address global memory, so
// every access through it is a global (not flat) instruction.  Flat
// instructions also count against the LDS wait counter, so every LDS wait of
// the z-march would wait for the HBM loads as well.
typedef const __attribute__((address_space(4))) BatchItem ConstItem;
__device__ __forceinline__ CoeffsD item_coeffs(ConstItem &it) {
    CoeffsD c;
    for (int q = 0; q < 3; ++q) { c.im[q] = it.c.im[q]; c.sg[q] = it.c.sg[q]; }
    c.cls = it.c.cls;
    c.ctab = it.c.ctab;
    return c;
}

// One segment of a batched launch.  STG as for op_column.
template <int TX, bool CLS, bool STG>
__device__ __forceinline__ void batch_segment(const BatchItem *__restrict__ items, const BatchSeg &sg, double *lds) {
    const BatchItem &it = items[sg.d];
    ConstItem &ci = ((ConstItem *)items)[sg.d];
    const ColSpan cs{sg.col % (unsigned)it.t.ntx, sg.col / (unsigned)it.t.ntx, sg.kb, sg.ke};
    const CoeffsD c = item_coeffs(ci);
    const double *x = ci.x;
    double *y = ci.y;
    if constexpr (TX != 0) {
        op_column<double, double, OP_APPLY, TX, CLS, STG>(it.g, cs, c, it.k, x, (const double *)nullptr, y, lds);
    } else if (c.cls) {
        if (it.t.tx == 32)
            op_column<double, double, OP_APPLY, 32, true>(it.g, cs, c, it.k, x, (const double *)nullptr, y, lds);
        else
            op_column<double, double, OP_APPLY, 64, true>(it.g, cs, c, it.k, x, (const double *)nullptr, y, lds);
    } else {
        if (it.t.tx == 32)
            op_column<double, double, OP_APPLY, 32, false>(it.g, cs, c, it.k, x, (const double *)nullptr, y, lds);
        else
            op_column<double, double, OP_APPLY, 64, false>(it.g, cs, c, it.k, x, (const double *)nullptr, y, lds);
    }
}

// Persistent form: slots workgroups, each running its list of segments.
template <int TX, bool CLS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) k_operator_batch(const BatchItem *__restrict__ items,
                                                        const BatchSeg *__restrict__ segs,
                                                        const unsigned *__restrict__ slot_off) {
    __shared__ double lds[4 * 3 * (TX ? ZShape<TX ? TX : 64>::PL : ZPL_MAX)];
    const unsigned slot = (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    const unsigned e1 = slot_off[slot + 1];
    for (unsigned e = slot_off[slot]; e < e1; ++e) batch_segment<TX, CLS, false>(items, segs[e], lds);
}

// One segment per workgroup (the hardware dispatcher balances the load):
// with no loop around the z-march, the 32-wide class-table variant fits the
// unrolled staging in 128 VGPRs, as k_operator does.  Segment order is
// XCD-aware as in k_operator.
template <int TX, bool CLS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) k_operator_batch1(const BatchItem *__restrict__ items,
                                                         const BatchSeg *__restrict__ segs) {
    __shared__ double lds[4 * 3 * (TX ? ZShape<TX ? TX : 64>::PL : ZPL_MAX)];
    batch_segment<TX, CLS, TX == 32 && CLS>(items, segs[xcd_tile(blockIdx.x, gridDim.x)], lds);
}

// Row diagonal D for every entry (1 on fixed rows, 0 where no unknown exists).
__global__ void k_diag(Grid g, CoeffsD c, OpK k, double *__restrict__ D) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.n1) return;
    const int i = (int)(p % g.nx), j = (int)((p / g.nx) % g.ny), kk = (int)(p / ((int64_t)g.nx * g.ny));
    const View<double> xv{nullptr, nullptr, nullptr, nullptr, k.sy, k.sz};
    const CView cv = cview_at(c, p, k);
    for (int f = 0; f < 4; ++f) {
        double d = 0.0;
        if (exists_f(g, f, i, j, kk)) {
            if (!free_f(g, f, i, j, kk)) d = 1.0;
            else if (f == FAX) (void)row_ax(xv, cv, k, 0, d);
            else if (f == FAY) (void)row_ay(xv, cv, k, 0, d);
            else if (f == FAZ) (void)row_az(xv, cv, k, 0, d);
            else (void)row_v(xv, cv, k, 0, d);
        }
        D[f * g.n1 + p] = d;
    }
}

// ---------------------------------------------------------------------------
// Right-hand side b(A^n), QMS_3D.f90:1459-1475, :1624, :1772, :1906-1928
// (steady: :459-470 etc. with zero interior RHS).
// ---------------------------------------------------------------------------
template <bool KP>
__global__ void __launch_bounds__(256) k_rhs(Grid g, Tiling t, CoeffsD c, BC bc, OpK k0, const OpK *__restrict__ kp,
                                             const double *__restrict__ aprev, double *__restrict__ b) {
    const OpK k = op_consts<KP>(k0, kp);
    int i, j, kk;
    if (!tile_node(g, t, i, j, kk)) return;
    const int64_t p = gidx(g, i, j, kk);
    const int64_t n1 = g.n1;
    for (int f = 0; f < 4; ++f) {
        double r = 0.0;
        if (exists_f(g, f, i, j, kk)) {
            if (!free_f(g, f, i, j, kk)) {
                r = bc_value(g, bc, k.v, f, i, j, kk);
            } else if (k.tr) {
                // sigma streamed from the arrays: here the class-table gathers
                // (one node per thread, so each gather's latency is exposed)
                // measured slower than the coalesced array reads, 31 against
                // 27 us at 128^3
                if (f < 3) {
                    r = (c.sg[f][p] * k.idt) * aprev[f * n1 + p];
                } else {
                    const double *Ax = aprev + p, *Ay = aprev + n1 + p, *Az = aprev + 2 * n1 + p;
                    const double *sx = c.sg[0] + p, *sy = c.sg[1] + p, *sz = c.sg[2] + p;
                    r = (sx[0] * Ax[0] - sx[-1] * Ax[-1]) * k.iht[0]
                      + (sy[0] * Ay[0] - sy[-k.sy] * Ay[-k.sy]) * k.iht[1]
                      + (sz[0] * Az[0] - sz[-k.sz] * Az[-k.sz]) * k.iht[2];
                }
            }
        }
        b[f * n1 + p] = r;
    }
}

// x <- Dirichlet values on fixed entries, 0 where no unknown exists; free
// entries keep their value (warm start).
template <bool KP>
__global__ void __launch_bounds__(256) k_setbc(Grid g, BC bc, OpK k0, const OpK *__restrict__ kp,
                                               double *__restrict__ x) {
    const OpK k = op_consts<KP>(k0, kp);
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.n1) return;
    const int i = (int)(p % g.nx), j = (int)((p / g.nx) % g.ny), kk = (int)(p / ((int64_t)g.nx * g.ny));
    for (int f = 0; f < 4; ++f) {
        const int64_t q = f * g.n1 + p;
        if (!exists_f(g, f, i, j, kk)) x[q] = 0.0;
        else if (!free_f(g, f, i, j, kk)) x[q] = bc_value(g, bc, k.v, f, i, j, kk);
    }
}

// ---------------------------------------------------------------------------
// Cell-centred H (or B when unit_mu), QMS_3D.f90:1346-1376.
// Output: 3 arrays of cx*cy*cz, x fastest.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_hfield(Grid g, CoeffsD c, const double *__restrict__ x,
                                                double *__restrict__ H, int unit_mu, double ihx, double ihy,
                                                double ihz) {
    const int cx = g.nx - 1, cy = g.ny - 1, cz = g.nz - 1;
    const int64_t nc = (int64_t)cx * cy * cz;
    const int64_t q = (int64_t)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (q >= nc) return;
    const int i = (int)(q % cx), j = (int)((q / cx) % cy), k = (int)(q / ((int64_t)cx * cy));
    const int64_t n1 = g.n1, p = gidx(g, i, j, k);
    const int X = 1, Y = g.nx, Z = g.nx * g.ny;
    const double *Ax = x + p, *Ay = x + n1 + p, *Az = x + 2 * n1 + p;
    // Face curls (B on the x-, y- and z-faces) at the two faces of the cell
    // along each axis, multiplied by the stored 1/Mu (the reference divides
    // by Mu; 1/(1/Mu) is not always Mu bit for bit).
    const double bx0 = (Az[Y] - Az[0]) * ihy - (Ay[Z] - Ay[0]) * ihz;
    const double bx1 = (Az[X + Y] - Az[X]) * ihy - (Ay[X + Z] - Ay[X]) * ihz;
    const double by0 = (Ax[Z] - Ax[0]) * ihz - (Az[X] - Az[0]) * ihx;
    const double by1 = (Ax[Y + Z] - Ax[Y]) * ihz - (Az[X + Y] - Az[Y]) * ihx;
    const double bz0 = (Ay[X] - Ay[0]) * ihx - (Ax[Y] - Ax[0]) * ihy;
    const double bz1 = (Ay[X + Z] - Ay[Z]) * ihx - (Ax[Y + Z] - Ax[Z]) * ihy;
    if (unit_mu) {
        H[q] = 0.5 * (bx0 + bx1);
        H[nc + q] = 0.5 * (by0 + by1);
        H[2 * nc + q] = 0.5 * (bz0 + bz1);
    } else if (c.cls) {
        // 1/Mu from the class table (the same doubles as the arrays)
        const double *t0 = c.ctab + 6 * (int)c.cls[p];
        H[q] = 0.5 * (bx0 * t0[0] + bx1 * c.ctab[6 * (int)c.cls[p + X]]);
        H[nc + q] = 0.5 * (by0 * t0[1] + by1 * c.ctab[6 * (int)c.cls[p + Y] + 1]);
        H[2 * nc + q] = 0.5 * (bz0 * t0[2] + bz1 * c.ctab[6 * (int)c.cls[p + Z] + 2]);
    } else {
        const double *imx = c.im[0] + p, *imy = c.im[1] + p, *imz = c.im[2] + p;
        H[q] = 0.5 * (bx0 * imx[0] + bx1 * imx[X]);
        H[nc + q] = 0.5 * (by0 * imy[0] + by1 * imy[Y]);
        H[2 * nc + q] = 0.5 * (bz0 * imz[0] + bz1 * imz[Z]);
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static inline dim3 grid1(int64_t n, int bs = 256) { return dim3((unsigned)((n + bs - 1) / bs)); }

OpK make_opk(const Grid &g, const OpParams &op) {
    OpK k;
    const double h[3] = {g.hx, g.hy, g.hz};
    k.sy = g.nx;
    k.sz = g.nx * g.ny;
    k.tr = op.tr;
    k.adv = op.adv;
    k.idt = op.tr ? 1.0 / op.dt : 0.0;
    for (int a = 0; a < 3; ++a) {
        k.v[a] = op.v[a];
        k.hv[a] = 0.5 * op.v[a];
        k.ih2[a] = 1.0 / (h[a] * h[a]);
        k.ih[a] = 1.0 / h[a];
        k.iht[a] = op.tr ? 1.0 / (h[a] * op.dt) : 0.0;
        for (int b = 0; b < 3; ++b) k.v1[a][b] = op.v[a] / (2.0 * h[b]);
    }
    k.ixy = 1.0 / (h[0] * h[1]);
    k.ixz = 1.0 / (h[0] * h[2]);
    k.iyz = 1.0 / (h[1] * h[2]);
    // (a, b, c) of each v_a / (2 h_b h_c) term of the V row, in the order of OpK::w
    static const int W[12][3] = {{2, 2, 0}, {0, 2, 2}, {1, 1, 0}, {0, 1, 1}, {2, 2, 1}, {1, 2, 2},
                                 {1, 0, 0}, {0, 0, 1}, {2, 1, 1}, {1, 1, 2}, {2, 0, 0}, {0, 0, 2}};
    for (int q = 0; q < 12; ++q) k.w[q] = op.v[W[q][0]] / (2.0 * h[W[q][1]] * h[W[q][2]]);
    return k;
}

// The RHS kernel uses 128 x 2 tiles of one plane.
static Tiling rhs_tiling(const Grid &g, dim3 &grid) {
    Tiling t;
    t.txs = 7;
    t.ntx = (g.nx + 127) / 128;
    t.nty = (g.ny + 1) / 2;
    grid = dim3((unsigned)((int64_t)t.ntx * t.nty * g.nz));
    return t;
}

ZTiling ztiling(const Grid &g, int kz, int tx, unsigned &ncol) {
    ZTiling t;
    t.kz = kz >= 1 ? kz : 4;
    t.tx = tx == 32 ? 32 : 64;
    t.ntx = (g.nx + t.tx - 1) / t.tx;
    t.nty = (g.ny + 256 / t.tx - 1) / (256 / t.tx);
    ncol = (unsigned)((int64_t)t.ntx * t.nty * ((g.nz + t.kz - 1) / t.kz));
    return t;
}

template <typename TIN, typename TOUT, int MODE, typename TB = double>
static void launch_operator(const Grid &g, const CoeffsD &c, const OpParams &op, const TIN *x, const TB *b,
                            TOUT *out, hipStream_t s, int64_t offv = 0, int64_t bv = -1) {
    unsigned ncol;
    const ZTiling t = ztiling(g, op.zchunk, op.colw, ncol);
    const OpK k = make_opk(g, op);
    const OpK *kp = op.kdev;
    const dim3 grid(ncol), block(256);
    auto go = [&](auto tx, auto cls, auto kpc) {
        constexpr int TX = decltype(tx)::value;
        constexpr bool CLS = decltype(cls)::value, KP = decltype(kpc)::value;
        hipLaunchKernelGGL((k_operator<TIN, TOUT, MODE, TX, CLS, KP, TB>), grid, block, 0, s, g, t, c, k, kp, x, b,
                           out, offv, bv);
    };
    using T32 = std::integral_constant<int, 32>;
    using T64 = std::integral_constant<int, 64>;
    using Y = std::true_type;
    using N = std::false_type;
    if (kp) {
        if (c.cls) { if (t.tx == 32) go(T32(), Y(), Y()); else go(T64(), Y(), Y()); }
        else { if (t.tx == 32) go(T32(), N(), Y()); else go(T64(), N(), Y()); }
    } else {
        if (c.cls) { if (t.tx == 32) go(T32(), Y(), N()); else go(T64(), Y(), N()); }
        else { if (t.tx == 32) go(T32(), N(), N()); else go(T64(), N(), N()); }
    }
}

// slots: the grid size.  With slot offsets, a multiple of 8: one persistent
// workgroup per slot of the plan.  Without (null), one workgroup per segment.
// tx, cls: the column width and coefficient form every design of the launch
// shares, or tx = 0 when they differ between designs.
void launch_operator_batch(const BatchItem *items_dev, const BatchSeg *segs_dev, const unsigned *slot_off_dev,
                           unsigned slots, int tx, int cls, hipStream_t s) {
    const dim3 grid(slots), block(256);
    if (!slot_off_dev) {
        if (tx == 32 && cls) hipLaunchKernelGGL((k_operator_batch1<32, true>), grid, block, 0, s, items_dev, segs_dev);
        else if (tx == 32) hipLaunchKernelGGL((k_operator_batch1<32, false>), grid, block, 0, s, items_dev, segs_dev);
        else if (tx == 64 && cls) hipLaunchKernelGGL((k_operator_batch1<64, true>), grid, block, 0, s, items_dev, segs_dev);
        else if (tx == 64) hipLaunchKernelGGL((k_operator_batch1<64, false>), grid, block, 0, s, items_dev, segs_dev);
        else hipLaunchKernelGGL((k_operator_batch1<0, false>), grid, block, 0, s, items_dev, segs_dev);
        return;
    }
    if (tx == 32 && cls) hipLaunchKernelGGL((k_operator_batch<32, true>), grid, block, 0, s, items_dev, segs_dev, slot_off_dev);
    else if (tx == 32) hipLaunchKernelGGL((k_operator_batch<32, false>), grid, block, 0, s, items_dev, segs_dev, slot_off_dev);
    else if (tx == 64 && cls) hipLaunchKernelGGL((k_operator_batch<64, true>), grid, block, 0, s, items_dev, segs_dev, slot_off_dev);
    else if (tx == 64) hipLaunchKernelGGL((k_operator_batch<64, false>), grid, block, 0, s, items_dev, segs_dev, slot_off_dev);
    else hipLaunchKernelGGL((k_operator_batch<0, false>), grid, block, 0, s, items_dev, segs_dev, slot_off_dev);
}

// Workgroups of the batched kernel variant (tx, cls as for the launch) that
// the device holds at once: its occupancy per CU times the CU count, rounded
// down to a multiple of 8.
int batch_resident_blocks(int tx, int cls) {
    int dev = 0, ncu = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 1024;
    const void *f = tx == 32 && cls ? (const void *)&k_operator_batch<32, true>
                  : tx == 32        ? (const void *)&k_operator_batch<32, false>
                  : tx == 64 && cls ? (const void *)&k_operator_batch<64, true>
                  : tx == 64        ? (const void *)&k_operator_batch<64, false>
                                    : (const void *)&k_operator_batch<0, false>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, f, 256, 0) != hipSuccess || per < 1) return 1024;
    const int r = per * ncu / 8 * 8;
    return r >= 8 ? r : 8;
}

void launch_coeffs(const Grid &g, const uint16_t *slot, const double *mat, const Coeffs &c, hipStream_t s) {
    hipLaunchKernelGGL(k_coeffs, grid1(g.n1), dim3(256), 0, s, g, slot, mat, c);
}
void launch_diag(const Grid &g, const CoeffsD &c, const OpParams &op, double *D, hipStream_t s) {
    hipLaunchKernelGGL(k_diag, grid1(g.n1), dim3(256), 0, s, g, c, make_opk(g, op), D);
}
void launch_rhs(const Grid &g, const CoeffsD &c, const BC &bc, const OpParams &op, const double *aprev, double *b,
                hipStream_t s) {
    dim3 grid;
    const Tiling t = rhs_tiling(g, grid);
    if (op.kdev) hipLaunchKernelGGL(k_rhs<true>, grid, dim3(256), 0, s, g, t, c, bc, make_opk(g, op), op.kdev, aprev, b);
    else hipLaunchKernelGGL(k_rhs<false>, grid, dim3(256), 0, s, g, t, c, bc, make_opk(g, op), op.kdev, aprev, b);
}
void launch_setbc(const Grid &g, const BC &bc, const OpParams &op, double *x, hipStream_t s) {
    if (op.kdev) hipLaunchKernelGGL(k_setbc<true>, grid1(g.n1), dim3(256), 0, s, g, bc, make_opk(g, op), op.kdev, x);
    else hipLaunchKernelGGL(k_setbc<false>, grid1(g.n1), dim3(256), 0, s, g, bc, make_opk(g, op), op.kdev, x);
}
void launch_resid(const Grid &g, const CoeffsD &c, const OpParams &op, const double *x, const double *b, double *r,
                  hipStream_t s) {
    launch_operator<double, double, OP_RESID>(g, c, op, x, b, r, s);
}
void launch_apply(const Grid &g, const CoeffsD &c, const OpParams &op, const double *x, double *y, hipStream_t s) {
    launch_operator<double, double, OP_APPLY>(g, c, op, x, (const double *)nullptr, y, s);
}
void launch_apply_f32in(const Grid &g, const CoeffsD &c, const OpParams &op, const float *x, double *y, hipStream_t s) {
    launch_operator<float, double, OP_APPLY>(g, c, op, x, (const double *)nullptr, y, s);
}
void launch_vrhs(const Grid &g, const CoeffsD &c, const OpParams &op, const float *z, const double *v, double *out,
                 hipStream_t s) {
    launch_operator<float, double, OP_VRHS>(g, c, op, z, v, out, s);
}
void launch_vrhs(const Grid &g, const CoeffsD &c, const OpParams &op, const double *z, const double *v, double *out,
                 hipStream_t s) {
    launch_operator<double, double, OP_VRHS>(g, c, op, z, v, out, s);
}
void launch_vrhs_seg(const Grid &g, const CoeffsD &c, const OpParams &op, const float *z, const double *vv,
                     double *out, hipStream_t s) {
    launch_operator<float, double, OP_VRHS>(g, c, op, z, vv, out, s, 0, 0);
}
void launch_apply_f32(const Grid &g, const CoeffsD &c, const OpParams &op, const float *x, float *y, int64_t offv,
                      hipStream_t s) {
    launch_operator<float, float, OP_APPLY>(g, c, op, x, (const double *)nullptr, y, s, offv);
}
void launch_hfield(const Grid &g, const CoeffsD &c, const double *x, double *H, int unit_mu, hipStream_t s) {
    const int64_t nc = (int64_t)(g.nx - 1) * (g.ny - 1) * (g.nz - 1);
    hipLaunchKernelGGL(k_hfield, grid1(nc), dim3(256), 0, s, g, c, x, H, unit_mu, 1.0 / g.hx, 1.0 / g.hy, 1.0 / g.hz);
}

}  // namespace qms
