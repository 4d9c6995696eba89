// Fused matrix-free operator of the coupled A-V system on the Yee grid.
//
// One thread owns one node p = (i,j,k) and evaluates the four rows (Ax, Ay,
// Az, V) that live there, exactly the rows the reference assembles into CSR
// every step (steady QMS_3D.f90:448-1293, transient :1457-2189).  Arithmetic
// is FP64 whatever the storage type of the input vector: the V rows mix
// conductivities 1e-3..6e7 S/m, and only FP64 evaluation keeps the weak
// background couplings that pin each conductor's floating potential.
// Self couplings are evaluated in difference form c*(x_p - x_q) so that a
// constant on a conductor produces its true (tiny) residual.
#include <hip/hip_runtime.h>
#include "qms_kernels.h"

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
// diagonal (the scaling D used by the Krylov solver).  `tr` enables the
// transient sigma/dt terms, `adv` the frame-velocity terms.
// ---------------------------------------------------------------------------
template <typename T>
struct View {
    const T *ax, *ay, *az, *v;
    const Grid *g;
    __device__ __forceinline__ double AX(int i, int j, int k) const { return (double)ax[gidx(*g, i, j, k)]; }
    __device__ __forceinline__ double AY(int i, int j, int k) const { return (double)ay[gidx(*g, i, j, k)]; }
    __device__ __forceinline__ double AZ(int i, int j, int k) const { return (double)az[gidx(*g, i, j, k)]; }
    __device__ __forceinline__ double VV(int i, int j, int k) const { return (double)v[gidx(*g, i, j, k)]; }
};

struct CView {
    const double *imx, *imy, *imz, *sx, *sy, *sz;
    const Grid *g;
    __device__ __forceinline__ double IMX(int i, int j, int k) const { return imx[gidx(*g, i, j, k)]; }
    __device__ __forceinline__ double IMY(int i, int j, int k) const { return imy[gidx(*g, i, j, k)]; }
    __device__ __forceinline__ double IMZ(int i, int j, int k) const { return imz[gidx(*g, i, j, k)]; }
    __device__ __forceinline__ double SX(int i, int j, int k) const { return sx[gidx(*g, i, j, k)]; }
    __device__ __forceinline__ double SY(int i, int j, int k) const { return sy[gidx(*g, i, j, k)]; }
    __device__ __forceinline__ double SZ(int i, int j, int k) const { return sz[gidx(*g, i, j, k)]; }
};

// Which blocks of the row to include.
enum : int { B_AA = 1, B_AV = 2, B_VA = 4, B_VV = 8, B_ALL = 15 };

// Ax row, QMS_3D.f90:476-607 (+ :1504, :1475)
template <typename T>
__device__ __forceinline__ double row_ax(const View<T> &x, const CView &c, int i, int j, int k, const OpParams &op,
                                         int blocks, double &diag) {
    const Grid &g = *c.g;
    const double hx = g.hx, hy = g.hy, hz = g.hz;
    const double s = c.SX(i, j, k);
    const double imy_km = c.IMY(i, j, k - 1), imy_k = c.IMY(i, j, k);
    const double imz_jm = c.IMZ(i, j - 1, k), imz_j = c.IMZ(i, j, k);
    const double imx_im = c.IMX(i - 1, j, k), imx_i = c.IMX(i, j, k);
    double d = (imz_j + imz_jm) / (hy * hy) + (imy_k + imy_km) / (hz * hz) + (imx_i + imx_im) / (hx * hx);
    if (op.tr) d = d + s / op.dt;
    diag = d;
    double y = 0.0;
    if (blocks & B_AA) {
        const double x0 = x.AX(i, j, k);
        y += imy_km / (hz * hz) * (x0 - x.AX(i, j, k - 1)) + imz_jm / (hy * hy) * (x0 - x.AX(i, j - 1, k))
           + imx_im / (hx * hx) * (x0 - x.AX(i - 1, j, k)) + imx_i / (hx * hx) * (x0 - x.AX(i + 1, j, k))
           + imz_j / (hy * hy) * (x0 - x.AX(i, j + 1, k)) + imy_k / (hz * hz) * (x0 - x.AX(i, j, k + 1));
        if (op.tr) y += (s / op.dt) * x0;
        const double ay_m0 = x.AY(i, j - 1, k), ay_p0 = x.AY(i + 1, j - 1, k);
        const double ay_00 = x.AY(i, j, k), ay_10 = x.AY(i + 1, j, k);
        const double az_m0 = x.AZ(i, j, k - 1), az_p0 = x.AZ(i + 1, j, k - 1);
        const double az_00 = x.AZ(i, j, k), az_10 = x.AZ(i + 1, j, k);
        y += imz_jm / (hx * hy) * (ay_m0 - ay_p0) + imz_j / (hx * hy) * (ay_10 - ay_00);
        y += imy_km / (hx * hz) * (az_m0 - az_p0) + imy_k / (hx * hz) * (az_10 - az_00);
        if (op.adv) {
            const double vx = op.v[0], vy = op.v[1], vz = op.v[2];
            (void)vx;
            y += (s * vz) / (2.0 * hz) * (x.AX(i, j, k - 1) - x.AX(i, j, k + 1))
               + (s * vy) / (2.0 * hy) * (x.AX(i, j - 1, k) - x.AX(i, j + 1, k));
            y += (s * vy) / (2.0 * hx) * (-ay_m0 + ay_p0 - ay_00 + ay_10);
            y += (s * vz) / (2.0 * hx) * (-az_m0 + az_p0 - az_00 + az_10);
        }
    }
    if (blocks & B_AV) y += s / hx * (x.VV(i + 1, j, k) - x.VV(i, j, k));
    return y;
}

// Ay row, QMS_3D.f90:650-781 (+ :1681, :1624)
template <typename T>
__device__ __forceinline__ double row_ay(const View<T> &x, const CView &c, int i, int j, int k, const OpParams &op,
                                         int blocks, double &diag) {
    const Grid &g = *c.g;
    const double hx = g.hx, hy = g.hy, hz = g.hz;
    const double s = c.SY(i, j, k);
    const double imx_km = c.IMX(i, j, k - 1), imx_k = c.IMX(i, j, k);
    const double imy_jm = c.IMY(i, j - 1, k), imy_j = c.IMY(i, j, k);
    const double imz_im = c.IMZ(i - 1, j, k), imz_i = c.IMZ(i, j, k);
    double d = (imz_i + imz_im) / (hx * hx) + (imx_k + imx_km) / (hz * hz) + (imy_j + imy_jm) / (hy * hy);
    if (op.tr) d = d + s / op.dt;
    diag = d;
    double y = 0.0;
    if (blocks & B_AA) {
        const double x0 = x.AY(i, j, k);
        y += imx_km / (hz * hz) * (x0 - x.AY(i, j, k - 1)) + imy_jm / (hy * hy) * (x0 - x.AY(i, j - 1, k))
           + imz_im / (hx * hx) * (x0 - x.AY(i - 1, j, k)) + imz_i / (hx * hx) * (x0 - x.AY(i + 1, j, k))
           + imy_j / (hy * hy) * (x0 - x.AY(i, j + 1, k)) + imx_k / (hz * hz) * (x0 - x.AY(i, j, k + 1));
        if (op.tr) y += (s / op.dt) * x0;
        const double ax_m0 = x.AX(i - 1, j, k), ax_00 = x.AX(i, j, k);
        const double ax_m1 = x.AX(i - 1, j + 1, k), ax_01 = x.AX(i, j + 1, k);
        const double az_m0 = x.AZ(i, j, k - 1), az_m1 = x.AZ(i, j + 1, k - 1);
        const double az_00 = x.AZ(i, j, k), az_01 = x.AZ(i, j + 1, k);
        y += imz_im / (hx * hy) * (ax_m0 - ax_m1) + imz_i / (hx * hy) * (ax_01 - ax_00);
        y += imx_km / (hy * hz) * (az_m0 - az_m1) + imx_k / (hy * hz) * (az_01 - az_00);
        if (op.adv) {
            const double vx = op.v[0], vz = op.v[2];
            y += (s * vz) / (2.0 * hz) * (x.AY(i, j, k - 1) - x.AY(i, j, k + 1))
               + (s * vx) / (2.0 * hx) * (x.AY(i - 1, j, k) - x.AY(i + 1, j, k));
            y += (s * vx) / (2.0 * hy) * (-ax_m0 - ax_00 + ax_m1 + ax_01);
            y += (s * vz) / (2.0 * hy) * (-az_m0 + az_m1 - az_00 + az_01);
        }
    }
    if (blocks & B_AV) y += s / hy * (x.VV(i, j + 1, k) - x.VV(i, j, k));
    return y;
}

// Az row, QMS_3D.f90:824-955 (+ :1857, :1772)
template <typename T>
__device__ __forceinline__ double row_az(const View<T> &x, const CView &c, int i, int j, int k, const OpParams &op,
                                         int blocks, double &diag) {
    const Grid &g = *c.g;
    const double hx = g.hx, hy = g.hy, hz = g.hz;
    const double s = c.SZ(i, j, k);
    const double imz_km = c.IMZ(i, j, k - 1), imz_k = c.IMZ(i, j, k);
    const double imx_jm = c.IMX(i, j - 1, k), imx_j = c.IMX(i, j, k);
    const double imy_im = c.IMY(i - 1, j, k), imy_i = c.IMY(i, j, k);
    double d = (imy_i + imy_im) / (hx * hx) + (imx_j + imx_jm) / (hy * hy) + (imz_k + imz_km) / (hz * hz);
    if (op.tr) d = d + s / op.dt;
    diag = d;
    double y = 0.0;
    if (blocks & B_AA) {
        const double x0 = x.AZ(i, j, k);
        y += imz_km / (hz * hz) * (x0 - x.AZ(i, j, k - 1)) + imx_jm / (hy * hy) * (x0 - x.AZ(i, j - 1, k))
           + imy_im / (hx * hx) * (x0 - x.AZ(i - 1, j, k)) + imy_i / (hx * hx) * (x0 - x.AZ(i + 1, j, k))
           + imx_j / (hy * hy) * (x0 - x.AZ(i, j + 1, k)) + imz_k / (hz * hz) * (x0 - x.AZ(i, j, k + 1));
        if (op.tr) y += (s / op.dt) * x0;
        const double ax_m0 = x.AX(i - 1, j, k), ax_00 = x.AX(i, j, k);
        const double ax_m1 = x.AX(i - 1, j, k + 1), ax_01 = x.AX(i, j, k + 1);
        const double ay_m0 = x.AY(i, j - 1, k), ay_00 = x.AY(i, j, k);
        const double ay_m1 = x.AY(i, j - 1, k + 1), ay_01 = x.AY(i, j, k + 1);
        y += imy_im / (hx * hz) * (ax_m0 - ax_m1) + imy_i / (hx * hz) * (ax_01 - ax_00);
        y += imx_jm / (hy * hz) * (ay_m0 - ay_m1) + imx_j / (hy * hz) * (ay_01 - ay_00);
        if (op.adv) {
            const double vx = op.v[0], vy = op.v[1];
            y += (s * vy) / (2.0 * hy) * (x.AZ(i, j - 1, k) - x.AZ(i, j + 1, k))
               + (s * vx) / (2.0 * hx) * (x.AZ(i - 1, j, k) - x.AZ(i + 1, j, k));
            y += (s * vx) / (2.0 * hz) * (-ax_m0 - ax_00 + ax_m1 + ax_01);
            y += (s * vy) / (2.0 * hz) * (-ay_m0 - ay_00 + ay_m1 + ay_01);
        }
    }
    if (blocks & B_AV) y += s / hz * (x.VV(i, j, k + 1) - x.VV(i, j, k));
    return y;
}

// V row, QMS_3D.f90:1003-1285 (+ :1975-2126 transient terms)
template <typename T>
__device__ __forceinline__ double row_v(const View<T> &x, const CView &c, int i, int j, int k, const OpParams &op,
                                        int blocks, double &diag) {
    const Grid &g = *c.g;
    const double hx = g.hx, hy = g.hy, hz = g.hz;
    const double hx2 = hx * hx, hy2 = hy * hy, hz2 = hz * hz;
    const double x0 = c.SX(i, j, k), x1 = c.SX(i - 1, j, k);
    const double y0 = c.SY(i, j, k), y1 = c.SY(i, j - 1, k);
    const double z0 = c.SZ(i, j, k), z1 = c.SZ(i, j, k - 1);
    const double fx0 = fmax(x0, QMS_SIGFLOOR), fx1 = fmax(x1, QMS_SIGFLOOR);
    const double fy0 = fmax(y0, QMS_SIGFLOOR), fy1 = fmax(y1, QMS_SIGFLOOR);
    const double fz0 = fmax(z0, QMS_SIGFLOOR), fz1 = fmax(z1, QMS_SIGFLOOR);
    diag = (fx0 + fx1) / hx2 + (fy0 + fy1) / hy2 + (fz0 + fz1) / hz2;
    double y = 0.0;
    if (blocks & B_VV) {
        const double v0 = x.VV(i, j, k);
        y += fz1 / hz2 * (v0 - x.VV(i, j, k - 1)) + fy1 / hy2 * (v0 - x.VV(i, j - 1, k))
           + fx1 / hx2 * (v0 - x.VV(i - 1, j, k)) + fx0 / hx2 * (v0 - x.VV(i + 1, j, k))
           + fy0 / hy2 * (v0 - x.VV(i, j + 1, k)) + fz0 / hz2 * (v0 - x.VV(i, j, k + 1));
    }
    if (blocks & B_VA) {
        if (op.tr) {
            const double t = op.dt;
            y += x0 / (hx * t) * x.AX(i, j, k) - x1 / (hx * t) * x.AX(i - 1, j, k)
               + y0 / (hy * t) * x.AY(i, j, k) - y1 / (hy * t) * x.AY(i, j - 1, k)
               + z0 / (hz * t) * x.AZ(i, j, k) - z1 / (hz * t) * x.AZ(i, j, k - 1);
        }
        if (op.adv) {
            const double vx = op.v[0], vy = op.v[1], vz = op.v[2];
            double w = 0.0;
            // Ax columns
            w += (x1 * vz / (2.0 * hz * hx) - z1 * vx / (2.0 * hz2)) * x.AX(i - 1, j, k - 1);
            w += (-x0 * vz / (2.0 * hz * hx) - z1 * vx / (2.0 * hz2)) * x.AX(i, j, k - 1);
            w += (x1 * vy / (2.0 * hy * hx) - y1 * vx / (2.0 * hy2)) * x.AX(i - 1, j - 1, k);
            w += (-x0 * vy / (2.0 * hy * hx) - y1 * vx / (2.0 * hy2)) * x.AX(i, j - 1, k);
            const double axc = (y0 + y1) * vx / (2.0 * hy2) + (z0 + z1) * vx / (2.0 * hz2);
            w += axc * (x.AX(i - 1, j, k) + x.AX(i, j, k));
            w += (-x1 * vy / (2.0 * hy * hx) - y0 * vx / (2.0 * hy2)) * x.AX(i - 1, j + 1, k);
            w += (x0 * vy / (2.0 * hy * hx) - y0 * vx / (2.0 * hy2)) * x.AX(i, j + 1, k);
            w += (-x1 * vz / (2.0 * hz * hx) - z0 * vx / (2.0 * hz2)) * x.AX(i - 1, j, k + 1);
            w += (x0 * vz / (2.0 * hz * hx) - z0 * vx / (2.0 * hz2)) * x.AX(i, j, k + 1);
            // Ay columns
            w += (y1 * vz / (2.0 * hz * hy) - z1 * vy / (2.0 * hz2)) * x.AY(i, j - 1, k - 1);
            w += (-y0 * vz / (2.0 * hz * hy) - z1 * vy / (2.0 * hz2)) * x.AY(i, j, k - 1);
            w += (-x1 * vy / (2.0 * hx2) + y1 * vx / (2.0 * hx * hy)) * x.AY(i - 1, j - 1, k);
            const double ayc = (x0 + x1) * vy / (2.0 * hx2) + (z0 + z1) * vy / (2.0 * hz2);
            w += ayc * (x.AY(i, j - 1, k) + x.AY(i, j, k));
            w += (-x0 * vy / (2.0 * hx2) - y1 * vx / (2.0 * hx * hy)) * x.AY(i + 1, j - 1, k);
            w += (-x1 * vy / (2.0 * hx2) - y0 * vx / (2.0 * hx * hy)) * x.AY(i - 1, j, k);
            w += (-x0 * vy / (2.0 * hx2) + y0 * vx / (2.0 * hx * hy)) * x.AY(i + 1, j, k);
            w += (-y1 * vz / (2.0 * hz * hy) - z0 * vy / (2.0 * hz2)) * x.AY(i, j - 1, k + 1);
            w += (y0 * vz / (2.0 * hz * hy) - z0 * vy / (2.0 * hz2)) * x.AY(i, j, k + 1);
            // Az columns
            w += (-y1 * vz / (2.0 * hy2) + z1 * vy / (2.0 * hy * hz)) * x.AZ(i, j - 1, k - 1);
            w += (-x1 * vz / (2.0 * hx2) + z1 * vx / (2.0 * hx * hz)) * x.AZ(i - 1, j, k - 1);
            const double azc = (x0 + x1) * vz / (2.0 * hx2) + (y0 + y1) * vz / (2.0 * hy2);
            w += azc * (x.AZ(i, j, k - 1) + x.AZ(i, j, k));
            w += (-x0 * vz / (2.0 * hx2) - z1 * vx / (2.0 * hx * hz)) * x.AZ(i + 1, j, k - 1);
            w += (-y0 * vz / (2.0 * hy2) - z1 * vy / (2.0 * hy * hz)) * x.AZ(i, j + 1, k - 1);
            w += (-y1 * vz / (2.0 * hy2) - z0 * vy / (2.0 * hy * hz)) * x.AZ(i, j - 1, k);
            w += (-x1 * vz / (2.0 * hx2) - z0 * vx / (2.0 * hx * hz)) * x.AZ(i - 1, j, k);
            w += (-x0 * vz / (2.0 * hx2) + z0 * vx / (2.0 * hx * hz)) * x.AZ(i + 1, j, k);
            w += (-y0 * vz / (2.0 * hy2) + z0 * vy / (2.0 * hy * hz)) * x.AZ(i, j + 1, k);
            y += w;
        }
    }
    return y;
}

// ---------------------------------------------------------------------------
// Fused operator kernel.
//   OP_RESID: out = (b - M x)/D on free rows, b - x on fixed rows (x: FP64)
//   OP_APPLY: out = (M x)/D on free rows, x on fixed rows
//   OP_VA:    out_V = M_VA x_A on free V rows (A-part of the V rows only),
//             other entries untouched; used by the block-triangular
//             preconditioner.
// The entries of `out` that do not exist in the reference's unknown vector
// are written as 0.
// ---------------------------------------------------------------------------
template <typename TIN, typename TOUT, int MODE>
__global__ void __launch_bounds__(256) k_operator(Grid g, CoeffsD c, OpParams op, const TIN *__restrict__ x,
                                                  const double *__restrict__ b, TOUT *__restrict__ out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.n1) return;
    const int i = (int)(p % g.nx), j = (int)((p / g.nx) % g.ny), k = (int)(p / ((int64_t)g.nx * g.ny));
    View<TIN> xv{x, x + g.n1, x + 2 * g.n1, x + 3 * g.n1, &g};
    CView cv{c.im[0], c.im[1], c.im[2], c.sg[0], c.sg[1], c.sg[2], &g};
    if (MODE == OP_VA) {
        if (free_v(g, i, j, k)) {
            double d;
            out[3 * g.n1 + p] = (TOUT)row_v(xv, cv, i, j, k, op, B_VA, d);
        }
        return;
    }
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const int64_t q = f * g.n1 + p;
        double r;
        if (!exists_f(g, f, i, j, k)) {
            r = 0.0;
        } else if (!free_f(g, f, i, j, k)) {
            r = (MODE == OP_RESID) ? b[q] - (double)x[q] : (double)x[q];
        } else {
            double d, mx;
            if (f == FAX) mx = row_ax(xv, cv, i, j, k, op, B_ALL, d);
            else if (f == FAY) mx = row_ay(xv, cv, i, j, k, op, B_ALL, d);
            else if (f == FAZ) mx = row_az(xv, cv, i, j, k, op, B_ALL, d);
            else mx = row_v(xv, cv, i, j, k, op, B_ALL, d);
            r = (MODE == OP_RESID) ? (b[q] - mx) / d : mx / d;
        }
        out[q] = (TOUT)r;
    }
}

// Row diagonal D for every entry (1 on fixed rows, 0 where no unknown exists).
__global__ void k_diag(Grid g, CoeffsD c, OpParams op, double *__restrict__ D) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.n1) return;
    const int i = (int)(p % g.nx), j = (int)((p / g.nx) % g.ny), k = (int)(p / ((int64_t)g.nx * g.ny));
    View<double> xv{nullptr, nullptr, nullptr, nullptr, &g};
    CView cv{c.im[0], c.im[1], c.im[2], c.sg[0], c.sg[1], c.sg[2], &g};
    for (int f = 0; f < 4; ++f) {
        double d = 0.0;
        if (exists_f(g, f, i, j, k)) {
            if (!free_f(g, f, i, j, k)) d = 1.0;
            else if (f == FAX) (void)row_ax(xv, cv, i, j, k, op, 0, d);
            else if (f == FAY) (void)row_ay(xv, cv, i, j, k, op, 0, d);
            else if (f == FAZ) (void)row_az(xv, cv, i, j, k, op, 0, d);
            else (void)row_v(xv, cv, i, j, k, op, 0, d);
        }
        D[f * g.n1 + p] = d;
    }
}

// ---------------------------------------------------------------------------
// Right-hand side b(A^n), QMS_3D.f90:1459-1475, :1624, :1772, :1906-1928
// (steady: :459-470 etc. with zero interior RHS).
// ---------------------------------------------------------------------------
__global__ void k_rhs(Grid g, CoeffsD c, BC bc, OpParams op, const double *__restrict__ aprev,
                      double *__restrict__ b) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.n1) return;
    const int i = (int)(p % g.nx), j = (int)((p / g.nx) % g.ny), k = (int)(p / ((int64_t)g.nx * g.ny));
    const int64_t n1 = g.n1;
    for (int f = 0; f < 4; ++f) {
        double r = 0.0;
        if (exists_f(g, f, i, j, k)) {
            if (!free_f(g, f, i, j, k)) {
                r = bc_value(g, bc, op.v, f, i, j, k);
            } else if (op.tr) {
                const double t = op.dt;
                if (f < 3) {
                    r = (c.sg[f][p] / t) * aprev[f * n1 + p];
                } else {
                    const double *Ax = aprev, *Ay = aprev + n1, *Az = aprev + 2 * n1;
                    r = (c.sg[0][p] * Ax[p] - c.sg[0][gidx(g, i - 1, j, k)] * Ax[gidx(g, i - 1, j, k)]) / (t * g.hx)
                      + (c.sg[1][p] * Ay[p] - c.sg[1][gidx(g, i, j - 1, k)] * Ay[gidx(g, i, j - 1, k)]) / (t * g.hy)
                      + (c.sg[2][p] * Az[p] - c.sg[2][gidx(g, i, j, k - 1)] * Az[gidx(g, i, j, k - 1)]) / (t * g.hz);
                }
            }
        }
        b[f * n1 + p] = r;
    }
}

// x <- Dirichlet values on fixed entries, 0 where no unknown exists; free
// entries keep their value (warm start).
__global__ void k_setbc(Grid g, BC bc, OpParams op, double *__restrict__ x) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.n1) return;
    const int i = (int)(p % g.nx), j = (int)((p / g.nx) % g.ny), k = (int)(p / ((int64_t)g.nx * g.ny));
    for (int f = 0; f < 4; ++f) {
        const int64_t q = f * g.n1 + p;
        if (!exists_f(g, f, i, j, k)) x[q] = 0.0;
        else if (!free_f(g, f, i, j, k)) x[q] = bc_value(g, bc, op.v, f, i, j, k);
    }
}

// ---------------------------------------------------------------------------
// Cell-centred H (or B when unit_mu), QMS_3D.f90:1346-1376.
// Output: 3 arrays of cx*cy*cz, x fastest.
// ---------------------------------------------------------------------------
__global__ void k_hfield(Grid g, CoeffsD c, const double *__restrict__ x, double *__restrict__ H, int unit_mu) {
    const int cx = g.nx - 1, cy = g.ny - 1, cz = g.nz - 1;
    const int64_t nc = (int64_t)cx * cy * cz;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nc) return;
    const int i = (int)(q % cx), j = (int)((q / cx) % cy), k = (int)(q / ((int64_t)cx * cy));
    const int64_t n1 = g.n1;
    const double *Ax = x, *Ay = x + n1, *Az = x + 2 * n1;
    const double hx = g.hx, hy = g.hy, hz = g.hz;
    auto F = [&](const double *a, int ii, int jj, int kk) { return a[gidx(g, ii, jj, kk)]; };
    auto MU = [&](int comp, int ii, int jj, int kk) { return unit_mu ? 1.0 : 1.0 / c.im[comp][gidx(g, ii, jj, kk)]; };
    // The reference divides by Mu; 1/(1/Mu) is not always Mu bit-for-bit, so
    // the multiplication form below uses the stored inverse directly.
    auto IM = [&](int comp, int ii, int jj, int kk) { return unit_mu ? 1.0 : c.im[comp][gidx(g, ii, jj, kk)]; };
    (void)MU;
    H[q] = 0.5 * (((F(Az, i, j + 1, k) - F(Az, i, j, k)) / hy - (F(Ay, i, j, k + 1) - F(Ay, i, j, k)) / hz) * IM(0, i, j, k)
                + ((F(Az, i + 1, j + 1, k) - F(Az, i + 1, j, k)) / hy - (F(Ay, i + 1, j, k + 1) - F(Ay, i + 1, j, k)) / hz) * IM(0, i + 1, j, k));
    H[nc + q] = 0.5 * (((F(Ax, i, j, k + 1) - F(Ax, i, j, k)) / hz - (F(Az, i + 1, j, k) - F(Az, i, j, k)) / hx) * IM(1, i, j, k)
                     + ((F(Ax, i, j + 1, k + 1) - F(Ax, i, j + 1, k)) / hz - (F(Az, i + 1, j + 1, k) - F(Az, i, j + 1, k)) / hx) * IM(1, i, j + 1, k));
    H[2 * nc + q] = 0.5 * (((F(Ay, i + 1, j, k) - F(Ay, i, j, k)) / hx - (F(Ax, i, j + 1, k) - F(Ax, i, j, k)) / hy) * IM(2, i, j, k)
                         + ((F(Ay, i + 1, j, k + 1) - F(Ay, i, j, k + 1)) / hx - (F(Ax, i, j + 1, k + 1) - F(Ax, i, j, k + 1)) / hy) * IM(2, i, j, k + 1));
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static inline dim3 grid1(int64_t n, int bs = 256) { return dim3((unsigned)((n + bs - 1) / bs)); }

void launch_coeffs(const Grid &g, const uint16_t *slot, const double *mat, const Coeffs &c, hipStream_t s) {
    hipLaunchKernelGGL(k_coeffs, grid1(g.n1), dim3(256), 0, s, g, slot, mat, c);
}
void launch_diag(const Grid &g, const CoeffsD &c, const OpParams &op, double *D, hipStream_t s) {
    hipLaunchKernelGGL(k_diag, grid1(g.n1), dim3(256), 0, s, g, c, op, D);
}
void launch_rhs(const Grid &g, const CoeffsD &c, const BC &bc, const OpParams &op, const double *aprev, double *b,
                hipStream_t s) {
    hipLaunchKernelGGL(k_rhs, grid1(g.n1), dim3(256), 0, s, g, c, bc, op, aprev, b);
}
void launch_setbc(const Grid &g, const BC &bc, const OpParams &op, double *x, hipStream_t s) {
    hipLaunchKernelGGL(k_setbc, grid1(g.n1), dim3(256), 0, s, g, bc, op, x);
}
void launch_resid(const Grid &g, const CoeffsD &c, const OpParams &op, const double *x, const double *b, double *r,
                  hipStream_t s) {
    hipLaunchKernelGGL((k_operator<double, double, OP_RESID>), grid1(g.n1), dim3(256), 0, s, g, c, op, x, b, r);
}
void launch_apply(const Grid &g, const CoeffsD &c, const OpParams &op, const double *x, double *y, hipStream_t s) {
    hipLaunchKernelGGL((k_operator<double, double, OP_APPLY>), grid1(g.n1), dim3(256), 0, s, g, c, op, x,
                       (const double *)nullptr, y);
}
void launch_va_f32(const Grid &g, const CoeffsD &c, const OpParams &op, const float *x, float *y, hipStream_t s) {
    hipLaunchKernelGGL((k_operator<float, float, OP_VA>), grid1(g.n1), dim3(256), 0, s, g, c, op, x,
                       (const double *)nullptr, y);
}
void launch_va_f64(const Grid &g, const CoeffsD &c, const OpParams &op, const double *x, double *y, hipStream_t s) {
    hipLaunchKernelGGL((k_operator<double, double, OP_VA>), grid1(g.n1), dim3(256), 0, s, g, c, op, x,
                       (const double *)nullptr, y);
}
void launch_hfield(const Grid &g, const CoeffsD &c, const double *x, double *H, int unit_mu, hipStream_t s) {
    const int64_t nc = (int64_t)(g.nx - 1) * (g.ny - 1) * (g.nz - 1);
    hipLaunchKernelGGL(k_hfield, grid1(nc), dim3(256), 0, s, g, c, x, H, unit_mu);
}

}  // namespace qms
