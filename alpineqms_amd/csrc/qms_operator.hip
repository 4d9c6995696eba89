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
// diagonal (the scaling D used by the Krylov solver).  `k.tr` enables the
// transient sigma/dt terms, `k.adv` the frame-velocity terms.
//
// Accessors take a node offset o = di + dj*sy + dk*sz relative to the row's
// node; the field and coefficient pointers are already advanced to it.
// ---------------------------------------------------------------------------
template <typename T>
struct View {
    const T *ax, *ay, *az, *v;
    __device__ __forceinline__ double AX(int o) const { return (double)ax[o]; }
    __device__ __forceinline__ double AY(int o) const { return (double)ay[o]; }
    __device__ __forceinline__ double AZ(int o) const { return (double)az[o]; }
    __device__ __forceinline__ double VV(int o) const { return (double)v[o]; }
};

struct CView {
    const double *imx, *imy, *imz, *sx, *sy, *sz;
    __device__ __forceinline__ double IMX(int o) const { return imx[o]; }
    __device__ __forceinline__ double IMY(int o) const { return imy[o]; }
    __device__ __forceinline__ double IMZ(int o) const { return imz[o]; }
    __device__ __forceinline__ double SX(int o) const { return sx[o]; }
    __device__ __forceinline__ double SY(int o) const { return sy[o]; }
    __device__ __forceinline__ double SZ(int o) const { return sz[o]; }
};

template <typename T>
__device__ __forceinline__ View<T> view_at(const T *x, int64_t n1, int64_t p) {
    return View<T>{x + p, x + n1 + p, x + 2 * n1 + p, x + 3 * n1 + p};
}
__device__ __forceinline__ CView cview_at(const CoeffsD &c, int64_t p) {
    return CView{c.im[0] + p, c.im[1] + p, c.im[2] + p, c.sg[0] + p, c.sg[1] + p, c.sg[2] + p};
}

// Which blocks of the row to include.
enum : int { B_AA = 1, B_AV = 2, B_VA = 4, B_VV = 8, B_ALL = 15 };

// Ax row, QMS_3D.f90:476-607 (+ :1504, :1475)
template <typename T>
__device__ __forceinline__ double row_ax(const View<T> &x, const CView &c, const OpK &k, int blocks, double &diag) {
    const int X = 1, Y = k.sy, Z = k.sz;
    const double s = c.SX(0);
    const double imy_km = c.IMY(-Z), imy_k = c.IMY(0);
    const double imz_jm = c.IMZ(-Y), imz_j = c.IMZ(0);
    const double imx_im = c.IMX(-X), imx_i = c.IMX(0);
    const double ihx2 = k.ih2[0], ihy2 = k.ih2[1], ihz2 = k.ih2[2];
    double d = (imz_j + imz_jm) * ihy2 + (imy_k + imy_km) * ihz2 + (imx_i + imx_im) * ihx2;
    if (k.tr) d = d + s * k.idt;
    diag = d;
    double y = 0.0;
    if (blocks & B_AA) {
        const double x0 = x.AX(0);
        const double xzm = x.AX(-Z), xym = x.AX(-Y), xxm = x.AX(-X), xxp = x.AX(X), xyp = x.AX(Y), xzp = x.AX(Z);
        y += imy_km * ihz2 * (x0 - xzm) + imz_jm * ihy2 * (x0 - xym) + imx_im * ihx2 * (x0 - xxm)
           + imx_i * ihx2 * (x0 - xxp) + imz_j * ihy2 * (x0 - xyp) + imy_k * ihz2 * (x0 - xzp);
        if (k.tr) y += (s * k.idt) * x0;
        const double ay_m0 = x.AY(-Y), ay_p0 = x.AY(X - Y), ay_00 = x.AY(0), ay_10 = x.AY(X);
        const double az_m0 = x.AZ(-Z), az_p0 = x.AZ(X - Z), az_00 = x.AZ(0), az_10 = x.AZ(X);
        const double ixy = k.ihh[0][1], ixz = k.ihh[0][2];
        y += imz_jm * ixy * (ay_m0 - ay_p0) + imz_j * ixy * (ay_10 - ay_00);
        y += imy_km * ixz * (az_m0 - az_p0) + imy_k * ixz * (az_10 - az_00);
        if (k.adv) {
            y += s * k.v1[2][2] * (xzm - xzp) + s * k.v1[1][1] * (xym - xyp);
            y += s * k.v1[1][0] * (-ay_m0 + ay_p0 - ay_00 + ay_10);
            y += s * k.v1[2][0] * (-az_m0 + az_p0 - az_00 + az_10);
        }
    }
    if (blocks & B_AV) y += s * k.ih[0] * (x.VV(X) - x.VV(0));
    return y;
}

// Ay row, QMS_3D.f90:650-781 (+ :1681, :1624)
template <typename T>
__device__ __forceinline__ double row_ay(const View<T> &x, const CView &c, const OpK &k, int blocks, double &diag) {
    const int X = 1, Y = k.sy, Z = k.sz;
    const double s = c.SY(0);
    const double imx_km = c.IMX(-Z), imx_k = c.IMX(0);
    const double imy_jm = c.IMY(-Y), imy_j = c.IMY(0);
    const double imz_im = c.IMZ(-X), imz_i = c.IMZ(0);
    const double ihx2 = k.ih2[0], ihy2 = k.ih2[1], ihz2 = k.ih2[2];
    double d = (imz_i + imz_im) * ihx2 + (imx_k + imx_km) * ihz2 + (imy_j + imy_jm) * ihy2;
    if (k.tr) d = d + s * k.idt;
    diag = d;
    double y = 0.0;
    if (blocks & B_AA) {
        const double x0 = x.AY(0);
        const double xzm = x.AY(-Z), xym = x.AY(-Y), xxm = x.AY(-X), xxp = x.AY(X), xyp = x.AY(Y), xzp = x.AY(Z);
        y += imx_km * ihz2 * (x0 - xzm) + imy_jm * ihy2 * (x0 - xym) + imz_im * ihx2 * (x0 - xxm)
           + imz_i * ihx2 * (x0 - xxp) + imy_j * ihy2 * (x0 - xyp) + imx_k * ihz2 * (x0 - xzp);
        if (k.tr) y += (s * k.idt) * x0;
        const double ax_m0 = x.AX(-X), ax_00 = x.AX(0), ax_m1 = x.AX(Y - X), ax_01 = x.AX(Y);
        const double az_m0 = x.AZ(-Z), az_m1 = x.AZ(Y - Z), az_00 = x.AZ(0), az_01 = x.AZ(Y);
        const double ixy = k.ihh[0][1], iyz = k.ihh[1][2];
        y += imz_im * ixy * (ax_m0 - ax_m1) + imz_i * ixy * (ax_01 - ax_00);
        y += imx_km * iyz * (az_m0 - az_m1) + imx_k * iyz * (az_01 - az_00);
        if (k.adv) {
            y += s * k.v1[2][2] * (xzm - xzp) + s * k.v1[0][0] * (xxm - xxp);
            y += s * k.v1[0][1] * (-ax_m0 - ax_00 + ax_m1 + ax_01);
            y += s * k.v1[2][1] * (-az_m0 + az_m1 - az_00 + az_01);
        }
    }
    if (blocks & B_AV) y += s * k.ih[1] * (x.VV(Y) - x.VV(0));
    return y;
}

// Az row, QMS_3D.f90:824-955 (+ :1857, :1772)
template <typename T>
__device__ __forceinline__ double row_az(const View<T> &x, const CView &c, const OpK &k, int blocks, double &diag) {
    const int X = 1, Y = k.sy, Z = k.sz;
    const double s = c.SZ(0);
    const double imz_km = c.IMZ(-Z), imz_k = c.IMZ(0);
    const double imx_jm = c.IMX(-Y), imx_j = c.IMX(0);
    const double imy_im = c.IMY(-X), imy_i = c.IMY(0);
    const double ihx2 = k.ih2[0], ihy2 = k.ih2[1], ihz2 = k.ih2[2];
    double d = (imy_i + imy_im) * ihx2 + (imx_j + imx_jm) * ihy2 + (imz_k + imz_km) * ihz2;
    if (k.tr) d = d + s * k.idt;
    diag = d;
    double y = 0.0;
    if (blocks & B_AA) {
        const double x0 = x.AZ(0);
        const double xzm = x.AZ(-Z), xym = x.AZ(-Y), xxm = x.AZ(-X), xxp = x.AZ(X), xyp = x.AZ(Y), xzp = x.AZ(Z);
        y += imz_km * ihz2 * (x0 - xzm) + imx_jm * ihy2 * (x0 - xym) + imy_im * ihx2 * (x0 - xxm)
           + imy_i * ihx2 * (x0 - xxp) + imx_j * ihy2 * (x0 - xyp) + imz_k * ihz2 * (x0 - xzp);
        if (k.tr) y += (s * k.idt) * x0;
        const double ax_m0 = x.AX(-X), ax_00 = x.AX(0), ax_m1 = x.AX(Z - X), ax_01 = x.AX(Z);
        const double ay_m0 = x.AY(-Y), ay_00 = x.AY(0), ay_m1 = x.AY(Z - Y), ay_01 = x.AY(Z);
        const double ixz = k.ihh[0][2], iyz = k.ihh[1][2];
        y += imy_im * ixz * (ax_m0 - ax_m1) + imy_i * ixz * (ax_01 - ax_00);
        y += imx_jm * iyz * (ay_m0 - ay_m1) + imx_j * iyz * (ay_01 - ay_00);
        if (k.adv) {
            y += s * k.v1[1][1] * (xym - xyp) + s * k.v1[0][0] * (xxm - xxp);
            y += s * k.v1[0][2] * (-ax_m0 - ax_00 + ax_m1 + ax_01);
            y += s * k.v1[1][2] * (-ay_m0 - ay_00 + ay_m1 + ay_01);
        }
    }
    if (blocks & B_AV) y += s * k.ih[2] * (x.VV(Z) - x.VV(0));
    return y;
}

// V row, QMS_3D.f90:1003-1285 (+ :1975-2126 transient terms)
template <typename T>
__device__ __forceinline__ double row_v(const View<T> &x, const CView &c, const OpK &k, int blocks, double &diag) {
    const int X = 1, Y = k.sy, Z = k.sz;
    const double ihx2 = k.ih2[0], ihy2 = k.ih2[1], ihz2 = k.ih2[2];
    const double x0 = c.SX(0), x1 = c.SX(-X);
    const double y0 = c.SY(0), y1 = c.SY(-Y);
    const double z0 = c.SZ(0), z1 = c.SZ(-Z);
    const double fx0 = fmax(x0, QMS_SIGFLOOR), fx1 = fmax(x1, QMS_SIGFLOOR);
    const double fy0 = fmax(y0, QMS_SIGFLOOR), fy1 = fmax(y1, QMS_SIGFLOOR);
    const double fz0 = fmax(z0, QMS_SIGFLOOR), fz1 = fmax(z1, QMS_SIGFLOOR);
    diag = (fx0 + fx1) * ihx2 + (fy0 + fy1) * ihy2 + (fz0 + fz1) * ihz2;
    double y = 0.0;
    if (blocks & B_VV) {
        const double v0 = x.VV(0);
        y += fz1 * ihz2 * (v0 - x.VV(-Z)) + fy1 * ihy2 * (v0 - x.VV(-Y)) + fx1 * ihx2 * (v0 - x.VV(-X))
           + fx0 * ihx2 * (v0 - x.VV(X)) + fy0 * ihy2 * (v0 - x.VV(Y)) + fz0 * ihz2 * (v0 - x.VV(Z));
    }
    if (blocks & B_VA) {
        if (k.tr) {
            y += x0 * k.iht[0] * x.AX(0) - x1 * k.iht[0] * x.AX(-X)
               + y0 * k.iht[1] * x.AY(0) - y1 * k.iht[1] * x.AY(-Y)
               + z0 * k.iht[2] * x.AZ(0) - z1 * k.iht[2] * x.AZ(-Z);
        }
        if (k.adv) {
            // W(a,b,c) = v_a / (2 h_b h_c); axes x=0, y=1, z=2
#define W(a, b, cc) k.v2[a][b][cc]
            double w = 0.0;
            // Ax columns
            w += (x1 * W(2, 2, 0) - z1 * W(0, 2, 2)) * x.AX(-X - Z);
            w += (-x0 * W(2, 2, 0) - z1 * W(0, 2, 2)) * x.AX(-Z);
            w += (x1 * W(1, 1, 0) - y1 * W(0, 1, 1)) * x.AX(-X - Y);
            w += (-x0 * W(1, 1, 0) - y1 * W(0, 1, 1)) * x.AX(-Y);
            const double axc = (y0 + y1) * W(0, 1, 1) + (z0 + z1) * W(0, 2, 2);
            w += axc * (x.AX(-X) + x.AX(0));
            w += (-x1 * W(1, 1, 0) - y0 * W(0, 1, 1)) * x.AX(Y - X);
            w += (x0 * W(1, 1, 0) - y0 * W(0, 1, 1)) * x.AX(Y);
            w += (-x1 * W(2, 2, 0) - z0 * W(0, 2, 2)) * x.AX(Z - X);
            w += (x0 * W(2, 2, 0) - z0 * W(0, 2, 2)) * x.AX(Z);
            // Ay columns
            w += (y1 * W(2, 2, 1) - z1 * W(1, 2, 2)) * x.AY(-Y - Z);
            w += (-y0 * W(2, 2, 1) - z1 * W(1, 2, 2)) * x.AY(-Z);
            w += (-x1 * W(1, 0, 0) + y1 * W(0, 0, 1)) * x.AY(-X - Y);
            const double ayc = (x0 + x1) * W(1, 0, 0) + (z0 + z1) * W(1, 2, 2);
            w += ayc * (x.AY(-Y) + x.AY(0));
            w += (-x0 * W(1, 0, 0) - y1 * W(0, 0, 1)) * x.AY(X - Y);
            w += (-x1 * W(1, 0, 0) - y0 * W(0, 0, 1)) * x.AY(-X);
            w += (-x0 * W(1, 0, 0) + y0 * W(0, 0, 1)) * x.AY(X);
            w += (-y1 * W(2, 2, 1) - z0 * W(1, 2, 2)) * x.AY(Z - Y);
            w += (y0 * W(2, 2, 1) - z0 * W(1, 2, 2)) * x.AY(Z);
            // Az columns
            w += (-y1 * W(2, 1, 1) + z1 * W(1, 1, 2)) * x.AZ(-Y - Z);
            w += (-x1 * W(2, 0, 0) + z1 * W(0, 0, 2)) * x.AZ(-X - Z);
            const double azc = (x0 + x1) * W(2, 0, 0) + (y0 + y1) * W(2, 1, 1);
            w += azc * (x.AZ(-Z) + x.AZ(0));
            w += (-x0 * W(2, 0, 0) - z1 * W(0, 0, 2)) * x.AZ(X - Z);
            w += (-y0 * W(2, 1, 1) - z1 * W(1, 1, 2)) * x.AZ(Y - Z);
            w += (-y1 * W(2, 1, 1) - z0 * W(1, 1, 2)) * x.AZ(-Y);
            w += (-x1 * W(2, 0, 0) - z0 * W(0, 0, 2)) * x.AZ(-X);
            w += (-x0 * W(2, 0, 0) + z0 * W(0, 0, 2)) * x.AZ(X);
            w += (-y0 * W(2, 1, 1) + z0 * W(1, 1, 2)) * x.AZ(Y);
#undef W
            y += w;
        }
    }
    return y;
}

// ---------------------------------------------------------------------------
// Tiled launch geometry.  A workgroup of 256 threads owns a tile of
// tx x (256/tx) nodes of one z-plane.  Tiles are numbered x-fastest, then y,
// then z, and the hardware workgroup index is remapped so that each of the
// 8 XCDs (which receive workgroups round-robin) works on one contiguous
// range of tiles: neighbouring rows and planes then share that XCD's L2.
// ---------------------------------------------------------------------------
struct Tiling {
    int txs;            // log2 of the tile width
    int ntx, nty;       // tiles per row / per plane column
};

__device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned nb) {
    const unsigned q = nb >> 3, r = nb & 7u, x = b & 7u;
    return x * q + min(x, r) + (b >> 3);
}

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
__global__ void __launch_bounds__(256) k_operator(Grid g, Tiling t, CoeffsD c, OpK k, const TIN *__restrict__ x,
                                                  const double *__restrict__ b, TOUT *__restrict__ out) {
    int i, j, kk;
    if (!tile_node(g, t, i, j, kk)) return;
    const int64_t p = gidx(g, i, j, kk);
    const View<TIN> xv = view_at(x, g.n1, p);
    const CView cv = cview_at(c, p);
    if (MODE == OP_VA) {
        if (free_v(g, i, j, kk)) {
            double d;
            out[3 * g.n1 + p] = (TOUT)row_v(xv, cv, k, B_VA, d);
        }
        return;
    }
    if (free_v(g, i, j, kk)) {
        // interior node: all four rows are free
        double d, mx;
        mx = row_ax(xv, cv, k, B_ALL, d);
        out[p] = (TOUT)((MODE == OP_RESID) ? (b[p] - mx) / d : mx / d);
        mx = row_ay(xv, cv, k, B_ALL, d);
        out[g.n1 + p] = (TOUT)((MODE == OP_RESID) ? (b[g.n1 + p] - mx) / d : mx / d);
        mx = row_az(xv, cv, k, B_ALL, d);
        out[2 * g.n1 + p] = (TOUT)((MODE == OP_RESID) ? (b[2 * g.n1 + p] - mx) / d : mx / d);
        mx = row_v(xv, cv, k, B_ALL, d);
        out[3 * g.n1 + p] = (TOUT)((MODE == OP_RESID) ? (b[3 * g.n1 + p] - mx) / d : mx / d);
        return;
    }
    for (int f = 0; f < 4; ++f) {
        const int64_t q = f * g.n1 + p;
        double r;
        if (!exists_f(g, f, i, j, kk)) {
            r = 0.0;
        } else if (!free_f(g, f, i, j, kk)) {
            r = (MODE == OP_RESID) ? b[q] - (double)x[q] : (double)x[q];
        } else {
            // only Ax (i = 1), Ay (j = 1) and Az (k = 1) rows are free outside the V interior
            double d, mx;
            if (f == FAX) mx = row_ax(xv, cv, k, B_ALL, d);
            else if (f == FAY) mx = row_ay(xv, cv, k, B_ALL, d);
            else mx = row_az(xv, cv, k, B_ALL, d);
            r = (MODE == OP_RESID) ? (b[q] - mx) / d : mx / d;
        }
        out[q] = (TOUT)r;
    }
}

// Row diagonal D for every entry (1 on fixed rows, 0 where no unknown exists).
__global__ void k_diag(Grid g, CoeffsD c, OpK k, double *__restrict__ D) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.n1) return;
    const int i = (int)(p % g.nx), j = (int)((p / g.nx) % g.ny), kk = (int)(p / ((int64_t)g.nx * g.ny));
    const View<double> xv{nullptr, nullptr, nullptr, nullptr};
    const CView cv = cview_at(c, p);
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
__global__ void __launch_bounds__(256) k_rhs(Grid g, Tiling t, CoeffsD c, BC bc, OpK k,
                                             const double *__restrict__ aprev, double *__restrict__ b) {
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
__global__ void k_setbc(Grid g, BC bc, OpK k, double *__restrict__ x) {
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
        k.ih2[a] = 1.0 / (h[a] * h[a]);
        k.ih[a] = 1.0 / h[a];
        k.iht[a] = op.tr ? 1.0 / (h[a] * op.dt) : 0.0;
        for (int b = 0; b < 3; ++b) {
            k.ihh[a][b] = 1.0 / (h[a] * h[b]);
            k.v1[a][b] = op.v[a] / (2.0 * h[b]);
            for (int c = 0; c < 3; ++c) k.v2[a][b][c] = op.v[a] / (2.0 * h[b] * h[c]);
        }
    }
    return k;
}

static Tiling tiling(const Grid &g, int tile_x, dim3 &grid) {
    if (tile_x != 32 && tile_x != 64 && tile_x != 128 && tile_x != 256) tile_x = 64;
    Tiling t;
    t.txs = tile_x == 32 ? 5 : tile_x == 64 ? 6 : tile_x == 128 ? 7 : 8;
    const int rows = 256 >> t.txs;
    t.ntx = (g.nx + tile_x - 1) / tile_x;
    t.nty = (g.ny + rows - 1) / rows;
    grid = dim3((unsigned)((int64_t)t.ntx * t.nty * g.nz));
    return t;
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
    const Tiling t = tiling(g, op.tile_x, grid);
    hipLaunchKernelGGL(k_rhs, grid, dim3(256), 0, s, g, t, c, bc, make_opk(g, op), aprev, b);
}
void launch_setbc(const Grid &g, const BC &bc, const OpParams &op, double *x, hipStream_t s) {
    hipLaunchKernelGGL(k_setbc, grid1(g.n1), dim3(256), 0, s, g, bc, make_opk(g, op), x);
}
void launch_resid(const Grid &g, const CoeffsD &c, const OpParams &op, const double *x, const double *b, double *r,
                  hipStream_t s) {
    dim3 grid;
    const Tiling t = tiling(g, op.tile_x, grid);
    hipLaunchKernelGGL((k_operator<double, double, OP_RESID>), grid, dim3(256), 0, s, g, t, c, make_opk(g, op), x, b,
                       r);
}
void launch_apply(const Grid &g, const CoeffsD &c, const OpParams &op, const double *x, double *y, hipStream_t s) {
    dim3 grid;
    const Tiling t = tiling(g, op.tile_x, grid);
    hipLaunchKernelGGL((k_operator<double, double, OP_APPLY>), grid, dim3(256), 0, s, g, t, c, make_opk(g, op), x,
                       (const double *)nullptr, y);
}
void launch_va_f32(const Grid &g, const CoeffsD &c, const OpParams &op, const float *x, float *y, hipStream_t s) {
    dim3 grid;
    const Tiling t = tiling(g, op.tile_x, grid);
    hipLaunchKernelGGL((k_operator<float, float, OP_VA>), grid, dim3(256), 0, s, g, t, c, make_opk(g, op), x,
                       (const double *)nullptr, y);
}
void launch_va_f64(const Grid &g, const CoeffsD &c, const OpParams &op, const double *x, double *y, hipStream_t s) {
    dim3 grid;
    const Tiling t = tiling(g, op.tile_x, grid);
    hipLaunchKernelGGL((k_operator<double, double, OP_VA>), grid, dim3(256), 0, s, g, t, c, make_opk(g, op), x,
                       (const double *)nullptr, y);
}
void launch_hfield(const Grid &g, const CoeffsD &c, const double *x, double *H, int unit_mu, hipStream_t s) {
    const int64_t nc = (int64_t)(g.nx - 1) * (g.ny - 1) * (g.nz - 1);
    hipLaunchKernelGGL(k_hfield, grid1(nc), dim3(256), 0, s, g, c, x, H, unit_mu);
}

}  // namespace qms
