// Grid geometry, unknown layout and Dirichlet shell of the QMS_3D Yee grid.
//
// Device layout (one design): every field is stored node-sized, x fastest:
//   p = i + nx*(j + ny*k),  0 <= i < nx, 0 <= j < ny, 0 <= k < nz
// with nx = cells_x + 1 etc. (QMS_3D.f90:172-174, node-sized storage
// :288-299).  The four unknown fields Ax, Ay, Az, V sit at offsets
// 0, n1, 2*n1, 3*n1 of a 4*n1 vector.  Entries outside a component's
// sub-box (Ax at i = nx-1, Ay at j = ny-1, Az at k = nz-1) are always 0.
//
// Indices here are 0-based; the reference's 1-based (i,j,k) is (i+1,j+1,k+1).
#pragma once
#include <stdint.h>

#ifndef QMS_HD
#ifdef __HIPCC__
#define QMS_HD __host__ __device__ __forceinline__
#else
#define QMS_HD inline
#endif
#endif

namespace qms {

enum Field { FAX = 0, FAY = 1, FAZ = 2, FV = 3 };

struct Grid {
    int nx, ny, nz;          // nodes per axis
    int64_t n1;              // nx*ny*nz
    double hx, hy, hz;       // del_x, del_y, del_z
};

QMS_HD int64_t gidx(const Grid &g, int i, int j, int k) {
    return (int64_t)i + (int64_t)g.nx * ((int64_t)j + (int64_t)g.ny * k);
}

// Free (non-Dirichlet) rows, QMS_3D.f90:361-396 restated 0-based:
//   Ax fixed if j<=1 | j>=ny-2 | k<=1 | k>=nz-2 | i==0 | i==nx-2
//   Ay fixed if i<=1 | i>=nx-2 | k<=1 | k>=nz-2 | j==0 | j==ny-2
//   Az fixed if i<=1 | i>=nx-2 | j<=1 | j>=ny-2 | k==0 | k==nz-2
//   V  fixed if any index <=1 or >= n-2
QMS_HD bool free_ax(const Grid &g, int i, int j, int k) {
    return i >= 1 && i <= g.nx - 3 && j >= 2 && j <= g.ny - 3 && k >= 2 && k <= g.nz - 3;
}
QMS_HD bool free_ay(const Grid &g, int i, int j, int k) {
    return i >= 2 && i <= g.nx - 3 && j >= 1 && j <= g.ny - 3 && k >= 2 && k <= g.nz - 3;
}
QMS_HD bool free_az(const Grid &g, int i, int j, int k) {
    return i >= 2 && i <= g.nx - 3 && j >= 2 && j <= g.ny - 3 && k >= 1 && k <= g.nz - 3;
}
QMS_HD bool free_v(const Grid &g, int i, int j, int k) {
    return i >= 2 && i <= g.nx - 3 && j >= 2 && j <= g.ny - 3 && k >= 2 && k <= g.nz - 3;
}
QMS_HD bool free_f(const Grid &g, int f, int i, int j, int k) {
    return f == FAX ? free_ax(g, i, j, k) : f == FAY ? free_ay(g, i, j, k)
         : f == FAZ ? free_az(g, i, j, k) : free_v(g, i, j, k);
}
// Entry exists in the reference's unknown vector (QMS_3D.f90:175, :1320-1343).
QMS_HD bool exists_f(const Grid &g, int f, int i, int j, int k) {
    return f == FAX ? i <= g.nx - 2 : f == FAY ? j <= g.ny - 2 : f == FAZ ? k <= g.nz - 2 : true;
}

// Boundary data of one design: B0, gradient rows and the float32 literals of
// the reference (QMS_3D.f90:465, :1237-1281).
struct BC {
    double B[3];
    double G[9];   // gxx gxy gxz gyx gyy gyz gzx gzy gzz
};

#define QMS_THIRD ((double)(1.0f / 3.0f))
#define QMS_SIGFLOOR ((double)1.0e-12f)
#define QMS_MU0 1.25663706E-6

// Dirichlet values at the unstaggered node position (i*dx, j*dy, k*dz),
// QMS_3D.f90:459-470, :633-644, :808-818, :982-996.
QMS_HD double bc_value(const Grid &g, const BC &bc, const double v[3], int f, int i, int j, int k) {
    const double x = i * g.hx, y = j * g.hy, z = k * g.hz;
    const double *G = bc.G;
    const double gx = G[0] * x + G[1] * y + G[2] * z;
    const double gy = G[3] * x + G[4] * y + G[5] * z;
    const double gz = G[6] * x + G[7] * y + G[8] * z;
    const double Bx = bc.B[0], By = bc.B[1], Bz = bc.B[2];
    if (f == FAX) return 0.5 * By * z - 0.5 * Bz * y + QMS_THIRD * (gy * z - gz * y);
    if (f == FAY) return 0.5 * Bz * x - 0.5 * Bx * z + QMS_THIRD * (gz * x - gx * z);
    if (f == FAZ) return 0.5 * Bx * y - 0.5 * By * x + QMS_THIRD * (gx * y - gy * x);
    const double vx = v[0], vy = v[1], vz = v[2];
    return -(1.0 * (vy * Bz - vz * By) * x) - 1.0 * (vz * Bx - vx * Bz) * y
           - 1.0 * (vx * By - vy * Bx) * z - 0.5 * (vy * gz - vz * gy) * x
           - 0.5 * (vz * gx - vx * gz) * y - 0.5 * (vx * gy - vy * gx) * z;
}

}  // namespace qms
