// Geometric multigrid for the two diagonal blocks of the block-triangular
// preconditioner:
//   A block: K + Sigma/dt restricted to the free Ax, Ay, Az unknowns
//            (curl-curl + component-wise gauge penalty + mass, no advection),
//            coupled across the three components.
//   V block: the sigma-floored Laplacian of the V rows (QMS_3D.f90:1237-1281).
// Coarse spaces: 2x2x2 aggregation of each component's free unknowns
// (coarse index = fine index >> 1), Galerkin coarse operators.  With this
// aggregation the coarse operator keeps exactly the fine sparsity pattern
// (7-point self stencil plus the 4+4 cross-component offsets), so every level
// is stored as the same structured stencil.  All couplings are kept in
// difference form with an explicit excess diagonal; in FP32 this preserves the
// relative accuracy of the 1e-3 S/m background links next to 6e7 S/m
// conductors.
#include <hip/hip_runtime.h>
#include <math.h>
#include <vector>
#include <stdexcept>
#include <string>
#include "qms_kernels.h"

namespace qms {

// Self-neighbour offsets: -x, +x, -y, +y, -z, +z
__device__ __forceinline__ void self_off(int d, int &di, int &dj, int &dk) {
    di = (d == 0) ? -1 : (d == 1) ? 1 : 0;
    dj = (d == 2) ? -1 : (d == 3) ? 1 : 0;
    dk = (d == 4) ? -1 : (d == 5) ? 1 : 0;
}

// Cross-coupling target (component, offset) for component c and slot s.
// Slots 0-3 couple to the first other component, 4-7 to the second, with the
// offsets of the reference's rows (QMS_3D.f90:546-592, :727-770, :844-888).
__device__ __forceinline__ void cross_off(int c, int s, int &tc, int &di, int &dj, int &dk) {
    // table[c][s] = {tc, di, dj, dk}
    const int T[3][8][4] = {
        {{1, 0, -1, 0}, {1, 1, -1, 0}, {1, 0, 0, 0}, {1, 1, 0, 0}, {2, 0, 0, -1}, {2, 1, 0, -1}, {2, 0, 0, 0}, {2, 1, 0, 0}},
        {{0, -1, 0, 0}, {0, 0, 0, 0}, {0, -1, 1, 0}, {0, 0, 1, 0}, {2, 0, 0, -1}, {2, 0, 1, -1}, {2, 0, 0, 0}, {2, 0, 1, 0}},
        {{0, -1, 0, 0}, {0, 0, 0, 0}, {0, -1, 0, 1}, {0, 0, 0, 1}, {1, 0, -1, 0}, {1, 0, 0, 0}, {1, 0, -1, 1}, {1, 0, 0, 1}}};
    tc = T[c][s][0]; di = T[c][s][1]; dj = T[c][s][2]; dk = T[c][s][3];
}

struct Dims {
    int nx, ny, nz;
    int64_t n1;
};
__device__ __forceinline__ int64_t didx(const Dims &d, int i, int j, int k) {
    return (int64_t)i + (int64_t)d.nx * ((int64_t)j + (int64_t)d.ny * k);
}
__device__ __forceinline__ bool inside(const Dims &d, int i, int j, int k) {
    return i >= 0 && j >= 0 && k >= 0 && i < d.nx && j < d.ny && k < d.nz;
}

// ---------------------------------------------------------------------------
// Level-0 providers: stencil rows computed on the fly from the material
// coefficients, with couplings to fixed (Dirichlet) unknowns eliminated
// (conductance moved to the excess diagonal, cross coefficient dropped) since
// the preconditioner acts on corrections that vanish there.
// ---------------------------------------------------------------------------
template <typename T>
struct FineA {
    Grid g;
    const T *im[3];
    const T *sg[3];
    T inv_dt;       // 1/dt, 0 for steady
    static constexpr int NC = 3;
    __device__ __forceinline__ Dims dims() const { return Dims{g.nx, g.ny, g.nz, g.n1}; }
    __device__ __forceinline__ bool active(int c, int i, int j, int k) const { return free_f(g, c, i, j, k); }
    __device__ __forceinline__ T IM(int a, int i, int j, int k) const { return im[a][gidx(g, i, j, k)]; }
    __device__ bool row(int c, int i, int j, int k, T cond[6], T &exc, T cross[8]) const {
        if (!free_f(g, c, i, j, k)) return false;
        const T hx2 = (T)(g.hx * g.hx), hy2 = (T)(g.hy * g.hy), hz2 = (T)(g.hz * g.hz);
        const T hxy = (T)(g.hx * g.hy), hxz = (T)(g.hx * g.hz), hyz = (T)(g.hy * g.hz);
        T cm;
        if (c == 0) {
            cond[0] = IM(0, i - 1, j, k) / hx2; cond[1] = IM(0, i, j, k) / hx2;
            cond[2] = IM(2, i, j - 1, k) / hy2; cond[3] = IM(2, i, j, k) / hy2;
            cond[4] = IM(1, i, j, k - 1) / hz2; cond[5] = IM(1, i, j, k) / hz2;
            const T a = IM(2, i, j - 1, k) / hxy, b = IM(2, i, j, k) / hxy;
            const T e = IM(1, i, j, k - 1) / hxz, f = IM(1, i, j, k) / hxz;
            cross[0] = a; cross[1] = -a; cross[2] = -b; cross[3] = b;
            cross[4] = e; cross[5] = -e; cross[6] = -f; cross[7] = f;
            cm = sg[0][gidx(g, i, j, k)];
        } else if (c == 1) {
            cond[0] = IM(2, i - 1, j, k) / hx2; cond[1] = IM(2, i, j, k) / hx2;
            cond[2] = IM(1, i, j - 1, k) / hy2; cond[3] = IM(1, i, j, k) / hy2;
            cond[4] = IM(0, i, j, k - 1) / hz2; cond[5] = IM(0, i, j, k) / hz2;
            const T a = IM(2, i - 1, j, k) / hxy, b = IM(2, i, j, k) / hxy;
            const T e = IM(0, i, j, k - 1) / hyz, f = IM(0, i, j, k) / hyz;
            cross[0] = a; cross[1] = -b; cross[2] = -a; cross[3] = b;
            cross[4] = e; cross[5] = -e; cross[6] = -f; cross[7] = f;
            cm = sg[1][gidx(g, i, j, k)];
        } else {
            cond[0] = IM(1, i - 1, j, k) / hx2; cond[1] = IM(1, i, j, k) / hx2;
            cond[2] = IM(0, i, j - 1, k) / hy2; cond[3] = IM(0, i, j, k) / hy2;
            cond[4] = IM(2, i, j, k - 1) / hz2; cond[5] = IM(2, i, j, k) / hz2;
            const T a = IM(1, i - 1, j, k) / hxz, b = IM(1, i, j, k) / hxz;
            const T e = IM(0, i, j - 1, k) / hyz, f = IM(0, i, j, k) / hyz;
            cross[0] = a; cross[1] = -b; cross[2] = -a; cross[3] = b;
            cross[4] = e; cross[5] = -f; cross[6] = -e; cross[7] = f;
            cm = sg[2][gidx(g, i, j, k)];
        }
        exc = cm * inv_dt;
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            int di, dj, dk;
            self_off(d, di, dj, dk);
            if (!free_f(g, c, i + di, j + dj, k + dk)) { exc += cond[d]; cond[d] = (T)0; }
        }
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            int tc, di, dj, dk;
            cross_off(c, s, tc, di, dj, dk);
            if (!free_f(g, tc, i + di, j + dj, k + dk)) cross[s] = (T)0;
        }
        return true;
    }
};

template <typename T>
struct FineV {
    Grid g;
    const T *sg[3];
    static constexpr int NC = 1;
    __device__ __forceinline__ Dims dims() const { return Dims{g.nx, g.ny, g.nz, g.n1}; }
    __device__ __forceinline__ bool active(int, int i, int j, int k) const { return free_v(g, i, j, k); }
    __device__ bool row(int, int i, int j, int k, T cond[6], T &exc, T *) const {
        if (!free_v(g, i, j, k)) return false;
        const T fl = (T)QMS_SIGFLOOR;
        const T hx2 = (T)(g.hx * g.hx), hy2 = (T)(g.hy * g.hy), hz2 = (T)(g.hz * g.hz);
        auto S = [&](int a, int ii, int jj, int kk) { const T s = sg[a][gidx(g, ii, jj, kk)]; return s > fl ? s : fl; };
        cond[0] = S(0, i - 1, j, k) / hx2; cond[1] = S(0, i, j, k) / hx2;
        cond[2] = S(1, i, j - 1, k) / hy2; cond[3] = S(1, i, j, k) / hy2;
        cond[4] = S(2, i, j, k - 1) / hz2; cond[5] = S(2, i, j, k) / hz2;
        exc = (T)0;
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            int di, dj, dk;
            self_off(d, di, dj, dk);
            if (!free_v(g, i + di, j + dj, k + dk)) { exc += cond[d]; cond[d] = (T)0; }
        }
        return true;
    }
};

// Explicit-stencil provider for levels >= 1.
template <typename T, int NCOMP>
struct Expl {
    Dims d;
    const T *cond, *exc, *cross, *dinv;
    static constexpr int NC = NCOMP;
    __device__ __forceinline__ Dims dims() const { return d; }
    __device__ __forceinline__ bool active(int c, int i, int j, int k) const {
        return dinv[(int64_t)c * d.n1 + didx(d, i, j, k)] != (T)0;
    }
    __device__ bool row(int c, int i, int j, int k, T cnd[6], T &ex, T crs[8]) const {
        const int64_t p = didx(d, i, j, k);
        if (dinv[(int64_t)c * d.n1 + p] == (T)0) return false;
#pragma unroll
        for (int q = 0; q < 6; ++q) cnd[q] = cond[((int64_t)c * 6 + q) * d.n1 + p];
        ex = exc[(int64_t)c * d.n1 + p];
        if (NCOMP == 3) {
#pragma unroll
            for (int s = 0; s < 8; ++s) crs[s] = cross[((int64_t)c * 8 + s) * d.n1 + p];
        }
        return true;
    }
};

// (A x) for component c at (i,j,k) given the row coefficients.
template <typename T, int NC>
__device__ __forceinline__ T apply_row(const Dims &d, const T *x, int c, int i, int j, int k, const T cond[6], T exc,
                                       const T cross[8]) {
    const int64_t p = didx(d, i, j, k);
    const T x0 = x[(int64_t)c * d.n1 + p];
    T y = exc * x0;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        if (cond[q] != (T)0) {
            int di, dj, dk;
            self_off(q, di, dj, dk);
            y += cond[q] * (x0 - x[(int64_t)c * d.n1 + didx(d, i + di, j + dj, k + dk)]);
        }
    }
    if (NC == 3) {
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            if (cross[s] != (T)0) {
                int tc, di, dj, dk;
                cross_off(c, s, tc, di, dj, dk);
                y += cross[s] * x[(int64_t)tc * d.n1 + didx(d, i + di, j + dj, k + dk)];
            }
        }
    }
    return y;
}

__device__ __forceinline__ bool node_of(const Dims &d, int64_t p, int &i, int &j, int &k) {
    if (p >= d.n1) return false;
    i = (int)(p % d.nx); j = (int)((p / d.nx) % d.ny); k = (int)(p / ((int64_t)d.nx * d.ny));
    return true;
}

// Damped Jacobi sweep: xo = x + w (b - A x)/diag; first sweep assumes x = 0.
template <typename T, typename P>
__global__ void __launch_bounds__(256) k_smooth(P pv, const T *__restrict__ x, const T *__restrict__ b,
                                                T *__restrict__ xo, T omega, int first) {
    const Dims d = pv.dims();
    int i, j, k;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (!node_of(d, p, i, j, k)) return;
#pragma unroll
    for (int c = 0; c < P::NC; ++c) {
        T cnd[6], ex, crs[8];
        const int64_t q = (int64_t)c * d.n1 + p;
        if (!pv.row(c, i, j, k, cnd, ex, crs)) { xo[q] = (T)0; continue; }
        T diag = ex;
#pragma unroll
        for (int s = 0; s < 6; ++s) diag += cnd[s];
        if (first) {
            xo[q] = omega * b[q] / diag;
        } else {
            const T ax = apply_row<T, P::NC>(d, x, c, i, j, k, cnd, ex, crs);
            xo[q] = x[q] + omega * (b[q] - ax) / diag;
        }
    }
}

template <typename T, typename P>
__global__ void __launch_bounds__(256) k_resid(P pv, const T *__restrict__ x, const T *__restrict__ b,
                                               T *__restrict__ r) {
    const Dims d = pv.dims();
    int i, j, k;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (!node_of(d, p, i, j, k)) return;
#pragma unroll
    for (int c = 0; c < P::NC; ++c) {
        T cnd[6], ex, crs[8];
        const int64_t q = (int64_t)c * d.n1 + p;
        if (!pv.row(c, i, j, k, cnd, ex, crs)) { r[q] = (T)0; continue; }
        r[q] = b[q] - apply_row<T, P::NC>(d, x, c, i, j, k, cnd, ex, crs);
    }
}

// bc = P^T r (sum over the 2x2x2 children)
template <typename T>
__global__ void k_restrict(Dims fd, Dims cd, int ncomp, const T *__restrict__ r, T *__restrict__ bc) {
    int I, J, K;
    const int64_t P = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (!node_of(cd, P, I, J, K)) return;
    for (int c = 0; c < ncomp; ++c) {
        T s = (T)0;
        for (int dk = 0; dk < 2; ++dk)
            for (int dj = 0; dj < 2; ++dj)
                for (int di = 0; di < 2; ++di) {
                    const int i = 2 * I + di, j = 2 * J + dj, k = 2 * K + dk;
                    if (inside(fd, i, j, k)) s += r[(int64_t)c * fd.n1 + didx(fd, i, j, k)];
                }
        bc[(int64_t)c * cd.n1 + P] = s;
    }
}

// x += alpha * P xc on active fine rows
template <typename T, typename PV>
__global__ void k_prolong(PV pv, Dims cd, const T *__restrict__ xc, T *__restrict__ x, T alpha) {
    const Dims d = pv.dims();
    int i, j, k;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (!node_of(d, p, i, j, k)) return;
    const int64_t P = didx(cd, i >> 1, j >> 1, k >> 1);
#pragma unroll
    for (int c = 0; c < PV::NC; ++c)
        if (pv.active(c, i, j, k)) x[(int64_t)c * d.n1 + p] += alpha * xc[(int64_t)c * cd.n1 + P];
}

// Galerkin coarse operator of level l (provider pv) into explicit arrays of
// level l+1.
template <typename T, typename P>
__global__ void k_galerkin(P pv, Dims cd, T *__restrict__ cond, T *__restrict__ exc, T *__restrict__ cross,
                           T *__restrict__ dinv) {
    const Dims fd = pv.dims();
    int I, J, K;
    const int64_t Pn = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (!node_of(cd, Pn, I, J, K)) return;
#pragma unroll 1
    for (int c = 0; c < P::NC; ++c) {
        T ccnd[6] = {0, 0, 0, 0, 0, 0}, cex = 0, ccrs[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        bool any = false;
#pragma unroll 1
        for (int ch = 0; ch < 8; ++ch) {
            const int i = 2 * I + (ch & 1), j = 2 * J + ((ch >> 1) & 1), k = 2 * K + (ch >> 2);
            if (!inside(fd, i, j, k)) continue;
            T cnd[6], ex, crs[8];
            if (!pv.row(c, i, j, k, cnd, ex, crs)) continue;
            any = true;
            cex += ex;
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                int di, dj, dk;
                self_off(q, di, dj, dk);
                const int ni = i + di, nj = j + dj, nk = k + dk;
                const bool internal = ((ni >> 1) == I) && ((nj >> 1) == J) && ((nk >> 1) == K);
                if (!internal) ccnd[q] += cnd[q];
            }
            if (P::NC == 3) {
#pragma unroll
                for (int s = 0; s < 8; ++s) {
                    if (crs[s] == (T)0) continue;
                    int tc, di, dj, dk;
                    cross_off(c, s, tc, di, dj, dk);
                    const int oi = ((i + di) >> 1) - I, oj = ((j + dj) >> 1) - J, ok = ((k + dk) >> 1) - K;
                    const int base = s & 4;
#pragma unroll
                    for (int s2 = 0; s2 < 4; ++s2) {
                        int tc2, ei, ej, ek;
                        cross_off(c, base + s2, tc2, ei, ej, ek);
                        if (ei == oi && ej == oj && ek == ok) ccrs[base + s2] += crs[s];
                    }
                }
            }
        }
        const int64_t q = (int64_t)c * cd.n1 + Pn;
        T diag = cex;
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            cond[((int64_t)c * 6 + s) * cd.n1 + Pn] = any ? ccnd[s] : (T)0;
            diag += ccnd[s];
        }
        exc[q] = any ? cex : (T)0;
        if (P::NC == 3) {
#pragma unroll
            for (int s = 0; s < 8; ++s) cross[((int64_t)c * 8 + s) * cd.n1 + Pn] = any ? ccrs[s] : (T)0;
        }
        dinv[q] = any ? (T)1 / diag : (T)0;
    }
}

// Coarsest level: x = Ainv * b over the active unknowns (one workgroup).
template <typename T>
__global__ void k_coarse_solve(int n, const int *__restrict__ idx, const double *__restrict__ inv,
                               const T *__restrict__ b, T *__restrict__ x) {
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        double s = 0.0;
        for (int c = 0; c < n; ++c) s += inv[(int64_t)r * n + c] * (double)b[idx[c]];
        x[idx[r]] = (T)s;
    }
}

template <typename T>
__global__ void k_zero(T *__restrict__ x, int64_t n) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) x[p] = (T)0;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static inline dim3 g1(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }
static inline Dims dims_of(const Level<float> &L) { return Dims{L.nx, L.ny, L.nz, L.n1}; }
template <typename T>
static inline Dims dims_of(const Level<T> &L) { return Dims{L.nx, L.ny, L.nz, L.n1}; }

template <typename T> struct FineCoef;
template <> struct FineCoef<float> {
    static const float *im(const Fine &f, int a) { return f.c32.im[a]; }
    static const float *sg(const Fine &f, int a) { return f.c32.sg[a]; }
};
template <> struct FineCoef<double> {
    static const double *im(const Fine &f, int a) { return f.c64.im[a]; }
    static const double *sg(const Fine &f, int a) { return f.c64.sg[a]; }
};

template <typename T>
static FineA<T> fineA(const Fine &f) {
    FineA<T> p;
    p.g = f.g;
    for (int a = 0; a < 3; ++a) { p.im[a] = FineCoef<T>::im(f, a); p.sg[a] = FineCoef<T>::sg(f, a); }
    p.inv_dt = f.dt > 0 ? (T)(1.0 / f.dt) : (T)0;
    return p;
}
template <typename T>
static FineV<T> fineV(const Fine &f) {
    FineV<T> p;
    p.g = f.g;
    for (int a = 0; a < 3; ++a) p.sg[a] = FineCoef<T>::sg(f, a);
    return p;
}
template <typename T, int NC>
static Expl<T, NC> expl(const Level<T> &L) {
    return Expl<T, NC>{dims_of(L), L.cond, L.exc, L.cross, L.dinv};
}

// Launch helpers dispatching on (level, block).
template <typename T>
static void smooth(const Fine &f, Hierarchy<T> &h, int l, const T *x, const T *b, T *xo, T w, int first,
                   hipStream_t s) {
    const Level<T> &L = h.lev[l];
    if (l == 0) {
        if (h.block == 0) hipLaunchKernelGGL((k_smooth<T, FineA<T>>), g1(L.n1), dim3(256), 0, s, fineA<T>(f), x, b, xo, w, first);
        else hipLaunchKernelGGL((k_smooth<T, FineV<T>>), g1(L.n1), dim3(256), 0, s, fineV<T>(f), x, b, xo, w, first);
    } else {
        if (h.block == 0) hipLaunchKernelGGL((k_smooth<T, Expl<T, 3>>), g1(L.n1), dim3(256), 0, s, expl<T, 3>(L), x, b, xo, w, first);
        else hipLaunchKernelGGL((k_smooth<T, Expl<T, 1>>), g1(L.n1), dim3(256), 0, s, expl<T, 1>(L), x, b, xo, w, first);
    }
}
template <typename T>
static void resid(const Fine &f, Hierarchy<T> &h, int l, const T *x, const T *b, T *r, hipStream_t s) {
    const Level<T> &L = h.lev[l];
    if (l == 0) {
        if (h.block == 0) hipLaunchKernelGGL((k_resid<T, FineA<T>>), g1(L.n1), dim3(256), 0, s, fineA<T>(f), x, b, r);
        else hipLaunchKernelGGL((k_resid<T, FineV<T>>), g1(L.n1), dim3(256), 0, s, fineV<T>(f), x, b, r);
    } else {
        if (h.block == 0) hipLaunchKernelGGL((k_resid<T, Expl<T, 3>>), g1(L.n1), dim3(256), 0, s, expl<T, 3>(L), x, b, r);
        else hipLaunchKernelGGL((k_resid<T, Expl<T, 1>>), g1(L.n1), dim3(256), 0, s, expl<T, 1>(L), x, b, r);
    }
}
template <typename T>
static void prolong(const Fine &f, Hierarchy<T> &h, int l, const T *xc, T *x, T alpha, hipStream_t s) {
    const Level<T> &L = h.lev[l];
    const Dims cd = dims_of(h.lev[l + 1]);
    if (l == 0) {
        if (h.block == 0) hipLaunchKernelGGL((k_prolong<T, FineA<T>>), g1(L.n1), dim3(256), 0, s, fineA<T>(f), cd, xc, x, alpha);
        else hipLaunchKernelGGL((k_prolong<T, FineV<T>>), g1(L.n1), dim3(256), 0, s, fineV<T>(f), cd, xc, x, alpha);
    } else {
        if (h.block == 0) hipLaunchKernelGGL((k_prolong<T, Expl<T, 3>>), g1(L.n1), dim3(256), 0, s, expl<T, 3>(L), cd, xc, x, alpha);
        else hipLaunchKernelGGL((k_prolong<T, Expl<T, 1>>), g1(L.n1), dim3(256), 0, s, expl<T, 1>(L), cd, xc, x, alpha);
    }
}

template <typename T>
void mg_galerkin_from(const Fine &f, Hierarchy<T> &h, int l, hipStream_t s) {
    Level<T> &C = h.lev[l + 1];
    const Dims cd = dims_of(C);
    if (l == 0) {
        if (h.block == 0) hipLaunchKernelGGL((k_galerkin<T, FineA<T>>), g1(C.n1), dim3(256), 0, s, fineA<T>(f), cd, C.cond, C.exc, C.cross, C.dinv);
        else hipLaunchKernelGGL((k_galerkin<T, FineV<T>>), g1(C.n1), dim3(256), 0, s, fineV<T>(f), cd, C.cond, C.exc, C.cross, C.dinv);
    } else {
        const Level<T> &L = h.lev[l];
        if (h.block == 0) hipLaunchKernelGGL((k_galerkin<T, Expl<T, 3>>), g1(C.n1), dim3(256), 0, s, expl<T, 3>(L), cd, C.cond, C.exc, C.cross, C.dinv);
        else hipLaunchKernelGGL((k_galerkin<T, Expl<T, 1>>), g1(C.n1), dim3(256), 0, s, expl<T, 1>(L), cd, C.cond, C.exc, C.cross, C.dinv);
    }
}

static void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Dense FP64 inverse by Gauss-Jordan with partial pivoting (host, n <= ~2000).
static void dense_inverse(std::vector<double> &A, int n, std::vector<double> &inv) {
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
        if (piv != col) {
            for (int c = 0; c < n; ++c) {
                std::swap(A[(size_t)col * n + c], A[(size_t)piv * n + c]);
                std::swap(inv[(size_t)col * n + c], inv[(size_t)piv * n + c]);
            }
        }
        const double d = 1.0 / A[(size_t)col * n + col];
        for (int c = 0; c < n; ++c) { A[(size_t)col * n + c] *= d; inv[(size_t)col * n + c] *= d; }
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

// Build levels 1..nlev-1 and the coarsest dense inverse.  Level storage must
// already be allocated by the caller (sizes from the level dims).
template <typename T>
void mg_build(const Fine &f, Hierarchy<T> &h, hipStream_t s) {
    for (int l = 0; l + 1 < h.nlev; ++l) mg_galerkin_from<T>(f, h, l, s);
    hip_check(hipStreamSynchronize(s), "galerkin");
    // coarsest level to host
    const Level<T> &C = h.lev[h.nlev - 1];
    const int nc = C.ncomp;
    const int64_t n1 = C.n1;
    std::vector<T> cond((size_t)nc * 6 * n1), exc((size_t)nc * n1), dinv((size_t)nc * n1), cross;
    // All copies are ordered on the handle's stream s (never the legacy null
    // stream, which a non-blocking stream does not wait for).
    hip_check(hipMemcpyAsync(cond.data(), C.cond, cond.size() * sizeof(T), hipMemcpyDeviceToHost, s), "d2h");
    hip_check(hipMemcpyAsync(exc.data(), C.exc, exc.size() * sizeof(T), hipMemcpyDeviceToHost, s), "d2h");
    hip_check(hipMemcpyAsync(dinv.data(), C.dinv, dinv.size() * sizeof(T), hipMemcpyDeviceToHost, s), "d2h");
    if (nc == 3) {
        cross.resize((size_t)nc * 8 * n1);
        hip_check(hipMemcpyAsync(cross.data(), C.cross, cross.size() * sizeof(T), hipMemcpyDeviceToHost, s), "d2h");
    }
    hip_check(hipStreamSynchronize(s), "d2h sync");
    std::vector<int> map((size_t)nc * n1, -1), idx;
    for (int64_t q = 0; q < nc * n1; ++q)
        if (dinv[q] != (T)0) { map[q] = (int)idx.size(); idx.push_back((int)q); }
    const int n = (int)idx.size();
    std::vector<double> A((size_t)n * n, 0.0), inv;
    static const int SO[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    static const int XO[3][8][4] = {
        {{1, 0, -1, 0}, {1, 1, -1, 0}, {1, 0, 0, 0}, {1, 1, 0, 0}, {2, 0, 0, -1}, {2, 1, 0, -1}, {2, 0, 0, 0}, {2, 1, 0, 0}},
        {{0, -1, 0, 0}, {0, 0, 0, 0}, {0, -1, 1, 0}, {0, 0, 1, 0}, {2, 0, 0, -1}, {2, 0, 1, -1}, {2, 0, 0, 0}, {2, 0, 1, 0}},
        {{0, -1, 0, 0}, {0, 0, 0, 0}, {0, -1, 0, 1}, {0, 0, 0, 1}, {1, 0, -1, 0}, {1, 0, 0, 0}, {1, 0, -1, 1}, {1, 0, 0, 1}}};
    for (int r = 0; r < n; ++r) {
        const int q = idx[r];
        const int c = (int)(q / n1);
        const int64_t p = q % n1;
        const int i = (int)(p % C.nx), j = (int)((p / C.nx) % C.ny), k = (int)(p / ((int64_t)C.nx * C.ny));
        double diag = (double)exc[q];
        for (int d = 0; d < 6; ++d) {
            const double cv = (double)cond[((size_t)c * 6 + d) * n1 + p];
            if (cv == 0.0) continue;
            diag += cv;
            const int ni = i + SO[d][0], nj = j + SO[d][1], nk = k + SO[d][2];
            const int64_t nq = (int64_t)c * n1 + ni + (int64_t)C.nx * (nj + (int64_t)C.ny * nk);
            if (map[nq] >= 0) A[(size_t)r * n + map[nq]] -= cv;
        }
        A[(size_t)r * n + r] += diag;
        if (nc == 3) {
            for (int sI = 0; sI < 8; ++sI) {
                const double av = (double)cross[((size_t)c * 8 + sI) * n1 + p];
                if (av == 0.0) continue;
                const int tc = XO[c][sI][0];
                const int ni = i + XO[c][sI][1], nj = j + XO[c][sI][2], nk = k + XO[c][sI][3];
                if (ni < 0 || nj < 0 || nk < 0 || ni >= C.nx || nj >= C.ny || nk >= C.nz) continue;
                const int64_t nq = (int64_t)tc * n1 + ni + (int64_t)C.nx * (nj + (int64_t)C.ny * nk);
                if (map[nq] >= 0) A[(size_t)r * n + map[nq]] += av;
            }
        }
    }
    if (n > 0) dense_inverse(A, n, inv);
    h.ncoarse = n;
    hip_check(hipMalloc(&h.coarse_idx, sizeof(int) * (size_t)(n > 0 ? n : 1)), "malloc");
    hip_check(hipMalloc(&h.coarse_inv, sizeof(double) * (size_t)(n > 0 ? (size_t)n * n : 1)), "malloc");
    if (n > 0) {
        // Uploaded on s, the stream k_coarse_solve runs on, and waited for
        // before the host vectors go out of scope.  A legacy-stream hipMemcpy
        // from pageable memory may return before the data reaches the device
        // and is not ordered with work on a non-blocking stream.
        hip_check(hipMemcpyAsync(h.coarse_idx, idx.data(), sizeof(int) * n, hipMemcpyHostToDevice, s), "h2d");
        hip_check(hipMemcpyAsync(h.coarse_inv, inv.data(), sizeof(double) * (size_t)n * n, hipMemcpyHostToDevice, s),
                  "h2d");
        hip_check(hipStreamSynchronize(s), "h2d sync");
    }
}

// One V-cycle: x0 <- MG(b0).  x0/b0 hold [ncomp][n1] of level 0.
template <typename T>
static void vcycle_rec(const Fine &f, Hierarchy<T> &h, const MGParams &mp, int l, T *x, const T *b, hipStream_t s) {
    Level<T> &L = h.lev[l];
    const int64_t nv = (int64_t)L.ncomp * L.n1;
    if (l == h.nlev - 1) {
        hipLaunchKernelGGL(k_zero<T>, g1(nv), dim3(256), 0, s, x, nv);
        if (h.ncoarse > 0)
            hipLaunchKernelGGL(k_coarse_solve<T>, dim3(1), dim3(256), 0, s, h.ncoarse, h.coarse_idx, h.coarse_inv, b, x);
        return;
    }
    const T w = (T)mp.omega;
    // pre-smoothing, ping-pong between x and L.x (scratch)
    T *cur = x, *oth = L.t;
    // arrange so that the last pre-smoothing sweep lands in x
    if ((mp.nu % 2) == 0) { cur = L.t; oth = x; }
    smooth<T>(f, h, l, nullptr, b, cur, w, 1, s);
    for (int it = 1; it < mp.nu; ++it) {
        smooth<T>(f, h, l, cur, b, oth, w, 0, s);
        T *t = cur; cur = oth; oth = t;
    }
    // cur == x here
    resid<T>(f, h, l, x, b, L.r, s);
    Level<T> &C = h.lev[l + 1];
    hipLaunchKernelGGL(k_restrict<T>, g1(C.n1), dim3(256), 0, s, dims_of(L), dims_of(C), L.ncomp,
                       (const T *)L.r, C.b);
    vcycle_rec<T>(f, h, mp, l + 1, C.x, C.b, s);
    prolong<T>(f, h, l, C.x, x, (T)mp.alpha, s);
    cur = x; oth = L.t;
    for (int it = 0; it < mp.nu; ++it) {
        smooth<T>(f, h, l, cur, b, oth, w, 0, s);
        T *t = cur; cur = oth; oth = t;
    }
    if (cur != x) hipMemcpyAsync(x, cur, sizeof(T) * nv, hipMemcpyDeviceToDevice, s);
}

template <typename T>
void mg_vcycle(const Fine &f, Hierarchy<T> &h, const MGParams &mp, T *x0, const T *b0, hipStream_t s) {
    vcycle_rec<T>(f, h, mp, 0, x0, b0, s);
}

template <typename T>
void mg_fine_sweep(const Fine &f, Hierarchy<T> &h, const MGParams &mp, const T *x, const T *b, T *xo, hipStream_t s) {
    smooth<T>(f, h, 0, x, b, xo, (T)mp.omega, 0, s);
}

// explicit instantiations
template void mg_fine_sweep<float>(const Fine &, Hierarchy<float> &, const MGParams &, const float *, const float *, float *, hipStream_t);
template void mg_fine_sweep<double>(const Fine &, Hierarchy<double> &, const MGParams &, const double *, const double *, double *, hipStream_t);
template void mg_build<float>(const Fine &, Hierarchy<float> &, hipStream_t);
template void mg_build<double>(const Fine &, Hierarchy<double> &, hipStream_t);
template void mg_vcycle<float>(const Fine &, Hierarchy<float> &, const MGParams &, float *, const float *, hipStream_t);
template void mg_vcycle<double>(const Fine &, Hierarchy<double> &, const MGParams &, double *, const double *, hipStream_t);

// ---------------------------------------------------------------------------
// Vector conversion kernels used at the Krylov/MG boundary.
// ---------------------------------------------------------------------------
template <typename TD>
__global__ void k_d2t(const double *__restrict__ src, const double *__restrict__ scale, TD *__restrict__ dst, int64_t n) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) dst[p] = (TD)(scale ? src[p] * scale[p] : src[p]);
}
template <typename TS>
__global__ void k_t2d(const TS *__restrict__ src, double *__restrict__ dst, int64_t n) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) dst[p] = (double)src[p];
}
template <typename TT>
__global__ void k_sub(TT *__restrict__ a, const TT *__restrict__ b, int64_t n) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) a[p] -= b[p];
}
void launch_d2t(const double *src, const double *scale, float *dst, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_d2t<float>, g1(n), dim3(256), 0, s, src, scale, dst, n);
}
void launch_d2t(const double *src, const double *scale, double *dst, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_d2t<double>, g1(n), dim3(256), 0, s, src, scale, dst, n);
}
void launch_t2d(const float *src, double *dst, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_t2d<float>, g1(n), dim3(256), 0, s, src, dst, n);
}
void launch_t2d(const double *src, double *dst, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_t2d<double>, g1(n), dim3(256), 0, s, src, dst, n);
}
void launch_sub(float *a, const float *b, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_sub<float>, g1(n), dim3(256), 0, s, a, b, n);
}
void launch_sub(double *a, const double *b, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_sub<double>, g1(n), dim3(256), 0, s, a, b, n);
}

}  // namespace qms
