// FP64 vector kernels of the Krylov solver.  Reductions are two-stage with a
// fixed block count (no atomics), so every result is bitwise reproducible
// for a given vector length: a design gives the same iterates alone or in a
// batch.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <mutex>
#include <string>
#include "qms_kernels.h"
#include "../../include/qms.h"

namespace qms {

static constexpr int RB = 256;          // threads per reduction block

int reduce_blocks(int64_t n) {
    int64_t b = (n + RB * 8 - 1) / (RB * 8);
    if (b < 1) b = 1;
    if (b > 1024) b = 1024;
    return (int)b;
}

__device__ __forceinline__ double wave_sum(double v) {
    // wave64 butterfly
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block-wide sum of K accumulators; thread 0 ends with the totals in acc.
template <int K>
__device__ __forceinline__ void block_sum(double acc[K], double *sh /* [RB/64][K] */) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    // one accumulator at a time: interleaving K butterflies would hold K sets
    // of temporaries and cut the kernel's occupancy
#pragma unroll
    for (int c = 0; c < K; ++c) {
        acc[c] = wave_sum(acc[c]);
        __builtin_amdgcn_sched_barrier(0);
    }
    if (lane == 0) {
#pragma unroll
        for (int c = 0; c < K; ++c) sh[wid * K + c] = acc[c];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < K; ++c) {
            double s = 0.0;
            for (int w = 0; w < RB / 64; ++w) s += sh[w * K + c];
            acc[c] = s;
        }
    }
}

// Krylov vectors have two segments: na values of type TE from byte 0 and nb
// FP64 values from byte offset offv.  In the solver the first segment holds
// the A components (na = 3 n1) and the second the V component (nb = n1):
// TE = double with offv = 24 n1 is a plain FP64 vector; TE = float is the
// mixed-precision basis (the V component stays FP64: the conductors'
// floating potentials are fixed only through the 1e-3 S/m background, far
// below FP32 resolution of the conductor rows).  Every sum is FP64.
//
// Basis vectors are read through pointers loaded from memory; marking them
// global (rather than generic) lets every vector load be a global load from
// a wave-uniform base.
template <typename TV>
using gcp = const __attribute__((address_space(1))) TV *;
template <typename TV>
using gp = __attribute__((address_space(1))) TV *;

// Element at byte offset off of base: I = uint32_t when every byte offset
// fits 32 bits (then all loads of an element share one 32-bit offset on top of
// wave-uniform base addresses), int64_t otherwise.
template <typename TE, typename P, typename I>
__device__ __forceinline__ TE ld(P base, I off) {
    return *(gcp<TE>)((const __attribute__((address_space(1))) char *)base + off);
}
template <typename TE, typename P, typename I>
__device__ __forceinline__ void st(P base, I off, TE v) {
    *(gp<TE>)((__attribute__((address_space(1))) char *)base + off) = v;
}

// partial[i*nblk + blk] = sum over this block's elements of V_i . w for the
// up-to-K vectors i in [c0, c0 + min(K, nv - c0)).  Every element of w is
// read once per chunk; the K loads of an element are independent and
// unconditional (slots past the chunk re-read vector c0, from cache, and
// their sums are dropped).
template <int K, typename I, typename TE>
__global__ void __launch_bounds__(RB) k_multidot(const TE *const *__restrict__ V, int nv, int c0,
                                                 const TE *__restrict__ w, int64_t na, int64_t nb, int64_t offv,
                                                 double *__restrict__ partial) {
    __shared__ double sh[(RB / 64) * K];
    double acc[K];
    gcp<char> vp[K];
    const int cn = min(K, nv - c0);
#pragma unroll
    for (int c = 0; c < K; ++c) {
        acc[c] = 0.0;
        vp[c] = (gcp<char>)V[c0 + (c < cn ? c : 0)];
    }
    const gcp<char> wg = (gcp<char>)w;
    const I t0 = (I)blockIdx.x * RB + threadIdx.x, nt = (I)gridDim.x * RB;
    {   // A components
        const I end = (I)na * (I)sizeof(TE), step = nt * (I)sizeof(TE);
        for (I off = t0 * (I)sizeof(TE); off < end; off += step) {
            const double we = (double)ld<TE>(wg, off);
#pragma unroll
            for (int c = 0; c < K; ++c) acc[c] = fma((double)ld<TE>(vp[c], off), we, acc[c]);
        }
    }
    {   // V component
        const I end = (I)offv + (I)nb * (I)sizeof(double), step = nt * (I)sizeof(double);
        for (I off = (I)offv + t0 * (I)sizeof(double); off < end; off += step) {
            const double we = ld<double>(wg, off);
#pragma unroll
            for (int c = 0; c < K; ++c) acc[c] = fma(ld<double>(vp[c], off), we, acc[c]);
        }
    }
    block_sum<K>(acc, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < K; ++c)
            if (c < cn) partial[(int64_t)(c0 + c) * gridDim.x + blockIdx.x] = acc[c];
    }
}

__global__ void k_reduce_partials(const double *__restrict__ partial, int nblk, double *__restrict__ out) {
    __shared__ double sh[RB / 64];
    double acc[1] = {0.0};
    const double *p = partial + (int64_t)blockIdx.x * nblk;
    for (int b = threadIdx.x; b < nblk; b += RB) acc[0] += p[b];
    block_sum<1>(acc, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = acc[0];
}

// w -= sum_{i in [c0, c0 + min(K, nv - c0))} h_i V_i, one chunk of up to K
// basis vectors per pass over w (the pointers and coefficients of the chunk
// sit in scalar registers; slots past the chunk re-read vector c0 with a
// zero coefficient).  The pass that ends the basis also writes
// partial[blk] = sum of w^2 over the block (of the stored values).  Chunks
// run in order, so the subtractions happen in basis order whatever the
// chunking.  The update is formed in FP64 and stored in each segment's type.
template <int K, typename I, typename TE>
__global__ void __launch_bounds__(RB) k_orth_update(TE *const *__restrict__ V, int nv, int c0,
                                                    const double *__restrict__ h, TE *__restrict__ w, int64_t na,
                                                    int64_t nb, int64_t offv, double *__restrict__ partial) {
    __shared__ double sh[RB / 64];
    gcp<char> vp[K];
    double hc[K];
    const int cn = min(K, nv - c0);
#pragma unroll
    for (int c = 0; c < K; ++c) {
        vp[c] = (gcp<char>)V[c0 + (c < cn ? c : 0)];
        hc[c] = c < cn ? h[c0 + c] : 0.0;
    }
    const bool last = c0 + cn == nv;
    double acc[1] = {0.0};
    const gp<char> wg = (gp<char>)w;
    const I t0 = (I)blockIdx.x * RB + threadIdx.x, nt = (I)gridDim.x * RB;
    {
        const I end = (I)na * (I)sizeof(TE), step = nt * (I)sizeof(TE);
        for (I off = t0 * (I)sizeof(TE); off < end; off += step) {
            double ve[K];
#pragma unroll
            for (int c = 0; c < K; ++c) ve[c] = (double)ld<TE>(vp[c], off);
            double we = (double)ld<TE>(wg, off);
#pragma unroll
            for (int c = 0; c < K; ++c) we -= hc[c] * ve[c];
            const TE ws = (TE)we;
            st<TE>(wg, off, ws);
            acc[0] += (double)ws * (double)ws;
        }
    }
    {
        const I end = (I)offv + (I)nb * (I)sizeof(double), step = nt * (I)sizeof(double);
        for (I off = (I)offv + t0 * (I)sizeof(double); off < end; off += step) {
            double ve[K];
#pragma unroll
            for (int c = 0; c < K; ++c) ve[c] = ld<double>(vp[c], off);
            double we = ld<double>(wg, off);
#pragma unroll
            for (int c = 0; c < K; ++c) we -= hc[c] * ve[c];
            st<double>(wg, off, we);
            acc[0] += we * we;
        }
    }
    // every launch reduces (one barrier whatever the chunk, so the kernel's
    // synchronisation does not depend on its arguments); the last one writes
    block_sum<1>(acc, sh);
    if (last && threadIdx.x == 0) partial[blockIdx.x] = acc[0];
}

__global__ void __launch_bounds__(RB) k_norm2(const double *__restrict__ w, int64_t n, double *__restrict__ partial) {
    __shared__ double sh[RB / 64];
    double acc[1] = {0.0};
    for (int64_t e = (int64_t)blockIdx.x * RB + threadIdx.x; e < n; e += (int64_t)gridDim.x * RB) {
        const double we = w[e];
        acc[0] += we * we;
    }
    block_sum<1>(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc[0];
}

// dst = src * a for a flat FP64 vector src of 4 n1 values into a two-segment
// vector dst (FP64 product, stored in the segment's type)
template <typename TE>
__global__ void k_scale(TE *__restrict__ dst, const double *__restrict__ src, double a, int64_t n1, int64_t offv) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 4 * n1) return;
    if (e < 3 * n1) dst[e] = (TE)(src[e] * a);
    else ((double *)((char *)dst + offv))[e - 3 * n1] = src[e] * a;
}

// x += sum_i y_i Z_i (Z in the preconditioner's precision; FP64 sums, added
// in basis order).  The loads of eight vectors are issued before their sums,
// so a thread has eight loads in flight instead of one.
template <typename TZ>
__global__ void __launch_bounds__(256) k_axpy_list(double *__restrict__ x, const TZ *const *__restrict__ Z, int nz,
                                                   const double *__restrict__ y, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double s = x[e];
    int i = 0;
    for (; i + 8 <= nz; i += 8) {
        TZ v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = Z[i + u][e];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += y[i + u] * (double)v[u];
    }
    for (; i < nz; ++i) s += y[i] * (double)Z[i][e];
    x[e] = s;
}

// dst = src * a[0] for two-segment vectors (a device scalar, so the launch
// needs no host value)
template <typename TE>
__global__ void __launch_bounds__(256) k_scale_dev(TE *__restrict__ dst, const TE *__restrict__ src,
                                                   const double *__restrict__ a, int64_t n1, int64_t offv) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 4 * n1) return;
    if (e < 3 * n1) {
        dst[e] = (TE)((double)src[e] * a[0]);
    } else {
        const int64_t q = e - 3 * n1;
        ((double *)((char *)dst + offv))[q] = ((const double *)((const char *)src + offv))[q] * a[0];
    }
}

// ---------------------------------------------------------------------------
// FGMRES bookkeeping on the device (one thread): the Hessenberg matrix H
// ((m+1) x m, row-major), the Givens rotations cs/sn and the rotated
// right-hand side g live in device memory, so an iteration needs no host
// round trip.  st holds the cycle's scalars (KS_* below).
// ---------------------------------------------------------------------------
// A new cycle from a residual of norm beta: g = beta e_1, no iteration done.
__global__ void k_cycle_start(int m, double beta, double bnorm, double tol, double floor, double *__restrict__ g,
                              double *__restrict__ st) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int i = 0; i <= m; ++i) g[i] = 0.0;
    g[0] = beta;
    st[KS_BNORM] = bnorm;
    st[KS_TOL] = tol;
    st[KS_INVH] = 0.0;
    st[KS_EST] = bnorm > 0 ? beta / bnorm : beta;
    st[KS_K] = 0.0;
    st[KS_ITERS] = 0.0;
    st[KS_FLOOR] = floor;
    st[KS_STALL] = 0.0;
    for (int i = 0; i < 4; ++i) st[KS_HIST + i] = 0.0;
}

// Iteration j: hcol[0..j] are the Gram-Schmidt coefficients of w against
// V_0..V_j and hcol[j+1] = |w|^2 after the orthogonalisation.  Stores column j
// of H, rotates it, extends the rotated right-hand side, records the residual
// estimate |g[j+1]| / bnorm and, the first time it is below tol/2 (or the
// Krylov space is exhausted), the iteration count k = j+1 of the solution.
// st[KS_INVH] = 1/|w| scales w into V_{j+1}.
__global__ void k_givens(int j, int m, const double *__restrict__ hcol, double *__restrict__ H,
                         double *__restrict__ cs, double *__restrict__ sn, double *__restrict__ g,
                         double *__restrict__ st) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int i = 0; i <= j; ++i) H[(int64_t)i * m + j] = hcol[i];
    const double hn = sqrt(fmax(hcol[j + 1], 0.0));
    for (int i = 0; i < j; ++i) {
        const double a = H[(int64_t)i * m + j], b = H[(int64_t)(i + 1) * m + j];
        H[(int64_t)i * m + j] = cs[i] * a + sn[i] * b;
        H[(int64_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * b;
    }
    const double a = H[(int64_t)j * m + j];
    const double d = hypot(a, hn);
    cs[j] = d > 0 ? a / d : 1.0;
    sn[j] = d > 0 ? hn / d : 0.0;
    H[(int64_t)j * m + j] = d;
    g[j + 1] = -sn[j] * g[j];
    g[j] = cs[j] * g[j];
    const double bnorm = st[KS_BNORM];
    const double est = fabs(g[j + 1]) / (bnorm > 0 ? bnorm : 1.0);
    st[KS_EST] = est;
    st[KS_ITERS] = (double)(j + 1);
    if (st[KS_K] == 0.0 && (est <= 0.5 * st[KS_TOL] || hn == 0.0 || d == 0.0)) st[KS_K] = (double)(j + 1);
    // stagnation below the floor: less than 10% progress over the last two
    // iterations ends the cycle here.  At the residual floor the estimate
    // creeps down by a few percent per iteration while the FP64 residual no
    // longer moves; on a 200-step 128^3 velocity ramp, waiting for less than
    // 20% over four iterations instead changed the fields less than solving
    // to a tenth of the tolerance does.  (Less than 20% over two iterations
    // also stops slow but real progress: a 24^3 steady solve then ended at
    // 9e-11 instead of 2e-12.)
    const double est2 = st[KS_HIST + ((j - 2) & 3)];   // the estimate of iteration j - 2
    if (st[KS_K] == 0.0 && j >= 2 && est <= st[KS_FLOOR] && est > 0.9 * est2) {
        st[KS_K] = (double)(j + 1);
        st[KS_STALL] = 1.0;
    }
    st[KS_HIST + (j & 3)] = est;
    st[KS_INVH] = hn > 0 ? 1.0 / hn : 0.0;
}

// y = R^-1 g over the first k columns (R: the rotated upper-triangular H).
__global__ void k_backsub(int k, int m, const double *__restrict__ H, const double *__restrict__ g,
                          double *__restrict__ y) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int i = k - 1; i >= 0; --i) {
        double t = g[i];
        for (int c = i + 1; c < k; ++c) t -= H[(int64_t)i * m + c] * y[c];
        const double r = H[(int64_t)i * m + i];
        y[i] = r != 0.0 ? t / r : 0.0;
    }
}

// Initial guess of a time step from the solutions of the last steps, and the
// history shifted by one: with x0 = x^n, x1 = x^(n-1), x2 = x^(n-2),
//   nh = 0: x0;  nh = 1: 2 x0 - x1;  nh >= 2: 3 x0 - 3 x1 + x2.
__global__ void __launch_bounds__(256) k_extrapolate(double *__restrict__ x, double *__restrict__ h1,
                                                     double *__restrict__ h2, int nh, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double x0 = x[e], x1 = h1[e], x2 = h2[e];
    const double gss = nh >= 2 ? 3.0 * (x0 - x1) + x2 : nh == 1 ? 2.0 * x0 - x1 : x0;
    h2[e] = x1;
    h1[e] = x0;
    x[e] = gss;
}

// dst = sqrt(max(src, 0)) (a reduced sum of squares to a norm, on the device)
__global__ void k_sqrt_to(const double *__restrict__ src, double *__restrict__ dst) {
    if (threadIdx.x == 0 && blockIdx.x == 0) dst[0] = sqrt(fmax(src[0], 0.0));
}

// Plain streaming copy (calibration reference for HBM byte counters).
__global__ void k_copy(double *__restrict__ dst, const double *__restrict__ src, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) dst[e] = src[e];
}
void launch_copy(double *dst, const double *src, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dst, src, n);
}


// Splits the element range of a Gram-Schmidt pass into na values of type TE
// and nb FP64 values from byte offv; uint32_t byte offsets when they fit.
template <typename TE>
void launch_multidot_seg(const TE *const *V, int nv, const TE *w, int64_t na, int64_t nb, int64_t offv,
                         double *partial, int nblk, hipStream_t s) {
    // chunks of up to 16 vectors; the narrower kernel when the chunk fits it
    const bool o32 = std::max<int64_t>(na * (int64_t)sizeof(TE), offv + 8 * nb) + (int64_t)nblk * RB * 8 <
                     ((int64_t)1 << 32);
    for (int c0 = 0; c0 < nv; c0 += 16) {
        const int cn = nv - c0 < 16 ? nv - c0 : 16;
        if (o32) {
            if (cn <= 8) hipLaunchKernelGGL((k_multidot<8, uint32_t, TE>), dim3(nblk), dim3(RB), 0, s, V, nv, c0, w, na, nb, offv, partial);
            else hipLaunchKernelGGL((k_multidot<16, uint32_t, TE>), dim3(nblk), dim3(RB), 0, s, V, nv, c0, w, na, nb, offv, partial);
        } else {
            if (cn <= 8) hipLaunchKernelGGL((k_multidot<8, int64_t, TE>), dim3(nblk), dim3(RB), 0, s, V, nv, c0, w, na, nb, offv, partial);
            else hipLaunchKernelGGL((k_multidot<16, int64_t, TE>), dim3(nblk), dim3(RB), 0, s, V, nv, c0, w, na, nb, offv, partial);
        }
    }
}
template <typename TE>
void launch_orth_update_seg(TE *const *V, int nv, const double *h, TE *w, int64_t na, int64_t nb, int64_t offv,
                            double *partial, int nblk, hipStream_t s) {
    const bool o32 = std::max<int64_t>(na * (int64_t)sizeof(TE), offv + 8 * nb) + (int64_t)nblk * RB * 8 <
                     ((int64_t)1 << 32);
    for (int c0 = 0; c0 < nv; c0 += 16) {
        const int cn = nv - c0 < 16 ? nv - c0 : 16;
        if (o32) {
            if (cn <= 8) hipLaunchKernelGGL((k_orth_update<8, uint32_t, TE>), dim3(nblk), dim3(RB), 0, s, V, nv, c0, h, w, na, nb, offv, partial);
            else hipLaunchKernelGGL((k_orth_update<16, uint32_t, TE>), dim3(nblk), dim3(RB), 0, s, V, nv, c0, h, w, na, nb, offv, partial);
        } else {
            if (cn <= 8) hipLaunchKernelGGL((k_orth_update<8, int64_t, TE>), dim3(nblk), dim3(RB), 0, s, V, nv, c0, h, w, na, nb, offv, partial);
            else hipLaunchKernelGGL((k_orth_update<16, int64_t, TE>), dim3(nblk), dim3(RB), 0, s, V, nv, c0, h, w, na, nb, offv, partial);
        }
    }
}
template <typename TE>
void launch_multidot(const TE *const *V, int nv, const TE *w, int64_t n1, int64_t offv, double *partial, int nblk,
                     hipStream_t s) {
    launch_multidot_seg<TE>(V, nv, w, 3 * n1, n1, offv, partial, nblk, s);
}
void launch_reduce_partials(const double *partial, int nv, int nblk, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_partials, dim3(nv), dim3(RB), 0, s, partial, nblk, out);
}
template <typename TE>
void launch_orth_update(TE *const *V, int nv, const double *h, TE *w, int64_t n1, int64_t offv, double *partial,
                        int nblk, hipStream_t s) {
    launch_orth_update_seg<TE>(V, nv, h, w, 3 * n1, n1, offv, partial, nblk, s);
}
template void launch_multidot<double>(const double *const *, int, const double *, int64_t, int64_t, double *, int,
                                      hipStream_t);
template void launch_multidot<float>(const float *const *, int, const float *, int64_t, int64_t, double *, int,
                                     hipStream_t);
template void launch_orth_update<double>(double *const *, int, const double *, double *, int64_t, int64_t, double *,
                                         int, hipStream_t);
template void launch_orth_update<float>(float *const *, int, const double *, float *, int64_t, int64_t, double *, int,
                                        hipStream_t);
void launch_norm2(const double *w, int64_t n, double *partial, int nblk, hipStream_t s) {
    hipLaunchKernelGGL(k_norm2, dim3(nblk), dim3(RB), 0, s, w, n, partial);
}
template <typename TE>
void launch_scale(TE *dst, const double *src, double a, int64_t n1, int64_t offv, hipStream_t s) {
    hipLaunchKernelGGL(k_scale<TE>, dim3((unsigned)((4 * n1 + 255) / 256)), dim3(256), 0, s, dst, src, a, n1, offv);
}
template void launch_scale<double>(double *, const double *, double, int64_t, int64_t, hipStream_t);
template void launch_scale<float>(float *, const double *, double, int64_t, int64_t, hipStream_t);
void launch_axpy_list(double *x, const double *const *Z, int nz, const double *y_dev, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_axpy_list<double>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, Z, nz, y_dev, n);
}
void launch_axpy_list(double *x, const float *const *Z, int nz, const double *y_dev, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_axpy_list<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, Z, nz, y_dev, n);
}
template <typename TE>
void launch_scale_dev(TE *dst, const TE *src, const double *a_dev, int64_t n1, int64_t offv, hipStream_t s) {
    hipLaunchKernelGGL(k_scale_dev<TE>, dim3((unsigned)((4 * n1 + 255) / 256)), dim3(256), 0, s, dst, src, a_dev, n1,
                       offv);
}
template void launch_scale_dev<double>(double *, const double *, const double *, int64_t, int64_t, hipStream_t);
template void launch_scale_dev<float>(float *, const float *, const double *, int64_t, int64_t, hipStream_t);
void launch_cycle_start(int m, double beta, double bnorm, double tol, double floor, double *g, double *st,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_cycle_start, dim3(1), dim3(64), 0, s, m, beta, bnorm, tol, floor, g, st);
}
void launch_givens(int j, int m, const double *hcol, double *H, double *cs, double *sn, double *g, double *st,
                   hipStream_t s) {
    hipLaunchKernelGGL(k_givens, dim3(1), dim3(64), 0, s, j, m, hcol, H, cs, sn, g, st);
}
void launch_backsub(int k, int m, const double *H, const double *g, double *y, hipStream_t s) {
    hipLaunchKernelGGL(k_backsub, dim3(1), dim3(64), 0, s, k, m, H, g, y);
}
void launch_extrapolate(double *x, double *h1, double *h2, int nh, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_extrapolate, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, h1, h2, nh, n);
}
void launch_sqrt_to(const double *src, double *dst, hipStream_t s) {
    hipLaunchKernelGGL(k_sqrt_to, dim3(1), dim3(64), 0, s, src, dst);
}

}  // namespace qms

// ---------------------------------------------------------------------------
// Allocation-free Gram-Schmidt step, qms_gram_schmidt_noalloc (declared in
// include/qms.h).  Its kernels are qms_gram_schmidt's; their scratch -- the
// device table of basis pointers and the per-block partial sums -- lives in
// `scratch`, a fixed array in device memory: every GPU the library runs on
// holds its own copy until the process ends.  Host code never passes the
// array's name to a kernel: the host-side symbol is not a device address.
// The device address is looked up once per device (hipGetSymbolAddress, with
// that device current) and kept.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[QMS_NOALLOC_SCRATCH_BYTES];

namespace qms {

// scratch layout: [basis pointer table, 1 KiB][partial sums: (nv+1) x nblk doubles]
constexpr size_t NOALLOC_TAB_BYTES = 1024;
static_assert(QMS_NOALLOC_MAX_VECTORS * sizeof(void *) <= NOALLOC_TAB_BYTES, "pointer table");
static_assert(NOALLOC_TAB_BYTES + (size_t)(QMS_NOALLOC_MAX_VECTORS + 1) * 1024 * sizeof(double) <=
                  QMS_NOALLOC_SCRATCH_BYTES,
              "partial sums of the largest call (reduce_blocks is at most 1024)");

static std::mutex g_scratch_mu;
static void *g_scratch_dev[256];       // device address of `scratch` per device id (null: not looked up yet)

static hipError_t scratch_address(int dev, void **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return hipSuccess;
}

// tab[i] = base + i*n for the nv basis vectors (one workgroup)
__global__ void k_basis_table(const double **__restrict__ tab, const double *base, int64_t n, int nv) {
    const int i = (int)threadIdx.x;
    if (i < nv) tab[i] = base + (int64_t)i * n;
}

}  // namespace qms

extern "C" int qms_gram_schmidt_noalloc(int64_t n, int nv, const double *basis_dev, double *w_dev, double *hcoef_dev,
                                        void *stream) {
    using namespace qms;
    if (n < 1 || nv < 1 || nv > QMS_NOALLOC_MAX_VECTORS || !basis_dev || !w_dev || !hcoef_dev) {
        set_last_error("qms_gram_schmidt_noalloc: need n >= 1, 1 <= nv <= QMS_NOALLOC_MAX_VECTORS and device buffers");
        return QMS_ERR_ARG;
    }
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= 256) {
        set_last_error("qms_gram_schmidt_noalloc: no current device");
        return QMS_ERR_HIP;
    }
    void *base = nullptr;
    e = scratch_address(dev, &base);
    if (e != hipSuccess) {
        set_last_error((std::string("qms_gram_schmidt_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e)).c_str());
        return QMS_ERR_HIP;
    }
    const double **tab = (const double **)base;
    double *partial = (double *)((char *)base + NOALLOC_TAB_BYTES);
    const int nblk = reduce_blocks(n);
    hipStream_t s = (hipStream_t)stream;
    // the same element split as the solver's FP64 basis when n is a multiple
    // of 4 (so the sums are those of qms_gram_schmidt), else one segment
    const int64_t na = n % 4 == 0 ? 3 * (n / 4) : n, nb = n - na, offv = 8 * na;
    hipLaunchKernelGGL(k_basis_table, dim3(1), dim3(64), 0, s, tab, basis_dev, n, nv);
    launch_multidot_seg<double>(tab, nv, w_dev, na, nb, offv, partial, nblk, s);
    launch_reduce_partials(partial, nv, nblk, hcoef_dev, s);
    launch_orth_update_seg<double>((double *const *)tab, nv, hcoef_dev, w_dev, na, nb, offv, partial, nblk, s);
    launch_reduce_partials(partial, 1, nblk, hcoef_dev + nv, s);
    e = hipGetLastError();
    if (e != hipSuccess) {
        set_last_error((std::string("qms_gram_schmidt_noalloc: launch: ") + hipGetErrorString(e)).c_str());
        return QMS_ERR_HIP;
    }
    return QMS_OK;
}
