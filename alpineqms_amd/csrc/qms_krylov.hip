// FP64 vector kernels of the Krylov solver.  Reductions are two-stage with a
// fixed block count (no atomics), so every result is bitwise reproducible
// for a given vector length: a design gives the same iterates alone or in a
// batch.
#include <hip/hip_runtime.h>
#include "qms_kernels.h"

namespace qms {

static constexpr int RB = 256;          // threads per reduction block
static constexpr int CH = 8;            // vectors per multi-dot chunk

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

// Block-wide sum of CH accumulators; thread 0 ends with the totals in acc.
__device__ __forceinline__ void block_sum(double acc[CH], double *sh /* [RB/64][CH] */) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < CH; ++c) acc[c] = wave_sum(acc[c]);
    if (lane == 0)
        for (int c = 0; c < CH; ++c) sh[wid * CH + c] = acc[c];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int c = 0; c < CH; ++c) {
            double s = 0.0;
            for (int w = 0; w < RB / 64; ++w) s += sh[w * CH + c];
            acc[c] = s;
        }
    }
}

// partial[i*nblk + blk] = sum over this block's elements of V_i . w, i in [c0, c0+CH)
__global__ void __launch_bounds__(RB) k_multidot(const double *const *__restrict__ V, int nv, int c0,
                                                 const double *__restrict__ w, int64_t n,
                                                 double *__restrict__ partial) {
    __shared__ double sh[(RB / 64) * CH];
    double acc[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) acc[c] = 0.0;
    const int cn = min(CH, nv - c0);
    const double *vp[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) vp[c] = (c < cn) ? V[c0 + c] : nullptr;
    for (int64_t e = (int64_t)blockIdx.x * RB + threadIdx.x; e < n; e += (int64_t)gridDim.x * RB) {
        const double we = w[e];
#pragma unroll
        for (int c = 0; c < CH; ++c)
            if (c < cn) acc[c] += vp[c][e] * we;
    }
    block_sum(acc, sh);
    if (threadIdx.x == 0)
        for (int c = 0; c < cn; ++c) partial[(int64_t)(c0 + c) * gridDim.x + blockIdx.x] = acc[c];
}

__global__ void k_reduce_partials(const double *__restrict__ partial, int nblk, double *__restrict__ out) {
    __shared__ double sh[(RB / 64) * CH];
    double acc[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) acc[c] = 0.0;
    const double *p = partial + (int64_t)blockIdx.x * nblk;
    for (int b = threadIdx.x; b < nblk; b += RB) acc[0] += p[b];
    block_sum(acc, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = acc[0];
}

// w -= sum_i h_i V_i; partial[blk] = sum of w^2 over the block
__global__ void __launch_bounds__(RB) k_orth_update(double *const *__restrict__ V, int nv,
                                                    const double *__restrict__ h, double *__restrict__ w, int64_t n,
                                                    double *__restrict__ partial) {
    __shared__ double sh[(RB / 64) * CH];
    double acc[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) acc[c] = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * RB + threadIdx.x; e < n; e += (int64_t)gridDim.x * RB) {
        double we = w[e];
        for (int i = 0; i < nv; ++i) we -= h[i] * V[i][e];
        w[e] = we;
        acc[0] += we * we;
    }
    block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc[0];
}

__global__ void __launch_bounds__(RB) k_norm2(const double *__restrict__ w, int64_t n, double *__restrict__ partial) {
    __shared__ double sh[(RB / 64) * CH];
    double acc[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) acc[c] = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * RB + threadIdx.x; e < n; e += (int64_t)gridDim.x * RB) {
        const double we = w[e];
        acc[0] += we * we;
    }
    block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc[0];
}

// dst = src * a
__global__ void k_scale(double *__restrict__ dst, const double *__restrict__ src, double a, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    dst[e] = src[e] * a;
}

// x += sum_i y_i Z_i
__global__ void k_axpy_list(double *__restrict__ x, const double *const *__restrict__ Z, int nz,
                            const double *__restrict__ y, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double s = x[e];
    for (int i = 0; i < nz; ++i) s += y[i] * Z[i][e];
    x[e] = s;
}

// Plain streaming copy (calibration reference for HBM byte counters).
__global__ void k_copy(double *__restrict__ dst, const double *__restrict__ src, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) dst[e] = src[e];
}
void launch_copy(double *dst, const double *src, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dst, src, n);
}

void launch_multidot(const double *const *V, int nv, const double *w, int64_t n, double *partial, int nblk,
                     hipStream_t s) {
    for (int c0 = 0; c0 < nv; c0 += CH)
        hipLaunchKernelGGL(k_multidot, dim3(nblk), dim3(RB), 0, s, V, nv, c0, w, n, partial);
}
void launch_reduce_partials(const double *partial, int nv, int nblk, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_partials, dim3(nv), dim3(RB), 0, s, partial, nblk, out);
}
void launch_orth_update(double *const *V, int nv, const double *h, double *w, int64_t n, double *partial, int nblk,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_orth_update, dim3(nblk), dim3(RB), 0, s, V, nv, h, w, n, partial);
}
void launch_norm2(const double *w, int64_t n, double *partial, int nblk, hipStream_t s) {
    hipLaunchKernelGGL(k_norm2, dim3(nblk), dim3(RB), 0, s, w, n, partial);
}
void launch_scale(double *dst, const double *src, double a, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_scale, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dst, src, a, n);
}
void launch_axpy_list(double *x, const double *const *Z, int nz, const double *y_dev, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_axpy_list, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, Z, nz, y_dev, n);
}

}  // namespace qms
