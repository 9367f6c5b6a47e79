// k_krylov.hpp -- FP64 vector kernels of the Krylov drivers (Arnoldi with
// DGKS re-orthogonalisation, implicit restarts, BiCGStab(L), Newton): the
// device BLAS-1/2 of Source/Arnoldi/cuda_supp.cu, written for gfx950.
//
// A basis V holds m vectors of length n, vector j at V + j*ldv.  All
// reductions use a fixed partition (KV_BLOCKS blocks, grid-stride, wave64
// butterflies, then one block over the partials in block order), so results
// are bitwise reproducible run to run.
#pragma once
#include "ns3d_common.hpp"

namespace ns3d {

constexpr int KV_THREADS = 256;
constexpr int KV_BLOCKS = 512;
constexpr int KV_COLS = 8;        // basis vectors per pass of the transposed product
constexpr int KV_MAXM = 128;      // largest basis handled by value

struct KvCoef { double h[KV_MAXM]; };

__device__ __forceinline__ double kv_wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block sum of one value per thread; thread 0 receives the result.
__device__ __forceinline__ double kv_block_sum(double v, double* sh) {
    v = kv_wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < KV_THREADS / 64; i++) r += sh[i];
    return r;
}

// partial[c*gridDim.x + b] = sum over block b's elements of V_{c0+c} . w (c < ncols)
__global__ void __launch_bounds__(KV_THREADS)
k_kv_gemv_t_partial(long n, int ncols, const double* __restrict__ V, long ldv, int c0, const double* __restrict__ w,
                    double* __restrict__ partial) {
    __shared__ double sh[KV_THREADS / 64];
    double acc[KV_COLS];
#pragma unroll
    for (int c = 0; c < KV_COLS; c++) acc[c] = 0.0;
    const long stride = (long)gridDim.x * KV_THREADS;
    for (long i = (long)blockIdx.x * KV_THREADS + threadIdx.x; i < n; i += stride) {
        const double wi = w[i];
#pragma unroll
        for (int c = 0; c < KV_COLS; c++)
            if (c < ncols) acc[c] += V[(long)(c0 + c) * ldv + i] * wi;
    }
#pragma unroll
    for (int c = 0; c < KV_COLS; c++) {
        if (c >= ncols) break;
        const double s = kv_block_sum(acc[c], sh);
        if (threadIdx.x == 0) partial[(long)c * gridDim.x + blockIdx.x] = s;
    }
}

// out[c] = sum_b partial[c*nb + b], one block per output, fixed order
__global__ void __launch_bounds__(KV_THREADS)
k_kv_final(const double* __restrict__ partial, int nb, double* __restrict__ out) {
    __shared__ double sh[KV_THREADS / 64];
    const int c = blockIdx.x;
    double a = 0.0;
    for (int b = threadIdx.x; b < nb; b += KV_THREADS) a += partial[(long)c * nb + b];
    const double s = kv_block_sum(a, sh);
    if (threadIdx.x == 0) out[c] = s;
}

// y = alpha * sum_j h[j] V_j + beta * y
__global__ void __launch_bounds__(KV_THREADS)
k_kv_gemv(long n, int m, const double* __restrict__ V, long ldv, KvCoef h, double alpha, double beta,
          double* __restrict__ y) {
    const long i = (long)blockIdx.x * KV_THREADS + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int j = 0; j < m; j++) s += V[(long)j * ldv + i] * h.h[j];
    y[i] = beta == 0.0 ? alpha * s : alpha * s + beta * y[i];
}

// W_c = sum_j Q[j + c*m] V_j for c < k (W and V distinct)
__global__ void __launch_bounds__(KV_THREADS)
k_kv_gemm(long n, int m, int k, const double* __restrict__ V, long ldv, const double* __restrict__ Q,
          double* __restrict__ W, long ldw) {
    const long i = (long)blockIdx.x * KV_THREADS + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= n || c >= k) return;
    double s = 0.0;
    for (int j = 0; j < m; j++) s += V[(long)j * ldv + i] * Q[j + (long)c * m];
    W[(long)c * ldw + i] = s;
}

// y = a*x + b*y
__global__ void __launch_bounds__(KV_THREADS)
k_kv_axpby(long n, double a, const double* __restrict__ x, double b, double* __restrict__ y) {
    const long i = (long)blockIdx.x * KV_THREADS + threadIdx.x;
    if (i >= n) return;
    y[i] = b == 0.0 ? a * x[i] : a * x[i] + b * y[i];
}

}  // namespace ns3d
