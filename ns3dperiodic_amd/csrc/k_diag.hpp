// k_diag.hpp -- diagnostics of the reference driver (Source/math_support.cu:712-987):
// vorticity by central differences and the two spectral LES filters.
#pragma once
#include "k_misc.hpp"
#include "k_krylov.hpp"

namespace ns3d {

// get_curl_device (Source/math_support.cu:712-718 kernel body): second-order
// central differences with periodic neighbours.  fix = 0 keeps the reference's
// y-component, which differences uz(j+1) against ux(j-1); fix = 1 uses uz(j-1).
__global__ void __launch_bounds__(256)
k_curl(const double* __restrict__ ux, const double* __restrict__ uy, const double* __restrict__ uz,
       double* __restrict__ rx, double* __restrict__ ry, double* __restrict__ rz, Grid3 G, double dx, double dy,
       double dz, int fix) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)G.Nx * G.Ny * G.Nz;
    if (g >= n) return;
    const int l = (int)(g % G.Nz);
    const long t = g / G.Nz;
    const int k = (int)(t % G.Ny), j = (int)(t / G.Ny);
    rx[g] = 0.5 * (uz[G.at(j, k + 1, l)] - uz[G.at(j, k - 1, l)]) / dy - 0.5 * (uy[G.at(j, k, l + 1)] - uy[G.at(j, k, l - 1)]) / dz;
    const double* xm = fix ? uz : ux;
    ry[g] = 0.5 * (ux[G.at(j, k, l + 1)] - ux[G.at(j, k, l - 1)]) / dz - 0.5 * (uz[G.at(j + 1, k, l)] - xm[G.at(j - 1, k, l)]) / dx;
    rz[g] = 0.5 * (uy[G.at(j + 1, k, l)] - uy[G.at(j - 1, k, l)]) / dx - 0.5 * (ux[G.at(j, k + 1, l)] - ux[G.at(j, k - 1, l)]) / dy;
}

// Helmholz_Fourier_Filter_Device (:911-972 kernel): 1/(1 + eps^2 |k|^2);
// CutOff_Fourier_Filter_Device (:975-987 kernel, called with Nz := Mz): keep
// m^2/Nx^2 + n^2/Ny^2 + q^2/Mz^2 <= r^2.
template <int KIND>
__global__ void __launch_bounds__(256)
k_spectral_filter(PtrArr<const cplx<double>, 3> in, PtrArr<cplx<double>, 3> out, long nc, int Nx, int Ny, int Mz,
                  double Lx, double Ly, double Lz, double p) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nc) return;
    const int l = (int)(g % Mz);
    const long t = g / Mz;
    const int k = (int)(t % Ny), j = (int)(t / Ny);
    const int m = j >= Nx / 2 ? j - Nx : j, n = k >= Ny / 2 ? k - Ny : k, q = l;
    double f;
    if (KIND == 0) {
        const double a = 2.0 * M_PI / Lx * m, b = 2.0 * M_PI / Ly * n, c = 2.0 * M_PI / Lz * q;
        f = 1.0 / (1.0 + p * p * (a * a + b * b + c * c));
    } else {
        f = (1.0 * m * m / (1.0 * Nx * Nx) + 1.0 * n * n / (1.0 * Ny * Ny) + 1.0 * q * q / (1.0 * Mz * Mz)) <= p * p ? 1.0 : 0.0;
    }
#pragma unroll
    for (int b = 0; b < 3; b++) {
        const cplx<double> v = in.p[b][g];
        out.p[b][g] = mk<double>(f * v.x, f * v.y);
    }
}

}  // namespace ns3d

namespace ns3d {

// Physical velocity at grid point (jp, kp, lp) straight from the half
// spectrum: u = (1/(NxNyNz)) sum w_l Re(u^_{jkl} exp(2 pi i (j jp/Nx + k kp/Ny + l lp/Nz))),
// w_l = 1 on the l = 0 and l = Nz/2 planes and 2 elsewhere (the C2R convention of
// the library's transforms).  One pass over the three fields replaces three
// inverse 3D transforms (return_vector3_solution, Source/return_map.cu:110-160).
// blockIdx.y: member of a batch (fields bs elements apart); partial[(3*member
// + c)*gridDim.x + b], c < 3.
__global__ void __launch_bounds__(KV_THREADS)
k_probe_partial(PtrArr<const cplx<double>, 3> u, long nc, int Nx, int Ny, int Nz, int Mz, int jp, int kp, int lp,
                double* __restrict__ partial, long bs = 0) {
    __shared__ double sh[KV_THREADS / 64];
    double acc[3] = {0.0, 0.0, 0.0};
    const long stride = (long)gridDim.x * KV_THREADS;
    for (long g = (long)blockIdx.x * KV_THREADS + threadIdx.x; g < nc; g += stride) {
        const int l = (int)(g % Mz);
        const long t = g / Mz;
        const int k = (int)(t % Ny), j = (int)(t / Ny);
        const long a = ((long)j * jp) % Nx, b = ((long)k * kp) % Ny, c = ((long)l * lp) % Nz;
        double s, co;
        sincospi(2.0 * ((double)a / Nx + (double)b / Ny + (double)c / Nz), &s, &co);
        const double w = (l == 0 || 2 * l == Nz) ? 1.0 : 2.0;
#pragma unroll
        for (int f = 0; f < 3; f++) {
            const cplx<double> v = u.p[f][(long)blockIdx.y * bs + g];
            acc[f] += w * (v.x * co - v.y * s);
        }
    }
#pragma unroll
    for (int f = 0; f < 3; f++) {
        const double r = kv_block_sum(acc[f], sh);
        if (threadIdx.x == 0) partial[((long)blockIdx.y * 3 + f) * gridDim.x + blockIdx.x] = r;
    }
}

}  // namespace ns3d
