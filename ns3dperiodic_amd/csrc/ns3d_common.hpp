// ns3d_common.hpp -- shared device types and spectral-table helpers for the
// gfx950 NS3DPeriodic hot path.
//
// Layout (reference Source/Macros.h:11,15): spectral fields are [Nx][Ny][Mz]
// interleaved complex, Mz = Nz/2+1, z fastest; physical fields [Nx][Ny][Nz].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ns3d {

template <typename T>
struct alignas(2 * sizeof(T)) cplx {
    T x, y;
};

template <typename T> __device__ __forceinline__ cplx<T> mk(T a, T b) { cplx<T> r; r.x = a; r.y = b; return r; }
template <typename T> __device__ __forceinline__ cplx<T> operator+(cplx<T> a, cplx<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T> __device__ __forceinline__ cplx<T> operator-(cplx<T> a, cplx<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T> __device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) { return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
template <typename T> __device__ __forceinline__ cplx<T> cscale(cplx<T> a, T s) { return mk<T>(a.x * s, a.y * s); }
template <typename T> __device__ __forceinline__ cplx<T> conjg(cplx<T> a) { return mk<T>(a.x, -a.y); }
// multiply by (sgn * i)
template <int SGN, typename T> __device__ __forceinline__ cplx<T> mul_i(cplx<T> a) {
    return SGN > 0 ? mk<T>(-a.y, a.x) : mk<T>(a.y, -a.x);
}
template <typename TO, typename TI> __device__ __forceinline__ cplx<TO> cvt(cplx<TI> a) { return mk<TO>((TO)a.x, (TO)a.y); }

// Buffer-resource (MUBUF) loads.  A raw buffer resource covers `bytes` bytes
// from `base`; a load whose byte offset is at or beyond that range returns
// zeros without a memory access.  An element that must read as zero (masked,
// truncated, a padding pencil) is loaded at offset BUF_OOB instead of being
// zero-filled and skipped by a branch, and offsets are 32-bit.  Word 3 of the
// resource (0x00020000) sets a 32-bit data format: with a zero format the
// range-checked loads of this GPU family return zeros.
constexpr int BUF_OOB = (int)0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, bytes, 0x00020000);
}
template <typename T> __device__ __forceinline__ cplx<T> buf_load(__amdgpu_buffer_rsrc_t r, int off);
template <> __device__ __forceinline__ cplx<double> buf_load<double>(__amdgpu_buffer_rsrc_t r, int off) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    cplx<double> a;
    __builtin_memcpy(&a, &v, sizeof(a));
    return a;
}
template <> __device__ __forceinline__ cplx<float> buf_load<float>(__amdgpu_buffer_rsrc_t r, int off) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
    cplx<float> a;
    __builtin_memcpy(&a, &v, sizeof(a));
    return a;
}

// Per-context spectral constants.  Per-axis tables are built on the host with
// the same floating-point expressions as the reference's host builders
// (Source/math_support.cu:374-519) so that on-the-fly recomposition on the
// device reproduces the reference tables exactly.
struct SpecTables {
    const double* kx;       // [Nx]  2*pi/Lx*m          (build_Nabla_Wavenumbers :393-414)
    const double* ky;       // [Ny]
    const double* kz;       // [Mz]
    const double* sx;       // [Nx]  kx*kx   (din_diffusion = sx+sy+sz, :374-390)
    const double* sy;       // [Ny]
    const double* sz;       // [Mz]
    const double* px;       // [Nx]  k1*k1*m*m (projection denominators, :418-466)
    const double* py;       // [Ny]
    const double* pz;       // [Mz]
    const double* fm;       // [Nx]  (double)m       (AM off-diagonal numerators)
    const double* fn;       // [Ny]  (double)n
    const double* fq;       // [Mz]  (double)q
    const double* mx;       // [Nx]  kx^2/kxMax^2  (build_mask_matrix :471-519, with Nz := Mz)
    const double* my;       // [Ny]
    const double* mz;       // [Mz]
    double k1k2, k1k3, k2k3;  // products of the base wavenumbers 2*pi/L
    double mask_smax;       // largest s with sqrt(s) < 2/3 (host libm sqrt)
    int lmax;               // largest l whose z-plane keeps any mode (mz[l] <= mask_smax)
};

// 2/3 mask of the reference (Q2): sqrt(mx+my+mz) < 2/3 and M(0)=0.
__device__ __forceinline__ bool mask_keep(const SpecTables& t, int j, int k, int l) {
    const double s = (t.mx[j] + t.my[k]) + t.mz[l];
    return (s <= t.mask_smax) && ((j | k | l) != 0);
}

// A whole x-pencil (k,l) is masked out iff its j=0 member is (mx >= 0, rounding monotone).
__device__ __forceinline__ bool column_kept(const SpecTables& t, int k, int l) {
    return (t.my[k] + t.mz[l]) <= t.mask_smax;
}

// XCD-aware block order: the dispatcher hands workgroup b to XCD b mod 8, so
// logical tile (b mod 8)*(G/8) + b/8 keeps neighbouring tiles (which share
// the 128-B lines their rows straddle) on one XCD and its L2.  G = gridDim.x
// is a multiple of 8; ids >= nb are padding.
__device__ __forceinline__ int xcd_tile() {
    const int b = blockIdx.x, g = gridDim.x;
    return (b & 7) * (g >> 3) + (b >> 3);
}

template <typename T, int N>
__device__ __forceinline__ void stage_twiddles(cplx<T>* __restrict__ dst, const cplx<T>* __restrict__ src) {
    for (int i = threadIdx.x; i < N; i += blockDim.x) dst[i] = src[i];
    __syncthreads();
}

// Reference AM_ab and din_diffusion at (j,k,l).  EXACT: the six divisions of
// the reference's host tables (Source/math_support.cu:418-466), bit for bit
// (ns3d_tables); otherwise one reciprocal of the denominator times each
// numerator (within an ulp of the table values; the fused epilogues).
struct ProjCoef {
    double a11, a22, a33, a12, a13, a23, din;
};
template <bool EXACT = false>
__device__ __forceinline__ ProjCoef proj_coef(const SpecTables& t, int j, int k, int l) {
    ProjCoef c;
    double den = (t.px[j] + t.py[k]) + t.pz[l];
    const bool zero = ((j | k | l) == 0);
    if (zero) den = 1.0;
    if constexpr (EXACT) {
        c.a11 = t.px[j] / den;
        c.a22 = t.py[k] / den;
        c.a33 = t.pz[l] / den;
        c.a12 = ((t.k1k2 * t.fm[j]) * t.fn[k]) / den;
        c.a13 = ((t.k1k3 * t.fm[j]) * t.fq[l]) / den;
        c.a23 = ((t.k2k3 * t.fn[k]) * t.fq[l]) / den;
    } else {
        const double inv = 1.0 / den;
        c.a11 = t.px[j] * inv;
        c.a22 = t.py[k] * inv;
        c.a33 = t.pz[l] * inv;
        c.a12 = ((t.k1k2 * t.fm[j]) * t.fn[k]) * inv;
        c.a13 = ((t.k1k3 * t.fm[j]) * t.fq[l]) * inv;
        c.a23 = ((t.k2k3 * t.fn[k]) * t.fq[l]) * inv;
    }
    if (zero) { c.a11 = c.a22 = c.a33 = c.a12 = c.a13 = c.a23 = 0.0; }
    c.din = (t.sx[j] + t.sy[k]) + t.sz[l];
    return c;
}

}  // namespace ns3d
