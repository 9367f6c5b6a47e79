// k_misc.hpp -- standalone transforms, pointwise operators, WENO, reductions.
#pragma once
#include "k_fused.hpp"

namespace ns3d {

// ------------------------------------------------------- z C2R / R2C (single fields)
// Two neighbouring (x,y) pencils share one complex FFT: Y = X_p0 + i X_p1.
template <typename TB, int NZ, int P>
__global__ void __launch_bounds__((P * ZPlanT<TB, NZ>::TPP))
k_z_c2r(const cplx<TB>* __restrict__ in, TB* __restrict__ out, long npairs, int Mz, int nb, SpecTables tab,
        const cplx<TB>* __restrict__ twg, TB scale) {
    using Z = ZPencil<TB, NZ, P>;
    constexpr int TPP = Z::TPP;
    __shared__ cplx<TB> lds[P * (NZ + Z::PAD)];
    __shared__ cplx<TB> tw[NZ];
    const int bid = xcd_tile();
    if (bid >= nb) return;
    stage_twiddles<TB, NZ>(tw, twg);
    const int tid = threadIdx.x, t = tid % TPP, p = tid / TPP;
    const long pair = (long)bid * P + p;
    const bool valid = pair < npairs;
    const long p0 = valid ? 2 * pair : 0, p1 = p0 + 1;
    TB x1[Z::NPT], x2[Z::NPT];
    Z::c2r_pair(x1, x2, in + p0 * Mz, false, in + p1 * Mz, false, valid, tab, lds, tw, t, p, scale);
    if (!valid) return;
#pragma unroll
    for (int q = 0; q < Z::FI::BL; q++) {
#pragma unroll
        for (int r = 0; r < Z::FI::RL; r++) {
            const int n = Z::FI::out_index(t, q, r);
            out[p0 * NZ + n] = x1[q * Z::FI::RL + r];
            out[p1 * NZ + n] = x2[q * Z::FI::RL + r];
        }
    }
}

template <typename TB, int NZ, int P>
__global__ void __launch_bounds__((P * ZPlanT<TB, NZ>::TPP))
k_z_r2c(const TB* __restrict__ in, cplx<TB>* __restrict__ out, long npairs, int Mz, int nb,
        const cplx<TB>* __restrict__ twg) {
    using Z = ZPencil<TB, NZ, P>;
    constexpr int TPP = Z::TPP;
    __shared__ cplx<TB> lds[P * (NZ + Z::PAD)];
    __shared__ cplx<TB> tw[NZ];
    const int bid = xcd_tile();
    if (bid >= nb) return;
    stage_twiddles<TB, NZ>(tw, twg);
    const int tid = threadIdx.x, t = tid % TPP, p = tid / TPP;
    const long pair = (long)bid * P + p;
    const bool valid = pair < npairs;
    const long p0 = valid ? 2 * pair : 0, p1 = p0 + 1;
    TB q1[Z::NPT], q2[Z::NPT];
#pragma unroll
    for (int q = 0; q < Z::FF::B0; q++) {
#pragma unroll
        for (int r = 0; r < Z::FF::R0; r++) {
            const int n = Z::FF::in_index(t, q, r);
            q1[q * Z::FF::R0 + r] = valid ? in[p0 * NZ + n] : (TB)0;
            q2[q * Z::FF::R0 + r] = valid ? in[p1 * NZ + n] : (TB)0;
        }
    }
    Z::r2c_pair(q1, q2, out + p0 * Mz, out + p1 * Mz, valid, lds, tw, t, p);
}

// ------------------------------------------------------- precision conversion
__global__ void k_d2f_c(const cplx<double>* __restrict__ in, cplx<float>* __restrict__ out, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = cvt<float>(in[i]);
}
__global__ void k_f2d_r(const float* __restrict__ in, double* __restrict__ out, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (double)in[i];
}
__global__ void k_d2f_r(const double* __restrict__ in, float* __restrict__ out, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}
__global__ void k_f2d_c(const cplx<float>* __restrict__ in, cplx<double>* __restrict__ out, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = cvt<double>(in[i]);
}
// FP64 field in the reference layout (rows of Mz) <-> FP32 field with rows padded to Mp
__global__ void k_d2f_rows(const cplx<double>* __restrict__ in, cplx<float>* __restrict__ out, long n, int Mz, int Mp) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long row = i / Mz;
    out[row * Mp + (i - row * Mz)] = cvt<float>(in[i]);
}
__global__ void k_f2d_rows(const cplx<float>* __restrict__ in, cplx<double>* __restrict__ out, long n, int Mz, int Mp) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long row = i / Mz;
    out[i] = cvt<double>(in[row * Mp + (i - row * Mz)]);
}

// FP64 field in the reference layout <-> the pair (hi, lo) of FP32 fields with
// rows padded to Mp: hi = fp32(u), lo = fp32(u - hi) (u = hi + lo to ~2^-48)
__global__ void k_split_rows(const cplx<double>* __restrict__ in, cplx<float>* __restrict__ hi, cplx<float>* __restrict__ lo,
                             long n, int Mz, int Mp) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long row = i / Mz, o = row * Mp + (i - row * Mz);
    const cplx<double> u = in[i];
    const cplx<float> h = cvt<float>(u);
    hi[o] = h;
    lo[o] = mk<float>((float)(u.x - (double)h.x), (float)(u.y - (double)h.y));
}
__global__ void k_join_rows(const cplx<float>* __restrict__ hi, const cplx<float>* __restrict__ lo,
                            cplx<double>* __restrict__ out, long n, int Mz, int Mp) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long row = i / Mz, o = row * Mp + (i - row * Mz);
    const cplx<float> h = hi[o], l = lo[o];
    out[i] = mk<double>((double)h.x + (double)l.x, (double)h.y + (double)l.y);
}

// ------------------------------------------------------- a7: copy / axpby on 3 fields
// single_RK_step_device (Source/RK_time_step.cu:82-115): dst = w0*s0 + w1*s1.
__global__ void __launch_bounds__(256)
k_axpby3(PtrArr<const cplx<double>, 3> s0, double w0, PtrArr<const cplx<double>, 3> s1, double w1,
         PtrArr<cplx<double>, 3> d, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const cplx<double> a = s0.p[c][i], b = s1.p[c][i];
        d.p[c][i] = mk<double>(w0 * a.x + w1 * b.x, w0 * a.y + w1 * b.y);
    }
}
// copy_arrays_device (Source/RK_time_step.cu:6-39).
__global__ void __launch_bounds__(256)
k_copy3(PtrArr<const cplx<double>, 3> s, PtrArr<cplx<double>, 3> d, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int c = 0; c < 3; c++) d.p[c][i] = s.p[c][i];
}

// ------------------------------------------------------- a8: divergence
// solve_divergence (Source/divergence.cu:8-30): div^ = i (k . u^) / dt.
__global__ void __launch_bounds__(256)
k_divergence(PtrArr<const cplx<double>, 3> u, cplx<double>* __restrict__ div, long nc, int Ny, int Mz,
             SpecTables tab, double dt) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nc) return;
    const int l = (int)(g % Mz);
    const long t1 = g / Mz;
    const int k = (int)(t1 % Ny), j = (int)(t1 / Ny);
    const cplx<double> a = u.p[0][g], b = u.p[1][g], c = u.p[2][g];
    const double kx = tab.kx[j], ky = tab.ky[k], kz = tab.kz[l];
    div[g] = mk<double>((-a.y * kx - b.y * ky - c.y * kz) / dt, (a.x * kx + b.x * ky + c.x * kz) / dt);
}

// ------------------------------------------------------- a10: Krylov packing
// velocities_from_A_vector_reduced_kernel / A_vector_from_velocities_reduced_kernel
// (Source/NS3D_periodic.cu:173-198, 238-250): vec[c*(N-1)+i] <-> Im v^_c[i+1]; v^_c[0] = 0.
__global__ void __launch_bounds__(256)
k_unpack_reduced(const double* __restrict__ vec, PtrArr<cplx<double>, 3> v, long N) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
#pragma unroll
    for (int c = 0; c < 3; c++) v.p[c][i] = mk<double>(0.0, i == 0 ? 0.0 : vec[c * (N - 1) + i - 1]);
}
__global__ void __launch_bounds__(256)
k_pack_reduced(PtrArr<const cplx<double>, 3> v, double* __restrict__ vec, long N) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N - 1) return;
#pragma unroll
    for (int c = 0; c < 3; c++) vec[c * (N - 1) + i] = v.p[c][i + 1].y;
}
// velocities_from_A_vector_kernel / A_vector_from_velocities_kernel (:110-125, :223-236): full length 3N.
__global__ void __launch_bounds__(256)
k_unpack_full(const double* __restrict__ vec, PtrArr<cplx<double>, 3> v, long N) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
#pragma unroll
    for (int c = 0; c < 3; c++) v.p[c][i] = mk<double>(0.0, vec[c * N + i]);
}
__global__ void __launch_bounds__(256)
k_pack_full(PtrArr<const cplx<double>, 3> v, double* __restrict__ vec, long N) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= 0 || i >= N) return;  // index 0 is left untouched, as in the reference
#pragma unroll
    for (int c = 0; c < 3; c++) vec[c * N + i] = v.p[c][i].y;
}
// A_vector_add_diagonal_kernel (:296-310) and the shifted pack (:253-270).
__global__ void __launch_bounds__(256)
k_pack_reduced_shifted(PtrArr<const cplx<double>, 3> v, const double* __restrict__ vin, double shift,
                       double* __restrict__ vec, long N) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N - 1) return;
#pragma unroll
    for (int c = 0; c < 3; c++) vec[c * (N - 1) + i] = v.p[c][i + 1].y - shift * vin[c * (N - 1) + i];
}

// ------------------------------------------------------- a9: WENO advection
// Reference: Source/advection_WENO.cu:47-101 (weno5/weno3), :105-301 (faces), :304-360 (flux).
// The nonlinear weights alpha_k = c_k / (eps + beta_k)^2 normalised by their
// sum are formed over the common denominator prod_k (eps + beta_k)^2, so a
// face costs one FP64 division instead of six; if those products leave the
// FP64 range (smoothness indicators beyond ~1e70) the per-weight form is used.
__device__ __forceinline__ double weno_combine3(double c0, double c1, double c2, double d0, double d1, double d2,
                                                double q0, double q1, double q2) {
    const double a0 = c0 * (d1 * d2), a1 = c1 * (d0 * d2), a2 = c2 * (d0 * d1);
    const double s = a0 + a1 + a2;
    if (s < 1.0e300 && s > 1.0e-300) return (a0 * q0 + a1 * q1 + a2 * q2) / s;
    const double b0 = c0 / d0, b1 = c1 / d1, b2 = c2 / d2, t = b0 + b1 + b2;
    return b0 / t * q0 + b1 / t * q1 + b2 / t * q2;
}
__device__ __forceinline__ double weno5(double f1, double f2, double f3, double f4, double f5) {
    const double eps = 1.0E-22;
    const double q0 = 1.0 / 3.0 * f1 - 7.0 / 6.0 * f2 + 11.0 / 6.0 * f3;
    const double q1 = -1.0 / 6.0 * f2 + 5.0 / 6.0 * f3 + 1.0 / 3.0 * f4;
    const double q2 = 1.0 / 3.0 * f3 + 5.0 / 6.0 * f4 - 1.0 / 6.0 * f5;
    const double t0 = f1 - 2.0 * f2 + f3, u0 = f1 - 4.0 * f2 + 3.0 * f3;
    const double t1 = f2 - 2.0 * f3 + f4, u1 = f2 - f4;
    const double t2 = f3 - 2.0 * f4 + f5, u2 = 3.0 * f3 - 4.0 * f4 + f5;
    const double b0 = 13.0 / 12.0 * t0 * t0 + 0.25 * u0 * u0;
    const double b1 = 13.0 / 12.0 * t1 * t1 + 0.25 * u1 * u1;
    const double b2 = 13.0 / 12.0 * t2 * t2 + 0.25 * u2 * u2;
    return weno_combine3(0.1, 0.6, 0.3, (eps + b0) * (eps + b0), (eps + b1) * (eps + b1), (eps + b2) * (eps + b2),
                         q0, q1, q2);
}
__device__ __forceinline__ double weno3(double f1, double f2, double f3) {
    const double eps = 1.0E-22;
    const double q0 = -0.5 * f1 + 1.5 * f2, q1 = 0.5 * f2 + 0.5 * f3;
    const double b0 = (f2 - f1) * (f2 - f1), b1 = (f3 - f2) * (f3 - f2);
    const double d0 = (eps + b0) * (eps + b0), d1 = (eps + b1) * (eps + b1);
    const double a0 = (1.0 / 3.0) * d1, a1 = (2.0 / 3.0) * d0, s = a0 + a1;
    if (s < 1.0e300 && s > 1.0e-300) return (a0 * q0 + a1 * q1) / s;
    const double c0 = (1.0 / 3.0) / d0, c1 = (2.0 / 3.0) / d1, t = c0 + c1;
    return c0 / t * q0 + c1 / t * q1;
}

struct Grid3 {
    int Nx, Ny, Nz;
    __device__ __forceinline__ long at(int j, int k, int l) const {
        j = j >= Nx ? j - Nx : (j < 0 ? j + Nx : j);
        k = k >= Ny ? k - Ny : (k < 0 ? k + Ny : k);
        l = l >= Nz ? l - Nz : (l < 0 ? l + Nz : l);
        return ((long)j * Ny + k) * Nz + l;
    }
};

// The two faces of a cell along one axis, upwinded by the face velocities
// Um (face -1/2) and Up (face +1/2): flux_{x,y,z}_weno_{L,R}
// (Source/advection_WENO.cu:105-301).  w[o] is the scalar at offset o along
// the axis (|o| <= 3 for WENO5, 2 for WENO3, 1 for the first-order scheme).
// Each face takes one reconstruction, its stencil selected by the upwind
// side, so diverging lanes run no second branch.  q9: the z right-biased
// reconstruction reads offset -1 where -2 belongs (Source/advection_WENO.cu:288).
template <int SCHEME>
__device__ __forceinline__ void weno_faces(const double* w, double Um, double Up, bool q9, double& sm, double& sp) {
    const bool lm = Um >= 0.0, lp = Up >= 0.0;
    if constexpr (SCHEME == 5) {
        sm = weno5(lm ? w[-3] : w[2], lm ? w[-2] : w[1], lm ? w[-1] : w[0], lm ? w[0] : w[-1],
                   lm ? w[1] : (q9 ? w[-1] : w[-2]));
        sp = weno5(lp ? w[-2] : w[3], lp ? w[-1] : w[2], lp ? w[0] : w[1], lp ? w[1] : w[0],
                   lp ? w[2] : (q9 ? w[0] : w[-1]));
    } else if constexpr (SCHEME == 3) {
        sm = weno3(lm ? w[-2] : w[1], lm ? w[-1] : w[0], lm ? w[0] : w[-1]);
        sp = weno3(lp ? w[-1] : w[2], lp ? w[0] : w[1], lp ? w[1] : w[0]);
    } else {
        sm = lm ? w[-1] : w[0];
        sp = lp ? w[0] : w[1];
    }
}

__device__ __forceinline__ int wrap_add(int i, int o, int n) {
    const int v = i + o;
    return v < 0 ? v + n : (v >= n ? v - n : v);
}

// (u.grad) x at (j,k,l) with upwind WENO faces chosen by the sign of the
// face-averaged velocity: flux_device + advection (Source/advection_WENO.cu:304-392).
template <int SCHEME, typename TU>
__device__ __forceinline__ double weno_point(const TU* __restrict__ ux, const TU* __restrict__ uy,
                                             const TU* __restrict__ uz, const double* __restrict__ x, const Grid3& G,
                                             long id, int j, int k, int l, double rdx, double rdy, double rdz, int zfix) {
    constexpr int R = SCHEME == 5 ? 3 : (SCHEME == 3 ? 2 : 1);
    const long sx = (long)G.Ny * G.Nz, sy = G.Nz;
    const long bx = id - (long)j * sx, by = id - (long)k * sy, bz = id - l;   // id with j, k or l removed
    long ox[2 * R + 1], oy[2 * R + 1], oz[2 * R + 1];
#pragma unroll
    for (int o = -R; o <= R; o++) {
        ox[o + R] = bx + wrap_add(j, o, G.Nx) * sx;
        oy[o + R] = by + (long)wrap_add(k, o, G.Ny) * sy;
        oz[o + R] = bz + wrap_add(l, o, G.Nz);
    }
    const double uc = (double)ux[id], vc = (double)uy[id], wc = (double)uz[id];
    const double Uw = 0.5 * ((double)ux[ox[R - 1]] + uc), Ue = 0.5 * (uc + (double)ux[ox[R + 1]]);
    const double Un = 0.5 * ((double)uy[oy[R - 1]] + vc), Us = 0.5 * (vc + (double)uy[oy[R + 1]]);
    const double Ur = 0.5 * ((double)uz[oz[R - 1]] + wc), Uf = 0.5 * (wc + (double)uz[oz[R + 1]]);
    double wx[2 * R + 1], wy[2 * R + 1], wz[2 * R + 1];
    const double xc = x[id];
#pragma unroll
    for (int o = 0; o <= 2 * R; o++) {
        wx[o] = o == R ? xc : x[ox[o]];
        wy[o] = o == R ? xc : x[oy[o]];
        wz[o] = o == R ? xc : x[oz[o]];
    }
    double sw, se, sn, ss, sr, sf;
    weno_faces<SCHEME>(wx + R, Uw, Ue, false, sw, se);
    weno_faces<SCHEME>(wy + R, Un, Us, false, sn, ss);
    weno_faces<SCHEME>(wz + R, Ur, Uf, !zfix, sr, sf);
    return uc * (se - sw) * rdx + vc * (ss - sn) * rdy + wc * (sf - sr) * rdz;
}

// Point index of this thread: blocks are remapped so that each XCD (block b
// runs on XCD b mod 8) sweeps one contiguous run of x-planes, keeping the
// +-3-plane stencil reuse inside one L2.  The grid is a multiple of 8 blocks.
__device__ __forceinline__ long weno_index() { return (long)xcd_tile() * blockDim.x + threadIdx.x; }

// advection kernel + flux_device (Source/advection_WENO.cu:304-392), generalised to any advected scalar x.
template <int SCHEME>
__global__ void __launch_bounds__(256)
k_weno(const double* __restrict__ ux, const double* __restrict__ uy, const double* __restrict__ uz,
       const double* __restrict__ x, double* __restrict__ out, Grid3 G, double rdx, double rdy, double rdz, int zfix) {
    const long id = weno_index();
    const long n = (long)G.Nx * G.Ny * G.Nz;
    if (id >= n) return;
    const int l = (int)(id % G.Nz);
    const long t1 = id / G.Nz;
    const int k = (int)(t1 % G.Ny), j = (int)(t1 / G.Ny);
    out[id] = weno_point<SCHEME>(ux, uy, uz, x, G, id, j, k, l, rdx, rdy, rdz, zfix);
}

// SSP stage of a passive scalar: out = a0*th_n + a1*(th_s - dt * (u.grad)_WENO th_s),
// u being the masked physical velocity of the same stage (TU = bulk precision).
template <int SCHEME, typename TU>
__global__ void __launch_bounds__(256)
k_weno_stage(const TU* __restrict__ u, long nr, const double* __restrict__ x, const double* __restrict__ xn,
             double* __restrict__ out, Grid3 G, double rdx, double rdy, double rdz, int zfix, double dt,
             double a0, double a1, int* nf) {
    const long id = weno_index();
    if (id >= nr) return;
    const int l = (int)(id % G.Nz);
    const long t1 = id / G.Nz;
    const int k = (int)(t1 % G.Ny), j = (int)(t1 / G.Ny);
    const double adv = weno_point<SCHEME>(u, u + nr, u + 2 * nr, x, G, id, j, k, l, rdx, rdy, rdz, zfix);
    const double e = x[id] - dt * adv;
    const double o = xn ? a0 * xn[id] + a1 * e : a1 * e;
    out[id] = o;
    flag_nonfinite(nf, !isfinite(o));
}

__global__ void k_set_int(int* __restrict__ p, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = v;
}

// ------------------------------------------------------- a11: reductions
// Wave64 butterfly reductions via cross-lane shuffles; block partials, then
// a single-block finish.  Sums are deterministic (fixed partition and order).
constexpr int RED_THREADS = 256;
constexpr int RED_BLOCKS = 1024;
// The reference seeds min/max with the integer macro inf = 0x7f800000 (Q8).
constexpr double REF_INF = 2139095040.0;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { const double w = __shfl_xor(v, o, 64); v = w < v ? w : v; }
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { const double w = __shfl_xor(v, o, 64); v = w > v ? w : v; }
    return v;
}

template <int OP>  // 0 sum, 1 min+max
__device__ __forceinline__ void block_finish(double a, double b, double* __restrict__ pa, double* __restrict__ pb) {
    __shared__ double sa[RED_THREADS / 64], sb[RED_THREADS / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (OP == 0) a = wave_sum(a);
    else { a = wave_min(a); b = wave_max(b); }
    if (lane == 0) { sa[wv] = a; sb[wv] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double x = sa[0], y = sb[0];
        for (int i = 1; i < RED_THREADS / 64; i++) {
            if (OP == 0) x += sa[i];
            else { x = sa[i] < x ? sa[i] : x; y = sb[i] > y ? sb[i] : y; }
        }
        *pa = x;
        if (OP == 1) *pb = y;
    }
}

// Pointwise producers feeding the reductions (value at flat physical index id).
// SRC_SPEC_*: over the nc half-spectrum modes of three spectral fields (a, b,
// c read as interleaved complex), each mode weighted by its Hermitian
// multiplicity (1 on the l = 0 and l = Nz/2 planes, else 2).
// SRC_NONFINITE: 1 for each NaN or Inf among n doubles, else 0 (check_nans_kernel,
// Source/cuda_supp.cu:242-325, which tests isnan only; Inf is counted here too).
enum RedSrc {
    SRC_ARRAY = 0, SRC_ENERGY = 1, SRC_DISSIPATION = 2, SRC_CFL = 3, SRC_SHAPIRO = 4,
    SRC_SPEC_ENERGY = 5, SRC_SPEC_DISSIPATION = 6, SRC_NONFINITE = 7
};
struct RedArgs {
    const double* a;
    const double* b;
    const double* c;
    Grid3 G;
    double dx, dy, dz, p0, p1;  // p0: CFL number or time; p1: Re
    const double* kx;           // wavenumber tables (spectral sources)
    const double* ky;
    const double* kz;
    int Mz;
};

template <int SRC>
__device__ __forceinline__ double red_value(const RedArgs& r, long id) {
    if (SRC == SRC_ARRAY) return r.a[id];
    if (SRC == SRC_NONFINITE) return isfinite(r.a[id]) ? 0.0 : 1.0;
    if (SRC == SRC_SPEC_ENERGY || SRC == SRC_SPEC_DISSIPATION) {
        const int l = (int)(id % r.Mz);
        const long t1 = id / r.Mz;
        const int k = (int)(t1 % r.G.Ny), j = (int)(t1 / r.G.Ny);
        const double w = (l == 0 || 2 * l == r.G.Nz) ? 1.0 : 2.0;
        const double ar = r.a[2 * id], ai = r.a[2 * id + 1];
        const double br = r.b[2 * id], bi = r.b[2 * id + 1];
        const double cr = r.c[2 * id], ci = r.c[2 * id + 1];
        const double e = (ar * ar + ai * ai) + (br * br + bi * bi) + (cr * cr + ci * ci);
        if (SRC == SRC_SPEC_ENERGY) return w * e;
        // |grad u|^2 + (div u)^2 = 2 S:S integrated: |k|^2 |u^|^2 + |k.u^|^2
        const double kx = r.kx[j], ky = r.ky[k], kz = r.kz[l];
        const double dr = kx * ar + ky * br + kz * cr, di = kx * ai + ky * bi + kz * ci;
        return w * (((kx * kx + ky * ky) + kz * kz) * e + (dr * dr + di * di));
    }
    const int l = (int)(id % r.G.Nz);
    const long t1 = id / r.G.Nz;
    const int k = (int)(t1 % r.G.Ny), j = (int)(t1 / r.G.Ny);
    if (SRC == SRC_ENERGY) {  // get_kinetic_energy_device (Source/math_support.cu:744)
        const double uy = r.b[id], uz = r.c[id];
        return 0.5 * (uy * uy + uz * uz) * r.dx * r.dy;
    }
    if (SRC == SRC_DISSIPATION) {  // get_dissipation_device (Source/math_support.cu:777-789)
        const Grid3& G = r.G;
        auto D = [&](const double* f, int ax) {
            if (ax == 0) return (f[G.at(j + 1, k, l)] - f[G.at(j - 1, k, l)]) / r.dx;
            if (ax == 1) return (f[G.at(j, k + 1, l)] - f[G.at(j, k - 1, l)]) / r.dy;
            return (f[G.at(j, k, l + 1)] - f[G.at(j, k, l - 1)]) / r.dz;
        };
        const double ux_x = D(r.a, 0), ux_y = D(r.a, 1), ux_z = D(r.a, 2);
        const double uy_x = D(r.b, 0), uy_y = D(r.b, 1), uy_z = D(r.b, 2);
        const double uz_x = D(r.c, 0), uz_y = D(r.c, 1), uz_z = D(r.c, 2);
        return r.dx * r.dy * r.dz *
               (2.0 * (ux_x * ux_x + uy_y * uy_y + uz_z * uz_z) + (uy_z + uz_y) * (uy_z + uz_y) +
                (ux_z + uz_x) * (ux_z + uz_x) + (ux_y + uy_x) * (ux_y + uy_x));
    }
    if (SRC == SRC_CFL) {  // calculate_CFL_device (Source/cuda_supp.cu:188)
        return r.p0 / ((1.0e-6 + fabs(r.a[id])) / r.dx + (1.0e-6 + fabs(r.b[id])) / r.dy + (1.0e-6 + fabs(r.c[id])) / r.dz);
    }
    // SRC_SHAPIRO: Shapiro_test_case_device (Source/Shapiro_test.cu:26-33)
    const double y = k * r.dy, x = j * r.dx, z = l * r.dz;
    const double f = sqrt(3.0);
    const double dec = exp(-(f * f) * r.p0 / r.p1);
    const double ex = r.a[id] - (-0.5 * (f * cos(x) * sin(y) * sin(z) + sin(x) * cos(y) * cos(z)) * dec);
    const double ey = r.b[id] - (0.5 * (f * sin(x) * cos(y) * sin(z) - cos(x) * sin(y) * cos(z)) * dec);
    const double ez = r.c[id] - (cos(x) * cos(y) * sin(z) * dec);
    return sqrt(ex * ex + ey * ey + ez * ez);
}

template <int OP, int SRC>
__global__ void __launch_bounds__(RED_THREADS)
k_reduce_partial(RedArgs r, long n, double* __restrict__ part) {
    double a = OP == 0 ? 0.0 : REF_INF, b = -REF_INF;
    const long stride = (long)gridDim.x * RED_THREADS;
    for (long id = (long)blockIdx.x * RED_THREADS + threadIdx.x; id < n; id += stride) {
        const double v = red_value<SRC>(r, id);
        if (OP == 0) a += v;
        else { a = v < a ? v : a; b = v > b ? v : b; }
    }
    block_finish<OP>(a, b, part + blockIdx.x, part + gridDim.x + blockIdx.x);
}

template <int OP>
__global__ void __launch_bounds__(RED_THREADS)
k_reduce_final(const double* __restrict__ part, int nb, double* __restrict__ out) {
    double a = OP == 0 ? 0.0 : REF_INF, b = -REF_INF;
    for (int i = threadIdx.x; i < nb; i += RED_THREADS) {
        if (OP == 0) a += part[i];
        else { const double x = part[i], y = part[nb + i]; a = x < a ? x : a; b = y > b ? y : b; }
    }
    block_finish<OP>(a, b, out, out + 1);
}

}  // namespace ns3d
