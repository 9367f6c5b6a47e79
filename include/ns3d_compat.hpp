// ns3d_compat.hpp -- the reference's hot-path entry-point names over the C ABI.
//
// For C++ host code written against NS3DPeriodic's headers (paths relative to
// the reference root): Source/RK_time_step.h, Source/advection_2_3.h,
// Source/advection_WENO.h, Source/diffusion.h, Source/divergence.h,
// Source/min_max_reduction.h, Source/math_support.h.  The same function names
// and argument meaning are kept; what the ABI makes unnecessary is dropped:
//   * launch geometry (dim3 dimGrid, dimBlock, dimGrid_C, dimBlock_C),
//   * the spectral tables (kx_nabla_d, din_diffusion_d, AM_*_d, mask) -- the
//     kernels recompute them from the context's box,
//   * grid sizes, dx/dy/dz and the RK scratch arrays -- held by the context.
// dt and Re stay explicit, as in the reference.  Calls run on the bound
// context's stream; status codes are returned instead of exit().
//
// Usage:  ns3d_ctx* c; ns3d_create(&cfg, &c); ns3d_compat::bind(c);
//         ns3d_compat::RK3_SSP(dt, Re, ux, uy, uz, fx, fy, fz);
#pragma once
#include "ns3d.h"

namespace ns3d_compat {

inline ns3d_ctx*& bound() {
    static ns3d_ctx* c = nullptr;
    return c;
}
inline void bind(ns3d_ctx* c) { bound() = c; }

typedef ns3d_cplx cudaComplex_t;  // layout of cufftDoubleComplex

inline int set_dt_Re(double dt, double Re) { return ns3d_set_params(bound(), dt, Re); }

// ---- Source/math_support.h:15-25 (FFT wrappers)
inline int FFTN_Device(const double* source, ns3d_cplx* destination) { return ns3d_fft_r2c(bound(), source, destination); }
inline int iFFTN_Device(const ns3d_cplx* source, double* destination) {
    return ns3d_fft_c2r_unscaled(bound(), source, destination);
}
inline int iFFTN_Device_scaled(const ns3d_cplx* source, double* destination) {  // iFFTN_Device(dimGrid, dimBlock, ...)
    return ns3d_fft_c2r(bound(), source, destination);
}

// ---- Source/advection_2_3.h:42-44
inline int calculate_convolution_2p3(const ns3d_cplx* ux, const ns3d_cplx* uy, const ns3d_cplx* uz, ns3d_cplx* Qx,
                                     ns3d_cplx* Qy, ns3d_cplx* Qz) {
    return ns3d_nonlinear(bound(), ux, uy, uz, Qx, Qy, Qz);
}
inline int calculate_convolution_2p3_UV(const ns3d_cplx* ux, const ns3d_cplx* uy, const ns3d_cplx* uz,
                                        const ns3d_cplx* vx, const ns3d_cplx* vy, const ns3d_cplx* vz, ns3d_cplx* Qx,
                                        ns3d_cplx* Qy, ns3d_cplx* Qz) {
    return ns3d_bilinear(bound(), ux, uy, uz, vx, vy, vz, Qx, Qy, Qz);
}

// ---- Source/diffusion.h:12-24
inline int solve_advection_diffusion_projection(ns3d_cplx* ux, ns3d_cplx* uy, ns3d_cplx* uz, const ns3d_cplx* fx,
                                                const ns3d_cplx* fy, const ns3d_cplx* fz, double Re, double dt,
                                                const ns3d_cplx* Qx, const ns3d_cplx* Qy, const ns3d_cplx* Qz) {
    int s = set_dt_Re(dt, Re);
    return s ? s : ns3d_project_update(bound(), NS3D_UPD_STAGE, ux, uy, uz, fx, fy, fz, Qx, Qy, Qz, 0, 0, 0);
}
inline int solve_advection_diffusion_projection_UV_RHS(ns3d_cplx* vx, ns3d_cplx* vy, ns3d_cplx* vz, double Re, double dt,
                                                       const ns3d_cplx* Qx, const ns3d_cplx* Qy, const ns3d_cplx* Qz) {
    int s = set_dt_Re(dt, Re);
    return s ? s : ns3d_project_update(bound(), NS3D_UPD_UV_RHS, vx, vy, vz, 0, 0, 0, Qx, Qy, Qz, 0, 0, 0);
}
inline int solve_advection_diffusion_projection_UV(ns3d_cplx* vx, ns3d_cplx* vy, ns3d_cplx* vz, double Re, double dt,
                                                   const ns3d_cplx* Qx, const ns3d_cplx* Qy, const ns3d_cplx* Qz) {
    int s = set_dt_Re(dt, Re);
    return s ? s : ns3d_project_update(bound(), NS3D_UPD_UV, vx, vy, vz, 0, 0, 0, Qx, Qy, Qz, 0, 0, 0);
}
inline int solve_advection_implicit_diffusion_projection_UV(ns3d_cplx* vx, ns3d_cplx* vy, ns3d_cplx* vz, double Re,
                                                            double dt, const ns3d_cplx* Qx, const ns3d_cplx* Qy,
                                                            const ns3d_cplx* Qz) {
    int s = set_dt_Re(dt, Re);
    return s ? s : ns3d_project_update(bound(), NS3D_UPD_IUV, vx, vy, vz, 0, 0, 0, Qx, Qy, Qz, 0, 0, 0);
}
inline int RHS_advection_diffusion_projection(const ns3d_cplx* ux, const ns3d_cplx* uy, const ns3d_cplx* uz,
                                              const ns3d_cplx* fx, const ns3d_cplx* fy, const ns3d_cplx* fz, double Re,
                                              const ns3d_cplx* Qx, const ns3d_cplx* Qy, const ns3d_cplx* Qz,
                                              ns3d_cplx* RHSx, ns3d_cplx* RHSy, ns3d_cplx* RHSz) {
    int s = ns3d_set_params(bound(), 0.0, Re);
    return s ? s
             : ns3d_project_update(bound(), NS3D_UPD_RHS, (ns3d_cplx*)ux, (ns3d_cplx*)uy, (ns3d_cplx*)uz, fx, fy, fz, Qx,
                                   Qy, Qz, RHSx, RHSy, RHSz);
}

// ---- Source/RK_time_step.h:17-34
inline int copy_arrays(const ns3d_cplx* s1, const ns3d_cplx* s2, const ns3d_cplx* s3, ns3d_cplx* d1, ns3d_cplx* d2,
                       ns3d_cplx* d3) {
    return ns3d_copy3(bound(), s1, s2, s3, d1, d2, d3);
}
inline int single_RK_step(const ns3d_cplx* a1, const ns3d_cplx* a2, const ns3d_cplx* a3, double w0, const ns3d_cplx* b1,
                          const ns3d_cplx* b2, const ns3d_cplx* b3, double w1, ns3d_cplx* d1, ns3d_cplx* d2, ns3d_cplx* d3) {
    return ns3d_axpby3(bound(), a1, a2, a3, w0, b1, b2, b3, w1, d1, d2, d3);
}
inline int single_time_step(double dt, double Re, ns3d_cplx* ux, ns3d_cplx* uy, ns3d_cplx* uz, const ns3d_cplx* fx,
                            const ns3d_cplx* fy, const ns3d_cplx* fz) {
    int s = set_dt_Re(dt, Re);
    return s ? s : ns3d_stage(bound(), ux, uy, uz, fx, fy, fz);
}
inline int return_RHS(double Re, const ns3d_cplx* ux, const ns3d_cplx* uy, const ns3d_cplx* uz, const ns3d_cplx* fx,
                      const ns3d_cplx* fy, const ns3d_cplx* fz, ns3d_cplx* RHSx, ns3d_cplx* RHSy, ns3d_cplx* RHSz) {
    int s = ns3d_set_params(bound(), 0.0, Re);
    return s ? s : ns3d_rhs(bound(), ux, uy, uz, fx, fy, fz, RHSx, RHSy, RHSz);
}
inline int RK3_SSP(double dt, double Re, ns3d_cplx* ux, ns3d_cplx* uy, ns3d_cplx* uz, const ns3d_cplx* fx,
                   const ns3d_cplx* fy, const ns3d_cplx* fz) {
    int s = set_dt_Re(dt, Re);
    return s ? s : ns3d_rk3_ssp(bound(), ux, uy, uz, fx, fy, fz);
}
inline int RK3_SSP_UV_RHS(double dt, double Re, const ns3d_cplx* ux, const ns3d_cplx* uy, const ns3d_cplx* uz,
                          ns3d_cplx* vx, ns3d_cplx* vy, ns3d_cplx* vz) {
    int s = set_dt_Re(dt, Re);
    return s ? s : ns3d_rk3_ssp_uv_rhs(bound(), ux, uy, uz, vx, vy, vz);
}
inline int single_time_step_UV(double dt, double Re, const ns3d_cplx* ux, const ns3d_cplx* uy, const ns3d_cplx* uz,
                               ns3d_cplx* vx, ns3d_cplx* vy, ns3d_cplx* vz) {
    int s = set_dt_Re(dt, Re);
    return s ? s : ns3d_stage_uv(bound(), ux, uy, uz, vx, vy, vz);
}
inline int single_time_step_iUV_RHS(double dt, double Re, const ns3d_cplx* ux, const ns3d_cplx* uy,
                                    const ns3d_cplx* uz, ns3d_cplx* vx, ns3d_cplx* vy, ns3d_cplx* vz) {
    int s = set_dt_Re(dt, Re);
    return s ? s : ns3d_stage_iuv_rhs(bound(), ux, uy, uz, vx, vy, vz);
}
inline int RK3_SSP_UV(double dt, double Re, ns3d_cplx* ux, ns3d_cplx* uy, ns3d_cplx* uz, ns3d_cplx* vx, ns3d_cplx* vy,
                      ns3d_cplx* vz, const ns3d_cplx* fx, const ns3d_cplx* fy, const ns3d_cplx* fz) {
    int s = set_dt_Re(dt, Re);
    return s ? s : ns3d_rk3_ssp_uv(bound(), ux, uy, uz, vx, vy, vz, fx, fy, fz);
}

// ---- Source/divergence.h:13-15 (dt from the bound context)
inline int divergence_device(ns3d_cplx* div_hat, const ns3d_cplx* ux, const ns3d_cplx* uy, const ns3d_cplx* uz) {
    return ns3d_divergence(bound(), ux, uy, uz, div_hat);
}

// ---- Source/advection_WENO.h:19-23
inline int WenoAdvection(int scheme, const ns3d_cplx* ux, const ns3d_cplx* uy, const ns3d_cplx* uz, ns3d_cplx* Qx,
                         ns3d_cplx* Qy, ns3d_cplx* Qz) {
    return ns3d_weno_advection(bound(), scheme, ux, uy, uz, Qx, Qy, Qz);
}

// ---- Source/min_max_reduction.h:33-37 (results on the host)
inline int compute_reduction(const double* d_in, double* out2_host, long long num_els) {
    return ns3d_reduce_minmax(bound(), num_els, d_in, out2_host);
}
inline double reduction_sum(long long N, const double* InputV) {
    double r = 0.0;
    ns3d_reduce_sum(bound(), N, InputV, &r);
    return r;
}

// ---- Source/Arnoldi/Matrix_Vector_emulator.h:11 callbacks (user_map_vector)
typedef void (*user_map_vector)(void* matctx, void* src, void* dst);
inline user_map_vector NSCallMatrixVector_reduced() { return &ns3d_jv_reduced; }
inline user_map_vector NSCallMatrixVector_reduced_Newton_Globalization() { return &ns3d_jv_reduced_newton; }
inline user_map_vector NSCallMatrixVector_exponential() { return &ns3d_jv_exponential; }

}  // namespace ns3d_compat
