/*
 * ns3d_oracle.h -- CPU FP64 restatement of NS3DPeriodic's time-stepping hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path (ns3dperiodic_amd/csrc) and the "port" CPU baseline timed beside it.
 * Nothing in the product path links, loads or calls it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Layout follows the reference (Source/Macros.h:11,15): physical arrays are
 * [Nx][Ny][Nz] doubles, spectral arrays are [Nx][Ny][Mz] interleaved complex
 * doubles with Mz = Nz/2+1, z fastest.
 *
 * Every entry point names the reference function it restates.  Pinning (what
 * checks each function independently of this restatement) is listed in the
 * header of ns3d_oracle.c.
 */
#ifndef NS3D_ORACLE_H
#define NS3D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int Nx, Ny, Nz;        /* physical grid; Mz = Nz/2+1 */
    double Lx, Ly, Lz;     /* box lengths */
    double Re, dt;         /* Reynolds number (nu = 1/Re), time step */
    int ipq;               /* 1: zero Re(Q^) and Q^(0) after every nonlinear term (reference, advection_2_3.cu:35-70) */
    int weno_zfix;         /* 0: keep the l-1 stencil typo of flux_z_weno_R (advection_WENO.cu:288); 1: use l-2 */
    int nthreads;          /* OpenMP threads (<=0: library default) */
    int mask_nyq;          /* 1: 2/3 mask relative to each axis's Nyquist wavenumber (not the reference) */
} or_params;

/* ---- FFT layer (math_support.cu:301-334) ---- */
void or_fft_c2c_1d(int n, int sign, const double *in, double *out);                   /* unnormalised, in/out interleaved complex */
void or_fft_r2c(const or_params *p, const double *in, double *out);                   /* FFTN_Device: D2Z, unnormalised */
void or_fft_c2r(const or_params *p, const double *in, double *out, int scale);        /* iFFTN_Device: Z2D (+ x 1/(NxNyNz) if scale) */

/* ---- host tables (math_support.cu:350-519), arrays sized as the reference sizes them ---- */
void or_tables(const or_params *p, double *kx, double *ky, double *kz,
               double *din_diffusion, double *din_poisson,
               double *AM11, double *AM22, double *AM33, double *AM12, double *AM13, double *AM23,
               double *mask);

/* ---- nonlinear terms (advection_2_3.cu:480-525) ---- */
void or_nonlinear(const or_params *p, const double *ux, const double *uy, const double *uz,
                  double *Qx, double *Qy, double *Qz);
void or_bilinear(const or_params *p, const double *ux, const double *uy, const double *uz,
                 const double *vx, const double *vy, const double *vz,
                 double *Qx, double *Qy, double *Qz);

/* ---- projection / diffusion / update (diffusion.cu:96-365) ---- */
enum { OR_UPD_STAGE = 0,   /* solve_advection_diffusion_projection_device          :96-142  */
       OR_UPD_UV_RHS = 1,  /* solve_advection_diffusion_projection_device_UV_RHS   :144-206 */
       OR_UPD_UV = 2,      /* solve_advection_diffusion_projection_device_UV       :209-259 */
       OR_UPD_IUV = 3,     /* solve_advection_implicit_diffusion_projection_device_UV :264-314 */
       OR_UPD_RHS = 4 };   /* RHS_advection_diffusion_projection_device           :319-365 */
void or_update(const or_params *p, int mode,
               double *ux, double *uy, double *uz,
               const double *fx, const double *fy, const double *fz,
               const double *Qx, const double *Qy, const double *Qz,
               double *Rx, double *Ry, double *Rz);

/* ---- time integrators (RK_time_step.cu) ---- */
void or_single_time_step(const or_params *p, double *ux, double *uy, double *uz,
                         const double *fx, const double *fy, const double *fz);          /* :248-263 */
void or_return_rhs(const or_params *p, const double *ux, const double *uy, const double *uz,
                   const double *fx, const double *fy, const double *fz,
                   double *Rx, double *Ry, double *Rz);                                 /* :231-245 */
void or_rk3_ssp(const or_params *p, double *ux, double *uy, double *uz,
                const double *fx, const double *fy, const double *fz);                  /* :310-327 */
void or_rk3_ssp_uv_rhs(const or_params *p, const double *ux, const double *uy, const double *uz,
                       double *vx, double *vy, double *vz);                            /* :330-338 */
void or_single_time_step_iuv_rhs(const or_params *p, const double *ux, const double *uy, const double *uz,
                                 double *vx, double *vy, double *vz);                  /* :296-306 */
void or_rk3_ssp_uv(const or_params *p, double *ux, double *uy, double *uz,
                   double *vx, double *vy, double *vz,
                   const double *fx, const double *fy, const double *fz);              /* :342-355 */

/* ---- divergence (divergence.cu:8-41) ---- */
void or_divergence(const or_params *p, const double *ux, const double *uy, const double *uz, double *div);

/* ---- WENO advection in physical space (advection_WENO.cu:304-392) ---- */
void or_weno_advect(const or_params *p, int scheme, const double *ux, const double *uy, const double *uz,
                    const double *x, double *out);
void or_weno_advection_spectral(const or_params *p, int scheme,
                                const double *ux_hat, const double *uy_hat, const double *uz_hat,
                                double *Qx_hat, double *Qy_hat, double *Qz_hat);        /* WenoAdvection :397-415 */

void or_rk3_ssp_scalar(const or_params *p, double *ux, double *uy, double *uz,
                       const double *fx, const double *fy, const double *fz, double *theta, int scheme);

/* ---- reductions (min_max_reduction.cu) ---- */
void or_reduce_minmax(long n, const double *in, double *out2);    /* compute_reduction :199-227 ; out2[0]=min, out2[1]=max */
double or_reduce_sum(long n, const double *in);                   /* reduction_sum :479-509 */

/* ---- diagnostics that feed the reductions ---- */
double or_kinetic_energy(const or_params *p, const double *ux, const double *uy, const double *uz);  /* math_support.cu:720-751,799-805 */
double or_dissipation(const or_params *p, const double *ux, const double *uy, const double *uz);     /* math_support.cu:755-796,808-816 */
double or_cfl_dt(const or_params *p, double cfl, const double *ux, const double *uy, const double *uz); /* cuda_supp.cu:170-238 */
double or_shapiro_error(const or_params *p, double t, const double *ux, const double *uy, const double *uz); /* Shapiro_test.cu:8-50 */

/* ---- Krylov vector packing (NS3D_periodic.cu:173-250) ---- */
void or_unpack_reduced(const or_params *p, const double *vec, double *vx, double *vy, double *vz);
void or_pack_reduced(const or_params *p, const double *vx, const double *vy, const double *vz, double *vec);
void or_jv_reduced(const or_params *p, const double *ux, const double *uy, const double *uz,
                   const double *vec_in, double *vec_out);                             /* NSCallMatrixVector_reduced :340-361 */
void or_unpack_full(const or_params *p, const double *vec, double *vx, double *vy, double *vz);       /* :111-125 */
void or_pack_full(const or_params *p, const double *vx, const double *vy, const double *vz, double *vec); /* :223-236 */
void or_jv_reduced_newton(const or_params *p, const double *ux, const double *uy, const double *uz,
                          double tau, const double *vec_in, double *vec_out);          /* :364-389 */
void or_jv_exponential(const or_params *p, double *ux, double *uy, double *uz,
                       const double *fx, const double *fy, const double *fz, int timesteps_period,
                       const double *vec_in, double *vec_out);                         /* :444-470 */

/* Test-only fault injection: plant mistake m (see ns3d_oracle.c), 0 = none. */
void or_set_mutation(int m);

#ifdef __cplusplus
}
#endif
#endif
