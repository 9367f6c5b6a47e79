/*
 * ns3d.h -- C ABI of the MI355X (gfx950) NS3DPeriodic time-stepping hot path.
 *
 * Plain C: opaque context, raw device pointers, sizes and scalars; every call
 * returns an int status (NS3D_OK = 0) except the Krylov callbacks, whose
 * signature is fixed by the reference (void(*)(void*,void*,void*)).
 *
 * Data layout is the reference's (Source/Macros.h:11,15):
 *   physical field  double [Nx][Ny][Nz]
 *   spectral field  ns3d_cplx [Nx][Ny][Mz], Mz = Nz/2+1, z fastest,
 *                   unnormalised forward DFT (sign -1).
 * ns3d_cplx has the layout of cufftDoubleComplex / hipDoubleComplex.
 * All array arguments are device pointers on the context's device unless
 * stated.  Calls are asynchronous on the context's stream unless they return
 * a host value.
 *
 * Reference entry points replaced (paths relative to the reference root):
 *   RK_time_step.h:17-34, advection_2_3.h:34-46, advection_WENO.h:19-23,
 *   diffusion.h:12-24, divergence.h:13-15, min_max_reduction.h:33-37,
 *   math_support.h:15-72, Arnoldi/Matrix_Vector_emulator.h:11.
 */
#ifndef NS3D_H
#define NS3D_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ns3d_ctx ns3d_ctx;
typedef struct { double x, y; } ns3d_cplx;

/* status codes */
#define NS3D_OK 0
#define NS3D_ERR_ARG (-1)          /* null pointer / bad value */
#define NS3D_ERR_SIZE (-2)         /* grid size without a compiled FFT plan */
#define NS3D_ERR_ALLOC (-3)        /* device allocation failed */
#define NS3D_ERR_HIP (-4)          /* HIP runtime error (launch or copy) */
#define NS3D_ERR_NONFINITE (-5)    /* NaN/Inf detected by a checking call */
#define NS3D_ERR_PRECISION (-6)    /* operation not offered at this precision */
#define NS3D_ERR_CAPACITY (-7)     /* grid too large for ns3d_rk3_ssp_steps_noalloc's fixed scratch */

/* precision of the transform / product bulk; state arrays at the interface are always FP64 */
#define NS3D_FP64 0
#define NS3D_FP32_BULK 1   /* FP32 transforms and products; state and epilogue FP64 */
#define NS3D_FP32_STATE 2  /* as FP32_BULK, and the RK3 loops (ns3d_rk3_ssp*, graph) keep the state and
                              stages in FP32, converting the caller's FP64 state once per call */
#define NS3D_FP32_INCR 3   /* as FP32_BULK, and the RK3 loops hold the state u^n as a pair of FP32 fields
                              hi + lo (converted from and to the caller's FP64 state once per call) and carry
                              the SSP stages as FP32 increments over it (u1 = u^n + d1, u2 = u^n + d2); the
                              transforms read hi: FP32 storage throughout, near-FP64 state accuracy */

/* behaviour flags (reference quirks are ON in reference mode) */
#define NS3D_FLAG_NO_IPQ 0x1u      /* disable zeroing Re(Q^) and Q^(0) (advection_2_3.cu:35-70) */
#define NS3D_FLAG_WENO_ZFIX 0x2u   /* use l-2 in flux_z_weno_R (advection_WENO.cu:288 reads l-1) */
#define NS3D_FLAG_MASK_NYQUIST 0x4u /* 2/3 mask relative to each axis's Nyquist wavenumber (alias-free);
                                       the reference normalises by (Nx, Ny, Mz) (math_support.cu:471-519),
                                       which leaves the x and y axes untruncated */
#define NS3D_FLAG_CHECK_FINITE 0x8u /* the RK3 entry points (ns3d_rk3_ssp, _steps, _batch, _graph, _scalar) end
                                       by synchronising the stream and reading (and clearing) the non-finite
                                       flag: NS3D_ERR_NONFINITE if any stage wrote a NaN or Inf.  The reference
                                       checks for NaN in its time loop (NS3D_periodic.cu:1357-1360) */

typedef struct {
    int Nx, Ny, Nz;        /* physical grid, each in {8,16,32,48,64,80,96,128,160,192,256,320,384,512} */
    double Lx, Ly, Lz;     /* box lengths (reference: 2*pi/alpha, 2*pi, 2*pi/beta, NS3D_periodic.cu:969-979) */
    double Re;             /* Reynolds number, nu = 1/Re */
    double dt;             /* time step */
    int precision;         /* NS3D_FP64, NS3D_FP32_BULK or NS3D_FP32_STATE */
    unsigned flags;        /* NS3D_FLAG_* */
    void *stream;          /* hipStream_t; NULL = default stream */
    int device;            /* HIP device ordinal; -1 = current device */
} ns3d_config;

/* ---- context (replaces struct_NS3D_RK3_Call, Structures.h:6-58, and
 *      init_dealiasing/clean_dealiasing, advection_2_3.cu:529-585) ---- */
int ns3d_create(const ns3d_config *cfg, ns3d_ctx **out);
int ns3d_destroy(ns3d_ctx *ctx);
int ns3d_set_params(ns3d_ctx *ctx, double dt, double Re);
int ns3d_set_stream(ns3d_ctx *ctx, void *stream);
int ns3d_synchronize(ns3d_ctx *ctx);
int ns3d_last_error(const ns3d_ctx *ctx);          /* status of the last void-returning callback */
const char *ns3d_status_string(int status);
int ns3d_supported_size(int n);                    /* 1 if an FFT plan for length n is compiled */
long long ns3d_workspace_bytes(const ns3d_ctx *ctx);
/* Count of reallocations of the buffers a captured RK3 step addresses (the fused passes' scratch,
 * the batch stage buffers): a graph captured from this context's steps stays valid while it is
 * unchanged.  Growth of other buffers (J.v unpacking, WENO scratch) does not count.  No reference
 * counterpart. */
long long ns3d_workspace_generation(const ns3d_ctx *ctx);

/* ---- a4: FFT layer (math_support.cu:301-345) ---- */
int ns3d_fft_r2c(ns3d_ctx *ctx, const double *in, ns3d_cplx *out);              /* FFTN_Device */
int ns3d_fft_c2r(ns3d_ctx *ctx, const ns3d_cplx *in, double *out);              /* iFFTN_Device(dimGrid,..) incl. scale_double */
int ns3d_fft_c2r_unscaled(ns3d_ctx *ctx, const ns3d_cplx *in, double *out);     /* iFFTN_Device(source,dest) */

/* ---- a5: spectral tables (math_support.cu:374-519).  The kernels recompute
 *      them on the fly; this writes the reference arrays for callers that
 *      need them.  Any pointer may be NULL.  kx:[Nx] ky:[Ny] kz:[Mz], others [Nx][Ny][Mz]. ---- */
int ns3d_tables(ns3d_ctx *ctx, double *kx, double *ky, double *kz, double *din_diffusion, double *din_poisson,
                double *AM11, double *AM22, double *AM33, double *AM12, double *AM13, double *AM23, double *mask);

/* ---- a3: nonlinear terms ---- */
int ns3d_nonlinear(ns3d_ctx *ctx, const ns3d_cplx *ux, const ns3d_cplx *uy, const ns3d_cplx *uz,
                   ns3d_cplx *Qx, ns3d_cplx *Qy, ns3d_cplx *Qz);                  /* calculate_convolution_2p3 */
int ns3d_bilinear(ns3d_ctx *ctx, const ns3d_cplx *ux, const ns3d_cplx *uy, const ns3d_cplx *uz,
                  const ns3d_cplx *vx, const ns3d_cplx *vy, const ns3d_cplx *vz,
                  ns3d_cplx *Qx, ns3d_cplx *Qy, ns3d_cplx *Qz);                   /* calculate_convolution_2p3_UV */

/* ---- a6: projection / diffusion / update family (diffusion.cu:96-416) ---- */
#define NS3D_UPD_STAGE 0    /* solve_advection_diffusion_projection:      u <- u + dt(AM Q - Q - |k|^2/Re u + f) */
#define NS3D_UPD_UV_RHS 1   /* solve_advection_diffusion_projection_UV_RHS: u <- AM Q - Q - |k|^2/Re u */
#define NS3D_UPD_UV 2       /* solve_advection_diffusion_projection_UV:     u <- u + dt(AM Q - Q - |k|^2/Re u) */
#define NS3D_UPD_IUV 3      /* solve_advection_implicit_diffusion_projection_UV: u <- (u + dt(AM Q - Q))/(1+dt|k|^2/Re) */
#define NS3D_UPD_RHS 4      /* RHS_advection_diffusion_projection:          R <- AM Q - Q - |k|^2/Re u + f */
int ns3d_project_update(ns3d_ctx *ctx, int mode, ns3d_cplx *ux, ns3d_cplx *uy, ns3d_cplx *uz,
                        const ns3d_cplx *fx, const ns3d_cplx *fy, const ns3d_cplx *fz,
                        const ns3d_cplx *Qx, const ns3d_cplx *Qy, const ns3d_cplx *Qz,
                        ns3d_cplx *Rx, ns3d_cplx *Ry, ns3d_cplx *Rz);

/* ---- a7: stage combination kernels (RK_time_step.cu:6-39,82-115) ---- */
int ns3d_copy3(ns3d_ctx *ctx, const ns3d_cplx *sx, const ns3d_cplx *sy, const ns3d_cplx *sz,
               ns3d_cplx *dx, ns3d_cplx *dy, ns3d_cplx *dz);                      /* copy_arrays */
int ns3d_axpby3(ns3d_ctx *ctx, const ns3d_cplx *ax, const ns3d_cplx *ay, const ns3d_cplx *az, double wa,
                const ns3d_cplx *bx, const ns3d_cplx *by, const ns3d_cplx *bz, double wb,
                ns3d_cplx *dx, ns3d_cplx *dy, ns3d_cplx *dz);                     /* single_RK_step */

/* ---- a2/a1: stages and the SSP-RK3 step (RK_time_step.cu:231-327).  Forcing
 *      pointers may all be NULL (zero forcing). ---- */
int ns3d_stage(ns3d_ctx *ctx, ns3d_cplx *ux, ns3d_cplx *uy, ns3d_cplx *uz,
               const ns3d_cplx *fx, const ns3d_cplx *fy, const ns3d_cplx *fz);    /* single_time_step */
int ns3d_rhs(ns3d_ctx *ctx, const ns3d_cplx *ux, const ns3d_cplx *uy, const ns3d_cplx *uz,
             const ns3d_cplx *fx, const ns3d_cplx *fy, const ns3d_cplx *fz,
             ns3d_cplx *Rx, ns3d_cplx *Ry, ns3d_cplx *Rz);                        /* return_RHS */
int ns3d_rk3_ssp(ns3d_ctx *ctx, ns3d_cplx *ux, ns3d_cplx *uy, ns3d_cplx *uz,
                 const ns3d_cplx *fx, const ns3d_cplx *fy, const ns3d_cplx *fz);  /* RK3_SSP */
int ns3d_rk3_ssp_steps(ns3d_ctx *ctx, int nsteps, ns3d_cplx *ux, ns3d_cplx *uy, ns3d_cplx *uz,
                       const ns3d_cplx *fx, const ns3d_cplx *fy, const ns3d_cplx *fz);
/* ns3d_rk3_ssp_steps without any device or host allocation inside the call.  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The fused passes' scratch and the RK stage
 * buffers come from a fixed 64 MiB device array, one per GPU, held until the process ends
 * (its address is looked up by ns3d_create); the context's own workspace is neither used nor
 * grown.  Grids whose scratch exceeds 64 MiB return NS3D_ERR_CAPACITY before any launch:
 * 15 bulk fields + 6 FP64 stage fields (9 FP32 fields in NS3D_FP32_STATE / NS3D_FP32_INCR) of padded rows,
 * i.e. every grid up to 64^3 in every precision (64^3 FP64: 52.5 MiB); 96^3 does not fit.
 * Results are bitwise those of ns3d_rk3_ssp_steps.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (two streams, two threads) race and must be ordered by
 * the caller.  Nothing in the array needs to start at zero: every element a call reads was
 * written earlier in the same call.  Per-pass profiling is off during the call. */
int ns3d_rk3_ssp_steps_noalloc(ns3d_ctx *ctx, int nsteps, ns3d_cplx *ux, ns3d_cplx *uy, ns3d_cplx *uz,
                               const ns3d_cplx *fx, const ns3d_cplx *fy, const ns3d_cplx *fz);

/* ---- a8: divergence (divergence.cu:8-47) ---- */
int ns3d_divergence(ns3d_ctx *ctx, const ns3d_cplx *ux, const ns3d_cplx *uy, const ns3d_cplx *uz,
                    ns3d_cplx *div_hat);                                          /* divergence_device */
int ns3d_divergence_physical(ns3d_ctx *ctx, const ns3d_cplx *ux, const ns3d_cplx *uy, const ns3d_cplx *uz,
                             ns3d_cplx *div_hat, double *div);                    /* + devergence_to_double */

/* ---- a9: WENO advection (advection_WENO.cu:304-415).  scheme 3 or 5;
 *      any other value gives the first-order upwind value, as in the reference. ---- */
int ns3d_weno_advect(ns3d_ctx *ctx, int scheme, const double *ux, const double *uy, const double *uz,
                     const double *x, double *out);                               /* advection kernel, scalar x */
int ns3d_weno_advection(ns3d_ctx *ctx, int scheme, const ns3d_cplx *ux, const ns3d_cplx *uy, const ns3d_cplx *uz,
                        ns3d_cplx *Qx, ns3d_cplx *Qy, ns3d_cplx *Qz);             /* WenoAdvection */

/* nsteps of RK3_SSP replayed from a HIP graph captured on the first call (and again
 * whenever the pointers, nsteps, dt or Re change).  Needs a created stream (not the
 * null stream).  Same results as ns3d_rk3_ssp_steps; one launch instead of 15*nsteps. */
int ns3d_rk3_ssp_graph(ns3d_ctx *ctx, int nsteps, ns3d_cplx *ux, ns3d_cplx *uy, ns3d_cplx *uz,
                       const ns3d_cplx *fx, const ns3d_cplx *fy, const ns3d_cplx *fz);

/* nsteps of RK3_SSP for nbatch independent fields of this context's grid, advanced by one set
 * of launches (the batch is the grid's z dimension).  u holds the members back to back in the
 * reference layout: member b's component k is u + (3b + k)*Nx*Ny*Mz.  The forcing (may be null)
 * is shared.  Each member's result is bitwise what ns3d_rk3_ssp_steps gives it alone (same
 * kernels, same order).  NS3D_FP32_STATE runs the members one after another.  No reference
 * counterpart: the reference advances one field per RK3_SSP call (RK_time_step.cu:310-327);
 * this serves its batched callers (many finite-difference J.v perturbations, parameter sweeps). */
#define NS3D_MAX_BATCH 1024
int ns3d_rk3_ssp_batch(ns3d_ctx *ctx, int nbatch, int nsteps, ns3d_cplx *u,
                       const ns3d_cplx *fx, const ns3d_cplx *fy, const ns3d_cplx *fz);
/* ns3d_rk3_ssp_batch with per-member parameters, for a batch of resident requests of one grid:
 * dt, Re (device arrays of nbatch doubles) and steps (device array of nbatch ints): member b takes
 * the first steps[b] of the nsteps steps (0: an idle slot, untouched) with its own dt and Re; its
 * blocks of the other steps return at once.  The arrays are read when the kernels run, so a caller
 * may rewrite them between calls.  graph != 0 replays a HIP graph of the nsteps steps captured on
 * the first call (and again whenever the pointers, nbatch or nsteps change; needs a created
 * stream).  Each member's result is bitwise what ns3d_rk3_ssp_steps gives it alone with that dt
 * and Re.  NS3D_FP64 and NS3D_FP32_BULK (NS3D_ERR_PRECISION otherwise).  No reference counterpart:
 * the batching of a request server (one launch per pass for every resident request of a grid). */
int ns3d_rk3_ssp_batch_members(ns3d_ctx *ctx, int nbatch, int nsteps, ns3d_cplx *u,
                               const ns3d_cplx *fx, const ns3d_cplx *fy, const ns3d_cplx *fz,
                               const double *dt, const double *Re, const int *steps, int graph);

/* RK3_SSP of the velocity plus a passive scalar theta (physical, double [Nx][Ny][Nz]) advected
 * with WENO (scheme 3/5) by each stage's 2/3-masked physical velocity, with the same SSP
 * weights: theta_t = -(u.grad) theta.  No reference counterpart as a whole: it couples
 * RK3_SSP (RK_time_step.cu:310-327) with the advection kernel of advection_WENO.cu:304-392
 * applied to a scalar (the kernel's `x` argument). */
int ns3d_rk3_ssp_scalar(ns3d_ctx *ctx, int nsteps, ns3d_cplx *ux, ns3d_cplx *uy, ns3d_cplx *uz,
                        const ns3d_cplx *fx, const ns3d_cplx *fy, const ns3d_cplx *fz, double *theta, int scheme);

/* ---- a10: J.v and the tangent map (RK_time_step.cu:266-355, NS3D_periodic.cu:95-470) ---- */
int ns3d_rk3_ssp_uv_rhs(ns3d_ctx *ctx, const ns3d_cplx *ux, const ns3d_cplx *uy, const ns3d_cplx *uz,
                        ns3d_cplx *vx, ns3d_cplx *vy, ns3d_cplx *vz);             /* RK3_SSP_UV_RHS */
int ns3d_stage_uv(ns3d_ctx *ctx, const ns3d_cplx *ux, const ns3d_cplx *uy, const ns3d_cplx *uz,
                  ns3d_cplx *vx, ns3d_cplx *vy, ns3d_cplx *vz);                   /* single_time_step_UV */
int ns3d_stage_iuv_rhs(ns3d_ctx *ctx, const ns3d_cplx *ux, const ns3d_cplx *uy, const ns3d_cplx *uz,
                       ns3d_cplx *vx, ns3d_cplx *vy, ns3d_cplx *vz);              /* single_time_step_iUV_RHS */
int ns3d_rk3_ssp_uv(ns3d_ctx *ctx, ns3d_cplx *ux, ns3d_cplx *uy, ns3d_cplx *uz,
                    ns3d_cplx *vx, ns3d_cplx *vy, ns3d_cplx *vz,
                    const ns3d_cplx *fx, const ns3d_cplx *fy, const ns3d_cplx *fz); /* RK3_SSP_UV (advances u too) */
int ns3d_unpack_reduced(ns3d_ctx *ctx, const double *vec, ns3d_cplx *vx, ns3d_cplx *vy, ns3d_cplx *vz);
int ns3d_pack_reduced(ns3d_ctx *ctx, const ns3d_cplx *vx, const ns3d_cplx *vy, const ns3d_cplx *vz, double *vec);
int ns3d_unpack_full(ns3d_ctx *ctx, const double *vec, ns3d_cplx *vx, ns3d_cplx *vy, ns3d_cplx *vz);
int ns3d_pack_full(ns3d_ctx *ctx, const ns3d_cplx *vx, const ns3d_cplx *vy, const ns3d_cplx *vz, double *vec);
/* Base flow, forcing and period used by the callbacks below (the fields of
 * struct_NS3D_RK3_Call the reference callbacks read). */
int ns3d_set_base_state(ns3d_ctx *ctx, ns3d_cplx *ux, ns3d_cplx *uy, ns3d_cplx *uz,
                        const ns3d_cplx *fx, const ns3d_cplx *fy, const ns3d_cplx *fz,
                        int timesteps_period, double tau);
/* user_map_vector callbacks (Arnoldi/Matrix_Vector_emulator.h:11); vectors are device double*. */
void ns3d_jv_reduced(void *ctx, void *vec_in, void *vec_out);                  /* NSCallMatrixVector_reduced */
void ns3d_jv_reduced_newton(void *ctx, void *vec_in, void *vec_out);           /* ..._reduced_Newton_Globalization */
void ns3d_jv_exponential(void *ctx, void *vec_in, void *vec_out);              /* NSCallMatrixVector_exponential */

/* ---- a11/a12: reductions and the diagnostics that feed them.  Results are
 *      returned on the host (synchronous, like the reference). ---- */
int ns3d_reduce_minmax(ns3d_ctx *ctx, long long n, const double *in, double *out2_host);  /* compute_reduction: [min,max] */
int ns3d_reduce_sum(ns3d_ctx *ctx, long long n, const double *in, double *out_host);      /* reduction_sum */
/* NS3D_ERR_NONFINITE if any of the n doubles is NaN or Inf, else NS3D_OK (a spectral field of
 * nc modes is 2*nc doubles).  check_nans_kernel (Source/cuda_supp.cu:242-325), which returns -1 /
 * exits on NaN; Inf is reported too. */
int ns3d_check_finite(ns3d_ctx *ctx, long long n, const double *x);
/* The fused stage epilogues (every ns3d_stage / RK3 / J.v / update call through the fused passes)
 * and the passive-scalar stage set a sticky per-context flag when they write a NaN or Inf.
 * ns3d_finite_status synchronises the stream, returns NS3D_ERR_NONFINITE if the flag is set
 * (else NS3D_OK) and clears it if `clear`; ns3d_finite_flag_async enqueues a copy of the flag
 * (0 or 1) to host memory on the context's stream (no synchronisation; pinned memory keeps it
 * asynchronous).  No reference counterpart beyond the NaN checks above. */
int ns3d_finite_status(ns3d_ctx *ctx, int clear);
int ns3d_finite_flag_async(ns3d_ctx *ctx, int *host_flag);
int ns3d_kinetic_energy(ns3d_ctx *ctx, const double *ux, const double *uy, const double *uz, double *out_host);
int ns3d_dissipation(ns3d_ctx *ctx, const double *ux, const double *uy, const double *uz, double *out_host);
int ns3d_cfl_dt(ns3d_ctx *ctx, double cfl, const double *ux, const double *uy, const double *uz, double *dt_host);
/* Physical variants of the two diagnostics above, exact from the spectral fields (no
 * transform): 0.5 * integral of |u|^2 over the box, and the integral of 2 S:S
 * (= |grad u|^2 + (div u)^2; the dissipation rate is this over Re).  The reference's
 * get_kinetic_energy (math_support.cu:744) integrates only uy, uz over one plane area and
 * get_dissipation (:777-789) uses central differences (f[j+1]-f[j-1])/dx without the 1/2, which
 * tends to 4x this integral; those are ns3d_kinetic_energy and ns3d_dissipation. */
int ns3d_kinetic_energy_spectral(ns3d_ctx *ctx, const ns3d_cplx *ux, const ns3d_cplx *uy, const ns3d_cplx *uz,
                                 double *out_host);
int ns3d_dissipation_spectral(ns3d_ctx *ctx, const ns3d_cplx *ux, const ns3d_cplx *uy, const ns3d_cplx *uz,
                              double *out_host);
int ns3d_shapiro_error(ns3d_ctx *ctx, double t, const double *ux, const double *uy, const double *uz,
                       double *err_host);                                          /* Shapiro_test_case: L-inf error */

/* ---- f1: FP64 vector kernels for the Krylov drivers (Arnoldi Gram-Schmidt with
 *      DGKS re-orthogonalisation, implicit restarts, BiCGStab(L), Newton), the device
 *      BLAS of Source/Arnoldi/cuda_supp.cu:180-990 (vector_dot_product_GPU,
 *      matrixDotVector_part_GPU, matrixMultVector_part_GPU, matrixMultMatrix_GPU,
 *      vectors_add_GPU).  A basis V is m device vectors of length n, vector j at
 *      V + j*ldv; m, k <= 128.  Reductions are deterministic (fixed partition) and
 *      return on the host after a stream synchronisation. ---- */
int ns3d_vec_dot(ns3d_ctx *ctx, long long n, const double *x, const double *y, double *out_host);
int ns3d_vec_gemv_t(ns3d_ctx *ctx, long long n, int m, const double *V, long long ldv, const double *w,
                    double *h_host);                                    /* h = V^T w */
int ns3d_vec_gemv(ns3d_ctx *ctx, long long n, int m, const double *V, long long ldv, const double *h_host,
                  double alpha, double beta, double *y);                /* y = alpha V h + beta y */
int ns3d_vec_gemm(ns3d_ctx *ctx, long long n, int m, int k, const double *V, long long ldv,
                  const double *Q_host, double *W, long long ldw);      /* W = V Q, Q m x k column-major; W, V disjoint */
int ns3d_vec_axpby(ns3d_ctx *ctx, long long n, double a, const double *x, double b, double *y);  /* y = a x + b y */

/* ---- f4: diagnostics of the reference driver (Source/math_support.cu:712-987) ---- */
int ns3d_curl(ns3d_ctx *ctx, const double *ux, const double *uy, const double *uz,
              double *rot_x, double *rot_y, double *rot_z, int fix);    /* get_curl; fix=0 keeps the
                                                                          reference's uz(j+1)-ux(j-1) in rot_y */
int ns3d_helmholtz_filter(ns3d_ctx *ctx, double eps, const ns3d_cplx *ux, const ns3d_cplx *uy, const ns3d_cplx *uz,
                          ns3d_cplx *fx, ns3d_cplx *fy, ns3d_cplx *fz);   /* Helmholz_Fourier_Filter (spectral part) */
int ns3d_cutoff_filter(ns3d_ctx *ctx, double radius, const ns3d_cplx *ux, const ns3d_cplx *uy, const ns3d_cplx *uz,
                       ns3d_cplx *fx, ns3d_cplx *fy, ns3d_cplx *fz);      /* CutOff_Fourier_Filter (spectral part) */

/* ---- f3: return map.  Velocity at grid point (j,k,l) evaluated from the half spectrum in one
 *      pass (return_vector3_solution, Source/return_map.cu:110-160, which inverse-transforms
 *      all three fields to read one point); out3_host = (ux, uy, uz). ---- */
int ns3d_probe(ns3d_ctx *ctx, const ns3d_cplx *ux, const ns3d_cplx *uy, const ns3d_cplx *uz, int j, int k, int l,
               double *out3_host);
/* ns3d_probe for nbatch members of a batch in ns3d_rk3_ssp_batch's layout (member b's component c
 * at u + (3b + c)*Nx*Ny*Mz): out_host[3b + c] = component c of member b at (j, k, l).  One launch
 * for the batch (execute_return_map's seed disc, Source/return_map.cu:502-690, advanced together). */
int ns3d_probe_batch(ns3d_ctx *ctx, int nbatch, const ns3d_cplx *u, int j, int k, int l, double *out_host);

/* ---- measurement: per-pass device time of the fused pipeline, recorded with
 *      HIP events on the context stream while enabled.  ms[5]/calls[5] are
 *      indexed 0 x-inverse, 1 y-inverse, 2 z product, 3 y-forward,
 *      4 x-forward + epilogue.  No reference counterpart (timing hook). ---- */
int ns3d_profile_enable(ns3d_ctx *ctx, int on);
int ns3d_profile_read(ns3d_ctx *ctx, double *ms, long long *calls);

#ifdef __cplusplus
}
#endif
#endif
