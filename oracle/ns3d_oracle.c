/*
 * ns3d_oracle.c -- CPU FP64 restatement of NS3DPeriodic's time-stepping hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the
 * "port" CPU baseline.  The product path never links or calls this file.
 *
 * Each function cites the reference source it restates (paths relative to the
 * reference root).  What pins each function independently of this restatement
 * (tests/test_oracle_*.py):
 *   FFT (or_fft_*)          brute-force O(N^2) DFT, numpy.fft rfftn/irfftn,
 *                           single-mode closed forms, round trip, Parseval.
 *   tables (or_tables)      closed forms, AM idempotent (AM^2=AM), AM k = k,
 *                           exact integer evaluation of the 2/3 ellipsoid.
 *   nonlinear/bilinear      analytic (u.grad)u of trigonometric fields
 *                           (fixes derivative sign, component indices, scale,
 *                           and the IPQ real-part rule), brute-force direct DFT
 *                           evaluation at N=8, B(u,u)=2N(u) and
 *                           B(u,v)=N(u+v)-N(u)-N(v) identities.
 *   update variants         linear closed forms per mode with Q given.
 *   RK3 / stage             linear-decay amplification (1+z+z^2/2+z^3/6)^n,
 *                           Shapiro Beltrami analytic solution, laminar
 *                           Kolmogorov fixed point.
 *   J.v / tangent map       central finite differences of the nonlinear map.
 *   divergence              closed form on trigonometric fields.
 *   WENO                    convergence order on smooth data; one Euler step
 *                           of a step function stays within its bounds along
 *                           each axis (fixes the upwind side); the two cells
 *                           of a face use the same face value, so
 *                           sum_j out_j / u_j telescopes to zero along a line
 *                           (fixes the face-averaged velocity and the face
 *                           indices).
 *   passive scalar RK3      theta advected by the exact decaying shear
 *                           (e^{-t/Re} sin y, 0, 0) against the closed form
 *                           theta0(x - sin y Re (1 - e^{-t/Re})), third order
 *                           in time (fixes each stage's velocity and weights).
 *
 * or_set_mutation() plants one known wiring mistake (test-only fault injection
 * on the CPU): the tests use it to show that each pinning check above fails
 * on the mistake it is meant to catch.  0 = the restatement as written.
 *   reductions              exact min/max; long double sums.
 *   energy/dissipation/CFL  closed forms on trigonometric fields.
 */
#include "ns3d_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cx;

/* The reference's PI literal (Source/Macros.h:75-77) rounds to M_PI. */
#define OR_PI 3.141592653589793238462643383279502884

#define IDX3(j, k, l, N2, N3) (((long)(j) * (N2) + (k)) * (long)(N3) + (l))

/* Planted mistakes (test-only): 1/2/3 downwind faces along x/y/z, 4 cell
 * velocity in place of the face average, 5 every scalar stage advected by
 * u^n's velocity, 6 the second scalar stage's SSP weights swapped. */
static int or_mut = 0;
void or_set_mutation(int m) { or_mut = m; }

static void set_threads(const or_params *p)
{
#ifdef _OPENMP
    if (p && p->nthreads > 0) omp_set_num_threads(p->nthreads);
#else
    (void)p;
#endif
}

/* ------------------------------------------------------------------------ */
/* 1D mixed-radix FFT: recursive decimation in time, n = p1*p2*...,         */
/* radix-4/2 butterflies plus a generic O(p^2) butterfly for other factors. */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n, sign, nf;
    int fac[40];
    cx *tw; /* tw[k] = exp(sign*2*pi*i*k/n) */
} plan1d;

static void plan_init(plan1d *pl, int n, int sign)
{
    pl->n = n;
    pl->sign = sign;
    pl->nf = 0;
    int m = n;
    while (m % 4 == 0) { pl->fac[pl->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { pl->fac[pl->nf++] = 2; m /= 2; }
    for (int f = 3; m > 1; f += 2)
        while (m % f == 0) { pl->fac[pl->nf++] = f; m /= f; }
    pl->tw = (cx *)malloc(sizeof(cx) * (size_t)n);
    for (int k = 0; k < n; k++) {
        long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
        pl->tw[k] = (double)cosl(a) + (double)sign * (double)sinl(a) * I;
    }
}

static void plan_free(plan1d *pl) { free(pl->tw); }

/* out[0..n-1] = DFT_n of in[0], in[is], ..., in[(n-1)*is]; tws = plan.n / n */
static void dit(const plan1d *pl, int n, cx *out, const cx *in, long is, const int *fac, int tws)
{
    if (n == 1) { out[0] = in[0]; return; }
    const int p = fac[0], m = n / p, N = pl->n;
    for (int q = 0; q < p; q++) dit(pl, m, out + (long)q * m, in + (long)q * is, is * p, fac + 1, tws * p);
    const cx *tw = pl->tw;
    if (p == 2) {
        for (int u = 0; u < m; u++) {
            cx a0 = out[u], a1 = out[m + u] * tw[(long)u * tws % N];
            out[u] = a0 + a1;
            out[u + m] = a0 - a1;
        }
    } else if (p == 4) {
        const cx jw = (double)pl->sign * I; /* W_4 = exp(sign*i*pi/2) */
        for (int u = 0; u < m; u++) {
            cx a0 = out[u];
            cx a1 = out[m + u] * tw[(long)u * tws % N];
            cx a2 = out[2 * m + u] * tw[(long)2 * u * tws % N];
            cx a3 = out[3 * m + u] * tw[(long)3 * u * tws % N];
            cx s02 = a0 + a2, d02 = a0 - a2, s13 = a1 + a3, d13 = (a1 - a3) * jw;
            out[u] = s02 + s13;
            out[u + m] = d02 + d13;
            out[u + 2 * m] = s02 - s13;
            out[u + 3 * m] = d02 - d13;
        }
    } else {
        cx tmp[64];
        cx *t = p <= 64 ? tmp : (cx *)malloc(sizeof(cx) * (size_t)p);
        for (int u = 0; u < m; u++) {
            for (int q = 0; q < p; q++) t[q] = out[(long)q * m + u] * tw[(long)q * u * tws % N];
            for (int s = 0; s < p; s++) {
                cx acc = 0;
                for (int q = 0; q < p; q++) acc += t[q] * tw[((long)q * s % p) * m * tws % N];
                out[u + (long)s * m] = acc;
            }
        }
        if (t != tmp) free(t);
    }
}

static void fft_line(const plan1d *pl, const cx *in, cx *out) { dit(pl, pl->n, out, in, 1, pl->fac, 1); }

void or_fft_c2c_1d(int n, int sign, const double *in, double *out)
{
    plan1d pl;
    plan_init(&pl, n, sign < 0 ? -1 : 1);
    fft_line(&pl, (const cx *)in, (cx *)out);
    plan_free(&pl);
}

/* C2C along one axis of a complex [A][B][C] array (in place). axis 0, 1 or 2.
 * Strided axes are gathered OR_W lines at a time (lines adjacent in the
 * fastest index), so every cache line read or written is used whole. */
#define OR_W 8
static void fft_axis(cx *a, int A, int B, int C, int axis, int sign)
{
    const int n = axis == 0 ? A : axis == 1 ? B : C;
    plan1d pl;
    plan_init(&pl, n, sign);
    if (axis == 2) {
        const long nlines = (long)A * B;
        #pragma omp parallel
        {
            cx *res = (cx *)malloc(sizeof(cx) * (size_t)n);
            #pragma omp for schedule(static)
            for (long ln = 0; ln < nlines; ln++) {
                fft_line(&pl, a + ln * C, res);
                memcpy(a + ln * C, res, sizeof(cx) * (size_t)n);
            }
            free(res);
        }
    } else {
        /* lines (o, c): o = b (axis 0) or i (axis 1); c = fastest index, grouped by OR_W */
        const long nouter = axis == 0 ? B : A;
        const long stride = axis == 0 ? (long)B * C : C;
        const long ngrp = (C + OR_W - 1) / OR_W;
        #pragma omp parallel
        {
            cx *buf = (cx *)malloc(sizeof(cx) * (size_t)n * OR_W);
            cx *res = (cx *)malloc(sizeof(cx) * (size_t)n);
            #pragma omp for schedule(static)
            for (long g = 0; g < nouter * ngrp; g++) {
                const long o = g / ngrp, c0 = (g % ngrp) * OR_W;
                const int w = (int)(C - c0 < OR_W ? C - c0 : OR_W);
                const long base = (axis == 0 ? o * C : o * (long)B * C) + c0;
                for (int q = 0; q < n; q++)
                    for (int e = 0; e < w; e++) buf[(long)e * n + q] = a[base + q * stride + e];
                for (int e = 0; e < w; e++) {
                    fft_line(&pl, buf + (long)e * n, res);
                    memcpy(buf + (long)e * n, res, sizeof(cx) * (size_t)n);
                }
                for (int q = 0; q < n; q++)
                    for (int e = 0; e < w; e++) a[base + q * stride + e] = buf[(long)e * n + q];
            }
            free(buf);
            free(res);
        }
    }
    plan_free(&pl);
}

/* FFTN_Device (Source/math_support.cu:301-312): cufftExecD2Z, forward sign -1, unnormalised. */
void or_fft_r2c(const or_params *p, const double *in, double *out_)
{
    set_threads(p);
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz, Mz = Nz / 2 + 1;
    cx *out = (cx *)out_;
    plan1d pz;
    plan_init(&pz, Nz, -1);
    #pragma omp parallel
    {
        cx *buf = (cx *)malloc(sizeof(cx) * (size_t)Nz * 2);
        cx *res = buf + Nz;
        #pragma omp for schedule(static)
        for (long ln = 0; ln < (long)Nx * Ny; ln++) {
            for (int l = 0; l < Nz; l++) buf[l] = in[ln * Nz + l];
            fft_line(&pz, buf, res);
            for (int l = 0; l < Mz; l++) out[ln * Mz + l] = res[l];
        }
        free(buf);
    }
    plan_free(&pz);
    fft_axis(out, Nx, Ny, Mz, 1, -1);
    fft_axis(out, Nx, Ny, Mz, 0, -1);
}

/* iFFTN_Device (Source/math_support.cu:315-334): cufftExecZ2D then, if scale,
 * scale_double (:123-150) multiplies by 1/(1.0*Nx*Ny*Nz).
 * Convention for the half-spectrum axis (documented, cuFFT leaves it open for
 * non-Hermitian input): the x and y inverse C2C passes run first, then the
 * z-C2R uses only the real part of the l=0 and l=Nz/2 coefficients. */
void or_fft_c2r(const or_params *p, const double *in_, double *out, int scale)
{
    set_threads(p);
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz, Mz = Nz / 2 + 1;
    const long nc = (long)Nx * Ny * Mz;
    cx *t = (cx *)malloc(sizeof(cx) * (size_t)nc);
    memcpy(t, in_, sizeof(cx) * (size_t)nc);
    fft_axis(t, Nx, Ny, Mz, 0, +1);
    fft_axis(t, Nx, Ny, Mz, 1, +1);
    plan1d pz;
    plan_init(&pz, Nz, +1);
    const double sc = 1.0 / (1.0 * Nx * Ny * Nz);
    #pragma omp parallel
    {
        cx *buf = (cx *)malloc(sizeof(cx) * (size_t)Nz * 2);
        cx *res = buf + Nz;
        #pragma omp for schedule(static)
        for (long ln = 0; ln < (long)Nx * Ny; ln++) {
            const cx *X = t + ln * Mz;
            buf[0] = creal(X[0]);
            for (int l = 1; l < Mz; l++) buf[l] = X[l];
            if (Nz % 2 == 0) buf[Nz / 2] = creal(X[Nz / 2]);
            for (int l = Mz; l < Nz; l++) buf[l] = conj(X[Nz - l]);
            fft_line(&pz, buf, res);
            for (int l = 0; l < Nz; l++) {
                double v = creal(res[l]);
                out[ln * Nz + l] = scale ? v * sc : v;
            }
        }
        free(buf);
    }
    plan_free(&pz);
    free(t);
}

/* ------------------------------------------------------------------------ */
/* Host tables (Source/math_support.cu:350-519), built with the dimensions  */
/* the driver passes: (Nx, Ny, Mz) (Source/NS3D_periodic.cu:1151-1161).     */
/* ------------------------------------------------------------------------ */
void or_tables(const or_params *p, double *kx, double *ky, double *kz,
               double *din_diffusion, double *din_poisson,
               double *AM11, double *AM22, double *AM33, double *AM12, double *AM13, double *AM23,
               double *mask)
{
    const int Nx = p->Nx, Ny = p->Ny, Mz = p->Nz / 2 + 1;
    const double Lx = p->Lx, Ly = p->Ly, Lz = p->Lz;
    /* build_Nabla_Wavenumbers (:393-414): m wraps at Nx/2 in x and y, not in z. */
    for (int j = 0; j < Nx; j++) { int m = j >= Nx / 2 ? j - Nx : j; if (kx) kx[j] = 2.0 * M_PI / Lx * m; }
    for (int k = 0; k < Ny; k++) { int n = k >= Ny / 2 ? k - Ny : k; if (ky) ky[k] = 2.0 * M_PI / Ly * n; }
    for (int l = 0; l < Mz; l++) { if (kz) kz[l] = 2.0 * M_PI / Lz * l; }
    for (int j = 0; j < Nx; j++) {
        const int m = j >= Nx / 2 ? j - Nx : j;
        const double kxl = 2.0 * M_PI / Lx * m;
        for (int k = 0; k < Ny; k++) {
            const int n = k >= Ny / 2 ? k - Ny : k;
            const double kyl = 2.0 * M_PI / Ly * n;
            for (int l = 0; l < Mz; l++) {
                const int q = l;
                const long id = IDX3(j, k, l, Ny, Mz);
                const double kzl = 2.0 * M_PI / Lz * q;
                /* build_Laplace_and_Diffusion (:374-390) */
                const double dp = -(kxl * kxl + kyl * kyl + kzl * kzl);
                if (din_poisson) din_poisson[id] = dp;
                if (din_diffusion) din_diffusion[id] = -dp;
                /* build_projection_matrix_elements (:418-466) */
                const double k1 = 2.0 * M_PI / Lx, k2 = 2.0 * M_PI / Ly, k3 = 2.0 * M_PI / Lz;
                double din = k1 * k1 * m * m + k2 * k2 * n * n + k3 * k3 * q * q;
                if ((m * m + n * n + q * q) == 0) din = 1.0;
                if (AM11) AM11[id] = k1 * k1 * m * m / din;
                if (AM22) AM22[id] = k2 * k2 * n * n / din;
                if (AM33) AM33[id] = k3 * k3 * q * q / din;
                if (AM12) AM12[id] = k1 * k2 * m * n / din;
                if (AM13) AM13[id] = k1 * k3 * m * q / din;
                if (AM23) AM23[id] = k2 * k3 * n * q / din;
                /* build_mask_matrix (:471-519), called with Nz := Mz (Q2). */
                if (mask) {
                    const double kxm = m * 2 * OR_PI / Lx, kym = n * 2 * OR_PI / Ly, kzm = q * 2 * OR_PI / Lz;
                    const double kxMax = p->mask_nyq ? Nx * OR_PI / Lx : 2 * Nx * OR_PI / Lx;
                    const double kyMax = p->mask_nyq ? Ny * OR_PI / Ly : 2 * Ny * OR_PI / Ly;
                    const double kzMax = p->mask_nyq ? p->Nz * OR_PI / Lz : 2 * Mz * OR_PI / Lz;
                    const double s2 = sqrt(kxm * kxm / (kxMax * kxMax) + kym * kym / (kyMax * kyMax) + kzm * kzm / (kzMax * kzMax));
                    mask[id] = s2 < 2.0 / 3.0 ? 1.0 : 0.0;
                }
            }
        }
    }
    if (din_poisson) din_poisson[0] = 1.0;
    if (AM11) { AM11[0] = 0.0; AM22[0] = 0.0; AM33[0] = 0.0; AM12[0] = 0.0; AM13[0] = 0.0; AM23[0] = 0.0; }
    if (mask) mask[0] = 0.0;
}

typedef struct {
    int Nx, Ny, Nz, Mz;
    long nc, nr;
    double *kx, *ky, *kz, *din, *AM[6], *mask;
} tables_t;

static void tables_make(const or_params *p, tables_t *t)
{
    t->Nx = p->Nx; t->Ny = p->Ny; t->Nz = p->Nz; t->Mz = p->Nz / 2 + 1;
    t->nc = (long)t->Nx * t->Ny * t->Mz;
    t->nr = (long)t->Nx * t->Ny * t->Nz;
    t->kx = (double *)malloc(sizeof(double) * t->Nx);
    t->ky = (double *)malloc(sizeof(double) * t->Ny);
    t->kz = (double *)malloc(sizeof(double) * t->Mz);
    t->din = (double *)malloc(sizeof(double) * t->nc);
    for (int i = 0; i < 6; i++) t->AM[i] = (double *)malloc(sizeof(double) * t->nc);
    t->mask = (double *)malloc(sizeof(double) * t->nc);
    double *poisson = (double *)malloc(sizeof(double) * t->nc);
    or_tables(p, t->kx, t->ky, t->kz, t->din, poisson, t->AM[0], t->AM[1], t->AM[2], t->AM[3], t->AM[4], t->AM[5], t->mask);
    free(poisson);
}

static void tables_free(tables_t *t)
{
    free(t->kx); free(t->ky); free(t->kz); free(t->din); free(t->mask);
    for (int i = 0; i < 6; i++) free(t->AM[i]);
}

/* build_vels_device + build_ders_device + filter_1p3_wavenumbers_device
 * (Source/advection_2_3.cu:146-217,284-316): w~ = M w^, d_a = M (i k_a w^). */
static void masked_field_and_derivs(const tables_t *t, const cx *w, cx *wm, cx *dx, cx *dy, cx *dz)
{
    const int Ny = t->Ny, Mz = t->Mz;
    #pragma omp parallel for schedule(static)
    for (long id = 0; id < t->nc; id++) {
        const int l = (int)(id % Mz), k = (int)((id / Mz) % Ny), j = (int)(id / ((long)Mz * Ny));
        const double m = t->mask[id];
        const cx s = w[id];
        /* der.x = -s.y*k, der.y = s.x*k, then x mask (advection_2_3.cu:206-213, :304-311) */
        if (wm) wm[id] = creal(s) * m + cimag(s) * m * I;
        if (dx) dx[id] = (-cimag(s) * t->kx[j]) * m + (creal(s) * t->kx[j]) * m * I;
        if (dy) dy[id] = (-cimag(s) * t->ky[k]) * m + (creal(s) * t->ky[k]) * m * I;
        if (dz) dz[id] = (-cimag(s) * t->kz[l]) * m + (creal(s) * t->kz[l]) * m * I;
    }
}

/* set_to_zero_real (Source/advection_2_3.cu:35-70): Re(Q^) = 0 everywhere, Q^(0) = 0. */
static void ipq(const tables_t *t, cx *Q)
{
    for (long id = 0; id < t->nc; id++) Q[id] = cimag(Q[id]) * I;
    Q[0] = 0;
}

/* calculate_convolution_2p3 (Source/advection_2_3.cu:480-497) with convolute_2p3 (:359-397):
 * 3 velocity C2R, then per component b: 3 derivative C2R, Q_b = ux d_x w_b + uy d_y w_b + uz d_z w_b
 * (calc_vels_nabla_vels :224-251), D2Z; finally IPQ (:493). */
void or_nonlinear(const or_params *p, const double *ux, const double *uy, const double *uz,
                  double *Qx, double *Qy, double *Qz)
{
    set_threads(p);
    tables_t t;
    tables_make(p, &t);
    const cx *u[3] = {(const cx *)ux, (const cx *)uy, (const cx *)uz};
    cx *Q[3] = {(cx *)Qx, (cx *)Qy, (cx *)Qz};
    cx *sp = (cx *)malloc(sizeof(cx) * (size_t)t.nc * 3);
    double *vel = (double *)malloc(sizeof(double) * (size_t)t.nr * 3);
    double *der = (double *)malloc(sizeof(double) * (size_t)t.nr * 3);
    double *q = (double *)malloc(sizeof(double) * (size_t)t.nr);
    for (int b = 0; b < 3; b++) {
        masked_field_and_derivs(&t, u[b], sp, NULL, NULL, NULL);
        or_fft_c2r(p, (const double *)sp, vel + b * t.nr, 1);
    }
    for (int b = 0; b < 3; b++) {
        masked_field_and_derivs(&t, u[b], NULL, sp, sp + t.nc, sp + 2 * t.nc);
        for (int a = 0; a < 3; a++) or_fft_c2r(p, (const double *)(sp + a * t.nc), der + a * t.nr, 1);
        #pragma omp parallel for schedule(static)
        for (long i = 0; i < t.nr; i++)
            q[i] = vel[i] * der[i] + vel[t.nr + i] * der[t.nr + i] + vel[2 * t.nr + i] * der[2 * t.nr + i];
        or_fft_r2c(p, q, (double *)Q[b]);
        if (p->ipq) ipq(&t, Q[b]);
    }
    free(sp); free(vel); free(der); free(q);
    tables_free(&t);
}

/* calculate_convolution_2p3_UV (Source/advection_2_3.cu:500-525), convolute_2p3_UV (:401-475):
 * Q_b = (v.grad) u_b + (u.grad) v_b, then IPQ. */
void or_bilinear(const or_params *p, const double *ux, const double *uy, const double *uz,
                 const double *vx, const double *vy, const double *vz,
                 double *Qx, double *Qy, double *Qz)
{
    set_threads(p);
    tables_t t;
    tables_make(p, &t);
    const cx *u[3] = {(const cx *)ux, (const cx *)uy, (const cx *)uz};
    const cx *v[3] = {(const cx *)vx, (const cx *)vy, (const cx *)vz};
    cx *Q[3] = {(cx *)Qx, (cx *)Qy, (cx *)Qz};
    cx *sp = (cx *)malloc(sizeof(cx) * (size_t)t.nc * 3);
    double *uvel = (double *)malloc(sizeof(double) * (size_t)t.nr * 3);
    double *vvel = (double *)malloc(sizeof(double) * (size_t)t.nr * 3);
    double *der = (double *)malloc(sizeof(double) * (size_t)t.nr * 3);
    double *q = (double *)malloc(sizeof(double) * (size_t)t.nr);
    for (int b = 0; b < 3; b++) {
        masked_field_and_derivs(&t, u[b], sp, NULL, NULL, NULL);
        or_fft_c2r(p, (const double *)sp, uvel + b * t.nr, 1);
        masked_field_and_derivs(&t, v[b], sp, NULL, NULL, NULL);
        or_fft_c2r(p, (const double *)sp, vvel + b * t.nr, 1);
    }
    for (int b = 0; b < 3; b++) {
        /* (v.grad) u_b  (calc_vels_nabla_vels with v velocities, u derivatives) */
        masked_field_and_derivs(&t, u[b], NULL, sp, sp + t.nc, sp + 2 * t.nc);
        for (int a = 0; a < 3; a++) or_fft_c2r(p, (const double *)(sp + a * t.nc), der + a * t.nr, 1);
        #pragma omp parallel for schedule(static)
        for (long i = 0; i < t.nr; i++)
            q[i] = vvel[i] * der[i] + vvel[t.nr + i] * der[t.nr + i] + vvel[2 * t.nr + i] * der[2 * t.nr + i];
        /* += (u.grad) v_b  (calc_vels_nabla_vels_plus :254-281) */
        masked_field_and_derivs(&t, v[b], NULL, sp, sp + t.nc, sp + 2 * t.nc);
        for (int a = 0; a < 3; a++) or_fft_c2r(p, (const double *)(sp + a * t.nc), der + a * t.nr, 1);
        #pragma omp parallel for schedule(static)
        for (long i = 0; i < t.nr; i++)
            q[i] += (uvel[i] * der[i] + uvel[t.nr + i] * der[t.nr + i] + uvel[2 * t.nr + i] * der[2 * t.nr + i]);
        or_fft_r2c(p, q, (double *)Q[b]);
        if (p->ipq) ipq(&t, Q[b]);
    }
    free(sp); free(uvel); free(vvel); free(der); free(q);
    tables_free(&t);
}

/* Projection/diffusion/update family (Source/diffusion.cu:96-365).
 * AM Q uses the off-diagonal-symmetric tables; the identity part of the
 * projector is written as the explicit "- Q" term exactly as in the source. */
static void update_tab(const or_params *p, const tables_t *t, int mode,
                       cx *ux, cx *uy, cx *uz, const cx *fx, const cx *fy, const cx *fz,
                       const cx *Qx, const cx *Qy, const cx *Qz, cx *Rx, cx *Ry, cx *Rz)
{
    const double dt = p->dt, Re = p->Re;
    #pragma omp parallel for schedule(static)
    for (long id = 0; id < t->nc; id++) {
        const double a11 = t->AM[0][id], a22 = t->AM[1][id], a33 = t->AM[2][id];
        const double a12 = t->AM[3][id], a13 = t->AM[4][id], a23 = t->AM[5][id];
        const cx qx = Qx[id], qy = Qy[id], qz = Qz[id];
        const double AQxr = a11 * creal(qx) + a12 * creal(qy) + a13 * creal(qz);
        const double AQxi = a11 * cimag(qx) + a12 * cimag(qy) + a13 * cimag(qz);
        const double AQyr = a12 * creal(qx) + a22 * creal(qy) + a23 * creal(qz);
        const double AQyi = a12 * cimag(qx) + a22 * cimag(qy) + a23 * cimag(qz);
        const double AQzr = a13 * creal(qx) + a23 * creal(qy) + a33 * creal(qz);
        const double AQzi = a13 * cimag(qx) + a23 * cimag(qy) + a33 * cimag(qz);
        const double din = t->din[id];
        const cx w[3] = {ux[id], uy[id], uz[id]};
        const double AQr[3] = {AQxr, AQyr, AQzr}, AQi[3] = {AQxi, AQyi, AQzi};
        const cx q[3] = {qx, qy, qz};
        cx res[3];
        for (int c = 0; c < 3; c++) {
            double rr, ri;
            if (mode == OR_UPD_STAGE) {
                const cx f = c == 0 ? fx[id] : c == 1 ? fy[id] : fz[id];
                const double G = (1.0 - din * dt / Re);
                rr = dt * AQr[c] + G * creal(w[c]) + dt * creal(f) - dt * creal(q[c]);
                ri = dt * AQi[c] + G * cimag(w[c]) + dt * cimag(f) - dt * cimag(q[c]);
            } else if (mode == OR_UPD_UV_RHS) {
                const double G = (-din / Re);
                rr = AQr[c] + G * creal(w[c]) - creal(q[c]);
                ri = AQi[c] + G * cimag(w[c]) - cimag(q[c]);
            } else if (mode == OR_UPD_UV) {
                const double G = (1.0 - din * dt / Re);
                rr = dt * AQr[c] + G * creal(w[c]) - dt * creal(q[c]);
                ri = dt * AQi[c] + G * cimag(w[c]) - dt * cimag(q[c]);
            } else if (mode == OR_UPD_IUV) {
                const double G = (1.0 + din * dt / Re);
                rr = (dt * AQr[c] + creal(w[c]) - dt * creal(q[c])) / G;
                ri = (dt * AQi[c] + cimag(w[c]) - dt * cimag(q[c])) / G;
            } else { /* OR_UPD_RHS */
                const cx f = c == 0 ? fx[id] : c == 1 ? fy[id] : fz[id];
                const double G = (-din / Re);
                rr = AQr[c] + G * creal(w[c]) + creal(f) - creal(q[c]);
                ri = AQi[c] + G * cimag(w[c]) + cimag(f) - cimag(q[c]);
            }
            res[c] = rr + ri * I;
        }
        if (mode == OR_UPD_RHS) { Rx[id] = res[0]; Ry[id] = res[1]; Rz[id] = res[2]; }
        else { ux[id] = res[0]; uy[id] = res[1]; uz[id] = res[2]; }
    }
}

void or_update(const or_params *p, int mode, double *ux, double *uy, double *uz,
               const double *fx, const double *fy, const double *fz,
               const double *Qx, const double *Qy, const double *Qz,
               double *Rx, double *Ry, double *Rz)
{
    set_threads(p);
    tables_t t;
    tables_make(p, &t);
    update_tab(p, &t, mode, (cx *)ux, (cx *)uy, (cx *)uz, (const cx *)fx, (const cx *)fy, (const cx *)fz,
               (const cx *)Qx, (const cx *)Qy, (const cx *)Qz, (cx *)Rx, (cx *)Ry, (cx *)Rz);
    tables_free(&t);
}

static long ncomplex(const or_params *p) { return (long)p->Nx * p->Ny * (p->Nz / 2 + 1); }

/* single_time_step (Source/RK_time_step.cu:248-263): nonlinear term then the fused update, in place. */
void or_single_time_step(const or_params *p, double *ux, double *uy, double *uz,
                         const double *fx, const double *fy, const double *fz)
{
    const long nc = ncomplex(p);
    double *Q = (double *)malloc(sizeof(cx) * (size_t)nc * 3);
    or_nonlinear(p, ux, uy, uz, Q, Q + 2 * nc, Q + 4 * nc);
    or_update(p, OR_UPD_STAGE, ux, uy, uz, fx, fy, fz, Q, Q + 2 * nc, Q + 4 * nc, NULL, NULL, NULL);
    free(Q);
}

/* return_RHS (Source/RK_time_step.cu:231-245): R = AM Q - Q - |k|^2/Re u + f. */
void or_return_rhs(const or_params *p, const double *ux, const double *uy, const double *uz,
                   const double *fx, const double *fy, const double *fz,
                   double *Rx, double *Ry, double *Rz)
{
    const long nc = ncomplex(p);
    double *Q = (double *)malloc(sizeof(cx) * (size_t)nc * 3);
    or_nonlinear(p, ux, uy, uz, Q, Q + 2 * nc, Q + 4 * nc);
    or_update(p, OR_UPD_RHS, (double *)ux, (double *)uy, (double *)uz, fx, fy, fz, Q, Q + 2 * nc, Q + 4 * nc, Rx, Ry, Rz);
    free(Q);
}

/* single_RK_step (Source/RK_time_step.cu:82-115): dst = w0*s0 + w1*s1 (component-wise). */
static void axpby(long nc, double w0, const double *s0, double w1, const double *s1, double *dst)
{
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < 2 * nc; i++) dst[i] = w0 * s0[i] + w1 * s1[i];
}

/* RK3_SSP (Source/RK_time_step.cu:310-327). */
void or_rk3_ssp(const or_params *p, double *ux, double *uy, double *uz,
                const double *fx, const double *fy, const double *fz)
{
    const long nc = ncomplex(p);
    double *u1 = (double *)malloc(sizeof(cx) * (size_t)nc * 3);
    double *u2 = (double *)malloc(sizeof(cx) * (size_t)nc * 3);
    double *u[3] = {ux, uy, uz};
    for (int c = 0; c < 3; c++) memcpy(u1 + 2 * c * nc, u[c], sizeof(cx) * (size_t)nc);
    or_single_time_step(p, u1, u1 + 2 * nc, u1 + 4 * nc, fx, fy, fz);
    or_single_time_step(p, u1, u1 + 2 * nc, u1 + 4 * nc, fx, fy, fz);
    for (int c = 0; c < 3; c++) axpby(nc, 3.0 / 4.0, u[c], 1.0 / 4.0, u1 + 2 * c * nc, u2 + 2 * c * nc);
    or_single_time_step(p, u2, u2 + 2 * nc, u2 + 4 * nc, fx, fy, fz);
    for (int c = 0; c < 3; c++) axpby(nc, 1.0 / 3.0, u[c], 2.0 / 3.0, u2 + 2 * c * nc, u[c]);
    free(u1); free(u2);
}

/* RK3_SSP_UV_RHS (Source/RK_time_step.cu:330-338) -> single_time_step_UV_RHS (:282-292):
 * v <- AM B - B - |k|^2/Re v with B = B(u, v). */
void or_rk3_ssp_uv_rhs(const or_params *p, const double *ux, const double *uy, const double *uz,
                       double *vx, double *vy, double *vz)
{
    const long nc = ncomplex(p);
    double *Q = (double *)malloc(sizeof(cx) * (size_t)nc * 3);
    or_bilinear(p, ux, uy, uz, vx, vy, vz, Q, Q + 2 * nc, Q + 4 * nc);
    or_update(p, OR_UPD_UV_RHS, vx, vy, vz, NULL, NULL, NULL, Q, Q + 2 * nc, Q + 4 * nc, NULL, NULL, NULL);
    free(Q);
}

/* single_time_step_iUV_RHS (Source/RK_time_step.cu:296-306). */
void or_single_time_step_iuv_rhs(const or_params *p, const double *ux, const double *uy, const double *uz,
                                 double *vx, double *vy, double *vz)
{
    const long nc = ncomplex(p);
    double *Q = (double *)malloc(sizeof(cx) * (size_t)nc * 3);
    or_bilinear(p, ux, uy, uz, vx, vy, vz, Q, Q + 2 * nc, Q + 4 * nc);
    or_update(p, OR_UPD_IUV, vx, vy, vz, NULL, NULL, NULL, Q, Q + 2 * nc, Q + 4 * nc, NULL, NULL, NULL);
    free(Q);
}

/* single_time_step_UV (Source/RK_time_step.cu:266-278). */
static void single_time_step_uv(const or_params *p, const double *ux, const double *uy, const double *uz,
                                double *vx, double *vy, double *vz)
{
    const long nc = ncomplex(p);
    double *Q = (double *)malloc(sizeof(cx) * (size_t)nc * 3);
    or_bilinear(p, ux, uy, uz, vx, vy, vz, Q, Q + 2 * nc, Q + 4 * nc);
    or_update(p, OR_UPD_UV, vx, vy, vz, NULL, NULL, NULL, Q, Q + 2 * nc, Q + 4 * nc, NULL, NULL, NULL);
    free(Q);
}

/* RK3_SSP_UV (Source/RK_time_step.cu:342-355): Heun step on v around the
 * moving base state; the base state u is advanced in place by one RK3 step (Q10). */
void or_rk3_ssp_uv(const or_params *p, double *ux, double *uy, double *uz,
                   double *vx, double *vy, double *vz,
                   const double *fx, const double *fy, const double *fz)
{
    const long nc = ncomplex(p);
    double *v1 = (double *)malloc(sizeof(cx) * (size_t)nc * 3);
    double *v[3] = {vx, vy, vz};
    for (int c = 0; c < 3; c++) memcpy(v1 + 2 * c * nc, v[c], sizeof(cx) * (size_t)nc);
    single_time_step_uv(p, ux, uy, uz, v1, v1 + 2 * nc, v1 + 4 * nc);
    or_rk3_ssp(p, ux, uy, uz, fx, fy, fz);
    single_time_step_uv(p, ux, uy, uz, v1, v1 + 2 * nc, v1 + 4 * nc);
    for (int c = 0; c < 3; c++) axpby(nc, 1.0 / 2.0, v[c], 1.0 / 2.0, v1 + 2 * c * nc, v[c]);
    free(v1);
}

/* solve_divergence (Source/divergence.cu:8-30): div^ = i (k . u^) / dt. */
void or_divergence(const or_params *p, const double *ux_, const double *uy_, const double *uz_, double *div_)
{
    tables_t t;
    tables_make(p, &t);
    const cx *ux = (const cx *)ux_, *uy = (const cx *)uy_, *uz = (const cx *)uz_;
    cx *div = (cx *)div_;
    const int Ny = t.Ny, Mz = t.Mz;
    const double dt = p->dt;
    for (long id = 0; id < t.nc; id++) {
        const int l = (int)(id % Mz), k = (int)((id / Mz) % Ny), j = (int)(id / ((long)Mz * Ny));
        const double re = (-cimag(ux[id]) * t.kx[j] - cimag(uy[id]) * t.ky[k] - cimag(uz[id]) * t.kz[l]) / dt;
        const double im = (creal(ux[id]) * t.kx[j] + creal(uy[id]) * t.ky[k] + creal(uz[id]) * t.kz[l]) / dt;
        div[id] = re + im * I;
    }
    tables_free(&t);
}

/* ------------------------------------------------------------------------ */
/* WENO (Source/advection_WENO.cu:5-392)                                     */
/* ------------------------------------------------------------------------ */
static double weno5(double f1, double f2, double f3, double f4, double f5)
{
    const double eps = 1.0E-22;
    const double q0 = 1.0 / 3.0 * f1 - 7.0 / 6.0 * f2 + 11.0 / 6.0 * f3;
    const double q1 = -1.0 / 6.0 * f2 + 5.0 / 6.0 * f3 + 1.0 / 3.0 * f4;
    const double q2 = 1.0 / 3.0 * f3 + 5.0 / 6.0 * f4 - 1.0 / 6.0 * f5;
    const double b0 = 13.0 / 12.0 * (f1 - 2.0 * f2 + f3) * (f1 - 2.0 * f2 + f3) + 0.25 * (f1 - 4.0 * f2 + 3.0 * f3) * (f1 - 4.0 * f2 + 3.0 * f3);
    const double b1 = 13.0 / 12.0 * (f2 - 2.0 * f3 + f4) * (f2 - 2.0 * f3 + f4) + 0.25 * (f2 - f4) * (f2 - f4);
    const double b2 = 13.0 / 12.0 * (f3 - 2.0 * f4 + f5) * (f3 - 2.0 * f4 + f5) + 0.25 * (3.0 * f3 - 4.0 * f4 + f5) * (3.0 * f3 - 4.0 * f4 + f5);
    const double a0 = 0.1 / ((eps + b0) * (eps + b0));
    const double a1 = 0.6 / ((eps + b1) * (eps + b1));
    const double a2 = 0.3 / ((eps + b2) * (eps + b2));
    const double s = a0 + a1 + a2;
    return a0 / s * q0 + a1 / s * q1 + a2 / s * q2;
}

static double weno3(double f1, double f2, double f3)
{
    const double eps = 1.0E-22;
    const double q0 = -0.5 * f1 + 1.5 * f2, q1 = 0.5 * f2 + 0.5 * f3;
    const double b0 = (f2 - f1) * (f2 - f1), b1 = (f3 - f2) * (f3 - f2);
    const double a0 = (1.0 / 3.0) / ((eps + b0) * (eps + b0));
    const double a1 = (2.0 / 3.0) / ((eps + b1) * (eps + b1));
    const double s = a0 + a1;
    return a0 / s * q0 + a1 / s * q1;
}

static inline int wrap(int i, int n) { return i > n - 1 ? i - n : i < 0 ? n + i : i; }

/* Value of x along axis `ax` at offset o from (j,k,l), periodic (I3 macro, Source/Macros.h:19). */
static inline double at(const double *x, int Nx, int Ny, int Nz, int j, int k, int l, int ax, int o)
{
    if (ax == 0) j = wrap(j + o, Nx);
    else if (ax == 1) k = wrap(k + o, Ny);
    else l = wrap(l + o, Nz);
    return x[IDX3(j, k, l, Ny, Nz)];
}

/* flux_{x,y,z}_weno_L / _R (:105-301) at cell (j,k,l) along axis ax. */
static double face(const or_params *p, const double *x, int j, int k, int l, int ax, int right, int scheme)
{
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
    if (scheme == 3) {
        const double f1 = at(x, Nx, Ny, Nz, j, k, l, ax, 1), f2 = at(x, Nx, Ny, Nz, j, k, l, ax, 0), f3 = at(x, Nx, Ny, Nz, j, k, l, ax, -1);
        return right ? weno3(f1, f2, f3) : weno3(f3, f2, f1);
    }
    if (scheme == 5) {
        const double f1 = at(x, Nx, Ny, Nz, j, k, l, ax, 2), f2 = at(x, Nx, Ny, Nz, j, k, l, ax, 1);
        const double f3 = at(x, Nx, Ny, Nz, j, k, l, ax, 0), f4 = at(x, Nx, Ny, Nz, j, k, l, ax, -1);
        /* Q9: flux_z_weno_R reads l-1 for its fifth point (:288). */
        const int o5 = (right && ax == 2 && !p->weno_zfix) ? -1 : -2;
        const double f5 = at(x, Nx, Ny, Nz, j, k, l, ax, o5);
        return right ? weno5(f1, f2, f3, f4, f5) : weno5(f5, f4, f3, f2, f1);
    }
    return at(x, Nx, Ny, Nz, j, k, l, ax, 0);
}

/* flux_device (:304-360): upwind faces chosen by the sign of the face-averaged velocity. */
void or_weno_advect(const or_params *p, int scheme, const double *ux, const double *uy, const double *uz,
                    const double *x, double *out)
{
    set_threads(p);
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
    const double dx = p->Lx / Nx, dy = p->Ly / Ny, dz = p->Lz / Nz;
    #pragma omp parallel for collapse(2) schedule(static)
    for (int j = 0; j < Nx; j++)
        for (int k = 0; k < Ny; k++)
            for (int l = 0; l < Nz; l++) {
                const long id = IDX3(j, k, l, Ny, Nz);
                const int cellv = or_mut == 4;
                double Uw = 0.5 * (at(ux, Nx, Ny, Nz, j, k, l, 0, -1) + at(ux, Nx, Ny, Nz, j, k, l, 0, 0));
                double Ue = 0.5 * (at(ux, Nx, Ny, Nz, j, k, l, 0, 0) + at(ux, Nx, Ny, Nz, j, k, l, 0, 1));
                double Un = 0.5 * (at(uy, Nx, Ny, Nz, j, k, l, 1, -1) + at(uy, Nx, Ny, Nz, j, k, l, 1, 0));
                double Us = 0.5 * (at(uy, Nx, Ny, Nz, j, k, l, 1, 0) + at(uy, Nx, Ny, Nz, j, k, l, 1, 1));
                double Ur = 0.5 * (at(uz, Nx, Ny, Nz, j, k, l, 2, -1) + at(uz, Nx, Ny, Nz, j, k, l, 2, 0));
                double Uf = 0.5 * (at(uz, Nx, Ny, Nz, j, k, l, 2, 0) + at(uz, Nx, Ny, Nz, j, k, l, 2, 1));
                if (cellv) { Uw = Ue = ux[id]; Un = Us = uy[id]; Ur = Uf = uz[id]; }
                const double gx = or_mut == 1 ? -1.0 : 1.0, gy = or_mut == 2 ? -1.0 : 1.0, gz = or_mut == 3 ? -1.0 : 1.0;
                Uw *= gx; Ue *= gx; Un *= gy; Us *= gy; Ur *= gz; Uf *= gz;
                const double sw = Uw >= 0.0 ? face(p, x, wrap(j - 1, Nx), k, l, 0, 0, scheme) : face(p, x, j, k, l, 0, 1, scheme);
                const double se = Ue >= 0.0 ? face(p, x, j, k, l, 0, 0, scheme) : face(p, x, wrap(j + 1, Nx), k, l, 0, 1, scheme);
                const double sn = Un >= 0.0 ? face(p, x, j, wrap(k - 1, Ny), l, 1, 0, scheme) : face(p, x, j, k, l, 1, 1, scheme);
                const double ss = Us >= 0.0 ? face(p, x, j, k, l, 1, 0, scheme) : face(p, x, j, wrap(k + 1, Ny), l, 1, 1, scheme);
                const double sr = Ur >= 0.0 ? face(p, x, j, k, wrap(l - 1, Nz), 2, 0, scheme) : face(p, x, j, k, l, 2, 1, scheme);
                const double sf = Uf >= 0.0 ? face(p, x, j, k, l, 2, 0, scheme) : face(p, x, j, k, wrap(l + 1, Nz), 2, 1, scheme);
                out[id] = ux[id] * (se - sw) / dx + uy[id] * (ss - sn) / dy + uz[id] * (sf - sr) / dz;
            }
}

/* WenoAdvection (:397-415): velocities to physical space, advect each velocity component, back to Fourier. */
void or_weno_advection_spectral(const or_params *p, int scheme,
                                const double *ux_hat, const double *uy_hat, const double *uz_hat,
                                double *Qx_hat, double *Qy_hat, double *Qz_hat)
{
    const long nr = (long)p->Nx * p->Ny * p->Nz;
    double *u = (double *)malloc(sizeof(double) * (size_t)nr * 4);
    or_fft_c2r(p, ux_hat, u, 1);
    or_fft_c2r(p, uy_hat, u + nr, 1);
    or_fft_c2r(p, uz_hat, u + 2 * nr, 1);
    double *Q[3] = {Qx_hat, Qy_hat, Qz_hat};
    for (int c = 0; c < 3; c++) {
        or_weno_advect(p, scheme, u, u + nr, u + 2 * nr, u + c * nr, u + 3 * nr);
        or_fft_r2c(p, u + 3 * nr, Q[c]);
    }
    free(u);
}

/* RK3_SSP (Source/RK_time_step.cu:310-327) with a passive scalar advected by the
 * WENO kernel (advection_WENO.cu:304-392, scalar `x` argument) using each stage's
 * 2/3-masked physical velocity C2R(M w^_s) and the same SSP weights. */
static void scalar_stage(const or_params *p, int scheme, const double *w[3], const double *th_s,
                         const double *th_n, double a0, double a1, double *out)
{
    tables_t t;
    tables_make(p, &t);
    const long nr = t.nr;
    cx *sp = (cx *)malloc(sizeof(cx) * (size_t)t.nc);
    double *up = (double *)malloc(sizeof(double) * (size_t)nr * 4);
    for (int b = 0; b < 3; b++) {
        masked_field_and_derivs(&t, (const cx *)w[b], sp, NULL, NULL, NULL);
        or_fft_c2r(p, (const double *)sp, up + b * nr, 1);
    }
    or_weno_advect(p, scheme, up, up + nr, up + 2 * nr, th_s, up + 3 * nr);
    for (long i = 0; i < nr; i++) {
        const double e = th_s[i] - p->dt * up[3 * nr + i];
        out[i] = th_n ? a0 * th_n[i] + a1 * e : a1 * e;
    }
    free(sp); free(up);
    tables_free(&t);
}

void or_rk3_ssp_scalar(const or_params *p, double *ux, double *uy, double *uz,
                       const double *fx, const double *fy, const double *fz, double *theta, int scheme)
{
    const long nc = ncomplex(p), nr = (long)p->Nx * p->Ny * p->Nz;
    double *th1 = (double *)malloc(sizeof(double) * (size_t)nr * 2), *th2 = th1 + nr;
    double *u1 = (double *)malloc(sizeof(cx) * (size_t)nc * 3);
    double *u2 = (double *)malloc(sizeof(cx) * (size_t)nc * 3);
    double *u[3] = {ux, uy, uz};
    const double *w0[3] = {ux, uy, uz};
    scalar_stage(p, scheme, w0, theta, NULL, 0.0, 1.0, th1);
    for (int c = 0; c < 3; c++) memcpy(u1 + 2 * c * nc, u[c], sizeof(cx) * (size_t)nc);
    or_single_time_step(p, u1, u1 + 2 * nc, u1 + 4 * nc, fx, fy, fz);
    const double *w1[3] = {u1, u1 + 2 * nc, u1 + 4 * nc};
    if (or_mut == 6) scalar_stage(p, scheme, or_mut == 5 ? w0 : w1, th1, theta, 1.0 / 4.0, 3.0 / 4.0, th2);
    else scalar_stage(p, scheme, or_mut == 5 ? w0 : w1, th1, theta, 3.0 / 4.0, 1.0 / 4.0, th2);
    or_single_time_step(p, u1, u1 + 2 * nc, u1 + 4 * nc, fx, fy, fz);
    for (int c = 0; c < 3; c++) axpby(nc, 3.0 / 4.0, u[c], 1.0 / 4.0, u1 + 2 * c * nc, u2 + 2 * c * nc);
    const double *w2[3] = {u2, u2 + 2 * nc, u2 + 4 * nc};
    scalar_stage(p, scheme, or_mut == 5 ? w0 : w2, th2, theta, 1.0 / 3.0, 2.0 / 3.0, th1);
    or_single_time_step(p, u2, u2 + 2 * nc, u2 + 4 * nc, fx, fy, fz);
    for (int c = 0; c < 3; c++) axpby(nc, 1.0 / 3.0, u[c], 2.0 / 3.0, u2 + 2 * c * nc, u[c]);
    memcpy(theta, th1, sizeof(double) * (size_t)nr);
    free(th1); free(u1); free(u2);
}

/* ------------------------------------------------------------------------ */
/* Reductions (Source/min_max_reduction.cu)                                  */
/* ------------------------------------------------------------------------ */
/* compute_reduction (:199-227): min and max seeded with the integer macro
 * inf = 0x7f800000 (Q8, min_max_reduction.h:16), i.e. +-2139095040.0;
 * comparisons `v < m ? v : m` skip NaN. */
void or_reduce_minmax(long n, const double *in, double *out2)
{
    double mn = (double)0x7f800000, mx = -(double)0x7f800000;
    for (long i = 0; i < n; i++) {
        const double v = in[i];
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
    }
    out2[0] = mn;
    out2[1] = mx;
}

/* reduction_sum (:479-509): plain sum; long double accumulation here. */
double or_reduce_sum(long n, const double *in)
{
    long double s = 0.0L;
    for (long i = 0; i < n; i++) s += in[i];
    return (double)s;
}

/* get_kinetic_energy_device (Source/math_support.cu:720-751): transverse energy 0.5(uy^2+uz^2) dx dy (Q7). */
double or_kinetic_energy(const or_params *p, const double *ux, const double *uy, const double *uz)
{
    (void)ux;
    const long nr = (long)p->Nx * p->Ny * p->Nz;
    const double dx = p->Lx / p->Nx, dy = p->Ly / p->Ny;
    long double s = 0.0L;
    for (long i = 0; i < nr; i++) s += 0.5 * (uy[i] * uy[i] + uz[i] * uz[i]) * dx * dy;
    return (double)s;
}

/* get_dissipation_device (Source/math_support.cu:755-796): undivided central differences (Q7). */
double or_dissipation(const or_params *p, const double *ux, const double *uy, const double *uz)
{
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
    const double dx = p->Lx / Nx, dy = p->Ly / Ny, dz = p->Lz / Nz;
    long double s = 0.0L;
    for (int j = 0; j < Nx; j++)
        for (int k = 0; k < Ny; k++)
            for (int l = 0; l < Nz; l++) {
#define D(f, ax) ((at(f, Nx, Ny, Nz, j, k, l, ax, 1) - at(f, Nx, Ny, Nz, j, k, l, ax, -1)) / (ax == 0 ? dx : ax == 1 ? dy : dz))
                const double ux_x = D(ux, 0), ux_y = D(ux, 1), ux_z = D(ux, 2);
                const double uy_x = D(uy, 0), uy_y = D(uy, 1), uy_z = D(uy, 2);
                const double uz_x = D(uz, 0), uz_y = D(uz, 1), uz_z = D(uz, 2);
#undef D
                s += dx * dy * dz * (2.0 * (ux_x * ux_x + uy_y * uy_y + uz_z * uz_z) + (uy_z + uz_y) * (uy_z + uz_y) +
                                     (ux_z + uz_x) * (ux_z + uz_x) + (ux_y + uy_x) * (ux_y + uy_x));
            }
    return (double)s;
}

/* calculate_CFL_device + calc_dt (Source/cuda_supp.cu:170-238): min over points of
 * CFL / sum_a (1e-6+|u_a|)/d_a, through compute_reduction (out[0] = min). */
double or_cfl_dt(const or_params *p, double cfl, const double *ux, const double *uy, const double *uz)
{
    const long nr = (long)p->Nx * p->Ny * p->Nz;
    const double dx = p->Lx / p->Nx, dy = p->Ly / p->Ny, dz = p->Lz / p->Nz;
    double *c = (double *)malloc(sizeof(double) * (size_t)nr);
    for (long i = 0; i < nr; i++)
        c[i] = cfl / ((1.0e-6 + fabs(ux[i])) / dx + (1.0e-6 + fabs(uy[i])) / dy + (1.0e-6 + fabs(uz[i])) / dz);
    double mm[2];
    or_reduce_minmax(nr, c, mm);
    free(c);
    return mm[0];
}

/* Shapiro_test_case (Source/Shapiro_test.cu:8-50): L-inf (max via compute_reduction, ret[1])
 * of the pointwise error against the Beltrami solution decaying as exp(-3t/Re). */
double or_shapiro_error(const or_params *p, double t, const double *ux, const double *uy, const double *uz)
{
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
    const double dx = p->Lx / Nx, dy = p->Ly / Ny, dz = p->Lz / Nz, Re = p->Re;
    const long nr = (long)Nx * Ny * Nz;
    double *e = (double *)malloc(sizeof(double) * (size_t)nr);
    for (int j = 0; j < Nx; j++)
        for (int k = 0; k < Ny; k++)
            for (int l = 0; l < Nz; l++) {
                const double y = k * dy, x = j * dx, z = l * dz, f = sqrt(3.0);
                const double dec = exp(-(f * f) * t / Re);
                const long id = IDX3(j, k, l, Ny, Nz);
                const double ex = ux[id] - (-0.5 * (f * cos(x) * sin(y) * sin(z) + sin(x) * cos(y) * cos(z)) * dec);
                const double ey = uy[id] - (0.5 * (f * sin(x) * cos(y) * sin(z) - cos(x) * sin(y) * cos(z)) * dec);
                const double ez = uz[id] - (cos(x) * cos(y) * sin(z) * dec);
                e[id] = sqrt(ex * ex + ey * ey + ez * ez);
            }
    double mm[2];
    or_reduce_minmax(nr, e, mm);
    free(e);
    return mm[1];
}

/* ------------------------------------------------------------------------ */
/* Krylov packing (Source/NS3D_periodic.cu:173-250, 340-361)                 */
/* ------------------------------------------------------------------------ */
void or_unpack_reduced(const or_params *p, const double *vec, double *vx, double *vy, double *vz)
{
    const long N = ncomplex(p);
    double *v[3] = {vx, vy, vz};
    for (int c = 0; c < 3; c++) {
        v[c][0] = 0.0; v[c][1] = 0.0;
        for (long i = 0; i < N - 1; i++) { v[c][2 * (i + 1)] = 0.0; v[c][2 * (i + 1) + 1] = vec[c * (N - 1) + i]; }
    }
}

void or_pack_reduced(const or_params *p, const double *vx, const double *vy, const double *vz, double *vec)
{
    const long N = ncomplex(p);
    const double *v[3] = {vx, vy, vz};
    for (int c = 0; c < 3; c++)
        for (long i = 0; i < N - 1; i++) vec[c * (N - 1) + i] = v[c][2 * (i + 1) + 1];
}

void or_jv_reduced(const or_params *p, const double *ux, const double *uy, const double *uz,
                   const double *vec_in, double *vec_out)
{
    const long nc = ncomplex(p);
    double *v = (double *)malloc(sizeof(cx) * (size_t)nc * 3);
    or_unpack_reduced(p, vec_in, v, v + 2 * nc, v + 4 * nc);
    or_rk3_ssp_uv_rhs(p, ux, uy, uz, v, v + 2 * nc, v + 4 * nc);
    or_pack_reduced(p, v, v + 2 * nc, v + 4 * nc, vec_out);
    free(v);
}

/* velocities_from_A_vector_kernel (Source/NS3D_periodic.cu:111-125): full-length
 * vectors, v_c[i] = i * vec[c*N + i] for every i < N. */
void or_unpack_full(const or_params *p, const double *vec, double *vx, double *vy, double *vz)
{
    const long N = ncomplex(p);
    double *v[3] = {vx, vy, vz};
    for (int c = 0; c < 3; c++)
        for (long i = 0; i < N; i++) {
            v[c][2 * i] = 0.0;
            v[c][2 * i + 1] = vec[c * N + i];
        }
}

/* A_vector_from_velocities_kernel (:223-236): vec[c*N + i] = Im v_c[i] for
 * 0 < i < N; the three i = 0 entries are left untouched. */
void or_pack_full(const or_params *p, const double *vx, const double *vy, const double *vz, double *vec)
{
    const long N = ncomplex(p);
    const double *v[3] = {vx, vy, vz};
    for (int c = 0; c < 3; c++)
        for (long i = 1; i < N; i++) vec[c * N + i] = v[c][2 * i + 1];
}

/* NSCallMatrixVector_reduced_Newton_Globalization (:364-389): the reduced J.v
 * followed by A_vector_add_diagonal_kernel (:296-310), out -= vec_in / tau. */
void or_jv_reduced_newton(const or_params *p, const double *ux, const double *uy, const double *uz,
                          double tau, const double *vec_in, double *vec_out)
{
    const long n = 3 * (ncomplex(p) - 1);
    or_jv_reduced(p, ux, uy, uz, vec_in, vec_out);
    const double one_over_tau = 1.0 / tau;
    for (long i = 0; i < n; i++) vec_out[i] = vec_out[i] - vec_in[i] * one_over_tau;
}

/* NSCallMatrixVector_exponential (:444-470): unpack the full-length vector,
 * take `timesteps_period` tangent-map steps RK3_SSP_UV (which advance the base
 * flow u too, in place), pack back with index 0 of each component untouched. */
void or_jv_exponential(const or_params *p, double *ux, double *uy, double *uz,
                       const double *fx, const double *fy, const double *fz, int timesteps_period,
                       const double *vec_in, double *vec_out)
{
    const long nc = ncomplex(p);
    double *v = (double *)malloc(sizeof(cx) * (size_t)nc * 3);
    or_unpack_full(p, vec_in, v, v + 2 * nc, v + 4 * nc);
    for (int t = 0; t < timesteps_period; t++)
        or_rk3_ssp_uv(p, ux, uy, uz, v, v + 2 * nc, v + 4 * nc, fx, fy, fz);
    or_pack_full(p, v, v + 2 * nc, v + 4 * nc, vec_out);
    free(v);
}
