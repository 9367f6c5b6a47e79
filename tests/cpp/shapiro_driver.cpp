// The Shapiro Beltrami flow (32^3, Re = 20) driven from host C++ through the
// reference-named wrappers of ns3d_compat.hpp: the drop-in as a C++ caller of
// the reference would use it (device arrays from the HIP runtime, the
// reference's FFTN_Device / RK3_SSP / iFFTN_Device calls).
// Prints the L-inf error against the exact solution after 100 steps.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <vector>

#include "ns3d_compat.hpp"

#define CHECK(x)                                                   \
    do {                                                           \
        int s_ = (int)(x);                                         \
        if (s_) {                                                  \
            std::fprintf(stderr, "%s failed: %d\n", #x, s_);       \
            return 1;                                              \
        }                                                          \
    } while (0)

int main() {
    const int N = 32, M = N / 2 + 1, steps = 100;
    const double Re = 20.0, dt = 1e-3, pi = 3.14159265358979323846;
    ns3d_config cfg = {};
    cfg.Nx = cfg.Ny = cfg.Nz = N;
    cfg.Lx = cfg.Ly = cfg.Lz = 2.0 * pi;
    cfg.Re = Re;
    cfg.dt = dt;
    cfg.precision = NS3D_FP64;
    cfg.flags = 0;
    cfg.stream = nullptr;
    cfg.device = 0;
    ns3d_ctx* ctx = nullptr;
    CHECK(ns3d_create(&cfg, &ctx));
    ns3d_compat::bind(ctx);

    // Shapiro_test.cu:26-31 at t = 0 on x_j = j*dx (NS3D_periodic.cu:810-829)
    const size_t nr = (size_t)N * N * N, nc = (size_t)N * N * M;
    std::vector<double> h(3 * nr);
    const double s3 = std::sqrt(3.0), dx = 2.0 * pi / N;
    for (int j = 0; j < N; j++)
        for (int k = 0; k < N; k++)
            for (int l = 0; l < N; l++) {
                const double x = j * dx, y = k * dx, z = l * dx;
                const size_t i = ((size_t)j * N + k) * N + l;
                h[i] = -0.5 * (s3 * std::cos(x) * std::sin(y) * std::sin(z) + std::sin(x) * std::cos(y) * std::cos(z));
                h[nr + i] = 0.5 * (s3 * std::sin(x) * std::cos(y) * std::sin(z) - std::cos(x) * std::sin(y) * std::cos(z));
                h[2 * nr + i] = std::cos(x) * std::cos(y) * std::sin(z);
            }
    double* phys = nullptr;
    ns3d_cplx* spec = nullptr;
    CHECK(hipMalloc((void**)&phys, sizeof(double) * 3 * nr));
    CHECK(hipMalloc((void**)&spec, sizeof(ns3d_cplx) * 3 * nc));
    CHECK(hipMemcpy(phys, h.data(), sizeof(double) * 3 * nr, hipMemcpyHostToDevice));
    ns3d_cplx *ux = spec, *uy = spec + nc, *uz = spec + 2 * nc;
    for (int b = 0; b < 3; b++) CHECK(ns3d_compat::FFTN_Device(phys + b * nr, spec + b * nc));
    for (int t = 0; t < steps; t++) CHECK(ns3d_compat::RK3_SSP(dt, Re, ux, uy, uz, nullptr, nullptr, nullptr));
    for (int b = 0; b < 3; b++) CHECK(ns3d_compat::iFFTN_Device_scaled(spec + b * nc, phys + b * nr));
    double err = -1.0;
    CHECK(ns3d_shapiro_error(ctx, steps * dt, phys, phys + nr, phys + 2 * nr, &err));
    std::printf("shapiro_linf %.17e\n", err);
    CHECK(hipFree(phys));
    CHECK(hipFree(spec));
    CHECK(ns3d_destroy(ctx));
    return 0;
}
