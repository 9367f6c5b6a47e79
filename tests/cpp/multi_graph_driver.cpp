// Several contexts of different grid sizes in one HIP graph, from host C++
// through the plain C ABI (ns3d.h): the RK3 steps of a 32^3 and a 16^3
// context are captured from two streams forked from one capture stream and
// joined back, so the graph has two concurrent branches.  Five replays of
// K steps, then the same 5*K steps eagerly on two other contexts; the states
// must agree bit for bit, and both must follow the Shapiro solution.
// Prints "bitwise_equal <0|1>" and "shapiro_linf <N> <error>" lines.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "ns3d.h"

#define CHECK(x)                                                   \
    do {                                                           \
        int s_ = (int)(x);                                         \
        if (s_) {                                                  \
            std::fprintf(stderr, "%s failed: %d\n", #x, s_);       \
            return 1;                                              \
        }                                                          \
    } while (0)

static const double kPi = 3.14159265358979323846;
static const double kRe = 20.0, kDt = 1e-3;

// Shapiro_test.cu:26-31 at t = 0 on x_j = j*dx
static std::vector<double> shapiro(int N) {
    const size_t nr = (size_t)N * N * N;
    std::vector<double> h(3 * nr);
    const double s3 = std::sqrt(3.0), dx = 2.0 * kPi / N;
    for (int j = 0; j < N; j++)
        for (int k = 0; k < N; k++)
            for (int l = 0; l < N; l++) {
                const double x = j * dx, y = k * dx, z = l * dx;
                const size_t i = ((size_t)j * N + k) * N + l;
                h[i] = -0.5 * (s3 * std::cos(x) * std::sin(y) * std::sin(z) + std::sin(x) * std::cos(y) * std::cos(z));
                h[nr + i] = 0.5 * (s3 * std::sin(x) * std::cos(y) * std::sin(z) - std::cos(x) * std::sin(y) * std::cos(z));
                h[2 * nr + i] = std::cos(x) * std::cos(y) * std::sin(z);
            }
    return h;
}

struct Field {
    int N = 0;
    size_t nr = 0, nc = 0;
    ns3d_ctx* ctx = nullptr;
    ns3d_cplx* u = nullptr;  // 3 spectral components back to back
    ns3d_cplx* c(int b) const { return u + b * nc; }
};

static int make(Field& f, int N, hipStream_t s, const double* phys_init) {
    f.N = N;
    f.nr = (size_t)N * N * N;
    f.nc = (size_t)N * N * (N / 2 + 1);
    ns3d_config cfg = {};
    cfg.Nx = cfg.Ny = cfg.Nz = N;
    cfg.Lx = cfg.Ly = cfg.Lz = 2.0 * kPi;
    cfg.Re = kRe;
    cfg.dt = kDt;
    cfg.precision = NS3D_FP64;
    cfg.flags = 0;
    cfg.stream = s;
    cfg.device = 0;
    CHECK(ns3d_create(&cfg, &f.ctx));
    CHECK(hipMalloc((void**)&f.u, sizeof(ns3d_cplx) * 3 * f.nc));
    for (int b = 0; b < 3; b++) CHECK(ns3d_fft_r2c(f.ctx, phys_init + b * f.nr, f.c(b)));
    CHECK(ns3d_synchronize(f.ctx));
    return 0;
}

int main() {
    const int sizes[2] = {32, 16};
    const int K = 10, replays = 5;
    hipStream_t cap, br[2], eager;
    CHECK(hipStreamCreate(&cap));
    CHECK(hipStreamCreate(&br[0]));
    CHECK(hipStreamCreate(&br[1]));
    CHECK(hipStreamCreate(&eager));
    Field g[2], e[2];
    double* phys[2] = {nullptr, nullptr};
    for (int i = 0; i < 2; i++) {
        const int N = sizes[i];
        std::vector<double> h = shapiro(N);
        CHECK(hipMalloc((void**)&phys[i], sizeof(double) * h.size()));
        CHECK(hipMemcpy(phys[i], h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
        CHECK(make(g[i], N, br[i], phys[i]));
        CHECK(make(e[i], N, eager, phys[i]));
        // one step on a copy first, so every buffer the step uses exists before capture
        ns3d_cplx* tmp = nullptr;
        CHECK(hipMalloc((void**)&tmp, sizeof(ns3d_cplx) * 3 * g[i].nc));
        CHECK(hipMemcpy(tmp, g[i].u, sizeof(ns3d_cplx) * 3 * g[i].nc, hipMemcpyDeviceToDevice));
        CHECK(ns3d_rk3_ssp_steps(g[i].ctx, 1, tmp, tmp + g[i].nc, tmp + 2 * g[i].nc, nullptr, nullptr, nullptr));
        CHECK(hipDeviceSynchronize());
        CHECK(ns3d_rk3_ssp_steps(e[i].ctx, 1, tmp, tmp + g[i].nc, tmp + 2 * g[i].nc, nullptr, nullptr, nullptr));
        CHECK(hipDeviceSynchronize());
        CHECK(hipFree(tmp));
    }

    hipEvent_t fork, join[2];
    CHECK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    CHECK(hipEventCreateWithFlags(&join[0], hipEventDisableTiming));
    CHECK(hipEventCreateWithFlags(&join[1], hipEventDisableTiming));
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    CHECK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
    CHECK(hipEventRecord(fork, cap));
    for (int i = 0; i < 2; i++) {
        CHECK(hipStreamWaitEvent(br[i], fork, 0));
        CHECK(ns3d_rk3_ssp_steps(g[i].ctx, K, g[i].c(0), g[i].c(1), g[i].c(2), nullptr, nullptr, nullptr));
        CHECK(hipEventRecord(join[i], br[i]));
        CHECK(hipStreamWaitEvent(cap, join[i], 0));
    }
    CHECK(hipStreamEndCapture(cap, &graph));
    CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));

    for (int r = 0; r < replays; r++) CHECK(hipGraphLaunch(exec, cap));
    CHECK(hipStreamSynchronize(cap));
    for (int i = 0; i < 2; i++)
        CHECK(ns3d_rk3_ssp_steps(e[i].ctx, K * replays, e[i].c(0), e[i].c(1), e[i].c(2), nullptr, nullptr, nullptr));
    CHECK(hipStreamSynchronize(eager));

    int equal = 1;
    for (int i = 0; i < 2; i++) {
        std::vector<ns3d_cplx> a(3 * g[i].nc), b(3 * g[i].nc);
        CHECK(hipMemcpy(a.data(), g[i].u, sizeof(ns3d_cplx) * a.size(), hipMemcpyDeviceToHost));
        CHECK(hipMemcpy(b.data(), e[i].u, sizeof(ns3d_cplx) * b.size(), hipMemcpyDeviceToHost));
        if (std::memcmp(a.data(), b.data(), sizeof(ns3d_cplx) * a.size()) != 0) equal = 0;
    }
    std::printf("bitwise_equal %d\n", equal);
    for (int i = 0; i < 2; i++) {
        for (int b = 0; b < 3; b++) CHECK(ns3d_fft_c2r(g[i].ctx, g[i].c(b), phys[i] + b * g[i].nr));
        double err = -1.0;
        CHECK(ns3d_shapiro_error(g[i].ctx, K * replays * kDt, phys[i], phys[i] + g[i].nr, phys[i] + 2 * g[i].nr,
                                 &err));
        std::printf("shapiro_linf %d %.17e\n", g[i].N, err);
    }

    CHECK(hipGraphExecDestroy(exec));
    CHECK(hipGraphDestroy(graph));
    CHECK(hipEventDestroy(fork));
    CHECK(hipEventDestroy(join[0]));
    CHECK(hipEventDestroy(join[1]));
    for (int i = 0; i < 2; i++) {
        CHECK(hipFree(g[i].u));
        CHECK(hipFree(e[i].u));
        CHECK(hipFree(phys[i]));
        CHECK(ns3d_destroy(g[i].ctx));
        CHECK(ns3d_destroy(e[i].ctx));
    }
    for (int i = 0; i < 2; i++) CHECK(hipStreamDestroy(br[i]));
    CHECK(hipStreamDestroy(cap));
    CHECK(hipStreamDestroy(eager));
    return 0;
}
