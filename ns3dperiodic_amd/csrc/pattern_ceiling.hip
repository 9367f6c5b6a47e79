// pattern_ceiling: the HBM bandwidth an access pattern of the fused RK3 passes
// allows on this GPU, measured with kernels that only move the bytes.
//
// The five passes read and write spectral fields (N x N x (N/2+1) complex
// modes, rows padded to Mi modes: 136 at N = 256) along three
// kinds of pencils:
//   x   the x-inverse and x-forward passes: P consecutive z-modes l of one
//       (k) row per block, every x-plane j (planes N*Mi modes apart);
//   y   the y-inverse and y-forward passes: the same tile along y (rows Mi
//       modes apart);
//   z   the z pass: whole z-rows, each lane reading its spectral elements in
//       the Hermitian-extended order of the paired C2R transform (index n and
//       N - n), truncated to the planes l <= lmax the 2/3 mask keeps.
// Each kernel here copies (or sums) nin fields into nout fields with those
// tiles and nothing else: the rate of moving the same bytes the same way with
// no transform in between, a reference for the pass's memory traffic.  (It is
// not a hard ceiling: a pass reads what the previous pass has just written,
// part of which the Infinity Cache still holds, while these kernels read
// fields written long before.)  "flat" is a plain contiguous copy of the same
// bytes.
//
// Usage: pattern_ceiling [N] [fp64|fp32] [reps]  -> one JSON line per pattern
//   {"pattern": ..., "precision": ..., "N": ..., "read_fields": ..., "write_fields": ...,
//    "bytes": ..., "ms": ..., "GBps": ...}
// bytes count the modes read and written (truncated reads count what they read).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

#define CK(x)                                                                                  \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) {                                                                \
            fprintf(stderr, "pattern_ceiling: HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            exit(1);                                                                           \
        }                                                                                      \
    } while (0)

template <typename T>
struct alignas(2 * sizeof(T)) c2 {
    T x, y;
};

// block b -> logical tile: consecutive tiles on one XCD (the passes' order)
__device__ __forceinline__ int tile_of_block() {
    const int b = blockIdx.x, g = gridDim.x;
    return (b & 7) * (g >> 3) + (b >> 3);
}

// Geometry of N^3 fields in the passes' scratch layout, all compile-time (a
// reference kernel should be the best code for its pattern: runtime strides
// and bounds cost these copies up to 12%).
constexpr int lmax_of(int mz) {
    // largest l whose z-plane keeps a mode under the reference's 2/3 mask with
    // (Nx, Ny, Mz) normalisation at kx = ky = 0: (l / Mz)^2 < 4/9
    int l = 0;
    while (9 * (l + 1) * (l + 1) < 4 * mz * mz) l++;
    return l;
}
template <typename T, int NX>
struct Geo {
    static constexpr int N = NX, Mz = NX / 2 + 1;
    static constexpr int pad = 8;
    static constexpr int Mi = (Mz + pad - 1) / pad * pad;  // rows padded as the passes' scratch
    static constexpr int lmax = lmax_of(Mz);
    static constexpr long nc = (long)NX * NX * Mi;          // modes per field
};

// x (AXIS 0) or y (AXIS 1) pencils: block = P z-modes x (N/E) threads, E elements per thread
template <typename T, int AXIS, int P, int E, int NI, int NO, int NX>
__global__ void __launch_bounds__(P * (NX / E)) k_pencil(const c2<T>* __restrict__ in, c2<T>* __restrict__ out, int nb) {
    using G = Geo<T, NX>;
    constexpr int TPP = NX / E, NLT = (G::Mz + P - 1) / P;
    const int bid = tile_of_block();
    if (bid >= nb) return;
    const int tid = threadIdx.x, p = tid % P, t = tid / P;
    const int row = bid / NLT, l = (bid % NLT) * P + p;
    if (l >= G::Mz) return;
    // AXIS 0: element j at (j * N + row) * Mi + l;  AXIS 1: element k at (row * N + k) * Mi + l
    constexpr long stride = AXIS == 0 ? (long)NX * G::Mi : (long)G::Mi;
    const long base = (AXIS == 0 ? (long)row * G::Mi : (long)row * NX * G::Mi) + l;
    c2<T> v[NI][E];
#pragma unroll
    for (int b = 0; b < NI; b++)
#pragma unroll
        for (int e = 0; e < E; e++) v[b][e] = in[b * G::nc + base + (t + e * TPP) * stride];
    // output b: input b mod nin, plus inputs b + nout, b + 2 nout, ... (every input is used)
#pragma unroll
    for (int b = 0; b < NO; b++)
#pragma unroll
        for (int e = 0; e < E; e++) {
            c2<T> s = v[b % NI][e];
#pragma unroll
            for (int i = b + NO; i < NI; i += NO) {
                s.x += v[i][e].x;
                s.y += v[i][e].y;
            }
            out[b * G::nc + base + (t + e * TPP) * stride] = s;
        }
}

// z rows: P rows per block, 64 lanes per row, NX/64 elements per lane of the NX-point extension
template <typename T, int P, int NI, int NO, int NX>
__global__ void __launch_bounds__(P * 64) k_zrows(const c2<T>* __restrict__ in, c2<T>* __restrict__ out, int nb) {
    using G = Geo<T, NX>;
    constexpr int EZ = NX / 64;
    const int bid = tile_of_block();
    if (bid >= nb) return;
    const int t = threadIdx.x % 64, p = threadIdx.x / 64;
    const long off = ((long)bid * P + p) * G::Mi;
    c2<T> acc[EZ];
#pragma unroll
    for (int e = 0; e < EZ; e++) acc[e] = c2<T>{0, 0};
#pragma unroll
    for (int b = 0; b < NI; b++)
#pragma unroll
        for (int e = 0; e < EZ; e++) {
            const int n = t + 64 * e;
            const int sl = n <= NX / 2 ? n : NX - n;
            if (sl <= G::lmax) {
                const c2<T> a = in[b * G::nc + off + sl];
                acc[e].x += a.x;
                acc[e].y += a.y;
            }
        }
#pragma unroll
    for (int b = 0; b < NO; b++)
#pragma unroll
        for (int e = 0; e < EZ; e++) {
            const int n = t + 64 * e;
            if (n < G::Mz) out[b * G::nc + off + n] = acc[e];
        }
}

template <typename T>
__global__ void __launch_bounds__(256) k_flat(const c2<T>* __restrict__ in, c2<T>* __restrict__ out, long n) {
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) out[i] = in[i];
}

struct Timer {
    hipEvent_t e0, e1;
    int reps;
    Timer(int r) : reps(r) {
        CK(hipEventCreate(&e0));
        CK(hipEventCreate(&e1));
    }
    template <typename F>
    float ms(F launch) {
        for (int i = 0; i < 3; i++) launch();
        CK(hipGetLastError());
        CK(hipDeviceSynchronize());
        CK(hipEventRecord(e0));
        for (int i = 0; i < reps; i++) launch();
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        CK(hipGetLastError());
        float t;
        CK(hipEventElapsedTime(&t, e0, e1));
        return t / reps;
    }
};

static void report(const char* pattern, const char* prec, int N, int nin, int nout, double bytes, float ms) {
    printf("{\"pattern\": \"%s\", \"precision\": \"%s\", \"N\": %d, \"read_fields\": %d, \"write_fields\": %d, "
           "\"bytes\": %.0f, \"ms\": %.5f, \"GBps\": %.1f}\n",
           pattern, prec, N, nin, nout, bytes, ms, bytes / (ms * 1e-3) / 1e9);
    fflush(stdout);
}

template <typename T, int NX>
static void run(const char* prec, int reps) {
    using G = Geo<T, NX>;
    const int N = NX;
    c2<T>*in, *out;
    CK(hipMalloc(&in, sizeof(c2<T>) * 9 * G::nc));
    CK(hipMalloc(&out, sizeof(c2<T>) * 6 * G::nc));
    CK(hipMemset(in, 0, sizeof(c2<T>) * 9 * G::nc));
    CK(hipMemset(out, 0, sizeof(c2<T>) * 6 * G::nc));
    Timer tm(reps);
    const double field = (double)N * N * G::Mz * sizeof(c2<T>);
    {
        const long n = 3L * N * N * G::Mz;
        const float ms = tm.ms([&] { hipLaunchKernelGGL(k_flat<T>, dim3(8192), dim3(256), 0, 0, in, out, n); });
        report("flat", prec, N, 3, 3, 6 * field, ms);
    }
    constexpr int P = 8;
    const int nlt = (G::Mz + P - 1) / P, nb = N * nlt, grid = (nb + 7) / 8 * 8;
#define PENCIL(NAME, AXIS, E, NI, NO)                                                                              \
    {                                                                                                              \
        const float ms = tm.ms([&] {                                                                               \
            hipLaunchKernelGGL((k_pencil<T, AXIS, P, E, NI, NO, NX>), dim3(grid), dim3(P * (NX / E)), 0, 0, in, out, nb); \
        });                                                                                                        \
        report(NAME, prec, N, NI, NO, (NI + NO) * field, ms);                                                      \
    }
    PENCIL("x_pencil", 0, 8, 3, 3)
    PENCIL("x_pencil", 0, 4, 3, 6)
    PENCIL("x_pencil", 0, 8, 6, 3)
    PENCIL("y_pencil", 1, 8, 3, 3)
    PENCIL("y_pencil", 1, 4, 3, 6)
    PENCIL("y_pencil", 1, 4, 6, 6)
#undef PENCIL
    {
        constexpr int PZ = 2;
        const int nbz = N * N / PZ, gz = (nbz + 7) / 8 * 8;
        const float ms = tm.ms([&] { hipLaunchKernelGGL((k_zrows<T, PZ, 9, 3, NX>), dim3(gz), dim3(PZ * 64), 0, 0, in, out, nbz); });
        report("z_rows_truncated", prec, N, 9, 3, (9.0 * (G::lmax + 1) / G::Mz + 3) * field, ms);
    }
    CK(hipFree(in));
    CK(hipFree(out));
}

int main(int argc, char** argv) {
    const int N = argc > 1 ? atoi(argv[1]) : 256;
    const char* prec = argc > 2 ? argv[2] : "fp64";
    const int reps = argc > 3 ? atoi(argv[3]) : 20;
    const bool f64 = strcmp(prec, "fp64") == 0;
    if ((!f64 && strcmp(prec, "fp32") != 0) || reps < 1) {
        fprintf(stderr, "usage: pattern_ceiling [128|256|512] [fp64|fp32] [reps]\n");
        return 2;
    }
    switch (N) {
        case 128: f64 ? run<double, 128>("fp64", reps) : run<float, 128>("fp32", reps); break;
        case 256: f64 ? run<double, 256>("fp64", reps) : run<float, 256>("fp32", reps); break;
        case 512: f64 ? run<double, 512>("fp64", reps) : run<float, 512>("fp32", reps); break;
        default: fprintf(stderr, "pattern_ceiling: N must be 128, 256 or 512\n"); return 2;
    }
    return 0;
}
