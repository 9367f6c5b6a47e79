// ns3d_api.hip -- context, launch plumbing and the extern "C" entry points of
// include/ns3d.h.  One translation unit; compiled for gfx950 only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <type_traits>
#include <vector>

#include "../../include/ns3d.h"
// tile shapes and plans chosen by ns3dperiodic_amd/autotune.py; the #ifndef
// defaults in the kernel headers apply to whatever it does not set
#if !defined(NS3D_NO_TUNED_HEADER) && __has_include("tuned.hpp")
#include "tuned.hpp"
#endif
#include "k_misc.hpp"
#include "k_krylov.hpp"
#include "k_diag.hpp"

using namespace ns3d;

namespace {

typedef cplx<double> zd;

template <class F>
int with_size(int n, F&& f) {
    switch (n) {
        case 8: return f(std::integral_constant<int, 8>{});
        case 16: return f(std::integral_constant<int, 16>{});
        case 32: return f(std::integral_constant<int, 32>{});
        case 48: return f(std::integral_constant<int, 48>{});
        case 64: return f(std::integral_constant<int, 64>{});
        case 80: return f(std::integral_constant<int, 80>{});
        case 96: return f(std::integral_constant<int, 96>{});
        case 128: return f(std::integral_constant<int, 128>{});
        case 160: return f(std::integral_constant<int, 160>{});
        case 192: return f(std::integral_constant<int, 192>{});
        case 256: return f(std::integral_constant<int, 256>{});
        case 320: return f(std::integral_constant<int, 320>{});
        case 384: return f(std::integral_constant<int, 384>{});
        case 512: return f(std::integral_constant<int, 512>{});
        default: return NS3D_ERR_SIZE;
    }
}

bool size_ok(int n) {
    return n == 8 || n == 16 || n == 32 || n == 48 || n == 64 || n == 80 || n == 96 || n == 128 || n == 160 ||
           n == 192 || n == 256 || n == 320 || n == 384 || n == 512;
}

// pencils per block in the x/y passes: rows of 128 B
#ifndef NS3D_PXY32
#define NS3D_PXY32 16
#endif
template <typename TB> constexpr int PXY0 = sizeof(TB) == 8 ? 8 : NS3D_PXY32;
// Blocks are capped at NS3D_XY_MAXT threads: 1024-thread blocks, which fill
// a CU's 16 waves with one block, lose to 512 (measured at 256^3 and 512^3).
#ifndef NS3D_XY_MAXT
#define NS3D_XY_MAXT 512
#endif
template <typename TB, int N, int E>
constexpr int PXY = PXY0<TB> * XYPlan<N, E>::TPP <= NS3D_XY_MAXT ? PXY0<TB>
                                                                 : (NS3D_XY_MAXT / XYPlan<N, E>::TPP > 0 ? NS3D_XY_MAXT / XYPlan<N, E>::TPP : 1);
// Pencils per block of each fused x/y pass, by precision.
#ifndef NS3D_PXINV32
#define NS3D_PXINV32 8
#endif
#ifndef NS3D_PYINV32
#define NS3D_PYINV32 8
#endif
#ifndef NS3D_PYFWD32
#define NS3D_PYFWD32 16
#endif
#ifndef NS3D_PXFWD32
#define NS3D_PXFWD32 8
#endif
#ifndef NS3D_PXINV64
#define NS3D_PXINV64 8
#endif
#ifndef NS3D_PYINV64
#define NS3D_PYINV64 8
#endif
template <typename TB, int N, int E, int P32, int P64>
constexpr int PXY_T = (sizeof(TB) == 8 ? P64 : P32) * XYPlan<N, E>::TPP <= NS3D_XY_MAXT
                          ? (sizeof(TB) == 8 ? P64 : P32)
                          : (NS3D_XY_MAXT / XYPlan<N, E>::TPP > 0 ? NS3D_XY_MAXT / XYPlan<N, E>::TPP : 1);
// Points per thread of each x/y pass, by precision (measured at 256^3 on MI355X):
// FP64 favours 4 points in the inverse passes, 8 in the forward ones; FP32 the mix below.
#ifndef NS3D_EXINV32
#define NS3D_EXINV32 8
#endif
#ifndef NS3D_EYINV32
#define NS3D_EYINV32 4
#endif
#ifndef NS3D_EXINV64
#define NS3D_EXINV64 4
#endif
#ifndef NS3D_EYINV64
#define NS3D_EYINV64 4
#endif
#ifndef NS3D_EXFWD64
#define NS3D_EXFWD64 8
#endif
template <typename TB> constexpr int E_XINV = sizeof(TB) == 8 ? NS3D_EXINV64 : NS3D_EXINV32;
template <typename TB> constexpr int E_YINV = sizeof(TB) == 8 ? NS3D_EYINV64 : NS3D_EYINV32;
template <typename TB> constexpr int E_C2C = 8;
#ifndef NS3D_EXFWD32
#define NS3D_EXFWD32 4
#endif
template <typename TB> constexpr int E_XFWD = sizeof(TB) == 8 ? NS3D_EXFWD64 : NS3D_EXFWD32;
// one output field per block in the inverse passes (input re-read by 2 or 3 blocks)
#ifndef NS3D_OS_XINV
#define NS3D_OS_XINV 0
#endif
#ifndef NS3D_OS_YINV
#define NS3D_OS_YINV 0
#endif
// pencils per block in the z pass: ~128 threads (FP32: NS3D_PZT32)
#ifndef NS3D_PZT32
#define NS3D_PZT32 128
#endif
// (measured: smaller blocks beat whole-wavefront ones, e.g. 96 vs 192 threads at 192^3)
#ifndef NS3D_PZT64
#define NS3D_PZT64 128
#endif
// Radix-3 lengths from 192 up have their own block shapes (NS3D_R3_*, and
// NS3D_R3_ZSWZ in the z pass), swept at 192^3 (FP64 366 -> 419, FP32 680 ->
// 700, FP32 state 725 -> 768 steps/s); at 48 and 96 these shapes were slower
// than the ones above (swept at 256^3), which those lengths keep.
#ifndef NS3D_R3_PXINV64
#define NS3D_R3_PXINV64 4
#endif
#ifndef NS3D_R3_PYINV64
#define NS3D_R3_PYINV64 4
#endif
#ifndef NS3D_R3_PZT64
#define NS3D_R3_PZT64 256
#endif
#ifndef NS3D_R3_PXFWD32
#define NS3D_R3_PXFWD32 16
#endif
#ifndef NS3D_R3_PYINV32
#define NS3D_R3_PYINV32 4
#endif
#define NS3D_BY_RADIX(N, A, R3) ((N) % 3 == 0 && (N) >= 192 ? (R3) : (A))
template <typename TB, int N>
constexpr int PZT = sizeof(TB) == 8 ? NS3D_BY_RADIX(N, NS3D_PZT64, NS3D_R3_PZT64) : NS3D_PZT32;
template <typename TB, int N>
constexpr int PZ = (PZT<TB, N> / ZPlanT<TB, N>::TPP) > 0 ? (PZT<TB, N> / ZPlanT<TB, N>::TPP) : 1;

#define HIPCHK(x)                                         \
    do {                                                  \
        hipError_t e_ = (x);                              \
        if (e_ != hipSuccess) return NS3D_ERR_HIP;        \
    } while (0)

}  // namespace

// Scratch of the allocation-free entry point ns3d_rk3_ssp_steps_noalloc: one
// copy per GPU, held until the process ends.  In host code `scratch` is only a
// symbol for hipGetSymbolAddress, never an address (see scratch_ptr).
constexpr size_t NS3D_NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[NS3D_NOALLOC_SCRATCH_BYTES];

struct ns3d_ctx {
    int Nx, Ny, Nz, Mz;
    long nc, nr, npen;
    long nci;  // modes per scratch field of the fused passes (rows padded from Mz to mi_of(c))
    int m8, m16;     // Mz rounded up to a multiple of NS3D_ROWPAD64 / NS3D_ROWPAD32 (FP64 / FP32 rows)
    long nc8, nc16;  // Nx*Ny*m8, Nx*Ny*m16
    double Lx, Ly, Lz, Re, dt;
    int precision;
    unsigned flags;
    hipStream_t stream;
    int device;
    int last_error;
    // device tables
    double* dtab = nullptr;
    SpecTables tab;
    void* tw64[3] = {nullptr, nullptr, nullptr};  // x, y, z twiddles (double)
    void* tw32[3] = {nullptr, nullptr, nullptr};  // (float)
    // bulk scratch (A|B regions) in bulk precision; capacity in fields
    void* scratch = nullptr;
    int scratch_fields = 0;
    // incremented whenever a buffer that a captured RK3 step addresses is reallocated
    long long generation = 0;
    // FP64 stage buffers and a Q^ buffer
    zd* stage = nullptr;  // 9 fields: u1[3], u2[3], q[3]
    zd* extra = nullptr;  // 6 fields: tangent map v1[3] + spare
    cplx<float>* s32 = nullptr;  // NS3D_FP32_STATE RK3 loop: state[3], u1[3], u2[3] in FP32
    double* red = nullptr;  // reduction partials (2*RED_BLOCKS + 2)
    double* red_host = nullptr;  // pinned
    void* real_scratch = nullptr;  // physical-space scratch (FP64 or FP32), 4 fields
    // registered base flow for the Krylov callbacks
    zd* base[3] = {nullptr, nullptr, nullptr};
    const zd* bforce[3] = {nullptr, nullptr, nullptr};
    int timesteps_period = 1;
    double tau = 1.0;
    zd* kry = nullptr;  // 3 fields for J.v unpacking
    // Krylov vector kernels: reduction partials, results, pinned host copies, small dense matrix
    double* kv_partial = nullptr;
    double* kv_out = nullptr;
    double* kv_host = nullptr;
    double* kq = nullptr;
    double* kq_host = nullptr;
    // a captured graph of nsteps RK3 steps for fixed state/forcing pointers, dt and Re
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    const void* graph_key[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int graph_steps = -1;
    double graph_dt = 0.0, graph_Re = 0.0;
    // passive scalar: physical velocity written by the z pass (bulk precision, 3 real fields)
    void* uphys = nullptr;
    void* uphys_out = nullptr;  // non-null only while a scalar step wants it
    double* theta_stage = nullptr;  // 2 real FP64 fields
    // batches of same-size fields advanced by one set of launches (grid.z = member):
    // members in flight, scratch stride per member (elements), stage buffers per member
    int batch = 1;
    long bsc = 0;
    zd* bstage = nullptr;  // batch members' u1[3], u2[3]: 6 fields of nc8 each per member
    int bstage_members = 0;
    // members sitting out steps, and per-member dt / Re (ns3d_rk3_ssp_batch_members)
    MemberMask mm = {nullptr, 0};
    const double* mdt = nullptr;
    const double* mre = nullptr;
    // graph of ns3d_rk3_ssp_batch_members for fixed pointers, nbatch and nsteps
    hipGraph_t mgraph = nullptr;
    hipGraphExec_t mgraph_exec = nullptr;
    const void* mgraph_key[8] = {};
    int mgraph_nb = -1, mgraph_steps = -1;
    // optional per-pass timing with HIP events on the context stream
    bool prof_on = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, int>> ev_used;  // (pass id, index of start event; stop = +1)
    hipEvent_t ev_open = nullptr;
    int ev_open_pass = -1;
    // sticky non-finite flag written by the fused epilogues and the scalar stage
    int* nf = nullptr;
    int* nf_host = nullptr;  // pinned
    // batched probe: partial sums and results (device), results (pinned host)
    double* pb = nullptr;
    double* pb_host = nullptr;
    size_t pb_cap = 0;
};

namespace {

int set_err(ns3d_ctx* c, int s) {
    if (c) c->last_error = s;
    return s;
}

// ---------------------------------------------------------------- pass timing
// Pass ids: 0 x-inverse, 1 y-inverse, 2 z product, 3 y-forward, 4 x-forward+epilogue.
constexpr int NPASS = 5;
hipEvent_t prof_event(ns3d_ctx* c, size_t i) {
    while (c->ev_pool.size() <= i) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[i];
}
void prof_begin(ns3d_ctx* c, int pass) {
    if (!c->prof_on) return;
    const size_t i = 2 * c->ev_used.size();
    hipEvent_t a = prof_event(c, i), b = prof_event(c, i + 1);
    if (!a || !b) return;
    hipEventRecord(a, c->stream);
    c->ev_used.push_back({pass, (int)i});
}
void prof_end(ns3d_ctx* c) {
    if (!c->prof_on || c->ev_used.empty()) return;
    hipEventRecord(c->ev_pool[c->ev_used.back().second + 1], c->stream);
}

// ---------------------------------------------------------------- host tables
int build_tables(ns3d_ctx* c) {
    const int Nx = c->Nx, Ny = c->Ny, Mz = c->Mz;
    const double Lx = c->Lx, Ly = c->Ly, Lz = c->Lz;
    const double PI = 3.141592653589793238462643383279502884;  // reference PI literal (Source/Macros.h:75)
    const int tot = 5 * (Nx + Ny + Mz);
    std::vector<double> h(tot);
    double* kx = h.data();
    double* ky = kx + Nx;
    double* kz = ky + Ny;
    double* sx = kz + Mz;
    double* sy = sx + Nx;
    double* sz = sy + Ny;
    double* px = sz + Mz;
    double* py = px + Nx;
    double* pz = py + Ny;
    double* fm = pz + Mz;
    double* fn = fm + Nx;
    double* fq = fn + Ny;
    double* mx = fq + Mz;
    double* my = mx + Nx;
    double* mz = my + Ny;
    (void)mz;
    std::vector<double> mxv(Nx), myv(Ny), mzv(Mz);
    const double k1 = 2.0 * M_PI / Lx, k2 = 2.0 * M_PI / Ly, k3 = 2.0 * M_PI / Lz;
    // build_mask_matrix is called with (Nx, Ny, Mz); NS3D_FLAG_MASK_NYQUIST uses the Nyquist wavenumbers
    const bool nyq = (c->flags & NS3D_FLAG_MASK_NYQUIST) != 0;
    const double kxMax = nyq ? Nx * PI / Lx : 2 * Nx * PI / Lx, kyMax = nyq ? Ny * PI / Ly : 2 * Ny * PI / Ly,
                 kzMax = nyq ? c->Nz * PI / Lz : 2 * Mz * PI / Lz;
    for (int j = 0; j < Nx; j++) {
        const int m = j >= Nx / 2 ? j - Nx : j;
        kx[j] = 2.0 * M_PI / Lx * m;            // build_Nabla_Wavenumbers (math_support.cu:393-414)
        sx[j] = kx[j] * kx[j];                  // build_Laplace_and_Diffusion (:374-390)
        px[j] = k1 * k1 * m * m;                // build_projection_matrix_elements (:418-466)
        fm[j] = (double)m;
        const double km = m * 2 * PI / Lx;      // build_mask_matrix (:471-519)
        mxv[j] = km * km / (kxMax * kxMax);
    }
    for (int k = 0; k < Ny; k++) {
        const int n = k >= Ny / 2 ? k - Ny : k;
        ky[k] = 2.0 * M_PI / Ly * n;
        sy[k] = ky[k] * ky[k];
        py[k] = k2 * k2 * n * n;
        fn[k] = (double)n;
        const double kn = n * 2 * PI / Ly;
        myv[k] = kn * kn / (kyMax * kyMax);
    }
    for (int l = 0; l < Mz; l++) {
        const int q = l;
        kz[l] = 2.0 * M_PI / Lz * q;
        sz[l] = kz[l] * kz[l];
        pz[l] = k3 * k3 * q * q;
        fq[l] = (double)q;
        const double kq = q * 2 * PI / Lz;
        mzv[l] = kq * kq / (kzMax * kzMax);
    }
    std::vector<double> all(h.begin(), h.begin() + 4 * (Nx + Ny + Mz));
    all.insert(all.end(), mxv.begin(), mxv.end());
    all.insert(all.end(), myv.begin(), myv.end());
    all.insert(all.end(), mzv.begin(), mzv.end());
    HIPCHK(hipMalloc(&c->dtab, sizeof(double) * all.size()));
    HIPCHK(hipMemcpy(c->dtab, all.data(), sizeof(double) * all.size(), hipMemcpyHostToDevice));
    const double* d = c->dtab;
    SpecTables& t = c->tab;
    t.kx = d; d += Nx; t.ky = d; d += Ny; t.kz = d; d += Mz;
    t.sx = d; d += Nx; t.sy = d; d += Ny; t.sz = d; d += Mz;
    t.px = d; d += Nx; t.py = d; d += Ny; t.pz = d; d += Mz;
    t.fm = d; d += Nx; t.fn = d; d += Ny; t.fq = d; d += Mz;
    t.mx = d; d += Nx; t.my = d; d += Ny; t.mz = d;
    t.k1k2 = k1 * k2;
    t.k1k3 = k1 * k3;
    t.k2k3 = k2 * k3;
    // largest s with sqrt(s) < 2/3 (the reference compares sqrt(sum) < 2.0/3.0)
    const double c23 = 2.0 / 3.0;
    double s = c23 * c23;
    while (std::sqrt(s) >= c23) s = std::nextafter(s, 0.0);
    while (std::sqrt(std::nextafter(s, 1.0)) < c23) s = std::nextafter(s, 1.0);
    t.mask_smax = s;
    t.lmax = -1;
    for (int l = 0; l < Mz; l++)
        if (mzv[l] <= s) t.lmax = l;
    return NS3D_OK;
}

int build_twiddles(ns3d_ctx* c) {
    const int ns[3] = {c->Nx, c->Ny, c->Nz};
    for (int a = 0; a < 3; a++) {
        const int n = ns[a];
        std::vector<zd> h64(n);
        std::vector<cplx<float>> h32(n);
        for (int m = 0; m < n; m++) {
            const long double ang = 2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)n;
            const long double cr = cosl(ang), si = -sinl(ang);
            h64[m].x = (double)cr;
            h64[m].y = (double)si;
            h32[m].x = (float)cr;
            h32[m].y = (float)si;
        }
        HIPCHK(hipMalloc(&c->tw64[a], sizeof(zd) * n));
        HIPCHK(hipMalloc(&c->tw32[a], sizeof(cplx<float>) * n));
        HIPCHK(hipMemcpy(c->tw64[a], h64.data(), sizeof(zd) * n, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->tw32[a], h32.data(), sizeof(cplx<float>) * n, hipMemcpyHostToDevice));
    }
    return NS3D_OK;
}

template <typename TB> const cplx<TB>* twp(const ns3d_ctx* c, int axis) {
    return (const cplx<TB>*)(sizeof(TB) == 8 ? c->tw64[axis] : c->tw32[axis]);
}

size_t bulk_bytes(const ns3d_ctx* c) { return c->precision == NS3D_FP64 ? sizeof(zd) : sizeof(cplx<float>); }

// Row stride of the fused passes' scratch fields: Mz rounded up so that every
// P-mode row tile of the x/y passes (128 B) starts on a 128-B line.
#ifndef NS3D_PAD_ROWS
#define NS3D_PAD_ROWS 1
#endif
// Row padding (modes) of the library's own FP64 and FP32 fields: 8 modes, so
// FP64 rows start on 128-B and FP32 rows on 64-B boundaries, which the x and y
// tiles' 8-mode segments need.  256^3, rows of 136 against 144: FP32 steps
// -1.0% (bulk), -1.8% (state), -2.1% (increments), chiefly the x-forward pass
// (-4%: x-planes 136*256 modes apart map onto HBM better than 144*256); FP64
// rows of 144 gained only 0.4% for 6% more memory.
#ifndef NS3D_ROWPAD64
#define NS3D_ROWPAD64 8
#endif
#ifndef NS3D_ROWPAD32
#define NS3D_ROWPAD32 8
#endif
int mi_of(const ns3d_ctx* c) {
    const int a = NS3D_PAD_ROWS ? (c->precision == NS3D_FP64 ? NS3D_ROWPAD64 : NS3D_ROWPAD32) : 1;
    return (c->Mz + a - 1) / a * a;
}

// Forget a captured RK3 graph (its nodes hold scratch pointers, table values and the stream).
void drop_graph(ns3d_ctx* c) {
    if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
    if (c->graph) hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
    c->graph_steps = -1;
    if (c->mgraph_exec) hipGraphExecDestroy(c->mgraph_exec);
    if (c->mgraph) hipGraphDestroy(c->mgraph);
    c->mgraph_exec = nullptr;
    c->mgraph = nullptr;
    c->mgraph_nb = c->mgraph_steps = -1;
}

int ensure_scratch(ns3d_ctx* c, int fields) {
    if (c->scratch_fields >= fields) return NS3D_OK;
    drop_graph(c);
    if (c->scratch) c->generation++;
    if (c->scratch) {
        hipStreamSynchronize(c->stream);
        hipFree(c->scratch);
        c->scratch = nullptr;
        c->scratch_fields = 0;
    }
    if (hipMalloc(&c->scratch, bulk_bytes(c) * (size_t)c->nci * fields) != hipSuccess) return NS3D_ERR_ALLOC;
    c->scratch_fields = fields;
    return NS3D_OK;
}

int ensure_real_scratch(ns3d_ctx* c) {
    if (c->real_scratch) return NS3D_OK;
    if (hipMalloc(&c->real_scratch, sizeof(double) * (size_t)c->nr * 4) != hipSuccess) return NS3D_ERR_ALLOC;
    return NS3D_OK;
}

int ensure_fp64(ns3d_ctx* c, zd** p, int fields) {
    if (*p) return NS3D_OK;
    // nc8 >= nc: room for fields in either the reference or the padded layout
    if (hipMalloc((void**)p, sizeof(zd) * (size_t)c->nc8 * fields) != hipSuccess) return NS3D_ERR_ALLOC;
    return NS3D_OK;
}

inline unsigned nblk(long n, int t) { return (unsigned)((n + t - 1) / t); }
// grid for the XCD-remapped pencil kernels: block count padded to a multiple of 8
inline dim3 xgrid(long nb, int nfy = 1, int nz = 1) {
    return dim3((unsigned)(((nb + 7) / 8) * 8), (unsigned)nfy, (unsigned)nz);
}

// ---------------------------------------------------------------- fused pipeline
// Runs P1..P4 for NF vector fields and leaves D (3 fields) at scratch + 3*nci.
// mw: row stride of the input fields (0: Mz, the reference layout)
// bin: batch member stride of the inputs (c->batch members, scratch c->bsc apart).
// in2 (nf = 1 only, may be null): a second field of in's type and layout added
// to the input before the transforms (the FP32-increment loop's base + increment).
template <typename TB, typename TS = double>
int run_nonlinear_passes(ns3d_ctx* c, const cplx<TS>* const* in, int nf, int mw = 0, long bin = 0,
                         const cplx<TS>* const* in2 = nullptr) {
    if (mw == 0) mw = c->Mz;
    int st = ensure_scratch(c, 15 * nf * c->batch);
    if (st) return st;
    const int nz = c->batch;
    const long bsc = c->bsc;
    // scratch fields: rows padded to Mi (128-B aligned pencil tiles), nci modes each
    const long nc = c->nci;
    const int Mi = mi_of(c);
    cplx<TB>* A = (cplx<TB>*)c->scratch;
    cplx<TB>* B = A + (long)6 * nf * nc;
    hipStream_t s = c->stream;
    // P1 x-inverse
    prof_begin(c, 0);
    st = with_size(c->Nx, [&](auto NXc) -> int {
        constexpr int NX = decltype(NXc)::value;
        constexpr int E = E_XINV<TB>;
        constexpr int P = PXY_T<TB, NX, E, NS3D_PXINV32, NS3D_BY_RADIX(NX, NS3D_PXINV64, NS3D_R3_PXINV64)>;
        const int nlt = (c->Mz + P - 1) / P;
        const int nb = c->Ny * nlt;
        constexpr int OS = NS3D_OS_XINV;
        dim3 grid = xgrid(nb, 3 * nf * (OS ? 2 : 1), nz), blk(P * XYPlan<NX, E>::TPP);
        if (nf == 1 && in2) {
            PtrArr<const cplx<TS>, 3> w{{in[0], in[1], in[2]}};
            PtrArr<const cplx<TS>, 3> w2{{in2[0], in2[1], in2[2]}};
            hipLaunchKernelGGL((k_x_inv_grad<TB, NX, P, 1, E, OS, TS, 1>), grid, blk, 0, s, w, A, nc, c->Ny, c->Mz, mw, Mi, nlt, nb,
                               c->tab, twp<TB>(c, 0), bin, bsc, w2, c->mm);
        } else if (nf == 1) {
            PtrArr<const cplx<TS>, 3> w{{in[0], in[1], in[2]}};
            hipLaunchKernelGGL((k_x_inv_grad<TB, NX, P, 1, E, OS, TS>), grid, blk, 0, s, w, A, nc, c->Ny, c->Mz, mw, Mi, nlt, nb, c->tab,
                               twp<TB>(c, 0), bin, bsc, PtrArr<const cplx<TS>, 3>{}, c->mm);
        } else if constexpr (std::is_same<TS, double>::value) {
            PtrArr<const zd, 6> w{{in[0], in[1], in[2], in[3], in[4], in[5]}};
            hipLaunchKernelGGL((k_x_inv_grad<TB, NX, P, 2, E, OS>), grid, blk, 0, s, w, A, nc, c->Ny, c->Mz, mw, Mi, nlt, nb, c->tab,
                               twp<TB>(c, 0), bin, bsc, PtrArr<const zd, 3>{}, c->mm);
        } else {
            return NS3D_ERR_ARG;
        }
        return NS3D_OK;
    });
    prof_end(c);
    if (st) return st;
    // P2 y-inverse
    prof_begin(c, 1);
    st = with_size(c->Ny, [&](auto NYc) -> int {
        constexpr int NY = decltype(NYc)::value;
        constexpr int E = E_YINV<TB>;
        constexpr int P = PXY_T<TB, NY, E, NS3D_BY_RADIX(NY, NS3D_PYINV32, NS3D_R3_PYINV32),
                                NS3D_BY_RADIX(NY, NS3D_PYINV64, NS3D_R3_PYINV64)>;
        const int nlt = (c->Mz + P - 1) / P;
        const int nb = c->Nx * nlt;
        constexpr int OS = NS3D_OS_YINV;
        dim3 grid = xgrid(nb, 3 * nf * (OS ? 3 : 1), nz), blk(P * XYPlan<NY, E>::TPP);
        if (nf == 1)
            hipLaunchKernelGGL((k_y_inv_grad<TB, NY, P, 1, E, OS>), grid, blk, 0, s, A, B, nc, c->Ny, c->Mz, Mi, nlt, nb, c->tab,
                               twp<TB>(c, 1), bsc, c->mm);
        else
            hipLaunchKernelGGL((k_y_inv_grad<TB, NY, P, 2, E, OS>), grid, blk, 0, s, A, B, nc, c->Ny, c->Mz, Mi, nlt, nb, c->tab,
                               twp<TB>(c, 1), bsc, c->mm);
        return NS3D_OK;
    });
    prof_end(c);
    if (st) return st;
    // P3 z: C2R, product, R2C -> C at A
    prof_begin(c, 2);
    st = with_size(c->Nz, [&](auto NZc) -> int {
        constexpr int NZ = decltype(NZc)::value;
        constexpr int P = PZ<TB, NZ>;
        const int nb = (int)nblk(c->npen, P);
        dim3 grid = xgrid(nb, 1, nz), blk(P * ZPlanT<TB, NZ>::TPP);
        const TB scale = (TB)(1.0 / (1.0 * c->Nx * c->Ny * c->Nz));
        TB* up = nf == 1 ? (TB*)c->uphys_out : nullptr;
        if (nf == 1)
            hipLaunchKernelGGL((k_z_product<TB, NZ, P, 1>), grid, blk, 0, s, B, A, nc, c->npen, Mi, nb, c->tab, twp<TB>(c, 2), scale,
                               up, c->nr, bsc, c->mm);
        else
            hipLaunchKernelGGL((k_z_product<TB, NZ, P, 2>), grid, blk, 0, s, B, A, nc, c->npen, Mi, nb, c->tab, twp<TB>(c, 2), scale,
                               (TB*)nullptr, c->nr, bsc, c->mm);
        return NS3D_OK;
    });
    prof_end(c);
    if (st) return st;
    // P4 y-forward C -> D
    prof_begin(c, 3);
    st = with_size(c->Ny, [&](auto NYc) -> int {
        constexpr int NY = decltype(NYc)::value;
        constexpr int E = E_C2C<TB>;
        constexpr int P = PXY_T<TB, NY, E, NS3D_PYFWD32, 8>;
        const int nlt = (c->Mz + P - 1) / P;
        const int nb = c->Nx * nlt;
        dim3 grid = xgrid(nb, 3, nz), blk(P * XYPlan<NY, E>::TPP);
        hipLaunchKernelGGL((k_axis_c2c<TB, NY, P, -1, 1, E>), grid, blk, 0, s, A, A + 3 * nc, 3, nc, c->Ny, c->Mz, Mi, nlt, nb,
                           twp<TB>(c, 1), bsc, c->mm);
        return NS3D_OK;
    });
    prof_end(c);
    return st;
}

template <typename TB, int MODE>
int run_x_forward_epilogue(ns3d_ctx* c, const EpiArgs& ea) {
    const cplx<TB>* D = (const cplx<TB>*)c->scratch + 3 * c->nci;
    prof_begin(c, 4);
    int st = with_size(c->Nx, [&](auto NXc) -> int {
        constexpr int NX = decltype(NXc)::value;
        constexpr int E = E_XFWD<TB>;
        constexpr int P = PXY_T<TB, NX, E, NS3D_BY_RADIX(NX, NS3D_PXFWD32, NS3D_R3_PXFWD32), 8>;
        const int nlt = (c->Mz + P - 1) / P;
        const int nb = c->Ny * nlt;
        dim3 grid = xgrid(nb, 1, c->batch), blk(P * XYPlan<NX, E>::TPP);
        hipLaunchKernelGGL((k_x_fwd_epi<TB, NX, P, MODE, E>), grid, blk, 0, c->stream, D, c->nci, c->Ny, c->Mz, mi_of(c), nlt, nb,
                           c->tab, twp<TB>(c, 0), ea, c->bsc, c->mm);
        return NS3D_OK;
    });
    prof_end(c);
    return st;
}

EpiArgs make_epi(const ns3d_ctx* c) {
    EpiArgs ea;
    memset(&ea, 0, sizeof(ea));
    ea.dt = c->dt;
    ea.Re = c->Re;
    ea.nu = 1.0 / c->Re;
    ea.dtnu = c->dt / c->Re;
    ea.a0 = 0.0;
    ea.a1 = 1.0;
    ea.ipq = (c->flags & NS3D_FLAG_NO_IPQ) ? 0 : 1;
    ea.ny = c->Ny;
    ea.mw = ea.mu0 = ea.mo = ea.mf = c->Mz;
    ea.nf = c->nf;
    ea.dtv = c->mdt;
    ea.rev = c->mre;
    return ea;
}

// nonlinear term of nf fields followed by the MODE epilogue
template <int MODE>
int fused(ns3d_ctx* c, const zd* const* in, int nf, const EpiArgs& ea) {
    int st;
    if (c->precision == NS3D_FP64) {
        st = run_nonlinear_passes<double>(c, in, nf, ea.mw, ea.bw);
        if (st) return st;
        st = run_x_forward_epilogue<double, MODE>(c, ea);
    } else {
        st = run_nonlinear_passes<float>(c, in, nf, ea.mw, ea.bw);
        if (st) return st;
        st = run_x_forward_epilogue<float, MODE>(c, ea);
    }
    if (st) return st;
    return hipGetLastError() == hipSuccess ? NS3D_OK : NS3D_ERR_HIP;
}

// E(w) stage: out = a0*u0 + a1*(w + dt*L(w)); mw/mu0/mo: row strides (0: Mz)
// bw, bu0, bo: batch member strides of w, u0 and out (when c->batch > 1)
int stage_into(ns3d_ctx* c, const zd* const w[3], const zd* const u0[3], double a0, double a1, zd* const out[3],
               const zd* const f[3], int mw = 0, int mu0 = 0, int mo = 0, long bw = 0, long bu0 = 0, long bo = 0) {
    EpiArgs ea = make_epi(c);
    if (mw) ea.mw = mw;
    if (mu0) ea.mu0 = mu0;
    if (mo) ea.mo = mo;
    ea.bw = bw;
    ea.bu0 = bu0;
    ea.bo = bo;
    for (int b = 0; b < 3; b++) {
        ea.w[b] = (zd*)w[b];
        ea.u0[b] = u0 ? u0[b] : nullptr;
        ea.f[b] = f ? f[b] : nullptr;
        ea.out[b] = out[b];
    }
    ea.a0 = a0;
    ea.a1 = a1;
    return fused<EPI_STAGE>(c, w, 1, ea);
}

int rk3_one(ns3d_ctx* c, zd* const u[3], const zd* const f[3]) {
    int st = ensure_fp64(c, &c->stage, 9);
    if (st) return st;
    // u1, u2 are the library's own: rows padded to M8 (128-B aligned), fields nc8 apart
    const long n8 = c->nc8;
    const int m8 = c->m8;
    zd* u1[3] = {c->stage, c->stage + n8, c->stage + 2 * n8};
    zd* u2[3] = {c->stage + 3 * n8, c->stage + 4 * n8, c->stage + 5 * n8};
    const zd* uc[3] = {u[0], u[1], u[2]};
    const zd* u1c[3] = {u1[0], u1[1], u1[2]};
    const zd* u2c[3] = {u2[0], u2[1], u2[2]};
    const int mz = c->Mz;
    // RK3_SSP (Source/RK_time_step.cu:310-327):
    //   u1 = E(u); u1 = E(u1); u2 = 3/4 u + 1/4 u1; u2 = E(u2); u = 1/3 u + 2/3 u2
    if ((st = stage_into(c, uc, nullptr, 0.0, 1.0, u1, f, mz, mz, m8))) return st;
    if ((st = stage_into(c, u1c, uc, 3.0 / 4.0, 1.0 / 4.0, u2, f, m8, mz, m8))) return st;
    if ((st = stage_into(c, u2c, uc, 1.0 / 3.0, 2.0 / 3.0, u, f, m8, mz, mz))) return st;
    return NS3D_OK;
}

// ---- NS3D_FP32_STATE: the RK3 loop keeps state and stages in FP32 (bulk
// transforms FP32, epilogue arithmetic and forcing FP64); the caller's FP64
// state is converted once on entry and once on exit of each call.
using cf = cplx<float>;

int stage_into32(ns3d_ctx* c, const cf* const w[3], const cf* const u0[3], double a0, double a1, cf* const out[3],
                 const zd* const f[3]) {
    EpiArgs ea = make_epi(c);
    for (int b = 0; b < 3; b++) {
        ea.w[b] = (zd*)w[b];
        ea.u0[b] = u0 ? (const zd*)u0[b] : nullptr;
        ea.f[b] = f ? f[b] : nullptr;
        ea.out[b] = (zd*)out[b];
    }
    ea.a0 = a0;
    ea.a1 = a1;
    ea.mw = ea.mu0 = ea.mo = c->m16;  // the FP32 state and stages use rows padded to m16
    int st = run_nonlinear_passes<float, float>(c, w, 1, c->m16);
    if (st) return st;
    st = run_x_forward_epilogue<float, EPI_STAGE32>(c, ea);
    if (st) return st;
    return hipGetLastError() == hipSuccess ? NS3D_OK : NS3D_ERR_HIP;
}

int rk3_one32(ns3d_ctx* c, const zd* const f[3]) {
    const long nc = c->nc16;
    cf* u[3] = {c->s32, c->s32 + nc, c->s32 + 2 * nc};
    cf* u1[3] = {c->s32 + 3 * nc, c->s32 + 4 * nc, c->s32 + 5 * nc};
    cf* u2[3] = {c->s32 + 6 * nc, c->s32 + 7 * nc, c->s32 + 8 * nc};
    const cf* uc[3] = {u[0], u[1], u[2]};
    const cf* u1c[3] = {u1[0], u1[1], u1[2]};
    const cf* u2c[3] = {u2[0], u2[1], u2[2]};
    int st;
    if ((st = stage_into32(c, uc, nullptr, 0.0, 1.0, u1, f))) return st;
    if ((st = stage_into32(c, u1c, uc, 3.0 / 4.0, 1.0 / 4.0, u2, f))) return st;
    if ((st = stage_into32(c, u2c, uc, 1.0 / 3.0, 2.0 / 3.0, u, f))) return st;
    return NS3D_OK;
}

int ensure_state32(ns3d_ctx* c) {
    if (c->s32) return NS3D_OK;
    if (hipMalloc((void**)&c->s32, sizeof(cf) * (size_t)c->nc16 * 9) != hipSuccess) return NS3D_ERR_ALLOC;
    return NS3D_OK;
}

// ---- NS3D_FP32_INCR: FP32 transforms and products over an FP64-accurate
// state.  Inside the RK3 loop the state u^n is the pair hi + lo of FP32
// fields (converted from and to the caller's FP64 field once per call) and
// the SSP stages are FP32 increments d over it (u1 = u^n + d1, u2 = u^n + d2;
// EPI_INC / EPI_INC_FINAL), so a stage value is rounded to eps32*|d|, not
// eps32*|u|.  The transforms and the stages' diffusion read hi alone; the
// last stage reads hi + lo and writes the new pair.  d2 overwrites d1 in place.
// s32 (rows m16): hi [3], lo [3], d [3].
int stage_incr(ns3d_ctx* c, bool first, double a1, bool last, const zd* const f[3]) {
    const long nc = c->nc16;
    cf* hi[3] = {c->s32, c->s32 + nc, c->s32 + 2 * nc};
    cf* lo[3] = {c->s32 + 3 * nc, c->s32 + 4 * nc, c->s32 + 5 * nc};
    cf* d[3] = {c->s32 + 6 * nc, c->s32 + 7 * nc, c->s32 + 8 * nc};
    const cf* hic[3] = {hi[0], hi[1], hi[2]};
    const cf* dc[3] = {d[0], d[1], d[2]};
    EpiArgs ea = make_epi(c);
    for (int b = 0; b < 3; b++) {
        ea.w[b] = (zd*)hi[b];
        ea.u0[b] = first ? nullptr : (const zd*)d[b];
        ea.f[b] = f ? f[b] : nullptr;
        ea.out[b] = (zd*)d[b];
        ea.out32[b] = lo[b];
    }
    ea.a1 = a1;
    ea.mw = ea.mo = ea.mu0 = c->m16;
    int st = run_nonlinear_passes<float, float>(c, hic, 1, c->m16, 0, first ? nullptr : dc);
    if (st) return st;
    st = last ? run_x_forward_epilogue<float, EPI_INC_FINAL>(c, ea) : run_x_forward_epilogue<float, EPI_INC>(c, ea);
    if (st) return st;
    return hipGetLastError() == hipSuccess ? NS3D_OK : NS3D_ERR_HIP;
}

int rk3_one_incr(ns3d_ctx* c, const zd* const f[3]) {
    int st;
    if ((st = stage_incr(c, true, 1.0, false, f))) return st;
    if ((st = stage_incr(c, false, 1.0 / 4.0, false, f))) return st;
    return stage_incr(c, false, 2.0 / 3.0, true, f);
}

// the RK3 loops of NS3D_FP32_STATE and NS3D_FP32_INCR keep 9 FP32 fields (s32)
bool uses_s32(const ns3d_ctx* c) { return c->precision == NS3D_FP32_STATE || c->precision == NS3D_FP32_INCR; }

// nsteps RK3 steps in the context's precision
// One RK3 step of nb same-size fields with one set of launches: member b's
// component k at u + (3b + k)*nc (reference layout); forcing shared.  The
// kernels and their order are those of rk3_one, so each member's result is
// bitwise the one rk3_one gives it alone.
int ensure_batch(ns3d_ctx* c, int nb) {
    if (c->bstage_members < nb) {
        drop_graph(c);
        if (c->bstage) {
            c->generation++;
            hipStreamSynchronize(c->stream);
            hipFree(c->bstage);
            c->bstage = nullptr;
            c->bstage_members = 0;
        }
        if (hipMalloc((void**)&c->bstage, sizeof(zd) * (size_t)c->nc8 * 6 * nb) != hipSuccess) return NS3D_ERR_ALLOC;
        c->bstage_members = nb;
    }
    return ensure_scratch(c, 15 * nb);
}

int rk3_one_batch(ns3d_ctx* c, int nb, zd* u, const zd* const f[3]) {
    const long n8 = c->nc8, nc = c->nc;
    int st = ensure_batch(c, nb);
    if (st) return st;
    zd* u1[3] = {c->bstage, c->bstage + n8, c->bstage + 2 * n8};
    zd* u2[3] = {c->bstage + 3 * n8, c->bstage + 4 * n8, c->bstage + 5 * n8};
    zd* uu[3] = {u, u + nc, u + 2 * nc};
    const zd* uc[3] = {uu[0], uu[1], uu[2]};
    const zd* u1c[3] = {u1[0], u1[1], u1[2]};
    const zd* u2c[3] = {u2[0], u2[1], u2[2]};
    const int mz = c->Mz, m8 = c->m8;
    const long bs = 3 * nc, bst = 6 * n8;
    c->batch = nb;
    c->bsc = 15 * c->nci;
    if (!(st = stage_into(c, uc, nullptr, 0.0, 1.0, u1, f, mz, mz, m8, bs, 0, bst)))
        if (!(st = stage_into(c, u1c, uc, 3.0 / 4.0, 1.0 / 4.0, u2, f, m8, mz, m8, bst, bs, bst)))
            st = stage_into(c, u2c, uc, 1.0 / 3.0, 2.0 / 3.0, uu, f, m8, mz, mz, bst, bs, bs);
    c->batch = 1;
    c->bsc = 0;
    return st;
}

// nsteps RK3 steps of nb batch members with per-member dt, Re (device arrays)
// and step counts: member b takes the first left[b] steps and sits out the rest.
int rk3_batch_members(ns3d_ctx* c, int nb, int nsteps, zd* u, const zd* const f[3], const double* dt, const double* re,
                      const int* left) {
    c->mm.left = left;
    c->mdt = dt;
    c->mre = re;
    int st = NS3D_OK;
    for (int i = 0; i < nsteps && !st; i++) {
        c->mm.step = i;
        st = rk3_one_batch(c, nb, u, f);
    }
    c->mm.left = nullptr;
    c->mm.step = 0;
    c->mdt = c->mre = nullptr;
    return st;
}

int rk3_steps(ns3d_ctx* c, int nsteps, zd* const u[3], const zd* const f[3]) {
    if (c->precision == NS3D_FP32_INCR) {
        if (nsteps == 0) return NS3D_OK;
        int st = ensure_state32(c);
        if (!st) st = ensure_scratch(c, 15);
        if (st) return st;
        const long n16 = c->nc16;
        for (int b = 0; b < 3; b++)
            hipLaunchKernelGGL(k_split_rows, dim3(nblk(c->nc, 256)), dim3(256), 0, c->stream, u[b], c->s32 + b * n16,
                               c->s32 + (3 + b) * n16, c->nc, c->Mz, c->m16);
        HIPCHK(hipGetLastError());
        for (int i = 0; i < nsteps; i++)
            if ((st = rk3_one_incr(c, f))) return st;
        for (int b = 0; b < 3; b++)
            hipLaunchKernelGGL(k_join_rows, dim3(nblk(c->nc, 256)), dim3(256), 0, c->stream, c->s32 + b * n16,
                               c->s32 + (3 + b) * n16, u[b], c->nc, c->Mz, c->m16);
        HIPCHK(hipGetLastError());
        return NS3D_OK;
    }
    if (c->precision != NS3D_FP32_STATE) {
        for (int i = 0; i < nsteps; i++) {
            int st = rk3_one(c, u, f);
            if (st) return st;
        }
        return NS3D_OK;
    }
    if (nsteps == 0) return NS3D_OK;
    int st = ensure_state32(c);
    if (!st) st = ensure_scratch(c, 15);
    if (st) return st;
    const long nc = c->nc, n16 = c->nc16;
    for (int b = 0; b < 3; b++)
        hipLaunchKernelGGL(k_d2f_rows, dim3(nblk(nc, 256)), dim3(256), 0, c->stream, u[b], c->s32 + b * n16, nc, c->Mz,
                           c->m16);
    for (int i = 0; i < nsteps; i++) {
        st = rk3_one32(c, f);
        if (st) return st;
    }
    for (int b = 0; b < 3; b++)
        hipLaunchKernelGGL(k_f2d_rows, dim3(nblk(nc, 256)), dim3(256), 0, c->stream, c->s32 + b * n16, u[b], nc, c->Mz,
                           c->m16);
    HIPCHK(hipGetLastError());
    return NS3D_OK;
}

// ---------------------------------------------------------------- standalone 3D transforms
template <typename TB>
int c2r_3d(ns3d_ctx* c, const cplx<TB>* in, TB* out, cplx<TB>* tmp, bool scale) {
    hipStream_t s = c->stream;
    const long nc = c->nc;
    int st = with_size(c->Nx, [&](auto N) -> int {
        constexpr int NX = decltype(N)::value;
        constexpr int E = E_C2C<TB>;
        constexpr int P = PXY<TB, NX, E>;
        const int nlt = (c->Mz + P - 1) / P;
        hipLaunchKernelGGL((k_axis_c2c<TB, NX, P, +1, 0, E>), xgrid(c->Ny * nlt), dim3(P * XYPlan<NX, E>::TPP), 0, s, in, tmp, 1, nc,
                           c->Ny, c->Mz, c->Mz, nlt, c->Ny * nlt, twp<TB>(c, 0));
        return NS3D_OK;
    });
    if (st) return st;
    st = with_size(c->Ny, [&](auto N) -> int {
        constexpr int NY = decltype(N)::value;
        constexpr int E = E_C2C<TB>;
        constexpr int P = PXY<TB, NY, E>;
        const int nlt = (c->Mz + P - 1) / P;
        hipLaunchKernelGGL((k_axis_c2c<TB, NY, P, +1, 1, E>), xgrid(c->Nx * nlt), dim3(P * XYPlan<NY, E>::TPP), 0, s, tmp, tmp, 1,
                           nc, c->Ny, c->Mz, c->Mz, nlt, c->Nx * nlt, twp<TB>(c, 1));
        return NS3D_OK;
    });
    if (st) return st;
    return with_size(c->Nz, [&](auto N) -> int {
        constexpr int NZ = decltype(N)::value;
        constexpr int P = PZ<TB, NZ>;
        const long npairs = c->npen / 2;
        const TB sc = scale ? (TB)(1.0 / (1.0 * c->Nx * c->Ny * c->Nz)) : (TB)1;
        const int nb = (int)nblk(npairs, P);
        hipLaunchKernelGGL((k_z_c2r<TB, NZ, P>), xgrid(nb), dim3(P * ZPlanT<TB, NZ>::TPP), 0, s, tmp, out, npairs,
                           c->Mz, nb, c->tab, twp<TB>(c, 2), sc);
        return NS3D_OK;
    });
}

template <typename TB>
int r2c_3d(ns3d_ctx* c, const TB* in, cplx<TB>* out) {
    hipStream_t s = c->stream;
    const long nc = c->nc;
    int st = with_size(c->Nz, [&](auto N) -> int {
        constexpr int NZ = decltype(N)::value;
        constexpr int P = PZ<TB, NZ>;
        const long npairs = c->npen / 2;
        const int nb = (int)nblk(npairs, P);
        hipLaunchKernelGGL((k_z_r2c<TB, NZ, P>), xgrid(nb), dim3(P * ZPlanT<TB, NZ>::TPP), 0, s, in, out, npairs,
                           c->Mz, nb, twp<TB>(c, 2));
        return NS3D_OK;
    });
    if (st) return st;
    st = with_size(c->Ny, [&](auto N) -> int {
        constexpr int NY = decltype(N)::value;
        constexpr int E = E_C2C<TB>;
        constexpr int P = PXY<TB, NY, E>;
        const int nlt = (c->Mz + P - 1) / P;
        hipLaunchKernelGGL((k_axis_c2c<TB, NY, P, -1, 1, E>), xgrid(c->Nx * nlt), dim3(P * XYPlan<NY, E>::TPP), 0, s, out, out, 1,
                           nc, c->Ny, c->Mz, c->Mz, nlt, c->Nx * nlt, twp<TB>(c, 1));
        return NS3D_OK;
    });
    if (st) return st;
    return with_size(c->Nx, [&](auto N) -> int {
        constexpr int NX = decltype(N)::value;
        constexpr int E = E_C2C<TB>;
        constexpr int P = PXY<TB, NX, E>;
        const int nlt = (c->Mz + P - 1) / P;
        hipLaunchKernelGGL((k_axis_c2c<TB, NX, P, -1, 0, E>), xgrid(c->Ny * nlt), dim3(P * XYPlan<NX, E>::TPP), 0, s, out, out, 1,
                           nc, c->Ny, c->Mz, c->Mz, nlt, c->Ny * nlt, twp<TB>(c, 0));
        return NS3D_OK;
    });
}

// FP64-interface 3D transforms (FP32 bulk: convert, transform, convert back)
int c2r_any(ns3d_ctx* c, const zd* in, double* out, bool scale) {
    int st = ensure_scratch(c, 15);
    if (st) return st;
    if (c->precision == NS3D_FP64) {
        st = c2r_3d<double>(c, in, out, (zd*)c->scratch, scale);
    } else {
        cplx<float>* a = (cplx<float>*)c->scratch;
        cplx<float>* b = a + c->nc;
        float* r = (float*)(b + c->nc);  // nr floats fit in 2 complex float fields (nr <= 2*nc)
        hipLaunchKernelGGL(k_d2f_c, dim3(nblk(c->nc, 256)), dim3(256), 0, c->stream, in, a, c->nc);
        st = c2r_3d<float>(c, a, r, b, scale);
        if (st) return st;
        hipLaunchKernelGGL(k_f2d_r, dim3(nblk(c->nr, 256)), dim3(256), 0, c->stream, r, out, c->nr);
    }
    if (st) return st;
    return hipGetLastError() == hipSuccess ? NS3D_OK : NS3D_ERR_HIP;
}

int r2c_any(ns3d_ctx* c, const double* in, zd* out) {
    int st;
    if (c->precision == NS3D_FP64) {
        st = r2c_3d<double>(c, in, out);
    } else {
        st = ensure_scratch(c, 15);
        if (st) return st;
        cplx<float>* a = (cplx<float>*)c->scratch;
        float* r = (float*)(a + c->nc);
        hipLaunchKernelGGL(k_d2f_r, dim3(nblk(c->nr, 256)), dim3(256), 0, c->stream, in, r, c->nr);
        st = r2c_3d<float>(c, r, a);
        if (st) return st;
        hipLaunchKernelGGL(k_f2d_c, dim3(nblk(c->nc, 256)), dim3(256), 0, c->stream, a, out, c->nc);
    }
    if (st) return st;
    return hipGetLastError() == hipSuccess ? NS3D_OK : NS3D_ERR_HIP;
}

// ---------------------------------------------------------------- reductions
template <int OP, int SRC>
int reduce(ns3d_ctx* c, const RedArgs& r, long n, double* host_out) {
    hipStream_t s = c->stream;
    int nb = (int)std::min<long>(RED_BLOCKS, (n + RED_THREADS - 1) / RED_THREADS);
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL((k_reduce_partial<OP, SRC>), dim3(nb), dim3(RED_THREADS), 0, s, r, n, c->red);
    hipLaunchKernelGGL((k_reduce_final<OP>), dim3(1), dim3(RED_THREADS), 0, s, c->red, nb, c->red + 2 * RED_BLOCKS);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->red_host, c->red + 2 * RED_BLOCKS, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    host_out[0] = c->red_host[0];
    if (OP == 1) host_out[1] = c->red_host[1];
    return NS3D_OK;
}

RedArgs red_args(const ns3d_ctx* c, const double* a, const double* b, const double* d) {
    RedArgs r;
    r.a = a;
    r.b = b;
    r.c = d;
    r.G = Grid3{c->Nx, c->Ny, c->Nz};
    r.dx = c->Lx / c->Nx;
    r.dy = c->Ly / c->Ny;
    r.dz = c->Lz / c->Nz;
    r.p0 = 0.0;
    r.p1 = c->Re;
    r.kx = c->tab.kx;
    r.ky = c->tab.ky;
    r.kz = c->tab.kz;
    r.Mz = c->Mz;
    return r;
}

// ---------------------------------------------------------------- passive scalar RK3
template <int SCHEME, typename TU>
void launch_weno_stage(ns3d_ctx* c, const double* xs, const double* xn, double* out, double a0, double a1) {
    const Grid3 G{c->Nx, c->Ny, c->Nz};
    const int zfix = (c->flags & NS3D_FLAG_WENO_ZFIX) ? 1 : 0;
    hipLaunchKernelGGL((k_weno_stage<SCHEME, TU>), xgrid(nblk(c->nr, 256)), dim3(256), 0, c->stream,
                       (const TU*)c->uphys, c->nr, xs, xn, out, G, c->Nx / c->Lx, c->Ny / c->Ly, c->Nz / c->Lz, zfix,
                       c->dt, a0, a1, c->nf);
}
template <typename TU>
void launch_weno_stage_s(ns3d_ctx* c, int scheme, const double* xs, const double* xn, double* out, double a0,
                         double a1) {
    if (scheme == 5) launch_weno_stage<5, TU>(c, xs, xn, out, a0, a1);
    else if (scheme == 3) launch_weno_stage<3, TU>(c, xs, xn, out, a0, a1);
    else launch_weno_stage<0, TU>(c, xs, xn, out, a0, a1);
}
int weno_stage(ns3d_ctx* c, int scheme, const double* xs, const double* xn, double* out, double a0, double a1) {
    if (c->precision == NS3D_FP64) launch_weno_stage_s<double>(c, scheme, xs, xn, out, a0, a1);
    else launch_weno_stage_s<float>(c, scheme, xs, xn, out, a0, a1);
    return hipGetLastError() == hipSuccess ? NS3D_OK : NS3D_ERR_HIP;
}

int rk3_scalar_one(ns3d_ctx* c, zd* const u[3], const zd* const f[3], double* th, int scheme) {
    int st = ensure_fp64(c, &c->stage, 9);
    if (st) return st;
    if (!c->uphys && hipMalloc(&c->uphys, bulk_bytes(c) / 2 * (size_t)c->nr * 3) != hipSuccess) return NS3D_ERR_ALLOC;
    if (!c->theta_stage && hipMalloc((void**)&c->theta_stage, sizeof(double) * (size_t)c->nr * 2) != hipSuccess)
        return NS3D_ERR_ALLOC;
    zd* u1[3] = {c->stage, c->stage + c->nc, c->stage + 2 * c->nc};
    zd* u2[3] = {c->stage + 3 * c->nc, c->stage + 4 * c->nc, c->stage + 5 * c->nc};
    const zd* uc[3] = {u[0], u[1], u[2]};
    const zd* u1c[3] = {u1[0], u1[1], u1[2]};
    const zd* u2c[3] = {u2[0], u2[1], u2[2]};
    double* th1 = c->theta_stage;
    double* th2 = c->theta_stage + c->nr;
    c->uphys_out = c->uphys;
    // the velocity stages write their physical flow; the scalar stage follows with the same SSP weights
    if (!(st = stage_into(c, uc, nullptr, 0.0, 1.0, u1, f))) st = weno_stage(c, scheme, th, nullptr, th1, 0.0, 1.0);
    if (!st && !(st = stage_into(c, u1c, uc, 3.0 / 4.0, 1.0 / 4.0, u2, f)))
        st = weno_stage(c, scheme, th1, th, th2, 3.0 / 4.0, 1.0 / 4.0);
    if (!st && !(st = stage_into(c, u2c, uc, 1.0 / 3.0, 2.0 / 3.0, u, f)))
        st = weno_stage(c, scheme, th2, th, th, 1.0 / 3.0, 2.0 / 3.0);
    c->uphys_out = nullptr;
    return st;
}

// Read the sticky non-finite flag (synchronises the context stream); clear it if asked.
int finite_status(ns3d_ctx* c, bool clear) {
    HIPCHK(hipMemcpyAsync(c->nf_host, c->nf, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const int v = *c->nf_host;
    if (v && clear) {
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, c->stream, c->nf, 0);
        HIPCHK(hipGetLastError());
    }
    return v ? NS3D_ERR_NONFINITE : NS3D_OK;
}

// Status of an RK3 entry point: with NS3D_FLAG_CHECK_FINITE, a successful call
// ends with a read (and clear) of the non-finite flag.
int checked(ns3d_ctx* c, int st) {
    if (st || !(c->flags & NS3D_FLAG_CHECK_FINITE)) return st;
    return finite_status(c, true);
}

// ---- allocation-free RK3 steps (ns3d_rk3_ssp_steps_noalloc)
// The fused passes' scratch (15 bulk fields) and the RK stage buffers (6 FP64
// fields, or the 9 FP32 fields of NS3D_FP32_STATE / NS3D_FP32_INCR) are carved out
// of the file-scope device array `scratch` instead of the context's own buffers.
// Host code never passes the array's name to a kernel: scratch_ptr[dev] is the
// device address of device dev's copy, looked up once per device by
// ns3d_create with that device current.
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex scratch_mu;
void* scratch_ptr[NOALLOC_MAX_DEVICES] = {};

size_t round256(size_t b) { return (b + 255) / 256 * 256; }

size_t noalloc_scratch_bytes(const ns3d_ctx* c) {
    const size_t passes = round256(bulk_bytes(c) * (size_t)c->nci * 15);
    if (uses_s32(c)) return passes + sizeof(cplx<float>) * (size_t)c->nc16 * 9;
    return passes + sizeof(zd) * (size_t)c->nc8 * 6;
}

// The context's scratch, RK stage and FP32 state pointers aimed at the static
// array for the duration of one call (restored on exit); per-pass timing, which
// creates events, is off meanwhile.  The fused passes then find every buffer
// already present and allocate nothing.
struct ScratchView {
    ns3d_ctx* c;
    void* scratch_;
    int fields_;
    zd* stage_;
    cplx<float>* s32_;
    bool prof_;
    ScratchView(ns3d_ctx* cc, unsigned char* base)
        : c(cc), scratch_(cc->scratch), fields_(cc->scratch_fields), stage_(cc->stage), s32_(cc->s32), prof_(cc->prof_on) {
        unsigned char* rk = base + round256(bulk_bytes(c) * (size_t)c->nci * 15);
        c->scratch = base;
        c->scratch_fields = 15;
        c->stage = (zd*)rk;
        c->s32 = (cplx<float>*)rk;
        c->prof_on = false;
    }
    ~ScratchView() {
        c->scratch = scratch_;
        c->scratch_fields = fields_;
        c->stage = stage_;
        c->s32 = s32_;
        c->prof_on = prof_;
    }
};

int ensure_kv(ns3d_ctx* c) {
    if (c->kv_partial) return NS3D_OK;
    if (hipMalloc((void**)&c->kv_partial, sizeof(double) * KV_COLS * KV_BLOCKS) != hipSuccess ||
        hipMalloc((void**)&c->kv_out, sizeof(double) * KV_MAXM) != hipSuccess ||
        hipMalloc((void**)&c->kq, sizeof(double) * KV_MAXM * KV_MAXM) != hipSuccess ||
        hipHostMalloc((void**)&c->kv_host, sizeof(double) * KV_MAXM) != hipSuccess ||
        hipHostMalloc((void**)&c->kq_host, sizeof(double) * KV_MAXM * KV_MAXM) != hipSuccess)
        return NS3D_ERR_ALLOC;
    return NS3D_OK;
}

}  // namespace

// =====================================================================================
extern "C" {

const char* ns3d_status_string(int s) {
    switch (s) {
        case NS3D_OK: return "ok";
        case NS3D_ERR_ARG: return "invalid argument";
        case NS3D_ERR_SIZE: return "grid size has no compiled FFT plan";
        case NS3D_ERR_ALLOC: return "device allocation failed";
        case NS3D_ERR_HIP: return "HIP runtime error";
        case NS3D_ERR_NONFINITE: return "non-finite value";
        case NS3D_ERR_PRECISION: return "operation not offered at this precision";
        case NS3D_ERR_CAPACITY: return "grid too large for the fixed scratch of the allocation-free entry point";
        default: return "unknown status";
    }
}

int ns3d_supported_size(int n) { return size_ok(n) ? 1 : 0; }

int ns3d_create(const ns3d_config* cfg, ns3d_ctx** out) {
    if (!cfg || !out) return NS3D_ERR_ARG;
    *out = nullptr;
    if (!size_ok(cfg->Nx) || !size_ok(cfg->Ny) || !size_ok(cfg->Nz)) return NS3D_ERR_SIZE;
    if (!(cfg->Lx > 0) || !(cfg->Ly > 0) || !(cfg->Lz > 0) || !(cfg->Re > 0)) return NS3D_ERR_ARG;
    if (cfg->precision != NS3D_FP64 && cfg->precision != NS3D_FP32_BULK && cfg->precision != NS3D_FP32_STATE &&
        cfg->precision != NS3D_FP32_INCR)
        return NS3D_ERR_ARG;
    if (cfg->device >= 0) HIPCHK(hipSetDevice(cfg->device));
    ns3d_ctx* c = new (std::nothrow) ns3d_ctx();
    if (!c) return NS3D_ERR_ALLOC;
    c->Nx = cfg->Nx;
    c->Ny = cfg->Ny;
    c->Nz = cfg->Nz;
    c->Mz = cfg->Nz / 2 + 1;
    c->nc = (long)c->Nx * c->Ny * c->Mz;
    c->nr = (long)c->Nx * c->Ny * c->Nz;
    c->npen = (long)c->Nx * c->Ny;
    c->Lx = cfg->Lx;
    c->Ly = cfg->Ly;
    c->Lz = cfg->Lz;
    c->Re = cfg->Re;
    c->dt = cfg->dt;
    c->precision = cfg->precision;
    c->flags = cfg->flags;
    c->nci = (long)c->Nx * c->Ny * mi_of(c);
    c->m8 = NS3D_PAD_ROWS ? (c->Mz + NS3D_ROWPAD64 - 1) / NS3D_ROWPAD64 * NS3D_ROWPAD64 : c->Mz;
    c->m16 = NS3D_PAD_ROWS ? (c->Mz + NS3D_ROWPAD32 - 1) / NS3D_ROWPAD32 * NS3D_ROWPAD32 : c->Mz;
    c->nc8 = (long)c->Nx * c->Ny * c->m8;
    c->nc16 = (long)c->Nx * c->Ny * c->m16;
    c->stream = (hipStream_t)cfg->stream;
    hipGetDevice(&c->device);
    c->last_error = NS3D_OK;
    int st = build_tables(c);
    if (!st) st = build_twiddles(c);
    if (!st && hipMalloc((void**)&c->red, sizeof(double) * (2 * RED_BLOCKS + 2)) != hipSuccess) st = NS3D_ERR_ALLOC;
    if (!st && hipHostMalloc((void**)&c->red_host, 2 * sizeof(double)) != hipSuccess) st = NS3D_ERR_ALLOC;
    if (!st && hipMalloc((void**)&c->nf, 64 * sizeof(int)) != hipSuccess) st = NS3D_ERR_ALLOC;
    if (!st && hipHostMalloc((void**)&c->nf_host, sizeof(int)) != hipSuccess) st = NS3D_ERR_ALLOC;
    if (!st) {
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, c->stream, c->nf, 0);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) st = NS3D_ERR_HIP;
    }
    if (!st) {
        // once per device, with the device current: the address of this GPU's
        // copy of the allocation-free entry point's scratch
        std::lock_guard<std::mutex> lock(scratch_mu);
        if (c->device >= 0 && c->device < NOALLOC_MAX_DEVICES && !scratch_ptr[c->device]) {
            void* p = nullptr;
            if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess && p)
                scratch_ptr[c->device] = p;
            else
                (void)hipGetLastError();  // the noalloc entry point then reports NS3D_ERR_HIP
        }
    }
    if (st) {
        ns3d_destroy(c);
        return st;
    }
    *out = c;
    return NS3D_OK;
}

int ns3d_destroy(ns3d_ctx* c) {
    if (!c) return NS3D_ERR_ARG;
    hipStreamSynchronize(c->stream);
    hipFree(c->dtab);
    for (int a = 0; a < 3; a++) {
        hipFree(c->tw64[a]);
        hipFree(c->tw32[a]);
    }
    hipFree(c->scratch);
    hipFree(c->stage);
    hipFree(c->bstage);
    hipFree(c->s32);
    hipFree(c->extra);
    hipFree(c->red);
    hipFree(c->real_scratch);
    hipFree(c->kry);
    hipFree(c->kv_partial);
    hipFree(c->kv_out);
    hipFree(c->kq);
    if (c->kv_host) hipHostFree(c->kv_host);
    if (c->kq_host) hipHostFree(c->kq_host);
    hipFree(c->uphys);
    hipFree(c->theta_stage);
    if (c->red_host) hipHostFree(c->red_host);
    hipFree(c->nf);
    if (c->nf_host) hipHostFree(c->nf_host);
    hipFree(c->pb);
    if (c->pb_host) hipHostFree(c->pb_host);
    for (hipEvent_t e : c->ev_pool) hipEventDestroy(e);
    drop_graph(c);
    delete c;
    return NS3D_OK;
}

int ns3d_set_params(ns3d_ctx* c, double dt, double Re) {
    if (!c || !(Re > 0)) return NS3D_ERR_ARG;
    c->dt = dt;
    c->Re = Re;
    return NS3D_OK;
}

int ns3d_set_stream(ns3d_ctx* c, void* stream) {
    if (!c) return NS3D_ERR_ARG;
    c->stream = (hipStream_t)stream;
    return NS3D_OK;
}

int ns3d_synchronize(ns3d_ctx* c) {
    if (!c) return NS3D_ERR_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    return NS3D_OK;
}

int ns3d_last_error(const ns3d_ctx* c) { return c ? c->last_error : NS3D_ERR_ARG; }

long long ns3d_workspace_generation(const ns3d_ctx* c) { return c ? c->generation : -1; }

long long ns3d_workspace_bytes(const ns3d_ctx* c) {
    if (!c) return -1;
    long long b = (long long)bulk_bytes(c) * c->nci * c->scratch_fields;
    if (c->stage) b += (long long)sizeof(zd) * c->nc8 * 9;
    if (c->extra) b += (long long)sizeof(zd) * c->nc8 * 6;
    if (c->kry) b += (long long)sizeof(zd) * c->nc8 * 3;
    if (c->s32) b += (long long)sizeof(cplx<float>) * c->nc16 * 9;
    if (c->real_scratch) b += (long long)sizeof(double) * c->nr * 4;
    return b;
}

// ---- a4
int ns3d_fft_r2c(ns3d_ctx* c, const double* in, ns3d_cplx* out) {
    if (!c || !in || !out) return NS3D_ERR_ARG;
    return r2c_any(c, in, (zd*)out);
}
int ns3d_fft_c2r(ns3d_ctx* c, const ns3d_cplx* in, double* out) {
    if (!c || !in || !out) return NS3D_ERR_ARG;
    return c2r_any(c, (const zd*)in, out, true);
}
int ns3d_fft_c2r_unscaled(ns3d_ctx* c, const ns3d_cplx* in, double* out) {
    if (!c || !in || !out) return NS3D_ERR_ARG;
    return c2r_any(c, (const zd*)in, out, false);
}

// ---- a5 (on-the-fly table recomposition written out, same device helpers as the kernels)
}  // extern "C"

namespace {
__global__ void k_tables(SpecTables tab, long nc, int Nx, int Ny, int Mz, double* din_diffusion, double* din_poisson,
                         double* a11, double* a22, double* a33, double* a12, double* a13, double* a23, double* mask,
                         double* kx, double* ky, double* kz) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < Nx && kx) kx[g] = tab.kx[g];
    if (g < Ny && ky) ky[g] = tab.ky[g];
    if (g < Mz && kz) kz[g] = tab.kz[g];
    if (g >= nc) return;
    const int l = (int)(g % Mz);
    const long t1 = g / Mz;
    const int k = (int)(t1 % Ny), j = (int)(t1 / Ny);
    const ProjCoef c = proj_coef<true>(tab, j, k, l);
    if (din_diffusion) din_diffusion[g] = c.din;
    if (din_poisson) din_poisson[g] = g == 0 ? 1.0 : -c.din;
    if (a11) {
        a11[g] = c.a11; a22[g] = c.a22; a33[g] = c.a33;
        a12[g] = c.a12; a13[g] = c.a13; a23[g] = c.a23;
    }
    if (mask) mask[g] = mask_keep(tab, j, k, l) ? 1.0 : 0.0;
}
}  // namespace

extern "C" {

int ns3d_tables(ns3d_ctx* c, double* kx, double* ky, double* kz, double* din_diffusion, double* din_poisson,
                double* AM11, double* AM22, double* AM33, double* AM12, double* AM13, double* AM23, double* mask) {
    if (!c) return NS3D_ERR_ARG;
    if (AM11 && !(AM22 && AM33 && AM12 && AM13 && AM23)) return NS3D_ERR_ARG;
    hipLaunchKernelGGL(k_tables, dim3(nblk(c->nc, 256)), dim3(256), 0, c->stream, c->tab, c->nc, c->Nx, c->Ny, c->Mz,
                       din_diffusion, din_poisson, AM11, AM22, AM33, AM12, AM13, AM23, mask, kx, ky, kz);
    HIPCHK(hipGetLastError());
    return NS3D_OK;
}

// ---- a3
int ns3d_nonlinear(ns3d_ctx* c, const ns3d_cplx* ux, const ns3d_cplx* uy, const ns3d_cplx* uz, ns3d_cplx* Qx,
                   ns3d_cplx* Qy, ns3d_cplx* Qz) {
    if (!c || !ux || !uy || !uz || !Qx || !Qy || !Qz) return NS3D_ERR_ARG;
    const zd* in[3] = {(const zd*)ux, (const zd*)uy, (const zd*)uz};
    EpiArgs ea = make_epi(c);
    ea.out[0] = (zd*)Qx;
    ea.out[1] = (zd*)Qy;
    ea.out[2] = (zd*)Qz;
    return fused<EPI_Q>(c, in, 1, ea);
}

int ns3d_bilinear(ns3d_ctx* c, const ns3d_cplx* ux, const ns3d_cplx* uy, const ns3d_cplx* uz, const ns3d_cplx* vx,
                  const ns3d_cplx* vy, const ns3d_cplx* vz, ns3d_cplx* Qx, ns3d_cplx* Qy, ns3d_cplx* Qz) {
    if (!c || !ux || !uy || !uz || !vx || !vy || !vz || !Qx || !Qy || !Qz) return NS3D_ERR_ARG;
    const zd* in[6] = {(const zd*)ux, (const zd*)uy, (const zd*)uz, (const zd*)vx, (const zd*)vy, (const zd*)vz};
    EpiArgs ea = make_epi(c);
    ea.out[0] = (zd*)Qx;
    ea.out[1] = (zd*)Qy;
    ea.out[2] = (zd*)Qz;
    return fused<EPI_Q>(c, in, 2, ea);
}

// ---- a6
int ns3d_project_update(ns3d_ctx* c, int mode, ns3d_cplx* ux, ns3d_cplx* uy, ns3d_cplx* uz, const ns3d_cplx* fx,
                        const ns3d_cplx* fy, const ns3d_cplx* fz, const ns3d_cplx* Qx, const ns3d_cplx* Qy,
                        const ns3d_cplx* Qz, ns3d_cplx* Rx, ns3d_cplx* Ry, ns3d_cplx* Rz) {
    if (!c || !ux || !uy || !uz || !Qx || !Qy || !Qz) return NS3D_ERR_ARG;
    EpiArgs ea = make_epi(c);
    zd* w[3] = {(zd*)ux, (zd*)uy, (zd*)uz};
    const zd* f[3] = {(const zd*)fx, (const zd*)fy, (const zd*)fz};
    for (int b = 0; b < 3; b++) {
        ea.w[b] = w[b];
        ea.f[b] = f[b];
        ea.out[b] = w[b];
    }
    dim3 g(nblk(c->nc, 256)), t(256);
    const zd *qx = (const zd*)Qx, *qy = (const zd*)Qy, *qz = (const zd*)Qz;
    switch (mode) {
        case NS3D_UPD_STAGE:
            hipLaunchKernelGGL(k_update<EPI_STAGE>, g, t, 0, c->stream, qx, qy, qz, c->nc, c->Ny, c->Mz, c->tab, ea);
            break;
        case NS3D_UPD_UV_RHS:
            hipLaunchKernelGGL(k_update<EPI_UV_RHS>, g, t, 0, c->stream, qx, qy, qz, c->nc, c->Ny, c->Mz, c->tab, ea);
            break;
        case NS3D_UPD_UV:
            hipLaunchKernelGGL(k_update<EPI_UV>, g, t, 0, c->stream, qx, qy, qz, c->nc, c->Ny, c->Mz, c->tab, ea);
            break;
        case NS3D_UPD_IUV:
            hipLaunchKernelGGL(k_update<EPI_IUV>, g, t, 0, c->stream, qx, qy, qz, c->nc, c->Ny, c->Mz, c->tab, ea);
            break;
        case NS3D_UPD_RHS:
            if (!Rx || !Ry || !Rz) return NS3D_ERR_ARG;
            ea.out[0] = (zd*)Rx;
            ea.out[1] = (zd*)Ry;
            ea.out[2] = (zd*)Rz;
            hipLaunchKernelGGL(k_update<EPI_RHS>, g, t, 0, c->stream, qx, qy, qz, c->nc, c->Ny, c->Mz, c->tab, ea);
            break;
        default: return NS3D_ERR_ARG;
    }
    HIPCHK(hipGetLastError());
    return NS3D_OK;
}

// ---- a7
int ns3d_copy3(ns3d_ctx* c, const ns3d_cplx* sx, const ns3d_cplx* sy, const ns3d_cplx* sz, ns3d_cplx* dx,
               ns3d_cplx* dy, ns3d_cplx* dz) {
    if (!c || !sx || !sy || !sz || !dx || !dy || !dz) return NS3D_ERR_ARG;
    PtrArr<const zd, 3> s{{(const zd*)sx, (const zd*)sy, (const zd*)sz}};
    PtrArr<zd, 3> d{{(zd*)dx, (zd*)dy, (zd*)dz}};
    hipLaunchKernelGGL(k_copy3, dim3(nblk(c->nc, 256)), dim3(256), 0, c->stream, s, d, c->nc);
    HIPCHK(hipGetLastError());
    return NS3D_OK;
}

int ns3d_axpby3(ns3d_ctx* c, const ns3d_cplx* ax, const ns3d_cplx* ay, const ns3d_cplx* az, double wa,
                const ns3d_cplx* bx, const ns3d_cplx* by, const ns3d_cplx* bz, double wb, ns3d_cplx* dx,
                ns3d_cplx* dy, ns3d_cplx* dz) {
    if (!c || !ax || !ay || !az || !bx || !by || !bz || !dx || !dy || !dz) return NS3D_ERR_ARG;
    PtrArr<const zd, 3> a{{(const zd*)ax, (const zd*)ay, (const zd*)az}};
    PtrArr<const zd, 3> b{{(const zd*)bx, (const zd*)by, (const zd*)bz}};
    PtrArr<zd, 3> d{{(zd*)dx, (zd*)dy, (zd*)dz}};
    hipLaunchKernelGGL(k_axpby3, dim3(nblk(c->nc, 256)), dim3(256), 0, c->stream, a, wa, b, wb, d, c->nc);
    HIPCHK(hipGetLastError());
    return NS3D_OK;
}

// ---- a2 / a1
int ns3d_stage(ns3d_ctx* c, ns3d_cplx* ux, ns3d_cplx* uy, ns3d_cplx* uz, const ns3d_cplx* fx, const ns3d_cplx* fy,
               const ns3d_cplx* fz) {
    if (!c || !ux || !uy || !uz) return NS3D_ERR_ARG;
    const zd* w[3] = {(const zd*)ux, (const zd*)uy, (const zd*)uz};
    zd* o[3] = {(zd*)ux, (zd*)uy, (zd*)uz};
    const zd* f[3] = {(const zd*)fx, (const zd*)fy, (const zd*)fz};
    return stage_into(c, w, nullptr, 0.0, 1.0, o, f);
}

int ns3d_rhs(ns3d_ctx* c, const ns3d_cplx* ux, const ns3d_cplx* uy, const ns3d_cplx* uz, const ns3d_cplx* fx,
             const ns3d_cplx* fy, const ns3d_cplx* fz, ns3d_cplx* Rx, ns3d_cplx* Ry, ns3d_cplx* Rz) {
    if (!c || !ux || !uy || !uz || !Rx || !Ry || !Rz) return NS3D_ERR_ARG;
    const zd* w[3] = {(const zd*)ux, (const zd*)uy, (const zd*)uz};
    EpiArgs ea = make_epi(c);
    const zd* f[3] = {(const zd*)fx, (const zd*)fy, (const zd*)fz};
    zd* R[3] = {(zd*)Rx, (zd*)Ry, (zd*)Rz};
    for (int b = 0; b < 3; b++) {
        ea.w[b] = (zd*)w[b];
        ea.f[b] = f[b];
        ea.out[b] = R[b];
    }
    return fused<EPI_RHS>(c, w, 1, ea);
}

int ns3d_rk3_ssp(ns3d_ctx* c, ns3d_cplx* ux, ns3d_cplx* uy, ns3d_cplx* uz, const ns3d_cplx* fx, const ns3d_cplx* fy,
                 const ns3d_cplx* fz) {
    return ns3d_rk3_ssp_steps(c, 1, ux, uy, uz, fx, fy, fz);
}

int ns3d_rk3_ssp_steps(ns3d_ctx* c, int nsteps, ns3d_cplx* ux, ns3d_cplx* uy, ns3d_cplx* uz, const ns3d_cplx* fx,
                       const ns3d_cplx* fy, const ns3d_cplx* fz) {
    if (!c || !ux || !uy || !uz || nsteps < 0) return NS3D_ERR_ARG;
    zd* u[3] = {(zd*)ux, (zd*)uy, (zd*)uz};
    const zd* f[3] = {(const zd*)fx, (const zd*)fy, (const zd*)fz};
    return checked(c, rk3_steps(c, nsteps, u, f));
}

int ns3d_rk3_ssp_steps_noalloc(ns3d_ctx* c, int nsteps, ns3d_cplx* ux, ns3d_cplx* uy, ns3d_cplx* uz,
                               const ns3d_cplx* fx, const ns3d_cplx* fy, const ns3d_cplx* fz) {
    if (!c || !ux || !uy || !uz || nsteps < 0) return NS3D_ERR_ARG;
    if (noalloc_scratch_bytes(c) > NS3D_NOALLOC_SCRATCH_BYTES) return NS3D_ERR_CAPACITY;
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(scratch_mu);
        if (c->device >= 0 && c->device < NOALLOC_MAX_DEVICES) base = scratch_ptr[c->device];
    }
    if (!base) return NS3D_ERR_HIP;
    zd* u[3] = {(zd*)ux, (zd*)uy, (zd*)uz};
    const zd* f[3] = {(const zd*)fx, (const zd*)fy, (const zd*)fz};
    int st;
    {
        ScratchView view(c, (unsigned char*)base);
        st = rk3_steps(c, nsteps, u, f);
    }
    return checked(c, st);
}

int ns3d_rk3_ssp_batch(ns3d_ctx* c, int nbatch, int nsteps, ns3d_cplx* u, const ns3d_cplx* fx, const ns3d_cplx* fy,
                       const ns3d_cplx* fz) {
    if (!c || !u || nbatch < 1 || nsteps < 0) return NS3D_ERR_ARG;
    if (nbatch > NS3D_MAX_BATCH) return NS3D_ERR_SIZE;
    const zd* f[3] = {(const zd*)fx, (const zd*)fy, (const zd*)fz};
    zd* ub = (zd*)u;
    if (uses_s32(c)) {
        // the FP32 state / increment loop keeps one member's stages: members run one after another
        for (int b = 0; b < nbatch; b++) {
            zd* m[3] = {ub + 3 * b * c->nc, ub + (3 * b + 1) * c->nc, ub + (3 * b + 2) * c->nc};
            int st = rk3_steps(c, nsteps, m, f);
            if (st) return st;
        }
        return checked(c, NS3D_OK);
    }
    for (int i = 0; i < nsteps; i++) {
        int st = rk3_one_batch(c, nbatch, ub, f);
        if (st) return st;
    }
    return checked(c, NS3D_OK);
}

int ns3d_rk3_ssp_batch_members(ns3d_ctx* c, int nbatch, int nsteps, ns3d_cplx* u, const ns3d_cplx* fx,
                               const ns3d_cplx* fy, const ns3d_cplx* fz, const double* dt, const double* Re,
                               const int* steps, int graph) {
    if (!c || !u || !dt || !Re || !steps || nbatch < 1 || nsteps < 0) return NS3D_ERR_ARG;
    if (nbatch > NS3D_MAX_BATCH) return NS3D_ERR_SIZE;
    if (c->precision != NS3D_FP64 && c->precision != NS3D_FP32_BULK) return NS3D_ERR_PRECISION;
    const zd* f[3] = {(const zd*)fx, (const zd*)fy, (const zd*)fz};
    if (!graph) return checked(c, rk3_batch_members(c, nbatch, nsteps, (zd*)u, f, dt, Re, steps));
    if (c->stream == nullptr || nsteps < 1) return NS3D_ERR_ARG;  // capture needs a created stream
    const void* key[8] = {u, fx, fy, fz, dt, Re, steps, nullptr};
    bool same = c->mgraph_exec && c->mgraph_nb == nbatch && c->mgraph_steps == nsteps;
    for (int i = 0; same && i < 8; i++) same = key[i] == c->mgraph_key[i];
    if (!same) {
        drop_graph(c);
        int st = ensure_batch(c, nbatch);  // every buffer the steps use exists before capture
        if (st) return st;
        HIPCHK(hipStreamSynchronize(c->stream));
        const bool prof = c->prof_on;
        c->prof_on = false;
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        st = rk3_batch_members(c, nbatch, nsteps, (zd*)u, f, dt, Re, steps);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        c->prof_on = prof;
        if (st || e != hipSuccess) {
            if (g) hipGraphDestroy(g);
            return st ? st : NS3D_ERR_HIP;
        }
        c->mgraph = g;
        if (hipGraphInstantiate(&c->mgraph_exec, g, nullptr, nullptr, 0) != hipSuccess) {
            hipGraphDestroy(g);
            c->mgraph = nullptr;
            c->mgraph_exec = nullptr;
            return NS3D_ERR_HIP;
        }
        for (int i = 0; i < 8; i++) c->mgraph_key[i] = key[i];
        c->mgraph_nb = nbatch;
        c->mgraph_steps = nsteps;
    }
    HIPCHK(hipGraphLaunch(c->mgraph_exec, c->stream));
    return checked(c, NS3D_OK);
}

int ns3d_rk3_ssp_scalar(ns3d_ctx* c, int nsteps, ns3d_cplx* ux, ns3d_cplx* uy, ns3d_cplx* uz, const ns3d_cplx* fx,
                        const ns3d_cplx* fy, const ns3d_cplx* fz, double* theta, int scheme) {
    if (!c || !ux || !uy || !uz || !theta || nsteps < 0) return NS3D_ERR_ARG;
    zd* u[3] = {(zd*)ux, (zd*)uy, (zd*)uz};
    const zd* f[3] = {(const zd*)fx, (const zd*)fy, (const zd*)fz};
    for (int i = 0; i < nsteps; i++) {
        int st = rk3_scalar_one(c, u, f, theta, scheme);
        if (st) return st;
    }
    return checked(c, NS3D_OK);
}

int ns3d_rk3_ssp_graph(ns3d_ctx* c, int nsteps, ns3d_cplx* ux, ns3d_cplx* uy, ns3d_cplx* uz, const ns3d_cplx* fx,
                       const ns3d_cplx* fy, const ns3d_cplx* fz) {
    if (!c || !ux || !uy || !uz || nsteps < 1) return NS3D_ERR_ARG;
    if (c->stream == nullptr) return NS3D_ERR_ARG;  // capture needs a created stream
    const void* key[6] = {ux, uy, uz, fx, fy, fz};
    bool same = c->graph_exec && c->graph_steps == nsteps && c->graph_dt == c->dt && c->graph_Re == c->Re;
    for (int i = 0; same && i < 6; i++) same = key[i] == c->graph_key[i];
    if (!same) {
        drop_graph(c);
        // every buffer the step uses is allocated before capture
        int st = ensure_scratch(c, 15);
        if (!st) st = uses_s32(c) ? ensure_state32(c) : ensure_fp64(c, &c->stage, 9);
        if (st) return st;
        HIPCHK(hipStreamSynchronize(c->stream));
        const bool prof = c->prof_on;
        c->prof_on = false;
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        zd* u[3] = {(zd*)ux, (zd*)uy, (zd*)uz};
        const zd* f[3] = {(const zd*)fx, (const zd*)fy, (const zd*)fz};
        st = rk3_steps(c, nsteps, u, f);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        c->prof_on = prof;
        if (st || e != hipSuccess) {
            if (g) hipGraphDestroy(g);
            return st ? st : NS3D_ERR_HIP;
        }
        c->graph = g;
        if (hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0) != hipSuccess) {
            hipGraphDestroy(g);
            c->graph = nullptr;
            c->graph_exec = nullptr;
            return NS3D_ERR_HIP;
        }
        for (int i = 0; i < 6; i++) c->graph_key[i] = key[i];
        c->graph_steps = nsteps;
        c->graph_dt = c->dt;
        c->graph_Re = c->Re;
    }
    HIPCHK(hipGraphLaunch(c->graph_exec, c->stream));
    return checked(c, NS3D_OK);
}

// ---- a8
int ns3d_divergence(ns3d_ctx* c, const ns3d_cplx* ux, const ns3d_cplx* uy, const ns3d_cplx* uz, ns3d_cplx* div_hat) {
    if (!c || !ux || !uy || !uz || !div_hat) return NS3D_ERR_ARG;
    PtrArr<const zd, 3> u{{(const zd*)ux, (const zd*)uy, (const zd*)uz}};
    hipLaunchKernelGGL(k_divergence, dim3(nblk(c->nc, 256)), dim3(256), 0, c->stream, u, (zd*)div_hat, c->nc, c->Ny,
                       c->Mz, c->tab, c->dt);
    HIPCHK(hipGetLastError());
    return NS3D_OK;
}

int ns3d_divergence_physical(ns3d_ctx* c, const ns3d_cplx* ux, const ns3d_cplx* uy, const ns3d_cplx* uz,
                             ns3d_cplx* div_hat, double* div) {
    int st = ns3d_divergence(c, ux, uy, uz, div_hat);
    if (st) return st;
    if (!div) return NS3D_ERR_ARG;
    return c2r_any(c, (const zd*)div_hat, div, true);
}

// ---- a9
int ns3d_weno_advect(ns3d_ctx* c, int scheme, const double* ux, const double* uy, const double* uz, const double* x,
                     double* out) {
    if (!c || !ux || !uy || !uz || !x || !out) return NS3D_ERR_ARG;
    const Grid3 G{c->Nx, c->Ny, c->Nz};
    const int zfix = (c->flags & NS3D_FLAG_WENO_ZFIX) ? 1 : 0;
    const double rdx = c->Nx / c->Lx, rdy = c->Ny / c->Ly, rdz = c->Nz / c->Lz;
    const dim3 grid = xgrid(nblk(c->nr, 256));
    if (scheme == 5)
        hipLaunchKernelGGL(k_weno<5>, grid, dim3(256), 0, c->stream, ux, uy, uz, x, out, G, rdx, rdy, rdz, zfix);
    else if (scheme == 3)
        hipLaunchKernelGGL(k_weno<3>, grid, dim3(256), 0, c->stream, ux, uy, uz, x, out, G, rdx, rdy, rdz, zfix);
    else
        hipLaunchKernelGGL(k_weno<0>, grid, dim3(256), 0, c->stream, ux, uy, uz, x, out, G, rdx, rdy, rdz, zfix);
    HIPCHK(hipGetLastError());
    return NS3D_OK;
}

int ns3d_weno_advection(ns3d_ctx* c, int scheme, const ns3d_cplx* ux, const ns3d_cplx* uy, const ns3d_cplx* uz,
                        ns3d_cplx* Qx, ns3d_cplx* Qy, ns3d_cplx* Qz) {
    if (!c || !ux || !uy || !uz || !Qx || !Qy || !Qz) return NS3D_ERR_ARG;
    int st = ensure_real_scratch(c);
    if (st) return st;
    double* r = (double*)c->real_scratch;
    const zd* u[3] = {(const zd*)ux, (const zd*)uy, (const zd*)uz};
    zd* Q[3] = {(zd*)Qx, (zd*)Qy, (zd*)Qz};
    for (int b = 0; b < 3; b++)
        if ((st = c2r_any(c, u[b], r + b * c->nr, true))) return st;
    for (int b = 0; b < 3; b++) {
        if ((st = ns3d_weno_advect(c, scheme, r, r + c->nr, r + 2 * c->nr, r + b * c->nr, r + 3 * c->nr))) return st;
        if ((st = r2c_any(c, r + 3 * c->nr, Q[b]))) return st;
    }
    return NS3D_OK;
}

// ---- a10
int ns3d_rk3_ssp_uv_rhs(ns3d_ctx* c, const ns3d_cplx* ux, const ns3d_cplx* uy, const ns3d_cplx* uz, ns3d_cplx* vx,
                        ns3d_cplx* vy, ns3d_cplx* vz) {
    if (!c || !ux || !uy || !uz || !vx || !vy || !vz) return NS3D_ERR_ARG;
    const zd* in[6] = {(const zd*)ux, (const zd*)uy, (const zd*)uz, (const zd*)vx, (const zd*)vy, (const zd*)vz};
    EpiArgs ea = make_epi(c);
    ea.w[0] = (zd*)vx;
    ea.w[1] = (zd*)vy;
    ea.w[2] = (zd*)vz;
    return fused<EPI_UV_RHS>(c, in, 2, ea);
}

int ns3d_stage_uv(ns3d_ctx* c, const ns3d_cplx* ux, const ns3d_cplx* uy, const ns3d_cplx* uz, ns3d_cplx* vx,
                  ns3d_cplx* vy, ns3d_cplx* vz) {
    if (!c || !ux || !uy || !uz || !vx || !vy || !vz) return NS3D_ERR_ARG;
    const zd* in[6] = {(const zd*)ux, (const zd*)uy, (const zd*)uz, (const zd*)vx, (const zd*)vy, (const zd*)vz};
    EpiArgs ea = make_epi(c);
    ea.w[0] = (zd*)vx;
    ea.w[1] = (zd*)vy;
    ea.w[2] = (zd*)vz;
    return fused<EPI_UV>(c, in, 2, ea);
}

int ns3d_stage_iuv_rhs(ns3d_ctx* c, const ns3d_cplx* ux, const ns3d_cplx* uy, const ns3d_cplx* uz, ns3d_cplx* vx,
                       ns3d_cplx* vy, ns3d_cplx* vz) {
    if (!c || !ux || !uy || !uz || !vx || !vy || !vz) return NS3D_ERR_ARG;
    const zd* in[6] = {(const zd*)ux, (const zd*)uy, (const zd*)uz, (const zd*)vx, (const zd*)vy, (const zd*)vz};
    EpiArgs ea = make_epi(c);
    ea.w[0] = (zd*)vx;
    ea.w[1] = (zd*)vy;
    ea.w[2] = (zd*)vz;
    return fused<EPI_IUV>(c, in, 2, ea);
}

int ns3d_rk3_ssp_uv(ns3d_ctx* c, ns3d_cplx* ux, ns3d_cplx* uy, ns3d_cplx* uz, ns3d_cplx* vx, ns3d_cplx* vy,
                    ns3d_cplx* vz, const ns3d_cplx* fx, const ns3d_cplx* fy, const ns3d_cplx* fz) {
    if (!c || !ux || !uy || !uz || !vx || !vy || !vz) return NS3D_ERR_ARG;
    int st = ensure_fp64(c, &c->extra, 6);
    if (st) return st;
    zd* v1[3] = {c->extra, c->extra + c->nc, c->extra + 2 * c->nc};
    // RK3_SSP_UV (Source/RK_time_step.cu:342-355)
    if ((st = ns3d_copy3(c, vx, vy, vz, (ns3d_cplx*)v1[0], (ns3d_cplx*)v1[1], (ns3d_cplx*)v1[2]))) return st;
    if ((st = ns3d_stage_uv(c, ux, uy, uz, (ns3d_cplx*)v1[0], (ns3d_cplx*)v1[1], (ns3d_cplx*)v1[2]))) return st;
    if ((st = ns3d_rk3_ssp(c, ux, uy, uz, fx, fy, fz))) return st;
    if ((st = ns3d_stage_uv(c, ux, uy, uz, (ns3d_cplx*)v1[0], (ns3d_cplx*)v1[1], (ns3d_cplx*)v1[2]))) return st;
    return ns3d_axpby3(c, vx, vy, vz, 0.5, (ns3d_cplx*)v1[0], (ns3d_cplx*)v1[1], (ns3d_cplx*)v1[2], 0.5, vx, vy, vz);
}

int ns3d_unpack_reduced(ns3d_ctx* c, const double* vec, ns3d_cplx* vx, ns3d_cplx* vy, ns3d_cplx* vz) {
    if (!c || !vec || !vx || !vy || !vz) return NS3D_ERR_ARG;
    PtrArr<zd, 3> v{{(zd*)vx, (zd*)vy, (zd*)vz}};
    hipLaunchKernelGGL(k_unpack_reduced, dim3(nblk(c->nc, 256)), dim3(256), 0, c->stream, vec, v, c->nc);
    HIPCHK(hipGetLastError());
    return NS3D_OK;
}
int ns3d_pack_reduced(ns3d_ctx* c, const ns3d_cplx* vx, const ns3d_cplx* vy, const ns3d_cplx* vz, double* vec) {
    if (!c || !vec || !vx || !vy || !vz) return NS3D_ERR_ARG;
    PtrArr<const zd, 3> v{{(const zd*)vx, (const zd*)vy, (const zd*)vz}};
    hipLaunchKernelGGL(k_pack_reduced, dim3(nblk(c->nc, 256)), dim3(256), 0, c->stream, v, vec, c->nc);
    HIPCHK(hipGetLastError());
    return NS3D_OK;
}
int ns3d_unpack_full(ns3d_ctx* c, const double* vec, ns3d_cplx* vx, ns3d_cplx* vy, ns3d_cplx* vz) {
    if (!c || !vec || !vx || !vy || !vz) return NS3D_ERR_ARG;
    PtrArr<zd, 3> v{{(zd*)vx, (zd*)vy, (zd*)vz}};
    hipLaunchKernelGGL(k_unpack_full, dim3(nblk(c->nc, 256)), dim3(256), 0, c->stream, vec, v, c->nc);
    HIPCHK(hipGetLastError());
    return NS3D_OK;
}
int ns3d_pack_full(ns3d_ctx* c, const ns3d_cplx* vx, const ns3d_cplx* vy, const ns3d_cplx* vz, double* vec) {
    if (!c || !vec || !vx || !vy || !vz) return NS3D_ERR_ARG;
    PtrArr<const zd, 3> v{{(const zd*)vx, (const zd*)vy, (const zd*)vz}};
    hipLaunchKernelGGL(k_pack_full, dim3(nblk(c->nc, 256)), dim3(256), 0, c->stream, v, vec, c->nc);
    HIPCHK(hipGetLastError());
    return NS3D_OK;
}

int ns3d_set_base_state(ns3d_ctx* c, ns3d_cplx* ux, ns3d_cplx* uy, ns3d_cplx* uz, const ns3d_cplx* fx,
                        const ns3d_cplx* fy, const ns3d_cplx* fz, int timesteps_period, double tau) {
    if (!c || !ux || !uy || !uz || timesteps_period < 0) return NS3D_ERR_ARG;
    c->base[0] = (zd*)ux;
    c->base[1] = (zd*)uy;
    c->base[2] = (zd*)uz;
    c->bforce[0] = (const zd*)fx;
    c->bforce[1] = (const zd*)fy;
    c->bforce[2] = (const zd*)fz;
    c->timesteps_period = timesteps_period;
    c->tau = tau;
    return ensure_fp64(c, &c->kry, 3);
}

}  // extern "C"

namespace {
int jv_reduced_impl(ns3d_ctx* c, const double* in, double* out, bool newton) {
    if (!c || !in || !out || !c->base[0] || !c->kry) return NS3D_ERR_ARG;
    ns3d_cplx* v[3] = {(ns3d_cplx*)c->kry, (ns3d_cplx*)(c->kry + c->nc), (ns3d_cplx*)(c->kry + 2 * c->nc)};
    ns3d_cplx* u[3] = {(ns3d_cplx*)c->base[0], (ns3d_cplx*)c->base[1], (ns3d_cplx*)c->base[2]};
    int st;
    // NSCallMatrixVector_reduced (Source/NS3D_periodic.cu:340-361)
    if ((st = ns3d_unpack_reduced(c, in, v[0], v[1], v[2]))) return st;
    if ((st = ns3d_rk3_ssp_uv_rhs(c, u[0], u[1], u[2], v[0], v[1], v[2]))) return st;
    if (!newton) return ns3d_pack_reduced(c, v[0], v[1], v[2], out);
    // ..._Newton_Globalization (:364-389): out = pack(J v) - v/tau
    PtrArr<const zd, 3> vv{{(const zd*)v[0], (const zd*)v[1], (const zd*)v[2]}};
    hipLaunchKernelGGL(k_pack_reduced_shifted, dim3(nblk(c->nc, 256)), dim3(256), 0, c->stream, vv, in, 1.0 / c->tau,
                       out, c->nc);
    HIPCHK(hipGetLastError());
    return NS3D_OK;
}
}  // namespace

extern "C" {

void ns3d_jv_reduced(void* ctx, void* vec_in, void* vec_out) {
    ns3d_ctx* c = (ns3d_ctx*)ctx;
    set_err(c, jv_reduced_impl(c, (const double*)vec_in, (double*)vec_out, false));
}

void ns3d_jv_reduced_newton(void* ctx, void* vec_in, void* vec_out) {
    ns3d_ctx* c = (ns3d_ctx*)ctx;
    set_err(c, jv_reduced_impl(c, (const double*)vec_in, (double*)vec_out, true));
}

void ns3d_jv_exponential(void* ctx, void* vec_in, void* vec_out) {
    ns3d_ctx* c = (ns3d_ctx*)ctx;
    if (!c || !vec_in || !vec_out || !c->base[0] || !c->kry) {
        set_err(c, NS3D_ERR_ARG);
        return;
    }
    ns3d_cplx* v[3] = {(ns3d_cplx*)c->kry, (ns3d_cplx*)(c->kry + c->nc), (ns3d_cplx*)(c->kry + 2 * c->nc)};
    ns3d_cplx* u[3] = {(ns3d_cplx*)c->base[0], (ns3d_cplx*)c->base[1], (ns3d_cplx*)c->base[2]};
    const ns3d_cplx* f[3] = {(const ns3d_cplx*)c->bforce[0], (const ns3d_cplx*)c->bforce[1], (const ns3d_cplx*)c->bforce[2]};
    // NSCallMatrixVector_exponential (Source/NS3D_periodic.cu:444-470)
    int st = ns3d_unpack_full(c, (const double*)vec_in, v[0], v[1], v[2]);
    for (int tt = 0; !st && tt < c->timesteps_period; tt++)
        st = ns3d_rk3_ssp_uv(c, u[0], u[1], u[2], v[0], v[1], v[2], f[0], f[1], f[2]);
    if (!st) st = ns3d_pack_full(c, v[0], v[1], v[2], (double*)vec_out);
    set_err(c, st);
}

// ---- pass timing
int ns3d_profile_enable(ns3d_ctx* c, int on) {
    if (!c) return NS3D_ERR_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    c->prof_on = on != 0;
    c->ev_used.clear();
    return NS3D_OK;
}

int ns3d_profile_read(ns3d_ctx* c, double* ms, long long* calls) {
    if (!c || !ms || !calls) return NS3D_ERR_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < NPASS; i++) {
        ms[i] = 0.0;
        calls[i] = 0;
    }
    for (auto& u : c->ev_used) {
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, c->ev_pool[u.second], c->ev_pool[u.second + 1]));
        ms[u.first] += t;
        calls[u.first] += 1;
    }
    c->ev_used.clear();
    return NS3D_OK;
}

// ---- a11 / a12
int ns3d_reduce_minmax(ns3d_ctx* c, long long n, const double* in, double* out2) {
    if (!c || !in || !out2 || n < 0) return NS3D_ERR_ARG;
    RedArgs r = red_args(c, in, nullptr, nullptr);
    return reduce<1, SRC_ARRAY>(c, r, (long)n, out2);
}
int ns3d_reduce_sum(ns3d_ctx* c, long long n, const double* in, double* out) {
    if (!c || !in || !out || n < 0) return NS3D_ERR_ARG;
    RedArgs r = red_args(c, in, nullptr, nullptr);
    return reduce<0, SRC_ARRAY>(c, r, (long)n, out);
}
int ns3d_check_finite(ns3d_ctx* c, long long n, const double* x) {
    if (!c || !x || n < 0) return NS3D_ERR_ARG;
    RedArgs r = red_args(c, x, nullptr, nullptr);
    double bad = 0.0;
    const int st = reduce<0, SRC_NONFINITE>(c, r, (long)n, &bad);
    if (st) return st;
    return bad > 0.0 ? NS3D_ERR_NONFINITE : NS3D_OK;
}
int ns3d_finite_status(ns3d_ctx* c, int clear) {
    if (!c) return NS3D_ERR_ARG;
    return finite_status(c, clear != 0);
}
int ns3d_finite_flag_async(ns3d_ctx* c, int* host_flag) {
    if (!c || !host_flag) return NS3D_ERR_ARG;
    HIPCHK(hipMemcpyAsync(host_flag, c->nf, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    return NS3D_OK;
}
int ns3d_kinetic_energy(ns3d_ctx* c, const double* ux, const double* uy, const double* uz, double* out) {
    if (!c || !ux || !uy || !uz || !out) return NS3D_ERR_ARG;
    RedArgs r = red_args(c, ux, uy, uz);
    return reduce<0, SRC_ENERGY>(c, r, c->nr, out);
}
int ns3d_dissipation(ns3d_ctx* c, const double* ux, const double* uy, const double* uz, double* out) {
    if (!c || !ux || !uy || !uz || !out) return NS3D_ERR_ARG;
    RedArgs r = red_args(c, ux, uy, uz);
    return reduce<0, SRC_DISSIPATION>(c, r, c->nr, out);
}
// Parseval with the unnormalised forward transform: integral over the box of
// |u|^2 = (Lx Ly Lz / (Nx Ny Nz)^2) * sum over all modes of |u^|^2.
int ns3d_kinetic_energy_spectral(ns3d_ctx* c, const ns3d_cplx* ux, const ns3d_cplx* uy, const ns3d_cplx* uz,
                                 double* out) {
    if (!c || !ux || !uy || !uz || !out) return NS3D_ERR_ARG;
    RedArgs r = red_args(c, (const double*)ux, (const double*)uy, (const double*)uz);
    const int st = reduce<0, SRC_SPEC_ENERGY>(c, r, c->nc, out);
    const double n3 = (double)c->Nx * c->Ny * c->Nz;
    if (!st) out[0] *= 0.5 * c->Lx * c->Ly * c->Lz / (n3 * n3);
    return st;
}
int ns3d_dissipation_spectral(ns3d_ctx* c, const ns3d_cplx* ux, const ns3d_cplx* uy, const ns3d_cplx* uz,
                              double* out) {
    if (!c || !ux || !uy || !uz || !out) return NS3D_ERR_ARG;
    RedArgs r = red_args(c, (const double*)ux, (const double*)uy, (const double*)uz);
    const int st = reduce<0, SRC_SPEC_DISSIPATION>(c, r, c->nc, out);
    const double n3 = (double)c->Nx * c->Ny * c->Nz;
    if (!st) out[0] *= c->Lx * c->Ly * c->Lz / (n3 * n3);
    return st;
}
int ns3d_cfl_dt(ns3d_ctx* c, double cfl, const double* ux, const double* uy, const double* uz, double* dt) {
    if (!c || !ux || !uy || !uz || !dt) return NS3D_ERR_ARG;
    RedArgs r = red_args(c, ux, uy, uz);
    r.p0 = cfl;
    double mm[2];
    int st = reduce<1, SRC_CFL>(c, r, c->nr, mm);
    if (!st) *dt = mm[0];
    return st;
}
int ns3d_shapiro_error(ns3d_ctx* c, double t, const double* ux, const double* uy, const double* uz, double* err) {
    if (!c || !ux || !uy || !uz || !err) return NS3D_ERR_ARG;
    RedArgs r = red_args(c, ux, uy, uz);
    r.p0 = t;
    double mm[2];
    int st = reduce<1, SRC_SHAPIRO>(c, r, c->nr, mm);
    if (!st) *err = mm[1];
    return st;
}

// ---- f1: FP64 vector kernels of the Krylov drivers (Source/Arnoldi/cuda_supp.cu)
int ns3d_vec_gemv_t(ns3d_ctx* c, long long n, int m, const double* V, long long ldv, const double* w, double* h) {
    if (!c || !V || !w || !h || n < 1 || m < 1 || m > KV_MAXM || ldv < n) return NS3D_ERR_ARG;
    int st = ensure_kv(c);
    if (st) return st;
    const int nb = (int)std::min<long long>(KV_BLOCKS, (n + KV_THREADS - 1) / KV_THREADS);
    for (int c0 = 0; c0 < m; c0 += KV_COLS) {
        const int nc = std::min(KV_COLS, m - c0);
        hipLaunchKernelGGL(k_kv_gemv_t_partial, dim3(nb), dim3(KV_THREADS), 0, c->stream, (long)n, nc, V, (long)ldv, c0, w,
                           c->kv_partial);
        hipLaunchKernelGGL(k_kv_final, dim3(nc), dim3(KV_THREADS), 0, c->stream, c->kv_partial, nb, c->kv_out + c0);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->kv_host, c->kv_out, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int j = 0; j < m; j++) h[j] = c->kv_host[j];
    return NS3D_OK;
}

int ns3d_vec_dot(ns3d_ctx* c, long long n, const double* x, const double* y, double* out) {
    return ns3d_vec_gemv_t(c, n, 1, x, n, y, out);
}

int ns3d_vec_gemv(ns3d_ctx* c, long long n, int m, const double* V, long long ldv, const double* h, double alpha,
                  double beta, double* y) {
    if (!c || !V || !h || !y || n < 1 || m < 1 || m > KV_MAXM || ldv < n) return NS3D_ERR_ARG;
    KvCoef hc;
    for (int j = 0; j < KV_MAXM; j++) hc.h[j] = j < m ? h[j] : 0.0;
    hipLaunchKernelGGL(k_kv_gemv, dim3(nblk(n, KV_THREADS)), dim3(KV_THREADS), 0, c->stream, (long)n, m, V, (long)ldv, hc,
                       alpha, beta, y);
    HIPCHK(hipGetLastError());
    return NS3D_OK;
}

int ns3d_vec_gemm(ns3d_ctx* c, long long n, int m, int k, const double* V, long long ldv, const double* Q, double* W,
                  long long ldw) {
    if (!c || !V || !Q || !W || n < 1 || m < 1 || k < 1 || m > KV_MAXM || k > KV_MAXM || ldv < n || ldw < n)
        return NS3D_ERR_ARG;
    if (W < V + (long long)(m - 1) * ldv + n && V < W + (long long)(k - 1) * ldw + n) return NS3D_ERR_ARG;  // overlap
    int st = ensure_kv(c);
    if (st) return st;
    for (long j = 0; j < (long)m * k; j++) c->kq_host[j] = Q[j];
    HIPCHK(hipMemcpyAsync(c->kq, c->kq_host, sizeof(double) * m * k, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_kv_gemm, dim3(nblk(n, KV_THREADS), k), dim3(KV_THREADS), 0, c->stream, (long)n, m, k, V, (long)ldv,
                       (const double*)c->kq, W, (long)ldw);
    HIPCHK(hipGetLastError());
    // kq_host is reused by the next call: wait for the copy (and the product)
    HIPCHK(hipStreamSynchronize(c->stream));
    return NS3D_OK;
}

int ns3d_vec_axpby(ns3d_ctx* c, long long n, double a, const double* x, double b, double* y) {
    if (!c || !x || !y || n < 1) return NS3D_ERR_ARG;
    hipLaunchKernelGGL(k_kv_axpby, dim3(nblk(n, KV_THREADS)), dim3(KV_THREADS), 0, c->stream, (long)n, a, x, b, y);
    HIPCHK(hipGetLastError());
    return NS3D_OK;
}

// ---- f4: diagnostics of the reference driver (Source/math_support.cu:712-987)
int ns3d_curl(ns3d_ctx* c, const double* ux, const double* uy, const double* uz, double* rx, double* ry, double* rz,
              int fix) {
    if (!c || !ux || !uy || !uz || !rx || !ry || !rz) return NS3D_ERR_ARG;
    Grid3 G{c->Nx, c->Ny, c->Nz};
    hipLaunchKernelGGL(k_curl, dim3(nblk(c->nr, 256)), dim3(256), 0, c->stream, ux, uy, uz, rx, ry, rz, G, c->Lx / c->Nx,
                       c->Ly / c->Ny, c->Lz / c->Nz, fix ? 1 : 0);
    HIPCHK(hipGetLastError());
    return NS3D_OK;
}

static int spectral_filter(ns3d_ctx* c, int kind, double p, const ns3d_cplx* ix, const ns3d_cplx* iy, const ns3d_cplx* iz,
                           ns3d_cplx* ox, ns3d_cplx* oy, ns3d_cplx* oz) {
    if (!c || !ix || !iy || !iz || !ox || !oy || !oz) return NS3D_ERR_ARG;
    PtrArr<const zd, 3> in{{(const zd*)ix, (const zd*)iy, (const zd*)iz}};
    PtrArr<zd, 3> out{{(zd*)ox, (zd*)oy, (zd*)oz}};
    if (kind == 0)
        hipLaunchKernelGGL(k_spectral_filter<0>, dim3(nblk(c->nc, 256)), dim3(256), 0, c->stream, in, out, c->nc, c->Nx,
                           c->Ny, c->Mz, c->Lx, c->Ly, c->Lz, p);
    else
        hipLaunchKernelGGL(k_spectral_filter<1>, dim3(nblk(c->nc, 256)), dim3(256), 0, c->stream, in, out, c->nc, c->Nx,
                           c->Ny, c->Mz, c->Lx, c->Ly, c->Lz, p);
    HIPCHK(hipGetLastError());
    return NS3D_OK;
}

int ns3d_helmholtz_filter(ns3d_ctx* c, double eps, const ns3d_cplx* ux, const ns3d_cplx* uy, const ns3d_cplx* uz,
                          ns3d_cplx* fx, ns3d_cplx* fy, ns3d_cplx* fz) {
    return spectral_filter(c, 0, eps, ux, uy, uz, fx, fy, fz);
}

int ns3d_cutoff_filter(ns3d_ctx* c, double radius, const ns3d_cplx* ux, const ns3d_cplx* uy, const ns3d_cplx* uz,
                       ns3d_cplx* fx, ns3d_cplx* fy, ns3d_cplx* fz) {
    return spectral_filter(c, 1, radius, ux, uy, uz, fx, fy, fz);
}

// ---- f3: the return map's probe (return_vector3_solution, Source/return_map.cu:110-160)
int ns3d_probe(ns3d_ctx* c, const ns3d_cplx* ux, const ns3d_cplx* uy, const ns3d_cplx* uz, int j, int k, int l,
               double* out3) {
    if (!c || !ux || !uy || !uz || !out3 || j < 0 || k < 0 || l < 0 || j >= c->Nx || k >= c->Ny || l >= c->Nz)
        return NS3D_ERR_ARG;
    int st = ensure_kv(c);
    if (st) return st;
    PtrArr<const zd, 3> u{{(const zd*)ux, (const zd*)uy, (const zd*)uz}};
    const int nb = (int)std::min<long>(KV_BLOCKS, (c->nc + KV_THREADS - 1) / KV_THREADS);
    hipLaunchKernelGGL(k_probe_partial, dim3(nb), dim3(KV_THREADS), 0, c->stream, u, c->nc, c->Nx, c->Ny, c->Nz, c->Mz, j,
                       k, l, c->kv_partial);
    hipLaunchKernelGGL(k_kv_final, dim3(3), dim3(KV_THREADS), 0, c->stream, c->kv_partial, nb, c->kv_out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->kv_host, c->kv_out, sizeof(double) * 3, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const double inv = 1.0 / (1.0 * c->Nx * c->Ny * c->Nz);
    for (int b = 0; b < 3; b++) out3[b] = c->kv_host[b] * inv;
    return NS3D_OK;
}

int ns3d_probe_batch(ns3d_ctx* c, int nbatch, const ns3d_cplx* u, int j, int k, int l, double* out_host) {
    if (!c || !u || !out_host || nbatch < 1 || nbatch > NS3D_MAX_BATCH || j < 0 || k < 0 || l < 0 || j >= c->Nx ||
        k >= c->Ny || l >= c->Nz)
        return NS3D_ERR_ARG;
    const int nb = (int)std::min<long>(64, (c->nc + KV_THREADS - 1) / KV_THREADS);
    const size_t need = (size_t)3 * nbatch * (nb + 1);
    if (c->pb_cap < need) {
        if (c->pb) {
            HIPCHK(hipStreamSynchronize(c->stream));
            hipFree(c->pb);
            hipHostFree(c->pb_host);
            c->pb = nullptr;
            c->pb_host = nullptr;
            c->pb_cap = 0;
        }
        if (hipMalloc((void**)&c->pb, sizeof(double) * need) != hipSuccess ||
            hipHostMalloc((void**)&c->pb_host, sizeof(double) * 3 * nbatch) != hipSuccess)
            return NS3D_ERR_ALLOC;
        c->pb_cap = need;
    }
    const zd* ub = (const zd*)u;
    PtrArr<const zd, 3> p{{ub, ub + c->nc, ub + 2 * c->nc}};
    double* partial = c->pb;
    double* out = c->pb + (size_t)3 * nbatch * nb;
    hipLaunchKernelGGL(k_probe_partial, dim3(nb, nbatch), dim3(KV_THREADS), 0, c->stream, p, c->nc, c->Nx, c->Ny, c->Nz,
                       c->Mz, j, k, l, partial, 3 * c->nc);
    hipLaunchKernelGGL(k_kv_final, dim3(3 * nbatch), dim3(KV_THREADS), 0, c->stream, partial, nb, out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->pb_host, out, sizeof(double) * 3 * nbatch, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const double inv = 1.0 / (1.0 * c->Nx * c->Ny * c->Nz);
    for (int i = 0; i < 3 * nbatch; i++) out_host[i] = c->pb_host[i] * inv;
    return NS3D_OK;
}

}  // extern "C"
