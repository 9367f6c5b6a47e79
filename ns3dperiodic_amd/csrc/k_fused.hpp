// k_fused.hpp -- the fused pseudo-spectral stage pipeline.
//
// One evaluation of the 2/3-masked nonlinear term (reference
// calculate_convolution_2p3, Source/advection_2_3.cu:480-497) plus the
// projection/diffusion/update epilogue (Source/diffusion.cu:96-365) runs as
// five passes over HBM:
//   P1 x-inverse : w^ (3, FP64 state) -> mask, i kx  -> A (6 fields)
//   P2 y-inverse : A (6)              -> i ky       -> B (9 fields)
//   P3 z         : B (9) -> C2R (+ i kz), (w.grad)w, R2C -> C (3 fields)
//   P4 y-forward : C (3) -> D (3)
//   P5 x-forward : D (3) -> IPQ -> epilogue (update / RHS / Q out)
// Real-space fields never touch HBM.  NF=2 evaluates B(u,v) of the J.v path
// (calculate_convolution_2p3_UV, :500-525) with the same passes.
#pragma once
#include <type_traits>

#include "fft_core.hpp"

namespace ns3d {

template <typename T, int N> struct PtrArr { T* p[N]; };

// Minimum waves per SIMD requested from the register allocator (the passes are
// latency bound: more resident waves hide HBM latency better than ILP).
#ifndef NS3D_WAVES_XY
#define NS3D_WAVES_XY 2
#endif
// Two transforms in step (PencilFFT::run2, a second exchange buffer) for the
// first two components of the x-forward pass: 0 off, 1 FP64 only, 2 both
// precisions (measured at 256^3: FP64 pass 0.407 -> 0.401 ms, its register
// spill gone; the x- and y-inverse passes, at 4-7 waves/SIMD, gained nothing).
#ifndef NS3D_XFWD_ILP
#define NS3D_XFWD_ILP 2
#endif
#ifndef NS3D_WAVES_Z
#define NS3D_WAVES_Z 2
#endif
// Tile order of each pass (1: last tile first).  The 256 MB Infinity Cache
// behind L2 keeps part of what a pass has just written; a pass that reads it
// back in the reverse of the order it was written finds the newest data first.
// The y-inverse -> z -> y-forward passes walk the x-planes in the same
// direction when all run forward, and the x-forward -> x-inverse passes the
// (k, l) columns; reversing the z and x-inverse passes alternates both chains
// (256^3 step: FP64 -0.3%, FP32 modes -0.4 to -0.5%; results bitwise unchanged).
#ifndef NS3D_REV_XINV
#define NS3D_REV_XINV 1
#endif
#ifndef NS3D_REV_YINV
#define NS3D_REV_YINV 0
#endif
#ifndef NS3D_REV_Z
#define NS3D_REV_Z 1
#endif
#ifndef NS3D_REV_YFWD
#define NS3D_REV_YFWD 0
#endif
#ifndef NS3D_REV_XFWD
#define NS3D_REV_XFWD 0
#endif
// The x- and y-inverse passes stage their twiddle table after their first
// global loads are issued (1) or before them (0): the table's own round trip
// and block barrier then overlap the HBM latency of the pencil loads.  The
// other passes stage it first: issuing their loads first made the 512^3 z pass
// 3-4% (FP64) and the FP32 y-forward pass 5% slower, with no gain at 256^3.
#ifndef NS3D_TW_LATE
#define NS3D_TW_LATE 1
#endif
template <int REV>
__device__ __forceinline__ int tile_order(int bid, int nb) {
    return REV ? nb - 1 - bid : bid;
}
// Minimum waves per SIMD of the FP32 z pass and of the FP32 x-forward
// epilogue pass (their register budgets: 2 -> 256 VGPRs, 3 -> 168, 4 -> 128).
#ifndef NS3D_WAVES_Z32
#define NS3D_WAVES_Z32 2
#endif
#ifndef NS3D_WAVES_XFWD32
#define NS3D_WAVES_XFWD32 2
#endif
// x/y-pass plan: E = 8 points per thread (Plan) or 4 (ZPlan); chosen per kernel
// and precision by the host from measurements.
template <int N, int E> struct XYSel { using type = Plan<N>; };
template <int N> struct XYSel<N, 4> { using type = ZPlan<N>; };
template <int N, int E> using XYPlan = typename XYSel<N, E>::type;
// The inverse passes and the y-forward pass handle one field per block
// (blockIdx.y): shorter blocks, more of them in flight.
// Batches: blockIdx.z is the member of a batch of same-size fields advanced
// by one launch; each kernel shifts its per-member pointers by z times the
// member stride it is given (bsc: scratch, bwi: input state; 0 when z is 1).
#define NS3D_FIELD_LOOP(f, n) for (int f = blockIdx.y; f < (int)blockIdx.y + 1; f++)

// Members of a batch that sit out a step (resident requests of one grid that
// need fewer steps than the launch carries): member z takes part in step
// `step` iff step < left[z]; left null: every member does.  The test is
// uniform over a block, so an idle member's blocks return before any barrier.
struct MemberMask {
    const int* left;
    int step;
};
__device__ __forceinline__ bool member_idle(const MemberMask& m) { return m.left && m.step >= m.left[blockIdx.z]; }

// ---------------------------------------------------------------- P1
// x-inverse of M w^_b and of M (i kx w^_b).  Block: one y index k and P
// consecutive z indices l; lanes p-fastest so each row load is contiguous.
// grid.y enumerates the fields; with OS = 1 each block produces one output
// (w_b or dx w_b) and the input is read by two blocks.
// TS: scalar type of the state w (double, or float for the FP32-state RK3 loop).
// ADD = 1: the transformed field is w + w2 (w2 of the same type and layout; the
// FP32-increment RK3 loop's stage value base + increment), summed in FP64.
template <typename TB, int NX, int P, int NF, int E, int OS, typename TS = double, int ADD = 0>
__global__ void __attribute__((amdgpu_waves_per_eu(NS3D_WAVES_XY, 8))) __launch_bounds__((P * XYPlan<NX, E>::TPP))
k_x_inv_grad(PtrArr<const cplx<TS>, 3 * NF> w, cplx<TB>* __restrict__ A, long nc, int Ny, int Mz, int Mw, int Ms,
             int nlt, int nb, SpecTables tab, const cplx<TB>* __restrict__ twg, long bwi = 0, long bsc = 0,
             PtrArr<const cplx<TS>, 3> w2 = {}, MemberMask mm = {}) {
    constexpr int TPP = XYPlan<NX, E>::TPP;
    using F = PencilFFT<TB, NX, TPP, +1, typename XYPlan<NX, E>::R, IndexMajor<P>>;
    __shared__ cplx<TB> lds[NX * P];
    __shared__ cplx<TB> tw[NX];
    const int bid0 = xcd_tile();
    if (bid0 >= nb || member_idle(mm)) return;
    const int bid = tile_order<NS3D_REV_XINV>(bid0, nb);
    const int tid = threadIdx.x, p = tid % P, t = tid / P;
    const int k = bid / nlt, lt = bid % nlt;
    const int l = lt * P + p;
    const bool lv = l < Mz;
    // a column (k,l) whose every x mode is masked stays unwritten: the y pass skips it
    const bool keep = lv && column_kept(tab, k, l);
    if (!__syncthreads_or(keep)) return;
    if (!NS3D_TW_LATE) stage_twiddles<TB, NX>(tw, twg);
    const long S = (long)Ny * Mw;        // state: row stride Mw (Mz: reference layout)
    const long base = (long)k * Mw + l;
    const long So = (long)Ny * Ms;       // scratch: rows padded to Ms
    const long baseo = (long)k * Ms + l;
    const int f = OS ? (int)blockIdx.y >> 1 : (int)blockIdx.y;
    const int which = OS ? (int)blockIdx.y & 1 : 0;  // OS: 0 -> w_b, 1 -> dx w_b
    const cplx<TS>* __restrict__ src = w.p[f] + (long)blockIdx.z * bwi;
    const cplx<TS>* __restrict__ src2 = ADD ? w2.p[f % 3] : nullptr;
    A += (long)blockIdx.z * bsc;
    cplx<TB> v[F::E], d[OS ? 1 : F::E];
    // the loads first (raw), then the twiddle table, then the arithmetic
    cplx<TS> raw[F::E], raw2[ADD != 0 ? F::E : 1];
#pragma unroll
    for (int q = 0; q < F::B0; q++) {
#pragma unroll
        for (int r = 0; r < F::R0; r++) {
            const int j = F::in_index(t, q, r);
            const int i = q * F::R0 + r;
            raw[i] = mk<TS>(0, 0);
            if constexpr (ADD != 0) raw2[i] = mk<TS>(0, 0);
            if (F::in_active(t, q) && keep && mask_keep(tab, j, k, l)) {
                raw[i] = src[base + j * S];
                if constexpr (ADD != 0) raw2[i] = src2[base + j * S];
            }
        }
    }
    if (NS3D_TW_LATE) stage_twiddles<TB, NX>(tw, twg);
#pragma unroll
    for (int q = 0; q < F::B0; q++) {
#pragma unroll
        for (int r = 0; r < F::R0; r++) {
            const int j = F::in_index(t, q, r);
            const int i = q * F::R0 + r;
            cplx<double> s = cvt<double>(raw[i]);
            if constexpr (ADD != 0) s = s + cvt<double>(raw2[i]);
            const double kx = F::in_active(t, q) ? tab.kx[j] : 0.0;
            const cplx<TB> dv = mk<TB>((TB)(-s.y * kx), (TB)(s.x * kx));
            if constexpr (OS) {
                v[q * F::R0 + r] = which ? dv : cvt<TB>(s);
            } else {
                v[q * F::R0 + r] = cvt<TB>(s);
                d[q * F::R0 + r] = dv;
            }
        }
    }
    const int s_ = f / 3, b = f % 3;
    cplx<TB>* __restrict__ ow = A + (long)(s_ * 6 + (OS ? 3 * which : 0) + b) * nc;
    cplx<TB>* __restrict__ od = A + (long)(s_ * 6 + 3 + b) * nc;
    F::run(v, lds, tw, t, p);
    if constexpr (!OS) F::run(d, lds, tw, t, p);
    if (keep) {
#pragma unroll
        for (int q = 0; q < F::BL; q++) {
            if (!F::out_active(t, q)) continue;
#pragma unroll
            for (int r = 0; r < F::RL; r++) {
                const long g = baseo + F::out_index(t, q, r) * So;
                ow[g] = v[q * F::RL + r];
                if constexpr (!OS) od[g] = d[q * F::RL + r];
            }
        }
    }
}

// ---------------------------------------------------------------- P2
// y-inverse: B(w_b) , B(dx w_b) , B(i ky w_b).  Block: one x index j and P z
// indices.  grid.y enumerates components; with OS = 1 each block produces one
// of the three outputs (A_b is then read by two blocks).
template <typename TB, int NY, int P, int NF, int E, int OS>
__global__ void __attribute__((amdgpu_waves_per_eu(NS3D_WAVES_XY, 8))) __launch_bounds__((P * XYPlan<NY, E>::TPP))
k_y_inv_grad(const cplx<TB>* __restrict__ A, cplx<TB>* __restrict__ B, long nc, int Ny, int Mz, int Ms, int nlt,
             int nb, SpecTables tab, const cplx<TB>* __restrict__ twg, long bsc = 0, MemberMask mm = {}) {
    constexpr int TPP = XYPlan<NY, E>::TPP;
    using F = PencilFFT<TB, NY, TPP, +1, typename XYPlan<NY, E>::R, IndexMajor<P>>;
    __shared__ cplx<TB> lds[NY * P];
    __shared__ cplx<TB> tw[NY];
    const int bid0 = xcd_tile();
    if (bid0 >= nb || member_idle(mm)) return;
    const int bid = tile_order<NS3D_REV_YINV>(bid0, nb);
    const int tid = threadIdx.x, p = tid % P, t = tid / P;
    const int j = bid / nlt, lt = bid % nlt;
    if (lt * P > tab.lmax) return;  // z-planes beyond lmax are zero; the z pass skips them
    A += (long)blockIdx.z * bsc;
    B += (long)blockIdx.z * bsc;
    const int l = lt * P + p;
    const bool lv = l < Mz && l <= tab.lmax;
    if (!NS3D_TW_LATE) stage_twiddles<TB, NY>(tw, twg);
    const long S = Ms;
    const long base = (long)j * Ny * Ms + l;
    const int f = OS ? (int)blockIdx.y / 3 : (int)blockIdx.y;
    const int which = OS ? (int)blockIdx.y % 3 : -1;  // OS: 0 -> w_b, 1 -> dy w_b, 2 -> dx w_b
    const int s_ = f / 3, b = f % 3;
#ifndef NS3D_HOIST
#define NS3D_HOIST 1
#endif
    // x-derivative input, loaded before the first transforms when both are done here
    cplx<TB> v2[(!OS && NS3D_HOIST) ? F::E : 1];
    if constexpr (!OS && NS3D_HOIST) {
        const cplx<TB>* __restrict__ src = A + (long)(s_ * 6 + 3 + b) * nc;
#pragma unroll
        for (int q = 0; q < F::B0; q++) {
#pragma unroll
            for (int r = 0; r < F::R0; r++) {
                const int kk = F::in_index(t, q, r);
                cplx<TB> s = mk<TB>(0, 0);
                if (F::in_active(t, q) && lv && column_kept(tab, kk, l)) s = src[base + kk * S];
                v2[q * F::R0 + r] = s;
            }
        }
    }
    // velocity component: produce w_b and dy w_b
    if (!OS || which < 2) {
        const cplx<TB>* __restrict__ src = A + (long)(s_ * 6 + b) * nc;
        cplx<TB> v[F::E], d[OS ? 1 : F::E];
#pragma unroll
        for (int q = 0; q < F::B0; q++) {
#pragma unroll
            for (int r = 0; r < F::R0; r++) {
                const int kk = F::in_index(t, q, r);
                cplx<TB> s = mk<TB>(0, 0);
                if (F::in_active(t, q) && lv && column_kept(tab, kk, l)) s = src[base + kk * S];
                v[q * F::R0 + r] = s;
            }
        }
        if (NS3D_TW_LATE) stage_twiddles<TB, NY>(tw, twg);
#pragma unroll
        for (int q = 0; q < F::B0; q++) {
#pragma unroll
            for (int r = 0; r < F::R0; r++) {
                const int kk = F::in_index(t, q, r);
                const cplx<TB> s = v[q * F::R0 + r];
                const TB ky = F::in_active(t, q) ? (TB)tab.ky[kk] : (TB)0;
                const cplx<TB> dv = mk<TB>(-s.y * ky, s.x * ky);
                if constexpr (OS) {
                    v[q * F::R0 + r] = which ? dv : s;
                } else {
                    v[q * F::R0 + r] = s;
                    d[q * F::R0 + r] = dv;
                }
            }
        }
        F::run(v, lds, tw, t, p);
        if constexpr (!OS) F::run(d, lds, tw, t, p);
        cplx<TB>* __restrict__ ow = B + (long)(s_ * 9 + (OS && which == 1 ? 6 : 0) + b) * nc;
        cplx<TB>* __restrict__ od = B + (long)(s_ * 9 + 6 + b) * nc;
        if (lv) {
#pragma unroll
            for (int q = 0; q < F::BL; q++) {
                if (!F::out_active(t, q)) continue;
#pragma unroll
                for (int r = 0; r < F::RL; r++) {
                    const long g = base + F::out_index(t, q, r) * S;
                    ow[g] = v[q * F::RL + r];
                    if constexpr (!OS) od[g] = d[q * F::RL + r];
                }
            }
        }
    }
    // x-derivative component: plain y-inverse
    if (!OS || which == 2) {
        cplx<TB> v[F::E];
        if constexpr (!OS && NS3D_HOIST) {
#pragma unroll
            for (int e = 0; e < F::E; e++) v[e] = v2[e];
        } else {
            const cplx<TB>* __restrict__ src = A + (long)(s_ * 6 + 3 + b) * nc;
#pragma unroll
            for (int q = 0; q < F::B0; q++) {
#pragma unroll
                for (int r = 0; r < F::R0; r++) {
                    const int kk = F::in_index(t, q, r);
                    cplx<TB> s = mk<TB>(0, 0);
                    if (F::in_active(t, q) && lv && column_kept(tab, kk, l)) s = src[base + kk * S];
                    v[q * F::R0 + r] = s;
                }
            }
        }
        // with OS these blocks skip the velocity branch, where the table is staged otherwise
        if (NS3D_TW_LATE && OS) stage_twiddles<TB, NY>(tw, twg);
        F::run(v, lds, tw, t, p);
        cplx<TB>* __restrict__ ow = B + (long)(s_ * 9 + 3 + b) * nc;
        if (lv) {
#pragma unroll
            for (int q = 0; q < F::BL; q++) {
                if (!F::out_active(t, q)) continue;
#pragma unroll
                for (int r = 0; r < F::RL; r++) ow[base + F::out_index(t, q, r) * S] = v[q * F::RL + r];
            }
        }
    }
}

// ---------------------------------------------------------------- P3
// z pass on one (x,y) pencil: paired C2R transforms (two real fields per
// complex FFT), the advective product, and paired R2C transforms.
// C2R convention: x,y passes already done; only Re of l=0 and l=Nz/2 is used.
template <typename TB, int NZ, int P, bool TRUNC = false>
struct ZPencil {
    static constexpr int TPP = ZPlanT<TB, NZ>::TPP;
    static constexpr int PAD = (sizeof(TB) == 8) ? 1 : 2;
#ifndef NS3D_ZSWZ
#define NS3D_ZSWZ 8
#endif
#ifndef NS3D_ZSWZ32
#define NS3D_ZSWZ32 16
#endif
    // LDS swizzle group in elements (0: none); measured gains: FP64 with the 4-point
    // plan, FP32 with the 8-point plan
#ifndef NS3D_R3_ZSWZ
#define NS3D_R3_ZSWZ 0
#endif
    static constexpr int SWG = sizeof(TB) == 8 ? (NZ % 3 == 0 && NZ >= 192 ? NS3D_R3_ZSWZ : NS3D_ZSWZ) : NS3D_ZSWZ32;
    using LA = typename std::conditional<SWG == 0, PencilMajor<NZ, PAD>, PencilMajorSw<NZ, PAD, (SWG == 0 ? 1 : SWG)>>::type;
    // Each pencil's LDS region is private to its TPP threads; when TPP divides
    // the wavefront size, those threads lie in one wavefront and every exchange
    // is ordered within the wave (exch_sync<true>) instead of by a block barrier.
#ifndef NS3D_ZWAVE
#define NS3D_ZWAVE 1
#endif
    static constexpr bool WAVE = NS3D_ZWAVE != 0 && TPP <= 64 && 64 % TPP == 0;
    using FI = PencilFFT<TB, NZ, TPP, +1, typename ZPlanT<TB, NZ>::R, LA, WAVE>;
    using FF = PencilFFT<TB, NZ, TPP, -1, typename SeqRev<typename ZPlanT<TB, NZ>::R>::type, LA, WAVE>;
    static_assert(FI::RL == FF::R0 && FI::BL == FF::B0, "mirror plans must share the product layout");
    static constexpr int E = FI::E > FF::E ? FI::E : FF::E;
    static constexpr int NPT = FI::BL * FI::RL;  // physical points per thread

    // Load the Hermitian extension of (X1 + i X2); dk selects i*kz multiplication.
    __device__ __forceinline__ static void load_pair(cplx<TB>* v, const cplx<TB>* __restrict__ s1, bool d1,
                                                     const cplx<TB>* __restrict__ s2, bool d2, bool valid,
                                                     const SpecTables& tab, int t) {
#pragma unroll
        for (int q = 0; q < FI::B0; q++) {
#pragma unroll
            for (int r = 0; r < FI::R0; r++) {
                const int n = FI::in_index(t, q, r);
                const int sl = n <= NZ / 2 ? n : NZ - n;
                cplx<TB> a = mk<TB>(0, 0), b = mk<TB>(0, 0);
                TB kz = 0;
                if (FI::in_active(t, q)) {
                    kz = (TB)tab.kz[sl];
                    if (valid && (!TRUNC || sl <= tab.lmax)) {
                        a = s1[sl];
                        b = s2[sl];
                    }
                }
                if (d1) a = mk<TB>(-a.y * kz, a.x * kz);
                if (d2) b = mk<TB>(-b.y * kz, b.x * kz);
                if (n == 0 || n == NZ / 2) {
                    a.y = 0;
                    b.y = 0;
                } else if (n > NZ / 2) {
                    a.y = -a.y;
                    b.y = -b.y;
                }
                v[q * FI::R0 + r] = mk<TB>(a.x - b.y, a.y + b.x);
            }
        }
    }

    // Split form of load_pair for software pipelining.  fetch_pair issues the
    // global loads of a pair's first-pass inputs (zeros where inactive,
    // invalid or truncated); finish_pair forms the Hermitian-extended
    // (X1 + i X2) from them.  kz_regs holds kz of each input slot (the same
    // for every pair).
    static constexpr int NIN = FI::B0 * FI::R0;
    __device__ __forceinline__ static void kz_regs(TB* kz, const SpecTables& tab, int t) {
#pragma unroll
        for (int q = 0; q < FI::B0; q++) {
#pragma unroll
            for (int r = 0; r < FI::R0; r++) {
                const int n = FI::in_index(t, q, r);
                const int sl = n <= NZ / 2 ? n : NZ - n;
                kz[q * FI::R0 + r] = FI::in_active(t, q) ? (TB)tab.kz[sl] : (TB)0;
            }
        }
    }
    __device__ __forceinline__ static void fetch_pair(cplx<TB>* r1, cplx<TB>* r2, const cplx<TB>* __restrict__ s1,
                                                      const cplx<TB>* __restrict__ s2, bool valid,
                                                      const SpecTables& tab, int t) {
#pragma unroll
        for (int q = 0; q < FI::B0; q++) {
#pragma unroll
            for (int r = 0; r < FI::R0; r++) {
                const int n = FI::in_index(t, q, r);
                const int sl = n <= NZ / 2 ? n : NZ - n;
                cplx<TB> a = mk<TB>(0, 0), b = mk<TB>(0, 0);
                // masked and truncated elements stay zero without a load (a branch-free
                // variant that loaded a clamped in-bounds element and zeroed it instead
                // made the 256^3 z pass 3.4x slower in FP64, 1.6x in FP32)
                if (FI::in_active(t, q) && valid && (!TRUNC || sl <= tab.lmax)) {
                    a = s1[sl];
                    b = s2[sl];
                }
                r1[q * FI::R0 + r] = a;
                r2[q * FI::R0 + r] = b;
            }
        }
    }
    // fetch_pair through buffer resources b1, b2 over the block's valid pencils:
    // prow = this pencil's first element in them (p * row stride); masked and
    // truncated elements and padding pencils fall outside the range and read 0.
    __device__ __forceinline__ static void fetch_pair_buf(cplx<TB>* r1, cplx<TB>* r2, __amdgpu_buffer_rsrc_t b1,
                                                          __amdgpu_buffer_rsrc_t b2, int prow, const SpecTables& tab,
                                                          int t) {
#pragma unroll
        for (int q = 0; q < FI::B0; q++) {
#pragma unroll
            for (int r = 0; r < FI::R0; r++) {
                const int n = FI::in_index(t, q, r);
                const int sl = n <= NZ / 2 ? n : NZ - n;
                const bool on = FI::in_active(t, q) && (!TRUNC || sl <= tab.lmax);
                const int o = on ? (prow + sl) * (int)sizeof(cplx<TB>) : BUF_OOB;
                r1[q * FI::R0 + r] = buf_load<TB>(b1, o);
                r2[q * FI::R0 + r] = buf_load<TB>(b2, o);
            }
        }
    }
    __device__ __forceinline__ static void finish_pair(cplx<TB>* v, const cplx<TB>* r1, bool d1, const cplx<TB>* r2,
                                                       bool d2, const TB* kz, int t) {
#pragma unroll
        for (int q = 0; q < FI::B0; q++) {
#pragma unroll
            for (int r = 0; r < FI::R0; r++) {
                const int i = q * FI::R0 + r;
                const int n = FI::in_index(t, q, r);
                cplx<TB> a = r1[i], b = r2[i];
                if (d1) a = mk<TB>(-a.y * kz[i], a.x * kz[i]);
                if (d2) b = mk<TB>(-b.y * kz[i], b.x * kz[i]);
                if (n == 0 || n == NZ / 2) {
                    a.y = 0;
                    b.y = 0;
                } else if (n > NZ / 2) {
                    a.y = -a.y;
                    b.y = -b.y;
                }
                v[i] = mk<TB>(a.x - b.y, a.y + b.x);
            }
        }
    }

    // load_pair through buffer resources over the block's valid pencils (as
    // fetch_pair_buf): out-of-range elements read as zero.
    __device__ __forceinline__ static void load_pair_buf(cplx<TB>* v, __amdgpu_buffer_rsrc_t b1, bool d1,
                                                         __amdgpu_buffer_rsrc_t b2, bool d2, int prow,
                                                         const SpecTables& tab, int t) {
#pragma unroll
        for (int q = 0; q < FI::B0; q++) {
#pragma unroll
            for (int r = 0; r < FI::R0; r++) {
                const int n = FI::in_index(t, q, r);
                const int sl = n <= NZ / 2 ? n : NZ - n;
                const bool on = FI::in_active(t, q) && (!TRUNC || sl <= tab.lmax);
                const int o = on ? (prow + sl) * (int)sizeof(cplx<TB>) : BUF_OOB;
                cplx<TB> a = buf_load<TB>(b1, o), b = buf_load<TB>(b2, o);
                const TB kz = FI::in_active(t, q) ? (TB)tab.kz[sl] : (TB)0;
                if (d1) a = mk<TB>(-a.y * kz, a.x * kz);
                if (d2) b = mk<TB>(-b.y * kz, b.x * kz);
                if (n == 0 || n == NZ / 2) {
                    a.y = 0;
                    b.y = 0;
                } else if (n > NZ / 2) {
                    a.y = -a.y;
                    b.y = -b.y;
                }
                v[q * FI::R0 + r] = mk<TB>(a.x - b.y, a.y + b.x);
            }
        }
    }
    __device__ __forceinline__ static void c2r_pair_buf(TB* x1, TB* x2, __amdgpu_buffer_rsrc_t b1, bool d1,
                                                        __amdgpu_buffer_rsrc_t b2, bool d2, int prow,
                                                        const SpecTables& tab, cplx<TB>* lds, const cplx<TB>* tw,
                                                        int t, int p, TB scale) {
        cplx<TB> v[E];
        load_pair_buf(v, b1, d1, b2, d2, prow, tab, t);
        FI::run(v, lds, tw, t, p);
#pragma unroll
        for (int e = 0; e < NPT; e++) {
            x1[e] = v[e].x * scale;
            x2[e] = v[e].y * scale;
        }
    }

    // Inverse-transform a pair; x1/x2 receive the scaled real fields.
    __device__ __forceinline__ static void c2r_pair(TB* x1, TB* x2, const cplx<TB>* s1, bool d1, const cplx<TB>* s2, bool d2,
                                                    bool valid, const SpecTables& tab, cplx<TB>* lds,
                                                    const cplx<TB>* tw, int t, int p, TB scale) {
        cplx<TB> v[E];
        load_pair(v, s1, d1, s2, d2, valid, tab, t);
        FI::run(v, lds, tw, t, p);
#pragma unroll
        for (int e = 0; e < NPT; e++) {
            x1[e] = v[e].x * scale;
            x2[e] = v[e].y * scale;
        }
    }

    // Forward-transform (q1 + i q2) and q3 at once (two exchange buffers) and
    // store the three half spectra: the results of r2c_pair(q1, q2) and
    // r2c_pair(q3, null), with one barrier per exchange for both.
    // With twr (the inverse plan's register twiddles; FI and FF must share
    // their radix sequence, as the palindromic power-of-two z plans do) the
    // forward transforms read their twiddles from registers too.
    static constexpr bool SAME_PLAN = std::is_same<typename ZPlanT<TB, NZ>::R,
                                                   typename SeqRev<typename ZPlanT<TB, NZ>::R>::type>::value;
    template <bool TWR>
    __device__ __forceinline__ static void r2c_two(const TB* q1, const TB* q2, const TB* q3, cplx<TB>* __restrict__ o1,
                                                   cplx<TB>* __restrict__ o2, cplx<TB>* __restrict__ o3, bool valid,
                                                   cplx<TB>* la, cplx<TB>* lb, const cplx<TB>* tw, int t, int p,
                                                   const typename FI::TwReg& twr) {
        cplx<TB> va[E], vb[E];
#pragma unroll
        for (int e = 0; e < NPT; e++) {
            va[e] = mk<TB>(q1[e], q2[e]);
            vb[e] = mk<TB>(q3[e], (TB)0);
        }
        if constexpr (TWR && SAME_PLAN)
            FF::run2_r(va, vb, la, lb, twr, t, p);
        else
            FF::run2(va, vb, la, lb, tw, t, p);
#pragma unroll
        for (int q = 0; q < FF::BL; q++) {
            if (!FF::out_active(t, q)) continue;
#pragma unroll
            for (int r = 0; r < FF::RL; r++) {
                la[LA::at(p, FF::out_index(t, q, r))] = va[q * FF::RL + r];
                lb[LA::at(p, FF::out_index(t, q, r))] = vb[q * FF::RL + r];
            }
        }
        exch_sync<WAVE>();
        if (valid) {
            for (int n = t; n <= NZ / 2; n += TPP) {
                const cplx<TB> zn = la[LA::at(p, n)];
                const cplx<TB> zm = la[LA::at(p, (NZ - n) % NZ)];
                o1[n] = mk<TB>((TB)0.5 * (zn.x + zm.x), (TB)0.5 * (zn.y - zm.y));
                o2[n] = mk<TB>((TB)0.5 * (zn.y + zm.y), (TB)0.5 * (zm.x - zn.x));
                o3[n] = lb[LA::at(p, n)];
            }
        }
        exch_sync<WAVE>();
    }

    // Forward-transform (q1 + i q2) and store the two half spectra (q2 may be null).
    __device__ __forceinline__ static void r2c_pair(const TB* q1, const TB* q2, cplx<TB>* __restrict__ o1,
                                                    cplx<TB>* __restrict__ o2, bool valid, cplx<TB>* lds,
                                                    const cplx<TB>* tw, int t, int p) {
        cplx<TB> v[E];
#pragma unroll
        for (int e = 0; e < NPT; e++) v[e] = mk<TB>(q1[e], q2 ? q2[e] : (TB)0);
        FF::run(v, lds, tw, t, p);
#pragma unroll
        for (int q = 0; q < FF::BL; q++) {
            if (!FF::out_active(t, q)) continue;
#pragma unroll
            for (int r = 0; r < FF::RL; r++) lds[LA::at(p, FF::out_index(t, q, r))] = v[q * FF::RL + r];
        }
        exch_sync<WAVE>();
        if (valid) {
            for (int n = t; n <= NZ / 2; n += TPP) {
                const cplx<TB> zn = lds[LA::at(p, n)];
                if (q2) {
                    const cplx<TB> zm = lds[LA::at(p, (NZ - n) % NZ)];
                    o1[n] = mk<TB>((TB)0.5 * (zn.x + zm.x), (TB)0.5 * (zn.y - zm.y));
                    o2[n] = mk<TB>((TB)0.5 * (zn.y + zm.y), (TB)0.5 * (zm.x - zn.x));
                } else {
                    o1[n] = zn;
                }
            }
        }
        exch_sync<WAVE>();
    }
};

template <typename TB, int NZ, int P, int NF>
// The bilinear z pass (24 real fields per point): with its inputs through
// buffer loads the FP32 pass fits 2 waves/SIMD in 240 VGPRs whatever the
// bound; in FP64 2 waves/SIMD cost 92 B/lane of scratch spills at 256^3 and
// still beat 1 wave with the overflow in AGPRs (256^3 tangent-map step
// 10.13 -> 9.46 ms, 128^3 1.35 -> 1.31 ms).
#ifndef NS3D_WAVES_Z2
#define NS3D_WAVES_Z2 2
#endif
#ifndef NS3D_WAVES_Z2_32
#define NS3D_WAVES_Z2_32 1
#endif
__global__ void __attribute__((amdgpu_waves_per_eu(NF == 1 ? (sizeof(TB) == 4 ? NS3D_WAVES_Z32 : NS3D_WAVES_Z)
                                                          : (sizeof(TB) == 4 ? NS3D_WAVES_Z2_32 : NS3D_WAVES_Z2), 8)))
__launch_bounds__((P * ZPlanT<TB, NZ>::TPP))
k_z_product(const cplx<TB>* __restrict__ Bf, cplx<TB>* __restrict__ C, long nc, long npen, int Mz, int nb,
            SpecTables tab, const cplx<TB>* __restrict__ twg, TB scale, TB* __restrict__ uphys, long nr, long bsc = 0,
            MemberMask mm = {}) {
    // B is zero on z-planes l > lmax (never written by the y pass): TRUNC skips them
    using Z = ZPencil<TB, NZ, P, true>;
    constexpr int TPP = Z::TPP, NPT = Z::NPT;
#ifndef NS3D_ZILP
#define NS3D_ZILP 1
#endif
#ifndef NS3D_ZILP32
#define NS3D_ZILP32 0
#endif
    // Two paired transforms at once (two exchange buffers): FP64 at power-of-two
    // lengths (256^3: z pass 0.458 -> 0.406 ms).  Measured no gain in FP32, and
    // register spills at the radix-3 lengths.
    constexpr bool ILP = NF == 1 && (NZ & (NZ - 1)) == 0 && (sizeof(TB) == 8 ? NS3D_ZILP != 0 : NS3D_ZILP32 != 0);
#ifndef NS3D_ZTWREG
#define NS3D_ZTWREG 1
#endif
#ifndef NS3D_ZTWREG32
#define NS3D_ZTWREG32 0
#endif
    // The inverse transforms' twiddles in registers, read from the table once
    // (NZ >= 32; below, the extra registers cost a wave per SIMD).  FP64 at
    // 256^3: z pass 0.375 -> 0.352 ms, at 512^3 3.68 -> 3.54 ms (+42 and +24
    // VGPRs, no spills); FP32 measured 1% slower and stays on the LDS table.
    constexpr bool TWR = NF == 1 && NZ >= 32 && (sizeof(TB) == 8 ? NS3D_ZTWREG != 0 : NS3D_ZTWREG32 != 0);
#ifndef NS3D_ZTWFWD
#define NS3D_ZTWFWD 1
#endif
    __shared__ cplx<TB> lds[(ILP ? 2 : 1) * P * (NZ + Z::PAD)];
    __shared__ cplx<TB> tw[NZ];
    const int bid0 = xcd_tile();
    if (bid0 >= nb || member_idle(mm)) return;
    const int bid = tile_order<NS3D_REV_Z>(bid0, nb);
    stage_twiddles<TB, NZ>(tw, twg);
    const int tid = threadIdx.x, t = tid % TPP, p = tid / TPP;
    const long pen = (long)bid * P + p;
    const bool valid = pen < npen;
    const long off = valid ? pen * Mz : 0;
    Bf += (long)blockIdx.z * bsc;
    C += (long)blockIdx.z * bsc;
    auto fld = [&](int f) { return Bf + (long)f * nc + off; };
#ifndef NS3D_ZBUF
#define NS3D_ZBUF 1
#endif
    // the first-pass inputs of fields fa and fb (the pipelined paths): with
    // NS3D_ZBUF through buffer resources over the block's valid pencils
    // (branch-free, no zero fill), else by predicated global loads
    const long pen0 = (long)bid * P;
    const int rbytes = (int)((npen - pen0 < P ? npen - pen0 : P) * Mz * (long)sizeof(cplx<TB>));
    auto rsrc = [&](int f) { return buf_rsrc(Bf + pen0 * Mz + (long)f * nc, rbytes); };
    auto fetch = [&](cplx<TB>* ra, cplx<TB>* rb, int fa, int fb) {
        if constexpr (NS3D_ZBUF != 0)
            Z::fetch_pair_buf(ra, rb, rsrc(fa), rsrc(fb), p * Mz, tab, t);
        else
            Z::fetch_pair(ra, rb, fld(fa), fld(fb), valid, tab, t);
    };
    // the bilinear path's paired inverse transform of fields fa, fb (d: times i kz)
    auto c2r = [&](TB* x1, TB* x2, int fa, bool d1, int fb, bool d2) {
        if constexpr (NS3D_ZBUF != 0)
            Z::c2r_pair_buf(x1, x2, rsrc(fa), d1, rsrc(fb), d2, p * Mz, tab, lds, tw, t, p, scale);
        else
            Z::c2r_pair(x1, x2, fld(fa), d1, fld(fb), d2, valid, tab, lds, tw, t, p, scale);
    };
    typename Z::FI::TwReg twi;
    if constexpr (TWR) Z::FI::load_tw(twi, tw, t);
    // field ids: s*9 + {0..2: w_b, 3..5: dx w_b, 6..8: dy w_b}
    TB Q[3][NPT];
    if constexpr (ILP) {
        // Two paired transforms at a time, the inputs of the next two loaded
        // before the current ones run; the same pairs and products as below.
        cplx<TB>* lb = lds + P * (NZ + Z::PAD);
        TB w[3][NPT], g0[NPT], g1[NPT], g2[NPT], g3[NPT], g4[NPT];
        TB kz[Z::NIN];
        cplx<TB> r1[Z::NIN], r2[Z::NIN], r3[Z::NIN], r4[Z::NIN];
        Z::kz_regs(kz, tab, t);
        fetch(r1, r2, 0, 1);
        fetch(r3, r4, 2, 3);
        // n1..n4: the fields of the next two pairs (-1: none)
        auto run_two = [&](TB* x1, TB* x2, bool d1, bool d2, TB* x3, TB* x4, bool d3, bool d4, int n1, int n2, int n3,
                           int n4) {
            cplx<TB> va[Z::E], vb[Z::E];
            Z::finish_pair(va, r1, d1, r2, d2, kz, t);
            Z::finish_pair(vb, r3, d3, r4, d4, kz, t);
            if (n1 >= 0) {
                fetch(r1, r2, n1, n2);
                fetch(r3, r4, n3, n4);
            }
            if constexpr (TWR)
                Z::FI::run2_r(va, vb, lds, lb, twi, t, p);
            else
                Z::FI::run2(va, vb, lds, lb, tw, t, p);
#pragma unroll
            for (int e = 0; e < NPT; e++) {
                x1[e] = va[e].x * scale;
                x2[e] = va[e].y * scale;
                x3[e] = vb[e].x * scale;
                x4[e] = vb[e].y * scale;
            }
        };
        // (w_x, w_y), (w_z, dx w_x)
        run_two(w[0], w[1], false, false, w[2], g0, false, false, 6, 0, 4, 7);
        if (uphys && valid) {
#pragma unroll
            for (int e = 0; e < NPT; e++) {
                const long g = pen * NZ + Z::FI::out_index(t, e / Z::FI::RL, e % Z::FI::RL);
                uphys[g] = w[0][e];
                uphys[nr + g] = w[1][e];
                uphys[2 * nr + g] = w[2][e];
            }
        }
        // (dy w_x, dz w_x), (dx w_y, dy w_y)
        run_two(g1, g2, false, true, g3, g4, false, false, 1, 5, 8, 2);
#pragma unroll
        for (int e = 0; e < NPT; e++) Q[0][e] = w[0][e] * g0[e] + w[1][e] * g1[e] + w[2][e] * g2[e];
        // (dz w_y, dx w_z), (dy w_z, dz w_z)
        run_two(g0, Q[2], true, false, g1, g2, false, true, -1, -1, -1, -1);
#pragma unroll
        for (int e = 0; e < NPT; e++) {
            Q[1][e] = w[0][e] * g3[e] + w[1][e] * g4[e] + w[2][e] * g0[e];
            Q[2][e] = w[0][e] * Q[2][e] + w[1][e] * g1[e] + w[2][e] * g2[e];
        }
    } else if constexpr (NF == 1) {
        // Software-pipelined: the loads of pair i+1 are issued before the
        // transform of pair i, so HBM latency overlaps the LDS passes.
        TB w[3][NPT], a[NPT], b[NPT], keep[NPT];
        TB kz[Z::NIN];
        cplx<TB> r1[Z::NIN], r2[Z::NIN];
        Z::kz_regs(kz, tab, t);
        fetch(r1, r2, 0, 1);
        auto run_pair = [&](TB* x1, TB* x2, bool d1, bool d2, int n1, int n2) {   // n1, n2: next pair (-1: none)
            cplx<TB> v[Z::E];
            Z::finish_pair(v, r1, d1, r2, d2, kz, t);
            if (n1 >= 0) fetch(r1, r2, n1, n2);
            if constexpr (TWR)
                Z::FI::run_r(v, lds, twi, t, p);
            else
                Z::FI::run(v, lds, tw, t, p);
#pragma unroll
            for (int e = 0; e < NPT; e++) {
                x1[e] = v[e].x * scale;
                x2[e] = v[e].y * scale;
            }
        };
        run_pair(w[0], w[1], false, false, 2, 3);   // w_x, w_y
        run_pair(w[2], keep, false, false, 6, 0);   // w_z, dx w_x
        if (uphys && valid) {
#pragma unroll
            for (int e = 0; e < NPT; e++) {
                const long g = pen * NZ + Z::FI::out_index(t, e / Z::FI::RL, e % Z::FI::RL);
                uphys[g] = w[0][e];
                uphys[nr + g] = w[1][e];
                uphys[2 * nr + g] = w[2][e];
            }
        }
        run_pair(a, b, false, true, 4, 7);          // dy w_x, dz w_x
#pragma unroll
        for (int e = 0; e < NPT; e++) Q[0][e] = w[0][e] * keep[e] + w[1][e] * a[e] + w[2][e] * b[e];
        run_pair(keep, a, false, false, 1, 5);      // dx w_y, dy w_y
        run_pair(b, Q[2], true, false, 8, 2);       // dz w_y, dx w_z
#pragma unroll
        for (int e = 0; e < NPT; e++) Q[1][e] = w[0][e] * keep[e] + w[1][e] * a[e] + w[2][e] * b[e];
        run_pair(a, b, false, true, -1, -1);        // dy w_z, dz w_z
#pragma unroll
        for (int e = 0; e < NPT; e++) Q[2][e] = w[0][e] * Q[2][e] + w[1][e] * a[e] + w[2][e] * b[e];
    } else {
        // Q_c = (v.grad) u_c + (u.grad) v_c  (convolute_2p3_UV, Source/advection_2_3.cu:401-475).
        // Each paired transform holds two u-derived or two v-derived fields,
        // never one of each: v (a Krylov vector) can be orders of magnitude
        // smaller than u, and a shared transform's rounding is relative to
        // the larger of its two fields.
        // ids: u_b = b, dx u_b = 3+b, dy u_b = 6+b; v_b = 9+b, dx v_b = 12+b, dy v_b = 15+b
        TB u[3][NPT], vv[3][NPT], a[NPT], b[NPT];
        c2r(u[0], u[1], 0, false, 1, false);
        c2r(vv[0], vv[1], 9, false, 10, false);
        c2r(u[2], a, 2, false, 3, false);      // u_z, dx u_x
        c2r(vv[2], b, 11, false, 12, false);   // v_z, dx v_x
#pragma unroll
        for (int e = 0; e < NPT; e++) Q[0][e] = vv[0][e] * a[e] + u[0][e] * b[e];
        c2r(a, b, 6, false, 0, true);          // dy u_x, dz u_x
#pragma unroll
        for (int e = 0; e < NPT; e++) Q[0][e] += vv[1][e] * a[e] + vv[2][e] * b[e];
        c2r(a, b, 15, false, 9, true);         // dy v_x, dz v_x
#pragma unroll
        for (int e = 0; e < NPT; e++) Q[0][e] += u[1][e] * a[e] + u[2][e] * b[e];
        c2r(a, b, 4, false, 7, false);         // dx u_y, dy u_y
#pragma unroll
        for (int e = 0; e < NPT; e++) Q[1][e] = vv[0][e] * a[e] + vv[1][e] * b[e];
        c2r(a, b, 1, true, 5, false);          // dz u_y, dx u_z
#pragma unroll
        for (int e = 0; e < NPT; e++) {
            Q[1][e] += vv[2][e] * a[e];
            Q[2][e] = vv[0][e] * b[e];
        }
        c2r(a, b, 13, false, 16, false);       // dx v_y, dy v_y
#pragma unroll
        for (int e = 0; e < NPT; e++) Q[1][e] += u[0][e] * a[e] + u[1][e] * b[e];
        c2r(a, b, 10, true, 14, false);        // dz v_y, dx v_z
#pragma unroll
        for (int e = 0; e < NPT; e++) {
            Q[1][e] += u[2][e] * a[e];
            Q[2][e] += u[0][e] * b[e];
        }
        c2r(a, b, 8, false, 2, true);          // dy u_z, dz u_z
#pragma unroll
        for (int e = 0; e < NPT; e++) Q[2][e] += vv[1][e] * a[e] + vv[2][e] * b[e];
        c2r(a, b, 17, false, 11, true);        // dy v_z, dz v_z
#pragma unroll
        for (int e = 0; e < NPT; e++) Q[2][e] += u[1][e] * a[e] + u[2][e] * b[e];
    }
    cplx<TB>* o0 = C + off;
    cplx<TB>* o1 = C + nc + off;
    cplx<TB>* o2 = C + 2 * nc + off;
    if constexpr (ILP) {
        Z::template r2c_two<TWR && NS3D_ZTWFWD != 0>(Q[0], Q[1], Q[2], o0, o1, o2, valid, lds, lds + P * (NZ + Z::PAD),
                                                     tw, t, p, twi);
    } else {
        Z::r2c_pair(Q[0], Q[1], o0, o1, valid, lds, tw, t, p);
        Z::r2c_pair(Q[2], nullptr, o2, nullptr, valid, lds, tw, t, p);
    }
}

// ---------------------------------------------------------------- P4 / generic
// C2C transform along x (AXIS 0) or y (AXIS 1) of nf spectral fields, out of
// place or in place (each block reads its pencils completely before writing).
template <typename TB, int NA, int P, int DIR, int AXIS, int E>
__global__ void __attribute__((amdgpu_waves_per_eu(NS3D_WAVES_XY, 8))) __launch_bounds__((P * XYPlan<NA, E>::TPP))
k_axis_c2c(const cplx<TB>* in, cplx<TB>* out, int nf, long nc, int Ny, int Mz, int Ms, int nlt, int nb,
           const cplx<TB>* __restrict__ twg, long bsc = 0, MemberMask mm = {}) {
    constexpr int TPP = XYPlan<NA, E>::TPP;
    using F = PencilFFT<TB, NA, TPP, DIR, typename XYPlan<NA, E>::R, IndexMajor<P>>;
    __shared__ cplx<TB> lds[NA * P];
    __shared__ cplx<TB> tw[NA];
    const int bid0 = xcd_tile();
    if (bid0 >= nb || member_idle(mm)) return;
    const int bid = tile_order<NS3D_REV_YFWD>(bid0, nb);
    stage_twiddles<TB, NA>(tw, twg);
    const int tid = threadIdx.x, p = tid % P, t = tid / P;
    const int o = bid / nlt, lt = bid % nlt;
    const int l = lt * P + p;
    const bool lv = l < Mz;
    const long S = AXIS == 0 ? (long)Ny * Ms : (long)Ms;
    const long base = AXIS == 0 ? (long)o * Ms + l : (long)o * Ny * Ms + l;
    in += (long)blockIdx.z * bsc;
    out += (long)blockIdx.z * bsc;
#pragma unroll 1
    NS3D_FIELD_LOOP(f, nf) {
        const cplx<TB>* src = in + (long)f * nc;
        cplx<TB>* dst = out + (long)f * nc;
        cplx<TB> v[F::E];
#pragma unroll
        for (int q = 0; q < F::B0; q++) {
#pragma unroll
            for (int r = 0; r < F::R0; r++) {
                cplx<TB> s = mk<TB>(0, 0);
                if (F::in_active(t, q) && lv) s = src[base + F::in_index(t, q, r) * S];
                v[q * F::R0 + r] = s;
            }
        }
        __syncthreads();  // in-place use: every pencil element is read before any is written
        F::run(v, lds, tw, t, p);
        if (lv) {
#pragma unroll
            for (int q = 0; q < F::BL; q++) {
                if (!F::out_active(t, q)) continue;
#pragma unroll
                for (int r = 0; r < F::RL; r++) dst[base + F::out_index(t, q, r) * S] = v[q * F::RL + r];
            }
        }
    }
}

// ---------------------------------------------------------------- P5 epilogues
// EPI_STAGE32: EPI_STAGE with w, u0 and out in FP32 (the FP32-state RK3 loop);
// the arithmetic and the forcing stay FP64.
// EPI_INC / EPI_INC_FINAL: the SSP stages in increment form over a base state
// u^n held as a pair of FP32 fields, u^n = hi + lo (the FP32-increment RK3
// loop).  With w = base + d (d: the previous stage's FP32 increment, u0; null
// in the first stage):
//   EPI_INC        out = a1*(d + dt*L(w)),  base = hi (w)            out FP32
//   EPI_INC_FINAL  r = base + a1*(d + dt*L(w)),  base = hi + lo (w, out32):
//                  hi = fp32(r) written to w, lo = fp32(r - hi) to out32
// which is RK3_SSP's u1 = u + dt L(u), u2 = 3/4 u + 1/4 (u1 + dt L(u1)),
// u = 1/3 u + 2/3 (u2 + dt L(u2)) with a1 = 1, 1/4, 2/3 and the stage values
// u1 = u + d1, u2 = u + d2.  out may alias u0 (each mode is read, then written,
// by one thread).
enum EpiMode {
    EPI_Q = 0, EPI_STAGE = 1, EPI_RHS = 2, EPI_UV = 3, EPI_UV_RHS = 4, EPI_IUV = 5, EPI_STAGE32 = 6,
    EPI_INC = 7, EPI_INC_FINAL = 8
};

struct EpiArgs {
    cplx<double>* w[3];          // state (read; written in place for UV/UV_RHS/IUV); cplx<float> for EPI_STAGE32
    const cplx<double>* u0[3];   // SSP combination base (EPI_STAGE, may be null); cplx<float> for EPI_STAGE32
    const cplx<double>* f[3];    // forcing (EPI_STAGE / EPI_RHS, may be null)
    cplx<double>* out[3];        // output (EPI_Q, EPI_STAGE, EPI_RHS); cplx<float> for EPI_STAGE32
    double dt, Re, a0, a1;       // EPI_STAGE: out = a0*u0 + a1*E(w)
    double nu, dtnu;             // 1/Re and dt/Re (formed once on the host)
    // per-member dt and Re of a batch (device arrays indexed by blockIdx.z; null: dt, Re above)
    const double* dtv;
    const double* rev;
    int ipq;                     // reference IPQ rule (Q1)
    // mode (j,k,l) of w / u0 / out / f sits at (j*ny + k)*m + l: m = Mz for the
    // reference layout, a padded row stride for the library's own stage buffers
    int ny, mw, mu0, mo, mf;
    // batch member strides (elements) of w, u0 and out; f is shared by the members
    long bw, bu0, bo;
    // sticky non-finite flag: set to 1 by any lane that writes a NaN or Inf (may be null)
    int* nf;
    // EPI_INC_FINAL: the lo half of the state (read and written), rows of stride mw
    cplx<float>* out32[3];
};

// Non-finite flag of the fused epilogues and the scalar stage: every lane that
// produced a NaN or Inf stores 1 (a plain vector store; all writers store the
// same value).  The flag is read and cleared by the host.
__device__ __forceinline__ void flag_nonfinite(int* nf, bool bad) {
    if (bad && nf) *(volatile int*)nf = 1;
}
template <typename T> __device__ __forceinline__ bool finite2(cplx<T> a) { return isfinite(a.x) && isfinite(a.y); }

// Projection/diffusion/update at one mode (Source/diffusion.cu:96-365).
template <int MODE>
__device__ __forceinline__ void epilogue(const EpiArgs& ea, const SpecTables& tab, int j, int k, int l,
                                         cplx<double> Qx, cplx<double> Qy, cplx<double> Qz) {
    const long row = (long)j * ea.ny + k;
    const long gw = row * ea.mw + l, gu = row * ea.mu0 + l, go = row * ea.mo + l, gf = row * ea.mf + l;
    if (MODE == EPI_Q) {
        ea.out[0][go] = Qx;
        ea.out[1][go] = Qy;
        ea.out[2][go] = Qz;
        flag_nonfinite(ea.nf, !(finite2(Qx) && finite2(Qy) && finite2(Qz)));
        return;
    }
    bool bad = false;
    const ProjCoef c = proj_coef(tab, j, k, l);
    const double AQr[3] = {c.a11 * Qx.x + c.a12 * Qy.x + c.a13 * Qz.x, c.a12 * Qx.x + c.a22 * Qy.x + c.a23 * Qz.x,
                           c.a13 * Qx.x + c.a23 * Qy.x + c.a33 * Qz.x};
    const double AQi[3] = {c.a11 * Qx.y + c.a12 * Qy.y + c.a13 * Qz.y, c.a12 * Qx.y + c.a22 * Qy.y + c.a23 * Qz.y,
                           c.a13 * Qx.y + c.a23 * Qy.y + c.a33 * Qz.y};
    const cplx<double> Q[3] = {Qx, Qy, Qz};
    const double dt = ea.dt;
#pragma unroll
    for (int b = 0; b < 3; b++) {
        cplx<double> r;
        if (MODE == EPI_STAGE32) {
            const cplx<double> w = cvt<double>(reinterpret_cast<const cplx<float>*>(ea.w[b])[gw]);
            const double G = (1.0 - c.din * ea.dtnu);
            cplx<double> f = mk<double>(0.0, 0.0);
            if (ea.f[b]) f = ea.f[b][gf];
            r.x = dt * AQr[b] + G * w.x + dt * f.x - dt * Q[b].x;
            r.y = dt * AQi[b] + G * w.y + dt * f.y - dt * Q[b].y;
            if (ea.u0[b]) {
                const cplx<double> u0 = cvt<double>(reinterpret_cast<const cplx<float>*>(ea.u0[b])[gu]);
                r.x = ea.a0 * u0.x + ea.a1 * r.x;
                r.y = ea.a0 * u0.y + ea.a1 * r.y;
            }
            const cplx<float> rf = cvt<float>(r);
            reinterpret_cast<cplx<float>*>(ea.out[b])[go] = rf;
            bad |= !finite2(rf);
            continue;
        }
        if (MODE == EPI_INC || MODE == EPI_INC_FINAL) {
            cplx<float>* hi = reinterpret_cast<cplx<float>*>(ea.w[b]);
            cplx<double> base = cvt<double>(hi[gw]);
            if (MODE == EPI_INC_FINAL) {
                const cplx<float> lo = ea.out32[b][gw];
                base.x += (double)lo.x;
                base.y += (double)lo.y;
            }
            const cplx<double> d =
                ea.u0[b] ? cvt<double>(reinterpret_cast<const cplx<float>*>(ea.u0[b])[gu]) : mk<double>(0.0, 0.0);
            cplx<double> f = mk<double>(0.0, 0.0);
            if (ea.f[b]) f = ea.f[b][gf];
            const double Gd = c.din * ea.dtnu;
            r.x = ea.a1 * (d.x + (dt * AQr[b] - Gd * (base.x + d.x) + dt * f.x - dt * Q[b].x));
            r.y = ea.a1 * (d.y + (dt * AQi[b] - Gd * (base.y + d.y) + dt * f.y - dt * Q[b].y));
            if (MODE == EPI_INC) {
                const cplx<float> rf = cvt<float>(r);
                reinterpret_cast<cplx<float>*>(ea.out[b])[go] = rf;
                bad |= !finite2(rf);
            } else {
                r.x += base.x;
                r.y += base.y;
                const cplx<float> h = cvt<float>(r);
                hi[gw] = h;
                ea.out32[b][gw] = mk<float>((float)(r.x - (double)h.x), (float)(r.y - (double)h.y));
                bad |= !finite2(r);
            }
            continue;
        }
        const cplx<double> w = ea.w[b][gw];
        if (MODE == EPI_STAGE) {
            const double G = (1.0 - c.din * ea.dtnu);
            cplx<double> f = mk<double>(0.0, 0.0);
            if (ea.f[b]) f = ea.f[b][gf];
            r.x = dt * AQr[b] + G * w.x + dt * f.x - dt * Q[b].x;
            r.y = dt * AQi[b] + G * w.y + dt * f.y - dt * Q[b].y;
            if (ea.u0[b]) {
                const cplx<double> u0 = ea.u0[b][gu];
                r.x = ea.a0 * u0.x + ea.a1 * r.x;
                r.y = ea.a0 * u0.y + ea.a1 * r.y;
            }
            ea.out[b][go] = r;
            bad |= !finite2(r);
        } else if (MODE == EPI_RHS) {
            const double G = (-c.din * ea.nu);
            cplx<double> f = mk<double>(0.0, 0.0);
            if (ea.f[b]) f = ea.f[b][gf];
            r.x = AQr[b] + G * w.x + f.x - Q[b].x;
            r.y = AQi[b] + G * w.y + f.y - Q[b].y;
            ea.out[b][go] = r;
            bad |= !finite2(r);
        } else if (MODE == EPI_UV) {
            const double G = (1.0 - c.din * ea.dtnu);
            r.x = dt * AQr[b] + G * w.x - dt * Q[b].x;
            r.y = dt * AQi[b] + G * w.y - dt * Q[b].y;
            ea.w[b][gw] = r;
            bad |= !finite2(r);
        } else if (MODE == EPI_UV_RHS) {
            const double G = (-c.din * ea.nu);
            r.x = AQr[b] + G * w.x - Q[b].x;
            r.y = AQi[b] + G * w.y - Q[b].y;
            ea.w[b][gw] = r;
            bad |= !finite2(r);
        } else {  // EPI_IUV
            const double G = (1.0 + c.din * ea.dtnu);
            r.x = (dt * AQr[b] + w.x - dt * Q[b].x) / G;
            r.y = (dt * AQi[b] + w.y - dt * Q[b].y) / G;
            ea.w[b][gw] = r;
            bad |= !finite2(r);
        }
    }
    flag_nonfinite(ea.nf, bad);
}

// set_to_zero_real (Source/advection_2_3.cu:35-70)
__device__ __forceinline__ void ipq_rule(cplx<double>& Qx, cplx<double>& Qy, cplx<double>& Qz, int ipq, bool origin) {
    if (ipq) {
        Qx.x = 0.0;
        Qy.x = 0.0;
        Qz.x = 0.0;
        if (origin) {
            Qx.y = 0.0;
            Qy.y = 0.0;
            Qz.y = 0.0;
        }
    }
}

template <typename TB, int NX, int P, int MODE, int E>
__global__ void __attribute__((amdgpu_waves_per_eu(sizeof(TB) == 4 ? NS3D_WAVES_XFWD32 : NS3D_WAVES_XY, 8))) __launch_bounds__((P * XYPlan<NX, E>::TPP))
k_x_fwd_epi(const cplx<TB>* __restrict__ D, long nc, int Ny, int Mz, int Ms, int nlt, int nb, SpecTables tab,
            const cplx<TB>* __restrict__ twg, EpiArgs ea, long bsc = 0, MemberMask mm = {}) {
    constexpr int TPP = XYPlan<NX, E>::TPP;
    using F = PencilFFT<TB, NX, TPP, -1, typename XYPlan<NX, E>::R, IndexMajor<P>>;
    constexpr bool ILP = NS3D_HOIST && (NS3D_XFWD_ILP == 2 || (NS3D_XFWD_ILP == 1 && sizeof(TB) == 8));
    __shared__ cplx<TB> lds[(ILP ? 2 : 1) * NX * P];
    __shared__ cplx<TB> tw[NX];
    const int bid0 = xcd_tile();
    if (bid0 >= nb || member_idle(mm)) return;
    const int bid = tile_order<NS3D_REV_XFWD>(bid0, nb);
    stage_twiddles<TB, NX>(tw, twg);
    const int tid = threadIdx.x, p = tid % P, t = tid / P;
    const int k = bid / nlt, lt = bid % nlt;
    const int l = lt * P + p;
    const bool lv = l < Mz;
    const long Si = (long)Ny * Ms;       // D: rows padded to Ms
    const long basei = (long)k * Ms + l;
    if (ea.dtv) {                        // per-member time step and Reynolds number
        const double dtm = ea.dtv[blockIdx.z], rem = ea.rev[blockIdx.z];
        ea.dt = dtm;
        ea.Re = rem;
        ea.nu = 1.0 / rem;
        ea.dtnu = dtm / rem;
    }
    if (blockIdx.z) {                    // batch member z
        const long z = blockIdx.z;
        D += z * bsc;
#pragma unroll
        for (int b = 0; b < 3; b++) {
            ea.w[b] += z * ea.bw;
            if (ea.u0[b]) ea.u0[b] += z * ea.bu0;
            ea.out[b] += z * ea.bo;
        }
    }
    cplx<TB> qv[3][F::E];
    // the three components are all live until the epilogue: load them together
#pragma unroll
    for (int b = 0; b < 3; b++) {
        const cplx<TB>* src = D + (long)b * nc;
#pragma unroll
        for (int q = 0; q < F::B0; q++) {
#pragma unroll
            for (int r = 0; r < F::R0; r++) {
                cplx<TB> s = mk<TB>(0, 0);
                if (F::in_active(t, q) && lv) s = src[basei + F::in_index(t, q, r) * Si];
                qv[b][q * F::R0 + r] = s;
            }
        }
        if (!NS3D_HOIST) F::run(qv[b], lds, tw, t, p);
    }
    if constexpr (ILP) {
        F::run2(qv[0], qv[1], lds, lds + NX * P, tw, t, p);
        F::run(qv[2], lds, tw, t, p);
    } else if (NS3D_HOIST) {
#pragma unroll
        for (int b = 0; b < 3; b++) F::run(qv[b], lds, tw, t, p);
    }
    if (!lv) return;
#pragma unroll
    for (int q = 0; q < F::BL; q++) {
        if (!F::out_active(t, q)) continue;
#pragma unroll
        for (int r = 0; r < F::RL; r++) {
            const int j = F::out_index(t, q, r);
            cplx<double> Qx = cvt<double>(qv[0][q * F::RL + r]);
            cplx<double> Qy = cvt<double>(qv[1][q * F::RL + r]);
            cplx<double> Qz = cvt<double>(qv[2][q * F::RL + r]);
            ipq_rule(Qx, Qy, Qz, ea.ipq, (j | k | l) == 0);
            epilogue<MODE>(ea, tab, j, k, l, Qx, Qy, Qz);
        }
    }
}

// Pointwise epilogue from stored Q^ (the standalone solve_advection_diffusion_projection*).
template <int MODE>
__global__ void __launch_bounds__(256)
k_update(const cplx<double>* __restrict__ Qx, const cplx<double>* __restrict__ Qy, const cplx<double>* __restrict__ Qz,
         long nc, int Ny, int Mz, SpecTables tab, EpiArgs ea) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nc) return;
    const int l = (int)(g % Mz);
    const long t1 = g / Mz;
    const int k = (int)(t1 % Ny), j = (int)(t1 / Ny);
    epilogue<MODE>(ea, tab, j, k, l, Qx[g], Qy[g], Qz[g]);
}

}  // namespace ns3d
