// fft_core.hpp -- register/LDS mixed-radix Stockham FFT building blocks for gfx950.
//
// A "pencil" is one 1D transform of length N.  A workgroup transforms P
// pencils at once; TPP threads cooperate on each pencil.  Every pass with
// radix R performs N/R butterflies; thread t owns butterflies t, t+TPP, ...
// Data live in registers; between passes they are exchanged through LDS.
//
// Pass i (Stockham autosort, Ns = product of the radices before i):
//   inputs  x[jj + r*N/R],  r < R
//   twiddle x_r *= W^(r*(jj mod Ns)*N/(Ns*R)),  W = exp(DIR*2*pi*i/N)
//   R-point DFT
//   outputs y[(jj/Ns)*Ns*R + (jj mod Ns) + r*Ns]
// so the first pass reads x[jj + r*N/R1] and the last pass writes
// y[jj + r*N/Rlast]: callers load and store those patterns directly from and
// to global memory, and only the inner exchanges touch LDS.
#pragma once
#include <type_traits>
#include "ns3d_common.hpp"

namespace ns3d {

// ---------------------------------------------------------------- radix DFTs
template <int R> struct Radix;

template <> struct Radix<2> {
    template <int DIR, typename T> __device__ __forceinline__ static void dft(cplx<T>* v) {
        cplx<T> a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    }
};

template <> struct Radix<3> {
    template <int DIR, typename T> __device__ __forceinline__ static void dft(cplx<T>* v) {
        const T h = (T)0.86602540378443864676372317075294;  // sqrt(3)/2
        cplx<T> a0 = v[0], a1 = v[1], a2 = v[2];
        cplx<T> s = a1 + a2;
        cplx<T> d = a1 - a2;
        cplx<T> m = mk<T>(a0.x - (T)0.5 * s.x, a0.y - (T)0.5 * s.y);
        cplx<T> r = mul_i<DIR>(cscale(d, h));
        v[0] = a0 + s;
        v[1] = m + r;
        v[2] = m - r;
    }
};

// 5-point DFT, W = exp(DIR 2 pi i / 5): with t1 = a1 + a4, t2 = a2 + a3,
// t3 = a1 - a4, t4 = a2 - a3 the outputs pair up as m +- DIR i n,
//   X1,4 = a0 + c1 t1 + c2 t2 +- DIR i (s1 t3 + s2 t4)
//   X2,3 = a0 + c2 t1 + c1 t2 +- DIR i (s2 t3 - s1 t4)
// (c_k = cos(2 pi k/5), s_k = sin(2 pi k/5)).
template <> struct Radix<5> {
    template <int DIR, typename T> __device__ __forceinline__ static void dft(cplx<T>* v) {
        const T c1 = (T)0.30901699437494742410229341718281906;   // cos(2 pi/5)
        const T c2 = (T)-0.80901699437494742410229341718281906;  // cos(4 pi/5)
        const T s1 = (T)0.95105651629515357211643933337938214;   // sin(2 pi/5)
        const T s2 = (T)0.58778525229247312916870595463907277;   // sin(4 pi/5)
        const cplx<T> a0 = v[0];
        const cplx<T> t1 = v[1] + v[4], t2 = v[2] + v[3];
        const cplx<T> t3 = v[1] - v[4], t4 = v[2] - v[3];
        const cplx<T> m1 = mk<T>(a0.x + c1 * t1.x + c2 * t2.x, a0.y + c1 * t1.y + c2 * t2.y);
        const cplx<T> m2 = mk<T>(a0.x + c2 * t1.x + c1 * t2.x, a0.y + c2 * t1.y + c1 * t2.y);
        const cplx<T> n1 = mul_i<DIR>(mk<T>(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
        const cplx<T> n2 = mul_i<DIR>(mk<T>(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
        v[0] = a0 + t1 + t2;
        v[1] = m1 + n1;
        v[4] = m1 - n1;
        v[2] = m2 + n2;
        v[3] = m2 - n2;
    }
};

template <> struct Radix<4> {
    template <int DIR, typename T> __device__ __forceinline__ static void dft(cplx<T>* v) {
        cplx<T> s02 = v[0] + v[2], d02 = v[0] - v[2];
        cplx<T> s13 = v[1] + v[3], d13 = mul_i<DIR>(v[1] - v[3]);
        v[0] = s02 + s13;
        v[1] = d02 + d13;
        v[2] = s02 - s13;
        v[3] = d02 - d13;
    }
};

template <> struct Radix<8> {
    template <int DIR, typename T> __device__ __forceinline__ static void dft(cplx<T>* v) {
        const T c = (T)0.70710678118654752440084436210485;  // 1/sqrt(2)
        cplx<T> e[4] = {v[0], v[2], v[4], v[6]};
        cplx<T> o[4] = {v[1], v[3], v[5], v[7]};
        Radix<4>::dft<DIR>(e);
        Radix<4>::dft<DIR>(o);
        // W8^1 = (1 + DIR i)/sqrt2, W8^2 = DIR i, W8^3 = (-1 + DIR i)/sqrt2
        cplx<T> o1 = cscale(o[1] + mul_i<DIR>(o[1]), c);
        cplx<T> o2 = mul_i<DIR>(o[2]);
        cplx<T> o3 = cscale(mul_i<DIR>(o[3]) - o[3], c);
        v[0] = e[0] + o[0];
        v[4] = e[0] - o[0];
        v[1] = e[1] + o1;
        v[5] = e[1] - o1;
        v[2] = e[2] + o2;
        v[6] = e[2] - o2;
        v[3] = e[3] + o3;
        v[7] = e[3] - o3;
    }
};

template <> struct Radix<16> {
    template <int DIR, typename T> __device__ __forceinline__ static void dft(cplx<T>* v) {
        const T c1 = (T)0.92387953251128675612818318939679;  // cos(pi/8)
        const T s1 = (T)0.38268343236508977172845998403040;  // sin(pi/8)
        const T c2 = (T)0.70710678118654752440084436210485;
        // columns n2: y[n2][k1] = DFT4 over n1 of v[4*n1+n2]
        cplx<T> y[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; n2++) {
            cplx<T> a[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
            Radix<4>::dft<DIR>(a);
#pragma unroll
            for (int k1 = 0; k1 < 4; k1++) y[n2][k1] = a[k1];
        }
        // twiddles W16^(n2*k1), W16 = exp(DIR*i*pi/8)
        const cplx<T> w1 = mk<T>(c1, (T)DIR * s1), w2 = mk<T>(c2, (T)DIR * c2), w3 = mk<T>(s1, (T)DIR * c1);
        const cplx<T> w6 = mk<T>(-c2, (T)DIR * c2), w9 = mk<T>(-c1, (T)(-DIR) * s1);
        y[1][1] = cmul(y[1][1], w1);
        y[1][2] = cmul(y[1][2], w2);
        y[1][3] = cmul(y[1][3], w3);
        y[2][1] = cmul(y[2][1], w2);
        y[2][2] = mul_i<DIR>(y[2][2]);
        y[2][3] = cmul(y[2][3], w6);
        y[3][1] = cmul(y[3][1], w3);
        y[3][2] = cmul(y[3][2], w6);
        y[3][3] = cmul(y[3][3], w9);
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) {
            cplx<T> a[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
            Radix<4>::dft<DIR>(a);
#pragma unroll
            for (int k2 = 0; k2 < 4; k2++) v[k1 + 4 * k2] = a[k2];
        }
    }
};

// ---------------------------------------------------------------- plans
template <int... R> struct Seq {};

template <typename S> struct SeqInfo;
template <int... R> struct SeqInfo<Seq<R...>> {
    static constexpr int n = sizeof...(R);
    static constexpr int rad[n] = {R...};
    static constexpr int prod_before(int i) {
        int p = 1;
        for (int m = 0; m < i; m++) p *= rad[m];
        return p;
    }
};

template <typename S> struct SeqRev;
template <int... R> struct SeqRevAcc;
template <> struct SeqRev<Seq<>> { using type = Seq<>; };
template <int A, int... R> struct SeqRev<Seq<A, R...>> {
    template <typename S2> struct Append;
    template <int... B> struct Append<Seq<B...>> { using type = Seq<B..., A>; };
    using type = typename Append<typename SeqRev<Seq<R...>>::type>::type;
};

// Plans: radices with at most 8 elements per thread per pass (TPP = N/8 for
// the powers of two; the radix-3 and radix-5 passes of other lengths leave
// some threads idle in their pass, the only pass whose butterfly count N/R
// is not a multiple of TPP).  The inverse plan of the fused z-pass and the forward
// plan are mirror images so the product stage needs no exchange.
template <int N> struct Plan;
template <> struct Plan<8> { using R = Seq<8>; static constexpr int TPP = 1; };
template <> struct Plan<16> { using R = Seq<2, 8>; static constexpr int TPP = 2; };
template <> struct Plan<24> { using R = Seq<3, 8>; static constexpr int TPP = 3; };
template <> struct Plan<32> { using R = Seq<4, 8>; static constexpr int TPP = 4; };
template <> struct Plan<48> { using R = Seq<3, 2, 8>; static constexpr int TPP = 6; };
template <> struct Plan<64> { using R = Seq<8, 8>; static constexpr int TPP = 8; };
template <> struct Plan<80> { using R = Seq<5, 2, 8>; static constexpr int TPP = 10; };
template <> struct Plan<96> { using R = Seq<3, 4, 8>; static constexpr int TPP = 12; };
template <> struct Plan<128> { using R = Seq<2, 8, 8>; static constexpr int TPP = 16; };
template <> struct Plan<160> { using R = Seq<5, 4, 8>; static constexpr int TPP = 20; };
template <> struct Plan<192> { using R = Seq<3, 8, 8>; static constexpr int TPP = 24; };
template <> struct Plan<256> { using R = Seq<4, 8, 8>; static constexpr int TPP = 32; };
template <> struct Plan<320> { using R = Seq<5, 8, 8>; static constexpr int TPP = 40; };
template <> struct Plan<384> { using R = Seq<3, 2, 8, 8>; static constexpr int TPP = 48; };
template <> struct Plan<512> { using R = Seq<8, 8, 8>; static constexpr int TPP = 64; };

// z-pass plans: 4 points per thread (TPP = N/4), radix-4 last so that the
// inverse plan and its mirror-image forward plan share the product layout with
// few live registers per thread.
template <int N> struct ZPlan;
template <> struct ZPlan<8> { using R = Seq<2, 4>; static constexpr int TPP = 2; };
template <> struct ZPlan<16> { using R = Seq<4, 4>; static constexpr int TPP = 4; };
template <> struct ZPlan<24> { using R = Seq<3, 2, 4>; static constexpr int TPP = 6; };
template <> struct ZPlan<32> { using R = Seq<2, 4, 4>; static constexpr int TPP = 8; };
template <> struct ZPlan<48> { using R = Seq<3, 4, 4>; static constexpr int TPP = 12; };
template <> struct ZPlan<64> { using R = Seq<4, 4, 4>; static constexpr int TPP = 16; };
template <> struct ZPlan<80> { using R = Seq<5, 4, 4>; static constexpr int TPP = 20; };
template <> struct ZPlan<96> { using R = Seq<3, 2, 4, 4>; static constexpr int TPP = 24; };
template <> struct ZPlan<128> { using R = Seq<2, 4, 4, 4>; static constexpr int TPP = 32; };
template <> struct ZPlan<160> { using R = Seq<5, 2, 4, 4>; static constexpr int TPP = 40; };
template <> struct ZPlan<192> { using R = Seq<3, 4, 4, 4>; static constexpr int TPP = 48; };
template <> struct ZPlan<256> { using R = Seq<4, 4, 4, 4>; static constexpr int TPP = 64; };
template <> struct ZPlan<320> { using R = Seq<5, 4, 4, 4>; static constexpr int TPP = 80; };
template <> struct ZPlan<384> { using R = Seq<3, 2, 4, 4, 4>; static constexpr int TPP = 96; };
template <> struct ZPlan<512> { using R = Seq<2, 4, 4, 4, 4>; static constexpr int TPP = 128; };

// z-pass plan by bulk precision: ZPlan (4 points per thread), or for FP32
// with NS3D_ZPLAN32 = 8 the 8-point Plan (one LDS exchange fewer per transform).
#ifndef NS3D_ZPLAN32
#define NS3D_ZPLAN32 8
#endif
template <typename TB, int N> struct ZSel {
    using type = typename std::conditional<(sizeof(TB) == 4 && NS3D_ZPLAN32 == 8), Plan<N>, ZPlan<N>>::type;
};
template <typename TB, int N> using ZPlanT = typename ZSel<TB, N>::type;

// Registers per thread for a plan: max over passes of ceil((N/R)/TPP)*R.
template <int N, int TPP, typename S> struct PlanRegs;
template <int N, int TPP, int... R> struct PlanRegs<N, TPP, Seq<R...>> {
    static constexpr int one(int r) { return ((N / r + TPP - 1) / TPP) * r; }
    static constexpr int value() {
        int e = 0;
        const int rr[] = {R...};
        for (int i = 0; i < (int)sizeof...(R); i++) e = one(rr[i]) > e ? one(rr[i]) : e;
        return e;
    }
};

// ------------------------------------------------------------- pencil FFT
// LDS addressing policies.  IndexMajor: element idx of pencil p at idx*P+p
// (pencils are consecutive z-columns: conflict-free when lanes are p-fastest).
// PencilMajor: p*(N+PAD)+idx (pencils contiguous; lanes t-fastest).
template <int P> struct IndexMajor {
    template <int NS = 0> __device__ __forceinline__ static int at(int p, int idx) { return idx * P + p; }
};
template <int N, int PAD> struct PencilMajor {
    template <int NS = 0> __device__ __forceinline__ static int at(int p, int idx) { return p * (N + PAD) + idx; }
};
// PencilMajor with a bank swizzle.  The exchange after a pass with stride Ns
// writes runs of Ns consecutive elements separated by Ns*R; XOR-ing the low
// log2(G) index bits with (idx/G)*Ns spreads such writes over all banks of a
// G-element group, and stays a permutation within each aligned group, so the
// next pass's contiguous reads remain conflict-free.  NS = 0: no swizzle.
template <int N, int PAD, int G> struct PencilMajorSw {
    template <int NS = 0> __device__ __forceinline__ static int at(int p, int idx) {
        if constexpr (NS > 0 && NS < G && N % G == 0) idx ^= ((idx / G) * NS) & (G - 1);
        return p * (N + PAD) + idx;
    }
};

// Ordering of an LDS exchange.  WAVE = false: a workgroup barrier.  WAVE =
// true, for layouts where each pencil's LDS region is touched only by that
// pencil's threads and a pencil's threads lie in one wavefront: LDS operations
// of one wavefront are performed in program order, so the exchange needs no
// hardware barrier, only a wavefront-scope fence that keeps the compiler from
// moving LDS accesses across it; the other waves of the block run on.
template <bool WAVE> __device__ __forceinline__ void exch_sync() {
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// Per-thread register twiddles of a plan (PencilFFT::load_tw); plans with the
// same radix sequence share the slot layout, and then one set serves both
// transform directions.
template <typename T, int K> struct TwRegs {
    cplx<T> w[K];
};

template <typename T, int N, int TPP, int DIR, typename S, typename Addr, bool WAVE = false>
struct PencilFFT {
    using Info = SeqInfo<S>;
    static constexpr int NP = Info::n;
    static constexpr int E = PlanRegs<N, TPP, S>::value();
    static constexpr int R0 = Info::rad[0];
    static constexpr int RL = Info::rad[NP - 1];
    // first-pass input pattern: v[q*R0 + r] <-> x[t + q*TPP + r*N/R0]
    static constexpr int B0 = (N / R0 + TPP - 1) / TPP;
    // last-pass output pattern: v[q*RL + r] <-> y[t + q*TPP + r*N/RL]
    static constexpr int BL = (N / RL + TPP - 1) / TPP;

    __device__ __forceinline__ static bool in_active(int t, int q) { return (N / R0) % TPP == 0 || t + q * TPP < N / R0; }
    __device__ __forceinline__ static int in_index(int t, int q, int r) { return t + q * TPP + r * (N / R0); }
    __device__ __forceinline__ static bool out_active(int t, int q) { return (N / RL) % TPP == 0 || t + q * TPP < N / RL; }
    __device__ __forceinline__ static int out_index(int t, int q, int r) { return t + q * TPP + r * (N / RL); }

    // Register twiddles: a thread that runs many transforms of one plan (the
    // z pass) reads its twiddles of every pass from the table once (load_tw)
    // instead of in every transform.  Slot tw_off(I) + q*(R-1) + r-1 holds
    // tw[r*jm*N/(Ns*R)] of pass I, as stored in the table.
    static constexpr int tw_count(int i) {
        return Info::prod_before(i) > 1 ? ((N / Info::rad[i] + TPP - 1) / TPP) * (Info::rad[i] - 1) : 0;
    }
    static constexpr int tw_off(int i) {
        int o = 0;
        for (int m = 0; m < i; m++) o += tw_count(m);
        return o;
    }
    using TwReg = TwRegs<T, (tw_off(NP) > 0 ? tw_off(NP) : 1)>;
    template <int I>
    __device__ __forceinline__ static void load_tw_pass(TwReg& tr, const cplx<T>* __restrict__ tw, int t) {
        constexpr int R = Info::rad[I];
        constexpr int Ns = Info::prod_before(I);
        constexpr int NB = N / R;
        constexpr int B = (NB + TPP - 1) / TPP;
        if constexpr (Ns > 1) {
#pragma unroll
            for (int q = 0; q < B; q++) {
                const int jj = t + q * TPP;
                const int jm = (NB % TPP == 0 || jj < NB) ? jj % Ns : 0;
#pragma unroll
                for (int r = 1; r < R; r++) tr.w[tw_off(I) + q * (R - 1) + r - 1] = tw[r * jm * (N / (Ns * R))];
            }
        }
        if constexpr (I + 1 < NP) load_tw_pass<I + 1>(tr, tw, t);
    }
    __device__ __forceinline__ static void load_tw(TwReg& tr, const cplx<T>* __restrict__ tw, int t) {
        load_tw_pass<0>(tr, tw, t);
    }

    // tw[m] = exp(-2*pi*i*m/N); DIR=+1 uses the conjugate.
    // Butterflies of pass I (twiddles, then the radix-R DFTs) on this thread's
    // values; twiddles from the table tw, or with tr from registers.
    template <int I>
    __device__ __forceinline__ static void butterflies(cplx<T>* v, const cplx<T>* __restrict__ tw, int t,
                                                       const TwReg* tr = nullptr) {
        constexpr int R = Info::rad[I];
        constexpr int Ns = Info::prod_before(I);
        constexpr int NB = N / R;
        constexpr int B = (NB + TPP - 1) / TPP;
#pragma unroll
        for (int q = 0; q < B; q++) {
            const int jj = t + q * TPP;
            if (NB % TPP == 0 || jj < NB) {
                if constexpr (Ns > 1) {
                    const int jm = jj % Ns;
#pragma unroll
                    for (int r = 1; r < R; r++) {
                        cplx<T> w = tr ? tr->w[tw_off(I) + q * (R - 1) + r - 1] : tw[r * jm * (N / (Ns * R))];
                        if (DIR > 0) w.y = -w.y;
                        v[q * R + r] = cmul(v[q * R + r], w);
                    }
                }
                Radix<R>::template dft<DIR>(v + q * R);
            }
        }
    }
    // Stockham exchange after pass I: store in the order pass I+1 reads.
    template <int I>
    __device__ __forceinline__ static void exch_write(const cplx<T>* v, cplx<T>* lds, int t, int p) {
        constexpr int R = Info::rad[I];
        constexpr int Ns = Info::prod_before(I);
        constexpr int NB = N / R;
        constexpr int B = (NB + TPP - 1) / TPP;
#pragma unroll
        for (int q = 0; q < B; q++) {
            const int jj = t + q * TPP;
            if (NB % TPP == 0 || jj < NB) {
                const int base = (jj / Ns) * Ns * R + (jj % Ns);
#pragma unroll
                for (int r = 0; r < R; r++) lds[Addr::template at<Ns>(p, base + r * Ns)] = v[q * R + r];
            }
        }
    }
    template <int I>
    __device__ __forceinline__ static void exch_read(cplx<T>* v, const cplx<T>* lds, int t, int p) {
        constexpr int Ns = Info::prod_before(I);
        constexpr int R2 = Info::rad[I + 1];
        constexpr int NB2 = N / R2;
        constexpr int B2 = (NB2 + TPP - 1) / TPP;
#pragma unroll
        for (int q = 0; q < B2; q++) {
            const int jj = t + q * TPP;
            if (NB2 % TPP == 0 || jj < NB2) {
#pragma unroll
                for (int r = 0; r < R2; r++) v[q * R2 + r] = lds[Addr::template at<Ns>(p, jj + r * NB2)];
            }
        }
    }

    template <int I>
    __device__ __forceinline__ static void pass(cplx<T>* v, cplx<T>* lds, const cplx<T>* __restrict__ tw, int t, int p,
                                                const TwReg* tr) {
        butterflies<I>(v, tw, t, tr);
        if constexpr (I + 1 < NP) {
            exch_write<I>(v, lds, t, p);
            exch_sync<WAVE>();
            exch_read<I>(v, lds, t, p);
            exch_sync<WAVE>();
            pass<I + 1>(v, lds, tw, t, p, tr);
        }
    }
    // Two independent transforms in step: one barrier per exchange for both,
    // and each one's LDS round trip overlaps the other's butterflies.
    template <int I>
    __device__ __forceinline__ static void pass2(cplx<T>* va, cplx<T>* vb, cplx<T>* la, cplx<T>* lb,
                                                 const cplx<T>* __restrict__ tw, int t, int p, const TwReg* tr) {
        butterflies<I>(va, tw, t, tr);
        butterflies<I>(vb, tw, t, tr);
        if constexpr (I + 1 < NP) {
            exch_write<I>(va, la, t, p);
            exch_write<I>(vb, lb, t, p);
            exch_sync<WAVE>();
            exch_read<I>(va, la, t, p);
            exch_read<I>(vb, lb, t, p);
            exch_sync<WAVE>();
            pass2<I + 1>(va, vb, la, lb, tw, t, p, tr);
        }
    }

    // Full transform: v holds the first-pass input pattern on entry and the
    // last-pass output pattern on exit.  All threads of the block must call.
    __device__ __forceinline__ static void run(cplx<T>* v, cplx<T>* lds, const cplx<T>* __restrict__ tw, int t, int p) {
        pass<0>(v, lds, tw, t, p, nullptr);
    }
    // Two transforms at once; la and lb are disjoint exchange buffers of the block's layout.
    __device__ __forceinline__ static void run2(cplx<T>* va, cplx<T>* vb, cplx<T>* la, cplx<T>* lb,
                                                const cplx<T>* __restrict__ tw, int t, int p) {
        pass2<0>(va, vb, la, lb, tw, t, p, nullptr);
    }
    // run / run2 with the twiddles in registers (load_tw)
    __device__ __forceinline__ static void run_r(cplx<T>* v, cplx<T>* lds, const TwReg& tr, int t, int p) {
        pass<0>(v, lds, nullptr, t, p, &tr);
    }
    __device__ __forceinline__ static void run2_r(cplx<T>* va, cplx<T>* vb, cplx<T>* la, cplx<T>* lb, const TwReg& tr,
                                                  int t, int p) {
        pass2<0>(va, vb, la, lb, nullptr, t, p, &tr);
    }
};

}  // namespace ns3d
