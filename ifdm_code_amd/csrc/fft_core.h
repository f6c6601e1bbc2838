// fft_core.h -- device-side building blocks of the pencil FFT passes for
// gfx950: complex arithmetic, radix-2/4/8 butterflies and a mixed-radix
// Stockham line transform whose first and last stages work on registers.
//
// A line of N points is transformed by P = N/E threads (E = 8 points per
// thread).  Thread t holds the points at positions t + e*P, e = 0..E-1
// ("natural register order") on entry and holds the transform at the same
// positions on exit.  Input and output therefore have the same coalesced
// pattern in global memory (consecutive threads, consecutive points), and a
// forward transform can be followed by an elementwise multiply and an inverse
// transform without any exchange between them.  Intermediate stages exchange
// through LDS (one write + one read per stage boundary).
//
// Stockham autosort, stage s with radix R and stride Ns = product of the
// earlier radices: butterfly j reads x[j + r*N/R], multiplies point r by
// W_N^(r*(j mod Ns)*N/(Ns*R)), takes an R-point DFT and writes
// y[(j/Ns)*Ns*R + (j mod Ns) + r*Ns].
#pragma once
#include <hip/hip_runtime.h>

namespace ifdm {

template <typename T> struct cx;
template <> struct __attribute__((aligned(16))) cx<double> { double x, y; };
template <> struct __attribute__((aligned(8))) cx<float> { float x, y; };

template <typename T> __device__ __forceinline__ cx<T> mkc(T a, T b) { cx<T> r; r.x = a; r.y = b; return r; }
template <typename T> __device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return mkc<T>(a.x + b.x, a.y + b.y); }
template <typename T> __device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return mkc<T>(a.x - b.x, a.y - b.y); }
template <typename T> __device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
    return mkc<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T> __device__ __forceinline__ cx<T> cscale(cx<T> a, T s) { return mkc<T>(a.x * s, a.y * s); }
// a * (DIR * i)
template <int DIR, typename T> __device__ __forceinline__ cx<T> crot(cx<T> a) {
    return DIR < 0 ? mkc<T>(a.y, -a.x) : mkc<T>(-a.y, a.x);
}

// ---------------------------------------------------------------- schedules
// Radix per stage (first stage first); E = points per thread (a stage of
// radix R runs E/R butterflies per thread).  V: the schedule variant, 0
// (radices <= 8) or 1 (radix-16 stages, defined for 1024 points at 16 per
// thread: two LDS exchanges instead of three; other sizes fall back to 0).
template <int N, int E, int V = 0> struct Sched;
#define IFDM_SCHED(NN, EE, VV, NSTG, R0, R1, R2, R3)                       \
    template <> struct Sched<NN, EE, VV> {                                 \
        static constexpr int ns = NSTG;                                    \
        static constexpr int E = EE;                                       \
        static constexpr int r[4] = {R0, R1, R2, R3};                      \
    };
IFDM_SCHED(8, 8, 0, 1, 8, 1, 1, 1)
IFDM_SCHED(16, 8, 0, 2, 8, 2, 1, 1)
IFDM_SCHED(32, 8, 0, 2, 8, 4, 1, 1)
IFDM_SCHED(64, 8, 0, 2, 8, 8, 1, 1)
IFDM_SCHED(128, 8, 0, 3, 8, 8, 2, 1)
IFDM_SCHED(256, 8, 0, 3, 8, 8, 4, 1)
IFDM_SCHED(512, 8, 0, 3, 8, 8, 8, 1)
IFDM_SCHED(1024, 8, 0, 4, 8, 8, 8, 2)
IFDM_SCHED(1024, 16, 0, 4, 8, 8, 8, 2)
IFDM_SCHED(1024, 16, 1, 3, 16, 16, 4, 1)
#undef IFDM_SCHED

// Points per thread: 8 (one 64-lane wavefront per 512-point line), and 16
// for complex floats at N = 1024, where a line then stays one wavefront and
// a strided tile of 16 columns (128-byte rows) fits 1024 threads.  Complex
// doubles keep 8 at 1024: 16 halves their occupancy (measured slower).
constexpr int pts_of(bool fp32, int n) { return (n == 1024 && fp32) ? 16 : 8; }
template <typename T, int N> constexpr int pts() { return pts_of(sizeof(T) == 4, N); }
template <int N, int E, int V> constexpr int sched_v() { return (V == 1 && N == 1024 && E == 16) ? 1 : 0; }
template <typename T, int N, int V = 0> using SchedT = Sched<N, pts<T, N>(), sched_v<N, pts<T, N>(), V>()>;

template <typename T, int N, int V> constexpr int ns_before(int s) {
    int p = 1;
    for (int i = 0; i < s; ++i) p *= SchedT<T, N, V>::r[i];
    return p;
}

// LDS position of point p of a line: one pad slot per 2^SH points.  SH = 4
// where the first stage has radix 16 (its scatter writes 16 consecutive
// points per thread: lanes then start 17 slots apart, which covers the
// banks once per 16 lanes; 18 apart would pair them), else 3.  Variant
// V = 2 keeps the radices of V = 0 with SH = 4 (a shorter line buffer).
template <typename T, int N, int V = 0> constexpr int pad_shift_t() {
    return (V == 2 || SchedT<T, N, V>::r[0] >= 16) ? 4 : 3;
}
template <int SH> __device__ __forceinline__ int lpos(int p) { return p + (p >> SH); }
// LDS line stride in elements of the z passes: congruent to 2 mod 16 so that
// the lines of a wavefront holding several short lines fall on different
// banks.
template <int N, int SH = 3> constexpr int lds_line() { return ((N + (N >> SH) + 15) / 16) * 16 + 2; }
// Line stride of the strided passes' column tiles, whose wavefront lanes run
// over the columns first: congruent to 9 mod 16.  Measured at 512^3 (LDS
// bank-conflict cycles per conflict-free LDS cycle): 1.0 (FP64) and 0.67
// (complex floats) at 2 mod 16, 0.67 and 0.33 at 1 mod 16, 0.33 in both
// at 9 mod 16.  NARROW: tiles of 64-byte rows (4 complex doubles or 8
// complex floats per row, so a wavefront spans 16 or 8 line positions of
// each column), which conflict at 9 mod 16 (measured 1.00 at 1024^3) and
// need a stride congruent to 2 mod 4 in elements: the lanes of one
// ds_read_b128 / ds_read_b64 group then cover every bank once.
template <int N, bool NARROW = false, int SH = 3> constexpr int lds_col_line() {
    return ((N + (N >> SH) + 15) / 16) * 16 + (NARROW ? 2 : 9);
}

// ---------------------------------------------------------------- butterflies
template <int DIR, typename T> __device__ __forceinline__ void bfly2(cx<T> &a0, cx<T> &a1) {
    cx<T> t = a0;
    a0 = cadd(t, a1);
    a1 = csub(t, a1);
}
template <int DIR, typename T> __device__ __forceinline__ void bfly4(cx<T> &a0, cx<T> &a1, cx<T> &a2, cx<T> &a3) {
    cx<T> b0 = cadd(a0, a2), b1 = csub(a0, a2), b2 = cadd(a1, a3), b3 = crot<DIR>(csub(a1, a3));
    a0 = cadd(b0, b2);
    a2 = csub(b0, b2);
    a1 = cadd(b1, b3);
    a3 = csub(b1, b3);
}
// x / sqrt(2); ACC (complex floats): the constant as a sum of two floats,
// since a rounded 1/sqrt(2) alone scales |W8^1|, |W8^3| by 1 - 3.4e-8
template <bool ACC, typename T> __device__ __forceinline__ T mul_rsqrt2(T x) {
    if constexpr (ACC && sizeof(T) == 4) {
        constexpr float hi = 0.707106769084930419921875f;        // float(1/sqrt 2)
        constexpr float lo = 1.21016175e-8f;                     // 1/sqrt 2 - hi
        return fmaf(hi, x, lo * x);
    } else {
        return (T)0.707106781186547524400844362104849 * x;
    }
}
template <int DIR, bool ACC = false, typename T> __device__ __forceinline__ void bfly8(cx<T> *a) {
    bfly4<DIR>(a[0], a[2], a[4], a[6]);
    bfly4<DIR>(a[1], a[3], a[5], a[7]);
    cx<T> o1 = a[3], o3 = a[7];
    a[3] = mkc<T>(mul_rsqrt2<ACC>(o1.x - DIR * o1.y), mul_rsqrt2<ACC>(o1.y + DIR * o1.x));     // * W8^1
    a[5] = crot<DIR>(a[5]);                                                                     // * W8^2
    a[7] = mkc<T>(mul_rsqrt2<ACC>(-o3.x - DIR * o3.y), mul_rsqrt2<ACC>(-o3.y + DIR * o3.x));   // * W8^3
    cx<T> e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
    cx<T> p0 = a[1], p1 = a[3], p2 = a[5], p3 = a[7];
    a[0] = cadd(e0, p0); a[4] = csub(e0, p0);
    a[1] = cadd(e1, p1); a[5] = csub(e1, p1);
    a[2] = cadd(e2, p2); a[6] = csub(e2, p2);
    a[3] = cadd(e3, p3); a[7] = csub(e3, p3);
}
// x * W16^m for the odd m, with W16^m = cos(2 pi m/16) + DIR i sin(2 pi m/16)
// given as (c, s) = (cos, sin) of the angle
template <int DIR, typename T> __device__ __forceinline__ cx<T> rot16(cx<T> x, T c, T s) {
    return mkc<T>(x.x * c - DIR * x.y * s, x.y * c + DIR * x.x * s);
}
// 16-point DFT as 4 x 4 (n = 4 n1 + n2, k = k1 + 4 k2): 4-point DFTs over n1,
// the factors W16^(n2 k1), 4-point DFTs over n2, then the transposition to
// natural order (register renaming).  Rounded constants: the radix is used
// by transforms whose twiddles are rounded to T as well (not ACC).
template <int DIR, bool ACC = false, typename T> __device__ __forceinline__ void bfly16(cx<T> *a) {
    static_assert(!ACC, "radix 16 has no accurate-constant variant");
    constexpr T c1 = (T)0.923879532511286756128183189396788933;   // cos(pi/8)
    constexpr T s1 = (T)0.382683432365089771728459984030398866;   // sin(pi/8)
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) bfly4<DIR>(a[n2], a[4 + n2], a[8 + n2], a[12 + n2]);
    // a[4 k1 + n2] *= W16^(n2 k1)
    a[5] = rot16<DIR>(a[5], c1, s1);                                                                    // m = 1
    a[6] = mkc<T>(mul_rsqrt2<false>(a[6].x - DIR * a[6].y), mul_rsqrt2<false>(a[6].y + DIR * a[6].x));         // m = 2
    a[7] = rot16<DIR>(a[7], s1, c1);                                                                    // m = 3
    a[9] = mkc<T>(mul_rsqrt2<false>(a[9].x - DIR * a[9].y), mul_rsqrt2<false>(a[9].y + DIR * a[9].x));         // m = 2
    a[10] = crot<DIR>(a[10]);                                                                           // m = 4
    a[11] = mkc<T>(mul_rsqrt2<false>(-a[11].x - DIR * a[11].y), mul_rsqrt2<false>(-a[11].y + DIR * a[11].x));  // m = 6
    a[13] = rot16<DIR>(a[13], s1, c1);                                                                  // m = 3
    a[14] = mkc<T>(mul_rsqrt2<false>(-a[14].x - DIR * a[14].y), mul_rsqrt2<false>(-a[14].y + DIR * a[14].x));  // m = 6
    a[15] = rot16<DIR>(a[15], -c1, -s1);                                                                // m = 9
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) bfly4<DIR>(a[4 * k1], a[4 * k1 + 1], a[4 * k1 + 2], a[4 * k1 + 3]);
    cx<T> t[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) t[i] = a[i];
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) a[k1 + 4 * k2] = t[4 * k1 + k2];
}
template <int R, int DIR, bool ACC = false, typename T> __device__ __forceinline__ void dft(cx<T> *a) {
    if constexpr (R == 2) bfly2<DIR>(a[0], a[1]);
    else if constexpr (R == 4) bfly4<DIR>(a[0], a[1], a[2], a[3]);
    else if constexpr (R == 8) bfly8<DIR, ACC>(a);
    else if constexpr (R == 16) bfly16<DIR, ACC>(a);
}

// ---------------------------------------------------------------- line FFT
// Base twiddles of one thread: for stage S >= 1 and its butterfly q
// (j = t + q*P) the factor W_N^((j mod Ns) * N/(Ns*R)); the powers
// W^r, r < R, are formed by complex products of depth <= 3 (r = 2: w^2;
// 3: w^2 w; 4: (w^2)^2; 5: w^4 w; 6: w^4 w^2; 7: w^4 w^3).  The base
// values are loaded once per thread at the start of a pass, so their
// latency hides behind the data loads instead of stalling each stage.
// W: the type of the base values (FP64 for complex floats where the
// twiddles must be accurate, see apply_twiddles); V: the schedule variant
// (Sched), which the transforms taking these bases follow
template <typename T, int N, typename W = T, int V = 0> struct TwBase {
    static constexpr int E = SchedT<T, N, V>::E;
    cx<W> w[4][E];  // [stage][butterfly]; stage 0 unused
};

// tw is the table W_M^k, k < M, of a transform of length M = TS * N
// (TS = 1: the line's own table; TS = 2: a half-length transform reading the
// full-length table)
template <typename T, int N, int TS = 1, typename W, int V>
__device__ __forceinline__ void load_twiddles(TwBase<T, N, W, V> &b, int t, const cx<W> *__restrict__ tw) {
    using SC = SchedT<T, N, V>;
    constexpr int E = SC::E, P = N / E;
#pragma unroll
    for (int S = 1; S < SC::ns; ++S) {
        const int R = SC::r[S], Q = E / R, NS = ns_before<T, N, V>(S);
#pragma unroll
        for (int q = 0; q < Q; ++q) b.w[S][q] = tw[TS * (((t + q * P) & (NS - 1)) * (N / (NS * R)))];
        // radix 16 (one butterfly per thread): w^4 beside w, so every power
        // up to w^15 is a product of depth <= 3 (apply_twiddles16)
        if (R == 16) b.w[S][1] = tw[TS * 4 * ((t & (NS - 1)) * (N / (NS * R)))];
    }
}

// Complex floats, ACC: the powers of the FP64 base twiddle are formed in
// FP64 and each product a * w is taken in FP64 and rounded once.  Rounded
// FP32 twiddles carry a fixed error per mode that the kick-drift sequence
// compounds step after step (a 7e-4 density error after 1000 steps at
// 512^3, against 2e-6 from FP32 storage alone).
__device__ __forceinline__ cx<float> twmul64(cx<float> a, cx<double> w) {
    const double ax = a.x, ay = a.y;
    return mkc<float>((float)(ax * w.x - ay * w.y), (float)(ax * w.y + ay * w.x));
}
template <int R, int DIR, bool ACC>
__device__ __forceinline__ void apply_twiddles(cx<float> *a, cx<float> w1) {
    using T = float;
    if constexpr (DIR > 0) w1.y = -w1.y;
    a[1] = cmul(a[1], w1);
    if constexpr (R >= 4) {
        const cx<T> w2 = cmul(w1, w1);
        a[2] = cmul(a[2], w2);
        a[3] = cmul(a[3], cmul(w2, w1));
        if constexpr (R == 8) {
            const cx<T> w4 = cmul(w2, w2), w3 = cmul(w2, w1);
            a[4] = cmul(a[4], w4);
            a[5] = cmul(a[5], cmul(w4, w1));
            a[6] = cmul(a[6], cmul(w4, w2));
            a[7] = cmul(a[7], cmul(w4, w3));
        }
    }
}

template <int R, int DIR, bool ACC>
__device__ __forceinline__ void apply_twiddles(cx<float> *a, cx<double> w) {
    if constexpr (!ACC) {
        apply_twiddles<R, DIR, false>(a, mkc<float>((float)w.x, (float)w.y));
    } else {
        if constexpr (DIR > 0) w.y = -w.y;
        a[1] = twmul64(a[1], w);
        if constexpr (R >= 4) {
            const cx<double> w2 = cmul(w, w);
            a[2] = twmul64(a[2], w2);
            a[3] = twmul64(a[3], cmul(w2, w));
            if constexpr (R == 8) {
                const cx<double> w4 = cmul(w2, w2), w3 = cmul(w2, w);
                a[4] = twmul64(a[4], w4);
                a[5] = twmul64(a[5], cmul(w4, w));
                a[6] = twmul64(a[6], cmul(w4, w2));
                a[7] = twmul64(a[7], cmul(w4, w3));
            }
        }
    }
}

template <int R, int DIR, bool ACC>
__device__ __forceinline__ void apply_twiddles(cx<double> *a, cx<double> w1) {
    using T = double;
    if constexpr (DIR > 0) w1.y = -w1.y;
    a[1] = cmul(a[1], w1);
    if constexpr (R >= 4) {
        const cx<T> w2 = cmul(w1, w1);
        a[2] = cmul(a[2], w2);
        a[3] = cmul(a[3], cmul(w2, w1));
        if constexpr (R == 8) {
            const cx<T> w4 = cmul(w2, w2), w3 = cmul(w2, w1);
            a[4] = cmul(a[4], w4);
            a[5] = cmul(a[5], cmul(w4, w1));
            a[6] = cmul(a[6], cmul(w4, w2));
            a[7] = cmul(a[7], cmul(w4, w3));
        }
    }
}

// Radix-16 stage twiddles from the base w and v = w^4 (both from the table,
// conjugated for the inverse): w^2, w^3 from w; w^5..w^7 as v w^(1..3);
// w^8 = v^2; w^9..w^11 as w^8 w^(1..3); w^12 = w^8 v; w^13..w^15 as
// w^12 w^(1..3).
template <int DIR, typename T>
__device__ __forceinline__ void apply_twiddles16(cx<T> *a, cx<T> w, cx<T> v) {
    if constexpr (DIR > 0) {
        w.y = -w.y;
        v.y = -v.y;
    }
    const cx<T> w2 = cmul(w, w), w3 = cmul(w2, w), w8 = cmul(v, v), w12 = cmul(w8, v);
    a[1] = cmul(a[1], w);
    a[2] = cmul(a[2], w2);
    a[3] = cmul(a[3], w3);
    a[4] = cmul(a[4], v);
    a[5] = cmul(a[5], cmul(v, w));
    a[6] = cmul(a[6], cmul(v, w2));
    a[7] = cmul(a[7], cmul(v, w3));
    a[8] = cmul(a[8], w8);
    a[9] = cmul(a[9], cmul(w8, w));
    a[10] = cmul(a[10], cmul(w8, w2));
    a[11] = cmul(a[11], cmul(w8, w3));
    a[12] = cmul(a[12], w12);
    a[13] = cmul(a[13], cmul(w12, w));
    a[14] = cmul(a[14], cmul(w12, w2));
    a[15] = cmul(a[15], cmul(w12, w3));
}

// Ordering of the LDS exchanges of one line.  WAVE: every point of the line
// is held by one wavefront, whose LDS operations are performed in issue
// order, so only the compiler must be kept from moving LDS accesses across
// the exchange (no s_barrier; the wavefronts of a work-group run
// independently).  Otherwise the line spans wavefronts and the work-group
// synchronises.
template <bool WAVE> __device__ __forceinline__ void line_sync() {
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// One Stockham stage S of an N-point transform, direction DIR (-1 forward,
// +1 inverse), for the thread at line position t.  L is the line's LDS
// region; b the thread's base twiddles.  FRESH: the work-group has not yet
// used its LDS, so the first scatter needs no preceding synchronisation.
template <typename T, int N, int DIR, int S, bool WAVE = false, bool FRESH = false, bool ACC = false, typename W, int V>
__device__ __forceinline__ void fft_stage(cx<T> *v, cx<T> *L, int t, const TwBase<T, N, W, V> &b) {
    using SC = SchedT<T, N, V>;
    constexpr int E = SC::E, P = N / E, R = SC::r[S], Q = E / R;
    constexpr int NS = ns_before<T, N, V>(S), SH = pad_shift_t<T, N, V>();
    constexpr bool FIRST = (S == 0), LAST = (S == SC::ns - 1);
    static_assert(R < 16 || (Q == 1 && (FIRST || sizeof(W) == sizeof(T))), "radix 16: one butterfly per thread, twiddles in T");
    cx<T> a[Q][R];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int j = t + q * P;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if constexpr (FIRST) a[q][r] = v[q + r * Q];
            else a[q][r] = L[lpos<SH>(j + r * (N / R))];
        }
        if constexpr (!FIRST) {
            if constexpr (R == 16) apply_twiddles16<DIR>(a[q], b.w[S][0], b.w[S][1]);
            else apply_twiddles<R, DIR, ACC>(a[q], b.w[S][q]);
        }
        dft<R, DIR, ACC>(a[q]);
    }
    if constexpr (LAST) {
#pragma unroll
        for (int q = 0; q < Q; ++q)
#pragma unroll
            for (int r = 0; r < R; ++r) v[q + r * Q] = a[q][r];
    } else {
        if constexpr (!(FRESH && FIRST)) line_sync<WAVE>();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int j = t + q * P;
            const int base = (j / NS) * NS * R + (j & (NS - 1));
#pragma unroll
            for (int r = 0; r < R; ++r) L[lpos<SH>(base + r * NS)] = a[q][r];
        }
        line_sync<WAVE>();
    }
}

// ACC: accurate twiddles for complex floats (see apply_twiddles)
template <typename T, int N, int DIR, bool WAVE = false, bool FRESH = false, bool ACC = false, typename W, int V>
__device__ __forceinline__ void fft_line(cx<T> *v, cx<T> *L, int t, const TwBase<T, N, W, V> &b0) {
    using SC = SchedT<T, N, V>;
    // FP64 bases of complex-float transforms pass an empty asm per
    // transform, so the compiler forms their powers again instead of keeping
    // a pass's earlier powers live in registers across its other transforms
    TwBase<T, N, W, V> b = b0;
    if constexpr (sizeof(T) == 4 && sizeof(W) == 8) {
#pragma unroll
        for (int S = 1; S < SC::ns; ++S)
#pragma unroll
            for (int q = 0; q < SC::E / SC::r[S]; ++q)
                asm volatile("" : "+v"(b.w[S][q].x), "+v"(b.w[S][q].y));
    }
    fft_stage<T, N, DIR, 0, WAVE, FRESH, ACC>(v, L, t, b);
    if constexpr (SC::ns > 1) fft_stage<T, N, DIR, 1, WAVE, false, ACC>(v, L, t, b);
    if constexpr (SC::ns > 2) fft_stage<T, N, DIR, 2, WAVE, false, ACC>(v, L, t, b);
    if constexpr (SC::ns > 3) fft_stage<T, N, DIR, 3, WAVE, false, ACC>(v, L, t, b);
}

// convenience: load the base twiddles and transform
template <typename T, int N, int DIR, bool WAVE = false>
__device__ __forceinline__ void fft_line(cx<T> *v, cx<T> *L, int t, const cx<T> *__restrict__ tw) {
    TwBase<T, N> b;
    load_twiddles<T, N>(b, t, tw);
    fft_line<T, N, DIR, WAVE>(v, L, t, b);
}

// true when a line of N points (N/E consecutive threads) lies inside one
// wavefront of 64 lanes
template <typename T, int N> constexpr bool wave_line() { return N / SchedT<T, N>::E <= 64; }

}  // namespace ifdm
