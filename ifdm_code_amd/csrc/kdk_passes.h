// kdk_passes.h -- the fused pencil passes of one kick-drift-kick step.
//
// Fields in HBM (x, y, z with z fastest, as the reference's opr.wfc):
//   psi   N^3 complex, x-planes of N*N cells ps cells apart (ps = N*N + pad)
//   phi   N^3 real, the same plane stride
//   shat  half spectrum of the Poisson source: rows of HS cells (N/2+1
//         valid, HS a multiple of 16, so rows start on 128-byte
//         boundaries), N rows per x-plane, planes sps cells apart
//
// One step is a drift table and eight passes, each one read and one write
// of a field (the drift factor separates by axis, so each axis makes its
// own forward - factor - inverse round trip):
//   T      drift table E[i] = exp(-i c k_i^2 dKE) / N
//   A  y:  kick (phase from phi and |psi|^2), forward, E[k_y]/N, inverse  psi
//   B  x:  forward, E[k_x]/N, inverse                                      psi
//   C  z:  forward, E[k_z]/N, inverse -> psi; m_s|psi|^2 - rho_c -> real
//          FFT -> shat; block max |psi|^2 and sum |psi|^2
//   F  y:  forward FFT                                                     shat
//   G  x:  forward FFT, Green -4 pi G / k^2 / N^3 (k = 0 -> 0), inverse    shat
//   H  y:  inverse FFT                                                     shat
//   I  z:  complex-to-real inverse FFT -> phi; block max |phi|
// The reference's order is src/CosmoBEC_MPI_stan.cpp:130-152 with
// poisson() at :80-119; its drift multiplies by fftnorm at :141 and its
// Poisson solve by fftnorm after the c2r at :116.
#pragma once
#include <type_traits>

#include "criteria.h"
#include "fastmath.h"
#include "fft_core.h"

namespace ifdm {

// half-spectrum row length: N/2+1 rounded up to 16 elements, so every row
// starts on a 128-byte boundary in both precisions (8 complex doubles or 16
// complex floats per 128 bytes) and a column tile never straddles lines
__host__ __device__ constexpr int hs_of(int n) { return ((n / 2 + 1 + 15) / 16) * 16; }

struct ZArgs {
    double aphi;      // kick: phase coefficient of phi
    double arho;      // kick: phase coefficient of |psi|^2
    double ms;        // Poisson source: m_s (times 2^(2s) for a scaled psi)
    double rhoc;      // Poisson source: subtracted density (rho_c; the FP64
                      // mean m_s <|psi|^2> in mixed mode)
    int ref_range;    // max|phi| over the reference's one-rank scan range
    int blk0;         // first block of this launch
    void *dtab;       // drift table E[k_z] (the z drift of pass C)
    int rev;          // blocks walk the lines in descending order
    long ps;          // plane stride (elements) of psi and phi: N*N + padding
    long sps;         // plane stride (elements) of shat: N*HS + padding
    const double *drhoc;  // when set: rhoc from device memory (the on-device step loop)
};

struct SArgs {
    int other0;        // first line index of the fixed axis in this launch
    const void *dtab;  // drift table E[i] (passes A, B)
    double ke_coef;    // 0.5 hbar / (m / c^2)   (src/struct3D.h:121)
    double dKE;        // time integral of the drift
    double green;      // -4 pi G
    double fftnorm;    // 1 / N^3
    double kunit2;     // (2 pi / L)^2: k_i^2 = kunit2 m_i^2, m_i the signed mode index
    int rev;           // blocks walk the fixed axis in descending order
    // the kick of S_KICK_DRIFT1: phase aphi phi + arho |psi|^2, phi a real
    // field of psi's shape
    double aphi;
    double arho;
    const void *phi;
    const double *dk;  // when set: aphi, arho from device memory (the on-device step loop)
};

// Block order.  A pass that walks its lines opposite to the pass before it
// starts on the planes that pass wrote last, which may still sit in the
// memory-side cache.
__device__ __forceinline__ int block_order(int rev, unsigned b, unsigned nb) {
    return rev ? (int)(nb - 1 - b) : (int)b;
}

// Z_DRIFT: forward, E[k_z] / N, inverse; Z_SRC: the Poisson source of psi
// (m_s |psi|^2 - rho_c), its real-to-complex transform and the density
// reductions; Z_DRIFT_SRC: Z_DRIFT, then Z_SRC on the result; Z_RSRC: the
// real-to-complex transform of a given real field; Z_C2R: complex-to-real
// (small N)
enum ZMode { Z_FWD = 0, Z_INV = 1, Z_SRC = 2, Z_C2R = 3, Z_RSRC = 4, Z_DRIFT = 5, Z_DRIFT_SRC = 6 };
// S_DRIFT1: forward, the line axis's drift factor E[i] / N, inverse (the
// drift factorises by axis, so each axis runs its own round trip);
// S_KICK_DRIFT1: the kick exp(-i (aphi phi + arho |psi|^2)) first;
// S_GREEN: forward, Green's function, inverse
enum SMode { S_FWD = 0, S_INV = 1, S_GREEN = 2, S_DRIFT1 = 3, S_KICK_DRIFT1 = 4 };

// Offset of z-line l = x N + y in a field of plane stride ps (elements):
// psi and phi keep each x-plane's N*N cells contiguous and start planes ps
// elements apart.  A plane stride of N*N (a power of two) maps the rows of
// an x-line to the same HBM channels; a small pad spreads them.
template <int N> __device__ __forceinline__ long zline_off(long l, long ps) {
    return (l / N) * ps + (l % N) * (long)N;
}

// Offset of half-spectrum row l = x N + y (rows of HS cells, planes sps apart)
template <int N, int HS> __device__ __forceinline__ long hline_off(long l, long sps) {
    return (l / N) * sps + (l % N) * (long)HS;
}

template <typename T, int N> constexpr int z_lines() {
    constexpr int P = N / SchedT<T, N>::E;
    return (256 / P) < (N * N) ? (256 / P) : (N * N);
}
// N = 1024 strided passes: 64-byte row segments ("narrow" tiles), so a
// tile's exchange buffer takes 74 KB of LDS and two work-groups share a CU,
// one loading while the other transforms (a 128-byte tile of 1024-point
// lines needs 149 KB: one work-group per CU, whose loads and transforms
// never overlap).  The two tiles of a 128-byte line run on one XCD at about
// the same time (xcd_group_rows), so the second finds the line in L2.
// Measured at 1024^3 with the grouping (without): A 8.03 -> 7.59 (8.51) ms
// mixed, 13.09 -> 12.34 (14.07) ms FP64; the half-spectrum y passes F, H
// 1.88, 1.91 (1.97, 2.01) ms mixed; in FP64 they keep 128-byte rows (3.57,
// 3.46 against 3.64, 3.65 ms narrow).  The x passes B and G take narrow
// tiles too, grouped the same way (B 5.40 -> 4.98 ms mixed, 10.3 -> 9.6 ms
// FP64; G 2.98 -> 2.63 ms mixed, 5.29 -> 4.41 ms FP64, where 5.67 ms either
// way while its Green factor loaded k^2 from a table); without the grouping
// and the addressing of the current build they had measured slower (B 5.8
// -> 6.9 ms mixed).  At N = 512, where two 128-byte tiles
// already share a CU, narrow x tiles are slower (FP64 B 0.83 -> 1.05 ms,
// mixed B 0.44 -> 0.46, G 0.248 -> 0.256).
template <typename T, int N, bool WIDE = false, int AXIS = 1, int MODE = -1> constexpr bool s_narrow() {
    return N == 1024 && !WIDE && ((AXIS == 1 && (sizeof(T) == 4 || MODE == S_KICK_DRIFT1 || MODE == S_DRIFT1)) ||
                                  (AXIS == 0 && (MODE == S_DRIFT1 || MODE == S_GREEN)));
}

// Schedule variant of a strided pass's transforms (Sched): radix-16 stages
// where they exist (1024 points at 16 per thread), except in the kick pass.
// Measured at 1024^3 mixed (ms, radix <= 8 -> radix 16): B 5.76 -> 5.30,
// G 3.13 -> 2.84, F 1.86 -> 1.85, H 1.91 -> 1.88; the kick pass A, capped
// at 128 VGPRs, 7.43 -> 7.72, so it keeps radices <= 8 (as do the z passes,
// whose pass C took 205 VGPRs with radix 16: 4.84 -> 5.09).
// The FP64 drift passes at N = 1024 take variant 2 (radices <= 8, one LDS
// pad slot per 16 points): their narrow tiles' exchange buffer shrinks from
// 74 to 70 KB, so the drift table fits beside it with two work-groups per CU.
template <typename T, int N, int MODE> constexpr int sp_sched() {
    return (sizeof(T) == 8 && N == 1024 && (MODE == S_DRIFT1 || MODE == S_KICK_DRIFT1)) ? 2
                                                                                      : (MODE == S_KICK_DRIFT1 ? 0 : 1);
}

// z-columns per strided-pass tile: 128-byte row segments (8 complex
// doubles, 16 complex floats; 64-byte ones for narrow tiles); WIDE doubles
// them for FP64 (256-byte rows, used by the 3-D transform's x pass), and a
// work-group never exceeds 1024 threads.
template <typename T, int N, bool WIDE = false, int MODE = -1, int AXIS = 1> constexpr int s_cols() {
    constexpr bool narrow = s_narrow<T, N, WIDE, AXIS, MODE>();
    constexpr int base = narrow ? 64 / (int)sizeof(cx<T>) : 128 / (int)sizeof(cx<T>);
    constexpr int want = (WIDE && sizeof(cx<T>) == 16) ? 2 * base : base;
    constexpr int cap = 1024 / (N / SchedT<T, N>::E);
    // and the exchange buffer within the 160 KB of LDS
    constexpr int lcap = 163840 / (lds_col_line<N, narrow, pad_shift_t<T, N, sp_sched<T, N, MODE>()>()>() * (int)sizeof(cx<T>));
    int c = want < cap ? want : cap;
    while (c > lcap) c /= 2;
    return c;
}

// exp(-i th) in the storage precision T: the FP64 argument is reduced mod
// pi/2 in FP64, then sin/cos are evaluated in T (fastmath.h)
template <typename T> __device__ __forceinline__ cx<T> expmi(double th) {
    cx<T> r;
    expmi_parts<T>(th, &r.x, &r.y);
    return r;
}

// The drift phase exp(-i KE(k) dKE), KE = ke_coef (kx^2 + ky^2 + kz^2), is a
// product of one factor per axis.  Pass T writes the per-step table
// E[i] = exp(-i ke_coef k_i^2 dKE) / N, i < N, which each axis's round trip
// multiplies in (the 1/N of its inverse transform folded in: a power of two,
// so the products are the same bits as scaling afterwards).
template <typename T, int N>
__device__ __forceinline__ void write_drift_table(cx<T> *tab, const double *k2, double ke_dke, int tid, int nt) {
    for (int i = tid; i < N; i += nt) tab[i] = cscale(expmi<T>((ke_dke * k2[i])), (T)(1.0 / N));
}

template <typename T, int N>
__global__ __launch_bounds__(256) void k_drift_table(cx<T> *__restrict__ tab, const double *__restrict__ k2,
                                                     double ke_dke) {
    write_drift_table<T, N>(tab, k2, ke_dke, threadIdx.x, 256);
}

template <bool MAX> __device__ __forceinline__ double red_op(double a, double b) { return MAX ? fmax(a, b) : a + b; }

// block reduction: butterfly within each 64-wide wavefront, then one LDS
// exchange across the block's wavefronts (sbuf: >= 16 doubles)
template <bool MAX>
__device__ __forceinline__ double block_reduce(double v, double *sbuf, int tid, int nthreads) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = red_op<MAX>(v, __shfl_xor(v, o, 64));
    const int nw = (nthreads + 63) / 64;
    __syncthreads();
    if ((tid & 63) == 0) sbuf[tid >> 6] = v;
    __syncthreads();
    double r = sbuf[0];
    for (int w = 1; w < nw; ++w) r = red_op<MAX>(r, sbuf[w]);
    return r;
}

// Complex floats: the drift's round trips on psi (passes A, B and the z
// round trip of C; also the 3-D transforms of psi) take accurate twiddles
// (FP64 base values, powers and products); the Poisson transforms keep FP32
// twiddles.  Measured at 512^3 (density error after 1000 steps against the
// FP64 path): FP32 twiddles everywhere 1.0e-3, accurate twiddles in the
// Poisson transforms only 1.0e-3, in the psi round trips only 1.1e-5, in
// both 1.1e-5.  The psi round trips repeat the same transform pair every
// step on the field that persists; an error in the potential does not
// carry over to the next step's.  Up to N = 512: the 16-point-per-thread
// transforms of N = 1024 have no register room for FP64 twiddles (measured
// 31 -> 25 steps/s at 1024^3), so they keep FP32 twiddles (the first 100
// steps then stay near 1e-5, the error of FP32 twiddles at 512^3).
template <typename T, int N> constexpr bool acc_tw() { return sizeof(T) == 4 && N <= 512; }
template <typename T, int N, bool PSI>
using tw_base_t = typename std::conditional<(acc_tw<T, N>() && PSI), double, T>::type;

// Stores of a pass's output field: non-temporal.  Each pass writes a field
// far larger than the caches that no kernel reads again before the next
// pass sweeps it, so the writes need not be kept in L2.  Measured at 512^3
// against plain stores: FP64 196.9 -> 203.1 steps/s (C 1.16 -> 1.11 ms,
// F, G, H 6-8 % faster), mixed 364.9 -> 370.7; pass I, whose phi the next
// step's kick pass reads, keeps plain stores (non-temporal there measured
// no faster).
template <typename T> struct vec2_of;
template <> struct vec2_of<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct vec2_of<float> { typedef float type __attribute__((ext_vector_type(2))); };
template <typename T> __device__ __forceinline__ void stream_store(cx<T> *p, cx<T> v) {
    typedef typename vec2_of<T>::type V;
    V w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<V *>(p));
}
template <typename T> __device__ __forceinline__ cx<T> stream_load(const cx<T> *p) {
    typedef typename vec2_of<T>::type V;
    const V w = __builtin_nontemporal_load(reinterpret_cast<const V *>(p));
    return mkc<T>(w.x, w.y);
}

// ------------------------------------------------------------------ z passes
// Minimum waves per SIMD asked of the compiler for a z pass: 3 for the FP64
// pass C at N = 512 (176 -> 168 VGPRs, 20 bytes of spill; measured 1.09 ->
// 1.03 ms; 4 waves spill 116 bytes, 1.32 ms).  At N = 1024 the same cap
// spills 52 bytes and is slower (10.1 -> 10.9 ms).
template <typename T, int N, int MODE> constexpr int zp_wpe() {
    return (sizeof(T) == 8 && N == 512 && MODE == Z_DRIFT_SRC) ? 3 : 1;
}

template <typename T, int N, int MODE>
__global__ __launch_bounds__((z_lines<T, N>() * (N / SchedT<T, N>::E)), (zp_wpe<T, N, MODE>()))
void k_zpass(cx<T> *__restrict__ psi, T *__restrict__ phi, cx<T> *__restrict__ shat,
             const cx<T> *__restrict__ tw, const cx<double> *__restrict__ tw64, ZArgs za, double *__restrict__ part) {
    constexpr int E = SchedT<T, N>::E, P = N / E, LZ = z_lines<T, N>(), LS = lds_line<N, pad_shift_t<T, N>()>();
    constexpr int HS = hs_of(N), NT = LZ * P;
    __shared__ cx<T> lds[LZ * LS];
    const int tid = threadIdx.x, l = tid / P, t = tid - l * P;
    const int blk = za.blk0 + block_order(za.rev, blockIdx.x, gridDim.x);
    const long line = (long)blk * LZ + l;
    const bool valid = line < (long)N * N;
    cx<T> *L = lds + l * LS;
    cx<T> v[E];
    constexpr bool PSI = MODE == Z_DRIFT || MODE == Z_DRIFT_SRC || MODE == Z_FWD || MODE == Z_INV;
    TwBase<T, N, tw_base_t<T, N, PSI>> tb;
    if constexpr (acc_tw<T, N>() && PSI) load_twiddles<T, N>(tb, t, tw64);
    else load_twiddles<T, N>(tb, t, tw);

    if constexpr (MODE == Z_C2R) {
        const cx<T> *row = shat + hline_off<N, HS>(line, za.sps);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int p = t + e * P;
            cx<T> s = mkc<T>(0, 0);
            if (valid) {
                if (p <= N / 2) s = row[p];
                else { s = row[N - p]; s.y = -s.y; }
            }
            v[e] = s;
        }
        fft_line<T, N, +1, wave_line<T, N>(), false, false>(v, L, t, tb);
        double mx = 0.0;
        T *out = phi + zline_off<N>(line, za.ps);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int p = t + e * P;
            if (valid) {
                out[p] = v[e].x;
                const bool in = !za.ref_range || (line * (N + 2) + p < (long)N * N * N);
                if (in) mx = fmax(mx, fabs((double)v[e].x));
            }
        }
        double *sb = reinterpret_cast<double *>(lds);
        const double bm = block_reduce<true>(mx, sb, tid, NT);
        if (tid == 0) part[blk] = bm;
        return;
    } else if constexpr (MODE == Z_RSRC) {
        // Poisson source from a real field r (read through `phi`): ms * r
        const T *rrow = phi + zline_off<N>(line, za.ps);
        cx<T> s[E];
#pragma unroll
        for (int e = 0; e < E; ++e) s[e] = mkc<T>(valid ? (T)(za.ms * (double)rrow[t + e * P]) : (T)0, (T)0);
        fft_line<T, N, -1, wave_line<T, N>(), false, false>(s, L, t, tb);
        if (valid) {
            cx<T> *srow = shat + hline_off<N, HS>(line, za.sps);
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int p = t + e * P;
                if (p <= N / 2) stream_store(&srow[p], s[e]);
            }
        }
        return;
    } else {
        // psi is read once per pass: non-temporal loads (measured at 512^3:
        // +0.5 % FP64, +0.9 % mixed, mostly in pass I, whose input then
        // stays in the memory-side cache)
        cx<T> *row = psi + zline_off<N>(line, za.ps);
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = valid ? stream_load(&row[t + e * P]) : mkc<T>(0, 0);

        // the z drift's table in LDS, filled once per work-group: from global
        // memory its loads, issued after the forward transform, cost L2
        // round trips in the middle of the line (measured at 512^3: FP64
        // pass C 1.02 -> 0.98 ms, mixed 0.545 -> 0.540)
        constexpr bool ZT = (MODE == Z_DRIFT || MODE == Z_DRIFT_SRC) && N >= 256;
        __shared__ cx<T> ztab[ZT ? N : 1];
        if constexpr (ZT) {
            for (int i = tid; i < N; i += NT) ztab[i] = ((const cx<T> *)za.dtab)[i];
            __syncthreads();
        }

        if constexpr (MODE == Z_DRIFT || MODE == Z_DRIFT_SRC) {
            // the z factor of the drift: forward, E[k_z] / N, inverse
            const cx<T> *tab = ZT ? (const cx<T> *)ztab : (const cx<T> *)za.dtab;
            fft_line<T, N, -1, wave_line<T, N>(), false, acc_tw<T, N>()>(v, L, t, tb);
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = cmul(v[e], tab[t + e * P]);
            fft_line<T, N, +1, wave_line<T, N>(), false, acc_tw<T, N>()>(v, L, t, tb);
            if (valid) {
#pragma unroll
                for (int e = 0; e < E; ++e) stream_store(&row[t + e * P], v[e]);
            }
            if constexpr (MODE == Z_DRIFT) return;
        }
        if constexpr (MODE == Z_FWD) {
            fft_line<T, N, -1, wave_line<T, N>(), false, acc_tw<T, N>()>(v, L, t, tb);
            if (valid) {
#pragma unroll
                for (int e = 0; e < E; ++e) stream_store(&row[t + e * P], v[e]);
            }
            return;
        }
        if constexpr (MODE == Z_INV) {
            fft_line<T, N, +1, wave_line<T, N>(), false, acc_tw<T, N>()>(v, L, t, tb);
            if (valid) {
#pragma unroll
                for (int e = 0; e < E; ++e) stream_store(&row[t + e * P], v[e]);
            }
        }
        if constexpr (MODE == Z_SRC || MODE == Z_DRIFT_SRC) {
            // the source replaces psi in the registers (psi is stored already)
            const double rhoc = za.drhoc ? *za.drhoc : za.rhoc;
            double mx = 0.0, sm = 0.0;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const double rr = (double)v[e].x * v[e].x + (double)v[e].y * v[e].y;
                if (valid) { mx = fmax(mx, rr); sm += rr; }
                v[e] = mkc<T>(valid ? (T)(za.ms * rr - rhoc) : (T)0, (T)0);
            }
            fft_line<T, N, -1, wave_line<T, N>(), false, false>(v, L, t, tb);
            if (valid) {
                cx<T> *srow = shat + hline_off<N, HS>(line, za.sps);
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const int p = t + e * P;
                    if (p <= N / 2) stream_store(&srow[p], v[e]);
                }
            }
            double *sb = reinterpret_cast<double *>(lds);
            const double bm = block_reduce<true>(mx, sb, tid, NT);
            const double bs = block_reduce<false>(sm, sb, tid, NT);
            if (tid == 0) { part[2 * blk] = bm; part[2 * blk + 1] = bs; }
        }
    }
}

// ------------------------------------------------------------- pass I (c2r)
// Complex-to-real along z with a half-length transform: from the Hermitian
// half line X[k], k <= N/2 (Re X[0], Re X[N/2] taken), form for k < N/2
//   Z[k] = (X[k] + conj X[N/2-k]) + i W^-k (X[k] - conj X[N/2-k]),
// then z = IDFT_{N/2}(Z) (unnormalised) gives y[2m] = Re z[m],
// y[2m+1] = Im z[m] with y the unnormalised length-N inverse of X.
// A block has the thread count of a full-length z-pass block, so it holds
// (N/E_N) / (N/2 / E_{N/2}) times as many (half-length) lines.
template <typename T, int N> constexpr int c2r_lines() {
    return z_lines<T, N>() * (N / SchedT<T, N>::E) / ((N / 2) / SchedT<T, N / 2>::E);
}

template <typename T, int N>
__global__ __launch_bounds__((z_lines<T, N>() * (N / SchedT<T, N>::E)))
void k_c2r_half(const cx<T> *__restrict__ shat, T *__restrict__ phi, const cx<T> *__restrict__ tw, ZArgs za,
                double *__restrict__ part) {
    constexpr int H = N / 2, E = SchedT<T, H>::E, P = H / E, LZ = c2r_lines<T, N>(), LS = lds_line<H, pad_shift_t<T, H>()>();
    constexpr int HS = hs_of(N), NT = LZ * P;
    static_assert(NT == z_lines<T, N>() * (N / SchedT<T, N>::E), "block shape");
    __shared__ cx<T> lds[(LZ * LS > NT / 2 + 1) ? LZ * LS : NT / 2 + 1];
    const int tid = threadIdx.x, l = tid / P, t = tid - l * P;
    const int blk = za.blk0 + block_order(za.rev, blockIdx.x, gridDim.x);
    const long line = (long)blk * LZ + l;
    const bool valid = line < (long)N * N;
    cx<T> *L = lds + l * LS;
    TwBase<T, H> tb;
    load_twiddles<T, H, 2>(tb, t, tw);
    const cx<T> *row = shat + hline_off<N, HS>(line, za.sps);
    const cx<T> wt = tw[t];                                    // W^t; W^(t+eP) = W^t W^(eP)
    cx<T> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int k = t + e * P;
        cx<T> a = mkc<T>(0, 0), b = mkc<T>(0, 0);
        if (valid) {
            a = row[k];
            b = row[H - k];
        }
        if (k == 0) { a.y = 0; b.y = 0; }                      // Re X[0], Re X[N/2]
        const cx<T> sum = mkc<T>(a.x + b.x, a.y - b.y);        // X[k] + conj X[N/2-k]
        cx<T> dif = mkc<T>(a.x - b.x, a.y + b.y);              // X[k] - conj X[N/2-k]
        cx<T> w = e == 0 ? wt : cmul(wt, tw[e * P]);         // W^k
        w.y = -w.y;                                            // W^-k
        dif = cmul(dif, w);
        v[e] = mkc<T>(sum.x - dif.y, sum.y + dif.x);           // sum + i dif
    }
    fft_line<T, H, +1, wave_line<T, H>(), false, false>(v, L, t, tb);
    double mx = 0.0;
    cx<T> *out = reinterpret_cast<cx<T> *>(phi + zline_off<N>(line, za.ps));
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int m = t + e * P;
        if (valid) {
            out[m] = v[e];                                      // y[2m], y[2m+1]
            const long f0 = line * (N + 2) + 2 * m;
            const bool in0 = !za.ref_range || f0 < (long)N * N * N;
            const bool in1 = !za.ref_range || f0 + 1 < (long)N * N * N;
            if (in0) mx = fmax(mx, fabs((double)v[e].x));
            if (in1) mx = fmax(mx, fabs((double)v[e].y));
        }
    }
    double *sb = reinterpret_cast<double *>(lds);
    const double bm = block_reduce<true>(mx, sb, tid, NT);
    if (tid == 0) {
        // partials are indexed per full-length z-pass block: LZ / z_lines of
        // them per block here
        constexpr int RATIO = LZ / z_lines<T, N>();
        part[RATIO * blk] = bm;
#pragma unroll
        for (int i = 1; i < RATIO; ++i) part[RATIO * blk + i] = 0.0;
    }
}

// ------------------------------------------------------------ strided passes
// AXIS 1: lines along y (stride rl); AXIS 0: lines along x (stride N*rl).
// A work-group takes TC consecutive z-columns so each global access of a
// wavefront covers 128-byte row segments.

// the transform (and the spectral multiply of passes C and G) of one tile;
// FRESH: the first transform of the work-group (no earlier LDS use).  tab
// is the drift table (global memory, or an LDS copy).
// HALF: tab holds the drift table for i <= N/2 only, and point i takes entry
// min(i, N - i) (k^2 of i and N - i are the same double: kvec(i) for i >= N/2
// is the exact negation of kvec(N - i), so the table entries are equal).
template <typename T, int N, int MODE, bool FRESH, int AXIS, typename W, int V, bool HALF = false>
__device__ __forceinline__ void spass_compute(cx<T> *v, cx<T> *L, int t, const TwBase<T, N, W, V> &tw,
                                              const SArgs &sa, int other, int k, bool valid,
                                              const cx<T> *tab) {
    constexpr int E = SchedT<T, N>::E, P = N / E;
    constexpr bool ACC = acc_tw<T, N>() && (MODE == S_DRIFT1 || MODE == S_KICK_DRIFT1);
    if constexpr (MODE == S_FWD) fft_line<T, N, -1, false, FRESH, ACC>(v, L, t, tw);
    else if constexpr (MODE == S_INV) fft_line<T, N, +1, false, FRESH, ACC>(v, L, t, tw);
    else {
        fft_line<T, N, -1, false, FRESH, ACC>(v, L, t, tw);
        // spectral indices of this thread's points: along the line i, fixed `other` and k
        if constexpr (MODE == S_DRIFT1 || MODE == S_KICK_DRIFT1) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int i = t + e * P;
                v[e] = cmul(v[e], tab[HALF ? (i <= N / 2 ? i : N - i) : i]);
            }
        } else {
            if constexpr (sizeof(T) == 4) {
                // complex floats: the Green factor in FP32 (an approximate
                // reciprocal, ~1e-7; the potential is recomputed each step)
                // from the integer |m|^2 = m_x^2 + m_y^2 + m_z^2 (exact), no
                // table loads: k^2 = (2 pi / L)^2 |m|^2
                const float g = (float)(sa.green * sa.fftnorm / sa.kunit2);
                const int mo = other < N / 2 ? other : other - N, mk = k < N / 2 ? k : k - N;
                const int mok = mo * mo + (valid ? mk * mk : 0);
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const int i = t + e * P, mi = i < N / 2 ? i : i - N;
                    const int m2 = mi * mi + mok;
                    v[e] = cscale(v[e], m2 == 0 ? 0.f : g * __builtin_amdgcn_rcpf((float)m2));
                }
            } else {
                // complex doubles: the same integer |m|^2, one FP64 quotient
                // per point, no table loads (k runs over the half spectrum,
                // or over all N columns in the initial conditions' solve)
                const double g = sa.green * sa.fftnorm / sa.kunit2;
                const int mo = other < N / 2 ? other : other - N, mk = k < N / 2 ? k : k - N;
                const int mok = mo * mo + (valid ? mk * mk : 0);
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const int i = t + e * P, mi = i < N / 2 ? i : i - N;
                    const int m2 = mi * mi + mok;
                    v[e] = cscale(v[e], m2 == 0 ? 0.0 : g / (double)m2);
                }
            }
        }
        fft_line<T, N, +1, false, false, ACC>(v, L, t, tw);
    }
}

template <int N, int AXIS>
__device__ __forceinline__ void spass_addr(int other, int k, int rl, long ps, long &base, long &stride) {
    if constexpr (AXIS == 1) { base = (long)other * ps + k; stride = rl; }
    else { base = (long)other * rl + k; stride = ps; }
}

// Element access within one tile: a wave-uniform tile base (scalar
// registers) plus a per-lane 32-bit byte offset, which maps to the
// scalar-base + vector-offset form of global loads and stores.  Up to
// N = 512 every offset inside an array fits 32 bits; N = 1024 uses 64-bit
// addresses.  NTS, NTL: non-temporal stores, loads (each tile is read and
// written once per pass).  Measured at 512^3, loads on top of the stores:
// FP64 202.4 -> 205.9 steps/s, mixed 370.4 -> 384; every strided pass
// gains except the 256-byte tiles of the FP64 x pass (0.93 -> 0.99 ms),
// which keep plain loads.  Tiles with rows narrower than a 128-byte line
// (complex floats on F and H at N = 1024) keep plain accesses: two
// work-groups share each line there, and the second finds it in L2 only
// if the first kept it (non-temporal: F 1.95 -> 3.0 ms at 1024^3).
template <typename T, int N, bool NTL = false, bool NTS = true> struct TileIO {
    static constexpr bool k32 = N <= 512;
    const char *base;   // wave-uniform
    __device__ __forceinline__ TileIO(const cx<T> *a, long uniform_elem) {
        const long u = __builtin_amdgcn_readfirstlane((int)(uniform_elem >> 32)) * (1L << 32) +
                       (unsigned)__builtin_amdgcn_readfirstlane((int)(uniform_elem & 0xffffffffL));
        base = reinterpret_cast<const char *>(a + u);
    }
    __device__ __forceinline__ cx<T> load(long elem) const {
        if constexpr (NTL) {
            if constexpr (k32) return stream_load(reinterpret_cast<const cx<T> *>(base + (unsigned)((unsigned)elem * sizeof(cx<T>))));
            else return stream_load(reinterpret_cast<const cx<T> *>(base) + elem);
        } else if constexpr (k32) {
            return *reinterpret_cast<const cx<T> *>(base + (unsigned)((unsigned)elem * sizeof(cx<T>)));
        } else {
            return reinterpret_cast<const cx<T> *>(base)[elem];
        }
    }
    __device__ __forceinline__ void store(long elem, cx<T> v) const {
        cx<T> *p;
        if constexpr (k32) p = reinterpret_cast<cx<T> *>(const_cast<char *>(base) + (unsigned)((unsigned)elem * sizeof(cx<T>)));
        else p = reinterpret_cast<cx<T> *>(const_cast<char *>(base)) + elem;
        if constexpr (NTS) stream_store(p, v);
        else *p = v;
    }
};

// Element e (row t + e P) of this lane's column of a tile: a wave-uniform
// address per e (the tile's row 0 plus e P rows, formed in scalar
// registers) plus one per-lane 32-bit byte offset (column c, row t) shared
// by every e.  It serves the kick pass, whose registers are the scarcest:
// there a 64-bit address per point (N = 1024, where whole-array offsets
// exceed 32 bits; and phi's at every N), held from the loads to the stores,
// took 2 VGPRs per point and spilled (12-96 bytes per lane; the spill
// traffic was 8 % of the pass's bytes at 512^3 and 23-40 % at 1024^3).
// The lane offset is at most (P - 1) rows of the largest stride, under
// 2.2 GB for every pass here.

// a wave-uniform 64-bit value moved to scalar registers
__device__ __forceinline__ long uniform64(long v) {
    return (long)__builtin_amdgcn_readfirstlane((int)(v >> 32)) * (1L << 32) +
           (unsigned)__builtin_amdgcn_readfirstlane((int)(v & 0xffffffffL));
}

template <typename V, bool NTL = false, bool NTS = false> struct ColIO {
    const char *a0;     // the array (a kernel argument: global memory)
    long ub;            // wave-uniform: byte offset of the tile's row 0
    long step;          // wave-uniform: bytes from row t + e P to row t + (e + 1) P
    unsigned off;       // per lane
    __device__ __forceinline__ ColIO(const V *a, long uniform_elem, long stride, int rows, long lane_elem) {
        a0 = reinterpret_cast<const char *>(a);
        ub = uniform64(uniform_elem * (long)sizeof(V));
        step = uniform64((long)rows * stride * (long)sizeof(V));
        off = (unsigned)lane_elem * (unsigned)sizeof(V);
    }
    __device__ __forceinline__ V *ptr(int e) const {
        // the row's offset formed in scalar registers; the lane adds only
        // its 32-bit offset
        return reinterpret_cast<V *>(const_cast<char *>(a0) + uniform64(ub + e * step) + off);
    }
    __device__ __forceinline__ V load(int e) const {
        if constexpr (NTL) return stream_load(ptr(e));
        else return *ptr(e);
    }
    __device__ __forceinline__ void store(int e, V v) const {
        if constexpr (NTS) stream_store(ptr(e), v);
        else *ptr(e) = v;
    }
};

// Work-groups start round-robin over the 8 XCDs (linear id mod 8), each
// with its own L2.  The kick pass's column tiles 2j and 2j+1 each read half
// of every 128-byte segment of phi; this maps them to linear ids 8 apart,
// i.e. to one XCD at about the same time, so the second half is an L2 hit.
// Within each run of 16 linear ids: id 16g + 8h + x -> tile 16g + 2x + h.
// Needs gridDim.x (tiles per line) to be a multiple of 16.
__device__ __forceinline__ void xcd_pair(int &bx, int &by) {
    const int L = blockIdx.x + gridDim.x * blockIdx.y;
    const int r = L & 15;
    const int Ll = (L & ~15) + 2 * (r & 7) + (r >> 3);
    by = Ll / (int)gridDim.x;
    bx = Ll - by * (int)gridDim.x;
}

// Minimum waves per SIMD asked of the compiler for a strided pass: 8 caps
// the complex-float Green pass at 64 VGPRs, so two 1024-thread work-groups
// share a CU (measured at 512^3: 0.39 -> 0.35 ms), for the sizes where two
// such work-groups fit the LDS, and the complex-float kick pass at N = 512
// (68 VGPRs uncapped, one work-group per CU: 0.68 -> 1.02 ms; capped, 12
// bytes of spill); 4 caps the narrow 512-thread tiles of N = 1024 at 128
// VGPRs, so two of them share a CU; the other passes fit without a cap
// (capping the z-pass C spilled and slowed).
template <typename T, int MODE, int N, bool WIDE = false, int AXIS = 1> constexpr int sp_wpe() {
    return (sizeof(T) == 4 && ((MODE == S_GREEN && N >= 128 && N <= 512) || (MODE == S_KICK_DRIFT1 && N == 512))) ? 8
                                                                          : (s_narrow<T, N, WIDE, AXIS, MODE>() ? 4 : 1);
}
// The kick pass's phi loads.  Batched (0 < PB): each batch of PB loads is
// issued together and unconditionally, then its kicks; the batches are kept
// apart by a scheduling barrier.  A load under a condition is a branch region
// of its own, and the compiler then waits for each before issuing the next
// (eight or sixteen memory round trips in series per tile).  Measured: FP64
// A 1.12 -> 1.00 ms at 512^3; complex floats at 1024^3 with batches of 8
// (of 16) 7.45 -> 6.70 ms (batches of 16: 6.92).  Complex floats up to
// N = 512 load one value per kick (PB = 0), unconditionally where every
// column is valid (ALLCOLS in k_spass), and leave the schedule to the
// compiler: batches of 8 took 78 VGPRs (one work-group per CU: 0.71 -> 0.84
// ms; capped at 64, they spilled: 0.80).
template <typename T, int N> constexpr int kick_pb() {
    return sizeof(T) == 8 ? SchedT<T, N>::E : (N == 1024 ? 8 : 0);
}

// Narrow tiles: the grid is (a multiple of G column tiles, lines of the
// fixed axis).  Within each run of 8G linear ids, id 8Gg + 8h + x takes the
// virtual tile 8Gg + Gx + h (row-major over the grid), so the G column
// tiles Gm .. Gm + G - 1 of a row start 8 ids apart, on one XCD (ids go
// round-robin over the 8 XCDs): G = 2 for the two 64-byte halves of a
// 128-byte psi segment, G = 4 where four tiles also share a 128-byte phi
// segment (the kick pass: 32-byte phi rows).  Needs gridDim.x * gridDim.y
// to be a multiple of 8G.
template <int G> __device__ __forceinline__ void xcd_group_rows(int &bx, int &by) {
    const int L = blockIdx.x + gridDim.x * blockIdx.y;
    const int r = L & (8 * G - 1);
    const int V = (L & ~(8 * G - 1)) + G * (r & 7) + (r >> 3);
    by = V / (int)gridDim.x;
    bx = V - by * (int)gridDim.x;
}
// tiles per XCD group of a narrow pass
template <int MODE> constexpr int xcd_group() { return MODE == S_KICK_DRIFT1 ? 4 : 2; }

// one tile per work-group: grid (z-column tiles, lines of the fixed axis)
template <typename T, int N, int AXIS, int MODE, bool WIDE>
__global__ __launch_bounds__((s_cols<T, N, WIDE, MODE, AXIS>() * (N / SchedT<T, N>::E)), (sp_wpe<T, MODE, N, WIDE, AXIS>()))
void k_spass(cx<T> *__restrict__ a, int rl, long ps, int kv, const cx<T> *__restrict__ tw, const cx<double> *__restrict__ tw64, SArgs sa) {
    constexpr int E = SchedT<T, N>::E, P = N / E, TC = s_cols<T, N, WIDE, MODE, AXIS>();
    constexpr int SV = sp_sched<T, N, MODE>(), LS = lds_col_line<N, s_narrow<T, N, WIDE, AXIS, MODE>(), pad_shift_t<T, N, SV>()>();
    __shared__ cx<T> lds[TC * LS];
    const int tid = threadIdx.x, c = tid % TC, t = tid / TC;
    int bx = blockIdx.x, by = blockIdx.y;
    if constexpr (s_narrow<T, N, WIDE, AXIS, MODE>()) {
        xcd_group_rows<xcd_group<MODE>()>(bx, by);
        if (bx * TC >= kv) return;            // a padding tile of the rounded-up count (whole work-group)
    } else if constexpr (MODE == S_KICK_DRIFT1 && (N / TC) % 16 == 0) {
        xcd_pair(bx, by);
    }
    // The drift passes read the drift table from LDS (at N = 1024 in FP64
    // through schedule variant 2, whose shorter line buffer leaves the room).
    // From global memory its loads, issued after the forward transform, cost
    // an L2 round trip per tile, and under the 128-VGPR cap of the 1024-point
    // complex-float tiles each of the 16 was awaited before the next.  Narrow
    // tiles and complex doubles keep entries 0..N/2 (see spass_compute): two
    // work-groups per CU leave no room for all N.  Measured (ms, global ->
    // LDS table) at 512^3: mixed A 0.60 -> 0.54, B 0.48 -> 0.44; FP64 A 0.96
    // -> 0.95, B 0.88 -> 0.83; at 1024^3 FP64: A 9.77 -> 9.19, B 9.54 -> 8.98.
    // The forward transform's exchanges order these writes before the reads.
    constexpr bool LTAB = (MODE == S_DRIFT1 || MODE == S_KICK_DRIFT1) && (N == 512 || (N == 1024 && (sizeof(T) == 4 || SV == 2)));
    constexpr bool HALF = LTAB && (s_narrow<T, N, WIDE, AXIS, MODE>() || sizeof(T) == 8);
    constexpr int TABN = LTAB ? (HALF ? N / 2 + 1 : N) : 1;
    __shared__ cx<T> tabL[TABN];
    if constexpr (LTAB) {
        for (int i = (int)threadIdx.x; i < TABN; i += TC * P) tabL[i] = ((const cx<T> *)sa.dtab)[i];
    }
    const int k0 = bx * TC, k = k0 + c;
    const int other = sa.other0 + block_order(sa.rev, by, gridDim.y);
    constexpr bool PSI = MODE == S_DRIFT1 || MODE == S_KICK_DRIFT1;
    // The psi passes cover every column of every row (kv = N, checked at the
    // launch), so where TC divides N no tile holds a column past the row.
    // The kick pass then loads psi and phi without a condition: a load under
    // a condition is a branch region of its own, awaited before the next one
    // is issued (measured at 512^3: mixed A 0.68 -> 0.60 ms with the
    // branch-free quadrant selection of expmi_parts, FP64 A 1.04 -> 0.96).
    // The drift-only passes keep the condition (without it, mixed B took 74
    // VGPRs, one work-group per CU: 0.47 -> 0.53 ms).
    constexpr bool ALLCOLS = MODE == S_KICK_DRIFT1 && N % TC == 0;
    const bool valid = ALLCOLS || k < kv;
    long base0, stride;
    spass_addr<N, AXIS>(other, k0, rl, ps, base0, stride);
    constexpr bool ROW128 = TC * sizeof(cx<T>) >= 128;
    // plain loads on the FP64 256-byte tiles and on rows narrower than a line
    constexpr bool NTL = ROW128 && !(WIDE && sizeof(T) == 8);
    const TileIO<T, N, NTL, ROW128> io(a, base0);
    // the kick pass: per-e uniform addresses and one lane offset (ColIO;
    // measured at 1024^3, mixed A 6.25 -> 4.93 ms, FP64 A 12.1 -> 9.6 ms; the
    // other passes, which do not spill, the same either way)
    constexpr bool COL = MODE == S_KICK_DRIFT1;
    const ColIO<cx<T>, NTL, ROW128> cio(a, base0, stride, P, c + (long)t * stride);
    cx<T> *L = lds + c * LS;
    TwBase<T, N, tw_base_t<T, N, PSI>, SV> tb;
    if constexpr (acc_tw<T, N>() && PSI) load_twiddles<T, N>(tb, t, tw64);
    else load_twiddles<T, N>(tb, t, tw);
    cx<T> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e)
        v[e] = valid ? (COL ? cio.load(e) : io.load(c + (long)(t + e * P) * stride)) : mkc<T>(0, 0);
    if constexpr (MODE == S_KICK_DRIFT1) {
        const double kaphi = sa.dk ? sa.dk[0] : sa.aphi, karho = sa.dk ? sa.dk[1] : sa.arho;
        // phi loads: see kick_pb.  Batched, a column past kv reads the
        // tile's first column, which lies inside the row (k0 < kv), and its
        // value is not used.
        constexpr int PB = kick_pb<T, N>();
        const int cp = valid ? c : 0;
        // phi through ColIO as psi
        const ColIO<T> pio((const T *)sa.phi, base0, stride, P, cp + (long)t * stride);
        if constexpr (PB == 0) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const double f = valid ? (double)pio.load(e) : 0.0;
                const double rr = (double)v[e].x * v[e].x + (double)v[e].y * v[e].y;
                v[e] = cmul(v[e], expmi<T>(kaphi * f + karho * rr));
            }
        } else {
#pragma unroll
            for (int e0 = 0; e0 < E; e0 += PB) {
                T f[PB];
#pragma unroll
                for (int i = 0; i < PB; ++i) f[i] = pio.load(e0 + i);
#pragma unroll
                for (int i = 0; i < PB; ++i) {
                    const int e = e0 + i;
                    const double rr = (double)v[e].x * v[e].x + (double)v[e].y * v[e].y;
                    v[e] = cmul(v[e], expmi<T>(kaphi * (valid ? (double)f[i] : 0.0) + karho * rr));
                }
                if (PB < E) __builtin_amdgcn_sched_barrier(0);
            }
        }
    }
    spass_compute<T, N, MODE, true, AXIS, tw_base_t<T, N, PSI>, SV, HALF>(v, L, t, tb, sa, other, k, valid,
                                                                        LTAB ? (const cx<T> *)tabL : (const cx<T> *)sa.dtab);
    if (valid) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if constexpr (COL) cio.store(e, v[e]);
            else io.store(c + (long)(t + e * P) * stride, v[e]);
        }
    }
}

// Persistent variant: each work-group walks tiles blockIdx.x, +gridDim.x, ...
// (tile = local line of the fixed axis * ntz + z-column tile) and issues the
// loads of its next tile before transforming the current one, so HBM reads
// overlap the butterflies and LDS exchanges.  The first LDS write of every
// transform is preceded by a barrier, so consecutive tiles never race on LDS.
template <typename T, int N, int AXIS, int MODE>
__global__ __launch_bounds__((s_cols<T, N, false, MODE, AXIS>() * (N / SchedT<T, N>::E)))
void k_spass_pp(cx<T> *__restrict__ a, int rl, long ps, int kv, const cx<T> *__restrict__ tw, const cx<double> *__restrict__ tw64, SArgs sa, int ntz,
                int ntiles) {
    constexpr int E = SchedT<T, N>::E, P = N / E, TC = s_cols<T, N, false, MODE, AXIS>(), SV = sp_sched<T, N, MODE>();
    constexpr int LS = lds_col_line<N, false, pad_shift_t<T, N, SV>()>();
    constexpr int NT = TC * P;
    __shared__ cx<T> lds[TC * LS];
    // the spectral multiply's table in LDS, so no global load is issued
    // after a tile's prefetch (loads complete in order: waiting for a table
    // value would also wait for the prefetch)
    // (up to N = 512; the 1024-point exchange buffer leaves no room)
    constexpr bool LT = N <= 512;
    constexpr bool DR = MODE == S_DRIFT1;
    static_assert(MODE != S_KICK_DRIFT1, "the kick pass runs one tile per work-group");
    __shared__ cx<T> tabL[(LT && DR) ? N : 1];
    const int tid = threadIdx.x, c = tid % TC, t = tid / TC;
    cx<T> *L = lds + c * LS;
    int tile = blockIdx.x;
    if (tile >= ntiles) return;
    if constexpr (LT && DR) {
        for (int i = tid; i < N; i += NT) tabL[i] = ((const cx<T> *)sa.dtab)[i];
    }
    const cx<T> *tab = LT ? (const cx<T> *)tabL : (const cx<T> *)sa.dtab;
    const long stride = AXIS == 1 ? (long)rl : ps;
    // With at least TC columns per line every column of every tile lies
    // inside its row (rl >= the tile span; columns k >= kv are the half
    // spectrum's zero padding), so loads and stores need no condition and
    // the count of memory operations in flight is the same on every path:
    // the wait for the prefetch then lets this tile's stores stay in flight.
    constexpr bool FULL = N >= TC;
    constexpr bool PSI = MODE == S_DRIFT1;
    TwBase<T, N, tw_base_t<T, N, PSI>, SV> tb;
    if constexpr (acc_tw<T, N>() && PSI) load_twiddles<T, N>(tb, t, tw64);
    else load_twiddles<T, N>(tb, t, tw);
    cx<T> v[E], nx[E];
    {
        const int other = sa.other0 + tile / ntz, k0 = (tile % ntz) * TC;
        long b0, st;
        spass_addr<N, AXIS>(other, k0, rl, ps, b0, st);
        const TileIO<T, N, true> io(a, b0);
        const bool valid = k0 + c < kv;
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = (FULL || valid) ? io.load(c + (long)(t + e * P) * stride) : mkc<T>(0, 0);
    }
    // nothing in flight on entry to the loop: the waits the compiler places
    // in the loop then follow the steady state alone (this tile's stores may
    // stay in flight while the next tile's loads are awaited)
    __builtin_amdgcn_s_waitcnt(0);
    for (;;) {
        const int nt = tile + (int)gridDim.x;
        const bool more = nt < ntiles;
        {
            // the last tile re-reads itself, so every iteration issues the same
            // number of loads (the compiler's waits then count exactly)
            const int lt = more ? nt : tile;
            const int o2 = sa.other0 + lt / ntz, k2 = (lt % ntz) * TC;
            long b2, st;
            spass_addr<N, AXIS>(o2, k2, rl, ps, b2, st);
            const TileIO<T, N, true> io(a, b2);
            const bool valid2 = k2 + c < kv;
#pragma unroll
            for (int e = 0; e < E; ++e) nx[e] = (FULL || valid2) ? io.load(c + (long)(t + e * P) * stride) : mkc<T>(0, 0);
        }
        __builtin_amdgcn_sched_barrier(0);   // issue the prefetch before the transform
        const int other = sa.other0 + tile / ntz, k0 = (tile % ntz) * TC, k = k0 + c;
        const bool valid = k < kv;
        spass_compute<T, N, MODE, false, AXIS>(v, L, t, tb, sa, other, k, valid, tab);
        {
            long b0, st;
            spass_addr<N, AXIS>(other, k0, rl, ps, b0, st);
            const TileIO<T, N, true> io(a, b0);
            if (FULL || valid) {
#pragma unroll
                for (int e = 0; e < E; ++e) io.store(c + (long)(t + e * P) * stride, v[e]);
            }
        }
        __builtin_amdgcn_sched_barrier(0);   // take over the prefetch only after the stores
        if (!more) break;
        for (int e = 0; e < E; ++e) v[e] = nx[e];   // register moves (becomes an array copy: no unroll pragma)
        tile = nt;
    }
}

// --------------------------------------------------------------- elementwise
template <typename T>
__global__ __launch_bounds__(256) void k_kick(cx<T> *__restrict__ psi, const T *__restrict__ phi, long n3,
                                              double aphi, double arho) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (long)gridDim.x * blockDim.x) {
        cx<T> z = psi[i];
        const double rr = (double)z.x * z.x + (double)z.y * z.y;
        psi[i] = cmul(z, expmi<T>(aphi * (double)phi[i] + arho * rr));
    }
}

// ------------------------------------------------------------------ adjoint
// s = Im(conj(g) psi): the real field whose Green's-function image enters
// the adjoint of a kick
template <typename T>
__global__ __launch_bounds__(256) void k_adj_src(const cx<T> *__restrict__ psi, const cx<T> *__restrict__ g,
                                                 T *__restrict__ s, long n3) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (long)gridDim.x * blockDim.x) {
        const cx<T> a = g[i], b = psi[i];
        s[i] = a.x * b.y - a.y * b.x;
    }
}

// Adjoint of the kick psi' = psi exp(-i theta), theta = aphi phi + arho |psi|^2
// (phi = G(m_s |psi|^2 - rho_c)), applied in reverse: with psi' and the
// gradient g' of the kicked field,
//   psi = psi' e^{+i theta},  g = e^{+i theta} g' + 2 psi (aphi G(m_s s) + arho s),
// s = Im(conj(g') psi'), phs = G(m_s s) (a Poisson solve of s).
template <typename T>
__global__ __launch_bounds__(256) void k_adj_kick(cx<T> *__restrict__ psi, cx<T> *__restrict__ g,
                                                  const T *__restrict__ phi, const T *__restrict__ phs,
                                                  const T *__restrict__ s, long n3, double aphi, double arho) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (long)gridDim.x * blockDim.x) {
        const cx<T> p1 = psi[i], g1 = g[i];
        const double rr = (double)p1.x * p1.x + (double)p1.y * p1.y;
        cx<T> e = expmi<T>(aphi * (double)phi[i] + arho * rr);
        e.y = -e.y;                                            // e^{+i theta}
        const cx<T> p0 = cmul(p1, e);
        const T as = (T)(aphi * (double)phs[i] + arho * (double)s[i]);
        const cx<T> eg = cmul(e, g1);
        psi[i] = p0;
        g[i] = mkc<T>(eg.x + 2 * p0.x * as, eg.y + 2 * p0.y * as);
    }
}

// ------------------------------------------------------------ diagnostics
// Matter power spectrum from the forward 3-D half spectrum of the Poisson
// source: per |n|-shell (n = integer wave vector, bin b = round(|n|),
// b = 1..nbins) the sums of w |s_k|^2, w |n| and w, with w = 1 on the
// kz = 0 and kz = N/2 planes and 2 elsewhere (Hermitian partners).  Each
// block bins into LDS, then adds its bins to the global sums.
template <typename T, int N>
__global__ __launch_bounds__(256) void k_pk_bin(const cx<T> *__restrict__ shat, int hs, long sps, int nbins,
                                                double *__restrict__ out) {
    extern __shared__ double sb[];   // 3 * nbins
    for (int i = threadIdx.x; i < 3 * nbins; i += blockDim.x) sb[i] = 0.0;
    __syncthreads();
    const long nh = (long)N * N * (N / 2 + 1);
    for (long o = (long)blockIdx.x * blockDim.x + threadIdx.x; o < nh; o += (long)gridDim.x * blockDim.x) {
        const int kz = (int)(o % (N / 2 + 1));
        const long r = o / (N / 2 + 1);
        const int j = (int)(r % N), i = (int)(r / N);
        const int nx = i < N / 2 ? i : i - N, ny = j < N / 2 ? j : j - N;
        const double nn = sqrt((double)nx * nx + (double)ny * ny + (double)kz * kz);
        const int b = (int)(nn + 0.5);
        if (b < 1 || b > nbins) continue;
        const cx<T> v = shat[(long)i * sps + (long)j * hs + kz];
        const double w = (kz == 0 || kz == N / 2) ? 1.0 : 2.0;
        atomicAdd(&sb[3 * (b - 1)], w * ((double)v.x * v.x + (double)v.y * v.y));
        atomicAdd(&sb[3 * (b - 1) + 1], w * nn);
        atomicAdd(&sb[3 * (b - 1) + 2], w);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 3 * nbins; i += blockDim.x)
        if (sb[i] != 0.0) atomicAdd(&out[i], sb[i]);
}

// zero fill (used where a caller may capture the stream in a graph, in
// place of a memset node)
__global__ __launch_bounds__(256) void k_fill0(double *__restrict__ p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0.0;
}

// P(k) from the binned sums (sb[3b] = sum w |s_k|^2, sb[3b+1] = sum w |n|,
// sb[3b+2] = sum w) and the device reductions red = {max|phi|, max|psi|^2,
// sum|psi|^2} of the same pass: out[b] = L^3 sum / (rho_mean^2 N^6 w),
// out[nb + b] = kf sum|n| / w, out[2 nb + b] = w, with rho_mean =
// m_s psc2 sum|psi|^2 / N^3 (psc2: the mixed mode's psi scale squared)
__global__ __launch_bounds__(256) void k_pk_finalize(const double *__restrict__ sb, int nb,
                                                     const double *__restrict__ red, double ms_psc2, double n3,
                                                     double l3, double n6, double kf, double *__restrict__ out) {
    const double rho_mean = ms_psc2 * red[2] / n3;
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
        const double w = sb[3 * b + 2];
        out[b] = w > 0 ? l3 * sb[3 * b] / (rho_mean * rho_mean * n6) / w : 0.0;
        out[nb + b] = w > 0 ? kf * sb[3 * b + 1] / w : kf * (b + 1);
        out[2 * nb + b] = w;
    }
}

// plain 16-byte-vector copy: the achievable streaming bandwidth reference
__global__ __launch_bounds__(256) void k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// final reduction of the per-block partials in two stages (latency-bound
// otherwise): stage 1 with 256-thread blocks, each thread one partial per
// stride; stage 2 one block.  out = {max|phi|, max|psi|^2, sum|psi|^2}.
__global__ __launch_bounds__(256) void k_reduce1(const double *__restrict__ prho, int nrho,
                                                 const double *__restrict__ pphi, int nphi,
                                                 double *__restrict__ stage) {
    __shared__ double sb[16];
    const int tid = threadIdx.x, g = blockIdx.x * 256 + tid, stride = gridDim.x * 256;
    double mr = 0.0, sm = 0.0, mp = 0.0;
    for (int i = g; i < nrho; i += stride) { mr = fmax(mr, prho[2 * i]); sm += prho[2 * i + 1]; }
    for (int i = g; i < nphi; i += stride) mp = fmax(mp, pphi[i]);
    const double a = block_reduce<true>(mp, sb, tid, 256);
    const double b = block_reduce<true>(mr, sb, tid, 256);
    const double c = block_reduce<false>(sm, sb, tid, 256);
    if (tid == 0) { stage[3 * blockIdx.x] = a; stage[3 * blockIdx.x + 1] = b; stage[3 * blockIdx.x + 2] = c; }
}

__global__ __launch_bounds__(256) void k_reduce2(const double *__restrict__ stage, int nst, double *__restrict__ out) {
    __shared__ double sb[16];
    const int tid = threadIdx.x;
    double mp = 0.0, mr = 0.0, sm = 0.0;
    for (int i = tid; i < nst; i += 256) {
        mp = fmax(mp, stage[3 * i]);
        mr = fmax(mr, stage[3 * i + 1]);
        sm += stage[3 * i + 2];
    }
    const double a = block_reduce<true>(mp, sb, tid, 256);
    const double b = block_reduce<true>(mr, sb, tid, 256);
    const double c = block_reduce<false>(sm, sb, tid, 256);
    if (tid == 0) { out[0] = a; out[1] = b; out[2] = c; }
}

// ------------------------------------------------------ on-device step loop
// The adaptive loop of Simulation (src/CosmoBEC_MPI_stan.cpp:212-304) with
// the time step decided on the device: after each step's passes, one
// work-group reduces the partials, evaluates Timestep_criteria (criteria.h,
// the same arithmetic as the host's ifdm_timestep_criteria), advances a,
// forms the kick coefficients of the next kick pass (the pending second
// half-kick plus the new first one) and writes the drift table.  The host
// then issues every step's launches back to back with no wait between
// steps (a wait, the copy of the reductions and the host's criteria took
// 30-40 us of idle GPU per step).

// loop state in device memory
struct DevStep {
    double a;                       // scale factor after the last prepared step
    double kphi, krho;              // kick coefficients of the next kick pass (krho times psc2)
    double pend_aphi, pend_arho;    // second half-kick of the last prepared step (pending)
    double rho_mean;                // m_s <|psi|^2> of the reductions read (mixed-mode source)
    double max_phi, max_rho, norm;  // the reductions read
};
// what a preparation takes from the host
struct StepIn {
    CritConst cc;
    double a0;                       // first step of a call: the scale factor
    double max_phi, max_rho, norm;   // first step: the reductions (host values)
    double pend_aphi, pend_arho;     // first step: a half-kick pending from before
    double psc2, ms, n3, m_over_hbar, g_over_hbar, ke_coef;
    int first, has_pend;
};
// per-step record: dKE, dPE, dQI, a after the step
static constexpr int IFDM_REC = 4;
static constexpr int IFDM_QAGS_LIMIT = 1000;

template <typename T, int N>
__global__ __launch_bounds__(256) void k_step_prep(const double *__restrict__ stage, int nst, StepIn in,
                                                   DevStep *__restrict__ st, double *__restrict__ rec,
                                                   cx<T> *__restrict__ tab, const double *__restrict__ k2,
                                                   Piece *__restrict__ pcs) {
    IFDM_FP_EXACT
    __shared__ double sb[16];
    __shared__ double fv[3][21];
    __shared__ double sh[5];   // a, a + da, dKE, dPE, dQI
    const int tid = threadIdx.x;
    double max_phi, max_rho, norm;
    if (in.first) {
        max_phi = in.max_phi;
        max_rho = in.max_rho;
        norm = in.norm;
    } else {
        // the final reduction of k_reduce2, then read_stats's scaling
        double mp = 0.0, mr = 0.0, sm = 0.0;
        for (int i = tid; i < nst; i += 256) {
            mp = fmax(mp, stage[3 * i]);
            mr = fmax(mr, stage[3 * i + 1]);
            sm += stage[3 * i + 2];
        }
        max_phi = block_reduce<true>(mp, sb, tid, 256);
        max_rho = block_reduce<true>(mr, sb, tid, 256) * in.psc2;
        norm = block_reduce<false>(sm, sb, tid, 256) * in.psc2;
    }
    if (tid == 0) {
        const double a = in.first ? in.a0 : st->a;
        sh[0] = a;
        sh[1] = a + criteria_da(in.cc, a, max_phi, max_rho, nullptr);
    }
    __syncthreads();
    const double a = sh[0], b = sh[1];
    // dKE, dPE, dQI: the integrals of 1/(H a^n), n = 3, 2, 4 (:66-68)
    auto power = [](int i) { return i == 0 ? 3 : (i == 1 ? 2 : 4); };
    if (tid < 63) fv[tid / 21][tid % 21] = integrand(in.cc, gk21_node(a, b, tid % 21), power(tid / 21));
    __syncthreads();
    if (tid < 3) {
        double err, rabs, rasc;
        double r = qk21_values(a, b, fv[tid], &err, &rabs, &rasc);
        if (!qags_first_accepted(r, err, rasc, 1e-10)) {
            bool ok;
            r = qags_bisect(in.cc, a, b, power(tid), r, err, 1e-10, IFDM_QAGS_LIMIT, pcs + IFDM_QAGS_LIMIT * tid, &ok);
        }
        sh[2 + tid] = r;
    }
    __syncthreads();
    const double dKE = sh[2];
    if (tid == 0) {
        const double dPE = sh[3], dQI = sh[4];
        // ifdm_step's coefficients: the half-kick of this step plus the pending one
        const double aphi = 0.5 * dPE * in.m_over_hbar, arho = 0.5 * dQI * in.g_over_hbar;
        const bool pend = in.first ? in.has_pend : true;
        double kphi = aphi, krho = arho;
        if (pend) {
            kphi += in.first ? in.pend_aphi : st->pend_aphi;
            krho += in.first ? in.pend_arho : st->pend_arho;
        }
        st->a = b;
        st->kphi = kphi;
        st->krho = krho * in.psc2;
        st->pend_aphi = aphi;
        st->pend_arho = arho;
        st->rho_mean = in.ms * norm / in.n3;
        st->max_phi = max_phi;
        st->max_rho = max_rho;
        st->norm = norm;
        rec[0] = dKE;
        rec[1] = dPE;
        rec[2] = dQI;
        rec[3] = b;
    }
    write_drift_table<T, N>(tab, k2, in.ke_coef * dKE, tid, 256);
}

}  // namespace ifdm
