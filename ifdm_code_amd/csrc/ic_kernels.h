// ic_kernels.h -- initial conditions on the device: cloud-in-cell deposit of
// Gadget particles, mass-weighted velocities, the velocity-divergence source
// of the phase, and the wave-function assembly.
//
// Reference: src/gadget_IC.h (CIC_scheme :45-156, read_gadget :583-670,
// initial_velocity_Poisson_Solver :157-247).  Two deposits:
//   * reference order (the default): every cell receives its contributions
//     in the order the reference's particle loop visits them, the density
//     accumulated in FP64 and the velocity grids in float (std::vector<float>,
//     :583; `vxg[c] += double` rounds to float after each add), so the grids
//     equal the reference's bit for bit and two runs are identical;
//   * FP64 atomics: unordered device-scope FP64 atomic adds into FP64 grids
//     (sums agree with a sequential FP64 deposit to round-off).
// Every double expression keeps the reference's association, and floating
// contraction is off in these kernels (a fused multiply-add would change the
// bits the reference computes).
#pragma once
#include <hip/hip_runtime.h>

#include "fft_core.h"

namespace ifdm {

struct CicArgs {
    int n;          // mesh points per axis
    double box;     // kpc/h (header BoxSize)
    double lg;      // box / n
    double pmass;   // Msun/h (header mass[1] * 1e10)
    double l3;      // lg^3
};

__device__ __forceinline__ double xg_of(int i, double lg) {
#pragma clang fp contract(off)
    return i * lg + 0.5 * lg;
}

// One thread per particle.  Returns nothing; particles outside
// [-lg/2, box) in some coordinate are counted in *skipped and not deposited.
static __global__ __launch_bounds__(256) void k_cic(const float *__restrict__ pos, const float *__restrict__ vel, long np,
                                             CicArgs a, double *__restrict__ rho, double *__restrict__ vx,
                                             double *__restrict__ vy, double *__restrict__ vz,
                                             unsigned long long *__restrict__ skipped) {
#pragma clang fp contract(off)
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    float xp = pos[3 * p], yp = pos[3 * p + 1], zp = pos[3 * p + 2];
    const float vp[3] = {vel[3 * p], vel[3 * p + 1], vel[3 * p + 2]};
    const int Ng = a.n;
    const double Lg = a.lg;
    int gi = (int)floor((xp - 0.5 * Lg) / Lg);
    int gj = (int)floor((yp - 0.5 * Lg) / Lg);
    int gk = (int)floor((zp - 0.5 * Lg) / Lg);
    // periodic wrap of the lower half cell (float, as the reference stores xp)
    if (gi == -1) { gi = Ng - 1; xp = (float)((double)xp + a.box); }
    if (gj < 0) { gj = Ng - 1; yp = (float)((double)yp + a.box); }
    if (gk < 0) { gk = Ng - 1; zp = (float)((double)zp + a.box); }
    if (gi < 0 || gi >= Ng || gj >= Ng || gk >= Ng) {
        atomicAdd(skipped, 1ull);
        return;
    }
    const double dx = fabs((double)xp - xg_of(gi, Lg)) / Lg;
    const double dy = fabs((double)yp - xg_of(gj, Lg)) / Lg;
    const double dz = fabs((double)zp - xg_of(gk, Lg)) / Lg;
    const double tx = 1. - dx, ty = 1. - dy, tz = 1. - dz;
    int gii = gi + 1, gjj = gj + 1, gkk = gk + 1;
    if (gii == Ng) gii = 0;
    if (gjj == Ng) gjj = 0;
    if (gkk == Ng) gkk = 0;
    const long n = Ng;
    const long c[8] = {gk + n * (gj + gi * n),  gk + n * (gjj + gi * n),  gkk + n * (gj + gi * n),
                       gkk + n * (gjj + gi * n), gk + n * (gjj + gii * n), gk + n * (gj + gii * n),
                       gkk + n * (gj + gii * n), gkk + n * (gjj + gii * n)};
    // weights in the reference's product order ((w_x w_y) w_z)
    const double w[8] = {tx * ty * tz, tx * dy * tz, tx * ty * dz, tx * dy * dz,
                         dx * dy * tz, dx * ty * tz, dx * ty * dz, dx * dy * dz};
    double *vg[3] = {vx, vy, vz};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        atomicAdd(&rho[c[q]], w[q] * a.pmass / a.l3);
#pragma unroll
        for (int d = 0; d < 3; ++d) atomicAdd(&vg[d][c[q]], w[q] * vp[d] / a.l3);
    }
}

// ---------------------------------------------------- reference-order deposit
// One deposit call visits the sequence s = 0 .. ng + np - 1: first the ng
// trailing "left ghost" copies of the chunk (particles np - ng .. np - 1,
// src/gadget_IC.h:445-448 at one rank), then all np particles in file order
// (:634-638).  A particle with base cell (gi, gj, gk) adds to the 8 cells
// (gi + a, gj + b, gk + c), a, b, c in {0, 1} (periodic), with weight
// w = ((wx wy) wz), wx = a ? dx : tx, and so on (:102-150).  Each cell sees
// each sequence entry at most once (N >= 2), so the reference's result is
// fixed by visiting, per cell, its contributions in increasing s.
//
// Device scheme per batch of the sequence: (1) the base cell of every entry
// (k_cico_keys); (2) a stable radix sort of (cell, s) by cell, so each
// cell's entries stay in increasing s; (3) the [first, last) range of each
// cell in the sorted order (k_cico_bounds); (4) one thread per target cell
// merges the ranges of its 8 neighbouring base cells by s and accumulates
// (k_cico_gather).  Batches run one after another, so order holds across
// them too.
struct CicOrd {
    int n;            // mesh points per axis
    double box;       // kpc/h
    double lg;        // box / n
    double hl;        // 0.5 lg
    double pmass;     // Msun/h
    double l3;        // pow(lg, 3), as the host's pow gives it
    const double *xg; // cell centres i lg + lg/2 (host-computed, :333)
    long long np;     // particles of the chunk
    long long ng;     // ghost copies visited first
};

__device__ __forceinline__ long long cico_particle(long long s, const CicOrd &a) {
    return s < a.ng ? a.np - a.ng + s : s - a.ng;
}

// base cell and offsets of a particle (the opening of CIC_scheme,
// :60-98); false when it lies outside [-lg/2, box) in some coordinate
__device__ __forceinline__ bool cico_locate(const float *__restrict__ pos, long long p, const CicOrd &a, int &gi,
                                            int &gj, int &gk, double &dx, double &dy, double &dz) {
#pragma clang fp contract(off)
    float xp = pos[3 * p], yp = pos[3 * p + 1], zp = pos[3 * p + 2];
    const int Ng = a.n;
    const double Lg = a.lg;
    gi = (int)floor((xp - a.hl) / Lg);
    gj = (int)floor((yp - a.hl) / Lg);
    gk = (int)floor((zp - a.hl) / Lg);
    if (gi == -1) { gi = Ng - 1; xp = (float)((double)xp + a.box); }
    if (gj < 0) { gj = Ng - 1; yp = (float)((double)yp + a.box); }
    if (gk < 0) { gk = Ng - 1; zp = (float)((double)zp + a.box); }
    if (gi < 0 || gi >= Ng || gj >= Ng || gk >= Ng) return false;
    dx = fabs((double)xp - a.xg[gi]) / Lg;
    dy = fabs((double)yp - a.xg[gj]) / Lg;
    dz = fabs((double)zp - a.xg[gk]) / Lg;
    return true;
}

// keys[s - s0] = base cell of entry s (n^3 for a skipped particle, which
// sorts last and is never gathered), vals[s - s0] = s
static __global__ __launch_bounds__(256) void k_cico_keys(const float *__restrict__ pos, CicOrd a, long long s0,
                                                   long long m, unsigned *__restrict__ keys,
                                                   unsigned *__restrict__ vals,
                                                   unsigned long long *__restrict__ skipped) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    const long long s = s0 + q;
    int gi, gj, gk;
    double dx, dy, dz;
    unsigned key;
    if (cico_locate(pos, cico_particle(s, a), a, gi, gj, gk, dx, dy, dz)) {
        key = (unsigned)gk + (unsigned)a.n * ((unsigned)gj + (unsigned)a.n * (unsigned)gi);
    } else {
        key = (unsigned)a.n * a.n * a.n;
        atomicAdd(skipped, 1ull);
    }
    keys[q] = key;
    vals[q] = (unsigned)s;
}

// first[c], last[c]: the range of base cell c in the sorted keys (both 0
// for a cell without entries)
static __global__ __launch_bounds__(256) void k_cico_bounds(const unsigned *__restrict__ keys, long long m, unsigned n3,
                                                     unsigned *__restrict__ first, unsigned *__restrict__ last) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    const unsigned k = keys[q];
    if (k >= n3) return;
    if (q == 0 || keys[q - 1] != k) first[k] = (unsigned)q;
    if (q == m - 1 || keys[q + 1] != k) last[k] = (unsigned)(q + 1);
}

// one thread per target cell: its contributions in increasing s, the
// density in FP64 and each velocity component rounded to G after every add
// (G = float: the reference's vector<float> grids)
template <typename G>
__global__ __launch_bounds__(256) void k_cico_gather(const float *__restrict__ pos, const float *__restrict__ vel,
                                                     CicOrd a, const unsigned *__restrict__ vals,
                                                     const unsigned *__restrict__ first,
                                                     const unsigned *__restrict__ last, double *__restrict__ rho,
                                                     G *__restrict__ vx, G *__restrict__ vy, G *__restrict__ vz) {
#pragma clang fp contract(off)
    const long n = a.n, n3 = n * n * n;
    for (long o = (long)blockIdx.x * blockDim.x + threadIdx.x; o < n3; o += (long)gridDim.x * blockDim.x) {
        const int k = (int)(o % n), j = (int)((o / n) % n), i = (int)(o / (n * n));
        unsigned cur[8], lim[8];
        int nonempty = 0;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            // corner (ca, cb, cc): this cell is base + (ca, cb, cc)
            const int ca = c >> 2, cb = (c >> 1) & 1, cc = c & 1;
            const long bi = (i - ca + n) % n, bj = (j - cb + n) % n, bk = (k - cc + n) % n;
            const long b = bk + n * (bj + n * bi);
            cur[c] = first[b];
            lim[c] = last[b];
            nonempty |= (cur[c] < lim[c]) << c;
        }
        if (!nonempty) continue;
        double r = rho[o];
        G u = vx[o], v = vy[o], w = vz[o];
        for (;;) {
            unsigned best = 0xffffffffu;
            int bc = -1;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                if (cur[c] < lim[c]) {
                    const unsigned sv = vals[cur[c]];
                    if (sv < best) { best = sv; bc = c; }
                }
            }
            if (bc < 0) break;
            ++cur[bc];
            const long long p = cico_particle(best, a);
            int gi, gj, gk;
            double dx, dy, dz;
            cico_locate(pos, p, a, gi, gj, gk, dx, dy, dz);
            const double wx = (bc >> 2) ? dx : 1. - dx;
            const double wy = ((bc >> 1) & 1) ? dy : 1. - dy;
            const double wz = (bc & 1) ? dz : 1. - dz;
            const double wt = wx * wy * wz;
            r += wt * a.pmass / a.l3;
            u = (G)((double)u + wt * (double)vel[3 * p] / a.l3);
            v = (G)((double)v + wt * (double)vel[3 * p + 1] / a.l3);
            w = (G)((double)w + wt * (double)vel[3 * p + 2] / a.l3);
        }
        rho[o] = r;
        vx[o] = u;
        vy[o] = v;
        vz[o] = w;
    }
}

// v *= pmass / rho where rho != 0 (src/gadget_IC.h:645-651); for float
// grids the product is taken in FP64 and rounded to float, as
// `vxg[i] *= double` does
template <typename G>
__global__ __launch_bounds__(256) void k_ic_velnorm(const double *__restrict__ rho, G *__restrict__ vx,
                                                    G *__restrict__ vy, G *__restrict__ vz, long n3, double pmass) {
#pragma clang fp contract(off)
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (long)gridDim.x * blockDim.x) {
        const double r = rho[i];
        if (r != 0) {
            const double f = pmass / r;
            vx[i] = (G)((double)vx[i] * f);
            vy[i] = (G)((double)vy[i] * f);
            vz[i] = (G)((double)vz[i] * f);
        }
    }
}

// theta source: 0.5 (dvx + dvy + dvz) * unit / dx with periodic central
// differences (src/gadget_IC.h:190-219 at one rank), as a complex field of
// plane stride ps.  The differences and their sum are taken in the grids'
// type (float dvx, dvy, dvz in the reference, :170), the rest in FP64 in the
// reference's association ((0.5 s) unit) / dx.
template <typename G>
__global__ __launch_bounds__(256) void k_ic_divergence(const G *__restrict__ vx, const G *__restrict__ vy,
                                                       const G *__restrict__ vz, cx<double> *__restrict__ out, int n,
                                                       long ps, double unit, double dx) {
#pragma clang fp contract(off)
    const long n3 = (long)n * n * n;
    for (long o = (long)blockIdx.x * blockDim.x + threadIdx.x; o < n3; o += (long)gridDim.x * blockDim.x) {
        const int k = (int)(o % n), j = (int)((o / n) % n), i = (int)(o / ((long)n * n));
        const int ip = (i + 1) % n, im = (i + n - 1) % n, jp = (j + 1) % n, jm = (j + n - 1) % n;
        const int kp = (k + 1) % n, km = (k + n - 1) % n;
        const long N = n;
        const G dvx = vx[k + N * (j + ip * N)] - vx[k + N * (j + im * N)];
        const G dvy = vy[k + N * (jp + i * N)] - vy[k + N * (jm + i * N)];
        const G dvz = vz[kp + N * (j + i * N)] - vz[km + N * (j + i * N)];
        const G sum = dvx + dvy + dvz;
        out[i * ps + (o - i * N * N)] = mkc<double>(0.5 * (double)sum * unit / dx, 0.0);
    }
}

// psi = 1.225 sqrt(rho uc / m_s) exp(i cph theta), theta complex
// (src/gadget_IC.h:666); psi overwrites theta in place (theta and psi with
// plane stride ps, rho dense)
static __global__ __launch_bounds__(256) void k_ic_assemble(const double *__restrict__ rho, cx<double> *__restrict__ th,
                                                     int n, long ps, double uc, double ms, double cph) {
    const long n2 = (long)n * n, n3 = n2 * n;
    for (long o = (long)blockIdx.x * blockDim.x + threadIdx.x; o < n3; o += (long)gridDim.x * blockDim.x) {
        const long i = (o / n2) * ps + o % n2;
        const cx<double> t = th[i];
        const double amp = 1.225 * sqrt(rho[o] * uc / ms);
        const double m = exp(-cph * t.y), ang = cph * t.x;
        double sn, cs;
        sincos(ang, &sn, &cs);
        th[i] = mkc<double>(amp * m * cs, amp * m * sn);
    }
}

// FP64 field -> storage precision with a power-of-two scale (mixed mode);
// n2 cells per x-plane, planes sps (source) and dps (destination) apart
template <typename T>
__global__ __launch_bounds__(256) void k_store_scaled(const cx<double> *__restrict__ src, cx<T> *__restrict__ dst,
                                                      long n3, long n2, long sps, long dps, int scale_exp) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (long)gridDim.x * blockDim.x) {
        const long x = i / n2, r = i - x * n2;
        const cx<double> z = src[x * sps + r];
        dst[x * dps + r] = mkc<T>((T)ldexp(z.x, -scale_exp), (T)ldexp(z.y, -scale_exp));
    }
}

// partial sums of |z|^2 (FP64) of a dense complex128 array, one per block,
// into part[b] (gridDim.x <= the length of part)
static __global__ __launch_bounds__(256) void k_norm_partial(const cx<double> *__restrict__ src, long n,
                                                      double *__restrict__ part) {
    __shared__ double sb[4];
    double s = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const cx<double> z = src[i];
        s += z.x * z.x + z.y * z.y;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) sb[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sb[0] + sb[1] + sb[2] + sb[3];
}

// storage precision -> FP64 field with the inverse power-of-two scale
static __global__ __launch_bounds__(256) void k_load_scaled(const cx<float> *__restrict__ src, cx<double> *__restrict__ dst,
                                                     long n3, long n2, long sps, long dps, int scale_exp) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (long)gridDim.x * blockDim.x) {
        const long x = i / n2, r = i - x * n2;
        const cx<float> z = src[x * sps + r];
        dst[x * dps + r] = mkc<double>(ldexp((double)z.x, scale_exp), ldexp((double)z.y, scale_exp));
    }
}

}  // namespace ifdm
