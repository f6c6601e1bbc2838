// ic_kernels.h -- initial conditions on the device: cloud-in-cell deposit of
// Gadget particles, mass-weighted velocities, the velocity-divergence source
// of the phase, and the wave-function assembly.
//
// Reference: src/gadget_IC.h (CIC_scheme :45-156, read_gadget :645-670,
// initial_velocity_Poisson_Solver :157-247).  The grids here are FP64 and
// accumulated with device-scope FP64 atomics (the reference accumulates the
// velocity grids in float, :583); the deposit order is therefore not fixed
// and sums agree with a sequential FP64 deposit to round-off.
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

__device__ __forceinline__ double xg_of(int i, double lg) { return i * lg + 0.5 * lg; }

// One thread per particle.  Returns nothing; particles outside
// [-lg/2, box) in some coordinate are counted in *skipped and not deposited.
static __global__ __launch_bounds__(256) void k_cic(const float *__restrict__ pos, const float *__restrict__ vel, long np,
                                             CicArgs a, double *__restrict__ rho, double *__restrict__ vx,
                                             double *__restrict__ vy, double *__restrict__ vz,
                                             unsigned long long *__restrict__ skipped) {
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

// v *= pmass / rho where rho != 0 (src/gadget_IC.h:645-651)
static __global__ __launch_bounds__(256) void k_ic_velnorm(const double *__restrict__ rho, double *__restrict__ vx,
                                                    double *__restrict__ vy, double *__restrict__ vz, long n3,
                                                    double pmass) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (long)gridDim.x * blockDim.x) {
        const double r = rho[i];
        if (r != 0) {
            const double f = pmass / r;
            vx[i] *= f;
            vy[i] *= f;
            vz[i] *= f;
        }
    }
}

// theta source: 0.5 (dvx + dvy + dvz) * unit / dx with periodic central
// differences (src/gadget_IC.h:190-219 at one rank), as a complex field of
// plane stride ps
static __global__ __launch_bounds__(256) void k_ic_divergence(const double *__restrict__ vx, const double *__restrict__ vy,
                                                       const double *__restrict__ vz, cx<double> *__restrict__ out,
                                                       int n, long ps, double scale) {
    const long n3 = (long)n * n * n;
    for (long o = (long)blockIdx.x * blockDim.x + threadIdx.x; o < n3; o += (long)gridDim.x * blockDim.x) {
        const int k = (int)(o % n), j = (int)((o / n) % n), i = (int)(o / ((long)n * n));
        const int ip = (i + 1) % n, im = (i + n - 1) % n, jp = (j + 1) % n, jm = (j + n - 1) % n;
        const int kp = (k + 1) % n, km = (k + n - 1) % n;
        const long N = n;
        const double dvx = vx[k + N * (j + ip * N)] - vx[k + N * (j + im * N)];
        const double dvy = vy[k + N * (jp + i * N)] - vy[k + N * (jm + i * N)];
        const double dvz = vz[kp + N * (j + i * N)] - vz[km + N * (j + i * N)];
        out[i * ps + (o - i * N * N)] = mkc<double>(0.5 * (dvx + dvy + dvz) * scale, 0.0);
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
