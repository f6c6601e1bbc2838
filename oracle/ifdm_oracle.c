/*
 * ifdm_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C, FP64 restatement of the reference's Schrodinger-Poisson
 * kick-drift-kick step (jchan192/IFDM_Code, src/CosmoBEC_MPI_stan.cpp and
 * src/struct3D.h), used as the correctness oracle for the HIP library and,
 * timed on the host cores, as the CPU baseline.  Nothing on the product path
 * links or calls this file: only tests/, the smoke check and the bench's
 * cpu_baseline leg load liboracle.so.
 *
 * Layout follows the reference exactly (one rank, P = 1):
 *   psi   N^3 complex<double> interleaved, offset k + N*(j + N*i)   (:133)
 *   phi   N*N*(2*(N/2+1)) doubles, FFTW padded real layout        (:88)
 *   phikk N*N*(N/2+1) complex<double> half spectrum               (:102)
 *
 * The FFTs are this file's own radix-2 transforms (FFTW conventions:
 * forward e^{-2 pi i nk/N}, backward e^{+2 pi i nk/N}, both unnormalised).
 * FFTW itself is not available here; the transforms are pinned by the tests
 * against a brute-force DFT and against numpy's pocketfft.
 *
 * Parity status of each function is listed in the oracle header of
 * oracle/oracle.py.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cd;

/* ---------------------------------------------------------------- FFT core */

static int ilog2i(int n) { int l = 0; while ((1 << l) < n) ++l; return l; }

/* twiddles w[k] = exp(sign * 2 pi i k / n), k < n/2 */
static cd *make_twiddles(int n, int sign) {
    cd *w = (cd *)malloc(sizeof(cd) * (size_t)(n / 2 + 1));
    for (int k = 0; k <= n / 2; ++k) {
        double a = 2.0 * M_PI * (double)k / (double)n;
        w[k].re = cos(a);
        w[k].im = sign * sin(a);
    }
    return w;
}

/* in-place iterative radix-2 decimation-in-time FFT of one contiguous line */
static void fft1d(cd *x, int n, const cd *w) {
    int lg = ilog2i(n);
    for (int i = 0; i < n; ++i) {
        int r = 0;
        for (int b = 0; b < lg; ++b) r |= ((i >> b) & 1) << (lg - 1 - b);
        if (r > i) { cd t = x[i]; x[i] = x[r]; x[r] = t; }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, step = n / len;
        for (int s = 0; s < n; s += len) {
            for (int q = 0; q < half; ++q) {
                cd ww = w[q * step];
                cd a = x[s + q], b = x[s + q + half];
                cd t = { b.re * ww.re - b.im * ww.im, b.re * ww.im + b.im * ww.re };
                x[s + q].re = a.re + t.re;        x[s + q].im = a.im + t.im;
                x[s + q + half].re = a.re - t.re; x[s + q + half].im = a.im - t.im;
            }
        }
    }
}

/* transform every line along one axis of a complex array of shape
 * (n0, n1, n2) with n2 contiguous; axis in {0,1,2}; `len` = extent along it */
static void fft_axis(cd *a, int n0, int n1, int n2, int axis, int sign) {
    int len = axis == 0 ? n0 : (axis == 1 ? n1 : n2);
    cd *w = make_twiddles(len, sign);
    long nlines = (long)n0 * n1 * n2 / len;
    #pragma omp parallel
    {
        cd *buf = (cd *)malloc(sizeof(cd) * (size_t)len);
        #pragma omp for schedule(static)
        for (long l = 0; l < nlines; ++l) {
            long base, stride;
            if (axis == 2) { base = l * n2; stride = 1; }
            else if (axis == 1) { long i = l / n2, k = l % n2; base = i * (long)n1 * n2 + k; stride = n2; }
            else { base = l; stride = (long)n1 * n2; }
            for (int m = 0; m < len; ++m) buf[m] = a[base + m * stride];
            fft1d(buf, len, w);
            for (int m = 0; m < len; ++m) a[base + m * stride] = buf[m];
        }
        free(buf);
    }
    free(w);
}

/* 3-D complex-to-complex, in place, unnormalised (fftw_mpi_plan_dft_3d) */
void oracle_fft3d_c2c(double *a, int n, int sign) {
    cd *c = (cd *)a;
    fft_axis(c, n, n, n, 2, sign);
    fft_axis(c, n, n, n, 1, sign);
    fft_axis(c, n, n, n, 0, sign);
}

/* 3-D real-to-complex from the padded real layout (row length 2*(n/2+1))
 * into the half spectrum n*n*(n/2+1) (fftw_mpi_plan_dft_r2c_3d). */
void oracle_fft3d_r2c(const double *in_pad, double *out_half, int n) {
    int h = n / 2 + 1, rp = 2 * h;
    cd *o = (cd *)out_half;
    cd *w = make_twiddles(n, -1);
    #pragma omp parallel
    {
        cd *buf = (cd *)malloc(sizeof(cd) * (size_t)n);
        #pragma omp for schedule(static)
        for (long row = 0; row < (long)n * n; ++row) {
            for (int k = 0; k < n; ++k) { buf[k].re = in_pad[row * rp + k]; buf[k].im = 0.0; }
            fft1d(buf, n, w);
            for (int k = 0; k < h; ++k) o[row * h + k] = buf[k];
        }
        free(buf);
    }
    free(w);
    fft_axis(o, n, n, h, 1, -1);
    fft_axis(o, n, n, h, 0, -1);
}

/* 3-D complex-to-real into the padded real layout; the half spectrum is
 * consumed (overwritten) as FFTW's c2r does. Only Re of the k=0 and k=n/2
 * entries of each z-line enter, as in a Hermitian c2r. */
void oracle_fft3d_c2r(double *in_half, double *out_pad, int n) {
    int h = n / 2 + 1, rp = 2 * h;
    cd *c = (cd *)in_half;
    fft_axis(c, n, n, h, 0, +1);
    fft_axis(c, n, n, h, 1, +1);
    cd *w = make_twiddles(n, +1);
    #pragma omp parallel
    {
        cd *buf = (cd *)malloc(sizeof(cd) * (size_t)n);
        #pragma omp for schedule(static)
        for (long row = 0; row < (long)n * n; ++row) {
            for (int k = 0; k < h; ++k) buf[k] = c[row * h + k];
            for (int k = h; k < n; ++k) { buf[k].re = c[row * h + (n - k)].re; buf[k].im = -c[row * h + (n - k)].im; }
            fft1d(buf, n, w);
            for (int k = 0; k < n; ++k) out_pad[row * rp + k] = buf[k].re;
        }
        free(buf);
    }
    free(w);
}

/* ------------------------------------------------------------- parameters */

/* Physical parameters of one run, restated from src/Cosmoinput3D.h:2-17 and
 * src/CosmoBEC_MPI_stan.cpp:329-332.  Filled by oracle.py. */
typedef struct {
    int n;
    double xmax;        /* Mpc, already divided by h_0 (:332) */
    double mfdm;        /* eV */
    double mfdm_solar;  /* Msun (:330) */
    double g;           /* eV Mpc^3 (:331) */
    double rhoc;        /* Msun / Mpc^3 (par.rhoc) */
    double hbar, c, bigG;
} oracle_params;

/* k-vector entry, src/struct3D.h:76-85 */
static double kvec(int i, int n, double xmax) {
    if (i < n / 2) return i * M_PI / xmax;
    return ((double)i - n) * M_PI / xmax;
}

/* ------------------------------------------------------------- hot path */

/* poisson(), src/CosmoBEC_MPI_stan.cpp:80-119.  phi: padded real layout.
 * phikk: half-spectrum scratch of n*n*(n/2+1) complex. */
void oracle_poisson(const oracle_params *p, const double *psi, double *phi, double *phikk) {
    int n = p->n, h = n / 2 + 1, rp = 2 * h;
    const cd *w = (const cd *)psi;
    cd *pk = (cd *)phikk;
    double fftnorm = 1. / (double)pow(n, 3);
    #pragma omp parallel for schedule(static)
    for (long row = 0; row < (long)n * n; ++row)
        for (int k = 0; k < n; ++k) {
            cd z = w[row * n + k];
            double a = hypot(z.re, z.im);
            phi[row * rp + k] = p->mfdm_solar * pow(a, 2) - p->rhoc;          /* :88 */
        }
    oracle_fft3d_r2c(phi, phikk, n);                                            /* :94 */
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < h; ++k) {
                double kx = kvec(i, n, p->xmax), ky = kvec(j, n, p->xmax), kz = kvec(k, n, p->xmax);
                long o = k + (long)h * (j + (long)i * n);
                double k2 = pow(kx, 2) + pow(ky, 2) + pow(kz, 2);
                if (k2 == 0) { pk[o].re = 0.0; pk[o].im = 0.0; }                  /* :101-102 */
                else {
                    double f = -4.0 * M_PI * p->bigG;                           /* :105 */
                    pk[o].re = f * pk[o].re / k2;
                    pk[o].im = f * pk[o].im / k2;
                }
            }
    oracle_fft3d_c2r(phikk, phi, n);                                            /* :111 */
    #pragma omp parallel for schedule(static)
    for (long row = 0; row < (long)n * n; ++row)
        for (int k = 0; k < rp; ++k) {
            if (k < n) phi[row * rp + k] *= fftnorm;                            /* :116 */
            else phi[row * rp + k] = 0.0;                                       /* :117 */
        }
}

/* one half-kick, src/CosmoBEC_MPI_stan.cpp:130-134 (and :148-152) */
void oracle_kick(const oracle_params *p, double *psi, const double *phi, double dPE, double dQI) {
    int n = p->n, rp = 2 * (n / 2 + 1);
    cd *w = (cd *)psi;
    #pragma omp parallel for schedule(static)
    for (long row = 0; row < (long)n * n; ++row)
        for (int k = 0; k < n; ++k) {
            cd z = w[row * n + k];
            double a = hypot(z.re, z.im);
            /* dt_PE*phi*mfdm/c^2/hbar + dt_QI*g/hbar*|psi|^2, left-to-right as written */
            double th = -0.5 * (dPE * phi[row * rp + k] * p->mfdm / pow(p->c, 2) / p->hbar
                                + dQI * p->g / p->hbar * pow(a, 2));
            double cs = cos(th), sn = sin(th);
            w[row * n + k].re = z.re * cs - z.im * sn;
            w[row * n + k].im = z.re * sn + z.im * cs;
        }
}

/* drift: forward c2c, phase exp(-i KE_phase dt_KE) * fftnorm, backward c2c.
 * src/CosmoBEC_MPI_stan.cpp:136-144 with KE_phase from src/struct3D.h:121 */
void oracle_drift(const oracle_params *p, double *psi, double dKE) {
    int n = p->n;
    cd *w = (cd *)psi;
    double fftnorm = 1. / (double)pow(n, 3);
    oracle_fft3d_c2c(psi, n, -1);
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k) {
                double kx = kvec(i, n, p->xmax), ky = kvec(j, n, p->xmax), kz = kvec(k, n, p->xmax);
                double ke = 0.5 * p->hbar / (p->mfdm / pow(p->c, 2)) * (pow(kx, 2) + pow(ky, 2) + pow(kz, 2));
                double th = -ke * dKE;
                double cs = cos(th) * fftnorm, sn = sin(th) * fftnorm;
                long o = k + (long)n * (j + (long)n * i);
                cd z = w[o];
                w[o].re = z.re * cs - z.im * sn;
                w[o].im = z.re * sn + z.im * cs;
            }
    oracle_fft3d_c2c(psi, n, +1);
}

/* Split_op_MPI, src/CosmoBEC_MPI_stan.cpp:121-159: K(1/2) D P K(1/2) */
void oracle_split_op(const oracle_params *p, double *psi, double *phi, double *phikk,
                     double dKE, double dPE, double dQI) {
    oracle_kick(p, psi, phi, dPE, dQI);
    oracle_drift(p, psi, dKE);
    oracle_poisson(p, psi, phi, phikk);
    oracle_kick(p, psi, phi, dPE, dQI);
}

/* The two scans of Timestep_criteria, src/CosmoBEC_MPI_stan.cpp:26-27.
 * ref_range != 0 scans the first n^3 entries of the padded phi array, the
 * range std::max_element covers in the reference at one rank (phi.size()
 * is n*n*local_n0 while the array is indexed with row length 2*(n/2+1));
 * ref_range == 0 scans every real-space entry. out[0]=max|phi|,
 * out[1]=max|psi|^2, out[2]=sum|psi|^2. */
void oracle_maxima(const oracle_params *p, const double *psi, const double *phi, int ref_range, double *out) {
    int n = p->n, rp = 2 * (n / 2 + 1);
    long n3 = (long)n * n * n;
    double mp = 0.0, mr = 0.0, s = 0.0;
    if (ref_range) {
        for (long f = 0; f < n3; ++f) { double v = fabs(phi[f]); if (v > mp) mp = v; }
    } else {
        for (long row = 0; row < (long)n * n; ++row)
            for (int k = 0; k < n; ++k) { double v = fabs(phi[row * rp + k]); if (v > mp) mp = v; }
    }
    const cd *w = (const cd *)psi;
    double ma = 0.0;
    for (long o = 0; o < n3; ++o) {
        double a = hypot(w[o].re, w[o].im);
        if (a > ma) ma = a;
        s += w[o].re * w[o].re + w[o].im * w[o].im;
    }
    mr = pow(ma, 2);
    out[0] = mp; out[1] = mr; out[2] = s;
}

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void oracle_set_threads(int t) {
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#else
    (void)t;
#endif
}

/* ------------------------------------------------------------- IC path */

/* The reference accumulates the three velocity grids in float
 * (vxg, vyg, vzg are std::vector<float>, src/gadget_IC.h:583) and forms the
 * divergence in float (:170).  The oracle keeps them in double and, with
 * f32 != 0, rounds every stored value to float exactly where the reference
 * stores into a float, reproducing its arithmetic. */
static inline double st32(double v, int f32) { return f32 ? (double)(float)v : v; }

/* CIC_scheme, src/gadget_IC.h:45-156, at one rank (my_rank = 0,
 * local_n0 = N).  rho: N^3 doubles (Re of opr.wfc).  Positions in kpc/h,
 * box in kpc/h, pmass in Msun/h (header mass[1] * 1e10, :317), xg[i] =
 * i*Lg + Lg/2 (:333).  Returns 0, or 1 when the particle lies outside
 * [-Lg/2, box) in some coordinate and is skipped (the reference would index
 * out of bounds there). */
static int cic_one(int Ng, double box, double pmass, double Lg, const double *xg,
                   float xp, float yp, float zp, float vxp, float vyp, float vzp,
                   double *rho, double *vx, double *vy, double *vz, int f32) {
    int gi = (int)floor((xp - 0.5 * Lg) / Lg);
    int gj = (int)floor((yp - 0.5 * Lg) / Lg);
    int gk = (int)floor((zp - 0.5 * Lg) / Lg);
    if (gi == -1) { gi = Ng - 1; xp += box; }
    if (gj < 0) { gj = Ng - 1; yp += box; }
    if (gk < 0) { gk = Ng - 1; zp += box; }
    if (gi < 0 || gi >= Ng || gj >= Ng || gk >= Ng) return 1;
    double dx = fabs((double)xp - xg[gi]) / Lg;
    double dy = fabs((double)yp - xg[gj]) / Lg;
    double dz = fabs((double)zp - xg[gk]) / Lg;
    double tx = 1. - dx, ty = 1. - dy, tz = 1. - dz;
    int gii = gi + 1, gjj = gj + 1, gkk = gk + 1;
    if (gii == Ng) gii = 0;
    if (gjj == Ng) gjj = 0;
    if (gkk == Ng) gkk = 0;
    double L3 = pow(Lg, 3);
    double *vs[3] = {vx, vy, vz};
    float vp[3] = {vxp, vyp, vzp};
#define IDX(i, j, k) ((long)(k) + (long)Ng * ((long)(j) + (long)(i) * Ng))
    /* the gi planes, then the gii planes, in the reference's order */
    {
        long o1 = IDX(gi, gj, gk), o2 = IDX(gi, gjj, gk), o3 = IDX(gi, gj, gkk), o4 = IDX(gi, gjj, gkk);
        rho[o1] += tx * ty * tz * pmass / L3;
        rho[o2] += tx * dy * tz * pmass / L3;
        rho[o3] += tx * ty * dz * pmass / L3;
        rho[o4] += tx * dy * dz * pmass / L3;
        for (int c = 0; c < 3; ++c) {
            vs[c][o1] = st32(vs[c][o1] + tx * ty * tz * vp[c] / L3, f32);
            vs[c][o2] = st32(vs[c][o2] + tx * dy * tz * vp[c] / L3, f32);
            vs[c][o3] = st32(vs[c][o3] + tx * ty * dz * vp[c] / L3, f32);
            vs[c][o4] = st32(vs[c][o4] + tx * dy * dz * vp[c] / L3, f32);
        }
    }
    {
        long o1 = IDX(gii, gjj, gk), o2 = IDX(gii, gj, gk), o3 = IDX(gii, gj, gkk), o4 = IDX(gii, gjj, gkk);
        rho[o1] += dx * dy * tz * pmass / L3;
        rho[o2] += dx * ty * tz * pmass / L3;
        rho[o3] += dx * ty * dz * pmass / L3;
        rho[o4] += dx * dy * dz * pmass / L3;
        for (int c = 0; c < 3; ++c) {
            vs[c][o1] = st32(vs[c][o1] + dx * dy * tz * vp[c] / L3, f32);
            vs[c][o2] = st32(vs[c][o2] + dx * ty * tz * vp[c] / L3, f32);
            vs[c][o3] = st32(vs[c][o3] + dx * ty * dz * vp[c] / L3, f32);
            vs[c][o4] = st32(vs[c][o4] + dx * dy * dz * vp[c] / L3, f32);
        }
    }
#undef IDX
    return 0;
}

/* Deposit one IC chunk of np particles (pos, vel: 3*np floats) the way
 * read_gadget does at one rank (src/gadget_IC.h:610-642).  ref_p1 != 0
 * also deposits the rank's "left ghost" copies: the trailing particles of
 * the x-sorted chunk with x > xg[N-1] (Pre_DistributeParticles_withGhost,
 * :445-448), which at one rank are thus deposited twice.  The extra particle
 * the reference reads past the end of the chunk (:431-432) is not
 * reproduced.  Returns the number of skipped particles. */
long oracle_cic_chunk(int Ng, double box, double pmass, const float *pos, const float *vel, long np,
                      int ref_p1, int f32, double *rho, double *vx, double *vy, double *vz) {
    double Lg = box / Ng;
    double *xg = (double *)malloc(sizeof(double) * Ng);
    long skipped = 0;
    for (int i = 0; i < Ng; ++i) xg[i] = i * Lg + 0.5 * Lg;
    if (ref_p1) {
        long ng = 0;
        for (long pn = np - 1; pn > 0; --pn) {
            if (xg[Ng - 1] < pos[3 * pn]) ++ng;
            else break;
        }
        for (long q = np - ng; q < np; ++q)
            skipped += cic_one(Ng, box, pmass, Lg, xg, pos[3 * q], pos[3 * q + 1], pos[3 * q + 2],
                               vel[3 * q], vel[3 * q + 1], vel[3 * q + 2], rho, vx, vy, vz, f32);
    }
    for (long q = 0; q < np; ++q)
        skipped += cic_one(Ng, box, pmass, Lg, xg, pos[3 * q], pos[3 * q + 1], pos[3 * q + 2],
                           vel[3 * q], vel[3 * q + 1], vel[3 * q + 2], rho, vx, vy, vz, f32);
    free(xg);
    return skipped;
}

/* After all chunks: mass-weighted velocity (src/gadget_IC.h:645-651), the
 * velocity-divergence phase (initial_velocity_Poisson_Solver, :157-247) and
 * the wave function (:659-670).  psi (out) N^3 complex; theta_out (may be
 * NULL) N^3 complex, the phase field opr.phik.  a: scale factor of the IC
 * header; dx: cell size in Mpc (par.dx); h0, m2Mpc, s2Myr: constants of
 * src/Cosmoinput3D.h. */
void oracle_ic_assemble(const oracle_params *p, double pmass, double a, double dx, double h0,
                        double m2Mpc_, double s2Myr_, int f32, double *rho, double *vx, double *vy, double *vz,
                        double *psi, double *theta_out) {
    int N = p->n;
    long n3 = (long)N * N * N;
    for (long i = 0; i < n3; ++i) {
        if (rho[i] != 0) {
            vx[i] = st32(vx[i] * (pmass / rho[i]), f32);
            vy[i] = st32(vy[i] * (pmass / rho[i]), f32);
            vz[i] = st32(vz[i] * (pmass / rho[i]), f32);
        }
    }
    cd *th = (cd *)malloc(sizeof(cd) * (size_t)n3);
    double unit = 1e3 / m2Mpc_ * s2Myr_;                         /* :171 */
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                int ip = (i + 1) % N, im = (i + N - 1) % N;     /* ghost planes at one rank */
                int jp = (j + 1) % N, jm = (j + N - 1) % N;
                int kp = (k + 1) % N, km = (k + N - 1) % N;
                double dvx = st32(vx[k + (long)N * (j + (long)ip * N)] - vx[k + (long)N * (j + (long)im * N)], f32);
                double dvy = st32(vy[k + (long)N * (jp + (long)i * N)] - vy[k + (long)N * (jm + (long)i * N)], f32);
                double dvz = st32(vz[kp + (long)N * (j + (long)i * N)] - vz[km + (long)N * (j + (long)i * N)], f32);
                double sum = st32(st32(dvx + dvy, f32) + dvz, f32);
                long o = k + (long)N * (j + (long)i * N);
                th[o].re = 0.5 * sum * unit / dx;                /* :216 */
                th[o].im = 0.0;
            }
    fft_axis(th, N, N, N, 2, -1);
    fft_axis(th, N, N, N, 1, -1);
    fft_axis(th, N, N, N, 0, -1);
    double a32 = pow(a, 3. / 2);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                double kx = kvec(i, N, p->xmax), ky = kvec(j, N, p->xmax), kz = kvec(k, N, p->xmax);
                double k2 = pow(kx, 2) + pow(ky, 2) + pow(kz, 2);
                long o = k + (long)N * (j + (long)i * N);
                if (k2 == 0) { th[o].re = 0.0; th[o].im = 0.0; }
                else { th[o].re = -a32 * th[o].re / k2; th[o].im = -a32 * th[o].im / k2; }   /* :233 */
            }
    fft_axis(th, N, N, N, 2, +1);
    fft_axis(th, N, N, N, 1, +1);
    fft_axis(th, N, N, N, 0, +1);
    double nn = (double)((long)N * N * N);
    for (long i = 0; i < n3; ++i) { th[i].re /= nn; th[i].im /= nn; }   /* :241-244 */
    double uc = pow(h0, 2) * pow(1e3, 3);                               /* :659 */
    double cph = p->mfdm / pow(p->c, 2) / p->hbar;
    cd *ps = (cd *)psi;
    for (long i = 0; i < n3; ++i) {
        /* 1.225 sqrt(rho u / m_s) exp(c theta i) with complex theta (:666):
         * exp(i c (re + i im)) = exp(-c im) (cos(c re) + i sin(c re)) */
        double amp = 1.225 * sqrt(rho[i] * uc / p->mfdm_solar);
        double m = exp(-cph * th[i].im), ang = cph * th[i].re;
        ps[i].re = amp * m * cos(ang);
        ps[i].im = amp * m * sin(ang);
    }
    if (theta_out) memcpy(theta_out, th, sizeof(cd) * (size_t)n3);
    free(th);
}
