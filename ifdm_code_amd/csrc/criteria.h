// criteria.h -- the adaptive time step of Timestep_criteria
// (src/CosmoBEC_MPI_stan.cpp:19-78) and its three scale-factor integrals of
// 1/(H a^n) (src/struct3D.h:127-146), written once for the host
// (ifdm_timestep_criteria, cosmo.cpp) and the device (the step preparation
// kernel of the on-device step loop, ifdm_capi.hip), so both give the same
// bits: integer powers by multiplication, floating contraction off, every
// sum in one fixed order, and only operations IEEE 754 rounds correctly on
// both (+, -, *, /, sqrt).
//
// The integrals follow GSL's gsl_integration_qags (GSL 2.4, the reference's
// call at :66-68: epsabs 0, epsrel 1e-10, limit 1000), a restatement of
// QUADPACK's QAGS: the 21-point Gauss-Kronrod rule QK21 (Piessens et al.
// 1983) with its error estimate, the first evaluation accepted under QAGS's
// test, otherwise globally adaptive bisection of the piece with the largest
// error.  QAGS's epsilon-algorithm extrapolation, which only engages on
// singular integrands, is not carried.  The error estimate's factor
// (200 err / resasc)^1.5 is formed as y sqrt(y).
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#else
#ifndef __host__
#define __host__
#endif
#ifndef __device__
#define __device__
#endif
#ifndef __forceinline__
#define __forceinline__ inline
#endif
#endif

#if defined(__clang__)
#define IFDM_FP_EXACT _Pragma("clang fp contract(off)")
#else
#define IFDM_FP_EXACT
#endif

namespace ifdm {

// 21-point Gauss-Kronrod rule on [-1,1], generated by gen_gk21.py.
// Positive Kronrod nodes, descending, then 0 (odd indices are the Gauss nodes)
#define IFDM_GK21_XGK                                                                                  \
    {9.956571630258080807355273e-1, 9.73906528517171720077964e-1,  9.301574913557082260012072e-1, \
     8.650633666889845107320967e-1, 7.808177265864168970637176e-1, 6.794095682990244062343274e-1, \
     5.627571346686046833390001e-1, 4.333953941292471907992659e-1, 2.943928627014601981311266e-1, \
     1.48874338981631210884826e-1,  0.0}
#define IFDM_GK21_WGK                                                                                  \
    {1.16946388673718742780644e-2,  3.255816230796472747881897e-2, 5.47558965743519960313813e-2,  \
     7.503967481091995276704314e-2, 9.312545458369760553506547e-2, 1.093871588022976418992106e-1, \
     1.234919762620658510779581e-1, 1.34709217311473325928054e-1,  1.427759385770600807970943e-1, \
     1.477391049013384913748415e-1, 1.494455540029169056649365e-1}
// Gauss weights of the 10-point rule for the nodes xgk[1], xgk[3], .., xgk[9]
#define IFDM_GK21_WG                                                                                   \
    {6.667134430868813759356881e-2, 1.494513491505805931457763e-1, 2.190863625159820439955349e-1, \
     2.692667193099963550912269e-1, 2.95524224714752870173893e-1}

struct CritConst {
    double H_0, Omega_m0;
    double ke;       // 4/(3 pi) (m/c^2)/hbar dx^2:  dt_KE = ke a^2        (:32)
    double pe;       // 2 pi hbar/(m/c^2):           dt_PE = pe/max_phi a  (:38)
    double qi;       // 2 pi hbar/|g|:               dt_QI = qi/max_rho a^3 (:41)
    int has_g;       // g != 0 (:42, :55)
};

struct Piece { double a, b, r, e; };

__host__ __device__ __forceinline__ double ipow(double x, int n) {
    IFDM_FP_EXACT
    double r = 1.0;
    for (int i = 0; i < n; ++i) r = r * x;
    return r;
}

// H(x) = H_0 sqrt(Omega_m x^-3 + 1 - Omega_m) and the integrand 1/(H x^n)
__host__ __device__ __forceinline__ double hubble(const CritConst &k, double x) {
    IFDM_FP_EXACT
    return k.H_0 * sqrt(k.Omega_m0 / ipow(x, 3) + (1.0 - k.Omega_m0));
}
__host__ __device__ __forceinline__ double integrand(const CritConst &k, double x, int n) {
    IFDM_FP_EXACT
    return 1. / (hubble(k, x) * ipow(x, n));
}

// the abscissa of node j (0..20) of the rule on [a, b]: j < 10 the left
// nodes centr - hlgth xgk[j], 10 the centre, j > 10 the right nodes
// centr + hlgth xgk[20 - j]
__host__ __device__ __forceinline__ double gk21_node(double a, double b, int j) {
    IFDM_FP_EXACT
    const double xgk[11] = IFDM_GK21_XGK;
    const double centr = 0.5 * (a + b), hlgth = 0.5 * (b - a);
    if (j == 10) return centr;
    return j < 10 ? centr - hlgth * xgk[j] : centr + hlgth * xgk[20 - j];
}

// QK21 from the integrand's values f[j] at gk21_node(a, b, j): result,
// QUADPACK's error estimate, resabs, resasc
__host__ __device__ __forceinline__ double qk21_values(double a, double b, const double *f, double *abserr,
                                                       double *resabs, double *resasc) {
    IFDM_FP_EXACT
    const double wgk[11] = IFDM_GK21_WGK;
    const double wg[5] = IFDM_GK21_WG;
    const double hlgth = 0.5 * (b - a), dhlgth = fabs(hlgth);
    const double fc = f[10];
    double resg = 0.0, resk = wgk[10] * fc, rabs = fabs(resk);
    for (int j = 0; j < 5; ++j) {               // Gauss nodes (odd indices)
        const int jt = 2 * j + 1;
        const double f1 = f[jt], f2 = f[20 - jt];
        resg += wg[j] * (f1 + f2);
        resk += wgk[jt] * (f1 + f2);
        rabs += wgk[jt] * (fabs(f1) + fabs(f2));
    }
    for (int j = 0; j < 5; ++j) {               // Kronrod-only nodes (even indices)
        const int jt = 2 * j;
        const double f1 = f[jt], f2 = f[20 - jt];
        resk += wgk[jt] * (f1 + f2);
        rabs += wgk[jt] * (fabs(f1) + fabs(f2));
    }
    const double reskh = 0.5 * resk;
    double rasc = wgk[10] * fabs(fc - reskh);
    for (int j = 0; j < 10; ++j) rasc += wgk[j] * (fabs(f[j] - reskh) + fabs(f[20 - j] - reskh));
    const double result = resk * hlgth;
    rabs *= dhlgth;
    rasc *= dhlgth;
    double err = fabs((resk - resg) * hlgth);
    if (rasc != 0.0 && err != 0.0) {
        const double y = 200.0 * err / rasc;
        err = rasc * fmin(1.0, y * sqrt(y));
    }
    const double eps = 2.220446049250313080847e-16, dmin = 2.2250738585072013830903e-308;
    if (rabs > dmin / (50.0 * eps)) err = fmax(eps * 50.0 * rabs, err);
    *abserr = err;
    *resabs = rabs;
    *resasc = rasc;
    return result;
}

__host__ __device__ __forceinline__ double qk21(const CritConst &k, double a, double b, int n, double *abserr,
                                                double *resabs, double *resasc) {
    double f[21];
    for (int j = 0; j < 21; ++j) f[j] = integrand(k, gk21_node(a, b, j), n);
    return qk21_values(a, b, f, abserr, resabs, resasc);
}

// QAGS's acceptance test of the first evaluation
__host__ __device__ __forceinline__ bool qags_first_accepted(double r0, double err, double rasc, double epsrel) {
    IFDM_FP_EXACT
    const double tol = epsrel * fabs(r0);
    return (err <= tol && err != rasc) || err == 0.0;
}

// Globally adaptive bisection from the first evaluation (r0, e0) over
// [a0, a1]: pcs holds room for `limit` pieces.  Returns the integral; *ok
// is false when the limit is reached before the tolerance.
__host__ __device__ __forceinline__ double qags_bisect(const CritConst &k, double a0, double a1, int n, double r0,
                                                       double e0, double epsrel, int limit, Piece *pcs, bool *ok) {
    IFDM_FP_EXACT
    int np = 1;
    pcs[0].a = a0; pcs[0].b = a1; pcs[0].r = r0; pcs[0].e = e0;
    double total = r0, terr = e0;
    int it = 1;
    for (; it < limit && terr > epsrel * fabs(total); ++it) {
        int mx = 0;                              // the first piece of largest error
        for (int i = 1; i < np; ++i)
            if (pcs[i].e > pcs[mx].e) mx = i;
        const Piece p = pcs[mx];
        const double m = 0.5 * (p.a + p.b);
        double e1, e2, t1, t2;
        const double r1 = qk21(k, p.a, m, n, &e1, &t1, &t2);
        const double r2 = qk21(k, m, p.b, n, &e2, &t1, &t2);
        pcs[mx].a = p.a; pcs[mx].b = m; pcs[mx].r = r1; pcs[mx].e = e1;
        pcs[np].a = m; pcs[np].b = p.b; pcs[np].r = r2; pcs[np].e = e2;
        ++np;
        total = 0.0;
        terr = 0.0;
        for (int i = 0; i < np; ++i) { total += pcs[i].r; terr += pcs[i].e; }
    }
    *ok = terr <= epsrel * fabs(total);
    return total;
}

// The step's time-step sizes and scale-factor increment (:32-61):
// dt = min(dt_KE, dt_PE, dt_QI if g != 0), da = H(a) a dt; *which = 3, 2, 4
// for the limiting criterion (KE, PE, QI)
__host__ __device__ __forceinline__ double criteria_da(const CritConst &k, double a, double max_phi, double max_rho,
                                                       int *which) {
    IFDM_FP_EXACT
    const double dt_KE = k.ke * (a * a);
    const double dt_PE = k.pe / max_phi * a;
    double dt_QI = k.qi / max_rho * (a * a * a);
    if (!k.has_g) dt_QI = 0.0;
    double max_dt = dt_KE;
    int w = 3;
    if (max_dt > dt_PE) { max_dt = dt_PE; w = 2; }
    if ((max_dt > dt_QI) && k.has_g) { max_dt = dt_QI; w = 4; }
    if (which) *which = w;
    return hubble(k, a) * a * max_dt;
}

}  // namespace ifdm
