// cosmo.cpp -- host-side physics of the KDK solver: the unit constants, the
// scale-factor integrals and the adaptive time-step criteria.
//
// Constants: src/Cosmoinput3D.h:2-17 of the reference (jchan192/IFDM_Code).
// Time step: Timestep_criteria, src/CosmoBEC_MPI_stan.cpp:19-78.
// Integrals: the reference calls gsl_integration_qags (GSL 2.4) with
// epsabs = 0, epsrel = 1e-10, limit 1000 on 1/(H a^n) (src/struct3D.h:127-146).
// This file integrates with its own 21-point Gauss-Kronrod rule (QUADPACK's
// QK21, Piessens et al. 1983) and globally adaptive bisection; the first
// evaluation is accepted under the same test QAGS applies, which is the case
// for every step interval of this smooth integrand.  QAGS's epsilon-algorithm
// extrapolation, which only engages on singular integrands, is not carried.
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>
#include <algorithm>

#include "../../include/ifdm.h"
#include "constants.h"

namespace ifdm {

// ---- src/Cosmoinput3D.h:2-17
extern const double m2kpc = 3.086e19;
extern const double m2Mpc = 3.086e22;
extern const double cm2Mpc = 3.086e24;
extern const double s2Myr = 60.0 * 60.0 * 24.0 * 365.0 * 1e6;
extern const double kg2eV = 1.783e-36;
extern const double kg2solar = 1.989e30;
extern const double rhoc = 3.51059e+10;
extern const double h_0 = 0.677;
extern const double H_0 = h_0 * 100 * s2Myr / m2Mpc * 1e3;
extern const double hbar = 6.58211928e-16 / s2Myr;
extern const double Omega_m0 = 0.276;
extern const double bigG = 6.674e-11 * kg2solar / std::pow(m2Mpc, 3) * std::pow(s2Myr, 2);
extern const double c_light = 2.99792458e8 / m2Mpc * s2Myr;

// 21-point Gauss-Kronrod rule on [-1,1], generated by gen_gk21.py
// xgk: positive Kronrod nodes, descending, then 0 (odd indices are the Gauss nodes)
static const double xgk[11] = {
    9.956571630258080807355273e-1,
    9.73906528517171720077964e-1,
    9.301574913557082260012072e-1,
    8.650633666889845107320967e-1,
    7.808177265864168970637176e-1,
    6.794095682990244062343274e-1,
    5.627571346686046833390001e-1,
    4.333953941292471907992659e-1,
    2.943928627014601981311266e-1,
    1.48874338981631210884826e-1,
    0.0,
};
static const double wgk[11] = {
    1.16946388673718742780644e-2,
    3.255816230796472747881897e-2,
    5.47558965743519960313813e-2,
    7.503967481091995276704314e-2,
    9.312545458369760553506547e-2,
    1.093871588022976418992106e-1,
    1.234919762620658510779581e-1,
    1.34709217311473325928054e-1,
    1.427759385770600807970943e-1,
    1.477391049013384913748415e-1,
    1.494455540029169056649365e-1,
};
// Gauss weights of the 10-point rule for the nodes xgk[1], xgk[3], .., xgk[9]
static const double wg[5] = {
    6.667134430868813759356881e-2,
    1.494513491505805931457763e-1,
    2.190863625159820439955349e-1,
    2.692667193099963550912269e-1,
    2.95524224714752870173893e-1,
};

// 1/(H(x) x^n), src/struct3D.h:127-146
static double Ha(double x, int n) {
    double H = H_0 * std::sqrt(Omega_m0 / std::pow(x, 3) + (1.0 - Omega_m0));
    return 1. / (H * std::pow(x, n));
}

struct Piece { double a, b, r, e; };

// QK21 with QUADPACK's error estimate; returns result, sets abserr, resabs, resasc
static double qk21(double a, double b, int n, double *abserr, double *resabs, double *resasc) {
    const double centr = 0.5 * (a + b), hlgth = 0.5 * (b - a), dhlgth = std::fabs(hlgth);
    double fv1[10], fv2[10];
    const double fc = Ha(centr, n);
    double resg = 0.0, resk = wgk[10] * fc, rabs = std::fabs(resk);
    for (int j = 0; j < 5; ++j) {               // Gauss nodes (odd indices)
        const int jt = 2 * j + 1;
        const double absc = hlgth * xgk[jt];
        const double f1 = Ha(centr - absc, n), f2 = Ha(centr + absc, n);
        fv1[jt] = f1; fv2[jt] = f2;
        resg += wg[j] * (f1 + f2);
        resk += wgk[jt] * (f1 + f2);
        rabs += wgk[jt] * (std::fabs(f1) + std::fabs(f2));
    }
    for (int j = 0; j < 5; ++j) {               // Kronrod-only nodes (even indices)
        const int jt = 2 * j;
        const double absc = hlgth * xgk[jt];
        const double f1 = Ha(centr - absc, n), f2 = Ha(centr + absc, n);
        fv1[jt] = f1; fv2[jt] = f2;
        resk += wgk[jt] * (f1 + f2);
        rabs += wgk[jt] * (std::fabs(f1) + std::fabs(f2));
    }
    const double reskh = 0.5 * resk;
    double rasc = wgk[10] * std::fabs(fc - reskh);
    for (int j = 0; j < 10; ++j) rasc += wgk[j] * (std::fabs(fv1[j] - reskh) + std::fabs(fv2[j] - reskh));
    const double result = resk * hlgth;
    rabs *= dhlgth;
    rasc *= dhlgth;
    double err = std::fabs((resk - resg) * hlgth);
    if (rasc != 0.0 && err != 0.0) err = rasc * std::min(1.0, std::pow(200.0 * err / rasc, 1.5));
    const double eps = std::numeric_limits<double>::epsilon();
    if (rabs > std::numeric_limits<double>::min() / (50.0 * eps)) err = std::max(eps * 50.0 * rabs, err);
    *abserr = err; *resabs = rabs; *resasc = rasc;
    return result;
}

double scale_integral(double a0, double a1, int n, double epsrel, int limit) {
    double err, rabs, rasc;
    const double r0 = qk21(a0, a1, n, &err, &rabs, &rasc);
    const double tol = epsrel * std::fabs(r0);
    // QAGS's acceptance test for the first evaluation
    if ((err <= tol && err != rasc) || err == 0.0) return r0;
    std::vector<Piece> pcs{{a0, a1, r0, err}};
    double total = r0, terr = err;
    for (int it = 1; it < limit && terr > epsrel * std::fabs(total); ++it) {
        auto mx = std::max_element(pcs.begin(), pcs.end(), [](const Piece &x, const Piece &y) { return x.e < y.e; });
        Piece p = *mx;
        const double m = 0.5 * (p.a + p.b);
        double e1, e2, t1, t2;
        const double r1 = qk21(p.a, m, n, &e1, &t1, &t2);
        const double r2 = qk21(m, p.b, n, &e2, &t1, &t2);
        *mx = Piece{p.a, m, r1, e1};
        pcs.push_back(Piece{m, p.b, r2, e2});
        total = 0.0; terr = 0.0;
        for (const Piece &q : pcs) { total += q.r; terr += q.e; }
    }
    return total;
}

}  // namespace ifdm

using namespace ifdm;

extern "C" double ifdm_scale_integral(double a0, double a1, int power) {
    return scale_integral(a0, a1, power, 1e-10, 1000);
}

extern "C" int ifdm_timestep_criteria(const ifdm_config *cfg, double a, const ifdm_stats *st,
                                      ifdm_dt *dt_out, double *a_next, int *which_out) {
    if (!cfg || !st || !dt_out || !a_next || !(a > 0.0) || cfg->n <= 0) return IFDM_E_ARG;
    const double mfdm = cfg->mfdm, g = cfg->g;
    const double dx = 2.0 * cfg->xmax / cfg->n;                                       // struct3D.h:52
    const double dt_KE = 4. / (3. * M_PI) * (mfdm / std::pow(c_light, 2)) / hbar * std::pow(dx, 2) * std::pow(a, 2);  // :32
    const double dt_PE = 2. * M_PI * hbar / (mfdm / std::pow(c_light, 2)) / st->max_phi * a;                     // :38
    double dt_QI = 2. * M_PI * hbar / std::fabs(g) / st->max_rho * std::pow(a, 3);                               // :41
    if (g == 0) dt_QI = 0.0;                                                                                      // :42
    double max_dt = dt_KE;                                                                                        // :49-58
    int which = 3;
    if (max_dt > dt_PE) { max_dt = dt_PE; which = 2; }
    if ((max_dt > dt_QI) && (g != 0)) { max_dt = dt_QI; which = 4; }
    const double H = H_0 * std::sqrt(Omega_m0 / std::pow(a, 3) + (1.0 - Omega_m0));                             // :61
    const double da = H * a * max_dt;
    dt_out->dKE = scale_integral(a, a + da, 3, 1e-10, 1000);                                                                   // :66
    dt_out->dPE = scale_integral(a, a + da, 2, 1e-10, 1000);                                                                   // :67
    dt_out->dQI = scale_integral(a, a + da, 4, 1e-10, 1000);                                                                   // :68
    *a_next = a + da;
    if (which_out) *which_out = which;
    return IFDM_OK;
}

// The snapshot trigger chain of Simulation (src/CosmoBEC_MPI_stan.cpp:238-284;
// variant 1: src/CosmoBEC_MPI.cpp, whose chain differs at :250 and :256-261).
// Evaluated after Timestep_criteria has advanced a (par.ai).
extern "C" int ifdm_output_trigger(int variant, int a_triggers, double a, int tn, int twrite, int t_stop,
                                   ifdm_trigger_state *st, int *t_out, int *stop) {
    if (!st || !t_out || !stop || twrite <= 0) return -1;
    *stop = 0;
    *t_out = -1;
    if (a_triggers) {
        if (a > 1) {                                                   // :238
            *t_out = 999999;
            st->write_num++;
            *stop = 1;
            return 1;
        } else if (a > 0.714 && st->write_num == 4) {                  // :244
            *t_out = 999998;
            st->write_num++;
            return 1;
        } else if (a > 0.5 && st->write_num == (variant == 1 ? 3 : 2)) {  // :250
            *t_out = 999997;
            st->write_num++;
            *stop = 1;
            return 1;
        } else if (variant == 1 && a > 0.333 && st->write_num == 2) {  // CosmoBEC_MPI.cpp:256
            *t_out = 999996;
            st->write_num++;
            return 1;
        } else if (a > 0.25 && st->write_num == 2) {                   // :262
            *t_out = st->write_sten_counter;
            st->write_sten_counter++;
            return 1;
        } else if (a > 0.25 && st->write_num == 1) {                   // :269
            *t_out = 999995;
            st->write_num++;
            st->sten_counter = 0;
            return 1;
        } else if (a > 0.105 && st->write_num == 0) {                  // :276
            *t_out = 999994;
            st->write_num++;
            return 1;
        }
    }
    if ((tn % twrite == 0) || tn == t_stop - 1) {                      // :282
        *t_out = tn + 1;
        return 1;
    }
    return 0;
}

extern "C" double ifdm_constant(const char *name) {
    if (!name) return std::nan("");
    struct { const char *k; double v; } tab[] = {
        {"hbar", hbar}, {"c", c_light}, {"bigG", bigG}, {"H_0", H_0}, {"h_0", h_0},
        {"Omega_m0", Omega_m0}, {"rhoc", rhoc}, {"cm2Mpc", cm2Mpc}, {"kg2eV", kg2eV},
        {"kg2solar", kg2solar}, {"s2Myr", s2Myr}, {"m2Mpc", m2Mpc}, {"m2kpc", m2kpc}};
    for (auto &e : tab) if (std::strcmp(e.k, name) == 0) return e.v;
    return std::nan("");
}

extern "C" int ifdm_abi_version(void) { return IFDM_ABI_VERSION; }
