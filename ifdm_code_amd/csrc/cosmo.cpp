// cosmo.cpp -- host-side physics of the KDK solver: the unit constants, the
// scale-factor integrals and the adaptive time-step criteria.
//
// Constants: src/Cosmoinput3D.h:2-17 of the reference (jchan192/IFDM_Code).
// Time step: Timestep_criteria, src/CosmoBEC_MPI_stan.cpp:19-78.
// Integrals: the reference calls gsl_integration_qags (GSL 2.4) with
// epsabs = 0, epsrel = 1e-10, limit 1000 on 1/(H a^n) (src/struct3D.h:127-146).
// The rule, the acceptance test and the bisection are in criteria.h, shared
// with the device's step preparation; the first evaluation is accepted for
// every step interval of this smooth integrand.
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>
#include <algorithm>

#include "../../include/ifdm.h"
#include "constants.h"
#include "criteria.h"

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

// The constants of the time-step criteria for a configuration, formed in
// the order of :32-41 so that dt_X = coefficient * (powers of a) keeps the
// reference's association
CritConst crit_const(const ifdm_config *cfg) {
    CritConst k;
    k.H_0 = H_0;
    k.Omega_m0 = Omega_m0;
    const double mt = cfg->mfdm / std::pow(c_light, 2);
    const double dx = 2.0 * cfg->xmax / cfg->n;                                  // struct3D.h:52
    k.ke = 4. / (3. * M_PI) * mt / hbar * (dx * dx);
    k.pe = 2. * M_PI * hbar / mt;
    k.qi = 2. * M_PI * hbar / std::fabs(cfg->g);
    k.has_g = cfg->g != 0;
    return k;
}

double scale_integral(double a0, double a1, int n, double epsrel, int limit) {
    CritConst k{};
    k.H_0 = H_0;
    k.Omega_m0 = Omega_m0;
    double err, rabs, rasc;
    const double r0 = qk21(k, a0, a1, n, &err, &rabs, &rasc);
    if (qags_first_accepted(r0, err, rasc, epsrel)) return r0;
    std::vector<Piece> pcs((size_t)std::max(limit, 1));
    bool ok;
    return qags_bisect(k, a0, a1, n, r0, err, epsrel, limit, pcs.data(), &ok);
}

}  // namespace ifdm

using namespace ifdm;

extern "C" double ifdm_scale_integral(double a0, double a1, int power) {
    return scale_integral(a0, a1, power, 1e-10, 1000);
}

extern "C" int ifdm_timestep_criteria(const ifdm_config *cfg, double a, const ifdm_stats *st,
                                      ifdm_dt *dt_out, double *a_next, int *which_out) {
    if (!cfg || !st || !dt_out || !a_next || !(a > 0.0) || cfg->n <= 0) return IFDM_E_ARG;
    const CritConst k = crit_const(cfg);
    int which = 3;
    const double da = criteria_da(k, a, st->max_phi, st->max_rho, &which);    // :32-61
    dt_out->dKE = scale_integral(a, a + da, 3, 1e-10, 1000);                     // :66
    dt_out->dPE = scale_integral(a, a + da, 2, 1e-10, 1000);                     // :67
    dt_out->dQI = scale_integral(a, a + da, 4, 1e-10, 1000);                     // :68
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
