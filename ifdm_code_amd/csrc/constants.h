// constants.h -- unit and cosmology constants of the reference
// (src/Cosmoinput3D.h:2-17), defined in cosmo.cpp.
#pragma once
#include "../../include/ifdm.h"
#include "criteria.h"
namespace ifdm {
extern const double m2kpc, m2Mpc, cm2Mpc, s2Myr, kg2eV, kg2solar, rhoc, h_0, H_0, hbar, Omega_m0, bigG, c_light;
double scale_integral(double a0, double a1, int n, double epsrel, int limit);
// the time-step criteria's constants of a configuration (cosmo.cpp)
CritConst crit_const(const ifdm_config *cfg);
}
