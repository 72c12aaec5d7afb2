// Host-side scalars of the KDK step (see hy_host_scalars.cpp for references).
#pragma once

#include <cstdint>
#include <vector>

namespace hy::host {

double gravitational_constant();
double rho_crit0();
double hubble(double a, double omega_m, double omega_lambda);
double drift_factor(double a1, double a2, double omega_m, double omega_lambda);
double kick_factor(double a1, double a2, double omega_m, double omega_lambda);
double geometric_mean(double lhs, double rhs);
std::vector<double> time_boundaries(double a_start, double a_final, double delta_ln_a,
                                    std::vector<double> stops);
double exact_sum(const double* values, uint64_t count);

}  // namespace hy::host
