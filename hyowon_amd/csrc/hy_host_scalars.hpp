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

// Matter power-spectrum estimator scalars (hy_pk_host.cpp).
std::vector<double> log_shell_edges(double minimum, double maximum, int bins);
long double cic_window_axis(double k, double dx);
double mass_square_fraction_uniform(uint64_t n);
// Validates masses (finite, positive); weight_sum receives sum(m / max m).
double mass_square_fraction(const double* masses, uint64_t n, double* weight_sum);
double volume_times(double L, double value);  // L^3 * value in a shared exponent domain

}  // namespace hy::host
