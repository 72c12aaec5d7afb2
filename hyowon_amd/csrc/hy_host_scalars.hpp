// Host-side scalars of the KDK step (see hy_host_scalars.cpp for references).
#pragma once

#include <cstdint>
#include <vector>

namespace hy::host {

double gravitational_constant();
double rho_crit0();
double matter_amplitude(double a, double omega_m);
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

// Initial-condition scalars (hy_ic_host.cpp).
struct Cosmology {
    double h, omega_m, omega_b, omega_lambda, sigma8, n_s;
};
// CosmologyModel growth: component 0 D1, 1 f1, 2 D2, 3 f2 at scale factor a.
void growth(const Cosmology& c, uint64_t n, const double* a, double* D1, double* f1, double* D2, double* f2);
// LinearPowerSpectrum over a table's text, evaluated at k (h/Mpc).
void linear_power(const Cosmology& c, const char* table, double table_redshift, double start_redshift,
                  uint64_t n, const double* k, double* out);
// RandomField per-mode variance N^6/L^3 P(|m| 2 pi/L), indexed by m^2 (3 K^2 + 1 entries).
void mode_variance(const Cosmology& c, const char* table, double table_redshift, double start_redshift,
                   uint64_t N, double L, uint64_t max_mode, double* out);
// Natural cubic spline (include/cosmo_nbody/math/cubic_spline.hpp) through (x, y), evaluated at q.
void cubic_spline(uint64_t n, const double* x, const double* y, uint64_t m, const double* q, double* out);

}  // namespace hy::host
