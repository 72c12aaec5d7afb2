// Host-side scalars of the matter power-spectrum estimator: logarithmic shell
// edges, the per-axis CIC window, mass-square fraction and shot noise.
// Long-double intermediates as in the reference, binary64 results, built with
// -ffp-contract=off.
//
//   shell edges          include/cosmo_nbody/analysis/spectral_numeric.hpp:28-122
//   CIC window (axis)    include/cosmo_nbody/analysis/cic_spectral_window.hpp:13-37
//   mass-square fraction src/analysis/fourier_density_field.cpp:27-85,160-225
//   L^3 * value          include/cosmo_nbody/analysis/spectral_numeric.hpp:170-205,224-236
#include "hy_host_scalars.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <vector>

namespace hy::host {

namespace {

// log(upper / lower) through log1p of the relative gap.
double log_ratio(double upper, double lower) {
    if (!std::isfinite(upper) || !std::isfinite(lower) || lower <= 0.0 || upper < lower)
        throw std::invalid_argument("Stable logarithmic ratio requires finite 0 < lower <= upper");
    if (upper == lower) return 0.0;
    const double rel = (upper - lower) / lower;
    const double v = std::isfinite(rel) ? std::log1p(rel) : std::log(upper) - std::log(lower);
    if (!std::isfinite(v) || v <= 0.0) throw std::overflow_error("Stable logarithmic ratio is not representable");
    return v;
}

// lower * (upper / lower)^f, widened to long double for spans above e.
double log_interpolate(double lower, double upper, double f) {
    if (!std::isfinite(lower) || !std::isfinite(upper) || lower <= 0.0 || upper < lower || !std::isfinite(f) ||
        f < 0.0 || f > 1.0)
        throw std::invalid_argument(
            "Stable logarithmic interpolation requires finite 0 < lower <= upper and fraction in [0,1]");
    if (f == 0.0 || lower == upper) return lower;
    if (f == 1.0) return upper;
    const double span = log_ratio(upper, lower);
    double r;
    if (span <= 1.0) {
        r = lower * std::exp(f * span);
    } else {
        const long double lv = std::log(static_cast<long double>(lower)) +
                               static_cast<long double>(f) * static_cast<long double>(span);
        const long double w = std::exp(lv);
        if (!std::isfinite(w) || w <= 0.0L || w > static_cast<long double>(std::numeric_limits<double>::max()))
            throw std::overflow_error("Stable logarithmic interpolation is not representable");
        r = static_cast<double>(w);
    }
    if (!std::isfinite(r) || r <= 0.0 || r < lower || r > upper)
        throw std::overflow_error("Stable logarithmic interpolation escaped its admitted endpoints");
    return r;
}

// Neumaier-compensated sum of non-negative terms.
struct CompensatedSum {
    long double sum = 0.0L, comp = 0.0L;
    void add(long double v) {
        if (!std::isfinite(v) || v < 0.0L)
            throw std::invalid_argument("FourierDensityBuilder moment contribution must be finite and non-negative");
        if (v == 0.0L) return;
        const long double u = sum + v;
        if (!std::isfinite(u)) throw std::overflow_error("FourierDensityBuilder moment accumulation overflowed");
        comp += std::fabs(sum) >= std::fabs(v) ? (sum - u) + v : (v - u) + sum;
        sum = u;
    }
    long double value() const {
        const long double r = sum + comp;
        if (!std::isfinite(r) || r <= 0.0L)
            throw std::overflow_error("FourierDensityBuilder compensated sum is not finite and positive");
        return r;
    }
};

double checked_fraction(long double v) {
    if (!std::isfinite(v) || v <= 0.0L || v > static_cast<long double>(std::numeric_limits<double>::max()))
        throw std::overflow_error("FourierDensityBuilder mass-square fraction is not representable in core::Real");
    const double r = static_cast<double>(v);
    if (!std::isfinite(r) || r <= 0.0)
        throw std::overflow_error("FourierDensityBuilder mass-square fraction rounded outside the finite positive range");
    if (r > 1.0 + 64.0 * std::numeric_limits<double>::epsilon())
        throw std::logic_error("FourierDensityBuilder mass-square fraction exceeds unity");
    return std::min(1.0, r);
}

}  // namespace

std::vector<double> log_shell_edges(double minimum, double maximum, int bins) {
    if (!std::isfinite(minimum) || !std::isfinite(maximum) || minimum <= 0.0 || maximum <= minimum || bins <= 0)
        throw std::invalid_argument(
            "Positive logarithmic bin grid requires finite 0 < minimum < maximum and positive bin_count");
    std::vector<double> e(static_cast<size_t>(bins) + 1);
    for (int i = 0; i <= bins; ++i) {
        e[i] = log_interpolate(minimum, maximum, static_cast<double>(i) / static_cast<double>(bins));
        if (i != 0 && !(e[i] > e[i - 1]))
            throw std::overflow_error("Positive logarithmic bin edges are not strictly representable");
    }
    return e;
}

long double cic_window_axis(double k, double dx) {
    if (!std::isfinite(k) || !std::isfinite(dx) || dx <= 0.0)
        throw std::invalid_argument("CIC spectral window requires finite k and positive cell size");
    const long double arg = 0.5L * static_cast<long double>(k) * static_cast<long double>(dx);
    if (arg == 0.0L) return 1.0L;
    const long double s = std::sin(arg) / arg;
    if (!std::isfinite(s)) throw std::overflow_error("CIC spectral sinc is not representable");
    const long double m = std::min(1.0L, std::fabs(s));
    return m * m;
}

double mass_square_fraction_uniform(uint64_t n) {
    return checked_fraction(1.0L / static_cast<long double>(n));
}

double mass_square_fraction(const double* masses, uint64_t n, double* weight_sum) {
    double unit = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        const double m = masses[i];
        if (!std::isfinite(m) || m <= 0.0)
            throw std::invalid_argument("FourierDensityBuilder masses must be finite and positive");
        unit = std::max(unit, m);
    }
    CompensatedSum s1, s2;
    for (uint64_t i = 0; i < n; ++i) {
        const double w = masses[i] / unit;
        if (!std::isfinite(w) || w <= 0.0 || w > 1.0)
            throw std::overflow_error("FourierDensityBuilder positive normalized mass weight is not representable");
        const long double lw = static_cast<long double>(w);
        s1.add(lw);
        s2.add(lw * lw);
    }
    const long double sw = s1.value(), sw2 = s2.value();
    if (weight_sum) *weight_sum = static_cast<double>(sw);
    return checked_fraction(sw2 / (sw * sw));
}

double volume_times(double L, double value) {
    if (!std::isfinite(L) || L <= 0.0 || !std::isfinite(value))
        throw std::invalid_argument("MatterPowerSpectrum shot noise requires finite positive box_size and finite value");
    if (value == 0.0) return 0.0;
    int be = 0, ve = 0, ce = 0;
    const long double bf = std::frexp(static_cast<long double>(L), &be);
    long double c = std::frexp(static_cast<long double>(value), &ve);
    for (int i = 0; i < 3; ++i) c *= bf;
    c = std::frexp(c, &ce);
    const long double v = std::scalbn(c, 3 * be + ve + ce);
    if (!std::isfinite(v) || std::fabs(v) > static_cast<long double>(std::numeric_limits<double>::max()))
        throw std::overflow_error("MatterPowerSpectrum shot noise is not representable in core::Real");
    const double r = static_cast<double>(v);
    if (!std::isfinite(r) || r == 0.0) throw std::overflow_error("MatterPowerSpectrum shot noise rounded outside core::Real");
    return r;
}

}  // namespace hy::host
