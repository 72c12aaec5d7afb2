// Host scalars of the initial conditions: the linear growth table, the
// tabulated linear power spectrum and the per-mode variance of the Gaussian
// field.  Binary64 results with the reference's operation order (built with
// -ffp-contract=off), long double where the reference uses it.
//
//   growth D1/f1/D2/f2     src/cosmology/cosmology_model.cpp:70-509
//   flat closure           include/cosmo_nbody/cosmology/flat_matter_lambda.hpp
//   natural cubic spline   include/cosmo_nbody/math/cubic_spline.hpp
//   power table + rescale  src/ic/linear_power_spectrum.cpp:27-230,414-664
//   mode wavenumber        include/cosmo_nbody/ic/particle_lattice_bandlimit.hpp:27-38
//   N^6/L^3 P(k)           src/ic/random_field.cpp:108-235 with
//                          include/cosmo_nbody/math/scaled_positive_product.hpp
#include "hy_host_scalars.hpp"

#include <algorithm>
#include <cctype>
#include <cfloat>
#include <cmath>
#include <limits>
#include <map>
#include <numbers>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace hy::host {
namespace {

// --------------------------------------------------------------------------
// Background and growth
// --------------------------------------------------------------------------
void require_cosmology(const Cosmology& c) {
    if (!std::isfinite(c.omega_m) || !std::isfinite(c.omega_lambda) || c.omega_m <= 0.0 || c.omega_m > 1.0 ||
        c.omega_lambda < 0.0)
        throw std::invalid_argument(
            "Flat matter-plus-Lambda cosmology requires finite Omega_m in (0,1] and finite Omega_lambda >= 0");
    const double terms[2] = {c.omega_m, c.omega_lambda};
    if (c.omega_lambda > 1.0 || exact_sum(terms, 2) != 1.0)
        throw std::invalid_argument("Cosmology must be flat (Omega_m + Omega_lambda = 1)");
}

double squared_component_fraction(double component, double complement) {
    if (component == 0.0) return 0.0;
    if (complement == 0.0) return 1.0;
    if (component >= complement) {
        const double ratio = complement / component;
        return 1.0 / (1.0 + ratio * ratio);
    }
    const double ratio = component / complement;
    const double ratio_squared = ratio * ratio;
    return ratio_squared / (1.0 + ratio_squared);
}

struct GrowthPoint {
    double a, D1, f1, D2, f2;
};

class Growth {
public:
    explicit Growth(const Cosmology& c) : om_(c.omega_m), ol_(c.omega_lambda) {
        require_cosmology(c);
        const double a_start = start_scale();
        if (a_start >= 1.0) {
            table_ = {{1.0, 1.0, 1.0, -3.0 / 7.0, 2.0}};
            return;
        }
        // the reference keeps the 2 x nominal grid; the coarser grids only
        // feed its convergence diagnostics
        table_ = integrate(a_start, 2 * balanced_steps(a_start));
    }

    double value(double a, int component) const {
        if (!std::isfinite(a) || a <= 0.0) throw std::invalid_argument("Scale factor must be finite and positive");
        if (a > 1.0) throw std::out_of_range("Growth table supports a <= 1 only; future-time extrapolation is disabled");
        const GrowthPoint& first = table_.front();
        if (a <= first.a) {
            const double ratio = a / first.a;
            if (component == 0) return first.D1 * ratio;
            if (component == 1) return 1.0;
            if (component == 2) return first.D2 * ratio * ratio;
            return 2.0;
        }
        auto pick = [component](const GrowthPoint& p) {
            return component == 0 ? p.D1 : component == 1 ? p.f1 : component == 2 ? p.D2 : p.f2;
        };
        if (a >= 1.0) return pick(table_.back());
        const auto right = std::lower_bound(table_.begin(), table_.end(), a,
                                            [](const GrowthPoint& p, double t) { return p.a < t; });
        if (right == table_.begin()) return pick(*right);
        if (right == table_.end()) return pick(table_.back());
        const auto left = std::prev(right);
        const double interval = right->a - left->a;
        const double t = (a - left->a) / interval;
        const bool second = component >= 2;
        const double gl = second ? left->D2 : left->D1;
        const double gr = second ? right->D2 : right->D1;
        const double rl = second ? left->f2 : left->f1;
        const double rr = second ? right->f2 : right->f1;
        const double sl = rl * gl / left->a;
        const double sr = rr * gr / right->a;
        // cubic Hermite in a with end slopes f D / a
        const double t2 = t * t;
        const double t3 = t2 * t;
        const double h00 = 2.0 * t3 - 3.0 * t2 + 1.0;
        const double h10 = t3 - 2.0 * t2 + t;
        const double h01 = -2.0 * t3 + 3.0 * t2;
        const double h11 = t3 - t2;
        const double g = h00 * gl + h10 * interval * sl + h01 * gr + h11 * interval * sr;
        if (!std::isfinite(g) || (second ? g >= 0.0 : g <= 0.0))
            throw std::runtime_error("Cubic growth interpolation violated the growing-mode sign invariant");
        if (component == 0 || component == 2) return g;
        const double dh00 = 6.0 * t2 - 6.0 * t;
        const double dh10 = 3.0 * t2 - 4.0 * t + 1.0;
        const double dh01 = -dh00;
        const double dh11 = 3.0 * t2 - 2.0 * t;
        const double derivative = (dh00 * gl + dh10 * interval * sl + dh01 * gr + dh11 * interval * sr) / interval;
        const double rate = a * derivative / g;
        if (!std::isfinite(rate))
            throw std::runtime_error("Cubic growth interpolation produced a non-finite logarithmic rate");
        return rate;
    }

private:
    double om_, ol_;
    std::vector<GrowthPoint> table_;

    // (Omega_L / Omega_m) a^3 <= binary64 epsilon at the start of the grid
    double start_scale() const {
        if (ol_ == 0.0) return 1.0;
        const long double eps = static_cast<long double>(std::numeric_limits<double>::epsilon());
        const long double lc =
            (std::log(eps) + std::log(static_cast<long double>(om_)) - std::log(static_cast<long double>(ol_))) / 3.0L;
        if (!std::isfinite(lc)) throw std::overflow_error("Growth ODE matter-era start scale is not representable");
        if (lc >= 0.0L) return 1.0;
        const double r = static_cast<double>(std::exp(lc));
        if (!std::isfinite(r) || r <= 0.0 || r >= 1.0)
            throw std::overflow_error("Growth ODE matter-era start scale is not representable in core::Real");
        return r;
    }

    // RK4 step ~ eps^(1/5) in ln a, even count
    static size_t balanced_steps(double a_start) {
        const long double span = -std::log(static_cast<long double>(a_start));
        const long double eps = static_cast<long double>(std::numeric_limits<double>::epsilon());
        const long double target = std::pow(eps, 1.0L / 5.0L);
        const long double required = std::ceil(span / target);
        if (!std::isfinite(required) || required < 1.0L)
            throw std::overflow_error("Growth ODE precision-derived step count is not representable");
        size_t steps = static_cast<size_t>(required);
        if (steps < 2) steps = 2;
        if (steps & 1u) ++steps;
        return steps;
    }

    // y = [D1, dD1/dlna, D2, dD2/dlna]
    void rhs(double lna, const double y[4], double d[4]) const {
        const double a = std::exp(lna);
        const double om = squared_component_fraction(matter_amplitude(a, om_), std::sqrt(ol_));
        const double drag = 2.0 - 1.5 * om;
        d[0] = y[1];
        d[1] = -drag * y[1] + 1.5 * om * y[0];
        d[2] = y[3];
        d[3] = -drag * y[3] + 1.5 * om * y[2] - 1.5 * om * y[0] * y[0];
    }

    std::vector<GrowthPoint> integrate(double a_start, size_t steps) const {
        const double lna0 = std::log(a_start);
        const double h = -lna0 / static_cast<double>(steps);
        if (!std::isfinite(h) || h <= 0.0) throw std::overflow_error("Growth ODE step is invalid");
        std::vector<GrowthPoint> t;
        t.reserve(steps + 1);
        double y[4] = {a_start, a_start, -(3.0 / 7.0) * a_start * a_start, -(6.0 / 7.0) * a_start * a_start};
        auto append = [&](double av) {
            if (!std::isfinite(av) || av <= 0.0 || !std::isfinite(y[0]) || y[0] <= 0.0 || !std::isfinite(y[1]) ||
                !std::isfinite(y[2]) || y[2] >= 0.0 || !std::isfinite(y[3]))
                throw std::runtime_error("Non-finite or non-growing cosmological growth state");
            t.push_back({av, y[0], y[1] / y[0], y[2], y[3] / y[2]});
        };
        append(a_start);
        for (size_t s = 0; s < steps; ++s) {
            const double lna = lna0 + static_cast<double>(s) * h;
            double k1[4], k2[4], k3[4], k4[4], tmp[4];
            rhs(lna, y, k1);
            for (int j = 0; j < 4; ++j) tmp[j] = y[j] + 0.5 * h * k1[j];
            rhs(lna + 0.5 * h, tmp, k2);
            for (int j = 0; j < 4; ++j) tmp[j] = y[j] + 0.5 * h * k2[j];
            rhs(lna + 0.5 * h, tmp, k3);
            for (int j = 0; j < 4; ++j) tmp[j] = y[j] + h * k3[j];
            rhs(lna + h, tmp, k4);
            for (int j = 0; j < 4; ++j) y[j] += (h / 6.0) * (k1[j] + 2.0 * k2[j] + 2.0 * k3[j] + k4[j]);
            const double next = s + 1 == steps ? 0.0 : lna0 + static_cast<double>(s + 1) * h;
            append(std::exp(next));
        }
        const double d1 = t.back().D1;
        const double d1sq = d1 * d1;
        if (!std::isfinite(d1) || d1 <= 0.0 || !std::isfinite(d1sq) || d1sq <= 0.0)
            throw std::runtime_error("Invalid present-day linear growth normalization");
        for (auto& p : t) {
            p.D1 /= d1;
            p.D2 /= d1sq;
        }
        return t;
    }
};

// --------------------------------------------------------------------------
// Natural cubic spline: the binary64 recurrence with long-double coefficient
// steps, or the knot- and value-normalised long-double build when a binary64
// coefficient would lose normal precision.
// --------------------------------------------------------------------------
bool representable(long double v) {
    constexpr long double mx = static_cast<long double>(std::numeric_limits<double>::max());
    return std::isfinite(v) && v <= mx && v >= -mx;
}

double checked(long double v) {
    if (!representable(v)) throw std::overflow_error("CubicSpline result is not representable.");
    const double r = static_cast<double>(v);
    if (!std::isfinite(r)) throw std::overflow_error("CubicSpline result is not representable.");
    return r;
}

class Spline {
public:
    Spline(std::vector<double> x, std::vector<double> y) : x_(std::move(x)), y_(std::move(y)) {
        if (x_.size() != y_.size() || x_.size() < 2)
            throw std::invalid_argument("CubicSpline requires at least 2 points of equal length.");
        for (size_t i = 0; i < x_.size(); ++i)
            if (!std::isfinite(x_[i]) || !std::isfinite(y_[i]))
                throw std::invalid_argument("CubicSpline knots and ordinates must be finite.");
        for (size_t i = 1; i < x_.size(); ++i)
            if (!(x_[i] > x_[i - 1])) throw std::invalid_argument("CubicSpline x values must be strictly increasing.");
        if (x_.size() == 2)
            throw std::invalid_argument("a two-knot spline (a straight line) is not supported; give at least 3 knots");
        if (!build_native()) build_normalized();
    }

    double operator()(double x) const {
        if (!std::isfinite(x)) throw std::invalid_argument("CubicSpline interpolation query must be finite.");
        const auto exact = std::lower_bound(x_.begin(), x_.end(), x);
        if (exact != x_.end() && *exact == x) return y_[static_cast<size_t>(exact - x_.begin())];
        if (normalized_) return eval_normalized(x);
        if (x < x_.front()) {
            const double dx = x_[1] - x_[0];
            const double dy = y_[1] - y_[0];
            const double d = dy / dx - dx * (2.0 * y2_[0] + y2_[1]) / 6.0;
            return checked(y_[0] + d * (x - x_[0]));
        }
        if (x > x_.back()) {
            const size_t n = x_.size() - 1;
            const double dx = x_[n] - x_[n - 1];
            const double dy = y_[n] - y_[n - 1];
            const double d = dy / dx + dx * (y2_[n - 1] + 2.0 * y2_[n]) / 6.0;
            return checked(y_[n] + d * (x - x_[n]));
        }
        const size_t hi = static_cast<size_t>(std::lower_bound(x_.begin(), x_.end(), x) - x_.begin());
        const size_t lo = hi - 1;
        const double h = x_[hi] - x_[lo];
        const double a = (x_[hi] - x) / h;
        const double b = (x - x_[lo]) / h;
        return checked(a * y_[lo] + b * y_[hi] + ((a * a * a - a) * y2_[lo] + (b * b * b - b) * y2_[hi]) * (h * h) / 6.0);
    }

private:
    std::vector<double> x_, y_, y2_;
    std::vector<long double> nx_, ny_, ny2_;
    long double yscale_ = 1.0L;
    bool normalized_ = false;

    bool build_native() {
        static_assert(LDBL_MANT_DIG > DBL_MANT_DIG, "long double must be wider than binary64");
        const size_t n = x_.size() - 1;
        std::vector<double> y2(n + 1, 0.0), u(n, 0.0);
        for (size_t i = 1; i < x_.size(); ++i) {
            const double iv = x_[i] - x_[i - 1];
            const double iv2 = iv * iv;
            if (!std::isfinite(iv) || !std::isfinite(iv2) ||
                (iv != 0.0 && (iv2 == 0.0 || std::fpclassify(iv2) == FP_SUBNORMAL)))
                return false;
        }
        for (size_t i = 1; i < n; ++i) {
            const double span = x_[i + 1] - x_[i - 1];
            const double rdy = y_[i + 1] - y_[i];
            const double ldy = y_[i] - y_[i - 1];
            if (!std::isfinite(span) || !std::isfinite(rdy) || !std::isfinite(ldy)) return false;
            const double sig = (x_[i] - x_[i - 1]) / span;
            const double p = sig * y2[i - 1] + 2.0;
            const double rs = rdy / (x_[i + 1] - x_[i]);
            const double ls = ldy / (x_[i] - x_[i - 1]);
            const double sd = rs - ls;
            if (!std::isfinite(sig) || !std::isfinite(p) || p == 0.0 || !std::isfinite(rs) || !std::isfinite(ls) ||
                !std::isfinite(sd) || (rdy != 0.0 && (rs == 0.0 || std::fpclassify(rs) == FP_SUBNORMAL)) ||
                (ldy != 0.0 && (ls == 0.0 || std::fpclassify(ls) == FP_SUBNORMAL)))
                return false;
            const long double ny2 = (static_cast<long double>(sig) - 1.0L) / static_cast<long double>(p);
            const long double nu = ((static_cast<long double>(sd) / static_cast<long double>(span)) * 6.0L -
                                    static_cast<long double>(sig) * static_cast<long double>(u[i - 1])) /
                                   static_cast<long double>(p);
            if (!representable(ny2) || !representable(nu)) return false;
            const double cy2 = static_cast<double>(ny2), cu = static_cast<double>(nu);
            if ((ny2 != 0.0L && cy2 == 0.0) || (nu != 0.0L && cu == 0.0) || std::fpclassify(cy2) == FP_SUBNORMAL ||
                std::fpclassify(cu) == FP_SUBNORMAL)
                return false;
            y2[i] = cy2;
            u[i] = cu;
        }
        y2[n] = 0.0;
        for (size_t i = n; i-- > 0;) {
            y2[i] = y2[i] * y2[i + 1] + u[i];
            if (!std::isfinite(y2[i])) return false;
        }
        y2_.swap(y2);
        return true;
    }

    void build_normalized() {
        const size_t n = x_.size() - 1;
        const double xs = std::max(std::abs(x_.front()), std::abs(x_.back()));
        int xe = 0;
        if (xs > 0.0) (void)std::frexp(xs, &xe);
        const long double first = std::scalbn(static_cast<long double>(x_.front()), -xe);
        const long double last = std::scalbn(static_cast<long double>(x_.back()), -xe);
        const long double full = last - first;
        if (!std::isfinite(full) || full <= 0.0L)
            throw std::overflow_error("CubicSpline normalized knot span is not representable.");
        nx_.resize(x_.size());
        for (size_t i = 0; i < x_.size(); ++i) {
            nx_[i] = (std::scalbn(static_cast<long double>(x_[i]), -xe) - first) / full;
            if (!std::isfinite(nx_[i]) || (i > 0 && !(nx_[i] > nx_[i - 1])))
                throw std::overflow_error("CubicSpline normalized knot spacing is not representable.");
        }
        yscale_ = 0.0L;
        for (double v : y_) yscale_ = std::max(yscale_, std::abs(static_cast<long double>(v)));
        if (yscale_ == 0.0L) yscale_ = 1.0L;
        ny_.resize(y_.size());
        for (size_t i = 0; i < y_.size(); ++i) ny_[i] = static_cast<long double>(y_[i]) / yscale_;
        ny2_.assign(n + 1, 0.0L);
        std::vector<long double> u(n, 0.0L);
        for (size_t i = 1; i < n; ++i) {
            const long double l = nx_[i] - nx_[i - 1];
            const long double r = nx_[i + 1] - nx_[i];
            const long double span = l + r;
            const long double sig = l / span;
            const long double p = sig * ny2_[i - 1] + 2.0L;
            const long double sd = (ny_[i + 1] - ny_[i]) / r - (ny_[i] - ny_[i - 1]) / l;
            const long double a = (sig - 1.0L) / p;
            const long double b = (6.0L * sd / span - sig * u[i - 1]) / p;
            if (!std::isfinite(a) || !std::isfinite(b))
                throw std::overflow_error("CubicSpline normalized coefficient construction is not representable.");
            ny2_[i] = a;
            u[i] = b;
        }
        for (size_t i = n; i-- > 0;) {
            ny2_[i] = ny2_[i] * ny2_[i + 1] + u[i];
            if (!std::isfinite(ny2_[i]))
                throw std::overflow_error("CubicSpline normalized back substitution is not representable.");
        }
        normalized_ = true;
    }

    long double coordinate(double x) const {
        const double scale = std::max({std::abs(x), std::abs(x_.front()), std::abs(x_.back())});
        int e = 0;
        if (scale > 0.0) (void)std::frexp(scale, &e);
        const long double q = std::scalbn(static_cast<long double>(x), -e);
        const long double f = std::scalbn(static_cast<long double>(x_.front()), -e);
        const long double l = std::scalbn(static_cast<long double>(x_.back()), -e);
        const long double r = (q - f) / (l - f);
        if (!std::isfinite(r))
            throw std::overflow_error("CubicSpline normalized interpolation coordinate is not representable.");
        return r;
    }

    double eval_normalized(double x) const {
        const long double q = coordinate(x);
        if (x < x_.front()) {
            const long double dt = nx_[1] - nx_[0];
            const long double d = (ny_[1] - ny_[0]) / dt - dt * (2.0L * ny2_[0] + ny2_[1]) / 6.0L;
            return checked((ny_[0] + d * (q - nx_[0])) * yscale_);
        }
        if (x > x_.back()) {
            const size_t n = x_.size() - 1;
            const long double dt = nx_[n] - nx_[n - 1];
            const long double d = (ny_[n] - ny_[n - 1]) / dt + dt * (ny2_[n - 1] + 2.0L * ny2_[n]) / 6.0L;
            return checked((ny_[n] + d * (q - nx_[n])) * yscale_);
        }
        const size_t hi = static_cast<size_t>(std::lower_bound(x_.begin(), x_.end(), x) - x_.begin());
        const size_t lo = hi - 1;
        const long double h = nx_[hi] - nx_[lo];
        const long double a = (nx_[hi] - q) / h;
        const long double b = (q - nx_[lo]) / h;
        return checked((a * ny_[lo] + b * ny_[hi] + ((a * a * a - a) * ny2_[lo] + (b * b * b - b) * ny2_[hi]) * h * h / 6.0L) *
                       yscale_);
    }
};

// --------------------------------------------------------------------------
// Tabulated linear power spectrum
// --------------------------------------------------------------------------
const char* const kRequiredKeys[] = {"k_unit", "power_unit", "redshift", "species", "gauge",
                                     "fidelity", "normalization", "normalization_value", "nonlinear",
                                     "massive_neutrino_sum_eV", "h", "omega_m", "omega_b", "omega_lambda", "n_s"};

std::string trim(const std::string& v) {
    size_t b = 0, e = v.size();
    while (b < e && std::isspace(static_cast<unsigned char>(v[b]))) ++b;
    while (e > b && std::isspace(static_cast<unsigned char>(v[e - 1]))) --e;
    return v.substr(b, e - b);
}

bool required_key(const std::string& k) {
    for (const char* r : kRequiredKeys)
        if (k == r) return true;
    return false;
}

double header_real(const std::map<std::string, std::string>& h, const char* key) {
    const auto it = h.find(key);
    if (it == h.end()) throw std::runtime_error(std::string("Power-spectrum semantic header is missing '") + key + "'");
    size_t used = 0;
    double v = 0.0;
    try {
        v = std::stod(it->second, &used);
    } catch (const std::exception&) {
        used = 0;
    }
    if (used != it->second.size() || !std::isfinite(v))
        throw std::runtime_error(std::string("Power-spectrum semantic binding '") + key + "' is not a finite scalar");
    return v;
}

void require_text(const std::map<std::string, std::string>& h, const char* key, const char* expected) {
    const auto it = h.find(key);
    if (it == h.end()) throw std::runtime_error(std::string("Power-spectrum semantic header is missing '") + key + "'");
    if (it->second != expected)
        throw std::runtime_error(std::string("Power-spectrum semantic binding '") + key + "' is '" + it->second +
                                 "' but expected '" + expected + "'");
}

void require_real(const std::map<std::string, std::string>& h, const char* key, double expected) {
    const double v = header_real(h, key);
    if (!std::isfinite(expected) || v != expected) {
        std::ostringstream m;
        m.precision(17);
        m << "Power-spectrum semantic binding '" << key << "' is " << v << " but expected " << expected;
        throw std::runtime_error(m.str());
    }
}

class LinearPower {
public:
    LinearPower(const Cosmology& c, const char* text, double table_z, double start_z)
        : spline_(load(c, text, table_z, start_z)) {}

    double operator()(double k) const {
        if (!std::isfinite(k)) throw std::invalid_argument("LinearPowerSpectrum query must be finite");
        if (k < 0.0) throw std::invalid_argument("LinearPowerSpectrum query must be a non-negative wavenumber magnitude");
        if (k == 0.0) return 0.0;
        if (k < kmin_ || k > kmax_) throw std::out_of_range("LinearPowerSpectrum query lies outside the tabulated support");
        const double r = std::exp(spline_(std::log(k)));
        if (!std::isfinite(r) || r <= 0.0) throw std::overflow_error("Interpolated linear power is not finite and positive");
        return r;
    }
    double kmin() const { return kmin_; }
    double kmax() const { return kmax_; }

private:
    double kmin_ = 0.0, kmax_ = 0.0;
    Spline spline_;

    Spline load(const Cosmology& c, const char* text, double table_z, double start_z) {
        if (!text) throw std::invalid_argument("LinearPowerSpectrum requires a table");
        const Growth growth(c);
        const double d_table = growth.value(1.0 / (1.0 + table_z), 0);
        const double d_start = growth.value(1.0 / (1.0 + start_z), 0);
        const double scale = std::pow(d_start / d_table, 2.0);
        if (!std::isfinite(scale) || scale <= 0.0)
            throw std::runtime_error("Power-spectrum growth rescaling produced a non-finite or non-positive factor");
        std::istringstream in{std::string(text)};
        std::map<std::string, std::string> header;
        std::vector<double> lk, lp;
        std::string line;
        size_t number = 0;
        bool data = false;
        while (std::getline(in, line)) {
            ++number;
            const std::string t = trim(line);
            if (t.empty() || t[0] == '#') {
                const size_t mark = line.find('#');
                if (mark == std::string::npos) continue;
                const std::string comment = trim(line.substr(mark + 1));
                const size_t eq = comment.find('=');
                if (eq == std::string::npos) continue;
                const std::string key = trim(comment.substr(0, eq));
                const std::string value = trim(comment.substr(eq + 1));
                if (!required_key(key)) continue;
                if (value.empty() || std::any_of(value.begin(), value.end(),
                                                 [](unsigned char b) { return std::isspace(b) != 0; }))
                    throw std::runtime_error("Power-spectrum semantic binding '" + key +
                                             "' must contain one non-empty value at line " + std::to_string(number));
                if (data)
                    throw std::runtime_error("Power-spectrum semantic binding appears after data at line " +
                                             std::to_string(number));
                if (!header.emplace(key, value).second)
                    throw std::runtime_error("Duplicate power-spectrum semantic binding '" + key + "' at line " +
                                             std::to_string(number));
                continue;
            }
            data = true;
            std::istringstream f(line);
            double k = 0.0, p = 0.0;
            if (!(f >> k >> p))
                throw std::runtime_error("Malformed power-spectrum row at line " + std::to_string(number) +
                                         "; expected exactly two numbers");
            f >> std::ws;
            if (f.peek() != std::char_traits<char>::eof() && f.peek() != '#')
                throw std::runtime_error("Unexpected trailing token in power-spectrum row at line " +
                                         std::to_string(number));
            if (!std::isfinite(k) || !std::isfinite(p) || k <= 0.0 || p <= 0.0)
                throw std::runtime_error("Power-spectrum k and P(k) must be finite and positive at line " +
                                         std::to_string(number));
            if (lk.empty()) kmin_ = k;
            kmax_ = k;
            const double sp = p * scale;
            if (!std::isfinite(sp) || sp <= 0.0)
                throw std::overflow_error("Growth-scaled power is non-finite or non-positive at line " +
                                          std::to_string(number));
            lk.push_back(std::log(k));
            lp.push_back(std::log(sp));
        }
        for (const char* key : kRequiredKeys)
            if (!header.count(key))
                throw std::runtime_error(std::string("Power-spectrum semantic header is missing '") + key + "'");
        require_text(header, "k_unit", "h/Mpc");
        require_text(header, "power_unit", "(Mpc/h)^3");
        require_text(header, "species", "cold_plus_baryon");
        require_text(header, "fidelity", "precision_boltzmann");
        require_text(header, "gauge", "synchronous");
        require_text(header, "normalization", "sigma8_z0");
        if (!(c.sigma8 > 0.0))
            throw std::runtime_error("Power-spectrum sigma8_z0 normalization requires positive configured sigma8");
        require_real(header, "normalization_value", c.sigma8);
        require_text(header, "nonlinear", "false");
        require_real(header, "massive_neutrino_sum_eV", 0.0);
        require_real(header, "redshift", table_z);
        require_real(header, "h", c.h);
        require_real(header, "omega_m", c.omega_m);
        require_real(header, "omega_b", c.omega_b);
        require_real(header, "omega_lambda", c.omega_lambda);
        require_real(header, "n_s", c.n_s);
        if (lk.size() < 2) throw std::runtime_error("Power spectrum table must have at least 2 valid positive entries");
        for (size_t i = 1; i < lk.size(); ++i)
            if (!(lk[i] > lk[i - 1])) throw std::runtime_error("Power spectrum k values must be strictly increasing");
        if (!std::isfinite(kmin_) || !std::isfinite(kmax_) || kmin_ <= 0.0 || kmax_ <= kmin_)
            throw std::overflow_error("Power spectrum support is not representable");
        return Spline(std::move(lk), std::move(lp));
    }
};

// |m| 2 pi / L with the integer square summed in long double
double mode_wavenumber(int64_t mx, int64_t my, int64_t mz, double kf) {
    const long double m2 = static_cast<long double>(mx) * static_cast<long double>(mx) +
                           static_cast<long double>(my) * static_cast<long double>(my) +
                           static_cast<long double>(mz) * static_cast<long double>(mz);
    return static_cast<double>(std::sqrt(m2)) * kf;
}

// Positive product held as a normalised long-double mantissa and a binary exponent.
struct ScaledProduct {
    long double mant = 1.0L;
    long long exp = 0;

    void consume(double factor, bool divide) {
        if (!std::isfinite(factor) || factor <= 0.0)
            throw std::invalid_argument("scaled product factors must be finite and strictly positive");
        int fe = 0;
        const long double fm = static_cast<long double>(std::frexp(factor, &fe));
        if (divide) {
            mant = mant / fm;
            exp -= fe;
        } else {
            mant = mant * fm;
            exp += fe;
        }
        int ne = 0;
        mant = std::frexp(mant, &ne);
        exp += ne;
    }

    double times(double factor) const {
        ScaledProduct p = *this;
        p.consume(factor, false);
        const long double wide = std::scalbn(p.mant, static_cast<int>(p.exp));
        if (wide == 0.0L) throw std::underflow_error("random-field mode variance underflowed");
        if (!representable(wide) || wide < 0.0L) throw std::overflow_error("random-field mode variance is not representable");
        const double r = static_cast<double>(wide);
        if (!std::isfinite(r) || r <= 0.0) throw std::underflow_error("random-field mode variance underflowed");
        return r;
    }
};

}  // namespace

void growth(const Cosmology& c, uint64_t n, const double* a, double* D1, double* f1, double* D2, double* f2) {
    const Growth g(c);
    for (uint64_t i = 0; i < n; ++i) {
        if (D1) D1[i] = g.value(a[i], 0);
        if (f1) f1[i] = g.value(a[i], 1);
        if (D2) D2[i] = g.value(a[i], 2);
        if (f2) f2[i] = g.value(a[i], 3);
    }
}

void linear_power(const Cosmology& c, const char* table, double table_redshift, double start_redshift, uint64_t n,
                  const double* k, double* out) {
    const LinearPower p(c, table, table_redshift, start_redshift);
    for (uint64_t i = 0; i < n; ++i) out[i] = p(k[i]);
}

void mode_variance(const Cosmology& c, const char* table, double table_redshift, double start_redshift, uint64_t N,
                   double L, uint64_t max_mode, double* out) {
    if (N < 2 || !std::isfinite(L) || L <= 0.0) throw std::invalid_argument("IC mesh and box size must be positive");
    if (max_mode == 0 || max_mode > N / 2)
        throw std::invalid_argument("IC Fourier band limit must lie in 1..N/2");
    const LinearPower p(c, table, table_redshift, start_redshift);
    const double kf = 2.0 * std::numbers::pi / L;
    const int64_t K = static_cast<int64_t>(max_mode);
    if (kf < p.kmin()) throw std::runtime_error("Power spectrum table does not cover the fundamental box mode");
    if (mode_wavenumber(K, K, K, kf) > p.kmax())
        throw std::runtime_error("Power spectrum table does not cover the full represented particle-lattice Fourier band");
    ScaledProduct norm;
    const double nr = static_cast<double>(N);
    for (int i = 0; i < 6; ++i) norm.consume(nr, false);
    for (int i = 0; i < 3; ++i) norm.consume(L, true);
    const uint64_t half = N / 2;
    const uint64_t entries = 3 * half * half + 1;
    const uint64_t limit = 3 * max_mode * max_mode;
    out[0] = 0.0;
    for (uint64_t m2 = 1; m2 < entries; ++m2) {
        if (m2 > limit) {
            out[m2] = 0.0;
            continue;
        }
        const long double r = std::sqrt(static_cast<long double>(m2));
        const double k = static_cast<double>(r) * kf;
        out[m2] = norm.times(p(k));
    }
}

void cubic_spline(uint64_t n, const double* x, const double* y, uint64_t m, const double* q, double* out) {
    const Spline s(std::vector<double>(x, x + n), std::vector<double>(y, y + n));
    for (uint64_t i = 0; i < m; ++i) out[i] = s(q[i]);
}

}  // namespace hy::host
