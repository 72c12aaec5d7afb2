// Host-side scalars of the KDK step: background expansion, drift/kick
// integrals, log-a boundary schedule, correctly rounded geometric mean and an
// exact non-negative sum.  Binary64 results, long-double quadrature, built with
// -ffp-contract=off so the operation order below is the one executed.
//
//   H(a)                 src/cosmology/cosmology_model.cpp:25-57,167-178
//   drift / kick factor  src/cosmology/drift_kick_factors.cpp:28-141
//   composite GL 8/16    include/cosmo_nbody/math/roundoff_resolved_gauss_legendre_16.hpp
//   boundaries           src/time/time_stepper.cpp:16-64,194-276
//   geometric mean       include/cosmo_nbody/math/exact_geometric_mean.hpp:168-275
//   exact sum            include/cosmo_nbody/math/exact_positive_sum.hpp (contract)
#include "hy_host_scalars.hpp"

#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <numbers>
#include <stdexcept>
#include <vector>

namespace hy::host {

double gravitational_constant() {
    // include/cosmo_nbody/cosmology/units.hpp:14-26 (same expression order)
    constexpr double g_si = 6.67430e-11;
    constexpr double mpc_m = 3.085677581491367e22;
    constexpr double msun_kg = 1.98847e30;
    constexpr double mass_unit_kg = 1e10 * msun_kg;
    constexpr double v_unit = 1000.0;
    return g_si * mass_unit_kg / (mpc_m * v_unit * v_unit);
}

double rho_crit0() {
    constexpr double h0 = 100.0;
    return 3.0 * h0 * h0 / (8.0 * std::numbers::pi * gravitational_constant());
}

// --------------------------------------------------------------------------
// Background
// --------------------------------------------------------------------------
// sqrt(Omega_m) a^(-3/2), the matter term of E(a) (src/cosmology/cosmology_model.cpp:25-57)
double matter_amplitude(double a, double omega_m) {
    if (!std::isfinite(a) || a <= 0.0) throw std::invalid_argument("Scale factor must be finite and positive");
    const double root_a = std::sqrt(a);
    const double root_m = std::sqrt(omega_m);
    if (a <= 1.0) {
        // sqrt(Om) / a / sqrt(a), one rounding to binary64 at the end
        long double q = static_cast<long double>(root_m);
        q /= static_cast<long double>(a);
        q /= static_cast<long double>(root_a);
        return static_cast<double>(q);
    }
    return root_m * ((1.0 / root_a) / a);
}

double hubble(double a, double omega_m, double omega_lambda) {
    const double matter = matter_amplitude(a, omega_m);
    return 100.0 * std::hypot(matter, std::sqrt(omega_lambda));
}

// --------------------------------------------------------------------------
// Gauss-Legendre rules (positive nodes ascending), from Newton iteration on
// the Legendre recurrence in long double, rounded once to binary64.
// --------------------------------------------------------------------------
namespace {

template <int ORDER>
struct GaussLegendre {
    std::array<double, ORDER / 2> node{};
    std::array<double, ORDER / 2> weight{};
    GaussLegendre() {
        constexpr long double pi = 3.141592653589793238462643383279502884L;
        for (int i = 0; i < ORDER / 2; ++i) {
            long double x = std::cos(pi * (static_cast<long double>(i) + 0.75L) / (ORDER + 0.5L));
            long double deriv = 0.0L;
            for (int iter = 0; iter < 64; ++iter) {
                auto [pn, dpn] = legendre(x);
                deriv = dpn;
                const long double step = pn / dpn;
                x -= step;
                if (std::fabs(step) < 1e-30L) break;
            }
            deriv = legendre(x).second;
            const int slot = ORDER / 2 - 1 - i;
            node[slot] = static_cast<double>(x);
            weight[slot] = static_cast<double>(2.0L / ((1.0L - x * x) * deriv * deriv));
        }
    }
    static std::pair<long double, long double> legendre(long double x) {
        long double p_prev = 1.0L, p = x;
        for (int k = 2; k <= ORDER; ++k) {
            const long double next = ((2.0L * k - 1.0L) * x * p - (k - 1.0L) * p_prev) / k;
            p_prev = p;
            p = next;
        }
        const long double dp = ORDER * (x * p - p_prev) / (x * x - 1.0L);
        return {p, dp};
    }
};

const GaussLegendre<16>& gl16() { static const GaussLegendre<16> r; return r; }
const GaussLegendre<8>& gl8() { static const GaussLegendre<8> r; return r; }

struct Kahan {
    long double sum = 0.0L;
    long double carry = 0.0L;
    void add(long double term) {
        const long double next = sum + term;
        if (std::fabs(sum) >= std::fabs(term)) carry += (sum - next) + term;
        else carry += (term - next) + sum;
        sum = next;
    }
    long double value() const { return sum + carry; }
};

template <typename F>
std::pair<long double, long double> composite(const F& f, double lo, double hi, std::size_t cells) {
    const double width = hi - lo;
    const double cell = width / static_cast<double>(cells);
    const double half = 0.5 * cell;
    if (!(width > 0.0) || !(cell > 0.0) || !(half > 0.0))
        throw std::underflow_error("quadrature cells are not representable");
    Kahan t8, t16;
    const auto& r8 = gl8();
    const auto& r16 = gl16();
    for (std::size_t c = 0; c < cells; ++c) {
        const double mid = lo + (static_cast<double>(c) + 0.5) * cell;
        Kahan l8, l16;
        auto rule = [&](const auto& r, Kahan& acc) {
            for (std::size_t i = 0; i < r.node.size(); ++i) {
                const double off = half * r.node[i];
                const double left = mid - off, right = mid + off;
                if (!(lo < left && left < right && right < hi))
                    throw std::underflow_error("quadrature nodes collapsed at binary64 resolution");
                const double fl = f(left), fr = f(right);
                if (!std::isfinite(fl) || !std::isfinite(fr))
                    throw std::overflow_error("quadrature integrand is non-finite");
                acc.add(static_cast<long double>(r.weight[i])
                        * (static_cast<long double>(fl) + static_cast<long double>(fr)));
            }
        };
        rule(r8, l8);
        rule(r16, l16);
        t8.add(static_cast<long double>(half) * l8.value());
        t16.add(static_cast<long double>(half) * l16.value());
    }
    return {t8.value(), t16.value()};
}

long double ulp_scale(double v) {
    long double s = 0.0L;
    for (double n : {std::nextafter(v, std::numeric_limits<double>::infinity()),
                     std::nextafter(v, -std::numeric_limits<double>::infinity())}) {
        if (std::isfinite(n)) s = std::max(s, std::fabs(static_cast<long double>(n) - static_cast<long double>(v)));
    }
    if (!(s > 0.0L)) throw std::underflow_error("quadrature could not resolve binary64 spacing");
    return s;
}

// Refine by cell doubling until the 8/16 gap and the 16-point change are both
// within one binary64 spacing of the returned value.
template <typename F>
double resolved_integral(const F& f, double lo, double hi) {
    auto prev = composite(f, lo, hi, 1);
    for (std::size_t cells = 2; cells != 0; cells *= 2) {
        const auto cur = composite(f, lo, hi, cells);
        const double value = static_cast<double>(cur.second);
        const long double tol = ulp_scale(value);
        if (std::fabs(cur.second - cur.first) <= tol && std::fabs(cur.second - prev.second) <= tol)
            return value;
        prev = cur;
    }
    throw std::overflow_error("quadrature subdivision count overflowed");
}

double log_interval(double upper, double lower) {
    const double rel = (upper - lower) / lower;
    const double v = std::isfinite(rel) ? std::log1p(rel) : std::log(upper) - std::log(lower);
    if (!std::isfinite(v)) throw std::overflow_error("logarithmic interval is non-finite");
    if (v <= 0.0) throw std::underflow_error("logarithmic interval was erased by rounding");
    return v;
}

// kind 0: drift 1/(a^2 H) da/a ; kind 1: kick 1/(a H) da/a  (u = ln(a/base))
double time_integral(double a1, double a2, double om, double ol, int kind) {
    if (!std::isfinite(a1) || !std::isfinite(a2) || a1 <= 0.0 || a2 <= 0.0)
        throw std::invalid_argument("Drift/kick scale factors must be finite and positive");
    if (a1 == a2) return 0.0;
    const bool forward = a2 > a1;
    const double base = forward ? a1 : a2;
    const double top = forward ? a2 : a1;
    const double width = log_interval(top, base);
    auto integrand = [&](double u) {
        const double ratio = std::exp(u);
        double a = base * ratio;
        if (!std::isfinite(ratio) || !std::isfinite(a) || a <= 0.0) a = std::exp(std::log(base) + u);
        const double ah = a * hubble(a, om, ol);
        return kind == 0 ? 1.0 / (a * ah) : 1.0 / ah;
    };
    const double v = resolved_integral(integrand, 0.0, width);
    if (!std::isfinite(v)) throw std::overflow_error("drift/kick quadrature produced a non-finite value");
    return forward ? v : -v;
}

}  // namespace

double drift_factor(double a1, double a2, double om, double ol) { return time_integral(a1, a2, om, ol, 0); }
double kick_factor(double a1, double a2, double om, double ol) { return time_integral(a1, a2, om, ol, 1); }

// --------------------------------------------------------------------------
// Correctly rounded sqrt(lhs*rhs): long-double estimate, then exact midpoint
// tests in binary128 (the products involved are exact there).
// --------------------------------------------------------------------------
double geometric_mean(double lhs, double rhs) {
    if (!std::isfinite(lhs) || !std::isfinite(rhs) || !(lhs > 0.0) || !(rhs > 0.0))
        throw std::invalid_argument("geometric mean needs finite positive inputs");
    using q = __float128;
    const q exact = static_cast<q>(lhs) * static_cast<q>(rhs);
    double r = static_cast<double>(std::sqrt(static_cast<long double>(lhs) * static_cast<long double>(rhs)));
    auto odd = [](double v) { uint64_t b; std::memcpy(&b, &v, 8); return (b & 1u) != 0; };
    for (int guard = 0; guard < 8; ++guard) {
        const double up = std::nextafter(r, std::numeric_limits<double>::infinity());
        const double dn = std::nextafter(r, 0.0);
        const q hi = (static_cast<q>(r) + static_cast<q>(up)) / 2;
        const q lo = (static_cast<q>(r) + static_cast<q>(dn)) / 2;
        if (hi * hi < exact) { r = up; continue; }
        if (lo * lo > exact) { r = dn; continue; }
        if (hi * hi == exact && odd(r)) r = up;
        else if (lo * lo == exact && odd(r)) r = dn;
        break;
    }
    return r;
}

// --------------------------------------------------------------------------
// Boundary schedule
// --------------------------------------------------------------------------
std::vector<double> time_boundaries(double a_start, double a_final, double dlna,
                                    std::vector<double> stops) {
    if (!std::isfinite(a_start) || !std::isfinite(a_final) || !std::isfinite(dlna))
        throw std::invalid_argument("TimeStepper inputs must be finite");
    if (a_start <= 0.0 || a_final <= 0.0) throw std::invalid_argument("Scale factors must be positive");
    if (dlna <= 0.0) throw std::invalid_argument("delta_ln_a must be positive");
    if (a_start > a_final) throw std::invalid_argument("TimeStepper production schedule must be non-descending");
    std::vector<double> out{a_start};
    if (a_start == a_final) return out;
    double last = a_start;
    for (double s : stops) {
        if (!std::isfinite(s) || s <= a_start || s > a_final)
            throw std::invalid_argument("TimeStepper stop scale factors must be in (a_start, a_final]");
        if (!(s > last)) throw std::invalid_argument("TimeStepper stop scale factors must be strictly increasing");
        last = s;
    }
    if (stops.empty() || stops.back() != a_final) stops.push_back(a_final);

    auto push = [&](double next) {
        if (!(next > out.back()) || !std::isfinite(next))
            throw std::underflow_error("TimeStepper boundary does not advance monotonically");
        const double mid = geometric_mean(out.back(), next);
        if (!(mid > out.back()) || !(mid < next))
            throw std::underflow_error("TimeStep has no distinct representable geometric half-step");
        out.push_back(next);
    };
    for (double stop : stops) {
        const double anchor = out.back();
        const double span = log_interval(stop, anchor);
        const double ratio = span / dlna;
        if (!std::isfinite(ratio) || ratio <= 0.0)
            throw std::overflow_error("TimeStepper logarithmic step count is not representable");
        const double count = std::max(1.0, std::ceil(std::nextafter(ratio, std::numeric_limits<double>::infinity())));
        const auto steps = static_cast<uint64_t>(count);
        bool placed = false;
        for (uint64_t k = 1; k < steps; ++k) {
            const double offset = dlna * static_cast<double>(k);
            const double inc = std::expm1(offset);
            const double next = std::isfinite(inc) ? std::fma(anchor, inc, anchor)
                                                   : std::exp(std::log(anchor) + offset);
            if (!std::isfinite(next)) throw std::overflow_error("TimeStepper scale-factor boundary is not representable");
            if (next >= stop) { push(stop); placed = true; break; }
            const double tail = geometric_mean(next, stop);
            if (!(tail > next && tail < stop)) { push(stop); placed = true; break; }
            push(next);
        }
        if (!placed) push(stop);
    }
    return out;
}

// --------------------------------------------------------------------------
// Exact sum: integer multiples of 2^-1074 in 32-bit digits, rounded once.
// --------------------------------------------------------------------------
double exact_sum(const double* v, uint64_t n) {
    constexpr int kDigits = 72;
    std::array<uint64_t, kDigits> d{};
    uint64_t pending = 0;
    auto normalise = [&] {
        uint64_t carry = 0;
        for (auto& w : d) {
            const uint64_t s = w + carry;
            w = s & 0xffffffffull;
            carry = s >> 32;
        }
        pending = 0;
    };
    for (uint64_t i = 0; i < n; ++i) {
        const double x = v[i];
        if (!std::isfinite(x) || x < 0.0) return std::numeric_limits<double>::quiet_NaN();
        if (x == 0.0) continue;
        uint64_t bits;
        std::memcpy(&bits, &x, 8);
        const uint32_t e = static_cast<uint32_t>((bits >> 52) & 0x7ff);
        const uint64_t frac = bits & ((1ull << 52) - 1);
        const uint64_t mant = e ? (frac | (1ull << 52)) : frac;
        const uint32_t shift = e ? e - 1 : 0;
        const uint32_t limb = shift / 32, off = shift % 32;
        unsigned __int128 w = static_cast<unsigned __int128>(mant) << off;
        for (uint32_t k = 0; k < 3 && limb + k < kDigits; ++k) {
            d[limb + k] += static_cast<uint64_t>(w & 0xffffffffu);
            w >>= 32;
        }
        if (++pending == (1u << 30)) normalise();
    }
    normalise();
    int top = -1;
    for (int i = kDigits - 1; i >= 0 && top < 0; --i)
        if (d[i]) top = i * 32 + (63 - __builtin_clzll(d[i]));
    if (top < 0) return 0.0;
    auto bit = [&](int b) -> uint64_t { return b < 0 ? 0 : (d[b / 32] >> (b % 32)) & 1u; };
    if (top < 53) {
        uint64_t m = 0;
        for (int b = 0; b <= top; ++b) m |= bit(b) << b;
        return std::ldexp(static_cast<double>(m), -1074);
    }
    const int lo = top - 52;
    uint64_t m = 0;
    for (int b = 0; b < 53; ++b) m |= bit(lo + b) << b;
    const bool round = bit(lo - 1) != 0;
    bool sticky = false;
    for (int b = 0; b < lo - 1 && !sticky; ++b) sticky = bit(b) != 0;
    if (round && (sticky || (m & 1u))) ++m;
    return std::ldexp(static_cast<double>(m), lo - 1074);
}

}  // namespace hy::host
