// Test-only C entry points over the reference's own hot-path classes.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference translation units listed there (straight from the reference
// source tree, unmodified) into oracle/_ref/libhyref.so.  It exists to pin the
// C oracle (oracle/oracle.c) against the reference implementation on the same
// inputs.  It is never linked into the product.
//
// Every entry point returns 0 on success, or a code naming the exception type
// the reference threw: 1 invalid_argument, 2 logic_error, 3 overflow_error,
// 4 underflow_error, 6 runtime_error, 7 other.
#include "cic_coordinate.hpp"

#include "cosmo_nbody/analysis/cic_spectral_window.hpp"
#include "cosmo_nbody/analysis/spectral_numeric.hpp"
#include "cosmo_nbody/core/particle_store.hpp"
#include "cosmo_nbody/cosmology/cosmology_model.hpp"
#include "cosmo_nbody/cosmology/drift_kick_factors.hpp"
#include "cosmo_nbody/cosmology/units.hpp"
#include "cosmo_nbody/config/simulation_parameters.hpp"
#include "cosmo_nbody/ic/linear_power_spectrum.hpp"
#include "cosmo_nbody/ic/particle_lattice_bandlimit.hpp"
#include "cosmo_nbody/ic/philox_rng.hpp"
#include "cosmo_nbody/math/cubic_spline.hpp"
#include "cosmo_nbody/math/scaled_positive_product.hpp"
#include "cosmo_nbody/math/cic_mass_conservation.hpp"
#include "cosmo_nbody/math/density_normalization.hpp"
#include "cosmo_nbody/math/exact_geometric_mean.hpp"
#include "cosmo_nbody/math/exact_positive_sum.hpp"
#include "cosmo_nbody/math/periodic_box.hpp"
#include "cosmo_nbody/mesh/green_function.hpp"
#include "cosmo_nbody/mesh/mass_assignment.hpp"
#include "cosmo_nbody/mesh/mesh_geometry.hpp"
#include "cosmo_nbody/mesh/pm_force_method.hpp"
#include "cosmo_nbody/time/leapfrog_integrator.hpp"
#include "cosmo_nbody/time/particle_update.hpp"
#include "cosmo_nbody/time/time_stepper.hpp"

#include <array>
#include <complex>
#include <numbers>
#include <string>
#include <cstdint>
#include <cstring>
#include <optional>
#include <span>
#include <stdexcept>
#include <vector>

namespace cn = cosmo_nbody;

namespace {

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::invalid_argument&) {
        return 1;
    } catch (const std::underflow_error&) {
        return 4;
    } catch (const std::overflow_error&) {
        return 3;
    } catch (const std::logic_error&) {
        return 2;
    } catch (const std::runtime_error&) {
        return 6;
    } catch (...) {
        return 7;
    }
}

cn::config::CosmologyParams cosmology(double om, double ol) {
    cn::config::CosmologyParams p;
    p.h = 0.7;
    p.omega_m = om;
    p.omega_lambda = ol;
    p.omega_b = 0.05;
    p.sigma8 = 0.8;
    p.n_s = 0.96;
    return p;
}

// Full cosmology (the IC spectrum header binds all six parameters).
cn::config::CosmologyParams cosmology_full(const double* c) {
    cn::config::CosmologyParams p;
    p.h = c[0];
    p.omega_m = c[1];
    p.omega_b = c[2];
    p.omega_lambda = c[3];
    p.sigma8 = c[4];
    p.n_s = c[5];
    return p;
}

// A generate-mode PM configuration carrying the IC spectrum inputs: lattice
// n_side^3 on an IC mesh ic_mesh, band limit max_mode, table at table_z, start z_start.
cn::config::SimulationParameters ic_config(const double* c, const char* table_path, double table_z, double z_start,
                                           double L, uint64_t n_side, uint64_t ic_mesh, uint64_t max_mode) {
    cn::config::BoxParams b;
    b.L = L;
    b.N = n_side;
    b.N_mesh = ic_mesh;
    cn::config::GravityParams g;
    g.solver = "PM";
    g.deconvolve_cic = false;
    cn::config::TimeParams t;
    t.z_start = z_start;
    t.z_final = 0.0;
    t.delta_ln_a = std::log(1.0 + z_start) / 8.0;
    cn::config::ICParams i;
    i.mode = "generate";
    i.lpt_order = 1;
    i.seed = 1;
    i.mesh_per_dimension = ic_mesh;
    i.max_mode_per_axis = max_mode;
    i.power_spectrum_file = table_path;
    i.power_spectrum_redshift = table_z;
    i.power_spectrum_fidelity = "precision_boltzmann";
    i.amplitude_mode = "gaussian";
    i.phase_pairing = "independent";
    cn::config::OutputParams o;
    o.snapshot_scale_factors = {1.0};
    o.root_directory = "/tmp";
    o.run_label = "oracle";
    o.timestamped_run_directory = false;
    return cn::config::SimulationParameters(cosmology_full(c), b, g, t, i, o);
}

}  // namespace

extern "C" {

// CosmologyModel growth factors at scale factors a[0..n).
int ref_growth(double om, double ol, uint64_t n, const double* a, double* D1, double* f1, double* D2, double* f2) {
    return guarded([&] {
        const cn::cosmology::CosmologyModel model(cosmology(om, ol));
        for (uint64_t i = 0; i < n; ++i) {
            D1[i] = model.D1(a[i]);
            f1[i] = model.f1(a[i]);
            D2[i] = model.D2(a[i]);
            f2[i] = model.f2(a[i]);
        }
    });
}

// math::CubicSpline through (x, y) at q[0..m).
int ref_cubic_spline(uint64_t n, const double* x, const double* y, uint64_t m, const double* q, double* out) {
    return guarded([&] {
        cn::math::CubicSpline s(std::vector<double>(x, x + n), std::vector<double>(y, y + n));
        for (uint64_t i = 0; i < m; ++i) out[i] = s.interpolate(q[i]);
    });
}

// LinearPowerSpectrum::evaluate over a table file (c: h, Om, Ob, OL, sigma8, n_s).
int ref_linear_power(const double* c, const char* table_path, double table_z, double z_start, double L,
                     uint64_t n_side, uint64_t ic_mesh, uint64_t max_mode, uint64_t n, const double* k, double* out) {
    return guarded([&] {
        const auto cfg = ic_config(c, table_path, table_z, z_start, L, n_side, ic_mesh, max_mode);
        const cn::ic::LinearPowerSpectrum pk(cfg);
        for (uint64_t i = 0; i < n; ++i) out[i] = pk.evaluate(k[i]);
    });
}

// RandomField's per-mode variance N^6/L^3 P(k_mag(m)), indexed by m^2 (3 K^2 + 1 entries).
int ref_mode_variance(const double* c, const char* table_path, double table_z, double z_start, double L,
                      uint64_t n_side, uint64_t ic_mesh, uint64_t max_mode, double* out) {
    return guarded([&] {
        const auto cfg = ic_config(c, table_path, table_z, z_start, L, n_side, ic_mesh, max_mode);
        const cn::ic::LinearPowerSpectrum pk(cfg);
        const double n_real = static_cast<double>(ic_mesh);
        const std::array<double, 6> num{n_real, n_real, n_real, n_real, n_real, n_real};
        const std::array<double, 3> den{L, L, L};
        const auto norm = cn::math::ScaledPositiveProduct::from_quotient(num, den, "variance");
        const double kf = 2.0 * std::numbers::pi / L;
        const uint64_t K = max_mode;
        out[0] = 0.0;
        // every m^2 <= 3 K^2 is reached by some (mx, my, mz) with |m_axis| <= K
        for (uint64_t m2 = 1; m2 <= 3 * K * K; ++m2) out[m2] = -1.0;
        for (int64_t mx = 0; mx <= (int64_t)K; ++mx)
            for (int64_t my = 0; my <= (int64_t)K; ++my)
                for (int64_t mz = 0; mz <= (int64_t)K; ++mz) {
                    const uint64_t m2 = (uint64_t)(mx * mx + my * my + mz * mz);
                    if (m2 == 0 || out[m2] >= 0.0) continue;
                    const double kmag = cn::ic::represented_mode_wavenumber(mx, my, mz, kf);
                    out[m2] = norm.multiplied_by(pk.evaluate(kmag), "variance");
                }
    });
}

double ref_gravitational_constant() { return cn::cosmology::units::G; }

// Philox4x32-10 block and its open-interval uniform / Box-Muller maps
// (src/ic/philox_rng.cpp).
void ref_philox4x32_10(const uint32_t* counter, const uint32_t* key, uint32_t* out) {
    const auto r = cn::ic::rng::philox4x32_10({counter[0], counter[1], counter[2], counter[3]}, {key[0], key[1]});
    for (int i = 0; i < 4; ++i) out[i] = r[i];
}
void ref_uniform_pair(uint64_t key, int32_t kx, int32_t ky, int32_t kz, uint32_t stream, double* out) {
    const auto [a, b] = cn::ic::rng::uniform_pair_open01(key, kx, ky, kz, stream);
    out[0] = a;
    out[1] = b;
}
void ref_normal_pair(uint64_t key, int32_t kx, int32_t ky, int32_t kz, uint32_t stream, double* out) {
    const auto [a, b] = cn::ic::rng::normal_pair(key, kx, ky, kz, stream);
    out[0] = a;
    out[1] = b;
}
double ref_rho_crit0() { return cn::cosmology::units::rho_crit0; }

double ref_wrap(double x, double L) { return cn::math::wrap(x, L); }
double ref_minimum_image(double d, double L) { return cn::math::minimum_image(d, L); }

int ref_cic_coordinate(double position, double L, uint64_t N, int64_t* base, double* fraction) {
    return guarded([&] {
        const cn::mesh::MeshGeometry geom(L, N);
        const auto c = cn::mesh::detail::cic_coordinate_1d(position, L, geom.cell_size(), N);
        *base = c.base;
        *fraction = c.fraction;
    });
}

int ref_deposit(uint64_t n, const double* x, const double* y, const double* z,
                const double* masses, double uniform_mass, double L, uint64_t N,
                double* mesh) {
    return guarded([&] {
        const cn::mesh::MeshGeometry geom(L, N);
        cn::mesh::CICMassAssignment cic(geom);
        cn::mesh::RealField field(geom.padded_real_size());
        std::optional<double> uniform;
        std::span<const double> m;
        if (masses) m = std::span<const double>(masses, n);
        else uniform = uniform_mass;
        cic.deposit(std::span<const double>(x, n), std::span<const double>(y, n),
                    std::span<const double>(z, n), m, uniform, field);
        std::memcpy(mesh, field.data(), geom.real_size() * sizeof(double));
    });
}

// Deposit in units of a positive mass (src/mesh/mass_assignment_deposit.inc:1-60),
// as FourierDensityBuilder::build uses it (src/analysis/fourier_density_field.cpp:227-232).
int ref_deposit_unit(uint64_t n, const double* x, const double* y, const double* z,
                     const double* masses, double uniform_mass, double unit, double L, uint64_t N,
                     double* mesh) {
    return guarded([&] {
        const cn::mesh::MeshGeometry geom(L, N);
        cn::mesh::CICMassAssignment cic(geom, {}, unit);
        cn::mesh::RealField field(geom.padded_real_size());
        std::optional<double> uniform;
        std::span<const double> m;
        if (masses) m = std::span<const double>(masses, n);
        else uniform = uniform_mass;
        cic.deposit(std::span<const double>(x, n), std::span<const double>(y, n),
                    std::span<const double>(z, n), m, uniform, field);
        std::memcpy(mesh, field.data(), geom.real_size() * sizeof(double));
    });
}

// Header-only estimator helpers (analysis/spectral_numeric.hpp,
// analysis/cic_spectral_window.hpp).
int ref_pk_edges(double minimum, double maximum, int bins, double* edges) {
    return guarded([&] {
        const cn::analysis::detail::PositiveLogBinGrid grid(minimum, maximum, bins);
        for (std::size_t i = 0; i <= grid.bin_count(); ++i) edges[i] = grid.edge(i);
    });
}

int ref_pk_bin_index(double minimum, double maximum, int bins, double value, int* out) {
    return guarded([&] {
        const cn::analysis::detail::PositiveLogBinGrid grid(minimum, maximum, bins);
        *out = grid.inclusive_bin_index(value);
    });
}

int ref_pk_window_axis(double k, double dx, long double* out) {
    return guarded([&] { *out = cn::analysis::detail::cic_window_axis_wide(k, dx); });
}

int ref_pk_raw_power(double L, double re, double im, double* out) {
    return guarded([&] {
        *out = cn::analysis::detail::scaled_box_volume_times_complex_norm_squared(
            L, std::complex<double>(re, im), "raw power");
    });
}

int ref_pk_raw_power_minus(double L, double re, double im, double s, double* out) {
    return guarded([&] {
        *out = cn::analysis::detail::scaled_box_volume_times_complex_norm_squared_minus(
            L, std::complex<double>(re, im), s, "raw power minus");
    });
}

int ref_pk_volume_times(double L, double value, double* out) {
    return guarded([&] {
        *out = cn::analysis::detail::scaled_box_volume_times(L, static_cast<long double>(value), "volume");
    });
}

// Density source exactly as src/gravity/pm_solver.cpp:700-759.
int ref_density_contrast(double* mesh, uint64_t N, double dx, double total_mass) {
    return guarded([&] {
        const std::size_t cells = N * N * N;
        const cn::math::DensityNormalization dn(total_mass, cells, dx, "ref density");
        if (dn.direct_available()) {
            const double mean = dn.mean_mass_per_cell();
            const double inv = dn.inverse_cell_volume();
            for (std::size_t i = 0; i < cells; ++i) mesh[i] = (mesh[i] - mean) * inv;
        } else {
            for (std::size_t i = 0; i < cells; ++i) mesh[i] = dn.normalize(mesh[i], "ref density");
        }
    });
}

int ref_density_direct_available(uint64_t N, double dx, double total_mass, int* direct) {
    return guarded([&] {
        const cn::math::DensityNormalization dn(total_mass, N * N * N, dx, "ref density");
        *direct = dn.direct_available() ? 1 : 0;
    });
}

int ref_green_apply(uint64_t N, double L, int deconvolve, double* cplx) {
    return guarded([&] {
        const cn::mesh::MeshGeometry geom(L, N);
        const cn::mesh::GreenFunction green(geom, cn::mesh::PMForceMethod::pure_pm(deconvolve != 0));
        cn::mesh::ComplexField field(reinterpret_cast<std::complex<double>*>(cplx),
                                     geom.complex_size(), cn::mesh::external_mesh_storage);
        green.apply(field);
    });
}

int ref_green_apply_treepm(uint64_t N, double L, double split_scale, double* cplx) {
    return guarded([&] {
        const cn::mesh::MeshGeometry geom(L, N);
        const cn::mesh::GreenFunction green(geom, cn::mesh::PMForceMethod::treepm_long_range(split_scale));
        cn::mesh::ComplexField field(reinterpret_cast<std::complex<double>*>(cplx),
                                     geom.complex_size(), cn::mesh::external_mesh_storage);
        green.apply(field);
    });
}

int ref_interpolate_add(uint64_t n, const double* mesh, const double* x, const double* y,
                        const double* z, double L, uint64_t N, double* out) {
    return guarded([&] {
        const cn::mesh::MeshGeometry geom(L, N);
        cn::mesh::CICMassAssignment cic(geom);
        cn::mesh::RealField field(const_cast<double*>(mesh), geom.real_size(),
                                  cn::mesh::external_mesh_storage);
        cic.interpolate_add(field, nullptr, std::span<const double>(x, n), std::span<const double>(y, n),
                            std::span<const double>(z, n), std::span<double>(out, n));
    });
}

// 0.5 * CICMassAssignment::interpolate_mass_weighted_sum_full_mesh, as PMSolver reports it.
int ref_potential_energy(uint64_t n, const double* mesh, const double* x, const double* y, const double* z,
                         const double* masses, double uniform_mass, double L, uint64_t N, double* out) {
    return guarded([&] {
        const cn::mesh::MeshGeometry geom(L, N);
        cn::mesh::CICMassAssignment cic(geom);
        cn::mesh::RealField field(const_cast<double*>(mesh), geom.real_size(), cn::mesh::external_mesh_storage);
        std::optional<double> um;
        if (!masses) um = uniform_mass;
        const double s = cic.interpolate_mass_weighted_sum_full_mesh(
            field, std::span<const double>(x, n), std::span<const double>(y, n), std::span<const double>(z, n),
            masses ? std::span<const double>(masses, n) : std::span<const double>(), um);
        *out = 0.5 * s;
    });
}

int ref_gather(uint64_t n, const double* phi, const double* x, const double* y, const double* z,
               double L, uint64_t N, double* ax, double* ay, double* az) {
    return guarded([&] {
        const cn::mesh::MeshGeometry geom(L, N);
        cn::mesh::CICMassAssignment cic(geom);
        cn::mesh::RealField field(const_cast<double*>(phi), geom.real_size(),
                                  cn::mesh::external_mesh_storage);
        cic.interpolate_centered_gradient3_add_full_mesh(
            field, std::span<const double>(x, n), std::span<const double>(y, n),
            std::span<const double>(z, n), std::span<double>(ax, n),
            std::span<double>(ay, n), std::span<double>(az, n));
    });
}

int ref_mass_conservation(double particle_mass, double deposited_mass, uint64_t count) {
    return guarded([&] { cn::math::require_cic_mass_conservation(particle_mass, deposited_mass, count); });
}

double ref_exact_sum(const double* v, uint64_t n) {
    cn::math::ExactPositiveDoubleSum s;
    for (uint64_t i = 0; i < n; ++i) s.add(v[i]);
    return s.value();
}

double ref_exact_repeated(double v, uint64_t n) {
    cn::math::ExactPositiveDoubleSum s;
    s.add_repeated(v, n);
    return s.value();
}

// One particle through particle_update.hpp (K1 kick + periodic drift).
int ref_kick_drift_one(double* pos, double* mom, const double* acc, double K, double D, double L) {
    cn::time::KickDriftCandidate c;
    const bool ok = cn::time::finite_kick_drift_candidate(
        pos[0], pos[1], pos[2], mom[0], mom[1], mom[2], acc[0], acc[1], acc[2], K, D, L, c);
    if (!ok) return 3;
    pos[0] = c.position_x; pos[1] = c.position_y; pos[2] = c.position_z;
    mom[0] = c.momentum_x; mom[1] = c.momentum_y; mom[2] = c.momentum_z;
    return 0;
}

double ref_hubble(double a, double om, double ol) {
    const cn::cosmology::CosmologyModel model(cosmology(om, ol));
    return model.H(a);
}

int ref_drift_kick(double a1, double a2, double om, double ol, double* drift, double* kick) {
    return guarded([&] {
        const cn::cosmology::CosmologyModel model(cosmology(om, ol));
        cn::cosmology::DriftKickIntegrator dk(model);
        *drift = dk.drift_factor(a1, a2);
        *kick = dk.kick_factor(a1, a2);
    });
}

double ref_geometric_mean(double a, double b) { return cn::math::exact_geometric_mean(a, b); }

int ref_time_boundaries(double a_start, double a_final, double dlna, const double* stops,
                        uint64_t n_stops, double* out, uint64_t cap, uint64_t* count) {
    return guarded([&] {
        cn::time::TimeStepper ts(a_start, a_final, dlna,
                                 std::vector<double>(stops, stops + n_stops));
        const auto& b = ts.boundaries();
        if (b.size() > cap) throw std::overflow_error("capacity");
        std::memcpy(out, b.data(), b.size() * sizeof(double));
        *count = b.size();
    });
}

// Reference LeapfrogIntegrator + TimeStepper + DriftKickIntegrator driving a
// caller-supplied force (C callback adds g into ax/ay/az, which the integrator
// has zeroed).  Arrays are updated in place; n_steps_out reports the steps run.
typedef int (*ref_force_cb)(uint64_t n, const double* x, const double* y, const double* z,
                            double* ax, double* ay, double* az, double a, void* user);

int ref_leapfrog_run(uint64_t n, double* x, double* y, double* z, double* px, double* py,
                     double* pz, double* ax, double* ay, double* az, double uniform_mass,
                     double L, double om, double ol, double a_start, double a_final,
                     double dlna, ref_force_cb cb, void* user, uint64_t* n_steps_out) {
    return guarded([&] {
        cn::core::ParticleStore ps;
        ps.set_uniform_mass(uniform_mass);
        ps.resize(n);
        std::memcpy(ps.get_positions_x().data(), x, n * sizeof(double));
        std::memcpy(ps.get_positions_y().data(), y, n * sizeof(double));
        std::memcpy(ps.get_positions_z().data(), z, n * sizeof(double));
        std::memcpy(ps.get_momenta_x().data(), px, n * sizeof(double));
        std::memcpy(ps.get_momenta_y().data(), py, n * sizeof(double));
        std::memcpy(ps.get_momenta_z().data(), pz, n * sizeof(double));
        const cn::cosmology::CosmologyModel model(cosmology(om, ol));
        cn::cosmology::DriftKickIntegrator dk(model);
        auto force = [&](std::span<const double> fx, std::span<const double> fy,
                         std::span<const double> fz, std::span<const double>,
                         std::span<double> gx, std::span<double> gy, std::span<double> gz,
                         double a) {
            const int rc = cb(fx.size(), fx.data(), fy.data(), fz.data(), gx.data(), gy.data(),
                              gz.data(), a, user);
            if (rc != 0) throw std::runtime_error("force callback failed");
        };
        cn::time::LeapfrogIntegrator integ(ps, dk, force, L);
        cn::time::TimeStepper stepper(a_start, a_final, dlna, {a_final});
        uint64_t steps = 0;
        while (auto step = stepper.next_step()) {
            integ.step(*step);
            ++steps;
        }
        *n_steps_out = steps;
        std::memcpy(x, ps.get_positions_x().data(), n * sizeof(double));
        std::memcpy(y, ps.get_positions_y().data(), n * sizeof(double));
        std::memcpy(z, ps.get_positions_z().data(), n * sizeof(double));
        std::memcpy(px, ps.get_momenta_x().data(), n * sizeof(double));
        std::memcpy(py, ps.get_momenta_y().data(), n * sizeof(double));
        std::memcpy(pz, ps.get_momenta_z().data(), n * sizeof(double));
        std::memcpy(ax, ps.get_accelerations_x().data(), n * sizeof(double));
        std::memcpy(ay, ps.get_accelerations_y().data(), n * sizeof(double));
        std::memcpy(az, ps.get_accelerations_z().data(), n * sizeof(double));
    });
}

}  // extern "C"
