// hyowon/pm_solver.hpp -- C++20 drop-in over libhyowon_pm's C ABI.
//
// Mirrors the reference's interfaces for the pure-PM step so that a caller of
//   cosmo_nbody::gravity::PMSolver            (include/cosmo_nbody/gravity/pm_solver.hpp:53-70)
//   cosmo_nbody::time::LeapfrogIntegrator::ForceSolverFunc
//                                              (include/cosmo_nbody/time/leapfrog_integrator.hpp:18-26)
//   cosmo_nbody::cosmology::DriftKickIntegrator (include/cosmo_nbody/cosmology/drift_kick_factors.hpp:14-27)
//   cosmo_nbody::time::TimeStepper             (include/cosmo_nbody/time/time_stepper.hpp)
// can switch to the MI355X path by changing the namespace.  Spans are host
// memory (the reference's ParticleStore); the library stages them to the GPU.
// Errors are rethrown as the same std:: exception classes the reference uses.
#pragma once

#include "hyowon_pm.h"

#include <cstddef>
#include <cstdint>
#include <limits>
#include <optional>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

namespace hyowon {

inline void throw_status(int rc) {
    if (rc == HY_OK) return;
    const std::string msg = hy_last_error();
    switch (rc) {
        case HY_EINVAL: throw std::invalid_argument(msg);
        case HY_ELOGIC: throw std::logic_error(msg);
        case HY_EOVERFLOW: throw std::overflow_error(msg);
        case HY_EUNDERFLOW: throw std::underflow_error(msg);
        case HY_ENOMEM: throw std::bad_alloc();
        default: throw std::runtime_error(msg);
    }
}

enum class Precision { FP64 = HY_PREC_FP64, FP32Mixed = HY_PREC_FP32_MIXED };

namespace mesh {

// mesh::PMForceMethod (include/cosmo_nbody/mesh/pm_force_method.hpp:10-60):
// pure PM, or the TreePM long-range force (continuum Green, Gaussian split,
// spectral gradient, CIC compensation).
class PMForceMethod {
public:
    static PMForceMethod pure_pm(bool deconvolve_cic) noexcept {
        return PMForceMethod(HY_METHOD_PURE_PM, 0.0, deconvolve_cic);
    }
    static PMForceMethod treepm_long_range(double split_scale) {
        if (!(split_scale > 0.0) || split_scale == std::numeric_limits<double>::infinity())
            throw std::invalid_argument("TreePM long-range PM method requires a finite positive split scale");
        return PMForceMethod(HY_METHOD_TREEPM_LONG_RANGE, split_scale, true);
    }
    int kind() const noexcept { return kind_; }
    double split_scale() const noexcept { return split_scale_; }
    bool deconvolves_cic() const noexcept { return deconvolve_cic_; }

private:
    PMForceMethod(int kind, double split_scale, bool deconvolve_cic)
        : kind_(kind), split_scale_(split_scale), deconvolve_cic_(deconvolve_cic) {}
    int kind_;
    double split_scale_;
    bool deconvolve_cic_;
};

// mesh::MeshGeometry (include/cosmo_nbody/mesh/mesh_geometry.hpp:18-72): the
// periodic box and its cells per axis.
class MeshGeometry {
public:
    MeshGeometry(double box_size, std::size_t grid_size) : L_(box_size), N_(grid_size) {
        if (!(box_size > 0.0) || box_size == std::numeric_limits<double>::infinity())
            throw std::invalid_argument("MeshGeometry box size must be finite and positive");
        if (grid_size == 0) throw std::invalid_argument("MeshGeometry grid size must be positive");
    }
    double box_size() const noexcept { return L_; }
    std::size_t grid_size() const noexcept { return N_; }
    double cell_size() const noexcept { return L_ / static_cast<double>(N_); }

private:
    double L_;
    std::size_t N_;
};

}  // namespace mesh

namespace config {

// config::MemoryPolicyParams (include/cosmo_nbody/config/memory_policy.hpp):
// scratch placement.  The GPU workspace lives in HBM whatever the mode says.
enum class ScratchMode { Memory, Disk };
struct MemoryPolicyParams {
    ScratchMode ic_scratch_mode{ScratchMode::Memory};
    ScratchMode evolution_scratch_mode{ScratchMode::Memory};
    std::string scratch_directory{};
};

}  // namespace config

namespace gravity {

// gravity::PMForceDiagnostics (include/cosmo_nbody/gravity/pm_solver.hpp:30-35).
struct PMForceDiagnostics {
    double potential_energy_comoving{0.0};
    std::optional<double> cic_self_energy_comoving;
};

// gravity::PMSolver: pure PM (discrete-Laplacian Green function,
// centered-gradient CIC gather) or the TreePM long-range force.
class PMSolver {
public:
    PMSolver(double box_size, std::size_t mesh_size, bool deconvolve_cic,
             std::size_t max_particles, Precision precision = Precision::FP32Mixed,
             int device = 0, void* stream = nullptr)
        : PMSolver(box_size, mesh_size, mesh::PMForceMethod::pure_pm(deconvolve_cic), max_particles, precision,
                   device, stream) {}
    PMSolver(double box_size, std::size_t mesh_size, mesh::PMForceMethod method, std::size_t max_particles,
             Precision precision = Precision::FP32Mixed, int device = 0, void* stream = nullptr) {
        desc_ = describe(box_size, mesh_size, method, precision, device, stream);
        create(max_particles);
    }
    // The reference's constructor (include/cosmo_nbody/gravity/pm_solver.hpp:53-57).
    // As the reference's ensure_workspace (src/gravity/pm_solver.cpp:388-492),
    // the device workspace is allocated on the first force call, for that
    // call's particle count, and grown when a later call needs more.  The
    // geometry is checked here.  mpi_global_reduce selects the reference's
    // multi-rank force, which this single-GPU build does not provide.  The
    // precision defaults to binary64, the reference's arithmetic.
    PMSolver(const mesh::MeshGeometry& geom, mesh::PMForceMethod method, bool mpi_global_reduce,
             config::MemoryPolicyParams /*memory_policy: the workspace is device memory*/,
             Precision precision = Precision::FP64, int device = 0, void* stream = nullptr) {
        if (mpi_global_reduce)
            throw std::invalid_argument("PMSolver MPI-global reduction is not available on this single-GPU solver");
        desc_ = describe(geom.box_size(), geom.grid_size(), method, precision, device, stream);
        create(1);  // validates the geometry now; resized on the first call
        capacity_ = 0;
    }
    ~PMSolver() {
        if (ctx_) hy_pm_destroy(ctx_);
    }
    PMSolver(const PMSolver&) = delete;
    PMSolver& operator=(const PMSolver&) = delete;
    PMSolver(PMSolver&& o) noexcept : desc_(o.desc_), ctx_(o.ctx_), capacity_(o.capacity_) { o.ctx_ = nullptr; }

    // Adds the PM acceleration into acc_* (reference semantics: transactional,
    // explicit masses when `masses` is non-empty, else the uniform mass).  With
    // collect_potential_energy the result carries 0.5 sum m phi
    // (src/gravity/pm_solver.cpp:817-841); otherwise it is empty.
    std::optional<PMForceDiagnostics> compute_forces(std::span<const double> pos_x, std::span<const double> pos_y,
                                                     std::span<const double> pos_z, std::span<const double> masses,
                                                     std::optional<double> uniform_mass, std::span<double> acc_x,
                                                     std::span<double> acc_y, std::span<double> acc_z,
                                                     bool collect_potential_energy = false) {
        const std::size_t n = pos_x.size();
        if (pos_y.size() != n || pos_z.size() != n || acc_x.size() != n || acc_y.size() != n ||
            acc_z.size() != n)
            throw std::invalid_argument("PMSolver position and acceleration component sizes must match");
        if (!uniform_mass && masses.size() != n)
            throw std::invalid_argument("PMSolver explicit mass count must match particle count");
        if (n > capacity_) create(n);
        const uint32_t flags = HY_HOST_POINTERS | (collect_potential_energy ? HY_COLLECT_POTENTIAL : 0u);
        throw_status(hy_pm_force(ctx_, n, pos_x.data(), pos_y.data(), pos_z.data(),
                                 uniform_mass ? nullptr : masses.data(),
                                 uniform_mass.value_or(0.0), acc_x.data(), acc_y.data(),
                                 acc_z.data(), flags));
        if (!collect_potential_energy) return std::nullopt;
        PMForceDiagnostics d;
        throw_status(hy_pm_potential_energy(ctx_, &d.potential_energy_comoving));
        return d;
    }

    hy_pm_ctx* handle() const noexcept { return ctx_; }

private:
    static hy_pm_desc describe(double box_size, std::size_t mesh_size, const mesh::PMForceMethod& method,
                               Precision precision, int device, void* stream) {
        hy_pm_desc d{};
        d.box_size = box_size;
        d.mesh_size = mesh_size;
        d.deconvolve_cic = method.deconvolves_cic() ? 1 : 0;
        d.precision = static_cast<int>(precision);
        d.device = device;
        d.stream = stream;
        d.method = method.kind();
        d.split_scale = method.split_scale();
        return d;
    }
    // (re)creates the context for `particles` (the old one is kept on failure)
    void create(std::size_t particles) {
        hy_pm_desc d = desc_;
        d.max_particles = particles;
        hy_pm_ctx* c = nullptr;
        throw_status(hy_pm_create(&d, &c));
        if (ctx_) hy_pm_destroy(ctx_);
        ctx_ = c;
        capacity_ = particles;
    }

    hy_pm_desc desc_{};
    hy_pm_ctx* ctx_ = nullptr;
    std::size_t capacity_ = 0;
};

}  // namespace gravity

namespace cosmology {

class DriftKickIntegrator {
public:
    DriftKickIntegrator(double omega_m, double omega_lambda) : om_(omega_m), ol_(omega_lambda) {}
    double drift_factor(double a1, double a2) const {
        double v = 0.0;
        throw_status(hy_drift_factor(a1, a2, om_, ol_, &v));
        return v;
    }
    double kick_factor(double a1, double a2) const {
        double v = 0.0;
        throw_status(hy_kick_factor(a1, a2, om_, ol_, &v));
        return v;
    }

private:
    double om_, ol_;
};

}  // namespace cosmology

namespace time {

struct TimeStep {
    double a_begin, a_half, a_end;
};

class TimeStepper {
public:
    TimeStepper(double a_start, double a_final, double delta_ln_a, std::vector<double> stops = {}) {
        std::vector<double> out(1u << 20);
        uint64_t count = 0;
        throw_status(hy_time_boundaries(a_start, a_final, delta_ln_a, stops.data(), stops.size(),
                                        out.data(), out.size(), &count));
        out.resize(count);
        boundaries_ = std::move(out);
    }
    std::optional<TimeStep> next_step() {
        if (current_ + 1 >= boundaries_.size()) return std::nullopt;
        const double b = boundaries_[current_], e = boundaries_[current_ + 1];
        ++current_;
        return TimeStep{b, hy_geometric_mean(b, e), e};
    }
    const std::vector<double>& boundaries() const { return boundaries_; }
    std::size_t num_steps() const { return boundaries_.empty() ? 0 : boundaries_.size() - 1; }

private:
    std::vector<double> boundaries_;
    std::size_t current_ = 0;
};

// ForceSolverFunc-shaped adapter: returns a callable with the reference's
// callback signature (x, y, z, masses, ax, ay, az, a) that adds the PM force.
template <typename Solver>
auto force_solver(Solver& solver, std::optional<double> uniform_mass) {
    return [&solver, uniform_mass](std::span<const double> x, std::span<const double> y,
                                   std::span<const double> z, std::span<const double> masses,
                                   std::span<double> ax, std::span<double> ay, std::span<double> az,
                                   double /*scale factor: PM force does not depend on a*/) {
        solver.compute_forces(x, y, z, masses, masses.empty() ? uniform_mass : std::nullopt, ax, ay,
                              az);
    };
}

}  // namespace time
}  // namespace hyowon
