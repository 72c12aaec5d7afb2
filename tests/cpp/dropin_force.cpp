// Reference-shaped C++ caller of the MI355X PM step through include/hyowon/pm_solver.hpp:
// reads particles from a binary file, computes PM forces with
// hyowon::gravity::PMSolver::compute_forces on host spans (as a caller of
// cosmo_nbody::gravity::PMSolver would), runs the time schedule through
// hyowon::time::TimeStepper / hyowon::cosmology::DriftKickIntegrator and
// writes the accelerations and the per-step (K1, D, K2) factors.
//
// usage: dropin_force <in.bin> <out.bin> <L> <N> <precision 0|1> <uniform_mass>
//   in.bin : uint64 n, then x[n], y[n], z[n] (binary64)
//   out.bin: ax[n], ay[n], az[n], uint64 steps, then steps x (K1, D, K2), then the
//            forces of the reference-shaped solver (two calls, halved) and its
//            collected potential energy
#include "hyowon/pm_solver.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <vector>

int main(int argc, char** argv) {
    if (argc != 7) {
        std::fprintf(stderr, "usage: %s in.bin out.bin L N precision uniform_mass\n", argv[0]);
        return 2;
    }
    try {
        std::FILE* f = std::fopen(argv[1], "rb");
        if (!f) throw std::runtime_error("cannot open input");
        uint64_t n = 0;
        if (std::fread(&n, 8, 1, f) != 1) throw std::runtime_error("short input");
        std::vector<double> x(n), y(n), z(n), ax(n, 0.0), ay(n, 0.0), az(n, 0.0);
        if (std::fread(x.data(), 8, n, f) != n || std::fread(y.data(), 8, n, f) != n ||
            std::fread(z.data(), 8, n, f) != n)
            throw std::runtime_error("short input");
        std::fclose(f);
        const double L = std::atof(argv[3]);
        const std::size_t N = std::strtoul(argv[4], nullptr, 10);
        const auto prec = std::atoi(argv[5]) ? hyowon::Precision::FP32Mixed : hyowon::Precision::FP64;
        const double m = std::atof(argv[6]);

        hyowon::gravity::PMSolver solver(L, N, /*deconvolve_cic=*/false, n, prec);
        auto force = hyowon::time::force_solver(solver, m);
        force(x, y, z, {}, ax, ay, az, 1.0);

        hyowon::cosmology::DriftKickIntegrator dk(0.3, 0.7);
        hyowon::time::TimeStepper stepper(1.0 / 50.0, 1.0, std::log(50.0) / 32.0, {1.0});
        std::vector<double> factors;
        while (auto s = stepper.next_step()) {
            factors.push_back(dk.kick_factor(s->a_begin, s->a_half));
            factors.push_back(dk.drift_factor(s->a_begin, s->a_end));
            factors.push_back(dk.kick_factor(s->a_half, s->a_end));
        }
        // the reference-shaped constructor (geometry, method, MPI flag, memory
        // policy): workspace allocated on the first call, potential energy collected
        hyowon::gravity::PMSolver shaped(hyowon::mesh::MeshGeometry(L, N),
                                         hyowon::mesh::PMForceMethod::pure_pm(false), false,
                                         hyowon::config::MemoryPolicyParams{}, prec);
        std::vector<double> bx(n, 0.0), by(n, 0.0), bz(n, 0.0);
        const auto diag = shaped.compute_forces(x, y, z, {}, m, bx, by, bz, true);
        if (!diag) throw std::logic_error("potential energy was not collected");
        if (shaped.compute_forces(x, y, z, {}, m, bx, by, bz)) throw std::logic_error("unrequested diagnostics");
        for (std::size_t i = 0; i < n; ++i) {
            bx[i] *= 0.5;
            by[i] *= 0.5;
            bz[i] *= 0.5;
        }
        bool mpi_threw = false;
        try {
            hyowon::gravity::PMSolver mpi(hyowon::mesh::MeshGeometry(L, N),
                                          hyowon::mesh::PMForceMethod::pure_pm(false), true,
                                          hyowon::config::MemoryPolicyParams{});
        } catch (const std::invalid_argument&) {
            mpi_threw = true;
        }
        if (!mpi_threw) throw std::logic_error("an MPI-global solver was accepted");

        // an invalid mesh size must surface as the reference's exception class
        bool threw = false;
        try {
            hyowon::gravity::PMSolver bad(L, 10, false, 4);
        } catch (const std::invalid_argument&) {
            threw = true;
        }
        if (!threw) throw std::logic_error("invalid mesh size was accepted");

        std::FILE* o = std::fopen(argv[2], "wb");
        if (!o) throw std::runtime_error("cannot open output");
        std::fwrite(ax.data(), 8, n, o);
        std::fwrite(ay.data(), 8, n, o);
        std::fwrite(az.data(), 8, n, o);
        const uint64_t steps = factors.size() / 3;
        std::fwrite(&steps, 8, 1, o);
        std::fwrite(factors.data(), 8, factors.size(), o);
        std::fwrite(bx.data(), 8, n, o);
        std::fwrite(by.data(), 8, n, o);
        std::fwrite(bz.data(), 8, n, o);
        std::fwrite(&diag->potential_energy_comoving, 8, 1, o);
        std::fclose(o);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "dropin_force: %s\n", e.what());
        return 1;
    }
    return 0;
}
