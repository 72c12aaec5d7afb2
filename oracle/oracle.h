/*
 * CPU oracle for the HYOWON particle-mesh gravity step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product in
 * hyowon_amd/: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  Nothing in the product links, loads or calls it.
 *
 * It restates, in plain C (binary64, -ffp-contract=off, OpenMP over independent
 * planes/lines), the reference's pure-PM force refresh and kick-drift-kick
 * update, the mode-keyed Gaussian IC field, the matter power spectrum and the
 * TreePM long-range PM force.  Every function cites the reference file:line it follows (paths are
 * relative to the reference source tree).
 *
 * Pinning (what makes each function trustworthy; see tests/test_oracle_*.py):
 *   or_wrap / or_minimum_image / or_cic_coordinate : reference build (oracle/_ref)
 *       on boundary cases + brute force invariants.
 *   or_xplane_order / or_bucket_order  : brute-force stable sort (numpy argsort
 *       kind="stable") on the same keys; bit-exact.
 *   or_deposit        : reference CICMassAssignment::deposit via oracle/_ref,
 *       bit-exact; single-particle closed form (8 weights).
 *   or_density_contrast : reference DensityNormalization via oracle/_ref.
 *   or_fft_r2c_3d / or_fft_c2r_3d : numpy.fft.rfftn/irfftn (pocketfft) and an
 *       O(N^4) long-double DFT on tiny meshes.
 *   or_green_apply    : reference GreenFunction::apply via oracle/_ref, bit-exact.
 *   or_gather_centered_gradient_add : reference
 *       interpolate_centered_gradient3_add_full_mesh via oracle/_ref, bit-exact;
 *       plane-wave closed form.
 *   or_kick_drift / or_kick : reference particle_update.hpp via oracle/_ref.
 *   or_drift_factor / or_kick_factor : reference DriftKickIntegrator via
 *       oracle/_ref (bit-exact) and Einstein-de Sitter closed forms.
 *   or_time_boundaries / or_geometric_mean : reference TimeStepper via oracle/_ref.
 *   or_pm_force       : composition of the above; checked against the _ref
 *       composition (reference deposit/Green/gather + this FFT).
 *   or_philox4x32_10 / or_uniform_pair / or_normal_pair : Random123
 *       known-answer vectors and the reference's philox_rng.cpp via oracle/_ref,
 *       bit-exact; or_gaussian_modes: Hermitian symmetry, band limit and
 *       per-mode variance statistics.
 *   or_pk_edges / or_pk_bin_index / or_pk_window_axis / or_pk_raw_power /
 *   or_pk_raw_power_minus : the reference's header-only spectral helpers via
 *       oracle/_ref, bit-exact (the subtraction's exact branch included).
 *   or_power_spectrum : closed forms (alternating planes, period-4 pattern
 *       with its interlacing partner), brute-force shell counts and mean |k|,
 *       mass-scale and translation invariance, the Poisson expectation of the
 *       aliased shot noise; the normalised deposit via oracle/_ref.
 *   or_green_apply_treepm / or_interpolate_add : reference GreenFunction
 *       (TreePM branch) and CICMassAssignment::interpolate_add via oracle/_ref,
 *       bit-exact.
 *   or_pm_force_treepm : the Gaussian-filtered Newtonian pair force (2%) and
 *       momentum conservation.
 */
#ifndef HYOWON_ORACLE_H
#define HYOWON_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    OR_OK = 0,
    OR_EINVAL = 1,    /* std::invalid_argument in the reference */
    OR_ELOGIC = 2,    /* std::logic_error */
    OR_EOVERFLOW = 3, /* std::overflow_error */
    OR_EUNDERFLOW = 4,/* std::underflow_error */
    OR_ERUNTIME = 6   /* std::runtime_error */
};

/* Physical constants restated from include/cosmo_nbody/cosmology/units.hpp:14-30. */
double or_gravitational_constant(void);
double or_rho_crit0(void);

/* include/cosmo_nbody/math/periodic_box.hpp:27-41 and :44-61. */
double or_wrap(double x, double L);
double or_minimum_image(double d, double L);

/* src/mesh/cic_coordinate.hpp:95-121 (CICCoordinateAxis::operator()).
 * upper = nextafter((double)N, 0).  Returns OR_OK or OR_ELOGIC. */
int or_cic_coordinate(double position, double L, double dx, uint64_t N,
                      int64_t* base, double* fraction);

/* src/mesh/mass_assignment_workspace.inc:264-374 (build_x_plane_bins):
 * stable counting sort of particle indices by wrapped base x-cell.
 * offsets has N+1 entries, order has n entries. */
int or_xplane_order(uint64_t n, const double* x, double L, uint64_t N,
                    uint64_t* offsets, uint32_t* order);

/* Deposit bucket key used by the HIP deposit (no reference counterpart; the
 * reference sorts by x-plane only): tile edge T divides N, tile = base cell / T
 * per axis, key = ((tx*nT+ty)*nT+tz)*8 + class, class bit2/1/0 set when the
 * base cell is the last cell of its tile along x/y/z. */
uint64_t or_bucket_tile_edge(uint64_t N);
int or_bucket_order(uint64_t n, const double* x, const double* y, const double* z,
                    double L, uint64_t N, uint64_t T,
                    uint64_t* offsets /* nbuckets+1 */, uint32_t* order);

/* src/mesh/mass_assignment_deposit.inc:15-334 (deposit_impl), serial mesh,
 * reference per-cell accumulation order.  masses==NULL selects uniform_mass.
 * mesh has N^3 entries, row-major (ix*N+iy)*N+iz. */
int or_deposit(uint64_t n, const double* x, const double* y, const double* z,
               const double* masses, double uniform_mass, double L, uint64_t N,
               double* mesh);

/* Exact (single-rounding) sum of non-negative doubles; restates the contract of
 * include/cosmo_nbody/math/exact_positive_sum.hpp (value rounded once, ties to
 * even).  Returns NaN for a negative or non-finite term. */
double or_exact_sum(const double* values, uint64_t count);
double or_exact_repeated(double value, uint64_t count);

/* src/math/cic_mass_conservation.cpp:25-63.  Returns OR_OK or OR_ERUNTIME. */
int or_require_cic_mass_conservation(double particle_mass, double deposited_mass,
                                     uint64_t particle_count);

/* src/gravity/pm_solver.cpp:700-759 with
 * include/cosmo_nbody/math/density_normalization.hpp:32-120.
 * Converts a deposited mass mesh to (m - M/N_c)/dx^3 in place. */
int or_density_contrast(double* mesh, uint64_t N, double dx, double total_mass);

/* FFT conventions of src/mesh/fft_backend.cpp:1088-1145: forward r2c
 * unnormalized; inverse c2r multiplied by 1/((N*N)*N) in a separate pass.
 * Complex layout is interleaved re,im with index iz+(N/2+1)*(iy+N*ix)
 * (include/cosmo_nbody/mesh/mesh_geometry.hpp:90-96). */
int or_fft_r2c_3d(uint64_t N, const double* real, double* cplx);
int or_fft_c2r_3d(uint64_t N, const double* cplx, double* real);
/* 1-D complex DFT of length n (sign -1 forward, +1 backward), in place. */
int or_fft_1d(uint64_t n, double* data, int sign);

/* src/mesh/green_function.cpp:15-135 (tables), :165-207 (evaluate_kernel),
 * :209-241 (apply); pure PM (discrete Laplacian), optional CIC deconvolution. */
int or_green_apply(uint64_t N, double L, int deconvolve_cic, double* cplx);
double or_green_kernel(uint64_t N, double L, int deconvolve_cic,
                       uint64_t ix, uint64_t iy, uint64_t iz);

/* src/mesh/mass_assignment_gradient.cpp:46-226: fused centered-gradient CIC
 * gather; adds into ax/ay/az. */
int or_gather_centered_gradient_add(uint64_t n, const double* phi,
                                    const double* x, const double* y, const double* z,
                                    double L, uint64_t N,
                                    double* ax, double* ay, double* az);

/* src/gravity/pm_solver.cpp:569-852 (compute_forces_in_place, pure PM, serial
 * mesh): deposit, exact mass identity and conservation check, density source,
 * r2c, Green, c2r, centered-gradient gather.  Adds into ax/ay/az.
 * phi_out (optional, N^3) receives the real-space potential. */
int or_pm_force(uint64_t n, const double* x, const double* y, const double* z,
                const double* masses, double uniform_mass,
                double L, uint64_t N, int deconvolve_cic,
                double* ax, double* ay, double* az, double* phi_out);

/* include/cosmo_nbody/time/particle_update.hpp:20-116 and
 * src/time/leapfrog_integrator.cpp:143-279: K1 kick + periodic drift
 * (preflight then publish), and K2 kick. */
int or_kick_drift(uint64_t n, double* x, double* y, double* z,
                  double* px, double* py, double* pz,
                  const double* ax, const double* ay, const double* az,
                  double K, double D, double L);
int or_kick(uint64_t n, double* px, double* py, double* pz,
            const double* ax, const double* ay, const double* az, double K);

/* src/cosmology/cosmology_model.cpp:25-57,167-178 (H(a) for flat matter+Lambda)
 * and src/cosmology/drift_kick_factors.cpp:28-141 with
 * include/cosmo_nbody/math/roundoff_resolved_gauss_legendre_16.hpp. */
double or_hubble(double a, double omega_m, double omega_lambda);
int or_drift_factor(double a1, double a2, double omega_m, double omega_lambda, double* out);
int or_kick_factor(double a1, double a2, double omega_m, double omega_lambda, double* out);

/* include/cosmo_nbody/math/exact_geometric_mean.hpp:168-275 (correctly rounded
 * sqrt(lhs*rhs)). */
double or_geometric_mean(double lhs, double rhs);

/* src/time/time_stepper.cpp:194-276: boundary table. Writes at most cap
 * boundaries, returns the count through *count. */
int or_time_boundaries(double a_start, double a_final, double delta_ln_a,
                       const double* stops, uint64_t n_stops,
                       double* out, uint64_t cap, uint64_t* count);

/* Full KDK evolution (src/runtime/simulation_runner.cpp:1348-1527 step loop
 * with src/time/leapfrog_integrator.cpp:143-279), pure PM, diagnostics off:
 * initial force at boundaries[0], then per step K1+drift, force at a_end, K2.
 * Accelerations are left as the last force. */
int or_run(uint64_t n, double* x, double* y, double* z,
           double* px, double* py, double* pz, double* ax, double* ay, double* az,
           const double* masses, double uniform_mass, double L, uint64_t N,
           int deconvolve_cic, double omega_m, double omega_lambda,
           const double* boundaries, uint64_t n_boundaries);

/* Mode-keyed Gaussian initial field (src/ic/philox_rng.cpp:1-90,
 * src/ic/random_field.cpp:70-330).
 * or_philox4x32_10: Random123 Philox4x32 with ten rounds (key bumped by the
 * Weyl constants between rounds).
 * or_uniform_pair: the block of counter (kx, ky, kz, stream), key (seed lo,
 * seed hi), mapped to two uniforms in (0,1) on the 53-bit grid (0 -> 2^-54).
 * or_normal_pair: Box-Muller of that pair.
 * or_gaussian_modes: the r2c half spectrum [i][j][k] (k < N/2+1) of an N^3 IC
 * mesh: zero mode and modes with any |m_axis| > max_mode are zero; on the
 * k = 0 and Nyquist planes the mode with the lexicographically smaller
 * conjugate (i, j) is drawn as the conjugate of that mode's draw; a
 * self-conjugate mode is real with standard deviation sqrt(var); otherwise
 * both parts have standard deviation sqrt(var / 2).  var comes from
 * variance_by_m2[mx^2 + my^2 + mz^2].  flags bit 0: fixed amplitude (phase
 * from stream 1), bit 1: reversed pair (negated field).  out holds
 * N * N * (N/2+1) complex values (re, im). */
void or_philox4x32_10(const uint32_t* counter, const uint32_t* key, uint32_t* out);
void or_uniform_pair(uint64_t seed, int32_t kx, int32_t ky, int32_t kz, uint32_t stream, double* out);
void or_normal_pair(uint64_t seed, int32_t kx, int32_t ky, int32_t kz, uint32_t stream, double* out);
int or_gaussian_modes(uint64_t N, uint64_t seed, uint64_t max_mode, const double* variance_by_m2,
                      uint32_t flags, double* out);

/* TreePM long-range PM (PMForceMethod::treepm_long_range):
 * or_green_apply_treepm: continuum Green 4 pi G / (-k^2) * exp(-k^2 r_s^2) / W^2
 *   (src/mesh/green_function.cpp:165-203).
 * or_interpolate_add: CIC interpolation added into out
 *   (src/mesh/mass_assignment_gather.inc:95-181,402-451).
 * or_pm_force_treepm: the long-range force with the spectral gradient
 *   (src/gravity/pm_solver.cpp:768-810); mesh_out (optional) gets 3 N^3. */
int or_green_apply_treepm(uint64_t N, double L, double split_scale, double* cplx);
int or_interpolate_add(uint64_t n, const double* field, const double* x, const double* y,
                       const double* z, double L, uint64_t N, double* out);
int or_pm_force_treepm(uint64_t n, const double* x, const double* y, const double* z,
                       const double* masses, double uniform_mass, double L, uint64_t N,
                       double split_scale, double* ax, double* ay, double* az, double* mesh_out);

/* Matter power spectrum (src/analysis/matter_power_spectrum.cpp:57-241 over
 * src/analysis/fourier_density_field.cpp:112-343).  masses == NULL selects the
 * uniform mass.  out: num_bins rows of 8 doubles (k_low, k_mean, k_high, p_raw,
 * p_corrected, mean CIC power deconvolution factor, shot-noise reference,
 * mode count).  Helpers: the logarithmic bin edges and their inclusive lookup
 * (analysis/spectral_numeric.hpp:93-141), the per-axis CIC window
 * (analysis/cic_spectral_window.hpp:13-37), |v|^2 L^3 and |v|^2 L^3 - s
 * (spectral_numeric.hpp:304-391). */
int or_power_spectrum(uint64_t n, const double* x, const double* y, const double* z,
                      const double* masses, double uniform_mass, double L, uint64_t N, int num_bins,
                      int subtract_shot_noise, int interlaced, double max_fraction, double* out);
/* or_lpt_first_order: first-order LPT lattice ICs of an r2c density half
 * spectrum delta_k [N][N][N/2+1] (complex128): per axis psi = c2r(i k_d/k^2
 * delta_k) (k_d = 0 on the Nyquist index), particle (i,j,k) of the n_side^3
 * lattice at q = index * L/n_side moves to wrap(q + psi(i s, j s, k s)),
 * p = p_factor * psi.  Pinned by a single-mode closed form and the discrete
 * divergence against numpy's FFT (tests/test_oracle_ic_field.py). */
int or_lpt_first_order(uint64_t N, const double* delta_k, uint64_t n_side, double L, double p_factor,
                       double* x, double* y, double* z, double* px, double* py, double* pz);
/* or_potential_energy: 0.5 sum_p m_p phi(x_p), the reference's binary64
 * blocked Neumaier sum; bit-exact against the reference build
 * (tests/test_oracle_reference.py). */
int or_potential_energy(uint64_t n, const double* field, const double* x, const double* y, const double* z,
                        const double* masses, double uniform_mass, double L, uint64_t N, double* out);
int or_pk_edges(double minimum, double maximum, int bins, double* edges);
int or_pk_bin_index(const double* edges, int bins, double v);
long double or_pk_window_axis(double k, double dx);
double or_pk_raw_power(double L, double re, double im);
double or_pk_raw_power_minus(double L, double re, double im, double s);

/* Thread count used by OpenMP regions (0 = runtime default). */
void or_set_threads(int threads);
int or_get_threads(void);

/* Wall time per stage of or_deposit, or_pm_force and or_run, accumulated in
 * seconds since the last or_stage_reset: x-plane sort, deposit, exact mass
 * check and density contrast, forward FFT, Green multiply, inverse FFT,
 * gradient gather, kick/drift passes. */
#define OR_STAGES 8
enum {
    OR_STAGE_SORT = 0,
    OR_STAGE_DEPOSIT,
    OR_STAGE_MASS_CONTRAST,
    OR_STAGE_FFT_FORWARD,
    OR_STAGE_GREEN,
    OR_STAGE_FFT_INVERSE,
    OR_STAGE_GATHER,
    OR_STAGE_KDK
};
void or_stage_reset(void);
void or_stage_times(double* out);

#ifdef __cplusplus
}
#endif
#endif
