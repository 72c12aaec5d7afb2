/*
 * hyowon_pm.h -- C ABI of the MI355X-native particle-mesh gravity step.
 *
 * libhyowon_pm.so implements HYOWON's pure-PM force refresh and kick-drift-kick
 * particle update with hand-written HIP kernels for gfx950.  Signatures use
 * plain pointers and sizes only.  Reference interfaces are cited as
 * file:line relative to the HYOWON source tree.
 *
 * Conventions
 *  - Return value: HY_OK (0) or an hy_status code; the message of the last
 *    failure on the calling thread is returned by hy_last_error().  Codes map
 *    one-to-one onto the reference's exception classes (invalid_argument,
 *    logic_error, overflow_error, underflow_error, runtime_error).
 *  - Particle arrays are SoA, binary64 positions in [0, L) (any finite value is
 *    wrapped as the reference does).  Pointers are device pointers unless
 *    HY_HOST_POINTERS is set.
 *  - Device inputs are read in stream order on the context's stream: work on
 *    the legacy default stream is ordered before the private stream's; work on
 *    other streams must be ordered by the caller (or pass that stream).
 *  - One context owns one HIP stream and its mesh workspace; calls on a
 *    context must be serialised (as PMSolver, include/cosmo_nbody/mesh/
 *    mass_assignment.hpp:44-45).  All device memory is allocated by
 *    hy_pm_create; the force and step entry points allocate nothing (except the
 *    deposit fallback scratch growing after an HY_ERUNTIME-free retry).
 */
#ifndef HYOWON_PM_H
#define HYOWON_PM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HY_ABI_VERSION 3

typedef enum {
    HY_OK = 0,
    HY_EINVAL = 1,      /* std::invalid_argument */
    HY_ELOGIC = 2,      /* std::logic_error */
    HY_EOVERFLOW = 3,   /* std::overflow_error */
    HY_EUNDERFLOW = 4,  /* std::underflow_error */
    HY_EDEVICE = 5,     /* HIP runtime failure */
    HY_ERUNTIME = 6,    /* std::runtime_error */
    HY_ENOMEM = 7       /* std::bad_alloc */
} hy_status;

typedef enum {
    HY_PREC_FP64 = 0,        /* binary64 mesh, FFT, forces, momenta */
    HY_PREC_FP32_MIXED = 1   /* binary32 mesh/FFT/forces/momenta, binary64 positions */
} hy_precision;

/* PM discretisation (include/cosmo_nbody/mesh/pm_force_method.hpp:10-60) */
typedef enum {
    HY_METHOD_PURE_PM = 0,            /* discrete Laplacian Green, centered real-space gradient */
    HY_METHOD_TREEPM_LONG_RANGE = 1   /* continuum Green x exp(-k^2 r_s^2), W^-2, spectral gradient */
} hy_method;

/* Flags */
#define HY_HOST_POINTERS  0x1u  /* particle arrays live in host memory */
#define HY_HAS_PREV       0x2u  /* hy_pm_kdk: apply K2 of the previous step */
#define HY_HAS_NEXT       0x4u  /* hy_pm_kdk: apply K1 of the next step and drift */
#define HY_ASYNC          0x8u  /* do not synchronise; errors are reported by hy_pm_check */
#define HY_REUSE_KEYS     0x10u /* hy_pm_kdk: x_in/y_in/z_in are the previous hy_pm_kdk call's
                                   outputs on this context; reuse the bucket keys it made */
#define HY_COLLECT_POTENTIAL 0x20u /* hy_pm_force: also evaluate 0.5 sum m phi(x) (pure PM,
                                      single box); read it with hy_pm_potential_energy */

typedef struct {
    double box_size;        /* L (Mpc/h) */
    uint64_t mesh_size;     /* N cells per axis: 2^a (8..1024) or 3*2^a (12..384) */
    int deconvolve_cic;     /* divide the Green function by W(k)^2 */
    int precision;          /* hy_precision */
    int device;             /* HIP device ordinal */
    void* stream;           /* hipStream_t to use, or NULL for a private blocking stream
                               (ordered with the legacy default stream) */
    uint64_t max_particles; /* capacity (all boxes); particle counts above it are rejected */
    uint32_t boxes;         /* independent boxes of this geometry held by the context (0 or
                               1 = one box).  Particle arrays of hy_pm_force / hy_pm_kdk then
                               concatenate the boxes (n / boxes particles each, uniform mass);
                               every box has its own mesh and its own mass identity. */
    int method;             /* hy_method; HY_METHOD_TREEPM_LONG_RANGE deconvolves always, takes
                               a single box and serves hy_pm_force / hy_pm_force_meshes only */
    double split_scale;     /* r_s of HY_METHOD_TREEPM_LONG_RANGE (finite, positive) */
} hy_pm_desc;

typedef struct {
    uint64_t mesh_size;
    uint64_t tile_edge;        /* deposit tile edge T (divides N, <= 16) */
    uint64_t bucket_count;     /* (N/T)^3 * 8 */
    uint64_t device_bytes;     /* device memory held by the context */
    double last_input_mass;    /* exact input mass of the last deposit */
    double last_deposited_mass;/* deposited mass of the last deposit */
    int precision;
    uint32_t boxes;            /* boxes held by the context */
} hy_pm_info;

typedef struct hy_pm_ctx hy_pm_ctx;

/* Replaces gravity::PMSolver::PMSolver + ensure_workspace
 * (include/cosmo_nbody/gravity/pm_solver.hpp:53-57; src/gravity/pm_solver.cpp:388-492):
 * validates geometry, allocates every device buffer up front. */
int hy_pm_create(const hy_pm_desc* desc, hy_pm_ctx** out);
int hy_pm_destroy(hy_pm_ctx* ctx);
int hy_pm_get_info(const hy_pm_ctx* ctx, hy_pm_info* info);
const char* hy_last_error(void);
int hy_abi_version(void);

/* Replaces gravity::PMSolver::compute_forces
 * (include/cosmo_nbody/gravity/pm_solver.hpp:61-70; src/gravity/pm_solver.cpp:494-852):
 * adds the pure-PM acceleration into ax/ay/az.  masses == NULL selects the
 * uniform mass.  Transactional: on error ax/ay/az are untouched. */
int hy_pm_force(hy_pm_ctx* ctx, uint64_t n, const double* x, const double* y, const double* z,
                const double* masses, double uniform_mass, double* ax, double* ay, double* az,
                uint32_t flags);

/* One force refresh fused with the surrounding kicks and drift of
 * time::LeapfrogIntegrator::step (src/time/leapfrog_integrator.cpp:143-279,
 * include/cosmo_nbody/time/particle_update.hpp:20-116), device pointers only:
 *   g = PM force at (x_in, y_in, z_in)
 *   HY_HAS_PREV: p += g * k_prev          (K2 of the step that ends here)
 *   HY_HAS_NEXT: p += g * k_next          (K1 of the step that starts here)
 *                x = wrap(wrap(x) + minimum_image(p * d_next))
 * Momenta are binary64 (HY_PREC_FP64) or binary32 (HY_PREC_FP32_MIXED).
 * Outputs may alias inputs.  acc_x/y/z (optional, binary64) receive g.
 * With HY_HAS_NEXT the call also computes the deposit bucket keys of the
 * drifted positions; HY_REUSE_KEYS on the next call skips recomputing them.
 * Reused keys are checked against the positions: synchronous calls fall back
 * to fresh keys when they differ, HY_ASYNC calls report logic_error through
 * hy_pm_check (outputs of that call are then undefined).
 * HY_ASYNC calls with HY_HAS_NEXT on a binary32 context with uniform mass
 * carry more: the deposit records of the drifted positions, written by the
 * gather into per-bucket capacity slots (overflow placed before the next
 * deposit), and the next HY_ASYNC | HY_REUSE_KEYS call deposits them without
 * re-reading the positions.  Results are bit-identical to routing afresh; the
 * caller must pass that call's output positions unmodified, since a change
 * made to them in place is not detected on this path (synchronous calls route
 * from the positions). */
int hy_pm_kdk(hy_pm_ctx* ctx, uint64_t n, const double* x_in, const double* y_in,
              const double* z_in, const void* px_in, const void* py_in, const void* pz_in,
              double* x_out, double* y_out, double* z_out, void* px_out, void* py_out,
              void* pz_out, const double* masses, double uniform_mass, double k_prev,
              double k_next, double d_next, double* acc_x, double* acc_y, double* acc_z,
              uint32_t flags);

/* PMForceDiagnostics::potential_energy_comoving of the last hy_pm_force call
 * made with HY_COLLECT_POTENTIAL (include/cosmo_nbody/gravity/pm_solver.hpp:30-35;
 * src/gravity/pm_solver.cpp:817-841 over src/mesh/mass_assignment_gather.inc:
 * 283-399): 0.5 sum_i m_i phi(x_i), phi CIC-interpolated from the force's
 * potential, summed in binary64 in a fixed order.  logic_error if that call did
 * not collect it. */
int hy_pm_potential_energy(const hy_pm_ctx* ctx, double* potential_energy_comoving);

/* Stage entry points (device pointers; mesh arrays use the context precision,
 * row-major (ix*N+iy)*N+iz as include/cosmo_nbody/mesh/mesh_geometry.hpp:74-79). */

/* Deposit + density source (src/mesh/mass_assignment_deposit.inc:15-334,
 * src/gravity/pm_solver.cpp:649-759): writes (m - M/N_c)/dx^3 to mesh_out. */
int hy_pm_density(hy_pm_ctx* ctx, uint64_t n, const double* x, const double* y, const double* z,
                  const double* masses, double uniform_mass, void* mesh_out);

/* Poisson solve of a density-source mesh: r2c, Green, c2r with 1/N^3
 * (src/mesh/fft_backend.cpp:1088-1145, src/mesh/green_function.cpp:209-241). */
int hy_pm_potential(hy_pm_ctx* ctx, const void* source_in, void* phi_out);

/* TreePM long-range force meshes of a density-source mesh: r2c, continuum
 * Green x exp(-k^2 r_s^2) / W^2, then per axis -i k phi (k = 0 on the Nyquist
 * index) and c2r with 1/N^3 (src/mesh/green_function.cpp:165-203,
 * src/gravity/pm_solver.cpp:768-810).  HY_METHOD_TREEPM_LONG_RANGE contexts. */
int hy_pm_force_meshes(hy_pm_ctx* ctx, const void* source_in, void* fx_out, void* fy_out, void* fz_out);

/* CIC interpolation of a mesh (context precision) added into out (binary64):
 * CICMassAssignment::interpolate_add (src/mesh/mass_assignment_gather.inc:181-196,402-451). */
int hy_pm_interpolate(hy_pm_ctx* ctx, uint64_t n, const double* x, const double* y, const double* z,
                      const void* mesh, double* out);

/* Fused centered-gradient gather of a given potential, adding into ax/ay/az
 * (src/mesh/mass_assignment_gradient.cpp:46-226). */
int hy_pm_gather(hy_pm_ctx* ctx, uint64_t n, const double* x, const double* y, const double* z,
                 const void* phi, double* ax, double* ay, double* az);

/* Canonical deposit-bucket permutation: particles stably sorted by the
 * bucket key ((tile, boundary class); tile edge from hy_pm_get_info).
 * order has n entries, offsets bucket_count+1 entries (device, uint32). */
int hy_pm_bucket_order(hy_pm_ctx* ctx, uint64_t n, const double* x, const double* y,
                       const double* z, uint32_t* order, uint32_t* offsets);

/* Reference x-plane bins: stable counting sort of particle indices by the
 * wrapped base x-cell (src/mesh/mass_assignment_workspace.inc:264-374).
 * order has n entries, offsets N+1 entries (device, uint32). */
int hy_pm_xplane_order(hy_pm_ctx* ctx, uint64_t n, const double* x, uint32_t* order,
                       uint32_t* offsets);

/* Allocation-free variant of hy_pm_xplane_order.  No reference counterpart:
 * it serves callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * context: box_size and mesh_size (as hy_pm_create accepts them) give the
 * geometry; x (device, n binary64 values), order (n) and offsets (mesh_size+1,
 * uint32) are the caller's.  The kernels' only other memory is one 16 MiB
 * file-scope device array per GPU (plane x chunk counters, scan sums, status),
 * resolved once per device; nothing is allocated or freed.  n must be below
 * 2^32.  Output and errors equal hy_pm_xplane_order's; runs on `stream` (NULL:
 * the legacy default stream) on `device` and synchronises it.  Calls that
 * overlap on one GPU share the array and race: serialise them.  The message
 * of a failure is returned by hy_static_last_error(). */
int hy_xplane_order_static(double box_size, uint64_t mesh_size, uint64_t n, const double* x, uint32_t* order,
                           uint32_t* offsets, int device, void* stream);
const char* hy_static_last_error(void);

/* Waits for outstanding work and reports errors raised by HY_ASYNC calls
 * (each error once: the context's error word is cleared when reported). */
int hy_pm_check(hy_pm_ctx* ctx);

/* Live stage timing with HIP events recorded on the context's stream (the
 * stream every kernel of the context runs on).  No reference counterpart:
 * replaces the reference's steady_clock phase timing
 * (src/runtime/simulation_runner.cpp:772-988) at kernel granularity.
 * hy_pm_profile(enable, max_calls) arms a ring for max_calls force/step calls;
 * hy_pm_profile_read synchronises, returns the summed milliseconds and the
 * number of timed launches per stage, and clears the ring. */
enum {
    HY_STAGE_BUCKET = 0,      /* key + count/rank, scan, scatter */
    HY_STAGE_DEPOSIT = 1,     /* deterministic deposit + density source + mass partials */
    HY_STAGE_FFT_Z = 2,       /* r2c along z */
    HY_STAGE_FFT_Y = 3,       /* forward y */
    HY_STAGE_FFT_X_GREEN = 4, /* forward x, Green, backward x */
    HY_STAGE_FFT_YINV = 5,    /* backward y */
    HY_STAGE_FFT_ZINV = 6,    /* c2r along z */
    HY_STAGE_GATHER_KDK = 7,  /* centered-gradient gather fused with kick/drift */
    HY_STAGE_COUNT = 8
};
int hy_pm_profile(hy_pm_ctx* ctx, int enable, uint64_t max_calls);
int hy_pm_profile_read(hy_pm_ctx* ctx, double* stage_ms, uint64_t* stage_count, uint64_t* calls);

/* Measured launch plan of the FFT passes.  Counterpart of the reference's
 * FFTW_MEASURE planning (src/mesh/fft_backend.cpp:967-1003): each pass has
 * compiled launch variants (rows per workgroup, hence workgroups per CU and
 * the run length of its transposing stores); hy_pm_autotune times every
 * variant on the context's own buffers and keeps the fastest, then picks the
 * kz plane-group size of the y / x+Green / y^-1 passes.  Results do not depend
 * on the plan: every variant computes the same values bit for bit. */
#define HY_FFT_PASSES 5        /* z r2c, y, x+Green+x^-1, y^-1, z c2r */
#define HY_FFT_MAX_VARIANTS 4
typedef struct {
    uint32_t variant[HY_FFT_PASSES];  /* launch variant per pass */
    uint32_t plane_group;             /* kz planes per y/x/y^-1 launch group, 0 = all at once */
    uint32_t variants;                /* variants compiled for this mesh size (1 or 4) */
    uint32_t rows[HY_FFT_PASSES][HY_FFT_MAX_VARIANTS];  /* rows per workgroup of each variant */
    float pass_ms[HY_FFT_PASSES][HY_FFT_MAX_VARIANTS];  /* last autotune timings, 0 = not timed */
    float group_ms[4];                /* last autotune: plane groups 0, 16, 64, 128 */
} hy_fft_tuning;
int hy_pm_get_tuning(hy_pm_ctx* ctx, hy_fft_tuning* out);
int hy_pm_set_tuning(hy_pm_ctx* ctx, const hy_fft_tuning* in);
int hy_pm_autotune(hy_pm_ctx* ctx, uint32_t reps, hy_fft_tuning* out);

/* Gather schedule.  The fused gather walks particle arrays held in Lagrangian
 * lattice order (n = S^3, S a multiple of 256) in column tiles of TX x TY
 * lattice rows, so the blocks resident on one XCD share that XCD's L2.
 * hy_pm_autotune_gather deposits and solves for the given positions (single
 * box, pure PM, uniform mass), times the gather under each candidate number
 * of rows per tile (0 = particle order, 16, 64, 256, 1024; ms_out[5] receives
 * the times, 0 where the lattice does not admit one) and keeps the fastest in
 * the context; hy_pm_set_gather_rows sets it (0, a power of two, or
 * HY_GATHER_ROWS_DEFAULT = 256 / (S / 256)).  Every schedule gives identical
 * results.  No reference counterpart (the reference gathers on CPU threads);
 * the FFT launch shapes are tuned by hy_pm_autotune, as FFTW_MEASURE plans do
 * for the reference (src/mesh/fft_backend.cpp:967-1003). */
#define HY_GATHER_ROWS_DEFAULT 0xffffffffu
#define HY_GATHER_TUNE_CANDIDATES 5
int hy_pm_autotune_gather(hy_pm_ctx* ctx, uint64_t n, const double* x, const double* y,
                          const double* z, double uniform_mass, uint32_t reps,
                          uint32_t* rows_out, float* ms_out);
int hy_pm_set_gather_rows(hy_pm_ctx* ctx, uint32_t rows);

/* Adjoint (vector-Jacobian products) for differentiable runs.  No reference
 * counterpart: the reference is not differentiable.  Device pointers, single-box
 * contexts, uniform mass; adjoint arrays are binary64.
 *
 * hy_pm_kdk_vjp is the exact reverse of one hy_pm_kdk call evaluated at the
 * positions (x, y, z) that call read.  On entry lx.. / lpx.. hold dL/dx_out and
 * dL/dp_out; on return dL/dx_in and dL/dp_in:
 *   u   = dL/dp_out + d_next dL/dx_out         (HY_HAS_NEXT; else dL/dp_out)
 *   dL/dp_in = u
 *   dL/dx_in = dL/dx_out + (k_prev + k_next) J_g(x)^T u
 * where J_g is the Jacobian of the PM acceleration (deposit -> Poisson ->
 * centred-gradient gather).  Periodic wraps have unit derivative.
 *
 * hy_pm_density_vjp adds coef * d/dx_k sum_c u_c m_c(x) to lx.. for a mesh
 * u (context precision), m_c the CIC-deposited mass of cell c. */
int hy_pm_kdk_vjp(hy_pm_ctx* ctx, uint64_t n, const double* x, const double* y, const double* z,
                  double uniform_mass, double* lx, double* ly, double* lz, double* lpx, double* lpy,
                  double* lpz, double k_prev, double k_next, double d_next, uint32_t flags);
int hy_pm_density_vjp(hy_pm_ctx* ctx, uint64_t n, const double* x, const double* y, const double* z,
                      double uniform_mass, const void* u_mesh, double coef, double* lx, double* ly,
                      double* lz);

/* Ensemble step in one HIP graph launch.  No reference counterpart (the
 * reference advances one box per process): several contexts, usually one per
 * box geometry with many boxes each, that share one stream are advanced
 * together.  Each member's state is updated in place (x, p).  The first
 * hy_ensemble_step for a given (HY_HAS_PREV, HY_HAS_NEXT) combination records
 * the members' fused steps into a HIP graph, with the kick/drift factors read
 * from device memory, instantiates it and launches it; later steps only write
 * the three factors and relaunch the graph.  A run starts with one eager
 * hy_pm_kdk per member with HY_HAS_NEXT only (force at the initial positions,
 * K1, drift) before the first graph step.  A step is recorded only when every
 * member holds what a recorded step leaves for its successor (the deposit
 * records of binary32 members, the bucket keys of binary64 ones, made for the
 * members' current positions); until then, and for steps without HY_HAS_PREV,
 * the step runs eagerly.  The recorded steps assume the members' positions
 * change only through the ensemble: after positions are written by anything
 * else (new initial conditions in the same buffers), call hy_ensemble_reset,
 * whose next hy_ensemble_step runs eagerly and routes the particles afresh (it
 * also discards errors not yet reported); without it binary64 members' reused
 * keys are stale (hy_ensemble_check raises logic_error) and binary32 members
 * deposit stale records.  Calls do not synchronise;
 * hy_ensemble_check checks every member, clears their error words and returns
 * the first kernel-raised error. */
typedef struct hy_ensemble hy_ensemble;
typedef struct {
    hy_pm_ctx* ctx;
    uint64_t n;                 /* particles over all boxes of the context */
    double *x, *y, *z;          /* device positions (updated in place) */
    void *px, *py, *pz;         /* device momenta in the context precision (in place) */
    double uniform_mass;
} hy_ensemble_member;
int hy_ensemble_create(const hy_ensemble_member* members, uint32_t count, hy_ensemble** out);
int hy_ensemble_step(hy_ensemble* e, double k_prev, double k_next, double d_next, uint32_t flags);
int hy_ensemble_reset(hy_ensemble* e);
int hy_ensemble_check(hy_ensemble* e);
int hy_ensemble_destroy(hy_ensemble* e);

/* Mode-keyed Gaussian initial field (src/ic/random_field.cpp:70-330 with
 * src/ic/philox_rng.cpp:1-90), on the GPU.  Writes the r2c half spectrum
 * out[(i*N + j)*(N/2+1) + k] (binary64 complex, device memory) of an N^3 IC
 * mesh.  Each mode draws Philox4x32-10 with key (seed lo, seed hi) and counter
 * (m_x, m_y, m_z, stream) of its canonical mode: on the k = 0 and Nyquist
 * planes the mode whose conjugate (i, j) is lexicographically smaller takes the
 * conjugate draw; self-conjugate modes are real with standard deviation
 * sqrt(var), others have sqrt(var/2) per part; the zero mode and modes with
 * any |m_axis| > max_mode are zero.  var = variance_by_m2[m_x^2+m_y^2+m_z^2]
 * (device, 3 (N/2)^2 + 1 entries) is the caller's N^6/L^3 P(|m| 2 pi/L).
 * flags: HY_IC_FIXED_AMPLITUDE (|X|^2 = var, phase from stream 1),
 * HY_IC_REVERSED_PAIR (negated field).  Synchronises the stream. */
#define HY_IC_FIXED_AMPLITUDE 0x1u
#define HY_IC_REVERSED_PAIR   0x2u
int hy_ic_gaussian_modes(uint64_t N, uint64_t seed, uint64_t max_mode, const double* variance_by_m2,
                         void* out, uint32_t flags, int device, void* stream);
const char* hy_ic_last_error(void);

/* 1LPT lattice ICs (src/ic/lpt_displacement.cpp:202-384 first order, lattice
 * src/ic/initial_conditions.cpp:296-312) on a binary64 context whose mesh is
 * the IC mesh: per axis psi = c2r(i k_axis / k^2 delta_k) (k_axis = 0 on the
 * Nyquist index, 1/N^3), then for particle (i, j, k) of the n_side^3 lattice
 * (n_side divides N) x = wrap(q + psi(i s, j s, k s)), p = p_factor * psi with
 * q = index * L / n_side and p_factor = a^2 H(a) f(a).  delta_k: device
 * complex128 half spectrum [N][N][N/2+1] (hy_ic_gaussian_modes' layout);
 * outputs: device binary64, n_side^3 each.  Synchronous. */
int hy_ic_lpt(hy_pm_ctx* ctx, uint64_t n_side, const void* delta_k, double p_factor, double* x, double* y,
              double* z, double* px, double* py, double* pz);

/* Matter power spectrum on the context's mesh (N = the estimator's mesh_size,
 * L = the periodic domain's box size). */
#define HY_PK_MAX_BINS 1024

/* include/cosmo_nbody/analysis/matter_power_spectrum.hpp:30-38 */
typedef struct {
    int num_bins;                  /* 1..HY_PK_MAX_BINS logarithmic shells in [2 pi/L, k_max] */
    int subtract_shot_noise;       /* 0 or 1 (core::ExplicitBoolean: any other value is "not
                                      specified", a logic_error) */
    int interlaced;                /* 0 or 1, likewise */
    double max_k_fraction_nyquist; /* (0, 1]: k_max = fraction * pi N / L */
} hy_pk_options;

/* include/cosmo_nbody/analysis/matter_power_spectrum.hpp:15-28 */
typedef struct {
    double k_low;
    double k_mean;
    double k_high;
    double p_raw;
    double p_corrected;
    double mean_cic_power_deconvolution_factor;
    double shot_noise_reference;
    uint64_t mode_count;
} hy_pk_bin;

/* Replaces analysis::MatterPowerSpectrum::compute_pk
 * (include/cosmo_nbody/analysis/matter_power_spectrum.hpp:39-46;
 * src/analysis/matter_power_spectrum.cpp:57-241) over
 * FourierDensityBuilder::build (src/analysis/fourier_density_field.cpp:112-343):
 * CIC deposit and r2c transform on the context's mesh and workspace, then the
 * shell sums on the device.  masses == NULL selects the uniform mass.  bins
 * (host memory) receives opts->num_bins shells.  Single-box contexts only.
 * An interlaced estimate allocates a partner spectrum and staging for the
 * shifted positions on first use.  Synchronous. */
int hy_pm_power_spectrum(hy_pm_ctx* ctx, uint64_t n, const double* x, const double* y, const double* z,
                         const double* masses, double uniform_mass, const hy_pk_options* opts,
                         hy_pk_bin* bins, uint32_t flags);

/* Host scalars of the step (binary64, long-double quadrature).
 * H(a): src/cosmology/cosmology_model.cpp:25-57,167-178.
 * D = int da/(a^3 H), K = int da/(a^2 H):
 *   src/cosmology/drift_kick_factors.cpp:28-141 with
 *   include/cosmo_nbody/math/roundoff_resolved_gauss_legendre_16.hpp.
 * Geometric mean: include/cosmo_nbody/math/exact_geometric_mean.hpp:168-275.
 * Boundaries: src/time/time_stepper.cpp:194-276 (stops, exact final a). */
double hy_hubble(double a, double omega_m, double omega_lambda);
int hy_drift_factor(double a1, double a2, double omega_m, double omega_lambda, double* out);
int hy_kick_factor(double a1, double a2, double omega_m, double omega_lambda, double* out);
double hy_geometric_mean(double lhs, double rhs);
int hy_time_boundaries(double a_start, double a_final, double delta_ln_a, const double* stops,
                       uint64_t n_stops, double* out, uint64_t capacity, uint64_t* count);
/* Correctly rounded exact sum of non-negative binary64 values
 * (contract of include/cosmo_nbody/math/exact_positive_sum.hpp). */
double hy_exact_sum(const double* values, uint64_t count);
/* Initial-condition scalars (host, binary64).  Cosmology parameters as
 * config::CosmologyParams (include/cosmo_nbody/config/simulation_parameters.hpp:24-31);
 * the model is flat matter + Lambda. */
typedef struct {
    double h, omega_m, omega_b, omega_lambda, sigma8, n_s;
} hy_cosmology;
/* Linear growth D1, f1 = dlnD1/dlna, D2, f2 at scale factors a[0..n) (any output
 * may be NULL): CosmologyModel::D1/f1/D2/f2 (include/cosmo_nbody/cosmology/
 * cosmology_model.hpp:46-53; src/cosmology/cosmology_model.cpp:70-509): RK4 in
 * ln a from the matter era, cubic Hermite interpolation in a. */
int hy_growth(const hy_cosmology* c, uint64_t n, const double* a, double* D1, double* f1, double* D2, double* f2);
/* LinearPowerSpectrum::evaluate (include/cosmo_nbody/ic/linear_power_spectrum.hpp:19-28;
 * src/ic/linear_power_spectrum.cpp:414-664) for a table given as text: "# key=value"
 * semantic header checked against c and table_redshift, rows "k P(k)" in h/Mpc and
 * (Mpc/h)^3, growth-rescaled to start_redshift by (D1(a_start)/D1(a_table))^2, ln P
 * interpolated by a natural cubic spline in ln k.  At least 3 rows. */
int hy_linear_power(const hy_cosmology* c, const char* table, double table_redshift, double start_redshift,
                    uint64_t n, const double* k, double* out);
/* Per-mode variance of the Gaussian IC field, N^6/L^3 P(|m| 2 pi/L) as
 * src/ic/random_field.cpp:108-235 forms it (scaled mantissa/exponent product),
 * indexed by m^2 = mx^2+my^2+mz^2 for m^2 <= 3 max_mode^2 (0 beyond, and at 0):
 * the variance_by_m2 input of hy_ic_gaussian_modes, 3 (N/2)^2 + 1 entries. */
int hy_ic_mode_variance(const hy_cosmology* c, const char* table, double table_redshift, double start_redshift,
                        uint64_t mesh_size, double box_size, uint64_t max_mode, double* variance_by_m2);
/* math::CubicSpline (include/cosmo_nbody/math/cubic_spline.hpp): natural spline
 * through (x, y) (n >= 3, x increasing), evaluated at q[0..m). */
int hy_cubic_spline(uint64_t n, const double* x, const double* y, uint64_t m, const double* q, double* out);
/* Native HYOWON snapshot files (HDF5, the reference's single-file schema:
 * include/cosmo_nbody/io/output_schema.hpp:22-140).  Host memory, binary64.
 * The HDF5 C library is loaded at run time (libhdf5.so, or the path in the
 * environment variable HY_HDF5_LIBRARY); without it these calls return
 * HY_ERUNTIME.  Errors: hy_snapshot_last_error(). */
typedef struct {
    double box_size;               /* Header.BoxSize (Mpc/h) */
    double omega_m, omega_lambda, h, omega_b, sigma8, n_s;
    uint64_t seed;                 /* Parameters.Seed */
    uint32_t particles_per_dim, pm_mesh_per_dim, ic_mesh_per_dim;
    int32_t lpt_order;
    double softening;              /* Parameters.SofteningComoving (0 for pure PM) */
    double z_start, z_final, delta_ln_a, power_spectrum_redshift;
    int deconvolve_cic;
    const char* ic_amplitude_mode;       /* "gaussian" when NULL */
    const char* ic_phase_pairing;        /* "independent" when NULL */
    const char* power_spectrum_fidelity; /* "precision_boltzmann" when NULL */
} hy_snapshot_meta;
typedef struct {
    double a, redshift, box_size, omega_m, omega_lambda, h;
    double uniform_mass;           /* MassTable[1]; 0 when the file holds explicit masses */
    uint64_t particles;            /* NumPart_ThisFile[1] */
    int explicit_masses;           /* PartType1/Masses present */
} hy_snapshot_header;
#define HY_SNAPSHOT_REPLACE 0x1u   /* hy_snapshot_write: replace an existing file (else it must not exist) */
/* SnapshotIO::write_snapshot_to_path (include/cosmo_nbody/io/snapshot_io.hpp:97-110;
 * src/io/snapshot_io_write_impl.inc): /Header, /Parameters, /Config (identity,
 * physics text, run metadata, semantic bindings), /PartType1 Coordinates (n x 3),
 * Velocities = p / a (n x 3, km/s), ParticleIDs (ids or 0..n-1), Masses when
 * masses != NULL (else MassTable[1] = uniform_mass).  Positions must lie in
 * [0, L). */
int hy_snapshot_write(const char* path, const hy_snapshot_meta* meta, double a, uint64_t n, const double* x,
                      const double* y, const double* z, const double* px, const double* py, const double* pz,
                      const uint64_t* ids, const double* masses, double uniform_mass, uint32_t flags);
/* The header checks of SnapshotIO::read_snapshot (src/io/snapshot_io_read_impl.inc:33-190):
 * Redshift consistent with Time, one file, PartType1 only, native units, the
 * Config semantic bindings, non-empty provenance. */
int hy_snapshot_read_header(const char* path, hy_snapshot_header* out);
/* Reads the particles (capacity >= header.particles) with p = a v; expect
 * (optional) is the admitted box and cosmology, compared as the reference does.
 * ids and masses may be NULL (masses is required when the file holds them). */
int hy_snapshot_read(const char* path, const hy_snapshot_header* expect, uint64_t capacity, double* x, double* y,
                     double* z, double* px, double* py, double* pz, uint64_t* ids, double* masses);
const char* hy_snapshot_last_error(void);
/* G in (Mpc/h)(km/s)^2/(1e10 Msun/h) and rho_crit0 (include/cosmo_nbody/cosmology/units.hpp:26-30). */
double hy_gravitational_constant(void);
double hy_rho_crit0(void);

#ifdef __cplusplus
}
#endif
#endif
