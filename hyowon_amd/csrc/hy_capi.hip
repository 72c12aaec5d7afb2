// C ABI of libhyowon_pm.so: context, workspace and launch orchestration of the
// MI355X particle-mesh step.  Every entry point is extern "C" with plain
// pointers; HIP failures and kernel-raised error bits are mapped onto the
// reference's exception classes (see include/hyowon_pm.h).
#include "hyowon_pm.h"

#include "hy_host_scalars.hpp"
#include "hy_fft_table.hpp"
#include "hy_kernels.hpp"
#include "hy_pk.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <memory>
#include <numeric>
#include <limits>
#include <numbers>
#include <string>
#include <type_traits>
#include <vector>

namespace {

thread_local std::string g_last_error;

struct HyError {
    int code;
    std::string what;
};

[[noreturn]] void fail(int code, const std::string& what) { throw HyError{code, what}; }

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(HY_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename F>
int guarded(F&& f) {
    try {
        f();
        g_last_error.clear();
        return HY_OK;
    } catch (const HyError& e) {
        g_last_error = e.what;
        return e.code;
    } catch (const std::invalid_argument& e) {
        g_last_error = e.what();
        return HY_EINVAL;
    } catch (const std::underflow_error& e) {
        g_last_error = e.what();
        return HY_EUNDERFLOW;
    } catch (const std::overflow_error& e) {
        g_last_error = e.what();
        return HY_EOVERFLOW;
    } catch (const std::logic_error& e) {
        g_last_error = e.what();
        return HY_ELOGIC;
    } catch (const std::bad_alloc&) {
        g_last_error = "allocation failed";
        return HY_ENOMEM;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return HY_ERUNTIME;
    }
}

int tile_edge_for(int N) {
    for (int t = 16; t >= 1; --t)
        if (N % t == 0) return t;
    return 1;
}

uint64_t blocks_for(uint64_t n, uint64_t per) { return (n + per - 1) / per; }

}  // namespace

struct hy_pm_ctx {
    hy_pm_desc desc{};
    hy::Geom geom{};
    int N = 0;
    int prec = HY_PREC_FP64;
    size_t real_bytes = 1;  // sizeof(T)
    hipStream_t stream = nullptr;
    bool own_stream = false;
    uint64_t cap = 0;
    uint64_t ntiles = 0, nb = 0, nb_scan_blocks = 0;
    uint64_t device_bytes = 0;
    double green_num = 0.0;
    double last_input_mass = 0.0, last_deposited_mass = 0.0;
    bool energy_valid = false;       // the last hy_pm_force collected the potential energy
    double last_potential_energy = 0.0;
    double* energy_partials = nullptr;  // potential-energy workgroup sums (allocated on first use)
    uint64_t energy_cap = 0;            // workgroups the partials hold
    double mass_unit = 1.0;  // fixed-point deposit unit (uniform mass or max explicit mass)
    int plane_group = 0;     // kz planes per y/x/y^-1 launch group (0 = all planes at once)
    uint32_t gather_rows = HY_GATHER_ROWS_DEFAULT;  // lattice rows per gather column tile (0: particle order)
    // Bucket keys/ranks/counts computed by the previous fused step for its output
    // positions; reused when the next step's input is that output.
    bool keys_ready = false;
    const double* keys_x = nullptr;
    const double* keys_y = nullptr;
    const double* keys_z = nullptr;
    uint64_t keys_n = 0;
    hy::DepRecord* rec = nullptr;       // binary32 contexts: deposit records by bucket
    float* rec_mass = nullptr;
    // Binary32 uniform-mass HY_ASYNC steps hand the next step the deposit
    // records of their drifted positions in a capacity layout (see
    // k_cap_reduce); rec_ready marks them valid for exactly those positions.
    bool rec_layout = false;            // the buffers below exist (binary32 contexts)
    bool rec_ready = false;
    const double* rec_x = nullptr;
    const double* rec_y = nullptr;
    const double* rec_z = nullptr;
    uint64_t rec_n = 0;
    uint32_t* cap_cur = nullptr;        // nball+1: layout the carried records were written in
    uint32_t* cap_next = nullptr;       // nball+1: layout the next step's records go to
    uint32_t* ovf_start = nullptr;      // nball+1: overflow prefix of the carried layout
    uint32_t* ovf_count = nullptr;      // overflow records the last step appended
    hy::DepRecord* ovf_rec = nullptr;   // overflow records (aliases the binary64 fallback scratch)
    uint32_t* cap_block_sums = nullptr; // 3 per scan block
    hy::DepRecord64* rec64 = nullptr;   // binary64 contexts
    double* rec_mass64 = nullptr;

    // mesh
    void* R = nullptr;
    void* A = nullptr;
    void* B = nullptr;
    void* tw = nullptr;
    void* lap = nullptr;
    void* win = nullptr;
    // particles
    uint32_t* keys = nullptr;
    uint32_t* ranks = nullptr;
    uint32_t* order = nullptr;
    uint32_t* counts = nullptr;
    uint32_t* offsets = nullptr;
    uint32_t* block_sums = nullptr;
    uint32_t* grand_total = nullptr;
    uint32_t* xcounts = nullptr;   // x-plane bins (N)
    uint32_t* xoffsets = nullptr;  // N+1
    unsigned long long* mass_digits = nullptr;  // per box: exact deposited-mass digit counters
    unsigned long long* mass_waves = nullptr;   // per deposit wave: its exact mass (binary64: digit window,
                                                // binary32: two digit sums of its 32.32 cells)
    double* mass_total = nullptr;
    uint64_t* scratch = nullptr;
    uint64_t scratch_cap = 0;
    unsigned long long* scratch_cursor = nullptr;
    uint32_t* err = nullptr;
    void* gx = nullptr;
    void* gy = nullptr;
    void* gz = nullptr;
    double* hpos = nullptr;  // device staging for host-pointer calls (3*cap doubles + masses)
    double* hacc = nullptr;  // 3*cap

    hy::FftKernels<float> fft32{};
    hy::FftKernels<double> fft64{};
    int fft_var[hy::FFT_PASSES] = {0, 0, 0, 0, 0};  // launch variant per pass
    int max_lds = 0;                                 // device LDS bytes per workgroup
    uint32_t nbox = 1;                               // boxes of this geometry (box-major arrays)
    uint64_t nball = 0;                              // deposit buckets over all boxes
    const double* dfac = nullptr;                    // device step factors (graph capture)
    double* hmass = nullptr;                         // pinned per-box deposited masses
    void* adj_phi = nullptr;                         // adjoint workspace (allocated on first use)
    void* adj_a[3] = {nullptr, nullptr, nullptr};
    double* adj_max = nullptr;
    int method = HY_METHOD_PURE_PM;
    double split_scale = 0.0;
    void* B2 = nullptr;          // TreePM long range: x-derivative spectrum
    void* R2 = nullptr;          // TreePM long range: y and z force meshes (R holds x)
    void* R3 = nullptr;
    void* kc = nullptr;          // k component per index
    void* kd = nullptr;          // k component per index, 0 on the Nyquist index
    void* ic_kc = nullptr;       // binary64 k tables of the 1LPT displacements (allocated on first use)
    void* ic_kd = nullptr;
    void* pk_aux = nullptr;      // interlaced partner spectrum (allocated on first use)
    double* pk_tab = nullptr;    // k components, windows, alias factors (3N), shell edges
    double* pk_part = nullptr;   // workgroup shell partials
    uint64_t pk_part_cap = 0;    // doubles held by pk_part

    // pinned host mirror of the status words
    struct Status {
        uint32_t err;
        uint32_t total;
        double mass;
    };
    Status* hstatus = nullptr;
    Status* dstatus = nullptr;

    std::vector<void*> allocations;

    // Live per-stage timing: kHyMarks events per call, recorded on `stream`.
    static constexpr int kHyMarks = HY_STAGE_COUNT + 1;
    bool prof = false;
    uint64_t prof_cap = 0, prof_calls = 0;
    std::vector<hipEvent_t> prof_ev;
    std::vector<uint8_t> prof_used;

    void* alloc(size_t bytes) {
        void* p = nullptr;
        hip_check(hipMalloc(&p, bytes ? bytes : 16), "hipMalloc");
        allocations.push_back(p);
        device_bytes += bytes;
        return p;
    }
    ~hy_pm_ctx() {
        for (hipEvent_t e : prof_ev) (void)hipEventDestroy(e);
        for (void* p : allocations) (void)hipFree(p);
        if (hstatus) (void)hipHostFree(hstatus);
        if (hmass) (void)hipHostFree(hmass);
        if (own_stream && stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

void set_device(const hy_pm_ctx* c) { hip_check(hipSetDevice(c->desc.device), "hipSetDevice"); }

void launch_check(const char* what) { hip_check(hipGetLastError(), what); }

// Clears `words` 32-bit words with a kernel (no memset nodes: callers may
// capture these paths into a HIP graph).
void fill_u32(hy_pm_ctx* c, void* p, uint64_t words) {
    const uint64_t blocks = std::min<uint64_t>(std::max<uint64_t>(blocks_for(words, 256), 1), 4096);
    hipLaunchKernelGGL(hy::k_fill_u32, dim3((uint32_t)blocks), dim3(256), 0, c->stream, static_cast<uint32_t*>(p),
                       words, 0u);
    launch_check("k_fill_u32");
}

// Particles per box of a call (particle arrays concatenate the boxes).
uint64_t set_boxes(hy_pm_ctx* c, uint64_t n) {
    if (n % c->nbox != 0) fail(HY_EINVAL, "particle count must be a multiple of the context's box count");
    const uint64_t nper = n / c->nbox;
    c->geom.nper = (uint32_t)nper;
    return nper;
}

void require_single_box(const hy_pm_ctx* c, const char* what) {
    if (c->nbox != 1) fail(HY_EINVAL, std::string(what) + " takes a single-box context");
}

// Record stage boundary `mark` (0 = call start, k = end of stage k-1) of the
// current profiled call.  No-op when profiling is off or the ring is full.
void mark(hy_pm_ctx* c, int m) {
    if (!c->prof || c->prof_calls >= c->prof_cap) return;
    const uint64_t slot = c->prof_calls * hy_pm_ctx::kHyMarks + (uint64_t)m;
    hip_check(hipEventRecord(c->prof_ev[slot], c->stream), "hipEventRecord");
    c->prof_used[slot] = 1;
}

void end_call(hy_pm_ctx* c) {
    if (c->prof && c->prof_calls < c->prof_cap) ++c->prof_calls;
}

const char* const kMsgPositions = "CIC positions must be finite";

// Map the kernel error word onto the reference's exceptions.
void raise_errors(uint32_t e) {
    using namespace hy;
    if (e & ERR_POSITION_NONFINITE) fail(HY_EINVAL, kMsgPositions);
    if (e & ERR_MASS_INVALID) fail(HY_EINVAL, "CIC masses must be finite and non-negative");
    if (e & ERR_STALE_KEYS) fail(HY_ELOGIC, "reused bucket keys do not match the input positions");
    if (e & ERR_COORDINATE) fail(HY_ELOGIC, "CIC bounded coordinate escaped the mesh domain");
    if (e & ERR_SCRATCH) fail(HY_ERUNTIME, "deposit fallback scratch exhausted");
    if (e & ERR_FIELD_NONFINITE) fail(HY_EINVAL, "CIC centered-gradient potential must be finite");
    if (e & ERR_FORCE_NONFINITE) fail(HY_EOVERFLOW, "CIC centered-gradient gathered acceleration is non-finite");
    if (e & ERR_MOMENTUM_NONFINITE) fail(HY_EOVERFLOW, "Leapfrog kick input is not finite");
    if (e & ERR_UPDATE_NONFINITE) fail(HY_EOVERFLOW, "Leapfrog K1/drift update is not finite and representable");
    if (e & ERR_SPECTRUM_NONFINITE) fail(HY_ERUNTIME, "FourierDensityBuilder Fourier mode is non-finite");
    if (e & ERR_INTERP_UNDERFLOW)
        fail(HY_EUNDERFLOW, "CIC interpolation lost a non-zero CIC weight or field-weight product in core::Real");
    if (e & ERR_GATHER_UNDERFLOW)
        fail(HY_EUNDERFLOW, "CIC centered-gradient gather lost a non-zero CIC weight, gradient, or weighted "
                            "contribution in core::Real");
    if (e & ERR_ROUTING) fail(HY_ELOGIC, "CIC bucket routing does not cover each particle exactly once");
    if (e & ERR_MASS_CONSERVATION) fail(HY_ERUNTIME, "CIC deposited mass violates the forward-error bound");
}

// Input mass identity (src/gravity/pm_solver.cpp:120-130,665-684).
double input_mass(hy_pm_ctx* c, uint64_t n, const double* masses, double uniform, bool host_ptrs) {
    if (n == 0) fail(HY_ERUNTIME, "PMSolver exact input mass identity must be finite and positive");
    double M;
    c->mass_unit = uniform;
    if (!masses) {
        if (!std::isfinite(uniform) || uniform <= 0.0)
            fail(HY_EINVAL, "PMSolver uniform mass must be finite and positive");
        // exact m*n rounded once: one binary64 product is correctly rounded
        // because n (< 2^53) is exact.
        M = uniform * static_cast<double>(n);
    } else {
        std::vector<double> h(n);
        if (host_ptrs) std::memcpy(h.data(), masses, n * sizeof(double));
        else hip_check(hipMemcpyAsync(h.data(), masses, n * sizeof(double), hipMemcpyDeviceToHost, c->stream), "mass copy");
        hip_check(hipStreamSynchronize(c->stream), "sync");
        M = hy::host::exact_sum(h.data(), n);
        double mx = 0.0;
        for (double v : h) mx = v > mx ? v : mx;
        c->mass_unit = mx > 0.0 ? mx : 1.0;
        if (!std::isfinite(M)) fail(HY_EINVAL, "PMSolver explicit masses must be finite and non-negative");
    }
    if (!std::isfinite(M) || M <= 0.0)
        fail(HY_ERUNTIME, "PMSolver exact input mass identity must be finite and positive");
    return M;
}

// Three-channel scan of the deposit bucket counts (k_cap_reduce): dense
// offsets (or null), the overflow prefix of the carried layout cur (or null),
// the next step's capacity layout capn (or null), and the grand total.
void cap_scan(hy_pm_ctx* c, const uint32_t* cur, uint32_t* dense, uint32_t* capn) {
    hy::CapScan s{};
    s.counts = c->counts;
    s.nb = c->nball;
    s.cur = cur;
    s.dense = dense;
    s.ovf_start = c->ovf_start;
    s.capn = capn;
    s.block_sums = c->cap_block_sums;
    s.grand_total = c->grand_total;
    const uint64_t nblk = blocks_for(c->nball, hy::SCAN_TILE);
    hipLaunchKernelGGL(hy::k_cap_reduce, dim3(nblk), dim3(hy::SCAN_THREADS), 0, c->stream, s);
    hipLaunchKernelGGL(hy::k_cap_blocks, dim3(1), dim3(hy::SCAN_THREADS), 0, c->stream, s, (uint32_t)nblk);
    hipLaunchKernelGGL(hy::k_cap_apply, dim3(nblk), dim3(hy::SCAN_THREADS), 0, c->stream, s);
    launch_check("k_cap_scan");
}

// Bucket count/rank + scan + scatter.  mode 0: deposit buckets, 1: x-planes.
// carried: keys, ranks and counts were produced by the previous fused step.
// capn (binary32 deposit buckets): also lay out the capacities of the next
// step's carried records from these counts.
void bucketize(hy_pm_ctx* c, uint64_t n, const double* x, const double* y, const double* z, int mode,
               uint32_t* counts, uint32_t* offsets, uint64_t nb, uint32_t* order, bool carried = false,
               const double* masses = nullptr, uint32_t* capn = nullptr, double shift = 0.0) {
    if (!carried) {
        fill_u32(c, counts, nb);
        hipLaunchKernelGGL(hy::k_bucket_key, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, x, y, z,
                           c->geom, mode, c->keys, c->ranks, counts, c->err, shift);
        launch_check("k_bucket_key");
    }
    if (mode == 0 && c->prec == HY_PREC_FP32_MIXED && c->rec_layout) {
        cap_scan(c, nullptr, offsets, capn);
    } else {
        const uint64_t nblk = blocks_for(nb, hy::SCAN_TILE);
        hipLaunchKernelGGL(hy::k_scan_reduce, dim3(nblk), dim3(hy::SCAN_THREADS), 0, c->stream, counts, nb,
                           c->block_sums);
        hipLaunchKernelGGL(hy::k_scan_blocks, dim3(1), dim3(hy::SCAN_THREADS), 0, c->stream, c->block_sums,
                           (uint32_t)nblk, c->grand_total);
        hipLaunchKernelGGL(hy::k_scan_apply, dim3(nblk), dim3(hy::SCAN_THREADS), 0, c->stream, counts, nb,
                           c->block_sums, c->grand_total, offsets);
        launch_check("k_scan");
    }
    if (mode == 0 && c->prec == HY_PREC_FP32_MIXED) {
        hipLaunchKernelGGL((hy::k_scatter_records<hy::DepRecord, float>), dim3(blocks_for(n, 256)), dim3(256), 0,
                           c->stream, n, x, y, z, masses, 1.0 / c->mass_unit, c->geom, c->keys, c->ranks, offsets,
                           order, c->rec, c->rec_mass, c->err, shift);
        launch_check("k_scatter_records");
    } else if (mode == 0) {
        hipLaunchKernelGGL((hy::k_scatter_records<hy::DepRecord64, double>), dim3(blocks_for(n, 256)), dim3(256), 0,
                           c->stream, n, x, y, z, masses, 1.0, c->geom, c->keys, c->ranks, offsets, order,
                           c->rec64, c->rec_mass64, c->err, shift);
        launch_check("k_scatter_records");
    } else if (carried) {
        hipLaunchKernelGGL(hy::k_scatter_checked, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, x, y,
                           z, c->geom, c->keys, c->ranks, offsets, order, c->err);
        launch_check("k_scatter_checked");
    } else {
        hipLaunchKernelGGL(hy::k_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, c->keys,
                           c->ranks, offsets, order);
        launch_check("k_scatter");
    }
}

void reset_status(hy_pm_ctx* c) {
    fill_u32(c, c->err, 1);
    fill_u32(c, c->scratch_cursor, 2);
}

// Copy err / grand total / deposited mass to the host and synchronise.
hy_pm_ctx::Status fetch_status(hy_pm_ctx* c) {
    hip_check(hipMemcpyAsync(&c->hstatus->err, c->err, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream), "status");
    hip_check(hipMemcpyAsync(&c->hstatus->total, c->grand_total, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream), "status");
    hip_check(hipMemcpyAsync(&c->hstatus->mass, c->mass_total, sizeof(double), hipMemcpyDeviceToHost, c->stream), "status");
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return *c->hstatus;
}

struct DepositPlan {
    double M;
    hy::DepParams P;
};

DepositPlan plan_deposit(hy_pm_ctx* c, uint64_t n, const double* x, const double* y, const double* z,
                         const double* masses, double uniform, double Min, double shift = 0.0) {
    const double dx = c->geom.dx;
    const double cells = static_cast<double>(static_cast<uint64_t>(c->N) * c->N * c->N);
    const double mean = Min / cells;
    const double inv_volume = 1.0 / ((dx * dx) * dx);
    const bool exact_mean = std::isfinite(mean) && mean > 0.0 && std::fma(mean, cells, -Min) == 0.0;
    hy::DepParams P{};
    P.g = c->geom;
    P.x = x; P.y = y; P.z = z;
    P.masses = masses;
    P.uniform_mass = uniform;
    P.pos_shift = shift;
    P.offsets = c->offsets;
    P.order = c->order;
    P.rec64 = c->rec64;
    P.rec_mass64 = c->rec_mass64;
    P.n = n;
    P.mean_mass = mean;
    P.inv_volume = inv_volume;
    P.total_mass = Min;
    P.cells_real = cells;
    P.centred_scale = inv_volume * (1.0 / cells);
    {
        // 1/(N_c dx^3) as the reference forms it: (((1/N_c)/dx)/dx)/dx in long double
        long double S = 1.0L;
        S /= (long double)cells;
        S /= (long double)dx;
        S /= (long double)dx;
        S /= (long double)dx;
        int e = 0;
        const long double f = std::frexp(S, &e);
        P.norm_mant = (uint64_t)std::ldexp(f, 64);
        P.norm_exp = e - 64;
    }
    P.direct = (exact_mean && std::isfinite(inv_volume) && inv_volume > 0.0) ? 1 : 0;
    P.mass_digits = c->mass_digits;
    P.mass_waves = c->mass_waves;
    P.scratch = c->scratch;
    P.scratch_cap = c->scratch_cap;
    P.scratch_cursor = c->scratch_cursor;
    P.err = c->err;
    (void)n;
    return {Min, P};
}

// Forward-error factor of the mass-conservation bound for n particles per box
// (src/math/cic_mass_conservation.cpp:25-63): gamma(32 n + 1) + 64 eps, plus
// 1e-4 for the binary32 mesh.
long double mass_tol_factor(const hy_pm_ctx* c, uint64_t n) {
    const long double eps = std::numeric_limits<double>::epsilon();
    const long double nu = (32.0L * static_cast<long double>(n) + 1.0L) * eps;
    if (!(nu < 0.25L)) fail(HY_EOVERFLOW, "CIC mass conservation forward-error bound is not informative at this particle count");
    const long double gamma = nu / (1.0L - nu);
    long double f = gamma + 64.0L * eps;
    if (c->prec == HY_PREC_FP32_MIXED) f += 1e-4L;
    return f;
}

// status_check: also evaluate the routing identity and the mass bound on the
// device (error bits reported by the synchronous check or hy_pm_check).
// Binary32: the records lie in the layout start (dense offsets unless
// carried in a capacity layout, whose overflow prefix is ovf).
void launch_deposit(hy_pm_ctx* c, const hy::DepParams& P, bool status_check = true,
                    const uint32_t* start = nullptr, const uint32_t* ovf = nullptr) {
    fill_u32(c, c->mass_digits, (uint64_t)hy::SA_WORDS * c->nbox * 2);
    if (c->prec == HY_PREC_FP64)
        hipLaunchKernelGGL((c->geom.T == 16 ? hy::k_deposit<double, 16> : hy::k_deposit<double, 0>),
                           dim3(c->ntiles * c->nbox), dim3(hy::DEP_THREADS), 0, c->stream, P,
                           static_cast<double*>(c->R));
    else {
        hy::DepFxParams Q{P, c->mass_unit, std::ldexp(c->mass_unit, -32), c->rec,
                          P.masses ? c->rec_mass : nullptr, start ? start : c->offsets, c->counts, ovf,
                          c->nball};
        hipLaunchKernelGGL(c->geom.T == 16 ? hy::k_deposit_fixed<16> : hy::k_deposit_fixed<0>,
                           dim3(c->ntiles * c->nbox), dim3(hy::DEPFX_THREADS), 0, c->stream, Q,
                           static_cast<float*>(c->R));
    }
    launch_check("k_deposit");
    // the exact sum of every cell's deposited mass (ExactPositiveDoubleSum), per box
    const uint64_t waves = c->ntiles * hy::SA_WAVES_PER_TILE;
    const dim3 rgrid((uint32_t)std::min<uint64_t>(256, blocks_for(waves, 256)), c->nbox);
    if (c->prec == HY_PREC_FP64)
        hipLaunchKernelGGL(hy::k_mass_reduce, rgrid, dim3(256), 0, c->stream, c->mass_waves, waves, c->mass_digits);
    else
        hipLaunchKernelGGL(hy::k_mass_reduce_fixed, rgrid, dim3(256), 0, c->stream, c->mass_waves, waves,
                           c->mass_unit, c->mass_digits);
    hipLaunchKernelGGL(hy::k_mass_exact, dim3((uint32_t)blocks_for(c->nbox, 64)), dim3(64), 0, c->stream,
                       c->mass_digits, c->nbox, c->mass_total);
    launch_check("k_mass_exact");
    if (status_check) {
        const double tol = static_cast<double>(mass_tol_factor(c, P.n / c->nbox));
        hipLaunchKernelGGL(hy::k_deposit_status, dim3((uint32_t)blocks_for(c->nbox, 256)), dim3(256), 0, c->stream,
                           c->grand_total, P.n, c->mass_total, c->nbox, P.total_mass, tol, c->err);
        launch_check("k_deposit_status");
    }
}

template <typename T>
hy::FftArgs<T> fft_args(hy_pm_ctx* c) {
    hy::FftArgs<T> a{};
    a.R = static_cast<T*>(c->R);
    a.A = static_cast<hy::cplx<T>*>(c->A);
    a.B = static_cast<hy::cplx<T>*>(c->B);
    a.tw = static_cast<const hy::cplx<T>*>(c->tw);
    a.green_lap = static_cast<const T*>(c->lap);
    a.green_win = static_cast<const T*>(c->win);
    a.green_num = c->green_num;
    a.norm = 1.0 / (static_cast<double>(c->N) * static_cast<double>(c->N) * static_cast<double>(c->N));
    a.deconvolve = c->desc.deconvolve_cic ? 1 : 0;
    a.box_real = (uint64_t)c->N * c->N * c->N;
    a.box_cplx = (uint64_t)c->N * (c->N / 2 + 1) * c->N;
    a.err = c->err;
    return a;
}

// One launch of pass p (variant v) over kz planes [kz0, kz0+g) (y and x
// passes) or over the whole mesh (z passes).
template <typename T>
void launch_pass(hy_pm_ctx* c, const hy::FftKernels<T>& K, hy::FftArgs<T>& a, int p, int v, int kz0, int g) {
    static const char* names[hy::FFT_PASSES] = {"k_fft_z_r2c", "k_fft_y_fwd", "k_fft_x_green", "k_fft_y_bwd",
                                                "k_fft_z_c2r"};
    const hy::FftPass<T>& P = K.pass[p][v];
    const uint64_t N = (uint64_t)c->N;
    uint64_t blocks;
    if (p == hy::FFT_Z_R2C || p == hy::FFT_Z_C2R) blocks = N * (N / P.rows);
    else if (p == hy::FFT_X_GREEN) blocks = (uint64_t)g * N / P.rows;
    else blocks = (uint64_t)g * (N / P.rows);
    a.kz0 = kz0;
    hipLaunchKernelGGL(P.fn, dim3((uint32_t)blocks, c->nbox), dim3(P.threads), P.lds, c->stream, a);
    launch_check(names[p]);
}

template <typename T>
void launch_fft(hy_pm_ctx* c, const hy::FftKernels<T>& K) {
    hy::FftArgs<T> a = fft_args<T>(c);
    const int NH = c->N / 2 + 1;
    const int* v = c->fft_var;
    launch_pass(c, K, a, hy::FFT_Z_R2C, v[hy::FFT_Z_R2C], 0, NH);
    mark(c, HY_STAGE_FFT_Z + 1);
    // The y, x+Green and y^-1 passes touch one kz plane (N^2 complex) at a
    // time.  Running them over groups of kz planes keeps a group (<= the
    // Infinity Cache) resident between the three passes.
    const int group = c->plane_group > 0 ? std::min(c->plane_group, NH) : NH;
    for (int kz0 = 0; kz0 < NH; kz0 += group) {
        const int g = std::min(group, NH - kz0);
        launch_pass(c, K, a, hy::FFT_Y_FWD, v[hy::FFT_Y_FWD], kz0, g);
        if (kz0 + g >= NH) mark(c, HY_STAGE_FFT_Y + 1);
        launch_pass(c, K, a, hy::FFT_X_GREEN, v[hy::FFT_X_GREEN], kz0, g);
        if (kz0 + g >= NH) mark(c, HY_STAGE_FFT_X_GREEN + 1);
        launch_pass(c, K, a, hy::FFT_Y_BWD, v[hy::FFT_Y_BWD], kz0, g);
    }
    mark(c, HY_STAGE_FFT_YINV + 1);
    launch_pass(c, K, a, hy::FFT_Z_C2R, v[hy::FFT_Z_C2R], 0, NH);
    mark(c, HY_STAGE_FFT_ZINV + 1);
}

void run_fft(hy_pm_ctx* c) {
    if (c->prec == HY_PREC_FP64) launch_fft<double>(c, c->fft64);
    else launch_fft<float>(c, c->fft32);
}

// TreePM long-range force meshes from the density source in R: r2c, y, then
// the spectral-gradient passes (src/gravity/pm_solver.cpp:768-810):
//   x: B2 = X^-1(-i k_x G phi), B = X^-1(G phi)
//   F_x = Z^-1 Y^-1 B2 -> R,  F_y = Z^-1 Y^-1(-i k_y B) -> R2,  F_z = Z^-1(-i k_z Y^-1 B) -> R3
// The dual passes (binary32 two-level meshes) form both x outputs from one
// read of B, and Y^-1(-i k_y B) -> A with -i k_z Y^-1 B -> B2 (free once F_x
// is done) from one read of B; all three z c2r passes then run without a
// derivative.
template <typename T>
void launch_fft_sg(hy_pm_ctx* c, const hy::FftKernels<T>& K) {
    using C = hy::cplx<T>;
    hy::FftArgs<T> a = fft_args<T>(c);
    a.kc = static_cast<const T*>(c->kc);
    a.kd = static_cast<const T*>(c->kd);
    a.rs = c->split_scale;
    const uint64_t N = (uint64_t)c->N, NH = N / 2 + 1;
    launch_pass(c, K, a, hy::FFT_Z_R2C, c->fft_var[hy::FFT_Z_R2C], 0, (int)NH);
    launch_pass(c, K, a, hy::FFT_Y_FWD, c->fft_var[hy::FFT_Y_FWD], 0, (int)NH);
    auto launch = [&](const hy::FftPass<T>& P, int which, const hy::FftArgs<T>& aa) {
        const uint64_t blocks = which == 0 ? NH * N / (uint64_t)P.rows
                                           : (which == 1 ? NH * (N / (uint64_t)P.rows) : N * (N / (uint64_t)P.rows));
        hipLaunchKernelGGL(P.fn, dim3((uint32_t)blocks, 1), dim3(P.threads), P.lds, c->stream, aa);
        launch_check(which == 0 ? "k_fft_x_green (spectral gradient)"
                                : (which == 1 ? "k_fft_y_bwd (spectral gradient)" : "k_fft_z_c2r (spectral gradient)"));
    };
    auto sg = [&](int which, const hy::FftArgs<T>& aa) { launch(K.sg[which], which, aa); };
    a.kz0 = 0;
    hy::FftArgs<T> b = a;
    b.xout = static_cast<C*>(c->B2);
    if (K.sg_x_dual) {
        sg(0, b);  // -i k_x spectrum -> B2, plain -> B, one read of B
    } else {
        b.deriv = 1;
        sg(0, b);
        sg(0, a);
    }
    b = a;
    b.B = static_cast<C*>(c->B2);
    sg(1, b);
    sg(2, b);  // F_x -> R
    if (K.sg_y_dual.fn) {
        b = a;
        b.xout = static_cast<C*>(c->B2);
        launch(K.sg_y_dual, 1, b);  // Y^-1(-i k_y B) -> A, -i k_z Y^-1 B -> B2
        b.R = static_cast<T*>(c->R2);
        sg(2, b);  // F_y -> R2
        b = a;
        b.A = static_cast<C*>(c->B2);
        b.R = static_cast<T*>(c->R3);
        sg(2, b);  // F_z -> R3
        return;
    }
    b = a;
    b.deriv = 2;
    sg(1, b);
    b.deriv = 0;
    b.R = static_cast<T*>(c->R2);
    sg(2, b);  // F_y -> R2
    b = a;
    sg(1, b);
    b.deriv = 3;
    b.R = static_cast<T*>(c->R3);
    sg(2, b);  // F_z -> R3
}

void run_fft_sg(hy_pm_ctx* c) {
    if (c->prec == HY_PREC_FP64) launch_fft_sg<double>(c, c->fft64);
    else launch_fft_sg<float>(c, c->fft32);
}

// CIC interpolation of `meshes` meshes (single box): out (written) or add (added).
template <typename T>
void launch_interp(hy_pm_ctx* c, hy::InterpArgs<T> a) {
    a.g = c->geom;
    a.err = c->err;
    const bool o32 = (uint64_t)c->N * c->N * c->N * sizeof(T) <= (1ull << 32);
    const dim3 grid((uint32_t)blocks_for(a.n, 256));
    if (o32) hipLaunchKernelGGL((hy::k_interp<T, true>), grid, dim3(256), 0, c->stream, a);
    else hipLaunchKernelGGL((hy::k_interp<T, false>), grid, dim3(256), 0, c->stream, a);
    launch_check("k_interp");
}

void require_pure_pm(const hy_pm_ctx* c, const char* what) {
    if (c->method != HY_METHOD_PURE_PM)
        fail(HY_EINVAL, std::string(what) + " takes a pure-PM context (the TreePM long-range force needs the "
                                            "short-range tree force beside it; use hy_pm_force)");
}

// Lattice block schedule for n = S^3 particles when a 256-particle block
// tiles z-rows exactly (S a multiple of 256); about 64K particles (the
// blocks resident on one XCD) per column tile.  Identity otherwise.
// rows_pref: lattice rows per column tile (HY_GATHER_ROWS_DEFAULT: 256 / ZC,
// about 256 blocks per tile; 0: particle order).
hy::LatticeTiling lattice_tiling(uint64_t n, uint32_t block, uint32_t rows_pref = HY_GATHER_ROWS_DEFAULT) {
    hy::LatticeTiling t{0, 0, 0, 0, -1, -1, -1, -1};
    if (rows_pref == 0) return t;
    const uint64_t S = (uint64_t)std::llround(std::cbrt((double)n));
    if (S == 0 || S * S * S != n || S % block != 0 || S > 2048) return t;
    const uint32_t ZC = (uint32_t)(S / block);
    uint32_t rows = rows_pref == HY_GATHER_ROWS_DEFAULT ? std::max<uint32_t>(1, 256 / ZC) : rows_pref;  // TX*TY
    uint32_t TX = 1, TY = 1;
    while (TX * TY < rows) {
        if (TY <= TX && S % (TY * 2) == 0) TY *= 2;
        else if (S % (TX * 2) == 0) TX *= 2;
        else break;
    }
    t.S = (uint32_t)S;
    t.TX = TX;
    t.TY = TY;
    t.ZC = ZC;
    auto lg = [](uint64_t v) { int l = 0; while ((1ull << l) < v) ++l; return (1ull << l) == v ? l : -1; };
    const int a = lg(S), b = lg(TX), cc = lg(TY), d = lg(ZC);
    if (a >= 0 && b >= 0 && cc >= 0 && d >= 0) {
        t.lS = a; t.lTX = b; t.lTY = cc; t.lZC = d;
    }
    return t;
}

template <typename T, typename PT>
void launch_gather(hy_pm_ctx* c, hy::GatherArgs<T, PT> a) {
    a.lat = c->nbox == 1 ? lattice_tiling(a.n, 256, c->gather_rows) : hy::LatticeTiling{0, 0, 0, 0, -1, -1, -1, -1};
    a.dfac = c->dfac;
    // 32-bit stencil offsets from a per-wave box base: no wave may span two boxes
    const bool wave_in_box = c->nbox == 1 || (a.n % c->nbox == 0 && (a.n / c->nbox) % 64 == 0);
    const bool o32 = wave_in_box && (uint64_t)c->N * c->N * c->N * sizeof(T) <= (1ull << 32);
    const dim3 grid((uint32_t)blocks_for(a.n, 256)), block(256);
    const bool out = a.acc_x || a.g_x;
    void (*kern)(hy::GatherArgs<T, PT>) =
        out ? (o32 ? hy::k_gather_kdk<T, PT, true, false, true> : hy::k_gather_kdk<T, PT, false, false, true>)
            : (o32 ? hy::k_gather_kdk<T, PT, true, false, false> : hy::k_gather_kdk<T, PT, false, false, false>);
    if constexpr (std::is_same_v<T, float>) {
        if (a.rec_out)
            kern = out ? (o32 ? hy::k_gather_kdk<T, PT, true, true, true> : hy::k_gather_kdk<T, PT, false, true, true>)
                       : (o32 ? hy::k_gather_kdk<T, PT, true, true, false>
                              : hy::k_gather_kdk<T, PT, false, true, false>);
    }
    hipLaunchKernelGGL(kern, grid, block, 0, c->stream, a);
    launch_check("k_gather_kdk");
}

// Potential-energy diagnostic of the potential in R (pure PM, single box):
// workgroup partial sums, then 0.5 * their ordered sum into
// energy_partials[energy_cap].
void launch_potential_energy(hy_pm_ctx* c, uint64_t n, const double* x, const double* y, const double* z,
                             const double* masses, double uniform) {
    const uint64_t blocks = blocks_for(n, hy::ENERGY_BLOCK);
    if (blocks > c->energy_cap) {
        const uint64_t cap = std::max<uint64_t>(blocks, blocks_for(c->cap, hy::ENERGY_BLOCK));
        c->energy_partials = static_cast<double*>(c->alloc((cap + 1) * sizeof(double)));
        c->energy_cap = cap;
    }
    if (c->prec == HY_PREC_FP64)
        hipLaunchKernelGGL(hy::k_potential_partials<double>, dim3((uint32_t)blocks), dim3(hy::ENERGY_THREADS), 0,
                           c->stream, c->geom, n, x, y, z, masses, uniform, static_cast<const double*>(c->R),
                           c->energy_partials, c->err);
    else
        hipLaunchKernelGGL(hy::k_potential_partials<float>, dim3((uint32_t)blocks), dim3(hy::ENERGY_THREADS), 0,
                           c->stream, c->geom, n, x, y, z, masses, uniform, static_cast<const float*>(c->R),
                           c->energy_partials, c->err);
    hipLaunchKernelGGL(hy::k_potential_total, dim3(1), dim3(64), 0, c->stream, c->energy_partials, blocks,
                       c->energy_partials + c->energy_cap, c->err);
    launch_check("k_potential_partials");
}

// Mass conservation (src/math/cic_mass_conservation.cpp:25-63).  The FP32
// deposit sums cells in binary32, so its bound uses binary32 epsilon per cell
// plus the binary64 bound of the reference.
void check_mass(hy_pm_ctx* c, double input, double deposited, uint64_t n) {
    if (!std::isfinite(deposited) || deposited <= 0.0)
        fail(HY_ERUNTIME, "PMSolver total deposited mass must be finite and positive");
    const long double scale = std::max<long double>(input, deposited);
    const long double tol = mass_tol_factor(c, n) * scale;
    if (std::fabs(static_cast<long double>(deposited) - static_cast<long double>(input)) > tol)
        fail(HY_ERUNTIME, "CIC deposited mass violates the forward-error bound");
}

// Deposit stage: bucketize, deposit (+ retry with a larger fallback scratch).
// carried_rec (binary32, asynchronous): deposit the records the previous
// fused step wrote into its capacity layout for these positions.  rec_out:
// also lay out the capacities of the records this step's gather writes.
// shift: deposit the positions x - shift (the power spectrum's interlacing
// partner), routed and wrapped as the reference's shifted copy would be.
void run_deposit(hy_pm_ctx* c, uint64_t n, const double* x, const double* y, const double* z,
                 const double* masses, double uniform, double Min, bool sync, bool carried = false,
                 bool carried_rec = false, bool rec_out = false, double shift = 0.0) {
    mark(c, 0);
    // the binary32 deposit reads records only: no particle order is written
    uint32_t* order = nullptr;
    uint32_t* capn = rec_out ? c->cap_next : nullptr;
    const uint32_t* start = nullptr;
    const uint32_t* ovf = nullptr;
    if (carried_rec) {
        // the layout the records were written in, then the overflow prefix and
        // the next layout from the final counts; overflow records are placed
        // after the capacity slots
        hipLaunchKernelGGL(hy::k_copy_u32, dim3(512), dim3(256), 0, c->stream, c->cap_next, c->cap_cur,
                           c->nball + 1);
        launch_check("k_copy_u32");
        cap_scan(c, c->cap_cur, nullptr, capn);
        hipLaunchKernelGGL(hy::k_place_overflow, dim3(1024), dim3(256), 0, c->stream, c->ovf_count, c->keys,
                           c->ranks, c->ovf_rec, c->cap_cur, c->ovf_start, c->nball, c->rec);
        launch_check("k_place_overflow");
        start = c->cap_cur;
        ovf = c->ovf_start;
    } else {
        bucketize(c, n, x, y, z, 0, c->counts, c->offsets, c->nball, order, carried, masses, capn, shift);
    }
    mark(c, HY_STAGE_BUCKET + 1);
    auto plan = plan_deposit(c, n, x, y, z, masses, uniform, Min, shift);
    launch_deposit(c, plan.P, true, start, ovf);
    mark(c, HY_STAGE_DEPOSIT + 1);
    c->last_input_mass = Min;
    if (!sync) return;
    for (int attempt = 0;; ++attempt) {
        const auto st = fetch_status(c);
        if (carried && (st.err & hy::ERR_STALE_KEYS)) {
            // the positions changed since the keys were made: route them afresh
            // (coordinate and scratch errors of the stale routing are its artefacts)
            raise_errors(st.err & (hy::ERR_POSITION_NONFINITE | hy::ERR_MASS_INVALID));
            carried = false;
            reset_status(c);
            bucketize(c, n, x, y, z, 0, c->counts, c->offsets, c->nball, order, false, masses, nullptr, shift);
            plan = plan_deposit(c, n, x, y, z, masses, uniform, Min, shift);
            launch_deposit(c, plan.P);
            continue;
        }
        if ((st.err & hy::ERR_SCRATCH) && attempt < 3) {
            // grow the fallback scratch (records of the hot tiles) and redo the
            // deposit; errors the retry would clear are raised first
            raise_errors(st.err & ~(hy::ERR_SCRATCH | hy::ERR_MASS_CONSERVATION | hy::ERR_ROUTING));
            const uint64_t grow = c->scratch_cap * 2;
            void* p = nullptr;
            hip_check(hipMalloc(&p, grow * sizeof(uint64_t)), "hipMalloc scratch");
            c->allocations.push_back(p);
            c->device_bytes += grow * sizeof(uint64_t);
            c->scratch = static_cast<uint64_t*>(p);
            c->scratch_cap = grow;
            reset_status(c);
            plan = plan_deposit(c, n, x, y, z, masses, uniform, Min, shift);
            launch_deposit(c, plan.P);
            continue;
        }
        raise_errors(st.err);
        if (st.total != n) fail(HY_ELOGIC, "CIC bucket routing does not cover each particle exactly once");
        c->last_deposited_mass = st.mass;
        if (c->nbox == 1) {
            check_mass(c, Min, st.mass, n);
        } else {
            hip_check(hipMemcpyAsync(c->hmass, c->mass_total, c->nbox * sizeof(double), hipMemcpyDeviceToHost,
                                     c->stream), "mass copy");
            hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
            for (uint32_t b = 0; b < c->nbox; ++b) check_mass(c, Min, c->hmass[b], n / c->nbox);
        }
        return;
    }
}

// Device staging of host-pointer calls: positions and masses (4 * cap), accelerations (3 * cap).
void ensure_host_staging(hy_pm_ctx* c) {
    if (!c->hpos) {
        c->hpos = (double*)c->alloc(c->cap * 4 * sizeof(double));
        c->hacc = (double*)c->alloc(c->cap * 3 * sizeof(double));
    }
}

void require_particles(hy_pm_ctx* c, uint64_t n) {
    if (n > c->cap) fail(HY_EINVAL, "particle count exceeds the context capacity (max_particles)");
    if (n > 0xffffffffull) fail(HY_EOVERFLOW, "particle count exceeds the 32-bit index range");
}

static_assert(hy::FFT_PASSES == HY_FFT_PASSES && hy::kFftVariants == HY_FFT_MAX_VARIANTS,
              "FFT tuning table layout");

template <typename T>
void fill_tuning(const hy_pm_ctx* c, const hy::FftKernels<T>& K, hy_fft_tuning* t) {
    t->variants = (uint32_t)K.variants;
    t->plane_group = (uint32_t)c->plane_group;
    for (int p = 0; p < hy::FFT_PASSES; ++p) {
        t->variant[p] = (uint32_t)c->fft_var[p];
        for (int v = 0; v < HY_FFT_MAX_VARIANTS; ++v)
            t->rows[p][v] = v < K.variants ? (uint32_t)K.pass[p][v].rows : 0u;
    }
}

// Times each launch variant of every pass on zero-filled mesh buffers (all
// passes stay finite), keeps the fastest, then times whole transforms over
// plane-group sizes.
template <typename T>
void autotune(hy_pm_ctx* c, const hy::FftKernels<T>& K, uint32_t reps, hy_fft_tuning* t) {
    const uint64_t N = (uint64_t)c->N, NH = N / 2 + 1;
    hip_check(hipMemsetAsync(c->R, 0, c->nbox * N * N * N * sizeof(T), c->stream), "memset R");
    hip_check(hipMemsetAsync(c->A, 0, c->nbox * N * NH * N * 2 * sizeof(T), c->stream), "memset A");
    hip_check(hipMemsetAsync(c->B, 0, c->nbox * N * NH * N * 2 * sizeof(T), c->stream), "memset B");
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    const bool prof = c->prof;
    c->prof = false;
    auto timed = [&](auto&& launch) {
        launch();  // warm
        hip_check(hipEventRecord(e0, c->stream), "hipEventRecord");
        for (uint32_t r = 0; r < reps; ++r) launch();
        hip_check(hipEventRecord(e1, c->stream), "hipEventRecord");
        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        return ms / (float)reps;
    };
    hy::FftArgs<T> a = fft_args<T>(c);
    for (int p = 0; p < hy::FFT_PASSES; ++p) {
        int best = 0;
        for (int v = 0; v < HY_FFT_MAX_VARIANTS; ++v) t->pass_ms[p][v] = 0.0f;
        for (int v = 0; v < K.variants; ++v) {
            if (K.pass[p][v].lds > c->max_lds) {
                t->pass_ms[p][v] = -1.0f;  // not launchable on this device
                continue;
            }
            t->pass_ms[p][v] = timed([&] { launch_pass(c, K, a, p, v, 0, (int)NH); });
            if (t->pass_ms[p][v] < t->pass_ms[p][best]) best = v;
        }
        c->fft_var[p] = best;
    }
    static const int groups[4] = {0, 16, 64, 128};
    int best_g = 0;
    for (int i = 0; i < 4; ++i) {
        t->group_ms[i] = 0.0f;
        if (groups[i] != 0 && groups[i] >= (int)NH) continue;
        c->plane_group = groups[i];
        t->group_ms[i] = timed([&] { launch_fft<T>(c, K); });
        // a plane group must win by 2% to replace whole-mesh launches
        if (i > 0 && t->group_ms[i] < 0.98f * t->group_ms[best_g]) best_g = i;
    }
    c->plane_group = groups[best_g];
    c->prof = prof;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    hip_check(hipMemsetAsync(c->err, 0, sizeof(uint32_t), c->stream), "memset err");
    hip_check(hipStreamSynchronize(c->stream), "sync");
    c->keys_ready = c->rec_ready = false;
    fill_tuning(c, K, t);
}

}  // namespace

extern "C" {

int hy_abi_version(void) { return HY_ABI_VERSION; }

const char* hy_last_error(void) { return g_last_error.c_str(); }

int hy_pm_create(const hy_pm_desc* desc, hy_pm_ctx** out) {
    return guarded([&] {
        if (!desc || !out) fail(HY_EINVAL, "hy_pm_create: null argument");
        *out = nullptr;
        const double L = desc->box_size;
        const uint64_t Nu = desc->mesh_size;
        if (!std::isfinite(L) || L <= 0.0) fail(HY_EINVAL, "MeshGeometry box size must be finite and positive");
        if (desc->precision != HY_PREC_FP64 && desc->precision != HY_PREC_FP32_MIXED)
            fail(HY_EINVAL, "unknown precision");
        hy::FftKernels<float> k32{};
        hy::FftKernels<double> k64{};
        if (Nu > 4096 || !hy::lookup_fft((int)Nu, k32) || !hy::lookup_fft((int)Nu, k64))
            fail(HY_EINVAL, "mesh size must be 2^a (8..1024) or 3*2^a (12..384)");
        const int N = (int)Nu;
        auto c = std::make_unique<hy_pm_ctx>();
        c->desc = *desc;
        c->N = N;
        c->prec = desc->precision;
        c->real_bytes = c->prec == HY_PREC_FP64 ? sizeof(double) : sizeof(float);
        c->fft32 = k32;
        c->fft64 = k64;
        const double dx = L / (double)N;
        if (!std::isfinite(dx) || dx <= 0.0) fail(HY_EOVERFLOW, "MeshGeometry cell size is not representable");
        c->geom.L = L;
        c->geom.dx = dx;
        {
            int e2 = 0;
            const double mant = std::frexp(dx, &e2);
            c->geom.inv_exact = (mant == 0.5 && std::isfinite(1.0 / dx)) ? 1 : 0;
            c->geom.inv_dx = 1.0 / dx;
        }
        c->geom.upper = std::nextafter((double)N, 0.0);
        c->nbox = desc->boxes == 0 ? 1u : desc->boxes;
        if (c->nbox > 65535) fail(HY_EINVAL, "box count exceeds 65535");
        // PMForceMethod named constructors (include/cosmo_nbody/mesh/pm_force_method.hpp:20-34)
        if (desc->method == HY_METHOD_TREEPM_LONG_RANGE) {
            if (!std::isfinite(desc->split_scale) || desc->split_scale <= 0.0)
                fail(HY_EINVAL, "TreePM long-range PM method requires a finite positive split scale");
            if (c->nbox != 1) fail(HY_EINVAL, "TreePM long-range contexts take a single box");
            c->desc.deconvolve_cic = 1;
        } else if (desc->method != HY_METHOD_PURE_PM) {
            fail(HY_EINVAL, "unknown PM method");
        }
        c->method = desc->method;
        c->split_scale = desc->split_scale;
        c->geom.nbox = c->nbox;
        c->geom.nper = 0;
        c->geom.N = N;
        c->geom.T = tile_edge_for(N);
        c->geom.nT = N / c->geom.T;
        c->cap = desc->max_particles;
        if (const char* env = std::getenv("HY_PLANE_GROUP")) c->plane_group = std::atoi(env);
        set_device(c.get());
        hip_check(hipDeviceGetAttribute(&c->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, desc->device),
                  "hipDeviceGetAttribute");
        for (int p = 0; p < hy::FFT_PASSES; ++p)
            if (std::max(k32.pass[p][0].lds, k64.pass[p][0].lds) > c->max_lds)
                fail(HY_EDEVICE, "device LDS per workgroup is below the FFT passes' requirement");
        if (desc->stream) {
            c->stream = static_cast<hipStream_t>(desc->stream);
        } else {
            // a blocking stream: ordered with the legacy default stream, on
            // which callers (PyTorch's default stream among them) produce
            // the device arrays they pass in
            hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamDefault), "hipStreamCreate");
            c->own_stream = true;
        }

        const uint64_t cells = (uint64_t)N * N * N;
        const uint64_t NH = N / 2 + 1;
        const uint64_t ccount = (uint64_t)N * NH * N;
        const size_t rb = c->real_bytes;
        c->R = c->alloc(c->nbox * cells * rb);
        c->A = c->alloc(c->nbox * ccount * 2 * rb);
        c->B = c->alloc(c->nbox * ccount * 2 * rb);
        c->tw = c->alloc((size_t)N * 2 * rb);
        c->lap = c->alloc((size_t)N * rb);
        c->win = c->alloc((size_t)N * rb);

        // Tables: twiddles exp(-2 pi i m/N) in long double; Laplacian symbol and
        // CIC window exactly as src/mesh/green_function.cpp:59-87,137-163.
        const long double pil = 3.141592653589793238462643383279502884L;
        std::vector<double> tw(2 * N), lap(N), win(N);
        const double inv = 1.0 / dx;
        const double inv2 = inv * inv;
        for (int m = 0; m < N; ++m) {
            const long double ang = -2.0L * pil * (long double)m / (long double)N;
            tw[2 * m] = (double)std::cos(ang);
            tw[2 * m + 1] = (double)std::sin(ang);
            const int64_t mode = m <= N / 2 ? m : (int64_t)m - N;
            const double h = std::numbers::pi * (double)mode / (double)N;
            const double s = std::sin(h);
            lap[m] = -4.0 * inv2 * s * s;
            const long double arg = pil * (long double)mode / (long double)N;
            if (arg == 0.0L) {
                win[m] = 1.0;
            } else {
                long double sc = std::fabs(std::sin(arg) / arg);
                if (sc > 1.0L) sc = 1.0L;
                win[m] = (double)(sc * sc);
            }
        }
        c->green_num = 4.0 * std::numbers::pi * hy::host::gravitational_constant();
        auto upload = [&](void* dst, const std::vector<double>& v) {
            if (c->prec == HY_PREC_FP64) {
                hip_check(hipMemcpy(dst, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice), "upload");
            } else {
                std::vector<float> f(v.begin(), v.end());
                hip_check(hipMemcpy(dst, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice), "upload");
            }
        };
        upload(c->tw, tw);
        upload(c->lap, lap);
        upload(c->win, win);
        if (c->method == HY_METHOD_TREEPM_LONG_RANGE) {
            // spectral-gradient workspace and k tables (k_component of
            // include/cosmo_nbody/mesh/mesh_geometry.hpp:112-115; the derivative
            // axis drops its Nyquist index, src/gravity/pm_solver.cpp:781-787)
            c->B2 = c->alloc(ccount * 2 * rb);
            c->R2 = c->alloc(cells * rb);
            c->R3 = c->alloc(cells * rb);
            c->kc = c->alloc((size_t)N * rb);
            c->kd = c->alloc((size_t)N * rb);
            std::vector<double> kc(N), kd(N);
            for (int m = 0; m < N; ++m) {
                const int64_t mode = m <= N / 2 ? m : (int64_t)m - N;
                kc[m] = 2.0 * std::numbers::pi * (double)mode / L;
                kd[m] = (N % 2 == 0 && m == N / 2) ? 0.0 : kc[m];
            }
            upload(c->kc, kc);
            upload(c->kd, kd);
        }

        const uint64_t cap = std::max<uint64_t>(c->cap, 1);
        c->keys = (uint32_t*)c->alloc(cap * 4);
        c->ranks = (uint32_t*)c->alloc(cap * 4);
        c->order = (uint32_t*)c->alloc(cap * 4);
        c->ntiles = (uint64_t)c->geom.nT * c->geom.nT * c->geom.nT;
        c->nb = c->ntiles * 8;
        c->nball = c->nb * c->nbox;
        if (c->nball > 0xffffffffull) fail(HY_EINVAL, "bucket count over all boxes exceeds the 32-bit key range");
        const uint64_t nb_max = std::max<uint64_t>(c->nball, (uint64_t)N);
        c->counts = (uint32_t*)c->alloc(nb_max * 4);
        c->offsets = (uint32_t*)c->alloc((nb_max + 1) * 4);
        c->block_sums = (uint32_t*)c->alloc((blocks_for(nb_max, hy::SCAN_TILE) + 1) * 4);
        c->grand_total = (uint32_t*)c->alloc(16);
        c->xcounts = (uint32_t*)c->alloc((size_t)N * 4);
        c->xoffsets = (uint32_t*)c->alloc(((size_t)N + 1) * 4);
        c->mass_digits = (unsigned long long*)c->alloc((uint64_t)hy::SA_WORDS * c->nbox * 8);
        const uint64_t wave_words = c->prec == HY_PREC_FP64 ? hy::SA_WAVE_WORDS : 2;  // window, or digit-sum pair
        c->mass_waves = (unsigned long long*)c->alloc(wave_words * hy::SA_WAVES_PER_TILE * c->ntiles * c->nbox * 8);
        c->mass_total = (double*)c->alloc(std::max<uint64_t>(2, c->nbox) * 8);
        c->scratch_cap = 2 * cap + (1u << 16);
        c->scratch = (uint64_t*)c->alloc(c->scratch_cap * 8);
        c->scratch_cursor = (unsigned long long*)c->alloc(16);
        c->err = (uint32_t*)c->alloc(16);
        if (c->prec == HY_PREC_FP64) {
            c->rec64 = (hy::DepRecord64*)c->alloc(cap * sizeof(hy::DepRecord64));
            c->rec_mass64 = (double*)c->alloc(cap * sizeof(double));
        } else {
            // capacity layout of carried records: the capacities (count + count/8
            // + slack per bucket) and room for every record to overflow
            const uint64_t slots = 2 * cap + cap / 8 + (uint64_t)hy::CAP_SLACK * c->nball + 64;
            c->rec_layout = slots <= 0xffffffffull;
            c->rec = (hy::DepRecord*)c->alloc((c->rec_layout ? slots : cap) * sizeof(hy::DepRecord));
            c->rec_mass = (float*)c->alloc(cap * sizeof(float));
            if (c->rec_layout) {
                c->cap_cur = (uint32_t*)c->alloc((c->nball + 1) * 4);
                c->cap_next = (uint32_t*)c->alloc((c->nball + 1) * 4);
                c->ovf_start = (uint32_t*)c->alloc((c->nball + 1) * 4);
                c->ovf_count = (uint32_t*)c->alloc(16);
                c->cap_block_sums = (uint32_t*)c->alloc((3 * blocks_for(c->nball, hy::SCAN_TILE) + 3) * 4);
                // the binary32 deposit never uses the binary64 fallback scratch
                static_assert(sizeof(hy::DepRecord) == 2 * sizeof(uint64_t), "overflow records in the scratch");
                if (c->scratch_cap < 2 * cap) fail(HY_ELOGIC, "deposit scratch smaller than the overflow list");
                c->ovf_rec = reinterpret_cast<hy::DepRecord*>(c->scratch);
            }
        }
        c->gx = c->alloc(cap * rb);
        c->gy = c->alloc(cap * rb);
        c->gz = c->alloc(cap * rb);
        hip_check(hipHostMalloc((void**)&c->hstatus, sizeof(hy_pm_ctx::Status), hipHostMallocDefault), "hipHostMalloc");
        hip_check(hipHostMalloc((void**)&c->hmass, c->nbox * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
        hip_check(hipMemset(c->err, 0, 16), "memset");
        hip_check(hipMemset(c->mass_total, 0, 16), "memset");
        hip_check(hipMemset(c->grand_total, 0, 16), "memset");
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        *out = c.release();
    });
}

int hy_pm_destroy(hy_pm_ctx* ctx) {
    return guarded([&] {
        if (!ctx) return;
        (void)hipSetDevice(ctx->desc.device);
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        delete ctx;
    });
}

int hy_pm_get_info(const hy_pm_ctx* c, hy_pm_info* info) {
    return guarded([&] {
        if (!c || !info) fail(HY_EINVAL, "hy_pm_get_info: null argument");
        info->mesh_size = c->N;
        info->tile_edge = c->geom.T;
        info->bucket_count = c->nb;
        info->device_bytes = c->device_bytes;
        info->last_input_mass = c->last_input_mass;
        info->last_deposited_mass = c->last_deposited_mass;
        info->precision = c->prec;
        info->boxes = c->nbox;
    });
}

int hy_pm_potential_energy(const hy_pm_ctx* c, double* out) {
    return guarded([&] {
        if (!c || !out) fail(HY_EINVAL, "hy_pm_potential_energy: null argument");
        if (!c->energy_valid)
            fail(HY_ELOGIC, "the last hy_pm_force call did not collect the potential energy");
        *out = c->last_potential_energy;
    });
}

int hy_pm_check(hy_pm_ctx* c) {
    return guarded([&] {
        if (!c) fail(HY_EINVAL, "null context");
        set_device(c);
        const auto st = fetch_status(c);
        if (st.err) {
            // report once; the context stays usable
            c->keys_ready = c->rec_ready = false;
            reset_status(c);
            hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        }
        raise_errors(st.err);
    });
}

int hy_pm_density(hy_pm_ctx* c, uint64_t n, const double* x, const double* y, const double* z,
                  const double* masses, double uniform, void* mesh_out) {
    return guarded([&] {
        if (!c || !x || !y || !z || !mesh_out) fail(HY_EINVAL, "hy_pm_density: null argument");
        require_single_box(c, "hy_pm_density");
        set_device(c);
        c->keys_ready = c->rec_ready = false;
        require_particles(c, n);
        set_boxes(c, n);
        const double M = input_mass(c, n, masses, uniform, false);
        reset_status(c);
        if (masses) {
            hipLaunchKernelGGL(hy::k_check_mass, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, masses, c->err);
        }
        run_deposit(c, n, x, y, z, masses, uniform, M, true);
        const size_t bytes = (size_t)c->N * c->N * c->N * c->real_bytes;
        hip_check(hipMemcpyAsync(mesh_out, c->R, bytes, hipMemcpyDeviceToDevice, c->stream), "copy mesh");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

int hy_pm_potential(hy_pm_ctx* c, const void* source_in, void* phi_out) {
    return guarded([&] {
        if (!c || !source_in || !phi_out) fail(HY_EINVAL, "hy_pm_potential: null argument");
        require_single_box(c, "hy_pm_potential");
        require_pure_pm(c, "hy_pm_potential");
        set_device(c);
        const size_t bytes = (size_t)c->N * c->N * c->N * c->real_bytes;
        reset_status(c);
        hip_check(hipMemcpyAsync(c->R, source_in, bytes, hipMemcpyDeviceToDevice, c->stream), "copy source");
        run_fft(c);
        hip_check(hipMemcpyAsync(phi_out, c->R, bytes, hipMemcpyDeviceToDevice, c->stream), "copy phi");
        const auto st = fetch_status(c);
        raise_errors(st.err);
    });
}

int hy_pm_force_meshes(hy_pm_ctx* c, const void* source_in, void* fx_out, void* fy_out, void* fz_out) {
    return guarded([&] {
        if (!c || !source_in || !fx_out || !fy_out || !fz_out) fail(HY_EINVAL, "hy_pm_force_meshes: null argument");
        if (c->method != HY_METHOD_TREEPM_LONG_RANGE)
            fail(HY_EINVAL, "hy_pm_force_meshes takes a TreePM long-range context");
        set_device(c);
        const size_t bytes = (size_t)c->N * c->N * c->N * c->real_bytes;
        reset_status(c);
        hip_check(hipMemcpyAsync(c->R, source_in, bytes, hipMemcpyDeviceToDevice, c->stream), "copy source");
        run_fft_sg(c);
        hip_check(hipMemcpyAsync(fx_out, c->R, bytes, hipMemcpyDeviceToDevice, c->stream), "copy F_x");
        hip_check(hipMemcpyAsync(fy_out, c->R2, bytes, hipMemcpyDeviceToDevice, c->stream), "copy F_y");
        hip_check(hipMemcpyAsync(fz_out, c->R3, bytes, hipMemcpyDeviceToDevice, c->stream), "copy F_z");
        const auto st = fetch_status(c);
        raise_errors(st.err);
    });
}

int hy_ic_lpt(hy_pm_ctx* c, uint64_t n_side, const void* delta_k, double p_factor, double* x, double* y,
              double* z, double* px, double* py, double* pz) {
    return guarded([&] {
        if (!c || !delta_k || !x || !y || !z || !px || !py || !pz) fail(HY_EINVAL, "hy_ic_lpt: null argument");
        require_single_box(c, "hy_ic_lpt");
        if (c->prec != HY_PREC_FP64) fail(HY_EINVAL, "hy_ic_lpt takes a binary64 context");
        const uint64_t N = (uint64_t)c->N;
        if (n_side == 0 || n_side > N || N % n_side != 0)
            fail(HY_EINVAL, "1LPT particle lattice must divide the IC mesh");
        if (!std::isfinite(p_factor)) fail(HY_EINVAL, "Invalid start-scaled LPT normalization");
        set_device(c);
        if (!c->ic_kc) {
            std::vector<double> kc(N), kd(N);
            for (uint64_t m = 0; m < N; ++m) {
                const int64_t mode = m <= N / 2 ? (int64_t)m : (int64_t)m - (int64_t)N;
                kc[m] = 2.0 * std::numbers::pi * (double)mode / c->geom.L;
                kd[m] = (N % 2 == 0 && m == N / 2) ? 0.0 : kc[m];
            }
            c->ic_kc = c->alloc(N * sizeof(double));
            c->ic_kd = c->alloc(N * sizeof(double));
            hip_check(hipMemcpy(c->ic_kc, kc.data(), N * sizeof(double), hipMemcpyHostToDevice), "upload");
            hip_check(hipMemcpy(c->ic_kd, kd.data(), N * sizeof(double), hipMemcpyHostToDevice), "upload");
        }
        reset_status(c);
        const uint64_t NH = N / 2 + 1;
        const double L = c->geom.L;
        const double dxp = L / (double)n_side;
        const double norm = 1.0 / ((double)N * (double)N * (double)N);
        const uint64_t np = n_side * n_side * n_side;
        hy::FftArgs<double> a = fft_args<double>(c);
        const hy::FftKernels<double>& K = c->fft64;
        double* pos[3] = {x, y, z};
        double* mom[3] = {px, py, pz};
        for (int axis = 0; axis < 3; ++axis) {
            hipLaunchKernelGGL(hy::k_lpt_spectrum, dim3((uint32_t)blocks_for(NH * N * N, 256)), dim3(256), 0,
                               c->stream, static_cast<const double2*>(delta_k), axis, (int)N,
                               static_cast<const double*>(c->ic_kc), static_cast<const double*>(c->ic_kd), norm,
                               static_cast<double2*>(c->B));
            launch_check("k_lpt_spectrum");
            const hy::FftPass<double>& P = K.x_inv;
            a.kz0 = 0;
            hipLaunchKernelGGL(P.fn, dim3((uint32_t)(NH * N / (uint64_t)P.rows), 1), dim3(P.threads), P.lds,
                               c->stream, a);
            launch_check("k_fft_x_dft (inverse)");
            launch_pass(c, K, a, hy::FFT_Y_BWD, c->fft_var[hy::FFT_Y_BWD], 0, (int)NH);
            launch_pass(c, K, a, hy::FFT_Z_C2R, c->fft_var[hy::FFT_Z_C2R], 0, (int)NH);
            hipLaunchKernelGGL(hy::k_lpt_update, dim3((uint32_t)blocks_for(np, 256)), dim3(256), 0, c->stream,
                               static_cast<const double*>(c->R), axis, (int)N, (int)n_side, (int)(N / n_side), dxp,
                               L, p_factor, pos[axis], mom[axis], c->err);
            launch_check("k_lpt_update");
        }
        const auto st = fetch_status(c);
        if (st.err & (hy::ERR_UPDATE_NONFINITE | hy::ERR_FIELD_NONFINITE))
            fail(HY_ERUNTIME, "1LPT particle update is non-finite or unrepresentable");
        raise_errors(st.err);
    });
}

int hy_pm_interpolate(hy_pm_ctx* c, uint64_t n, const double* x, const double* y, const double* z,
                      const void* mesh, double* out) {
    return guarded([&] {
        if (!c || !x || !y || !z || !mesh || !out) fail(HY_EINVAL, "hy_pm_interpolate: null argument");
        require_single_box(c, "hy_pm_interpolate");
        set_device(c);
        reset_status(c);
        if (n > 0) {
            auto interp = [&](auto tag) {
                using T = decltype(tag);
                hy::InterpArgs<T> a{};
                a.n = n;
                a.x = x; a.y = y; a.z = z;
                a.f[0] = static_cast<const T*>(mesh);
                a.add[0] = out;
                a.meshes = 1;
                launch_interp(c, a);
            };
            if (c->prec == HY_PREC_FP64) interp(double{});
            else interp(float{});
        }
        const auto st = fetch_status(c);
        raise_errors(st.err);
    });
}

int hy_pm_gather(hy_pm_ctx* c, uint64_t n, const double* x, const double* y, const double* z,
                 const void* phi, double* ax, double* ay, double* az) {
    return guarded([&] {
        if (!c || !x || !y || !z || !phi || !ax || !ay || !az) fail(HY_EINVAL, "hy_pm_gather: null argument");
        require_single_box(c, "hy_pm_gather");
        set_device(c);
        require_particles(c, n);
        set_boxes(c, n);
        reset_status(c);
        auto fill = [&](auto& a) {
            a.g = c->geom; a.n = n;
            a.x_in = x; a.y_in = y; a.z_in = z;
            a.acc_x = ax; a.acc_y = ay; a.acc_z = az;
            a.err = c->err;
        };
        if (c->prec == HY_PREC_FP64) {
            hy::GatherArgs<double, double> a{};
            fill(a);
            a.phi = static_cast<const double*>(phi);
            launch_gather(c, a);
        } else {
            hy::GatherArgs<float, float> a{};
            fill(a);
            a.phi = static_cast<const float*>(phi);
            launch_gather(c, a);
        }
        const auto st = fetch_status(c);
        raise_errors(st.err);
    });
}

int hy_pm_force(hy_pm_ctx* c, uint64_t n, const double* x, const double* y, const double* z,
                const double* masses, double uniform, double* ax, double* ay, double* az,
                uint32_t flags) {
    return guarded([&] {
        if (!c) fail(HY_EINVAL, "hy_pm_force: null context");
        if (n == 0) fail(HY_ERUNTIME, "PMSolver exact input mass identity must be finite and positive");
        if (!x || !y || !z || !ax || !ay || !az) fail(HY_EINVAL, "hy_pm_force: null argument");
        set_device(c);
        c->keys_ready = c->rec_ready = false;
        require_particles(c, n);
        const bool host = (flags & HY_HOST_POINTERS) != 0;
        if (host) {
            // reject overlapping output/input spans (src/gravity/pm_solver.cpp:231-263)
            auto overlap = [](const double* a, const double* b, uint64_t k) {
                return k && a < b + k && b < a + k;
            };
            const double* outs[3] = {ax, ay, az};
            const double* ins[4] = {x, y, z, masses};
            for (int i = 0; i < 3; ++i) {
                for (int j = i + 1; j < 3; ++j)
                    if (overlap(outs[i], outs[j], n)) fail(HY_EINVAL, "PMSolver force spans overlap");
                for (int j = 0; j < 4; ++j)
                    if (ins[j] && overlap(outs[i], ins[j], n)) fail(HY_EINVAL, "PMSolver force spans overlap");
            }
        }
        const uint64_t nper = set_boxes(c, n);
        if (masses && c->nbox > 1) fail(HY_EINVAL, "explicit masses take a single-box context");
        const bool collect = (flags & HY_COLLECT_POTENTIAL) != 0;
        c->energy_valid = false;
        if (collect) {
            require_single_box(c, "potential-energy collection");
            if (c->method != HY_METHOD_PURE_PM)
                fail(HY_EINVAL, "potential-energy collection takes a pure-PM context");
        }
        const double M = input_mass(c, nper, masses, uniform, host);
        const double *dx_ = x, *dy_ = y, *dz_ = z, *dm = masses;
        if (host) {
            ensure_host_staging(c);
            double* s = c->hpos;
            hip_check(hipMemcpyAsync(s, x, n * 8, hipMemcpyHostToDevice, c->stream), "h2d");
            hip_check(hipMemcpyAsync(s + c->cap, y, n * 8, hipMemcpyHostToDevice, c->stream), "h2d");
            hip_check(hipMemcpyAsync(s + 2 * c->cap, z, n * 8, hipMemcpyHostToDevice, c->stream), "h2d");
            dx_ = s; dy_ = s + c->cap; dz_ = s + 2 * c->cap;
            if (masses) {
                hip_check(hipMemcpyAsync(s + 3 * c->cap, masses, n * 8, hipMemcpyHostToDevice, c->stream), "h2d");
                dm = s + 3 * c->cap;
            }
        }
        reset_status(c);
        if (dm) hipLaunchKernelGGL(hy::k_check_mass, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, dm, c->err);
        run_deposit(c, n, dx_, dy_, dz_, dm, uniform, M, true);
        auto fill = [&](auto& a) {
            a.g = c->geom; a.n = n;
            a.x_in = dx_; a.y_in = dy_; a.z_in = dz_;
            a.err = c->err;
        };
        if (c->method == HY_METHOD_TREEPM_LONG_RANGE) {
            run_fft_sg(c);
            auto interp = [&](auto tag) {
                using T = decltype(tag);
                hy::InterpArgs<T> a{};
                a.n = n;
                a.x = dx_; a.y = dy_; a.z = dz_;
                a.f[0] = static_cast<const T*>(c->R);
                a.f[1] = static_cast<const T*>(c->R2);
                a.f[2] = static_cast<const T*>(c->R3);
                a.out[0] = static_cast<T*>(c->gx);
                a.out[1] = static_cast<T*>(c->gy);
                a.out[2] = static_cast<T*>(c->gz);
                a.meshes = 3;
                launch_interp(c, a);
            };
            if (c->prec == HY_PREC_FP64) interp(double{});
            else interp(float{});
        } else {
            run_fft(c);
            if (collect) launch_potential_energy(c, n, dx_, dy_, dz_, dm, uniform);
            if (c->prec == HY_PREC_FP64) {
                hy::GatherArgs<double, double> a{};
                fill(a);
                a.phi = static_cast<const double*>(c->R);
                a.g_x = (double*)c->gx; a.g_y = (double*)c->gy; a.g_z = (double*)c->gz;
                launch_gather(c, a);
            } else {
                hy::GatherArgs<float, float> a{};
                fill(a);
                a.phi = static_cast<const float*>(c->R);
                a.g_x = (float*)c->gx; a.g_y = (float*)c->gy; a.g_z = (float*)c->gz;
                launch_gather(c, a);
            }
        }
        mark(c, HY_STAGE_GATHER_KDK + 1);
        end_call(c);
        const auto st = fetch_status(c);
        raise_errors(st.err);
        if (collect) {
            hip_check(hipMemcpyAsync(&c->last_potential_energy, c->energy_partials + c->energy_cap, sizeof(double),
                                     hipMemcpyDeviceToHost, c->stream), "energy copy");
            hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
            c->energy_valid = true;
        }
        // commit: acc += g (only after every check passed)
        double *cx = ax, *cy = ay, *cz = az;
        if (host) {
            cx = c->hacc; cy = c->hacc + c->cap; cz = c->hacc + 2 * c->cap;
            hip_check(hipMemcpyAsync(cx, ax, n * 8, hipMemcpyHostToDevice, c->stream), "h2d");
            hip_check(hipMemcpyAsync(cy, ay, n * 8, hipMemcpyHostToDevice, c->stream), "h2d");
            hip_check(hipMemcpyAsync(cz, az, n * 8, hipMemcpyHostToDevice, c->stream), "h2d");
        }
        if (c->prec == HY_PREC_FP64)
            hipLaunchKernelGGL(hy::k_commit_force<double>, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n,
                               (const double*)c->gx, (const double*)c->gy, (const double*)c->gz, cx, cy, cz);
        else
            hipLaunchKernelGGL(hy::k_commit_force<float>, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n,
                               (const float*)c->gx, (const float*)c->gy, (const float*)c->gz, cx, cy, cz);
        launch_check("k_commit_force");
        if (host) {
            hip_check(hipMemcpyAsync(ax, cx, n * 8, hipMemcpyDeviceToHost, c->stream), "d2h");
            hip_check(hipMemcpyAsync(ay, cy, n * 8, hipMemcpyDeviceToHost, c->stream), "d2h");
            hip_check(hipMemcpyAsync(az, cz, n * 8, hipMemcpyDeviceToHost, c->stream), "d2h");
        }
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

int hy_pm_kdk(hy_pm_ctx* c, uint64_t n, const double* x_in, const double* y_in, const double* z_in,
              const void* px_in, const void* py_in, const void* pz_in, double* x_out, double* y_out,
              double* z_out, void* px_out, void* py_out, void* pz_out, const double* masses,
              double uniform, double k_prev, double k_next, double d_next, double* acc_x,
              double* acc_y, double* acc_z, uint32_t flags) {
    return guarded([&] {
        if (!c) fail(HY_EINVAL, "hy_pm_kdk: null context");
        require_pure_pm(c, "hy_pm_kdk");
        if (n == 0) fail(HY_ERUNTIME, "PMSolver exact input mass identity must be finite and positive");
        if (!x_in || !y_in || !z_in) fail(HY_EINVAL, "hy_pm_kdk: null position");
        const bool has_prev = (flags & HY_HAS_PREV) != 0, has_next = (flags & HY_HAS_NEXT) != 0;
        if ((has_prev || has_next) && (!px_in || !py_in || !pz_in || !px_out || !py_out || !pz_out))
            fail(HY_EINVAL, "hy_pm_kdk: null momentum");
        if (has_next && (!x_out || !y_out || !z_out)) fail(HY_EINVAL, "hy_pm_kdk: null output position");
        if (flags & HY_HOST_POINTERS) fail(HY_EINVAL, "hy_pm_kdk takes device pointers only");
        if ((has_prev && !std::isfinite(k_prev)) || (has_next && (!std::isfinite(k_next) || !std::isfinite(d_next))))
            fail(HY_EOVERFLOW, "Leapfrog drift/kick factors must be finite");
        set_device(c);
        require_particles(c, n);
        const bool sync = (flags & HY_ASYNC) == 0;
        const uint64_t nper = set_boxes(c, n);
        if (masses && c->nbox > 1) fail(HY_EINVAL, "explicit masses take a single-box context");
        const double M = input_mass(c, nper, masses, uniform, false);
        if (sync) reset_status(c);
        else fill_u32(c, c->scratch_cursor, 2);
        if (masses) hipLaunchKernelGGL(hy::k_check_mass, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, masses, c->err);
        // keys (or records) of the previous fused step are valid for exactly its
        // output positions
        const bool reuse = (flags & HY_REUSE_KEYS) != 0;
        const bool carried = reuse && c->keys_ready && c->keys_n == n && c->keys_x == x_in && c->keys_y == y_in &&
                             c->keys_z == z_in;
        const bool carried_rec = reuse && !sync && c->rec_ready && c->rec_n == n && c->rec_x == x_in &&
                                 c->rec_y == y_in && c->rec_z == z_in;
        // binary32 uniform-mass asynchronous steps carry deposit records, others keys
        const bool rec_out = has_next && !sync && !masses && c->rec_layout;
        c->keys_ready = c->rec_ready = false;
        run_deposit(c, n, x_in, y_in, z_in, masses, uniform, M, sync, carried, carried_rec, rec_out);
        // counts are consumed by the scan above; the gather below accumulates the
        // next step's counts into the zeroed buffer
        const bool make_keys = has_next;
        if (make_keys) fill_u32(c, c->counts, c->nball);
        if (rec_out) fill_u32(c, c->ovf_count, 1);
        run_fft(c);
        auto fill = [&](auto& a) {
            a.g = c->geom; a.n = n;
            a.x_in = x_in; a.y_in = y_in; a.z_in = z_in;
            a.x_out = x_out ? x_out : const_cast<double*>(x_in);
            a.y_out = y_out ? y_out : const_cast<double*>(y_in);
            a.z_out = z_out ? z_out : const_cast<double*>(z_in);
            a.acc_x = acc_x; a.acc_y = acc_y; a.acc_z = acc_z;
            a.k_prev = k_prev; a.k_next = k_next; a.d_next = d_next;
            a.has_prev = has_prev; a.has_next = has_next;
            a.err = c->err;
            if (make_keys) {
                a.next_counts = c->counts;
                a.next_keys = c->keys;
                a.next_ranks = c->ranks;
            }
        };
        if (c->prec == HY_PREC_FP64) {
            hy::GatherArgs<double, double> a{};
            fill(a);
            a.phi = static_cast<const double*>(c->R);
            a.px_in = (const double*)px_in; a.py_in = (const double*)py_in; a.pz_in = (const double*)pz_in;
            a.px_out = (double*)px_out; a.py_out = (double*)py_out; a.pz_out = (double*)pz_out;
            launch_gather(c, a);
        } else {
            hy::GatherArgs<float, float> a{};
            fill(a);
            if (rec_out) {
                a.rec_start = c->cap_next;
                a.rec_out = c->rec;
                a.ovf_count = c->ovf_count;
                a.ovf_rec = c->ovf_rec;
            }
            a.phi = static_cast<const float*>(c->R);
            a.px_in = (const float*)px_in; a.py_in = (const float*)py_in; a.pz_in = (const float*)pz_in;
            a.px_out = (float*)px_out; a.py_out = (float*)py_out; a.pz_out = (float*)pz_out;
            launch_gather(c, a);
        }
        mark(c, HY_STAGE_GATHER_KDK + 1);
        end_call(c);
        if (sync) {
            const auto st = fetch_status(c);
            raise_errors(st.err);
        }
        if (rec_out) {
            c->rec_ready = true;
            c->rec_x = x_out ? x_out : x_in;
            c->rec_y = y_out ? y_out : y_in;
            c->rec_z = z_out ? z_out : z_in;
            c->rec_n = n;
        } else if (make_keys) {
            c->keys_ready = true;
            c->keys_x = x_out ? x_out : x_in;
            c->keys_y = y_out ? y_out : y_in;
            c->keys_z = z_out ? z_out : z_in;
            c->keys_n = n;
        }
    });
}

int hy_pm_bucket_order(hy_pm_ctx* c, uint64_t n, const double* x, const double* y, const double* z,
                       uint32_t* order, uint32_t* offsets) {
    return guarded([&] {
        if (!c || !x || !y || !z || !order || !offsets) fail(HY_EINVAL, "hy_pm_bucket_order: null argument");
        require_single_box(c, "hy_pm_bucket_order");
        set_device(c);
        c->keys_ready = c->rec_ready = false;
        require_particles(c, n);
        set_boxes(c, n);
        reset_status(c);
        bucketize(c, n, x, y, z, 0, c->counts, offsets, c->nb, order);
        hipLaunchKernelGGL(hy::k_canonicalize, dim3(std::min<uint64_t>(c->nb, 65535)), dim3(hy::CANON_THREADS), 0,
                           c->stream, c->nb, offsets, order, c->scratch, c->scratch_cap, c->scratch_cursor, c->err);
        launch_check("k_canonicalize");
        const auto st = fetch_status(c);
        raise_errors(st.err);
        if (st.total != n) fail(HY_ELOGIC, "bucket routing does not cover each particle exactly once");
    });
}

int hy_pm_xplane_order(hy_pm_ctx* c, uint64_t n, const double* x, uint32_t* order, uint32_t* offsets) {
    return guarded([&] {
        if (!c || !x || !order || !offsets) fail(HY_EINVAL, "hy_pm_xplane_order: null argument");
        require_single_box(c, "hy_pm_xplane_order");
        set_device(c);
        c->keys_ready = c->rec_ready = false;
        require_particles(c, n);
        set_boxes(c, n);
        reset_status(c);
        bucketize(c, n, x, x, x, 1, c->xcounts, offsets, (uint64_t)c->N, order);
        hipLaunchKernelGGL(hy::k_canonicalize, dim3(c->N), dim3(hy::CANON_THREADS), 0, c->stream,
                           (uint64_t)c->N, offsets, order, c->scratch, c->scratch_cap, c->scratch_cursor, c->err);
        launch_check("k_canonicalize");
        const auto st = fetch_status(c);
        raise_errors(st.err);
        if (st.total != n) fail(HY_ELOGIC, "CIC x-plane binning does not cover the complete particle population");
    });
}

int hy_pm_profile(hy_pm_ctx* c, int enable, uint64_t max_calls) {
    return guarded([&] {
        if (!c) fail(HY_EINVAL, "null context");
        set_device(c);
        hip_check(hipStreamSynchronize(c->stream), "sync");
        for (hipEvent_t e : c->prof_ev) (void)hipEventDestroy(e);
        c->prof_ev.clear();
        c->prof_used.clear();
        c->prof = enable != 0 && max_calls > 0;
        c->prof_cap = c->prof ? max_calls : 0;
        c->prof_calls = 0;
        if (!c->prof) return;
        c->prof_ev.resize(max_calls * hy_pm_ctx::kHyMarks);
        c->prof_used.assign(c->prof_ev.size(), 0);
        for (auto& e : c->prof_ev) hip_check(hipEventCreate(&e), "hipEventCreate");
    });
}

int hy_pm_profile_read(hy_pm_ctx* c, double* stage_ms, uint64_t* stage_count, uint64_t* calls) {
    return guarded([&] {
        if (!c || !stage_ms || !stage_count) fail(HY_EINVAL, "null argument");
        set_device(c);
        hip_check(hipStreamSynchronize(c->stream), "sync");
        for (int k = 0; k < HY_STAGE_COUNT; ++k) { stage_ms[k] = 0.0; stage_count[k] = 0; }
        for (uint64_t call = 0; call < c->prof_calls; ++call) {
            const uint64_t base = call * hy_pm_ctx::kHyMarks;
            for (int k = 0; k < HY_STAGE_COUNT; ++k) {
                // stage k spans the latest recorded mark <= k to mark k+1
                if (!c->prof_used[base + k + 1]) continue;
                int s0 = k;
                while (s0 >= 0 && !c->prof_used[base + s0]) --s0;
                if (s0 < 0) continue;
                float ms = 0.0f;
                hip_check(hipEventElapsedTime(&ms, c->prof_ev[base + s0], c->prof_ev[base + k + 1]), "hipEventElapsedTime");
                stage_ms[k] += ms;
                stage_count[k] += 1;
            }
        }
        if (calls) *calls = c->prof_calls;
        c->prof_calls = 0;
        std::fill(c->prof_used.begin(), c->prof_used.end(), 0);
    });
}

int hy_pm_get_tuning(hy_pm_ctx* c, hy_fft_tuning* out) {
    return guarded([&] {
        if (!c || !out) fail(HY_EINVAL, "null argument");
        std::memset(out, 0, sizeof(*out));
        if (c->prec == HY_PREC_FP64) fill_tuning(c, c->fft64, out);
        else fill_tuning(c, c->fft32, out);
    });
}

int hy_pm_set_tuning(hy_pm_ctx* c, const hy_fft_tuning* in) {
    return guarded([&] {
        if (!c || !in) fail(HY_EINVAL, "null argument");
        const int nv = c->prec == HY_PREC_FP64 ? c->fft64.variants : c->fft32.variants;
        for (int p = 0; p < hy::FFT_PASSES; ++p) {
            const int v = (int)in->variant[p];
            if (v >= nv) fail(HY_EINVAL, "FFT launch variant not compiled for this mesh size");
            const int lds = c->prec == HY_PREC_FP64 ? c->fft64.pass[p][v].lds : c->fft32.pass[p][v].lds;
            if (lds > c->max_lds) fail(HY_EINVAL, "FFT launch variant exceeds the device LDS per workgroup");
        }
        if (in->plane_group > (uint32_t)c->N) fail(HY_EINVAL, "plane group exceeds the mesh size");
        for (int p = 0; p < hy::FFT_PASSES; ++p) c->fft_var[p] = (int)in->variant[p];
        c->plane_group = (int)in->plane_group;
    });
}

int hy_pm_autotune(hy_pm_ctx* c, uint32_t reps, hy_fft_tuning* out) {
    return guarded([&] {
        if (!c) fail(HY_EINVAL, "null context");
        if (reps == 0 || reps > 1000) fail(HY_EINVAL, "autotune repetitions must be 1..1000");
        set_device(c);
        hy_fft_tuning t{};
        if (c->prec == HY_PREC_FP64) autotune(c, c->fft64, reps, &t);
        else autotune(c, c->fft32, reps, &t);
        if (out) *out = t;
    });
}

int hy_pm_set_gather_rows(hy_pm_ctx* c, uint32_t rows) {
    return guarded([&] {
        if (!c) fail(HY_EINVAL, "null context");
        if (rows != HY_GATHER_ROWS_DEFAULT && rows != 0 && (rows & (rows - 1)) != 0)
            fail(HY_EINVAL, "gather rows per tile must be 0, a power of two or HY_GATHER_ROWS_DEFAULT");
        c->gather_rows = rows;
    });
}

int hy_pm_autotune_gather(hy_pm_ctx* c, uint64_t n, const double* x, const double* y, const double* z,
                          double uniform, uint32_t reps, uint32_t* rows_out, float* ms_out) {
    return guarded([&] {
        if (!c || !x || !y || !z) fail(HY_EINVAL, "hy_pm_autotune_gather: null argument");
        if (reps == 0 || reps > 1000) fail(HY_EINVAL, "autotune repetitions must be 1..1000");
        require_single_box(c, "hy_pm_autotune_gather");
        require_pure_pm(c, "hy_pm_autotune_gather");
        set_device(c);
        require_particles(c, n);
        set_boxes(c, n);
        c->keys_ready = c->rec_ready = false;
        const double M = input_mass(c, n, nullptr, uniform, false);
        reset_status(c);
        run_deposit(c, n, x, y, z, nullptr, uniform, M, true);  // the potential of these positions
        run_fft(c);
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        auto gather = [&]() {
            auto fill = [&](auto& a) {
                a.g = c->geom; a.n = n;
                a.x_in = x; a.y_in = y; a.z_in = z;
                a.err = c->err;
            };
            if (c->prec == HY_PREC_FP64) {
                hy::GatherArgs<double, double> a{};
                fill(a);
                a.phi = static_cast<const double*>(c->R);
                a.g_x = (double*)c->gx; a.g_y = (double*)c->gy; a.g_z = (double*)c->gz;
                launch_gather(c, a);
            } else {
                hy::GatherArgs<float, float> a{};
                fill(a);
                a.phi = static_cast<const float*>(c->R);
                a.g_x = (float*)c->gx; a.g_y = (float*)c->gy; a.g_z = (float*)c->gz;
                launch_gather(c, a);
            }
        };
        static const uint32_t cand[HY_GATHER_TUNE_CANDIDATES] = {0, 16, 64, 256, 1024};
        float best_ms = 0.0f;
        uint32_t best = 0;
        std::string err;
        for (int i = 0; i < HY_GATHER_TUNE_CANDIDATES; ++i) {
            c->gather_rows = cand[i];
            const hy::LatticeTiling t = lattice_tiling(n, 256, cand[i]);
            float ms = 0.0f;
            if (cand[i] == 0 || t.S != 0) {
                gather();  // warm
                hip_check(hipEventRecord(e0, c->stream), "hipEventRecord");
                for (uint32_t r = 0; r < reps; ++r) gather();
                hip_check(hipEventRecord(e1, c->stream), "hipEventRecord");
                hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
                hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
                ms /= (float)reps;
                if (best_ms == 0.0f || ms < best_ms) {
                    best_ms = ms;
                    best = cand[i];
                }
            }
            if (ms_out) ms_out[i] = ms;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        c->gather_rows = best;
        if (rows_out) *rows_out = best;
        const auto st = fetch_status(c);
        raise_errors(st.err);
    });
}

// ---------------------------------------------------------------------------
// Adjoint of the fused step and of the density
// ---------------------------------------------------------------------------
}  // extern "C"

namespace {

void ensure_adjoint_workspace(hy_pm_ctx* c) {
    if (c->adj_phi) return;
    const size_t bytes = (size_t)c->N * c->N * c->N * c->real_bytes;
    c->adj_phi = c->alloc(bytes);
    for (auto& a : c->adj_a) a = c->alloc(bytes);
    c->adj_max = (double*)c->alloc(1024 * sizeof(double));
}

// max |q| over n values (synchronises).
double abs_max(hy_pm_ctx* c, uint64_t n, const double* q) {
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(blocks_for(n, 256), 1), 1024);
    hipLaunchKernelGGL(hy::k_absmax, dim3(blocks), dim3(256), 0, c->stream, n, q, c->adj_max);
    launch_check("k_absmax");
    std::vector<double> h(blocks);
    hip_check(hipMemcpyAsync(h.data(), c->adj_max, blocks * sizeof(double), hipMemcpyDeviceToHost, c->stream),
              "d2h");
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    double m = 0.0;
    for (double v : h) m = std::max(m, v);
    return m;
}

// Raw CIC deposit of signed per-particle values q at positions whose buckets
// (and, for binary64, particle order) the last run_deposit left in place.
void deposit_values(hy_pm_ctx* c, uint64_t n, const double* x, const double* y, const double* z,
                    const double* q, double uniform, double Min) {
    const double unit = abs_max(c, n, q);
    if (unit == 0.0 || !std::isfinite(unit)) {
        if (!std::isfinite(unit)) fail(HY_EOVERFLOW, "adjoint values are not finite");
        fill_u32(c, c->R, (uint64_t)c->N * c->N * c->N * (c->real_bytes / 4));
        return;
    }
    auto plan = plan_deposit(c, n, x, y, z, q, uniform, Min);
    plan.P.raw = 1;
    c->mass_unit = unit;
    if (c->prec == HY_PREC_FP32_MIXED)
        hipLaunchKernelGGL(hy::k_scatter_values<float>, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, q,
                           1.0 / unit, c->keys, c->ranks, c->offsets, c->rec_mass);
    else
        hipLaunchKernelGGL(hy::k_scatter_values<double>, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, q,
                           1.0, c->keys, c->ranks, c->offsets, c->rec_mass64);
    launch_check("k_scatter_values");
    launch_deposit(c, plan.P, false);
}

}  // namespace

extern "C" {

int hy_pm_kdk_vjp(hy_pm_ctx* c, uint64_t n, const double* x, const double* y, const double* z,
                  double uniform, double* lx, double* ly, double* lz, double* lpx, double* lpy, double* lpz,
                  double k_prev, double k_next, double d_next, uint32_t flags) {
    return guarded([&] {
        if (!c || !x || !y || !z || !lx || !ly || !lz || !lpx || !lpy || !lpz)
            fail(HY_EINVAL, "hy_pm_kdk_vjp: null argument");
        require_single_box(c, "hy_pm_kdk_vjp");
        require_pure_pm(c, "hy_pm_kdk_vjp");
        if (n == 0) fail(HY_ERUNTIME, "PMSolver exact input mass identity must be finite and positive");
        const bool has_prev = (flags & HY_HAS_PREV) != 0, has_next = (flags & HY_HAS_NEXT) != 0;
        set_device(c);
        require_particles(c, n);
        set_boxes(c, n);
        c->keys_ready = c->rec_ready = false;
        const double M = input_mass(c, n, nullptr, uniform, false);
        ensure_adjoint_workspace(c);
        reset_status(c);
        // momentum adjoint entering the kick: u = lp + d_next lx
        hipLaunchKernelGGL(hy::k_adj_kick, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, lx, ly, lz, lpx,
                           lpy, lpz, d_next, has_next ? 1 : 0);
        launch_check("k_adj_kick");
        const double ksum = (has_prev ? k_prev : 0.0) + (has_next ? k_next : 0.0);
        if (ksum == 0.0) {
            hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
            return;
        }
        // forward potential at x
        run_deposit(c, n, x, y, z, nullptr, uniform, M, true);
        run_fft(c);
        const size_t bytes = (size_t)c->N * c->N * c->N * c->real_bytes;
        hip_check(hipMemcpyAsync(c->adj_phi, c->R, bytes, hipMemcpyDeviceToDevice, c->stream), "copy phi");
        // A_a = deposit of u_a, then div A into R, then chi = G div A
        const double* u[3] = {lpx, lpy, lpz};
        for (int d = 0; d < 3; ++d) {
            deposit_values(c, n, x, y, z, u[d], uniform, M);
            hip_check(hipMemcpyAsync(c->adj_a[d], c->R, bytes, hipMemcpyDeviceToDevice, c->stream), "copy A");
        }
        const uint64_t cells = (uint64_t)c->N * c->N * c->N;
        const double inv_2dx = 1.0 / (2.0 * c->geom.dx);
        if (c->prec == HY_PREC_FP64)
            hipLaunchKernelGGL(hy::k_div_stencil<double>, dim3(blocks_for(cells, 256)), dim3(256), 0, c->stream,
                               (const double*)c->adj_a[0], (const double*)c->adj_a[1], (const double*)c->adj_a[2],
                               (double*)c->R, c->N, inv_2dx);
        else
            hipLaunchKernelGGL(hy::k_div_stencil<float>, dim3(blocks_for(cells, 256)), dim3(256), 0, c->stream,
                               (const float*)c->adj_a[0], (const float*)c->adj_a[1], (const float*)c->adj_a[2],
                               (float*)c->R, c->N, inv_2dx);
        launch_check("k_div_stencil");
        run_fft(c);
        const double dx3 = (c->geom.dx * c->geom.dx) * c->geom.dx;
        auto gather = [&](auto tag) {
            using T = decltype(tag);
            hy::AdjGatherArgs<T> a{};
            a.g = c->geom; a.n = n; a.x = x; a.y = y; a.z = z;
            a.ux = lpx; a.uy = lpy; a.uz = lpz;
            a.phi = static_cast<const T*>(c->adj_phi);
            a.chi = static_cast<const T*>(c->R);
            a.coef_a = ksum;
            a.coef_b = ksum * uniform / dx3;
            a.lx = lx; a.ly = ly; a.lz = lz;
            a.err = c->err;
            hipLaunchKernelGGL(hy::k_adj_gather<T>, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, a);
            launch_check("k_adj_gather");
        };
        if (c->prec == HY_PREC_FP64) gather(double{});
        else gather(float{});
        const auto st = fetch_status(c);
        raise_errors(st.err);
    });
}

int hy_pm_density_vjp(hy_pm_ctx* c, uint64_t n, const double* x, const double* y, const double* z,
                      double uniform, const void* u_mesh, double coef, double* lx, double* ly, double* lz) {
    return guarded([&] {
        if (!c || !x || !y || !z || !u_mesh || !lx || !ly || !lz) fail(HY_EINVAL, "hy_pm_density_vjp: null argument");
        require_single_box(c, "hy_pm_density_vjp");
        set_device(c);
        require_particles(c, n);
        set_boxes(c, n);
        if (!std::isfinite(uniform) || uniform <= 0.0) fail(HY_EINVAL, "PMSolver uniform mass must be finite and positive");
        reset_status(c);
        auto gather = [&](auto tag) {
            using T = decltype(tag);
            hy::AdjGatherArgs<T> a{};
            a.g = c->geom; a.n = n; a.x = x; a.y = y; a.z = z;
            a.chi = static_cast<const T*>(u_mesh);
            a.coef_b = coef * uniform;
            a.lx = lx; a.ly = ly; a.lz = lz;
            a.err = c->err;
            hipLaunchKernelGGL(hy::k_adj_gather<T>, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, a);
            launch_check("k_adj_gather");
        };
        if (c->prec == HY_PREC_FP64) gather(double{});
        else gather(float{});
        const auto st = fetch_status(c);
        raise_errors(st.err);
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Matter power spectrum
// ---------------------------------------------------------------------------
namespace {

// r2c transform of the mesh in R into dst ([kz][ky][kx] after the forward x
// pass), for modes with |m_y|, m_z <= kmode only (kmode = N/2: all of them).
// The z pass stores the planes kz <= kmode (two-level passes), the y pass runs
// over those planes, the x pass over their rows with |m_y| <= kmode (whole
// workgroups of rows outside exit at once); every mode formed gets the same
// bits as in the full transform.
template <typename T>
void launch_forward(hy_pm_ctx* c, const hy::FftKernels<T>& K, void* dst, int kmode) {
    hy::FftArgs<T> a = fft_args<T>(c);
    a.B = static_cast<hy::cplx<T>*>(dst);
    const int NH = c->N / 2 + 1;
    const int planes = std::min(NH, kmode + 1);
    a.mode_limit = kmode < c->N / 2 ? kmode : 0;
    if (a.mode_limit) {
        const hy::FftPass<T>& Z = K.z_fwd;
        const uint64_t N = (uint64_t)c->N;
        a.kz0 = 0;
        hipLaunchKernelGGL(Z.fn, dim3((uint32_t)(N * (N / Z.rows)), 1), dim3(Z.threads), Z.lds, c->stream, a);
        launch_check("k_fft_z_r2c (pruned)");
    } else {
        launch_pass(c, K, a, hy::FFT_Z_R2C, c->fft_var[hy::FFT_Z_R2C], 0, NH);
    }
    launch_pass(c, K, a, hy::FFT_Y_FWD, c->fft_var[hy::FFT_Y_FWD], 0, planes);
    const hy::FftPass<T>& P = K.x_fwd;
    a.kz0 = 0;
    const uint64_t blocks = (uint64_t)planes * (uint64_t)c->N / (uint64_t)P.rows;
    hipLaunchKernelGGL(P.fn, dim3((uint32_t)blocks, 1), dim3(P.threads), P.lds, c->stream, a);
    launch_check("k_fft_x_fwd");
}

void run_forward(hy_pm_ctx* c, void* dst, int kmode) {
    if (c->prec == HY_PREC_FP64) launch_forward<double>(c, c->fft64, dst, kmode);
    else launch_forward<float>(c, c->fft32, dst, kmode);
}

// Workgroups of the shell-binning pass: a function of the mesh and the shell
// count only, so that the fixed-order sums do not depend on the device.
struct PkLaunch {
    int waves, groups;
    size_t lds;
    uint64_t row_step;
};

PkLaunch pk_launch(int N, int nb) {
    PkLaunch l{};
    const size_t per_wave = (size_t)nb * hy::PK_Q * sizeof(double);
    const size_t budget = 64 * 1024 - ((size_t)nb + 1) * sizeof(double);
    l.waves = (int)std::max<size_t>(1, std::min<size_t>(hy::PK_BLOCK / 64, budget / per_wave));
    const uint64_t rows = (uint64_t)(N / 2 + 1) * (uint64_t)N;
    uint64_t g = (rows + l.waves - 1) / l.waves;
    g = std::min<uint64_t>(g, 2048);
    g = std::min<uint64_t>(g, std::max<uint64_t>(1, (1u << 20) / (uint64_t)nb));
    l.groups = (int)g;
    l.lds = (size_t)l.waves * per_wave + ((size_t)nb + 1) * sizeof(double);
    // the first step coprime to rows from rows / golden ratio up: k -> k step
    // mod rows permutes the rows, and each contiguous block of k is spread
    // evenly over the half spectrum
    uint64_t step = std::max<uint64_t>(1, (uint64_t)((double)rows * 0.6180339887498949));
    while (std::gcd(step, rows) != 1) ++step;
    l.row_step = step;
    return l;
}

}  // namespace

extern "C" {

int hy_pm_power_spectrum(hy_pm_ctx* c, uint64_t n, const double* x, const double* y, const double* z,
                         const double* masses, double uniform, const hy_pk_options* o, hy_pk_bin* bins,
                         uint32_t flags) {
    return guarded([&] {
        if (!c || !o || !bins) fail(HY_EINVAL, "hy_pm_power_spectrum: null argument");
        // option checks in the order of src/analysis/matter_power_spectrum.cpp:60-70
        if (o->num_bins < 1) fail(HY_EINVAL, "MatterPowerSpectrum num_bins must be positive");
        if (o->num_bins > HY_PK_MAX_BINS) fail(HY_EINVAL, "MatterPowerSpectrum num_bins exceeds HY_PK_MAX_BINS");
        if ((o->subtract_shot_noise != 0 && o->subtract_shot_noise != 1) || (o->interlaced != 0 && o->interlaced != 1))
            fail(HY_ELOGIC, "Required scalar coordinate was not specified");
        const double frac = o->max_k_fraction_nyquist;
        if (!std::isfinite(frac) || frac <= 0.0 || frac > 1.0)
            fail(HY_EINVAL, "MatterPowerSpectrum max_k_fraction_nyquist must be in (0,1]");
        const bool interlaced = o->interlaced == 1, subtract = o->subtract_shot_noise == 1;
        if (interlaced && subtract)
            fail(HY_EINVAL,
                 "Interlaced shot-noise subtraction is disabled until the estimator uses an interlacing-matched "
                 "alias model");
        if (n == 0) fail(HY_EINVAL, "FourierDensityBuilder requires owned particles");
        if (!x || !y || !z) fail(HY_EINVAL, "hy_pm_power_spectrum: null argument");
        require_single_box(c, "hy_pm_power_spectrum");
        set_device(c);
        c->keys_ready = c->rec_ready = false;
        require_particles(c, n);
        set_boxes(c, n);
        const bool host = (flags & HY_HOST_POINTERS) != 0;
        const int N = c->N, nb = o->num_bins;
        const double L = c->geom.L, dx = c->geom.dx;

        // mass moments (src/analysis/fourier_density_field.cpp:160-225)
        double msf;
        if (!masses) {
            if (!std::isfinite(uniform) || uniform <= 0.0)
                fail(HY_EINVAL, "FourierDensityBuilder uniform mass must be finite and positive");
            msf = hy::host::mass_square_fraction_uniform(n);
        } else {
            std::vector<double> h(n);
            if (host) std::memcpy(h.data(), masses, n * sizeof(double));
            else hip_check(hipMemcpyAsync(h.data(), masses, n * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                           "mass copy");
            hip_check(hipStreamSynchronize(c->stream), "sync");
            msf = hy::host::mass_square_fraction(h.data(), n, nullptr);
        }
        const double M = input_mass(c, n, masses, uniform, host);

        const double *px = x, *py = y, *pz = z, *pm = masses;
        if (host) {
            ensure_host_staging(c);
            double* s = c->hpos;
            hip_check(hipMemcpyAsync(s, x, n * 8, hipMemcpyHostToDevice, c->stream), "h2d");
            hip_check(hipMemcpyAsync(s + c->cap, y, n * 8, hipMemcpyHostToDevice, c->stream), "h2d");
            hip_check(hipMemcpyAsync(s + 2 * c->cap, z, n * 8, hipMemcpyHostToDevice, c->stream), "h2d");
            px = s; py = s + c->cap; pz = s + 2 * c->cap;
            if (masses) {
                hip_check(hipMemcpyAsync(s + 3 * c->cap, masses, n * 8, hipMemcpyHostToDevice, c->stream), "h2d");
                pm = s + 3 * c->cap;
            }
        }
        const uint64_t ccount = (uint64_t)N * (uint64_t)(N / 2 + 1) * (uint64_t)N;
        if (interlaced && !c->pk_aux) c->pk_aux = c->alloc(ccount * 2 * c->real_bytes);
        auto deposit = [&](double shift) {
            reset_status(c);
            try {
                run_deposit(c, n, px, py, pz, pm, uniform, M, true, false, false, false, shift);
            } catch (const HyError& e) {
                if (e.code == HY_EINVAL && e.what == kMsgPositions)
                    fail(HY_EINVAL, "FourierDensityBuilder particle positions must be finite");
                throw;
            }
        };
        // contrast transforms (src/analysis/fourier_density_field.cpp:230-336).
        // The shells read modes with |k| <= frac k_Nyq only, i.e. |m| <= frac N/2
        // per axis; kmode bounds that with one mode to spare, so the rows the
        // transform leaves out are never read.
        const int kmode = std::min(N / 2, (int)(frac * (double)(N / 2) * (1.0 + 1e-9)) + 1);
        deposit(0.0);
        run_forward(c, c->B, kmode);
        const double shift = 0.5 * dx;
        if (interlaced) {
            // the partner positions wrap(x - dx/2) are formed inside the routing
            // and deposit kernels (src/analysis/fourier_density_field.cpp:300-311)
            deposit(shift);
            run_forward(c, c->pk_aux, kmode);
        }

        // shells (src/analysis/matter_power_spectrum.cpp:85-125)
        const double knyq = std::numbers::pi * static_cast<double>(N) / L;
        const double kmax = frac * knyq;
        const double kf = 2.0 * std::numbers::pi / L;
        if (!(kmax > kf)) fail(HY_EINVAL, "MatterPowerSpectrum requested range contains no nonzero shell");
        if (!(kf >= 0x1p-450) || !(kmax <= 0x1p450))
            fail(HY_EINVAL, "MatterPowerSpectrum wavenumbers outside [2^-450, 2^450] are not supported");
        const std::vector<double> edges = hy::host::log_shell_edges(kf, kmax, nb);
        const double shot = hy::host::volume_times(L, msf);
        std::vector<double> tab(3 * (size_t)N + (size_t)nb + 1);
        for (int i = 0; i < N; ++i) {
            const int m = i <= N / 2 ? i : i - N;
            const double k = 2.0 * std::numbers::pi * static_cast<double>(m) / L;
            const double arg = 0.5 * k * dx;
            const double s2 = std::sin(arg) * std::sin(arg);
            tab[i] = k;
            tab[N + i] = static_cast<double>(hy::host::cic_window_axis(k, dx));
            tab[2 * N + i] = 1.0 - 2.0 / 3.0 * s2;
        }
        std::copy(edges.begin(), edges.end(), tab.begin() + 3 * N);
        if (!c->pk_tab)
            c->pk_tab = static_cast<double*>(c->alloc((3 * (size_t)N + HY_PK_MAX_BINS + 1) * sizeof(double)));
        hip_check(hipMemcpyAsync(c->pk_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, c->stream),
                  "h2d tables");
        const PkLaunch l = pk_launch(N, nb);
        const uint64_t slab = (uint64_t)nb * hy::PK_Q;
        const uint64_t need = ((uint64_t)l.groups + 1) * slab;
        if (need > c->pk_part_cap) {
            void* p = nullptr;
            hip_check(hipMalloc(&p, need * sizeof(double)), "hipMalloc shell partials");
            c->allocations.push_back(p);
            c->device_bytes += need * sizeof(double);
            c->pk_part = static_cast<double*>(p);
            c->pk_part_cap = need;
        }
        hy::PkArgs a{};
        a.m1 = c->B;
        a.m2 = interlaced ? c->pk_aux : nullptr;
        a.kc = c->pk_tab;
        a.win = c->pk_tab + N;
        a.ali = c->pk_tab + 2 * N;
        a.edges = c->pk_tab + 3 * N;
        a.partials = c->pk_part;
        a.err = c->err;
        a.N = N;
        a.nb = nb;
        a.waves = l.waves;
        a.subtract = subtract ? 1 : 0;
        a.kmax = kmax;
        a.kf = kf;
        a.perp_skip = kmax * (1.0 + 0x1p-40);
        {
            // delta_k = FFT(source) L^3 / (M N^3): the density source (m - M/N^3)/dx^3
            // is the contrast times M / L^3, and the reference scales modes by 1/N^3
            const long double cells = (long double)N * (long double)N * (long double)N;
            const long double L3 = (long double)L * (long double)L * (long double)L;
            a.amp = static_cast<double>(L3 / ((long double)M * cells));
            a.volume = static_cast<double>(L3);
        }
        a.shot = shot;
        a.shift = shift;
        a.row_step = l.row_step;
        double* sums = c->pk_part + (uint64_t)l.groups * slab;
        if (c->prec == HY_PREC_FP64)
            hipLaunchKernelGGL(hy::k_pk_bin<double>, dim3(l.groups), dim3(hy::PK_BLOCK), l.lds, c->stream, a);
        else
            hipLaunchKernelGGL(hy::k_pk_bin<float>, dim3(l.groups), dim3(hy::PK_BLOCK), l.lds, c->stream, a);
        launch_check("k_pk_bin");
        hipLaunchKernelGGL(hy::k_pk_reduce, dim3((uint32_t)slab), dim3(256), 0, c->stream,
                           (const double*)c->pk_part, l.groups, (int)slab, sums);
        launch_check("k_pk_reduce");
        std::vector<double> h(slab);
        hip_check(hipMemcpyAsync(h.data(), sums, slab * sizeof(double), hipMemcpyDeviceToHost, c->stream), "d2h");
        const auto st = fetch_status(c);
        raise_errors(st.err);
        // shell means (src/analysis/matter_power_spectrum.cpp:219-240)
        for (int b = 0; b < nb; ++b) {
            hy_pk_bin& r = bins[b];
            r = hy_pk_bin{};
            r.k_low = edges[b];
            r.k_high = edges[b + 1];
            r.shot_noise_reference = shot;
            const double* q = h.data() + (size_t)b * hy::PK_Q;
            const uint64_t count = (uint64_t)q[4];
            if (count == 0) continue;
            const double cnt = static_cast<double>(count);
            r.k_mean = q[0] / cnt;
            r.p_raw = q[1] / cnt;
            r.p_corrected = q[2] / cnt;
            r.mean_cic_power_deconvolution_factor = q[3] / cnt;
            r.mode_count = count;
        }
    });
}

// ---------------------------------------------------------------------------
// Ensemble step graphs
// ---------------------------------------------------------------------------
}  // extern "C"

struct hy_ensemble {
    std::vector<hy_ensemble_member> members;
    hipStream_t stream = nullptr;
    int device = 0;
    double* dfac = nullptr;  // k_prev, k_next, d_next for the next replay
    hipGraphExec_t exec[4] = {nullptr, nullptr, nullptr, nullptr};
    bool fresh = false;  // next step re-keys the members eagerly (positions were replaced)
    // capture only: each member's step is recorded on its own stream, forked
    // from and joined back to `stream`, so the graph runs the members'
    // independent kernel chains side by side
    std::vector<hipStream_t> branch;
    std::vector<hipEvent_t> joined;
    hipEvent_t forked = nullptr;
    ~hy_ensemble() {
        for (hipGraphExec_t g : exec)
            if (g) (void)hipGraphExecDestroy(g);
        if (dfac) (void)hipFree(dfac);
        for (hipEvent_t ev : joined) (void)hipEventDestroy(ev);
        if (forked) (void)hipEventDestroy(forked);
        for (hipStream_t s : branch) (void)hipStreamDestroy(s);
    }
};

extern "C" {

int hy_ensemble_create(const hy_ensemble_member* members, uint32_t count, hy_ensemble** out) {
    return guarded([&] {
        if (!members || !out || count == 0) fail(HY_EINVAL, "hy_ensemble_create: null argument or no members");
        *out = nullptr;
        auto e = std::make_unique<hy_ensemble>();
        for (uint32_t i = 0; i < count; ++i) {
            const hy_ensemble_member& m = members[i];
            if (!m.ctx || !m.x || !m.y || !m.z || !m.px || !m.py || !m.pz || m.n == 0)
                fail(HY_EINVAL, "hy_ensemble_create: incomplete member");
            require_pure_pm(m.ctx, "hy_ensemble_create");
            if (i == 0) {
                e->stream = m.ctx->stream;
                e->device = m.ctx->desc.device;
            } else if (m.ctx->stream != e->stream || m.ctx->desc.device != e->device) {
                fail(HY_EINVAL, "ensemble members must share one stream and device");
            }
            e->members.push_back(m);
        }
        hip_check(hipSetDevice(e->device), "hipSetDevice");
        hip_check(hipMalloc(&e->dfac, 4 * sizeof(double)), "hipMalloc");
        hip_check(hipMemset(e->dfac, 0, 4 * sizeof(double)), "hipMemset");
        hip_check(hipEventCreateWithFlags(&e->forked, hipEventDisableTiming), "hipEventCreate");
        for (uint32_t i = 0; i < count; ++i) {
            hipStream_t s = nullptr;
            hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");  // default priority
            e->branch.push_back(s);
            hipEvent_t ev = nullptr;
            hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
            e->joined.push_back(ev);
        }
        *out = e.release();
    });
}

int hy_ensemble_step(hy_ensemble* e, double k_prev, double k_next, double d_next, uint32_t flags) {
    return guarded([&] {
        if (!e) fail(HY_EINVAL, "hy_ensemble_step: null ensemble");
        const bool has_prev = (flags & HY_HAS_PREV) != 0, has_next = (flags & HY_HAS_NEXT) != 0;
        if (!has_prev && !has_next) fail(HY_EINVAL, "hy_ensemble_step needs HY_HAS_PREV and/or HY_HAS_NEXT");
        if ((has_prev && !std::isfinite(k_prev)) || (has_next && (!std::isfinite(k_next) || !std::isfinite(d_next))))
            fail(HY_EOVERFLOW, "Leapfrog drift/kick factors must be finite");
        hip_check(hipSetDevice(e->device), "hipSetDevice");
        // A recorded step replays the routing its capture saw: it must start
        // from what a replay leaves, the deposit records (binary32, uniform
        // mass) or bucket keys of the members' current positions.  Until
        // every member holds them (and for a first step, which has nothing to
        // carry) the step runs eagerly.
        bool steady = has_prev && !e->fresh;
        for (const hy_ensemble_member& m : e->members) {
            const hy_pm_ctx* c = m.ctx;
            const bool recs = c->rec_ready && c->rec_n == m.n && c->rec_x == m.x && c->rec_y == m.y && c->rec_z == m.z;
            const bool keys = c->keys_ready && c->keys_n == m.n && c->keys_x == m.x && c->keys_y == m.y &&
                              c->keys_z == m.z;
            steady = steady && (c->rec_layout ? recs : keys);
        }
        if (!steady) {
            // positions replaced outside the graphs (fresh: route them anew),
            // or the carried routing is not yet the replays' own
            const uint32_t kflags = (has_prev ? HY_HAS_PREV : 0u) | (has_next ? HY_HAS_NEXT : 0u) | HY_ASYNC |
                                    (e->fresh ? 0u : HY_REUSE_KEYS);
            for (const hy_ensemble_member& m : e->members) {
                const int rc = hy_pm_kdk(m.ctx, m.n, m.x, m.y, m.z, m.px, m.py, m.pz, m.x, m.y, m.z, m.px, m.py,
                                         m.pz, nullptr, m.uniform_mass, k_prev, k_next, d_next, nullptr, nullptr,
                                         nullptr, kflags);
                if (rc != HY_OK) fail(rc, "ensemble eager step: " + g_last_error);
            }
            e->fresh = false;
            return;
        }
        const int idx = (has_prev ? 1 : 0) | (has_next ? 2 : 0);
        hipLaunchKernelGGL(hy::k_set_factors, dim3(1), dim3(64), 0, e->stream, e->dfac, k_prev, k_next, d_next);
        launch_check("k_set_factors");
        if (!e->exec[idx]) {
            const uint32_t kflags = (has_prev ? HY_HAS_PREV : 0u) | (has_next ? HY_HAS_NEXT : 0u) | HY_ASYNC |
                                    HY_REUSE_KEYS;
            std::vector<bool> prof(e->members.size());
            for (size_t i = 0; i < e->members.size(); ++i) {
                prof[i] = e->members[i].ctx->prof;
                e->members[i].ctx->prof = false;
                e->members[i].ctx->dfac = e->dfac;
            }
            hip_check(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            int rc = HY_OK;
            std::string msg;
            hipError_t eb = hipEventRecord(e->forked, e->stream);
            for (size_t i = 0; i < e->members.size() && eb == hipSuccess; ++i) {
                const hy_ensemble_member& m = e->members[i];
                eb = hipStreamWaitEvent(e->branch[i], e->forked, 0);
                if (eb != hipSuccess) break;
                m.ctx->stream = e->branch[i];
                rc = hy_pm_kdk(m.ctx, m.n, m.x, m.y, m.z, m.px, m.py, m.pz, m.x, m.y, m.z, m.px, m.py, m.pz,
                               nullptr, m.uniform_mass, 0.0, 0.0, 0.0, nullptr, nullptr, nullptr, kflags);
                m.ctx->stream = e->stream;
                if (rc != HY_OK) {
                    msg = g_last_error;
                    break;
                }
                eb = hipEventRecord(e->joined[i], e->branch[i]);
                if (eb == hipSuccess) eb = hipStreamWaitEvent(e->stream, e->joined[i], 0);
            }
            hipGraph_t graph = nullptr;
            const hipError_t ec = hipStreamEndCapture(e->stream, &graph);
            if (rc == HY_OK && eb != hipSuccess) {
                rc = HY_ERUNTIME;
                msg = std::string("ensemble capture fork/join: ") + hipGetErrorString(eb);
            }
            for (size_t i = 0; i < e->members.size(); ++i) {
                e->members[i].ctx->prof = prof[i];
                e->members[i].ctx->dfac = nullptr;
            }
            if (rc != HY_OK) {
                if (graph) (void)hipGraphDestroy(graph);
                fail(rc, "ensemble capture: " + msg);
            }
            hip_check(ec, "hipStreamEndCapture");
            const hipError_t ei = hipGraphInstantiate(&e->exec[idx], graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            hip_check(ei, "hipGraphInstantiate");
        }
        hip_check(hipGraphLaunch(e->exec[idx], e->stream), "hipGraphLaunch");
    });
}

int hy_ensemble_reset(hy_ensemble* e) {
    return guarded([&] {
        if (!e) fail(HY_EINVAL, "hy_ensemble_reset: null ensemble");
        hip_check(hipSetDevice(e->device), "hipSetDevice");
        hip_check(hipStreamSynchronize(e->stream), "hipStreamSynchronize");
        for (const hy_ensemble_member& m : e->members) {
            m.ctx->keys_ready = m.ctx->rec_ready = false;
            reset_status(m.ctx);  // errors of the replaced run are discarded
        }
        hip_check(hipStreamSynchronize(e->stream), "hipStreamSynchronize");
        e->fresh = true;
    });
}

int hy_ensemble_check(hy_ensemble* e) {
    if (!e) {
        g_last_error = "hy_ensemble_check: null ensemble";
        return HY_EINVAL;
    }
    // every member is checked (each report clears that member's error word);
    // the first error is returned
    int first = HY_OK;
    std::string msg;
    for (const hy_ensemble_member& m : e->members) {
        const int rc = hy_pm_check(m.ctx);
        if (rc != HY_OK && first == HY_OK) {
            first = rc;
            msg = g_last_error;
        }
    }
    g_last_error = msg;
    return first;
}

int hy_ensemble_destroy(hy_ensemble* e) {
    return guarded([&] {
        if (!e) return;
        (void)hipSetDevice(e->device);
        if (e->stream) (void)hipStreamSynchronize(e->stream);
        delete e;
    });
}

// ---------------------------------------------------------------------------
// Host scalars
// ---------------------------------------------------------------------------
double hy_hubble(double a, double om, double ol) {
    double r = std::numeric_limits<double>::quiet_NaN();
    guarded([&] { r = hy::host::hubble(a, om, ol); });
    return r;
}

int hy_drift_factor(double a1, double a2, double om, double ol, double* out) {
    return guarded([&] {
        if (!out) fail(HY_EINVAL, "null output");
        *out = hy::host::drift_factor(a1, a2, om, ol);
    });
}

int hy_kick_factor(double a1, double a2, double om, double ol, double* out) {
    return guarded([&] {
        if (!out) fail(HY_EINVAL, "null output");
        *out = hy::host::kick_factor(a1, a2, om, ol);
    });
}

double hy_geometric_mean(double a, double b) {
    double r = std::numeric_limits<double>::quiet_NaN();
    guarded([&] { r = hy::host::geometric_mean(a, b); });
    return r;
}

int hy_time_boundaries(double a_start, double a_final, double dlna, const double* stops, uint64_t n_stops,
                       double* out, uint64_t capacity, uint64_t* count) {
    return guarded([&] {
        if (!out || !count || (n_stops && !stops)) fail(HY_EINVAL, "null argument");
        const auto b = hy::host::time_boundaries(a_start, a_final, dlna,
                                                 std::vector<double>(stops, stops + n_stops));
        if (b.size() > capacity) fail(HY_EOVERFLOW, "boundary capacity too small");
        std::memcpy(out, b.data(), b.size() * sizeof(double));
        *count = b.size();
    });
}

double hy_exact_sum(const double* v, uint64_t n) { return hy::host::exact_sum(v, n); }

// ---------------------------------------------------------------------------
// Initial-condition scalars
// ---------------------------------------------------------------------------
namespace {
hy::host::Cosmology cosmology_of(const hy_cosmology* c) {
    if (!c) fail(HY_EINVAL, "null cosmology");
    return {c->h, c->omega_m, c->omega_b, c->omega_lambda, c->sigma8, c->n_s};
}
}  // namespace

int hy_growth(const hy_cosmology* c, uint64_t n, const double* a, double* D1, double* f1, double* D2, double* f2) {
    return guarded([&] {
        if (n && !a) fail(HY_EINVAL, "null scale factors");
        hy::host::growth(cosmology_of(c), n, a, D1, f1, D2, f2);
    });
}

int hy_linear_power(const hy_cosmology* c, const char* table, double table_redshift, double start_redshift,
                    uint64_t n, const double* k, double* out) {
    return guarded([&] {
        if (!table || (n && (!k || !out))) fail(HY_EINVAL, "null argument");
        hy::host::linear_power(cosmology_of(c), table, table_redshift, start_redshift, n, k, out);
    });
}

int hy_ic_mode_variance(const hy_cosmology* c, const char* table, double table_redshift, double start_redshift,
                        uint64_t mesh_size, double box_size, uint64_t max_mode, double* variance_by_m2) {
    return guarded([&] {
        if (!table || !variance_by_m2) fail(HY_EINVAL, "null argument");
        hy::host::mode_variance(cosmology_of(c), table, table_redshift, start_redshift, mesh_size, box_size,
                                max_mode, variance_by_m2);
    });
}

int hy_cubic_spline(uint64_t n, const double* x, const double* y, uint64_t m, const double* q, double* out) {
    return guarded([&] {
        if (!x || !y || (m && (!q || !out))) fail(HY_EINVAL, "null argument");
        hy::host::cubic_spline(n, x, y, m, q, out);
    });
}
double hy_gravitational_constant(void) { return hy::host::gravitational_constant(); }
double hy_rho_crit0(void) { return hy::host::rho_crit0(); }

}  // extern "C"
