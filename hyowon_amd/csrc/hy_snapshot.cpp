// Native HYOWON snapshot files (HDF5): writer and reader of the reference's
// single-file schema, so particle states computed here can be handed to the
// reference's tools and back.
//
//   schema names      include/cosmo_nbody/io/output_schema.hpp:22-140
//   writer            src/io/snapshot_io_write_impl.inc:1-286
//   reader checks     src/io/snapshot_io_read_impl.inc:1-260
//   vector datasets   src/io/snapshot_io_particle_helpers.inc:126-270
//   physics text      src/io/snapshot_physics.cpp:88-175
//   string attributes src/io/snapshot_io_hdf5_helpers.inc:119-150
//
// The HDF5 C library is loaded at run time (dlopen) through the few entry
// points declared below, so the library builds without HDF5 headers; a
// missing libhdf5 is reported as a runtime_error by the snapshot calls only.
#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <locale>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "hyowon_pm.h"

namespace {

// --------------------------------------------------------------------------
// HDF5 1.10+ C API subset (types and constants of the public API)
// --------------------------------------------------------------------------
using hid_t = int64_t;
using herr_t = int;
using htri_t = int;
using hsize_t = unsigned long long;
constexpr hid_t kP_DEFAULT = 0;
constexpr hid_t kS_ALL = 0;
constexpr unsigned kF_ACC_RDONLY = 0x0000u, kF_ACC_TRUNC = 0x0002u, kF_ACC_EXCL = 0x0004u;
constexpr int kS_SCALAR = 0;
constexpr int kT_CSET_UTF8 = 1;
constexpr int kT_STR_NULLTERM = 0;
constexpr int kF_SCOPE_GLOBAL = 1;

struct H5 {
    void* so = nullptr;
    herr_t (*open)();
    herr_t (*eset_auto2)(hid_t, void*, void*);
    hid_t (*fcreate)(const char*, unsigned, hid_t, hid_t);
    hid_t (*fopen)(const char*, unsigned, hid_t);
    herr_t (*fclose)(hid_t);
    herr_t (*fflush)(hid_t, int);
    hid_t (*gcreate2)(hid_t, const char*, hid_t, hid_t, hid_t);
    hid_t (*gopen2)(hid_t, const char*, hid_t);
    herr_t (*gclose)(hid_t);
    hid_t (*screate)(int);
    hid_t (*screate_simple)(int, const hsize_t*, const hsize_t*);
    int (*sget_simple_extent_dims)(hid_t, hsize_t*, hsize_t*);
    int (*sget_simple_extent_ndims)(hid_t);
    herr_t (*sclose)(hid_t);
    hid_t (*acreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t);
    hid_t (*aopen)(hid_t, const char*, hid_t);
    htri_t (*aexists)(hid_t, const char*);
    herr_t (*awrite)(hid_t, hid_t, const void*);
    herr_t (*aread)(hid_t, hid_t, void*);
    hid_t (*aget_type)(hid_t);
    hid_t (*aget_space)(hid_t);
    herr_t (*aclose)(hid_t);
    hid_t (*tcopy)(hid_t);
    herr_t (*tset_size)(hid_t, size_t);
    size_t (*tget_size)(hid_t);
    herr_t (*tset_cset)(hid_t, int);
    herr_t (*tset_strpad)(hid_t, int);
    herr_t (*tclose)(hid_t);
    hid_t (*dcreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
    hid_t (*dopen2)(hid_t, const char*, hid_t);
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*);
    herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*);
    hid_t (*dget_space)(hid_t);
    herr_t (*dclose)(hid_t);
    htri_t (*lexists)(hid_t, const char*, hid_t);
    const hid_t* t_double;
    const hid_t* t_u32;
    const hid_t* t_u64;
    const hid_t* t_i32;
    const hid_t* t_c_s1;
};

struct SnapError {
    int code;
    std::string what;
};
[[noreturn]] void fail(int code, const std::string& w) { throw SnapError{code, w}; }

thread_local std::string t_error;

const H5& h5() {
    static std::once_flag once;
    static H5 lib;
    static std::string load_error;
    std::call_once(once, [] {
        const char* env = std::getenv("HY_HDF5_LIBRARY");
        const char* names[] = {env, "libhdf5.so", "libhdf5.so.310", "libhdf5.so.200", "libhdf5.so.103",
                               "/opt/conda/lib/libhdf5.so"};
        for (const char* n : names) {
            if (!n || !*n) continue;
            lib.so = dlopen(n, RTLD_NOW | RTLD_LOCAL);
            if (lib.so) break;
        }
        if (!lib.so) {
            load_error = "the HDF5 C library (libhdf5.so) could not be loaded; set HY_HDF5_LIBRARY";
            return;
        }
        bool ok = true;
        auto sym = [&](auto& fn, const char* name) {
            fn = reinterpret_cast<std::remove_reference_t<decltype(fn)>>(dlsym(lib.so, name));
            if (!fn) ok = false;
        };
        sym(lib.open, "H5open");
        sym(lib.eset_auto2, "H5Eset_auto2");
        sym(lib.fcreate, "H5Fcreate");
        sym(lib.fopen, "H5Fopen");
        sym(lib.fclose, "H5Fclose");
        sym(lib.fflush, "H5Fflush");
        sym(lib.gcreate2, "H5Gcreate2");
        sym(lib.gopen2, "H5Gopen2");
        sym(lib.gclose, "H5Gclose");
        sym(lib.screate, "H5Screate");
        sym(lib.screate_simple, "H5Screate_simple");
        sym(lib.sget_simple_extent_dims, "H5Sget_simple_extent_dims");
        sym(lib.sget_simple_extent_ndims, "H5Sget_simple_extent_ndims");
        sym(lib.sclose, "H5Sclose");
        sym(lib.acreate2, "H5Acreate2");
        sym(lib.aopen, "H5Aopen");
        sym(lib.aexists, "H5Aexists");
        sym(lib.awrite, "H5Awrite");
        sym(lib.aread, "H5Aread");
        sym(lib.aget_type, "H5Aget_type");
        sym(lib.aget_space, "H5Aget_space");
        sym(lib.aclose, "H5Aclose");
        sym(lib.tcopy, "H5Tcopy");
        sym(lib.tset_size, "H5Tset_size");
        sym(lib.tget_size, "H5Tget_size");
        sym(lib.tset_cset, "H5Tset_cset");
        sym(lib.tset_strpad, "H5Tset_strpad");
        sym(lib.tclose, "H5Tclose");
        sym(lib.dcreate2, "H5Dcreate2");
        sym(lib.dopen2, "H5Dopen2");
        sym(lib.dwrite, "H5Dwrite");
        sym(lib.dread, "H5Dread");
        sym(lib.dget_space, "H5Dget_space");
        sym(lib.dclose, "H5Dclose");
        sym(lib.lexists, "H5Lexists");
        sym(lib.t_double, "H5T_NATIVE_DOUBLE_g");
        sym(lib.t_u32, "H5T_NATIVE_UINT32_g");
        sym(lib.t_u64, "H5T_NATIVE_UINT64_g");
        sym(lib.t_i32, "H5T_NATIVE_INT32_g");
        sym(lib.t_c_s1, "H5T_C_S1_g");
        if (!ok || lib.open() < 0) {
            load_error = "the loaded HDF5 library lacks an entry point this reader needs";
            lib.so = nullptr;
            return;
        }
        lib.eset_auto2(0, nullptr, nullptr);  // errors come back as return codes, not stderr
    });
    if (!lib.so) fail(HY_ERUNTIME, load_error);
    return lib;
}

// RAII owner of an HDF5 identifier
struct Id {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Id(hid_t i, herr_t (*c)(hid_t), const std::string& what) : id(i), close(c) {
        if (i < 0) fail(HY_ERUNTIME, "HDF5: cannot " + what);
    }
    Id(const Id&) = delete;
    ~Id() {
        if (id >= 0 && close) close(id);
    }
    operator hid_t() const { return id; }
};

void check(herr_t r, const std::string& what) {
    if (r < 0) fail(HY_ERUNTIME, "HDF5: cannot " + what);
}

// --------------------------------------------------------------------------
// Schema vocabulary (include/cosmo_nbody/io/output_schema.hpp)
// --------------------------------------------------------------------------
const char* const kHeader = "/Header";
const char* const kPartType1 = "/PartType1";
const char* const kParameters = "/Parameters";
const char* const kConfig = "/Config";
const char* const kBindings[8][2] = {
    {"ParticleSpecies", "cold_plus_baryon_collisionless_mnu0"},
    {"PhaseSpaceGauge", "newtonian_nbody"},
    {"MassNormalization", "omega_m_rho_crit0_box_volume"},
    {"CoordinateUnit", "Mpc_per_h_comoving"},
    {"VelocityUnit", "km_per_s_peculiar"},
    {"MassUnit", "1e10_Msun_per_h"},
    {"EpochConvention", "scale_factor_a_and_redshift_1_over_a_minus_1"},
    {"PeriodicBoundaryConvention", "wrapped_half_open_0_L"},
};

// code units in cgs (include/cosmo_nbody/cosmology/units.hpp:13-35)
constexpr double kUnitLength_cm = 3.085677581491367e22 * 100.0;
constexpr double kUnitMass_g = 1e10 * 1.98847e30 * 1000.0;
constexpr double kUnitVelocity_cms = 1000.0 * 100.0;

void write_attr(hid_t loc, const char* name, hid_t type, const void* v, hsize_t count = 0) {
    const H5& L = h5();
    Id space(count ? L.screate_simple(1, &count, nullptr) : L.screate(kS_SCALAR), L.sclose,
             std::string("make attribute space ") + name);
    Id attr(L.acreate2(loc, name, type, space, kP_DEFAULT, kP_DEFAULT), L.aclose,
            std::string("create attribute ") + name);
    check(L.awrite(attr, type, v), std::string("write attribute ") + name);
}

void write_string_attr(hid_t loc, const char* name, const std::string& value) {
    const H5& L = h5();
    Id space(L.screate(kS_SCALAR), L.sclose, std::string("make attribute space ") + name);
    Id type(L.tcopy(*L.t_c_s1), L.tclose, "copy the string type");
    check(L.tset_cset(type, kT_CSET_UTF8), "set UTF-8");
    check(L.tset_size(type, value.size() + 1), "set string size");
    check(L.tset_strpad(type, kT_STR_NULLTERM), "set string padding");
    Id attr(L.acreate2(loc, name, type, space, kP_DEFAULT, kP_DEFAULT), L.aclose,
            std::string("create attribute ") + name);
    check(L.awrite(attr, type, value.c_str()), std::string("write attribute ") + name);
}

void read_attr(hid_t loc, const char* name, hid_t type, void* v) {
    const H5& L = h5();
    if (L.aexists(loc, name) <= 0) fail(HY_ERUNTIME, std::string("Snapshot is missing attribute ") + name);
    Id attr(L.aopen(loc, name, kP_DEFAULT), L.aclose, std::string("open attribute ") + name);
    check(L.aread(attr, type, v), std::string("read attribute ") + name);
}

std::string read_string_attr(hid_t loc, const char* name) {
    const H5& L = h5();
    if (L.aexists(loc, name) <= 0) fail(HY_ERUNTIME, std::string("Snapshot is missing attribute ") + name);
    Id attr(L.aopen(loc, name, kP_DEFAULT), L.aclose, std::string("open attribute ") + name);
    Id ftype(L.aget_type(attr), L.tclose, "attribute type");
    const size_t size = L.tget_size(ftype);
    if (size == 0 || size > (1u << 24)) fail(HY_ERUNTIME, std::string("Snapshot attribute ") + name + " has no text");
    Id mtype(L.tcopy(*L.t_c_s1), L.tclose, "copy the string type");
    check(L.tset_cset(mtype, kT_CSET_UTF8), "set UTF-8");
    check(L.tset_size(mtype, size), "set string size");
    check(L.tset_strpad(mtype, kT_STR_NULLTERM), "set string padding");
    std::vector<char> buf(size + 1, '\0');
    check(L.aread(attr, mtype, buf.data()), std::string("read attribute ") + name);
    return std::string(buf.data());
}

bool nearly_equal(double a, double b) {
    if (a == b) return true;
    if (!std::isfinite(a) || !std::isfinite(b)) return false;
    const double scale = std::max({std::fabs(a), std::fabs(b), 1.0});
    return std::fabs(a - b) <= 64.0 * std::numeric_limits<double>::epsilon() * scale;
}

// The reference's physics-fingerprint text layout (key=value lines, strings as
// key=<bytes>:<text>, reals at 17 significant digits in the classic locale).
std::string physics_fingerprint(const hy_snapshot_meta& m) {
    std::ostringstream out;
    out.imbue(std::locale::classic());
    out.precision(17);
    auto str = [&](const char* k, const std::string& v) { out << k << '=' << v.size() << ':' << v << '\n'; };
    auto real = [&](const char* k, double v) { out << k << '=' << v << '\n'; };
    auto uint = [&](const char* k, uint64_t v) { out << k << '=' << v << '\n'; };
    out << "snapshot_physics\n";
    real("h", m.h);
    real("omega_m", m.omega_m);
    real("omega_lambda", m.omega_lambda);
    real("omega_b", m.omega_b);
    real("sigma8", m.sigma8);
    real("n_s", m.n_s);
    real("box_L", m.box_size);
    uint("box_N", m.particles_per_dim);
    uint("box_N_mesh", m.pm_mesh_per_dim);
    str("gravity_solver", "PM");
    uint("gravity_deconvolve_cic", m.deconvolve_cic ? 1u : 0u);
    str("gravity_force_resolution_model", "pm_mesh_cic_discrete_operator");
    real("time_z_start", m.z_start);
    real("time_z_final", m.z_final);
    real("time_delta_ln_a", m.delta_ln_a);
    str("time_step_policy", "global");
    str("ic_mode", "generate");
    uint("ic_lpt_order", (uint64_t)m.lpt_order);
    uint("ic_seed", m.seed);
    real("ic_power_spectrum_redshift", m.power_spectrum_redshift);
    str("ic_power_spectrum_fidelity", m.power_spectrum_fidelity ? m.power_spectrum_fidelity : "");
    uint("ic_mesh_per_dimension", m.ic_mesh_per_dim);
    str("ic_amplitude_mode", m.ic_amplitude_mode ? m.ic_amplitude_mode : "");
    str("ic_phase_pairing", m.ic_phase_pairing ? m.ic_phase_pairing : "");
    return out.str();
}

std::string run_metadata_json(const hy_snapshot_meta& m, double a) {
    std::ostringstream out;
    out.imbue(std::locale::classic());
    out.precision(17);
    out << "{\"producer\":\"hyowon_amd libhyowon_pm\",\"product_lineage\":\"direct_simulation\","
        << "\"scale_factor\":" << a << ",\"seed\":" << m.seed << ",\"particles_per_dimension\":" << m.particles_per_dim
        << ",\"pm_mesh_per_dimension\":" << m.pm_mesh_per_dim << ",\"ic_mesh_per_dimension\":" << m.ic_mesh_per_dim
        << ",\"lpt_order\":" << m.lpt_order << "}";
    return out.str();
}

void require_meta(const hy_snapshot_meta* m) {
    if (!m) fail(HY_EINVAL, "null snapshot metadata");
    if (!std::isfinite(m->box_size) || m->box_size <= 0.0 || !std::isfinite(m->h) || m->h <= 0.0 ||
        !std::isfinite(m->omega_m) || !std::isfinite(m->omega_lambda))
        fail(HY_EINVAL, "Snapshot box or cosmology metadata is not finite and positive");
}

template <typename F>
int guarded(F&& f) {
    try {
        f();
        t_error.clear();
        return HY_OK;
    } catch (const SnapError& e) {
        t_error = e.what;
        return e.code;
    } catch (const std::bad_alloc&) {
        t_error = "allocation failed";
        return HY_ENOMEM;
    } catch (const std::exception& e) {
        t_error = e.what();
        return HY_ERUNTIME;
    }
}

}  // namespace

extern "C" {

const char* hy_snapshot_last_error(void) { return t_error.c_str(); }

int hy_snapshot_write(const char* path, const hy_snapshot_meta* m, double a, uint64_t n, const double* x,
                      const double* y, const double* z, const double* px, const double* py, const double* pz,
                      const uint64_t* ids, const double* masses, double uniform_mass, uint32_t flags) {
    return guarded([&] {
        if (!path || !*path) fail(HY_EINVAL, "Snapshot output path must not be empty");
        require_meta(m);
        if (n > 0 && (!x || !y || !z || !px || !py || !pz))
            fail(HY_EINVAL, "hy_snapshot_write: null particle array");
        if (n > std::numeric_limits<uint32_t>::max())
            fail(HY_EOVERFLOW, "Single-file snapshot header cannot represent particle count above uint32 range");
        if (!std::isfinite(a) || a <= 0.0) fail(HY_EINVAL, "Snapshot scale factor must be finite and positive");
        if (!masses && (!std::isfinite(uniform_mass) || uniform_mass <= 0.0))
            fail(HY_EINVAL, "Snapshot uniform mass must be finite and positive");
        const double L = m->box_size;
        for (uint64_t i = 0; i < n; ++i) {
            if (!(x[i] >= 0.0 && x[i] < L) || !(y[i] >= 0.0 && y[i] < L) || !(z[i] >= 0.0 && z[i] < L))
                fail(HY_EINVAL, "Snapshot positions must be finite and wrapped into [0, L)");
            if (!std::isfinite(px[i]) || !std::isfinite(py[i]) || !std::isfinite(pz[i]))
                fail(HY_EINVAL, "Snapshot momenta must be finite");
            if (masses && !(std::isfinite(masses[i]) && masses[i] >= 0.0))
                fail(HY_EINVAL, "Snapshot masses must be finite and non-negative");
        }
        const H5& H = h5();
        const unsigned mode = (flags & HY_SNAPSHOT_REPLACE) ? kF_ACC_TRUNC : kF_ACC_EXCL;
        Id file(H.fcreate(path, mode, kP_DEFAULT, kP_DEFAULT), H.fclose, std::string("create snapshot ") + path);
        const hid_t dbl = *H.t_double, u32 = *H.t_u32, u64 = *H.t_u64, i32 = *H.t_i32;
        {
            Id header(H.gcreate2(file, kHeader, kP_DEFAULT, kP_DEFAULT, kP_DEFAULT), H.gclose, "create /Header");
            const uint32_t npart[6] = {0, (uint32_t)n, 0, 0, 0, 0};
            double table[6] = {0, 0, 0, 0, 0, 0};
            if (!masses) table[1] = uniform_mass;
            const double redshift = 1.0 / a - 1.0;
            const int32_t one = 1, dp = 1;
            write_attr(header, "NumPart_ThisFile", u32, npart, 6);
            write_attr(header, "NumPart_Total", u32, npart, 6);
            write_attr(header, "MassTable", dbl, table, 6);
            write_attr(header, "Time", dbl, &a);
            write_attr(header, "Redshift", dbl, &redshift);
            write_attr(header, "BoxSize", dbl, &m->box_size);
            write_attr(header, "Omega0", dbl, &m->omega_m);
            write_attr(header, "OmegaLambda", dbl, &m->omega_lambda);
            write_attr(header, "HubbleParam", dbl, &m->h);
            write_attr(header, "NumFilesPerSnapshot", i32, &one);
            write_attr(header, "Flag_DoublePrecision", i32, &dp);
            write_attr(header, "UnitLength_in_cm", dbl, &kUnitLength_cm);
            write_attr(header, "UnitMass_in_g", dbl, &kUnitMass_g);
            write_attr(header, "UnitVelocity_in_cm_per_s", dbl, &kUnitVelocity_cms);
        }
        {
            Id params(H.gcreate2(file, kParameters, kP_DEFAULT, kP_DEFAULT, kP_DEFAULT), H.gclose,
                      "create /Parameters");
            write_attr(params, "Seed", u64, &m->seed);
            write_attr(params, "ParticlesPerDim", u32, &m->particles_per_dim);
            write_attr(params, "PMMeshPerDim", u32, &m->pm_mesh_per_dim);
            write_attr(params, "ICMeshPerDim", u32, &m->ic_mesh_per_dim);
            write_string_attr(params, "ICAmplitudeMode", m->ic_amplitude_mode ? m->ic_amplitude_mode : "gaussian");
            write_string_attr(params, "ICPhasePairing", m->ic_phase_pairing ? m->ic_phase_pairing : "independent");
            write_attr(params, "SofteningComoving", dbl, &m->softening);
            write_attr(params, "Sigma8", dbl, &m->sigma8);
            write_attr(params, "n_s", dbl, &m->n_s);
            write_attr(params, "OmegaB", dbl, &m->omega_b);
            write_attr(params, "LptOrder", i32, &m->lpt_order);
            write_string_attr(params, "PowerSpectrumFidelity",
                              m->power_spectrum_fidelity ? m->power_spectrum_fidelity : "precision_boltzmann");
        }
        {
            Id config(H.gcreate2(file, kConfig, kP_DEFAULT, kP_DEFAULT, kP_DEFAULT), H.gclose, "create /Config");
            write_string_attr(config, "SoftwareName", "HYOWON");
            write_string_attr(config, "SnapshotSchema", "hyowon.snapshot");
            write_string_attr(config, "PhysicsFingerprint", physics_fingerprint(*m));
            write_string_attr(config, "RunMetadataJson", run_metadata_json(*m, a));
            for (const auto& b : kBindings) write_string_attr(config, b[0], b[1]);
        }
        if (n != 0) {
            Id part(H.gcreate2(file, kPartType1, kP_DEFAULT, kP_DEFAULT, kP_DEFAULT), H.gclose, "create /PartType1");
            const hsize_t dims2[2] = {(hsize_t)n, 3};
            std::vector<double> buf(3 * n);
            const double inv_a = 1.0 / a;
            for (int which = 0; which < 2; ++which) {
                const bool vel = which == 1;
                const char* name = vel ? "Velocities" : "Coordinates";
                for (uint64_t i = 0; i < n; ++i) {
                    if (vel) {
                        buf[3 * i + 0] = px[i] * inv_a;
                        buf[3 * i + 1] = py[i] * inv_a;
                        buf[3 * i + 2] = pz[i] * inv_a;
                    } else {
                        buf[3 * i + 0] = x[i];
                        buf[3 * i + 1] = y[i];
                        buf[3 * i + 2] = z[i];
                    }
                }
                Id space(H.screate_simple(2, dims2, nullptr), H.sclose, "make vector space");
                Id ds(H.dcreate2(part, name, dbl, space, kP_DEFAULT, kP_DEFAULT, kP_DEFAULT), H.dclose,
                      std::string("create dataset ") + name);
                write_string_attr(ds, "Unit", vel ? "km_per_s_peculiar" : "Mpc_per_h_comoving");
                if (vel) write_string_attr(ds, "VelocityConvention", "v_pec_kms");
                check(H.dwrite(ds, dbl, kS_ALL, kS_ALL, kP_DEFAULT, buf.data()), std::string("write ") + name);
            }
            const hsize_t dims1[1] = {(hsize_t)n};
            Id space1(H.screate_simple(1, dims1, nullptr), H.sclose, "make scalar space");
            std::vector<uint64_t> seq;
            if (!ids) {
                seq.resize(n);
                for (uint64_t i = 0; i < n; ++i) seq[i] = i;
            }
            {
                Id ds(H.dcreate2(part, "ParticleIDs", u64, space1, kP_DEFAULT, kP_DEFAULT, kP_DEFAULT), H.dclose,
                      "create dataset ParticleIDs");
                write_string_attr(ds, "Unit", "dimensionless_stable_particle_id");
                check(H.dwrite(ds, u64, kS_ALL, kS_ALL, kP_DEFAULT, ids ? ids : seq.data()), "write ParticleIDs");
            }
            if (masses) {
                Id ds(H.dcreate2(part, "Masses", dbl, space1, kP_DEFAULT, kP_DEFAULT, kP_DEFAULT), H.dclose,
                      "create dataset Masses");
                write_string_attr(ds, "Unit", "1e10_Msun_per_h");
                check(H.dwrite(ds, dbl, kS_ALL, kS_ALL, kP_DEFAULT, masses), "write Masses");
            }
        }
        check(H.fflush(file, kF_SCOPE_GLOBAL), "flush the snapshot file");
    });
}

int hy_snapshot_read_header(const char* path, hy_snapshot_header* out) {
    return guarded([&] {
        if (!path || !out) fail(HY_EINVAL, "hy_snapshot_read_header: null argument");
        const H5& H = h5();
        Id file(H.fopen(path, kF_ACC_RDONLY, kP_DEFAULT), H.fclose, std::string("open snapshot ") + path);
        if (H.lexists(file, kHeader, kP_DEFAULT) <= 0) fail(HY_ERUNTIME, "Native snapshot is missing /Header");
        Id header(H.gopen2(file, kHeader, kP_DEFAULT), H.gclose, "open /Header");
        hy_snapshot_header h{};
        double table[6];
        uint32_t np[6], nt[6];
        int32_t files = 0, dp = 0;
        read_attr(header, "Time", *H.t_double, &h.a);
        read_attr(header, "Redshift", *H.t_double, &h.redshift);
        read_attr(header, "BoxSize", *H.t_double, &h.box_size);
        read_attr(header, "Omega0", *H.t_double, &h.omega_m);
        read_attr(header, "OmegaLambda", *H.t_double, &h.omega_lambda);
        read_attr(header, "HubbleParam", *H.t_double, &h.h);
        read_attr(header, "MassTable", *H.t_double, table);
        read_attr(header, "NumPart_ThisFile", *H.t_u32, np);
        read_attr(header, "NumPart_Total", *H.t_u32, nt);
        read_attr(header, "NumFilesPerSnapshot", *H.t_i32, &files);
        read_attr(header, "Flag_DoublePrecision", *H.t_i32, &dp);
        if (!std::isfinite(h.a) || h.a <= 0.0)
            fail(HY_ERUNTIME, "Snapshot contains non-finite or non-positive scale factor");
        for (double v : table)
            if (!std::isfinite(v) || v < 0.0) fail(HY_ERUNTIME, "Snapshot MassTable values must be finite and non-negative");
        if (!std::isfinite(h.redshift) || !nearly_equal(h.redshift, 1.0 / h.a - 1.0))
            fail(HY_ERUNTIME, "Snapshot Header.Redshift is inconsistent with Header.Time");
        if (files != 1) fail(HY_ERUNTIME, "Serial snapshot reader requires NumFilesPerSnapshot=1");
        if (dp != 1) fail(HY_ERUNTIME, "Snapshot datasets are not binary64 (Flag_DoublePrecision)");
        for (int t = 0; t < 6; ++t) {
            if (np[t] != nt[t]) fail(HY_ERUNTIME, "Single-file snapshot NumPart_ThisFile and NumPart_Total disagree");
            if (t != 1 && np[t] != 0)
                fail(HY_ERUNTIME, "Native one-fluid snapshot contains a particle population outside the effective "
                                  "collisionless PartType1 cb slot");
        }
        const double units[3] = {kUnitLength_cm, kUnitMass_g, kUnitVelocity_cms};
        const char* unit_names[3] = {"UnitLength_in_cm", "UnitMass_in_g", "UnitVelocity_in_cm_per_s"};
        for (int u = 0; u < 3; ++u) {
            double v = 0.0;
            read_attr(header, unit_names[u], *H.t_double, &v);
            if (!nearly_equal(v, units[u]))
                fail(HY_ERUNTIME, std::string("Snapshot unit attribute disagrees with native code units: ") + unit_names[u]);
        }
        if (H.lexists(file, kConfig, kP_DEFAULT) <= 0)
            fail(HY_ERUNTIME, "Native snapshot is missing the Config semantic metadata group");
        Id config(H.gopen2(file, kConfig, kP_DEFAULT), H.gclose, "open /Config");
        for (const auto& b : kBindings)
            if (read_string_attr(config, b[0]) != b[1])
                fail(HY_ERUNTIME, std::string("Native snapshot Config attribute ") + b[0] + " has an unknown value");
        if (read_string_attr(config, "PhysicsFingerprint").empty() || read_string_attr(config, "RunMetadataJson").empty())
            fail(HY_ERUNTIME, "Native snapshot Config provenance attributes must be non-empty");
        h.particles = np[1];
        h.uniform_mass = table[1];
        h.explicit_masses = table[1] > 0.0 ? 0 : 1;
        *out = h;
    });
}

int hy_snapshot_read(const char* path, const hy_snapshot_header* expect, uint64_t capacity, double* x, double* y,
                     double* z, double* px, double* py, double* pz, uint64_t* ids, double* masses) {
    return guarded([&] {
        hy_snapshot_header h{};
        if (hy_snapshot_read_header(path, &h) != HY_OK) fail(HY_ERUNTIME, t_error);
        if (expect && (!nearly_equal(h.box_size, expect->box_size) || !nearly_equal(h.omega_m, expect->omega_m) ||
                       !nearly_equal(h.omega_lambda, expect->omega_lambda) || !nearly_equal(h.h, expect->h)))
            fail(HY_ERUNTIME, "Snapshot box or cosmology header disagrees with the admitted read context");
        const uint64_t n = h.particles;
        if (n > capacity) fail(HY_EINVAL, "Snapshot particle count exceeds the caller's capacity");
        if (n == 0) return;
        if (!x || !y || !z || !px || !py || !pz) fail(HY_EINVAL, "hy_snapshot_read: null particle array");
        if (h.explicit_masses && !masses) fail(HY_EINVAL, "Snapshot holds explicit masses: pass a masses array");
        const H5& H = h5();
        Id file(H.fopen(path, kF_ACC_RDONLY, kP_DEFAULT), H.fclose, std::string("open snapshot ") + path);
        if (H.lexists(file, kPartType1, kP_DEFAULT) <= 0) fail(HY_ERUNTIME, "Native snapshot is missing /PartType1");
        Id part(H.gopen2(file, kPartType1, kP_DEFAULT), H.gclose, "open /PartType1");
        const hid_t dbl = *H.t_double, u64 = *H.t_u64;
        auto open_checked = [&](const char* name, int rank, const char* unit) {
            if (H.lexists(part, name, kP_DEFAULT) <= 0)
                fail(HY_ERUNTIME, std::string("Snapshot is missing dataset ") + name);
            hid_t ds = H.dopen2(part, name, kP_DEFAULT);
            if (ds < 0) fail(HY_ERUNTIME, std::string("HDF5: cannot open dataset ") + name);
            Id space(H.dget_space(ds), H.sclose, "dataset space");
            hsize_t dims[2] = {0, 0};
            const int r = H.sget_simple_extent_ndims(space);
            if (r != rank || H.sget_simple_extent_dims(space, dims, nullptr) != rank || dims[0] != n ||
                (rank == 2 && dims[1] != 3)) {
                H.dclose(ds);
                fail(HY_ERUNTIME, std::string("Snapshot dataset ") + name + " has the wrong shape");
            }
            const std::string u = read_string_attr(ds, "Unit");
            if (u != unit) {
                H.dclose(ds);
                fail(HY_ERUNTIME, std::string("Snapshot dataset ") + name + " has an unknown unit");
            }
            return ds;
        };
        std::vector<double> buf(3 * n);
        for (int which = 0; which < 2; ++which) {
            const bool vel = which == 1;
            Id ds(open_checked(vel ? "Velocities" : "Coordinates", 2, vel ? "km_per_s_peculiar" : "Mpc_per_h_comoving"),
                  H.dclose, "dataset");
            check(H.dread(ds, dbl, kS_ALL, kS_ALL, kP_DEFAULT, buf.data()), "read a vector dataset");
            double* o[3] = {vel ? px : x, vel ? py : y, vel ? pz : z};
            for (uint64_t i = 0; i < n; ++i)
                for (int d = 0; d < 3; ++d) {
                    const double v = buf[3 * i + d];
                    if (!std::isfinite(v)) fail(HY_ERUNTIME, "Snapshot phase space is not finite");
                    if (!vel && !(v >= 0.0 && v < h.box_size))
                        fail(HY_ERUNTIME, "Snapshot coordinates lie outside [0, BoxSize)");
                    o[d][i] = vel ? v * h.a : v;  // p = a v (src/io/snapshot_io_particle_helpers.inc:255-265)
                }
        }
        {
            Id ds(open_checked("ParticleIDs", 1, "dimensionless_stable_particle_id"), H.dclose, "dataset");
            std::vector<uint64_t> tmp;
            uint64_t* dst = ids;
            if (!dst) {
                tmp.resize(n);
                dst = tmp.data();
            }
            check(H.dread(ds, u64, kS_ALL, kS_ALL, kP_DEFAULT, dst), "read ParticleIDs");
        }
        if (h.explicit_masses) {
            Id ds(open_checked("Masses", 1, "1e10_Msun_per_h"), H.dclose, "dataset");
            check(H.dread(ds, dbl, kS_ALL, kS_ALL, kP_DEFAULT, masses), "read Masses");
            for (uint64_t i = 0; i < n; ++i)
                if (!std::isfinite(masses[i]) || masses[i] < 0.0)
                    fail(HY_ERUNTIME, "Snapshot masses must be finite and non-negative");
        }
    });
}

}  // extern "C"
