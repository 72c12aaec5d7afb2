"""ctypes binding of libhyowon_pm.so (include/hyowon_pm.h)."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhyowon_pm.so")

HY_OK = 0
HY_PREC_FP64 = 0
HY_PREC_FP32_MIXED = 1
HY_HOST_POINTERS = 0x1
HY_HAS_PREV = 0x2
HY_HAS_NEXT = 0x4
HY_ASYNC = 0x8
HY_REUSE_KEYS = 0x10
HY_COLLECT_POTENTIAL = 0x20

EXPORTS = [
    "hy_pm_create", "hy_pm_destroy", "hy_pm_get_info", "hy_last_error", "hy_abi_version",
    "hy_pm_force", "hy_pm_kdk", "hy_pm_density", "hy_pm_potential", "hy_pm_gather",
    "hy_pm_bucket_order", "hy_pm_xplane_order", "hy_pm_check", "hy_hubble", "hy_drift_factor",
    "hy_kick_factor", "hy_geometric_mean", "hy_time_boundaries", "hy_exact_sum",
    "hy_gravitational_constant", "hy_rho_crit0", "hy_pm_profile", "hy_pm_profile_read",
    "hy_pm_get_tuning", "hy_pm_set_tuning", "hy_pm_autotune", "hy_ensemble_create", "hy_ensemble_step",
    "hy_ensemble_check", "hy_ensemble_reset", "hy_ensemble_destroy", "hy_pm_kdk_vjp", "hy_pm_density_vjp",
    "hy_ic_gaussian_modes", "hy_ic_last_error", "hy_pm_power_spectrum", "hy_pm_force_meshes",
    "hy_pm_interpolate", "hy_ic_lpt", "hy_xplane_order_static", "hy_static_last_error",
    "hy_growth", "hy_linear_power", "hy_ic_mode_variance", "hy_cubic_spline", "hy_pm_potential_energy",
    "hy_snapshot_write", "hy_snapshot_read_header", "hy_snapshot_read", "hy_snapshot_last_error",
    "hy_pm_autotune_gather", "hy_pm_set_gather_rows",
]

HY_FFT_PASSES = 5
HY_GATHER_ROWS_DEFAULT = 0xFFFFFFFF
HY_GATHER_TUNE_CANDIDATES = 5
GATHER_TUNE_ROWS = [0, 16, 64, 256, 1024]
HY_FFT_MAX_VARIANTS = 4
FFT_PASS_NAMES = ["z_r2c", "y_fwd", "x_green", "y_bwd", "z_c2r"]


class HyError(RuntimeError):
    """Failure reported by libhyowon_pm; `code` is the hy_status value."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


# hy_status -> Python exception class mirroring the reference's C++ exception
class HyInvalidArgument(HyError, ValueError):
    pass


class HyLogicError(HyError):
    pass


class HyOverflowError(HyError, OverflowError):
    pass


class HyUnderflowError(HyError, ArithmeticError):
    pass


class HyDeviceError(HyError):
    pass


class HyRuntimeError(HyError):
    pass


class HyMemoryError(HyError, MemoryError):
    pass


_BY_CODE = {1: HyInvalidArgument, 2: HyLogicError, 3: HyOverflowError, 4: HyUnderflowError,
            5: HyDeviceError, 6: HyRuntimeError, 7: HyMemoryError}


class hy_pm_desc(C.Structure):
    _fields_ = [("box_size", C.c_double), ("mesh_size", C.c_uint64), ("deconvolve_cic", C.c_int),
                ("precision", C.c_int), ("device", C.c_int), ("stream", C.c_void_p),
                ("max_particles", C.c_uint64), ("boxes", C.c_uint32), ("method", C.c_int),
                ("split_scale", C.c_double)]


class hy_pm_info(C.Structure):
    _fields_ = [("mesh_size", C.c_uint64), ("tile_edge", C.c_uint64), ("bucket_count", C.c_uint64),
                ("device_bytes", C.c_uint64), ("last_input_mass", C.c_double),
                ("last_deposited_mass", C.c_double), ("precision", C.c_int), ("boxes", C.c_uint32)]


class hy_fft_tuning(C.Structure):
    _fields_ = [("variant", C.c_uint32 * HY_FFT_PASSES), ("plane_group", C.c_uint32),
                ("variants", C.c_uint32),
                ("rows", (C.c_uint32 * HY_FFT_MAX_VARIANTS) * HY_FFT_PASSES),
                ("pass_ms", (C.c_float * HY_FFT_MAX_VARIANTS) * HY_FFT_PASSES),
                ("group_ms", C.c_float * 4)]


class hy_ensemble_member(C.Structure):
    _fields_ = [("ctx", C.c_void_p), ("n", C.c_uint64), ("x", C.c_void_p), ("y", C.c_void_p),
                ("z", C.c_void_p), ("px", C.c_void_p), ("py", C.c_void_p), ("pz", C.c_void_p),
                ("uniform_mass", C.c_double)]


HY_PK_MAX_BINS = 1024
HY_METHOD_PURE_PM = 0
HY_METHOD_TREEPM_LONG_RANGE = 1


class hy_cosmology(C.Structure):
    _fields_ = [("h", C.c_double), ("omega_m", C.c_double), ("omega_b", C.c_double),
                ("omega_lambda", C.c_double), ("sigma8", C.c_double), ("n_s", C.c_double)]


class hy_snapshot_meta(C.Structure):
    _fields_ = [("box_size", C.c_double), ("omega_m", C.c_double), ("omega_lambda", C.c_double),
                ("h", C.c_double), ("omega_b", C.c_double), ("sigma8", C.c_double), ("n_s", C.c_double),
                ("seed", C.c_uint64), ("particles_per_dim", C.c_uint32), ("pm_mesh_per_dim", C.c_uint32),
                ("ic_mesh_per_dim", C.c_uint32), ("lpt_order", C.c_int32), ("softening", C.c_double),
                ("z_start", C.c_double), ("z_final", C.c_double), ("delta_ln_a", C.c_double),
                ("power_spectrum_redshift", C.c_double), ("deconvolve_cic", C.c_int),
                ("ic_amplitude_mode", C.c_char_p), ("ic_phase_pairing", C.c_char_p),
                ("power_spectrum_fidelity", C.c_char_p)]


class hy_snapshot_header(C.Structure):
    _fields_ = [("a", C.c_double), ("redshift", C.c_double), ("box_size", C.c_double), ("omega_m", C.c_double),
                ("omega_lambda", C.c_double), ("h", C.c_double), ("uniform_mass", C.c_double),
                ("particles", C.c_uint64), ("explicit_masses", C.c_int)]


HY_SNAPSHOT_REPLACE = 0x1


class hy_pk_options(C.Structure):
    _fields_ = [("num_bins", C.c_int), ("subtract_shot_noise", C.c_int), ("interlaced", C.c_int),
                ("max_k_fraction_nyquist", C.c_double)]


class hy_pk_bin(C.Structure):
    _fields_ = [("k_low", C.c_double), ("k_mean", C.c_double), ("k_high", C.c_double),
                ("p_raw", C.c_double), ("p_corrected", C.c_double),
                ("mean_cic_power_deconvolution_factor", C.c_double),
                ("shot_noise_reference", C.c_double), ("mode_count", C.c_uint64)]


_lib = None


def lib():
    """Load libhyowon_pm.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m hyowon_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, u64, d, i, u32 = C.c_void_p, C.c_uint64, C.c_double, C.c_int, C.c_uint32
    pd = C.POINTER(C.c_double)

    def fn(name, res, *args):
        f = getattr(L, name)
        f.restype = res
        f.argtypes = list(args)

    fn("hy_pm_create", i, C.POINTER(hy_pm_desc), C.POINTER(vp))
    fn("hy_pm_destroy", i, vp)
    fn("hy_pm_get_info", i, vp, C.POINTER(hy_pm_info))
    fn("hy_last_error", C.c_char_p)
    fn("hy_abi_version", i)
    fn("hy_pm_force", i, vp, u64, vp, vp, vp, vp, d, vp, vp, vp, u32)
    fn("hy_pm_kdk", i, vp, u64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, d, d, d, d,
       vp, vp, vp, u32)
    fn("hy_pm_density", i, vp, u64, vp, vp, vp, vp, d, vp)
    fn("hy_pm_potential", i, vp, vp, vp)
    fn("hy_pm_gather", i, vp, u64, vp, vp, vp, vp, vp, vp, vp)
    fn("hy_pm_bucket_order", i, vp, u64, vp, vp, vp, vp, vp)
    fn("hy_pm_xplane_order", i, vp, u64, vp, vp, vp)
    fn("hy_pm_check", i, vp)
    fn("hy_pm_potential_energy", i, vp, pd)
    fn("hy_hubble", d, d, d, d)
    fn("hy_drift_factor", i, d, d, d, d, pd)
    fn("hy_kick_factor", i, d, d, d, d, pd)
    fn("hy_geometric_mean", d, d, d)
    fn("hy_time_boundaries", i, d, d, d, pd, u64, pd, u64, C.POINTER(C.c_uint64))
    fn("hy_exact_sum", d, pd, u64)
    fn("hy_gravitational_constant", d)
    fn("hy_rho_crit0", d)
    fn("hy_pm_profile", i, vp, i, u64)
    fn("hy_pm_profile_read", i, vp, pd, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64))
    fn("hy_pm_get_tuning", i, vp, C.POINTER(hy_fft_tuning))
    fn("hy_pm_set_tuning", i, vp, C.POINTER(hy_fft_tuning))
    fn("hy_pm_autotune", i, vp, u32, C.POINTER(hy_fft_tuning))
    fn("hy_pm_autotune_gather", i, vp, u64, vp, vp, vp, d, u32, C.POINTER(u32), C.POINTER(C.c_float))
    fn("hy_pm_set_gather_rows", i, vp, u32)
    fn("hy_ensemble_create", i, C.POINTER(hy_ensemble_member), u32, C.POINTER(vp))
    fn("hy_ensemble_step", i, vp, d, d, d, u32)
    fn("hy_ensemble_check", i, vp)
    fn("hy_ensemble_reset", i, vp)
    fn("hy_ensemble_destroy", i, vp)
    fn("hy_pm_kdk_vjp", i, vp, u64, vp, vp, vp, d, vp, vp, vp, vp, vp, vp, d, d, d, u32)
    fn("hy_pm_density_vjp", i, vp, u64, vp, vp, vp, d, vp, d, vp, vp, vp)
    fn("hy_ic_gaussian_modes", i, u64, u64, u64, vp, vp, u32, i, vp)
    fn("hy_ic_last_error", C.c_char_p)
    fn("hy_pm_power_spectrum", i, vp, u64, vp, vp, vp, vp, d, C.POINTER(hy_pk_options),
       C.POINTER(hy_pk_bin), u32)
    fn("hy_pm_force_meshes", i, vp, vp, vp, vp, vp)
    fn("hy_pm_interpolate", i, vp, u64, vp, vp, vp, vp, vp)
    fn("hy_ic_lpt", i, vp, u64, vp, d, vp, vp, vp, vp, vp, vp)
    fn("hy_xplane_order_static", i, d, u64, u64, vp, vp, vp, i, vp)
    fn("hy_static_last_error", C.c_char_p)
    pc = C.POINTER(hy_cosmology)
    fn("hy_growth", i, pc, u64, pd, pd, pd, pd, pd)
    fn("hy_linear_power", i, pc, C.c_char_p, d, d, u64, pd, pd)
    fn("hy_ic_mode_variance", i, pc, C.c_char_p, d, d, u64, d, u64, pd)
    fn("hy_cubic_spline", i, u64, pd, pd, u64, pd, pd)
    fn("hy_snapshot_write", i, C.c_char_p, C.POINTER(hy_snapshot_meta), d, u64, pd, pd, pd, pd, pd, pd,
       C.POINTER(C.c_uint64), pd, d, u32)
    fn("hy_snapshot_read_header", i, C.c_char_p, C.POINTER(hy_snapshot_header))
    fn("hy_snapshot_read", i, C.c_char_p, C.POINTER(hy_snapshot_header), u64, pd, pd, pd, pd, pd, pd,
       C.POINTER(C.c_uint64), pd)
    fn("hy_snapshot_last_error", C.c_char_p)
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != HY_OK:
        msg = lib().hy_last_error().decode(errors="replace")
        raise _BY_CODE.get(rc, HyError)(rc, msg)
