"""ctypes bindings for the CPU oracle (oracle/liboracle.so) and the reference
build (oracle/_ref/libhyref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (hyowon_amd) never
imports this package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libhyref.so")
REFERENCE_TREE = "/root/reference"

_d = C.c_double
_u64 = C.c_uint64
_i64 = C.c_int64
_pd = C.POINTER(C.c_double)
_pu32 = C.POINTER(C.c_uint32)
_pu64 = C.POINTER(C.c_uint64)
_pi64 = C.POINTER(C.c_int64)

ERRORS = {0: None, 1: ValueError, 2: RuntimeError, 3: OverflowError, 4: ArithmeticError,
          6: RuntimeError, 7: RuntimeError}


class OracleError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what} failed with code {code}")
        self.code = code


def build(ref: bool | None = None) -> None:
    """Compile liboracle.so and, when the reference tree is present, _ref."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if ref is None:
        ref = os.path.isdir(REFERENCE_TREE)
    if ref:
        subprocess.run(["make", "-s", "-j8", "-C", HERE, "ref", f"REF={REFERENCE_TREE}"],
                       check=True)


def _ptr(a, t=_pd):
    return a.ctypes.data_as(t) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class _Lib:
    def __init__(self, path: str):
        self.lib = C.CDLL(path)

    def fn(self, name, restype, *argtypes):
        f = getattr(self.lib, name)
        f.restype = restype
        f.argtypes = list(argtypes)
        return f


_oracle = None
_ref = None


def lib():
    global _oracle
    if _oracle is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is not built (make -C oracle)")
        L = _Lib(LIB_PATH)
        L.fn("or_wrap", _d, _d, _d)
        L.fn("or_minimum_image", _d, _d, _d)
        L.fn("or_cic_coordinate", C.c_int, _d, _d, _d, _u64, _pi64, _pd)
        L.fn("or_xplane_order", C.c_int, _u64, _pd, _d, _u64, _pu64, _pu32)
        L.fn("or_bucket_tile_edge", _u64, _u64)
        L.fn("or_bucket_order", C.c_int, _u64, _pd, _pd, _pd, _d, _u64, _u64, _pu64, _pu32)
        L.fn("or_deposit", C.c_int, _u64, _pd, _pd, _pd, _pd, _d, _d, _u64, _pd)
        L.fn("or_exact_sum", _d, _pd, _u64)
        L.fn("or_exact_repeated", _d, _d, _u64)
        L.fn("or_require_cic_mass_conservation", C.c_int, _d, _d, _u64)
        L.fn("or_density_contrast", C.c_int, _pd, _u64, _d, _d)
        L.fn("or_fft_r2c_3d", C.c_int, _u64, _pd, _pd)
        L.fn("or_fft_c2r_3d", C.c_int, _u64, _pd, _pd)
        L.fn("or_fft_1d", C.c_int, _u64, _pd, C.c_int)
        L.fn("or_green_apply", C.c_int, _u64, _d, C.c_int, _pd)
        L.fn("or_green_kernel", _d, _u64, _d, C.c_int, _u64, _u64, _u64)
        L.fn("or_gather_centered_gradient_add", C.c_int, _u64, _pd, _pd, _pd, _pd, _d, _u64,
             _pd, _pd, _pd)
        L.fn("or_pm_force", C.c_int, _u64, _pd, _pd, _pd, _pd, _d, _d, _u64, C.c_int,
             _pd, _pd, _pd, _pd)
        L.fn("or_kick_drift", C.c_int, _u64, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd,
             _d, _d, _d)
        L.fn("or_kick", C.c_int, _u64, _pd, _pd, _pd, _pd, _pd, _pd, _d)
        L.fn("or_hubble", _d, _d, _d, _d)
        L.fn("or_drift_factor", C.c_int, _d, _d, _d, _d, _pd)
        L.fn("or_kick_factor", C.c_int, _d, _d, _d, _d, _pd)
        L.fn("or_geometric_mean", _d, _d, _d)
        L.fn("or_time_boundaries", C.c_int, _d, _d, _d, _pd, _u64, _pd, _u64, _pu64)
        L.fn("or_run", C.c_int, _u64, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _d,
             _d, _u64, C.c_int, _d, _d, _pd, _u64)
        L.fn("or_philox4x32_10", None, _pu32, _pu32, _pu32)
        L.fn("or_uniform_pair", None, _u64, C.c_int32, C.c_int32, C.c_int32, C.c_uint32, _pd)
        L.fn("or_normal_pair", None, _u64, C.c_int32, C.c_int32, C.c_int32, C.c_uint32, _pd)
        L.fn("or_gaussian_modes", C.c_int, _u64, _u64, _u64, _pd, C.c_uint32, _pd)
        L.fn("or_power_spectrum", C.c_int, _u64, _pd, _pd, _pd, _pd, _d, _d, _u64, C.c_int,
             C.c_int, C.c_int, _d, _pd)
        L.fn("or_pk_edges", C.c_int, _d, _d, C.c_int, _pd)
        L.fn("or_pk_bin_index", C.c_int, _pd, C.c_int, _d)
        L.fn("or_pk_window_axis", C.c_longdouble, _d, _d)
        L.fn("or_pk_raw_power", _d, _d, _d, _d)
        L.fn("or_pk_raw_power_minus", _d, _d, _d, _d, _d)
        L.fn("or_green_apply_treepm", C.c_int, _u64, _d, _d, _pd)
        L.fn("or_interpolate_add", C.c_int, _u64, _pd, _pd, _pd, _pd, _d, _u64, _pd)
        L.fn("or_pm_force_treepm", C.c_int, _u64, _pd, _pd, _pd, _pd, _d, _d, _u64, _d, _pd, _pd,
             _pd, _pd)
        L.fn("or_potential_energy", C.c_int, _u64, _pd, _pd, _pd, _pd, _pd, _d, _d, _u64, _pd)
        L.fn("or_lpt_first_order", C.c_int, _u64, _pd, _u64, _d, _d, _pd, _pd, _pd, _pd, _pd, _pd)
        L.fn("or_set_threads", None, C.c_int)
        L.fn("or_get_threads", C.c_int)
        L.fn("or_stage_reset", None)
        L.fn("or_stage_times", None, _pd)
        L.fn("or_gravitational_constant", _d)
        L.fn("or_rho_crit0", _d)
        _oracle = L.lib
    return _oracle


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


FORCE_CB = C.CFUNCTYPE(C.c_int, _u64, _pd, _pd, _pd, _pd, _pd, _pd, _d, C.c_void_p)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise FileNotFoundError(f"{REF_PATH} is not built (make -C oracle ref)")
        L = _Lib(REF_PATH)
        L.fn("ref_wrap", _d, _d, _d)
        L.fn("ref_minimum_image", _d, _d, _d)
        L.fn("ref_cic_coordinate", C.c_int, _d, _d, _u64, _pi64, _pd)
        L.fn("ref_deposit", C.c_int, _u64, _pd, _pd, _pd, _pd, _d, _d, _u64, _pd)
        L.fn("ref_density_contrast", C.c_int, _pd, _u64, _d, _d)
        L.fn("ref_deposit_unit", C.c_int, _u64, _pd, _pd, _pd, _pd, _d, _d, _d, _u64, _pd)
        L.fn("ref_pk_edges", C.c_int, _d, _d, C.c_int, _pd)
        L.fn("ref_pk_bin_index", C.c_int, _d, _d, C.c_int, _d, C.POINTER(C.c_int))
        L.fn("ref_pk_window_axis", C.c_int, _d, _d, C.POINTER(C.c_longdouble))
        L.fn("ref_pk_raw_power", C.c_int, _d, _d, _d, _pd)
        L.fn("ref_pk_raw_power_minus", C.c_int, _d, _d, _d, _d, _pd)
        L.fn("ref_pk_volume_times", C.c_int, _d, _d, _pd)
        L.fn("ref_density_direct_available", C.c_int, _u64, _d, _d, C.POINTER(C.c_int))
        L.fn("ref_green_apply", C.c_int, _u64, _d, C.c_int, _pd)
        L.fn("ref_green_apply_treepm", C.c_int, _u64, _d, _d, _pd)
        L.fn("ref_interpolate_add", C.c_int, _u64, _pd, _pd, _pd, _pd, _d, _u64, _pd)
        L.fn("ref_gather", C.c_int, _u64, _pd, _pd, _pd, _pd, _d, _u64, _pd, _pd, _pd)
        L.fn("ref_mass_conservation", C.c_int, _d, _d, _u64)
        L.fn("ref_exact_sum", _d, _pd, _u64)
        L.fn("ref_exact_repeated", _d, _d, _u64)
        L.fn("ref_kick_drift_one", C.c_int, _pd, _pd, _pd, _d, _d, _d)
        L.fn("ref_hubble", _d, _d, _d, _d)
        L.fn("ref_drift_kick", C.c_int, _d, _d, _d, _d, _pd, _pd)
        L.fn("ref_geometric_mean", _d, _d, _d)
        L.fn("ref_time_boundaries", C.c_int, _d, _d, _d, _pd, _u64, _pd, _u64, _pu64)
        L.fn("ref_leapfrog_run", C.c_int, _u64, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd,
             _d, _d, _d, _d, _d, _d, _d, FORCE_CB, C.c_void_p, _pu64)
        L.fn("ref_gravitational_constant", _d)
        L.fn("ref_rho_crit0", _d)
        L.fn("ref_philox4x32_10", None, _pu32, _pu32, _pu32)
        L.fn("ref_potential_energy", C.c_int, _u64, _pd, _pd, _pd, _pd, _pd, _d, _d, _u64, _pd)
        L.fn("ref_growth", C.c_int, _d, _d, _u64, _pd, _pd, _pd, _pd, _pd)
        L.fn("ref_cubic_spline", C.c_int, _u64, _pd, _pd, _u64, _pd, _pd)
        L.fn("ref_linear_power", C.c_int, _pd, C.c_char_p, _d, _d, _d, _u64, _u64, _u64, _u64, _pd, _pd)
        L.fn("ref_mode_variance", C.c_int, _pd, C.c_char_p, _d, _d, _d, _u64, _u64, _u64, _pd)
        L.fn("ref_uniform_pair", None, _u64, C.c_int32, C.c_int32, C.c_int32, C.c_uint32, _pd)
        L.fn("ref_normal_pair", None, _u64, C.c_int32, C.c_int32, C.c_int32, C.c_uint32, _pd)
        _ref = L.lib
    return _ref


def _check(rc, what):
    if rc != 0:
        raise OracleError(rc, what)


# --------------------------------------------------------------------------
# numpy-level helpers (oracle)
# --------------------------------------------------------------------------
def cic_coordinate(pos, L, N):
    b = _i64(0)
    f = _d(0)
    rc = lib().or_cic_coordinate(pos, L, L / N, N, C.byref(b), C.byref(f))
    _check(rc, "or_cic_coordinate")
    return b.value, f.value


def xplane_order(x, L, N):
    x = _f64(x)
    off = np.zeros(N + 1, np.uint64)
    order = np.zeros(len(x), np.uint32)
    _check(lib().or_xplane_order(len(x), _ptr(x), L, N, _ptr(off, _pu64), _ptr(order, _pu32)),
           "or_xplane_order")
    return off, order


def bucket_tile_edge(N):
    return int(lib().or_bucket_tile_edge(N))


def bucket_order(x, y, z, L, N, T=None):
    T = bucket_tile_edge(N) if T is None else T
    x, y, z = _f64(x), _f64(y), _f64(z)
    nb = (N // T) ** 3 * 8
    off = np.zeros(nb + 1, np.uint64)
    order = np.zeros(len(x), np.uint32)
    _check(lib().or_bucket_order(len(x), _ptr(x), _ptr(y), _ptr(z), L, N, T,
                                 _ptr(off, _pu64), _ptr(order, _pu32)), "or_bucket_order")
    return off, order


def deposit(x, y, z, L, N, masses=None, uniform_mass=1.0):
    x, y, z = _f64(x), _f64(y), _f64(z)
    m = _f64(masses) if masses is not None else None
    mesh = np.zeros(N ** 3, np.float64)
    _check(lib().or_deposit(len(x), _ptr(x), _ptr(y), _ptr(z), _ptr(m), uniform_mass, L, N,
                            _ptr(mesh)), "or_deposit")
    return mesh.reshape(N, N, N)


def density_contrast(mesh, L, N, total_mass):
    out = _f64(mesh).copy().reshape(-1)
    _check(lib().or_density_contrast(_ptr(out), N, L / N, total_mass), "or_density_contrast")
    return out.reshape(N, N, N)


def rfftn(real):
    N = real.shape[0]
    r = _f64(real).reshape(-1)
    out = np.zeros(N * N * (N // 2 + 1), np.complex128)
    _check(lib().or_fft_r2c_3d(N, _ptr(r), out.ctypes.data_as(_pd)), "or_fft_r2c_3d")
    return out.reshape(N, N, N // 2 + 1)


def irfftn(cplx):
    N = cplx.shape[0]
    c = np.ascontiguousarray(cplx, np.complex128).reshape(-1)
    out = np.zeros(N ** 3, np.float64)
    _check(lib().or_fft_c2r_3d(N, c.ctypes.data_as(_pd), _ptr(out)), "or_fft_c2r_3d")
    return out.reshape(N, N, N)


def green_apply(cplx, L, deconvolve=False):
    N = cplx.shape[0]
    c = np.ascontiguousarray(cplx, np.complex128).copy()
    _check(lib().or_green_apply(N, L, int(deconvolve), c.ctypes.data_as(_pd)), "or_green_apply")
    return c


def gather(phi, x, y, z, L, acc=None):
    N = phi.shape[0]
    x, y, z = _f64(x), _f64(y), _f64(z)
    ph = _f64(phi).reshape(-1)
    if acc is None:
        acc = np.zeros((3, len(x)), np.float64)
    else:
        acc = np.array(acc, np.float64)
    _check(lib().or_gather_centered_gradient_add(len(x), _ptr(ph), _ptr(x), _ptr(y), _ptr(z),
                                                 L, N, _ptr(acc[0]), _ptr(acc[1]),
                                                 _ptr(acc[2])), "or_gather")
    return acc


def pm_force(x, y, z, L, N, masses=None, uniform_mass=1.0, deconvolve=False,
             return_phi=False):
    x, y, z = _f64(x), _f64(y), _f64(z)
    m = _f64(masses) if masses is not None else None
    acc = np.zeros((3, len(x)), np.float64)
    phi = np.zeros(N ** 3, np.float64) if return_phi else None
    _check(lib().or_pm_force(len(x), _ptr(x), _ptr(y), _ptr(z), _ptr(m), uniform_mass, L, N,
                             int(deconvolve), _ptr(acc[0]), _ptr(acc[1]), _ptr(acc[2]),
                             _ptr(phi)), "or_pm_force")
    if return_phi:
        return acc, phi.reshape(N, N, N)
    return acc


def drift_factor(a1, a2, om=0.3, ol=0.7):
    v = _d(0)
    _check(lib().or_drift_factor(a1, a2, om, ol, C.byref(v)), "or_drift_factor")
    return v.value


def kick_factor(a1, a2, om=0.3, ol=0.7):
    v = _d(0)
    _check(lib().or_kick_factor(a1, a2, om, ol, C.byref(v)), "or_kick_factor")
    return v.value


def time_boundaries(a_start, a_final, dlna, stops=()):
    s = _f64(list(stops)) if len(stops) else np.zeros(1)
    cap = 1 << 16
    out = np.zeros(cap, np.float64)
    cnt = _u64(0)
    _check(lib().or_time_boundaries(a_start, a_final, dlna, _ptr(s), len(stops), _ptr(out), cap,
                                    C.byref(cnt)), "or_time_boundaries")
    return out[: cnt.value].copy()


def kick_drift(x, y, z, px, py, pz, ax, ay, az, K, D, L):
    arrs = [np.array(a, np.float64) for a in (x, y, z, px, py, pz)]
    acc = [_f64(a) for a in (ax, ay, az)]
    _check(lib().or_kick_drift(len(arrs[0]), *[_ptr(a) for a in arrs], *[_ptr(a) for a in acc],
                               K, D, L), "or_kick_drift")
    return arrs


def kick(px, py, pz, ax, ay, az, K):
    arrs = [np.array(a, np.float64) for a in (px, py, pz)]
    acc = [_f64(a) for a in (ax, ay, az)]
    _check(lib().or_kick(len(arrs[0]), *[_ptr(a) for a in arrs], *[_ptr(a) for a in acc], K),
           "or_kick")
    return arrs


def run(x, y, z, px, py, pz, L, N, boundaries, uniform_mass, om=0.3, ol=0.7,
        deconvolve=False):
    st = [np.array(a, np.float64) for a in (x, y, z, px, py, pz)]
    acc = [np.zeros(len(st[0])) for _ in range(3)]
    b = _f64(boundaries)
    _check(lib().or_run(len(st[0]), *[_ptr(a) for a in st], *[_ptr(a) for a in acc], None,
                        uniform_mass, L, N, int(deconvolve), om, ol, _ptr(b), len(b)), "or_run")
    return st, acc


STAGES = ("sort", "deposit", "mass_contrast", "fft_forward", "green", "fft_inverse", "gather", "kdk")


def stage_reset() -> None:
    """Zero the per-stage wall-time accumulators of the force refresh and step loop."""
    lib().or_stage_reset()


def stage_times() -> dict:
    """Seconds per stage accumulated since stage_reset()."""
    out = np.zeros(len(STAGES))
    lib().or_stage_times(_ptr(out))
    return dict(zip(STAGES, out.tolist()))


def set_threads(n: int) -> None:
    lib().or_set_threads(int(n))


def get_threads() -> int:
    return int(lib().or_get_threads())


def philox(counter, key):
    c = np.ascontiguousarray(counter, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, np.uint32)
    lib().or_philox4x32_10(_ptr(c, _pu32), _ptr(k, _pu32), _ptr(out, _pu32))
    return out


def gaussian_modes(N, seed, max_mode, variance_by_m2, flags=0):
    """Mode-keyed Gaussian r2c half spectrum, complex128 [N][N][N/2+1]."""
    var = _f64(variance_by_m2)
    out = np.zeros((N, N, N // 2 + 1), dtype=np.complex128)
    _check(lib().or_gaussian_modes(N, seed, max_mode, _ptr(var), flags, out.ctypes.data_as(_pd)),
           "or_gaussian_modes")
    return out


PK_FIELDS = ("k_low", "k_mean", "k_high", "p_raw", "p_corrected",
             "mean_cic_power_deconvolution_factor", "shot_noise_reference", "mode_count")


def power_spectrum(x, y, z, L, N, num_bins, subtract_shot_noise, interlaced, max_fraction,
                   masses=None, uniform_mass=1.0):
    """MatterPowerSpectrum::compute_pk restated: (num_bins, 8) float64 rows in PK_FIELDS order."""
    x, y, z = _f64(x), _f64(y), _f64(z)
    m = _f64(masses) if masses is not None else None
    out = np.zeros((num_bins, 8), np.float64)
    _check(lib().or_power_spectrum(len(x), _ptr(x), _ptr(y), _ptr(z), _ptr(m), uniform_mass, L, N,
                                   int(num_bins), int(subtract_shot_noise), int(interlaced),
                                   max_fraction, _ptr(out)), "or_power_spectrum")
    return out


def pk_edges(minimum, maximum, bins):
    out = np.zeros(bins + 1, np.float64)
    _check(lib().or_pk_edges(minimum, maximum, bins, _ptr(out)), "or_pk_edges")
    return out


def green_apply_treepm(cplx, L, split_scale):
    N = cplx.shape[0]
    c = np.ascontiguousarray(cplx, np.complex128).copy()
    _check(lib().or_green_apply_treepm(N, L, split_scale, c.ctypes.data_as(_pd)), "or_green_apply_treepm")
    return c


def interpolate_add(mesh, x, y, z, L, out=None):
    N = mesh.shape[0]
    x, y, z = _f64(x), _f64(y), _f64(z)
    out = np.zeros(len(x)) if out is None else _f64(out).copy()
    m = _f64(mesh).reshape(-1)
    _check(lib().or_interpolate_add(len(x), _ptr(m), _ptr(x), _ptr(y), _ptr(z), L, N, _ptr(out)),
           "or_interpolate_add")
    return out


def pm_force_treepm(x, y, z, L, N, split_scale, masses=None, uniform_mass=1.0, meshes=False):
    """TreePM long-range PM force (spectral gradient); returns (ax, ay, az[, force meshes])."""
    x, y, z = _f64(x), _f64(y), _f64(z)
    n = len(x)
    m = _f64(masses) if masses is not None else None
    ax, ay, az = np.zeros(n), np.zeros(n), np.zeros(n)
    mo = np.zeros(3 * N ** 3) if meshes else None
    _check(lib().or_pm_force_treepm(n, _ptr(x), _ptr(y), _ptr(z), _ptr(m), uniform_mass, L, N, split_scale,
                                    _ptr(ax), _ptr(ay), _ptr(az), _ptr(mo)), "or_pm_force_treepm")
    if meshes:
        return ax, ay, az, mo.reshape(3, N, N, N)
    return ax, ay, az


def lpt_first_order(delta_k, L, n_side, p_factor):
    """First-order LPT lattice ICs of an r2c half spectrum (complex128 [N][N][N/2+1]):
    returns (pos (3, n^3), mom (3, n^3))."""
    dk = np.ascontiguousarray(delta_k, np.complex128)
    N = dk.shape[0]
    n = n_side ** 3
    out = [np.zeros(n) for _ in range(6)]
    _check(lib().or_lpt_first_order(N, dk.ctypes.data_as(_pd), n_side, L, p_factor, *[_ptr(o) for o in out]),
           "or_lpt_first_order")
    return np.stack(out[:3]), np.stack(out[3:])


def potential_energy(phi, x, y, z, L, masses=None, uniform_mass=1.0):
    """0.5 sum_p m_p phi(x_p) (PMForceDiagnostics::potential_energy_comoving)."""
    N = phi.shape[0]
    x, y, z = _f64(x), _f64(y), _f64(z)
    m = _f64(masses) if masses is not None else None
    out = _d(0)
    _check(lib().or_potential_energy(len(x), _ptr(_f64(phi).reshape(-1)), _ptr(x), _ptr(y), _ptr(z), _ptr(m),
                                     uniform_mass, L, N, C.byref(out)), "or_potential_energy")
    return out.value
