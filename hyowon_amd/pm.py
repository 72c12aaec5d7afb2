"""Host-side mirror of the reference's PM-step interfaces over libhyowon_pm.

Device memory and streams come from PyTorch (torch.cuda tensors); all compute
runs in the HIP kernels of libhyowon_pm.so.  Arrays passed as numpy arrays are
host memory (HY_HOST_POINTERS); torch CUDA tensors are device memory.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from ._lib import (FFT_PASS_NAMES, HY_ASYNC, HY_COLLECT_POTENTIAL, HY_FFT_PASSES, HY_HAS_NEXT, HY_HAS_PREV,
                   HY_HOST_POINTERS, HY_METHOD_PURE_PM, HY_METHOD_TREEPM_LONG_RANGE,
                   HY_PREC_FP32_MIXED, HY_PREC_FP64, HY_REUSE_KEYS, check,
                   hy_fft_tuning, hy_pk_bin, hy_pk_options, hy_pm_desc, hy_pm_info, lib)

PRECISIONS = {"fp64": HY_PREC_FP64, "fp32": HY_PREC_FP32_MIXED, "fp32_mixed": HY_PREC_FP32_MIXED}


def _torch():
    import torch  # plumbing only: device memory and streams
    return torch


def _ptr(a):
    """Device or host address of a torch tensor / numpy array (None -> NULL)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("arrays must be C-contiguous")
        return a.ctypes.data
    if not a.is_contiguous():
        raise ValueError("tensors must be contiguous")
    return a.data_ptr()


def _is_host(a) -> bool:
    return isinstance(a, np.ndarray)


# --------------------------------------------------------------------------
# Host scalars
# --------------------------------------------------------------------------
def hubble(a: float, omega_m: float = 0.3, omega_lambda: float = 0.7) -> float:
    return lib().hy_hubble(a, omega_m, omega_lambda)


def geometric_mean(lhs: float, rhs: float) -> float:
    return lib().hy_geometric_mean(lhs, rhs)


def gravitational_constant() -> float:
    return lib().hy_gravitational_constant()


def rho_crit0() -> float:
    return lib().hy_rho_crit0()


class DriftKickIntegrator:
    """cosmology::DriftKickIntegrator (src/cosmology/drift_kick_factors.cpp:91-141):
    D = int da/(a^3 H), K = int da/(a^2 H) resolved to binary64 spacing."""

    def __init__(self, omega_m: float = 0.3, omega_lambda: float = 0.7):
        self.omega_m = omega_m
        self.omega_lambda = omega_lambda

    def drift_factor(self, a1: float, a2: float) -> float:
        out = C.c_double()
        check(lib().hy_drift_factor(a1, a2, self.omega_m, self.omega_lambda, C.byref(out)))
        return out.value

    def kick_factor(self, a1: float, a2: float) -> float:
        out = C.c_double()
        check(lib().hy_kick_factor(a1, a2, self.omega_m, self.omega_lambda, C.byref(out)))
        return out.value


@dataclass(frozen=True)
class TimeStep:
    """time::TimeStep (src/time/time_stepper.cpp:148-151): half step is the
    correctly rounded geometric mean of the endpoints."""
    a_begin: float
    a_end: float

    @property
    def a_half(self) -> float:
        return geometric_mean(min(self.a_begin, self.a_end), max(self.a_begin, self.a_end))


class TimeStepper:
    """time::TimeStepper boundary table (src/time/time_stepper.cpp:194-310)."""

    def __init__(self, a_start: float, a_final: float, delta_ln_a: float,
                 stop_scale_factors: Sequence[float] = ()):
        stops = np.ascontiguousarray(list(stop_scale_factors), dtype=np.float64)
        cap = 1 << 20
        out = np.zeros(cap, np.float64)
        cnt = C.c_uint64()
        pd = C.POINTER(C.c_double)
        check(lib().hy_time_boundaries(a_start, a_final, abs(delta_ln_a),
                                       stops.ctypes.data_as(pd) if len(stops) else None,
                                       len(stops), out.ctypes.data_as(pd), cap, C.byref(cnt)))
        self.boundaries = out[: cnt.value].copy()
        self.current_step = 0

    @property
    def num_steps(self) -> int:
        return len(self.boundaries) - 1

    def next_step(self) -> Optional[TimeStep]:
        if self.current_step >= self.num_steps:
            return None
        s = TimeStep(float(self.boundaries[self.current_step]),
                     float(self.boundaries[self.current_step + 1]))
        self.current_step += 1
        return s

    def __iter__(self):
        while (s := self.next_step()) is not None:
            yield s


# --------------------------------------------------------------------------
# PM solver
def _tuning_dict(t) -> dict:
    nv = int(t.variants)
    return {"variant": {FFT_PASS_NAMES[p]: int(t.variant[p]) for p in range(HY_FFT_PASSES)},
            "plane_group": int(t.plane_group),
            "rows": {FFT_PASS_NAMES[p]: [int(t.rows[p][v]) for v in range(nv)]
                     for p in range(HY_FFT_PASSES)},
            "pass_ms": {FFT_PASS_NAMES[p]: [round(float(t.pass_ms[p][v]), 4) for v in range(nv)]
                        for p in range(HY_FFT_PASSES)},
            "group_ms": {g: round(float(t.group_ms[i]), 4) for i, g in enumerate((0, 16, 64, 128))}}


# --------------------------------------------------------------------------
@dataclass
class PMForceDiagnostics:
    """gravity::PMForceDiagnostics (include/cosmo_nbody/gravity/pm_solver.hpp:30-35):
    sum_i 0.5 m_i phi_i before the cosmological 1/a conversion."""
    potential_energy_comoving: float
    cic_self_energy_comoving: Optional[float] = None


class PMSolver:
    """gravity::PMSolver (include/cosmo_nbody/gravity/pm_solver.hpp:53-70),
    pure PM (discrete-Laplacian Green function, centered-gradient gather).

    precision: "fp64" (binary64 everywhere) or "fp32" (binary32 mesh, FFT,
    forces and momenta; binary64 positions and time integrals).

    method: "pure_pm" (default) or "treepm_long_range" with split_scale r_s
    (mesh::PMForceMethod::treepm_long_range: continuum Green x
    exp(-k^2 r_s^2), CIC compensation, spectral gradient; compute_forces,
    force_meshes and interpolate only)."""

    METHODS = {"pure_pm": HY_METHOD_PURE_PM, "treepm_long_range": HY_METHOD_TREEPM_LONG_RANGE}

    def __init__(self, box_size: float, mesh_size: int, deconvolve_cic: bool = False,
                 precision: str = "fp32", max_particles: int = 0, device: int = 0,
                 stream=None, boxes: int = 1, method: str = "pure_pm", split_scale: float = 0.0):
        self.box_size = float(box_size)
        self.mesh_size = int(mesh_size)
        self.precision = PRECISIONS[precision]
        self.method = method
        if stream is None:
            # run on the caller's current torch stream when it is not the
            # default stream (the private stream is ordered with the default one)
            try:
                torch = _torch()
                if torch.cuda.is_available():
                    cs = torch.cuda.current_stream(device).cuda_stream
                    stream = cs or None
            except ImportError:
                pass
        d = hy_pm_desc(box_size=self.box_size, mesh_size=self.mesh_size,
                       deconvolve_cic=int(bool(deconvolve_cic)), precision=self.precision,
                       device=device, stream=stream, max_particles=int(max_particles),
                       boxes=int(boxes), method=self.METHODS[method], split_scale=float(split_scale))
        h = C.c_void_p()
        check(lib().hy_pm_create(C.byref(d), C.byref(h)))
        self._h = h
        self.device = device
        self.boxes = int(boxes)

    # -- lifetime
    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().hy_pm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def info(self) -> hy_pm_info:
        inf = hy_pm_info()
        check(lib().hy_pm_get_info(self._h, C.byref(inf)))
        return inf

    # -- FFT launch plan (FFTW_MEASURE counterpart)
    def tuning(self) -> dict:
        t = hy_fft_tuning()
        check(lib().hy_pm_get_tuning(self._h, C.byref(t)))
        return _tuning_dict(t)

    def set_tuning(self, variants, plane_group: int = 0) -> None:
        t = hy_fft_tuning()
        check(lib().hy_pm_get_tuning(self._h, C.byref(t)))
        for p, v in enumerate(variants):
            t.variant[p] = int(v)
        t.plane_group = int(plane_group)
        check(lib().hy_pm_set_tuning(self._h, C.byref(t)))

    def autotune(self, reps: int = 3) -> dict:
        """Times every compiled launch variant of each FFT pass on this
        context's buffers and keeps the fastest (results are unchanged)."""
        t = hy_fft_tuning()
        check(lib().hy_pm_autotune(self._h, int(reps), C.byref(t)))
        return _tuning_dict(t)

    def autotune_gather(self, x, y, z, uniform_mass: float, reps: int = 3) -> dict:
        """Times the gather of the potential of these positions under each
        lattice column-tile size (lattice rows per tile; 0 = particle order)
        and keeps the fastest (results are unchanged)."""
        from ._lib import GATHER_TUNE_ROWS, HY_GATHER_TUNE_CANDIDATES
        rows = C.c_uint32(0)
        ms = (C.c_float * HY_GATHER_TUNE_CANDIDATES)()
        check(lib().hy_pm_autotune_gather(self._h, len(x), _ptr(x), _ptr(y), _ptr(z), float(uniform_mass),
                                          int(reps), C.byref(rows), ms))
        return {"rows": int(rows.value), "ms": {r: float(v) for r, v in zip(GATHER_TUNE_ROWS, ms)}}

    def set_gather_rows(self, rows: Optional[int]) -> None:
        """Lattice rows per gather column tile (0 = particle order, None = default)."""
        from ._lib import HY_GATHER_ROWS_DEFAULT
        check(lib().hy_pm_set_gather_rows(self._h, HY_GATHER_ROWS_DEFAULT if rows is None else int(rows)))

    @property
    def mesh_dtype(self):
        return np.float64 if self.precision == HY_PREC_FP64 else np.float32

    def _torch_mesh_dtype(self):
        torch = _torch()
        return torch.float64 if self.precision == HY_PREC_FP64 else torch.float32

    # -- reference-shaped force entry
    def compute_forces(self, x, y, z, masses, uniform_mass, acc_x, acc_y, acc_z,
                       collect_potential_energy: bool = False) -> Optional["PMForceDiagnostics"]:
        """Adds the PM acceleration into acc_* (PMSolver::compute_forces); with
        collect_potential_energy returns PMForceDiagnostics (0.5 sum m phi),
        else None."""
        n = len(x)
        host = _is_host(x)
        flags = HY_HOST_POINTERS if host else 0
        if collect_potential_energy:
            flags |= HY_COLLECT_POTENTIAL
        um = float(uniform_mass) if uniform_mass is not None else 0.0
        check(lib().hy_pm_force(self._h, n, _ptr(x), _ptr(y), _ptr(z), _ptr(masses), um,
                                _ptr(acc_x), _ptr(acc_y), _ptr(acc_z), flags))
        if not collect_potential_energy:
            return None
        e = C.c_double()
        check(lib().hy_pm_potential_energy(self._h, C.byref(e)))
        return PMForceDiagnostics(e.value)

    # -- stages (device tensors)
    def density(self, x, y, z, masses=None, uniform_mass: float = 1.0):
        torch = _torch()
        N = self.mesh_size
        out = torch.empty((N, N, N), dtype=self._torch_mesh_dtype(), device=x.device)
        check(lib().hy_pm_density(self._h, len(x), _ptr(x), _ptr(y), _ptr(z), _ptr(masses),
                                  float(uniform_mass), _ptr(out)))
        return out

    def potential(self, source):
        torch = _torch()
        src = source.to(self._torch_mesh_dtype()).contiguous()
        out = torch.empty_like(src)
        check(lib().hy_pm_potential(self._h, _ptr(src), _ptr(out)))
        return out

    def force_meshes(self, source):
        """TreePM long-range force meshes (F_x, F_y, F_z) of a density-source mesh."""
        torch = _torch()
        src = source.to(self._torch_mesh_dtype()).contiguous()
        out = [torch.empty_like(src) for _ in range(3)]
        check(lib().hy_pm_force_meshes(self._h, _ptr(src), *(_ptr(o) for o in out)))
        return out

    def interpolate(self, mesh, x, y, z, out) -> None:
        """out += CIC interpolation of mesh at the positions (CICMassAssignment::interpolate_add)."""
        m = mesh.to(self._torch_mesh_dtype()).contiguous()
        check(lib().hy_pm_interpolate(self._h, len(x), _ptr(x), _ptr(y), _ptr(z), _ptr(m), _ptr(out)))

    def gather(self, phi, x, y, z, acc_x, acc_y, acc_z) -> None:
        p = phi.to(self._torch_mesh_dtype()).contiguous()
        check(lib().hy_pm_gather(self._h, len(x), _ptr(x), _ptr(y), _ptr(z), _ptr(p),
                                 _ptr(acc_x), _ptr(acc_y), _ptr(acc_z)))

    def bucket_order(self, x, y, z):
        torch = _torch()
        nb = self.info().bucket_count
        order = torch.empty(len(x), dtype=torch.int32, device=x.device)
        offsets = torch.empty(nb + 1, dtype=torch.int32, device=x.device)
        check(lib().hy_pm_bucket_order(self._h, len(x), _ptr(x), _ptr(y), _ptr(z), _ptr(order),
                                       _ptr(offsets)))
        return order, offsets

    def xplane_order(self, x):
        torch = _torch()
        order = torch.empty(len(x), dtype=torch.int32, device=x.device)
        offsets = torch.empty(self.mesh_size + 1, dtype=torch.int32, device=x.device)
        check(lib().hy_pm_xplane_order(self._h, len(x), _ptr(x), _ptr(order), _ptr(offsets)))
        return order, offsets

    def kdk(self, src: "ParticleState", dst: "ParticleState", uniform_mass: float,
            k_prev: Optional[float], k_next: Optional[float], d_next: Optional[float],
            masses=None, acc=None, asynchronous: bool = False, reuse_keys: bool = False) -> None:
        """Force at src positions fused with the K2 kick of the previous step
        (k_prev), the K1 kick of the next step (k_next) and its drift (d_next).

        reuse_keys: src holds the previous kdk call's output positions, so the
        deposit bucket keys made by that call are reused (they are checked
        against the positions, and recomputed on a mismatch unless
        asynchronous)."""
        flags = HY_REUSE_KEYS if reuse_keys else 0
        if k_prev is not None:
            flags |= HY_HAS_PREV
        if k_next is not None:
            flags |= HY_HAS_NEXT
        if asynchronous:
            flags |= HY_ASYNC
        ax, ay, az = acc if acc is not None else (None, None, None)
        check(lib().hy_pm_kdk(self._h, src.n, _ptr(src.x), _ptr(src.y), _ptr(src.z),
                              _ptr(src.px), _ptr(src.py), _ptr(src.pz), _ptr(dst.x), _ptr(dst.y),
                              _ptr(dst.z), _ptr(dst.px), _ptr(dst.py), _ptr(dst.pz),
                              _ptr(masses), float(uniform_mass),
                              float(k_prev or 0.0), float(k_next or 0.0), float(d_next or 0.0),
                              _ptr(ax), _ptr(ay), _ptr(az), flags))

    def check(self) -> None:
        check(lib().hy_pm_check(self._h))

    def power_spectrum(self, x, y, z, num_bins: int, subtract_shot_noise, interlaced,
                       max_k_fraction_nyquist: float, masses=None, uniform_mass: float = 1.0):
        """Shell-averaged P(k) of the CIC density contrast on this context's
        mesh (MatterPowerSpectrum::compute_pk).  subtract_shot_noise and
        interlaced must be given explicitly (True/False); None is "not
        specified".  Returns a list of hy_pk_bin."""
        def explicit(v):
            return -1 if v is None else int(bool(v))
        o = hy_pk_options(num_bins=int(num_bins), subtract_shot_noise=explicit(subtract_shot_noise),
                          interlaced=explicit(interlaced),
                          max_k_fraction_nyquist=float(max_k_fraction_nyquist))
        nb = max(int(num_bins), 0)
        out = (hy_pk_bin * max(nb, 1))()
        flags = HY_HOST_POINTERS if _is_host(x) else 0
        check(lib().hy_pm_power_spectrum(self._h, len(x), _ptr(x), _ptr(y), _ptr(z), _ptr(masses),
                                         float(uniform_mass), C.byref(o), out, flags))
        return list(out)[:nb]


class ParticleState:
    """Device-resident SoA particles: binary64 positions, momenta in the
    solver precision (core::ParticleStore, include/cosmo_nbody/core/particle_store.hpp)."""

    def __init__(self, x, y, z, px, py, pz):
        self.x, self.y, self.z, self.px, self.py, self.pz = x, y, z, px, py, pz
        self.n = len(x)

    @classmethod
    def from_numpy(cls, pos, mom, precision: str = "fp32", device: str = "cuda"):
        torch = _torch()
        mdt = torch.float64 if PRECISIONS[precision] == HY_PREC_FP64 else torch.float32
        t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), device=device).to(dt).contiguous()
        return cls(t(pos[0], torch.float64), t(pos[1], torch.float64), t(pos[2], torch.float64),
                   t(mom[0], mdt), t(mom[1], mdt), t(mom[2], mdt))

    def empty_like(self) -> "ParticleState":
        torch = _torch()
        return ParticleState(*(torch.empty_like(a) for a in
                               (self.x, self.y, self.z, self.px, self.py, self.pz)))

    def to_numpy(self):
        pos = np.stack([a.double().cpu().numpy() for a in (self.x, self.y, self.z)])
        mom = np.stack([a.double().cpu().numpy() for a in (self.px, self.py, self.pz)])
        return pos, mom


class LeapfrogIntegrator:
    """time::LeapfrogIntegrator (src/time/leapfrog_integrator.cpp:143-279) over a
    device-resident ParticleState, one force refresh per step.

    The K2 kick of a step needs the force at a_end, which is also the force the
    next step's K1 kick uses; this integrator applies that K2 inside the next
    force refresh (or in finalize()), so each refresh is a single fused pass.
    After finalize() the state equals the reference's state after its last
    step."""

    def __init__(self, solver: PMSolver, factors: DriftKickIntegrator, uniform_mass: float,
                 masses=None):
        self.solver = solver
        self.factors = factors
        self.uniform_mass = uniform_mass
        self.masses = masses
        self._pending_k2: Optional[float] = None
        self._spare: Optional[ParticleState] = None
        self._last_out: Optional[ParticleState] = None

    def step(self, state: ParticleState, step: TimeStep, asynchronous: bool = False) -> ParticleState:
        D = self.factors.drift_factor(step.a_begin, step.a_end)
        K1 = self.factors.kick_factor(step.a_begin, step.a_half)
        K2 = self.factors.kick_factor(step.a_half, step.a_end)
        out = self._spare if self._spare is not None else state.empty_like()
        self.solver.kdk(state, out, self.uniform_mass, self._pending_k2, K1, D,
                        masses=self.masses, asynchronous=asynchronous,
                        reuse_keys=state is self._last_out)
        self._spare = state
        self._pending_k2 = K2
        self._last_out = out
        return out

    def finalize(self, state: ParticleState, acc=None) -> ParticleState:
        out = self._spare if self._spare is not None else state.empty_like()
        self.solver.kdk(state, out, self.uniform_mass, self._pending_k2, None, None,
                        masses=self.masses, acc=acc, reuse_keys=state is self._last_out)
        self._pending_k2 = None
        self._spare = state
        self._last_out = None
        # positions are unchanged by a pure kick; copy them across
        out.x.copy_(state.x)
        out.y.copy_(state.y)
        out.z.copy_(state.z)
        return out

    def run(self, state: ParticleState, stepper: TimeStepper, acc=None) -> ParticleState:
        for s in stepper:
            state = self.step(state, s)
        return self.finalize(state, acc=acc)
