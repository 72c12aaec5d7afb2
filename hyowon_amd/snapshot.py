"""Native HYOWON snapshot files (HDF5) through libhyowon_pm
(hy_snapshot_write / hy_snapshot_read): SnapshotIO::write_snapshot_to_path and
read_snapshot (include/cosmo_nbody/io/snapshot_io.hpp:97-110) for particle
states held here.  Device tensors are copied to host memory first."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from ._lib import (HY_OK, HY_SNAPSHOT_REPLACE, _BY_CODE, HyError, hy_snapshot_header, hy_snapshot_meta, lib)

_pd = C.POINTER(C.c_double)
_pu64 = C.POINTER(C.c_uint64)


def _check(rc: int) -> None:
    if rc != HY_OK:
        raise _BY_CODE.get(rc, HyError)(rc, lib().hy_snapshot_last_error().decode(errors="replace"))


@dataclass
class SnapshotMeta:
    """Run parameters stored with a snapshot (/Header, /Parameters, /Config)."""
    box_size: float
    particles_per_dim: int
    pm_mesh_per_dim: int
    ic_mesh_per_dim: int = 0
    omega_m: float = 0.3
    omega_lambda: float = 0.7
    h: float = 0.7
    omega_b: float = 0.05
    sigma8: float = 0.8
    n_s: float = 0.96
    seed: int = 1
    lpt_order: int = 1
    softening: float = 0.0
    z_start: float = 49.0
    z_final: float = 0.0
    delta_ln_a: float = 0.0
    power_spectrum_redshift: float = 0.0
    deconvolve_cic: bool = False
    ic_amplitude_mode: str = "gaussian"
    ic_phase_pairing: str = "independent"
    power_spectrum_fidelity: str = "precision_boltzmann"

    def c_struct(self) -> hy_snapshot_meta:
        return hy_snapshot_meta(self.box_size, self.omega_m, self.omega_lambda, self.h, self.omega_b, self.sigma8,
                                self.n_s, int(self.seed), int(self.particles_per_dim), int(self.pm_mesh_per_dim),
                                int(self.ic_mesh_per_dim or self.particles_per_dim), int(self.lpt_order),
                                self.softening, self.z_start, self.z_final, self.delta_ln_a,
                                self.power_spectrum_redshift, int(bool(self.deconvolve_cic)),
                                self.ic_amplitude_mode.encode(), self.ic_phase_pairing.encode(),
                                self.power_spectrum_fidelity.encode())


def _host(a):
    if a is None:
        return None
    if hasattr(a, "detach"):
        a = a.detach().to("cpu").double().numpy()
    return np.ascontiguousarray(a, dtype=np.float64)


def write(path: str, meta: SnapshotMeta, a: float, x, y, z, px, py, pz, ids=None, masses=None,
          uniform_mass: Optional[float] = None, replace: bool = False) -> None:
    """Write a snapshot of positions (Mpc/h, in [0, L)) and momenta p = a^2 dx/dt
    (stored as velocities p / a); masses (explicit) or uniform_mass."""
    arrs = [_host(v) for v in (x, y, z, px, py, pz)]
    n = len(arrs[0])
    idv = None if ids is None else np.ascontiguousarray(ids, dtype=np.uint64)
    mv = _host(masses)
    um = float(uniform_mass) if uniform_mass is not None else 0.0
    _check(lib().hy_snapshot_write(path.encode(), C.byref(meta.c_struct()), float(a), n,
                                   *[v.ctypes.data_as(_pd) for v in arrs],
                                   idv.ctypes.data_as(_pu64) if idv is not None else None,
                                   mv.ctypes.data_as(_pd) if mv is not None else None, um,
                                   HY_SNAPSHOT_REPLACE if replace else 0))


def read_header(path: str) -> hy_snapshot_header:
    h = hy_snapshot_header()
    _check(lib().hy_snapshot_read_header(path.encode(), C.byref(h)))
    return h


def read(path: str, expect: Optional[SnapshotMeta] = None) -> dict:
    """Read a snapshot: {a, positions (3, n), momenta (3, n) = a v, ids, masses or None, uniform_mass}."""
    h = read_header(path)
    n = int(h.particles)
    out = [np.zeros(n) for _ in range(6)]
    ids = np.zeros(n, np.uint64)
    masses = np.zeros(n) if h.explicit_masses else None
    exp = None
    if expect is not None:
        exp = hy_snapshot_header(0.0, 0.0, expect.box_size, expect.omega_m, expect.omega_lambda, expect.h, 0.0, 0, 0)
    _check(lib().hy_snapshot_read(path.encode(), C.byref(exp) if exp is not None else None, n,
                                  *[v.ctypes.data_as(_pd) for v in out], ids.ctypes.data_as(_pu64),
                                  masses.ctypes.data_as(_pd) if masses is not None else None))
    return {"a": h.a, "positions": np.stack(out[:3]), "momenta": np.stack(out[3:]), "ids": ids, "masses": masses,
            "uniform_mass": h.uniform_mass if not h.explicit_masses else None}
