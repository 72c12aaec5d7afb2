"""Matter power spectrum on the MI355X PM kernels.

Mirror of analysis::MatterPowerSpectrum
(include/cosmo_nbody/analysis/matter_power_spectrum.hpp:15-58): options and
shell records carry the reference's field names; the estimator runs the CIC
deposit, the r2c transform and the shell sums in libhyowon_pm
(hy_pm_power_spectrum).
"""
from __future__ import annotations

from dataclasses import dataclass, fields
from typing import List, Optional

from ._lib import HY_PK_MAX_BINS
from .pm import PMSolver


@dataclass
class PowerSpectrumOptions:
    """matter_power_spectrum.hpp:30-38.  The two estimator switches are
    explicit choices: None means "not specified" and is rejected."""
    num_bins: int = 0
    subtract_shot_noise: Optional[bool] = None
    interlaced: Optional[bool] = None
    max_k_fraction_nyquist: float = 0.0


@dataclass
class PowerSpectrumBin:
    """matter_power_spectrum.hpp:15-28."""
    k_low: float = 0.0
    k_mean: float = 0.0
    k_high: float = 0.0
    p_raw: float = 0.0
    p_corrected: float = 0.0
    mean_cic_power_deconvolution_factor: float = 0.0
    shot_noise_reference: float = 0.0
    mode_count: int = 0


class MatterPowerSpectrum:
    """MatterPowerSpectrum(domain, mesh_size): box_size is the periodic
    domain's edge.  The mesh sizes are those of the PM kernels (2^a from 8 to
    1024, 3*2^a from 12 to 384).  precision "fp64" (binary64 deposit and
    transform) or "fp32" (binary32 mesh and transform, binary64 shell sums)."""

    MAX_BINS = HY_PK_MAX_BINS

    def __init__(self, box_size: float, mesh_size: int, precision: str = "fp64", device: int = 0,
                 stream=None):
        if int(mesh_size) < 2:
            raise ValueError("MatterPowerSpectrum mesh_size must be >= 2")
        self.box_size = float(box_size)
        self.mesh_size = int(mesh_size)
        self.precision = precision
        self.device = device
        self.stream = stream
        self._solver: Optional[PMSolver] = None
        self._capacity = 0

    def _context(self, n: int) -> PMSolver:
        if self._solver is None or n > self._capacity:
            if self._solver is not None:
                self._solver.close()
            self._capacity = max(int(n), 1)
            self._solver = PMSolver(self.box_size, self.mesh_size, precision=self.precision,
                                    max_particles=self._capacity, device=self.device,
                                    stream=self.stream)
        return self._solver

    def compute_pk(self, x, y, z, options: PowerSpectrumOptions, masses=None,
                   uniform_mass: Optional[float] = None) -> List[PowerSpectrumBin]:
        """Positions as numpy arrays (host) or CUDA tensors (device); masses
        None selects uniform_mass (required then)."""
        n = len(x)
        ctx = self._context(n)
        um = float(uniform_mass) if uniform_mass is not None else float("nan")
        raw = ctx.power_spectrum(x, y, z, options.num_bins, options.subtract_shot_noise,
                                 options.interlaced, options.max_k_fraction_nyquist,
                                 masses=masses, uniform_mass=um)
        names = [f.name for f in fields(PowerSpectrumBin)]
        return [PowerSpectrumBin(**{k: getattr(b, k) for k in names}) for b in raw]

    def close(self) -> None:
        if self._solver is not None:
            self._solver.close()
            self._solver = None
