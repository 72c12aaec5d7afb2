"""hyowon_amd: MI355X-native particle-mesh gravity step for HYOWON.

The compute path is libhyowon_pm.so (hand-written HIP kernels for gfx950 behind
the C ABI declared in include/hyowon_pm.h).  This package loads it with ctypes
and mirrors the reference's C++ interfaces for the path:

  PMSolver            gravity::PMSolver (include/cosmo_nbody/gravity/pm_solver.hpp:53-70)
  DriftKickIntegrator cosmology::DriftKickIntegrator (drift_kick_factors.hpp:14-27)
  TimeStepper         time::TimeStepper (include/cosmo_nbody/time/time_stepper.hpp)
  LeapfrogIntegrator  time::LeapfrogIntegrator (include/cosmo_nbody/time/leapfrog_integrator.hpp)

There is no CPU fallback: importing the package without the built library
raises, and every compute call goes through the HIP library.
"""
from __future__ import annotations

from ._lib import (HY_ASYNC, HY_HAS_NEXT, HY_HAS_PREV, HY_HOST_POINTERS, HY_PREC_FP32_MIXED,
                   HY_PREC_FP64, HyError, LIB_PATH, lib)
from .pm import (DriftKickIntegrator, LeapfrogIntegrator, PMForceDiagnostics, PMSolver, ParticleState, TimeStep,
                 TimeStepper, geometric_mean, gravitational_constant, hubble, rho_crit0)

__all__ = [
    "HY_ASYNC", "HY_HAS_NEXT", "HY_HAS_PREV", "HY_HOST_POINTERS", "HY_PREC_FP32_MIXED",
    "HY_PREC_FP64", "HyError", "LIB_PATH", "lib", "DriftKickIntegrator", "LeapfrogIntegrator",
    "PMForceDiagnostics", "PMSolver", "ParticleState", "TimeStep", "TimeStepper", "geometric_mean",
    "gravitational_constant", "hubble", "rho_crit0",
]
