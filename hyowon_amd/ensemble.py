"""Batched ensembles of independent PM boxes advanced together.

Each box geometry (particles per side, mesh, box size) is one PMSolver
context holding many boxes: particle arrays concatenate the boxes, and every
kernel takes its box from the global particle, tile or row index.  The
contexts of all geometries share one stream and are advanced by one HIP graph
launch per KDK step (hy_ensemble_step), with the kick/drift factors written to
device memory ahead of each replay.  No reference counterpart: the reference
advances one box per process (src/runtime/simulation_runner.cpp:1348-1527);
each box here follows that step loop exactly.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import List, Optional, Sequence

from ._lib import HY_HAS_NEXT, HY_HAS_PREV, check, hy_ensemble_member, lib
from .pm import DriftKickIntegrator, PMSolver, ParticleState, TimeStepper, _ptr, _torch, geometric_mean


@dataclass
class BoxClass:
    """Boxes of one geometry: n_side^3 lattice particles, mesh^3 cells, side box_size."""
    n_side: int
    mesh: int
    box_size: float
    seeds: Sequence[int]

    @property
    def boxes(self) -> int:
        return len(self.seeds)

    @property
    def particles(self) -> int:
        return self.boxes * self.n_side ** 3


class Ensemble:
    """Graph-launched fused KDK steps over several contexts that share a stream.

    members: (solver, state, uniform_mass) per context; states are updated in
    place."""

    def __init__(self, members: Sequence[tuple]):
        arr = (hy_ensemble_member * len(members))()
        for i, (solver, st, m) in enumerate(members):
            arr[i] = hy_ensemble_member(ctx=solver.handle.value, n=st.n, x=_ptr(st.x), y=_ptr(st.y),
                                        z=_ptr(st.z), px=_ptr(st.px), py=_ptr(st.py), pz=_ptr(st.pz),
                                        uniform_mass=float(m))
        self._members = list(members)
        h = C.c_void_p()
        check(lib().hy_ensemble_create(arr, len(members), C.byref(h)))
        self._h = h

    def step(self, k_prev: Optional[float], k_next: Optional[float], d_next: Optional[float]) -> None:
        flags = (HY_HAS_PREV if k_prev is not None else 0) | (HY_HAS_NEXT if k_next is not None else 0)
        check(lib().hy_ensemble_step(self._h, float(k_prev or 0.0), float(k_next or 0.0),
                                     float(d_next or 0.0), flags))

    def check(self) -> None:
        check(lib().hy_ensemble_check(self._h))

    def reset(self) -> None:
        """Call after the members' positions were written outside the
        ensemble: the next step re-keys every member eagerly."""
        check(lib().hy_ensemble_reset(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().hy_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def schedule(a_start: float, a_final: float, steps: int, omega_m: float = 0.3, omega_lambda: float = 0.7):
    """K1, D, K2 per step of the log-a schedule (time::TimeStepper + DriftKickIntegrator)."""
    stepper = TimeStepper(a_start, a_final, math.log(a_final / a_start) / steps, [a_final])
    b = stepper.boundaries
    dk = DriftKickIntegrator(omega_m, omega_lambda)
    K1, D, K2 = [], [], []
    for i in range(len(b) - 1):
        ah = geometric_mean(b[i], b[i + 1])
        K1.append(dk.kick_factor(b[i], ah))
        D.append(dk.drift_factor(b[i], b[i + 1]))
        K2.append(dk.kick_factor(ah, b[i + 1]))
    return K1, D, K2


def class_ics(cls: BoxClass, a_start: float, precision: str, device):
    """Concatenated 1LPT initial conditions of a class's boxes (one seed per
    box): the reference's Philox mode-keyed field on an IC mesh of the
    lattice's size, displaced by hy_ic_lpt."""
    from . import ic
    torch = _torch()
    pos = [[], [], []]
    mom = [[], [], []]
    for seed in cls.seeds:
        p, q = ic.lpt_ics(cls.n_side, cls.box_size, a_start, seed=int(seed), device=device, rng="philox")
        for d in range(3):
            pos[d].append(p[d])
            mom[d].append(q[d])
    mdt = torch.float64 if precision == "fp64" else torch.float32
    return ParticleState(*(torch.cat(pos[d]).contiguous() for d in range(3)),
                         *(torch.cat(mom[d]).to(mdt).contiguous() for d in range(3)))


class EnsembleRun:
    """Config-3 style run: several box classes, one context per class, all
    advanced together.  step 0 (force at the ICs, K1, drift) runs eagerly per
    context so that its bucket keys exist; the steady steps and the final K2
    are graph launches."""

    def __init__(self, classes: Sequence[BoxClass], a_start: float, a_final: float, steps: int,
                 precision: str = "fp32", device: int = 0, autotune: bool = True):
        import hyowon_amd.ic as ic
        torch = _torch()
        self.classes = list(classes)
        self.K1, self.D, self.K2 = schedule(a_start, a_final, steps)
        self.steps = steps
        self.stream = torch.cuda.Stream(device=device)
        dev = torch.device("cuda", device)
        self.states: List[ParticleState] = []
        self.solvers: List[PMSolver] = []
        self.masses: List[float] = []
        for cls in self.classes:
            st = class_ics(cls, a_start, precision, dev)
            s = PMSolver(cls.box_size, cls.mesh, precision=precision, max_particles=cls.particles,
                         device=device, stream=self.stream.cuda_stream, boxes=cls.boxes)
            if autotune:
                s.autotune(reps=2)
            self.states.append(st)
            self.solvers.append(s)
            self.masses.append(ic.particle_mass(cls.n_side, cls.box_size))
        torch.cuda.synchronize(device)
        self.ensemble: Optional[Ensemble] = None
        self.done = 0

    def first_step(self) -> None:
        for s, st, m in zip(self.solvers, self.states, self.masses):
            s.kdk(st, st, m, None, self.K1[0], self.D[0])
        self.ensemble = Ensemble(list(zip(self.solvers, self.states, self.masses)))
        self.done = 1

    def next_step(self) -> None:
        i = self.done
        if i == 0:
            self.first_step()
        elif i < self.steps:
            self.ensemble.step(self.K2[i - 1], self.K1[i], self.D[i])
            self.done += 1
        elif i == self.steps:
            self.ensemble.step(self.K2[i - 1], None, None)
            self.done += 1
        else:
            raise RuntimeError("run already finished")

    def run(self) -> None:
        while self.done <= self.steps:
            self.next_step()
        self.ensemble.check()

    def close(self) -> None:
        if self.ensemble is not None:
            self.ensemble.close()
        for s in self.solvers:
            s.close()


# The build's config-3 workload: 192 boxes in three geometries, box side twice
# the mesh (in Mpc/h), seeds 1..192, z = 49 -> 0 in 32 KDK steps.
CONFIG3 = [BoxClass(32, 64, 128.0, range(1, 65)), BoxClass(64, 128, 256.0, range(65, 129)),
           BoxClass(96, 192, 384.0, range(129, 193))]
