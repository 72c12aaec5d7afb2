"""Differentiable PM runs: torch.autograd.Function wrappers over the HIP
forward step (hy_pm_kdk) and its hand-written adjoint (hy_pm_kdk_vjp,
hy_pm_density_vjp).

PMRun integrates S log-a KDK steps of time::LeapfrogIntegrator
(src/time/leapfrog_integrator.cpp:143-279) as S+1 fused force refreshes and
keeps the state only every `checkpoint_every` refreshes; its backward pass
recomputes each segment from its checkpoint and walks the refreshes in
reverse with hy_pm_kdk_vjp.  CICDensity is the CIC density contrast of a
particle set, differentiable in the positions.  No reference counterpart: the
reference is not differentiable.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Sequence

from ._lib import HY_HAS_NEXT, HY_HAS_PREV, check, lib
from .pm import PMSolver, ParticleState, _ptr, _torch


def _calls(K1: Sequence[float], D: Sequence[float], K2: Sequence[float]):
    """(k_prev, k_next, d_next, flags) of the S+1 fused refreshes of S steps."""
    S = len(K1)
    out = []
    for c in range(S + 1):
        kp = K2[c - 1] if c > 0 else 0.0
        kn = K1[c] if c < S else 0.0
        dn = D[c] if c < S else 0.0
        flags = (HY_HAS_PREV if c > 0 else 0) | (HY_HAS_NEXT if c < S else 0)
        out.append((kp, kn, dn, flags))
    return out


def _fused(solver: PMSolver, st: ParticleState, out: ParticleState, mass: float, call) -> None:
    kp, kn, dn, flags = call
    solver.kdk(st, out, mass, kp if flags & HY_HAS_PREV else None, kn if flags & HY_HAS_NEXT else None,
               dn if flags & HY_HAS_NEXT else None)
    if not flags & HY_HAS_NEXT:  # a pure kick leaves the positions in place
        out.x.copy_(st.x)
        out.y.copy_(st.y)
        out.z.copy_(st.z)


def _clone(st: ParticleState) -> ParticleState:
    return ParticleState(*(t.clone() for t in (st.x, st.y, st.z, st.px, st.py, st.pz)))


def _make_run():
    torch = _torch()

    class PMRun(torch.autograd.Function):
        """(x, y, z, px, py, pz) at a_start -> the same after the S steps."""

        @staticmethod
        def forward(ctx, x, y, z, px, py, pz, solver, mass, K1, D, K2, checkpoint_every):
            calls = _calls(K1, D, K2)
            k = max(1, int(checkpoint_every))
            st = ParticleState(*(t.detach().contiguous() for t in (x, y, z, px, py, pz)))
            st = _clone(st)
            tmp = st.empty_like()
            checkpoints = {}
            for c, call in enumerate(calls):
                if c % k == 0:
                    checkpoints[c] = _clone(st)
                _fused(solver, st, tmp, mass, call)
                st, tmp = tmp, st
            ctx.solver, ctx.mass, ctx.calls, ctx.k = solver, mass, calls, k
            ctx.checkpoints = checkpoints
            ctx.pdtype = px.dtype
            return st.x, st.y, st.z, st.px, st.py, st.pz

        @staticmethod
        def backward(ctx, gx, gy, gz, gpx, gpy, gpz):
            solver, mass, calls, k = ctx.solver, ctx.mass, ctx.calls, ctx.k
            ref = next(iter(ctx.checkpoints.values())).x
            n = ref.numel()

            def lam(g):
                return (g.to(torch.float64).contiguous().clone() if g is not None
                        else torch.zeros(n, dtype=torch.float64, device=ref.device))
            lx, ly, lz, lpx, lpy, lpz = (lam(g) for g in (gx, gy, gz, gpx, gpy, gpz))
            starts = sorted(ctx.checkpoints)
            for s in reversed(starts):
                end = min(s + k, len(calls))
                # recompute the inputs of the refreshes s .. end-1
                st = _clone(ctx.checkpoints[s])
                tmp = st.empty_like()
                inputs: List[ParticleState] = []
                for c in range(s, end):
                    inputs.append(_clone(st))
                    _fused(solver, st, tmp, mass, calls[c])
                    st, tmp = tmp, st
                for c in reversed(range(s, end)):
                    xin = inputs[c - s]
                    kp, kn, dn, flags = calls[c]
                    check(lib().hy_pm_kdk_vjp(solver.handle, n, _ptr(xin.x), _ptr(xin.y), _ptr(xin.z),
                                              float(mass), _ptr(lx), _ptr(ly), _ptr(lz), _ptr(lpx),
                                              _ptr(lpy), _ptr(lpz), float(kp), float(kn), float(dn), flags))
            pd = ctx.pdtype
            return lx, ly, lz, lpx.to(pd), lpy.to(pd), lpz.to(pd), None, None, None, None, None, None

    class CICDensity(torch.autograd.Function):
        """Positions -> CIC density contrast delta_c = m_c / mean - 1 on the
        solver's mesh (context precision)."""

        @staticmethod
        def forward(ctx, x, y, z, solver, mass):
            n = x.numel()
            src = solver.density(x.detach(), y.detach(), z.detach(), None, mass)
            N = solver.mesh_size
            dx = solver.box_size / N
            mean = mass * n / N ** 3
            ctx.save_for_backward(x, y, z)
            ctx.solver, ctx.mass, ctx.mean = solver, mass, mean
            return src * ((dx * dx * dx) / mean)

        @staticmethod
        def backward(ctx, g):
            x, y, z = ctx.saved_tensors
            solver, mass, mean = ctx.solver, ctx.mass, ctx.mean
            n = x.numel()
            u = (g / mean).to(solver._torch_mesh_dtype()).contiguous()
            lx, ly, lz = (torch.zeros(n, dtype=torch.float64, device=x.device) for _ in range(3))
            check(lib().hy_pm_density_vjp(solver.handle, n, _ptr(x), _ptr(y), _ptr(z), float(mass), _ptr(u),
                                          1.0, _ptr(lx), _ptr(ly), _ptr(lz)))
            return lx, ly, lz, None, None

    return PMRun, CICDensity


_FUNCS = None


def functions():
    """(PMRun, CICDensity) torch.autograd.Function classes."""
    global _FUNCS
    if _FUNCS is None:
        _FUNCS = _make_run()
    return _FUNCS


def pm_run(state_tensors, solver: PMSolver, mass: float, K1, D, K2, checkpoint_every: int = 4):
    PMRun, _ = functions()
    return PMRun.apply(*state_tensors, solver, mass, list(K1), list(D), list(K2), checkpoint_every)


def cic_density(x, y, z, solver: PMSolver, mass: float):
    _, CICDensity = functions()
    return CICDensity.apply(x, y, z, solver, mass)


def density_loss(noise, target, solver: PMSolver, box: float, a_start: float, K1, D, K2,
                 checkpoint_every: int = 4, precision: str = "fp32"):
    """Config-4 loss: white noise -> 1LPT ICs -> PM run -> CIC density
    contrast -> mean squared difference to `target`."""
    from . import ic
    torch = _torch()
    pos, mom = ic.lpt_from_noise(noise, box, a_start)
    n_side = noise.shape[0]
    mass = ic.particle_mass(n_side, box)
    pdt = torch.float64 if precision == "fp64" else torch.float32
    x, y, z, px, py, pz = pm_run((pos[0], pos[1], pos[2], mom[0].to(pdt), mom[1].to(pdt), mom[2].to(pdt)),
                                 solver, mass, K1, D, K2, checkpoint_every)
    delta = cic_density(x, y, z, solver, mass)
    return ((delta.to(torch.float64) - target) ** 2).mean()
