"""Hand-written adjoint of the PM step (hy_pm_kdk_vjp, hy_pm_density_vjp)
checked against central finite differences of the forward HIP path, and the
autograd wrappers (checkpointed recomputation) on a reduced differentiable run."""
import math

import numpy as np
import pytest

from conftest import uniform_particles

pytestmark = pytest.mark.gpu


def _away_from_faces(pos, L, N, h, count, seed):
    """Particle indices whose coordinates all sit at least 4h from a cell face."""
    dx = L / N
    f = np.mod(pos / dx, 1.0)
    ok = np.all((f > 4 * h / dx) & (f < 1 - 4 * h / dx), axis=0)
    idx = np.flatnonzero(ok)
    return np.random.default_rng(seed).choice(idx, size=count, replace=False)


def test_force_vjp_matches_finite_differences(gpu):
    import torch
    import hyowon_amd as H
    from hyowon_amd._lib import HY_HAS_PREV, check, lib
    from hyowon_amd.pm import _ptr
    N, L, n = 16, 32.0, 600
    pos = uniform_particles(n, L, 77)
    w = np.random.default_rng(5).normal(0, 1, (3, n))
    m = 1.7
    s = H.PMSolver(L, N, precision="fp64", max_particles=n)

    def F(p):
        acc = [torch.zeros(n, dtype=torch.float64, device=gpu) for _ in range(3)]
        s.compute_forces(*(torch.as_tensor(c, device=gpu) for c in p), None, m, *acc)
        g = np.stack([a.cpu().numpy() for a in acc])
        return float((w * g).sum())

    x, y, z = (torch.as_tensor(c, device=gpu) for c in pos)
    lx, ly, lz = (torch.zeros(n, dtype=torch.float64, device=gpu) for _ in range(3))
    lp = [torch.as_tensor(c.copy(), device=gpu) for c in w]
    check(lib().hy_pm_kdk_vjp(s.handle, n, _ptr(x), _ptr(y), _ptr(z), m, _ptr(lx), _ptr(ly), _ptr(lz),
                              _ptr(lp[0]), _ptr(lp[1]), _ptr(lp[2]), 1.0, 0.0, 0.0, HY_HAS_PREV))
    grad = np.stack([t.cpu().numpy() for t in (lx, ly, lz)])
    assert np.array_equal(np.stack([t.cpu().numpy() for t in lp]), w)  # dL/dp_in = u = w
    h = 1e-4 * (L / N)
    for j in _away_from_faces(pos, L, N, h, 8, 1):
        for a in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[a, j] += h
            pm[a, j] -= h
            fd = (F(pp) - F(pm)) / (2 * h)
            assert abs(fd - grad[a, j]) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(grad).max()), (j, a, fd, grad[a, j])


def test_density_vjp_matches_finite_differences(gpu):
    import torch
    import hyowon_amd as H
    from hyowon_amd._lib import check, lib
    from hyowon_amd.pm import _ptr
    N, L, n = 16, 32.0, 600
    pos = uniform_particles(n, L, 78)
    m = 0.9
    s = H.PMSolver(L, N, precision="fp64", max_particles=n)
    u = np.random.default_rng(6).normal(0, 1, (N, N, N))
    dx3 = (L / N) ** 3

    def F(p):
        src = s.density(*(torch.as_tensor(c, device=gpu) for c in p), None, m).cpu().numpy()
        mc = src * dx3 + m * n / N ** 3  # cell masses from the density source
        return float((u * mc).sum())

    x, y, z = (torch.as_tensor(c, device=gpu) for c in pos)
    lam = [torch.zeros(n, dtype=torch.float64, device=gpu) for _ in range(3)]
    ut = torch.as_tensor(u, device=gpu)
    check(lib().hy_pm_density_vjp(s.handle, n, _ptr(x), _ptr(y), _ptr(z), m, _ptr(ut), 1.0,
                                  *(_ptr(t) for t in lam)))
    grad = np.stack([t.cpu().numpy() for t in lam])
    h = 1e-4 * (L / N)
    for j in _away_from_faces(pos, L, N, h, 8, 2):
        for a in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[a, j] += h
            pm[a, j] -= h
            fd = (F(pp) - F(pm)) / (2 * h)
            assert abs(fd - grad[a, j]) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(grad).max()), (j, a, fd, grad[a, j])


def _reduced_setup(prec, gpu):
    import torch
    import hyowon_amd as H
    from hyowon_amd import ic
    from hyowon_amd.ensemble import schedule
    n_side, N, L, steps = 16, 32, 32.0, 4
    a0 = 1.0 / 20.0
    K1, D, K2 = schedule(a0, 1.0, steps)
    s = H.PMSolver(L, N, precision=prec, max_particles=n_side ** 3)
    noise = ic.white_noise(n_side, 7, gpu).to(torch.float64)
    target = ic.white_noise(N, 8, gpu).to(torch.float64) * 0.1
    return s, noise, target, L, a0, K1, D, K2


def _loss_grad(prec, gpu, every=2):
    import torch
    from hyowon_amd.adjoint import density_loss
    s, noise, target, L, a0, K1, D, K2 = _reduced_setup(prec, gpu)
    nz = noise.clone().requires_grad_(True)
    loss = density_loss(nz, target, s, L, a0, K1, D, K2, checkpoint_every=every, precision=prec)
    loss.backward()
    return float(loss), nz.grad.detach().cpu().numpy(), (s, noise, target, L, a0, K1, D, K2)


def test_adjoint_gradient_matches_finite_differences_fp64(gpu):
    """The loss is piecewise smooth (CIC weights have kinks at cell faces);
    with eps = 1e-4 no particle crosses a face within the differencing
    interval at this size and the discrete adjoint must agree to roundoff
    (eps = 1e-3 crossings give ~1e-3 deviations that shrink linearly with eps)."""
    import torch
    from hyowon_amd.adjoint import density_loss
    loss, grad, (s, noise, target, L, a0, K1, D, K2) = _loss_grad("fp64", gpu)
    eps = 1e-4
    rng = np.random.default_rng(11)
    flat = noise.reshape(-1)
    idx = rng.choice(flat.numel(), size=12, replace=False)
    errs = []
    for i in idx:
        lp = flat.clone()
        lp[i] += eps
        lm = flat.clone()
        lm[i] -= eps
        with torch.no_grad():
            Lp = float(density_loss(lp.view_as(noise), target, s, L, a0, K1, D, K2, precision="fp64"))
            Lm = float(density_loss(lm.view_as(noise), target, s, L, a0, K1, D, K2, precision="fp64"))
        fd = (Lp - Lm) / (2 * eps)
        g = grad.reshape(-1)[i]
        errs.append(abs(fd - g) / max(abs(fd), 1e-12))
        print(f"component {i}: adjoint {g:.8e}  central FD {fd:.8e}")
    assert np.median(errs) < 1e-6 and max(errs) < 1e-5, errs


def test_adjoint_checkpoint_interval_is_bit_identical(gpu):
    _, g1, _ = _loss_grad("fp64", gpu, every=1)
    _, g4, _ = _loss_grad("fp64", gpu, every=3)
    assert np.array_equal(g1, g4)


def test_fp32_adjoint_agrees_with_fp64_reference(gpu):
    l32, g32, _ = _loss_grad("fp32", gpu)
    l64, g64, _ = _loss_grad("fp64", gpu)
    assert abs(l32 - l64) <= 1e-4 * abs(l64)
    rel = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    print(f"FP32 vs FP64 gradient: relative L2 difference {rel:.3e}")
    assert rel < 1e-3, rel


def test_adjoint_gradient_matches_oracle_finite_differences(gpu, oracle_mod):
    """An FP64 reference gradient independent of the adjoint code: central
    differences of the oracle's binary64 forward (its KDK step loop and its
    CIC deposit, src/time/leapfrog_integrator.cpp:143-279 and
    src/mesh/mass_assignment_deposit.inc:15-334 restated) on the reduced run,
    against the GPU adjoint gradient.  The oracle steps one boundary interval
    per call (the same arithmetic as one call over all of them), so the CIC
    base index floor(x / dx) of every particle is seen at every step: per
    component the largest eps of a ladder whose +-eps runs keep every base
    index (no kink of the piecewise-linear CIC weights inside the interval)
    is used."""
    from hyowon_amd import ic
    import hyowon_amd as H
    _, grad, (s, noise, target, L, a0, K1, D, K2) = _loss_grad("fp64", gpu)
    n_side, N = noise.shape[0], s.mesh_size
    steps = len(K1)
    b = H.TimeStepper(a0, 1.0, math.log(1.0 / a0) / steps, [1.0]).boundaries
    m = ic.particle_mass(n_side, L)
    tgt = target.cpu().numpy()
    mean = m * n_side ** 3 / N ** 3

    def oracle_run(v):
        pos, mom = ic.lpt_from_noise(v, L, a0)
        st = [t.numpy() for t in (*pos, *mom)]
        corners = [np.floor(np.stack(st[:3]) * (N / L))]
        for k in range(len(b) - 1):
            st, _ = oracle_mod.run(*st, L, N, [b[k], b[k + 1]], m)
            corners.append(np.floor(np.stack(st[:3]) * (N / L)))
        mc = oracle_mod.deposit(st[0], st[1], st[2], L, N, uniform_mass=m)
        return float(((mc / mean - 1.0 - tgt) ** 2).mean()), np.stack(corners)

    flat = noise.cpu().reshape(-1)
    idx = np.random.default_rng(13).choice(flat.numel(), size=12, replace=False)
    errs = []
    for i in idx:
        for eps in (1e-4, 3e-5, 1e-5, 3e-6):
            lp, lm = flat.clone(), flat.clone()
            lp[i] += eps
            lm[i] -= eps
            (Lp, cp), (Lm, cm) = oracle_run(lp.view_as(noise)), oracle_run(lm.view_as(noise))
            if np.array_equal(cp, cm):
                break
        else:
            pytest.fail(f"component {i}: every eps of the ladder crosses a CIC kink")
        fd = (Lp - Lm) / (2 * eps)
        g = grad.reshape(-1)[i]
        errs.append(abs(fd - g) / max(abs(fd), 1e-12))
        print(f"component {i}: GPU adjoint {g:.8e}  oracle central difference {fd:.8e} (eps {eps:g})")
    assert np.median(errs) < 1e-6 and max(errs) < 1e-5, errs
