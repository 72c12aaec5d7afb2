"""The benchmark workloads at their stated sizes, through the C ABI.

- 64^3 lattice on a 128^3 mesh, box 128 Mpc/h, Philox seed 1, z = 49 -> 0 in
  32 log-a KDK steps: the binary64 GPU run against the oracle's step loop
  (src/time/leapfrog_integrator.cpp:143-279 over src/gravity/pm_solver.cpp:
  569-852), the binary32 run against the same oracle by stated tolerances.
- 512^3 lattice (1LPT, Philox seed 1) on a 1024^3 mesh, box 512 Mpc/h:
  superposed plane waves whose potential is known in closed form
  (src/mesh/green_function.cpp:165-241: phi_m = 4 pi G / lambda(m) delta_m),
  in both precisions; the reference x-plane permutation and the deposit-bucket
  permutation bit-exact against the oracle; momentum conservation of the
  force; binary32 against binary64 force RMS.
- The 192-box ensemble (64 x 32^3 on 64^3, 64 x 64^3 on 128^3, 64 x 96^3 on
  192^3, box 2 x mesh Mpc/h, seeds 1..192, 32 steps): graph-launched steps
  equal single-box eager runs for sampled boxes of every class, and one box
  per class has the oracle's binary64 force at its initial conditions.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A_START = 1.0 / 50.0


def _state_numpy(pos, mom):
    return np.stack([p.cpu().numpy() for p in pos]), np.stack([p.double().cpu().numpy() for p in mom])


def _periodic_cells(a, b, L, N):
    d = np.abs(a - b)
    d = np.minimum(d, L - d)
    return d / (L / N)


# ---------------------------------------------------------------------------
# 64^3 particles, 128^3 mesh, 32 steps
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def lattice64(gpu, oracle_mod):
    import hyowon_amd as H
    from hyowon_amd import ic
    n_side, N, L, steps = 64, 128, 128.0, 32
    pos, mom = ic.lpt_ics(n_side, L, A_START, seed=1, device=gpu, rng="philox", max_mode=n_side // 2 - 1)
    p0, q0 = _state_numpy(pos, mom)
    stepper = H.TimeStepper(A_START, 1.0, math.log(1.0 / A_START) / steps, [1.0])
    assert stepper.num_steps == steps
    m = ic.particle_mass(n_side, L)
    (rx, ry, rz, rpx, rpy, rpz), _ = oracle_mod.run(p0[0], p0[1], p0[2], q0[0], q0[1], q0[2], L, N,
                                                    stepper.boundaries, m)
    return dict(n_side=n_side, N=N, L=L, steps=steps, p0=p0, q0=q0, m=m, boundaries=stepper.boundaries,
                ref_pos=np.stack([rx, ry, rz]), ref_mom=np.stack([rpx, rpy, rpz]))


def _gpu_run(gpu, case, prec):
    import hyowon_amd as H
    n = case["n_side"] ** 3
    s = H.PMSolver(case["L"], case["N"], precision=prec, max_particles=n)
    st = H.ParticleState.from_numpy(case["p0"], case["q0"], precision=prec)
    stepper = H.TimeStepper(A_START, 1.0, math.log(1.0 / A_START) / case["steps"], [1.0])
    integ = H.LeapfrogIntegrator(s, H.DriftKickIntegrator(0.3, 0.7), case["m"])
    st = integ.run(st, stepper)
    gpos, gmom = st.to_numpy()
    s.close()
    return gpos, gmom


def test_lattice64_mesh128_fp64_run_matches_oracle(gpu, lattice64):
    gpos, gmom = _gpu_run(gpu, lattice64, "fp64")
    c = lattice64
    dcell = _periodic_cells(gpos, c["ref_pos"], c["L"], c["N"])
    dmom = np.abs(gmom - c["ref_mom"]).max() / np.abs(c["ref_mom"]).max()
    print(f"64^3/128^3 32-step binary64 run vs oracle: max position error {dcell.max():.3e} cells, "
          f"max momentum error {dmom:.3e} of max |p|")
    assert dcell.max() < 1e-9
    assert dmom < 1e-9


def test_lattice64_mesh128_fp32_run_tolerance(gpu, lattice64):
    gpos, gmom = _gpu_run(gpu, lattice64, "fp32")
    c = lattice64
    dcell = _periodic_cells(gpos, c["ref_pos"], c["L"], c["N"])
    rms = math.sqrt(float((dcell ** 2).sum(0).mean()))
    pm = math.sqrt(float(((gmom - c["ref_mom"]) ** 2).sum() / (c["ref_mom"] ** 2).sum()))
    print(f"64^3/128^3 32-step binary32 run vs binary64 oracle: position RMS {rms:.3e} cells "
          f"(max {dcell.max():.3e}), momentum RMS relative {pm:.3e}")
    assert rms < 1e-3
    assert pm < 1e-3


# ---------------------------------------------------------------------------
# 512^3 particles, 1024^3 mesh
# ---------------------------------------------------------------------------
N2, L2, NS2 = 1024, 512.0, 512


def _modes_field(gpu, modes, N, dtype):
    """sum_j a_j cos(2 pi (m_j . i) / N + theta_j) on the N^3 mesh (float64 on the GPU)."""
    import torch
    i = torch.arange(N, device=gpu, dtype=torch.int64)
    out = torch.zeros((N, N, N), dtype=torch.float64, device=gpu)
    for (mx, my, mz, amp, th) in modes:
        ph = ((mx * i).view(N, 1, 1) + (my * i).view(1, N, 1) + (mz * i).view(1, 1, N)) % N
        out += amp * torch.cos(ph.to(torch.float64) * (2.0 * math.pi / N) + th)
        del ph
    return out.to(dtype)


def _green_factor(m, N, L):
    """4 pi G / (lambda_x + lambda_y + lambda_z), lambda = (-4 (1/dx)^2 s) s, s = sin(pi m / N)."""
    import hyowon_amd as H
    dx = L / N
    lam = 0.0
    for ma in m:
        s = math.sin(math.pi * ma / N)
        lam += (-4.0 * (1.0 / dx) ** 2 * s) * s
    return 4.0 * math.pi * H.gravitational_constant() / lam


MODES = [(1, 0, 0, 1.0, 0.0), (3, 5, 7, 0.7, 0.4), (511, 2, 1, 0.3, 1.3), (0, 0, 512, 0.2, 0.0),
         (17, -200, 333, 0.5, -0.8), (512, 512, 511, 0.1, 2.0)]


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-11), ("fp32", 2e-6)])
def test_mesh1024_plane_waves_closed_form(gpu, prec, tol):
    import torch
    import hyowon_amd as H
    s = H.PMSolver(L2, N2, precision=prec, max_particles=1)
    dt = torch.float64 if prec == "fp64" else torch.float32
    src = _modes_field(gpu, MODES, N2, dt)
    phi = s.potential(src)
    del src
    s.close()
    expect = _modes_field(gpu, [(mx, my, mz, a * _green_factor((mx, my, mz), N2, L2), th)
                                for (mx, my, mz, a, th) in MODES], N2, torch.float64)
    d = phi.to(torch.float64) - expect
    del phi
    if prec == "fp64":
        err = float(d.abs().max() / expect.abs().max())
    else:
        err = float(torch.sqrt((d * d).mean() / (expect * expect).mean()))
    print(f"1024^3 plane waves, {prec}: error {err:.3e}")
    assert err < tol


@pytest.fixture(scope="module")
def lattice512(gpu):
    from hyowon_amd import ic
    pos, _ = ic.lpt_ics(NS2, L2, A_START, seed=1, device=gpu, rng="philox", max_mode=NS2 // 2 - 1)
    return pos, ic.particle_mass(NS2, L2)


def test_mesh1024_permutations_bit_exact(gpu, oracle_mod, lattice512):
    import hyowon_amd as H
    pos, _ = lattice512
    n = pos[0].numel()
    s = H.PMSolver(L2, N2, precision="fp32", max_particles=n)
    host = [p.cpu().numpy() for p in pos]
    order, offsets = s.xplane_order(pos[0])
    off_ref, ord_ref = oracle_mod.xplane_order(host[0], L2, N2)
    assert np.array_equal(offsets.cpu().numpy().astype(np.uint64), off_ref)
    assert np.array_equal(order.cpu().numpy().view(np.uint32), ord_ref)
    del order, offsets, ord_ref
    order, offsets = s.bucket_order(*pos)
    off_ref, ord_ref = oracle_mod.bucket_order(host[0], host[1], host[2], L2, N2, int(s.info().tile_edge))
    assert np.array_equal(offsets.cpu().numpy().astype(np.uint64), off_ref)
    assert np.array_equal(order.cpu().numpy().view(np.uint32), ord_ref)
    s.close()


def test_mesh1024_momentum_conservation_and_fp32_rms(gpu, lattice512):
    import torch
    import hyowon_amd as H
    pos, m = lattice512
    n = pos[0].numel()
    out = {}
    for prec in ("fp64", "fp32"):
        s = H.PMSolver(L2, N2, precision=prec, max_particles=n)
        acc = [torch.zeros(n, dtype=torch.float64, device=gpu) for _ in range(3)]
        s.compute_forces(*pos, None, m, *acc)
        s.close()
        g = torch.stack(acc)
        net = float(g.sum(1).norm() / g.norm(dim=0).sum())
        print(f"1024^3 {prec}: |sum g| / sum |g| = {net:.3e}")
        assert net < (1e-10 if prec == "fp64" else 1e-6), (prec, net)
        out[prec] = g
    d = out["fp32"] - out["fp64"]
    rms = float(torch.sqrt((d * d).sum() / (out["fp64"] ** 2).sum()))
    print(f"1024^3 force, binary32 vs binary64: RMS relative {rms:.3e}")
    assert rms < 1e-4


# ---------------------------------------------------------------------------
# 192-box ensemble
# ---------------------------------------------------------------------------
SAMPLED = {0: [0, 63], 1: [5], 2: [0, 63]}


def test_ensemble192_graph_steps_match_single_boxes_and_oracle(gpu, oracle_mod):
    import torch
    import hyowon_amd as H
    from hyowon_amd import ic
    from hyowon_amd.ensemble import CONFIG3, BoxClass, EnsembleRun, class_ics
    steps = 32
    assert sum(c.boxes for c in CONFIG3) == 192
    run = EnsembleRun(CONFIG3, A_START, 1.0, steps, precision="fp32")
    # initial conditions of the sampled boxes, before the run moves them
    singles = []
    for ci, cls in enumerate(CONFIG3):
        nb = cls.n_side ** 3
        st = run.states[ci]
        for j in SAMPLED[ci]:
            one = H.ParticleState(*(t[j * nb:(j + 1) * nb].clone() for t in (st.x, st.y, st.z, st.px, st.py, st.pz)))
            singles.append((ci, j, one))
    # one box per class: binary64 force at the initial conditions against the oracle
    for ci, cls in enumerate(CONFIG3):
        one = class_ics(BoxClass(cls.n_side, cls.mesh, cls.box_size, [cls.seeds[0]]), A_START, "fp64", gpu)
        n = cls.n_side ** 3
        m = ic.particle_mass(cls.n_side, cls.box_size)
        s = H.PMSolver(cls.box_size, cls.mesh, precision="fp64", max_particles=n)
        acc = [torch.zeros(n, dtype=torch.float64, device=gpu) for _ in range(3)]
        s.compute_forces(one.x, one.y, one.z, None, m, *acc)
        s.close()
        got = np.stack([a.cpu().numpy() for a in acc])
        ref = oracle_mod.pm_force(one.x.cpu().numpy(), one.y.cpu().numpy(), one.z.cpu().numpy(), cls.box_size,
                                  cls.mesh, uniform_mass=m)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print(f"class {cls.n_side}^3/{cls.mesh}^3 box seed {cls.seeds[0]}: binary64 force vs oracle {err:.3e}")
        assert err < 1e-10
    run.run()
    torch.cuda.synchronize()
    for ci, j, one in singles:
        cls = CONFIG3[ci]
        n = cls.n_side ** 3
        s = H.PMSolver(cls.box_size, cls.mesh, precision="fp32", max_particles=n)
        a, b = one, one.empty_like()
        m = ic.particle_mass(cls.n_side, cls.box_size)
        for i in range(steps + 1):
            if i < steps:
                s.kdk(a, b, m, run.K2[i - 1] if i else None, run.K1[i], run.D[i])
                a, b = b, a
            else:
                s.kdk(a, b, m, run.K2[i - 1], None, None)
                a = H.ParticleState(a.x, a.y, a.z, b.px, b.py, b.pz)
        torch.cuda.synchronize()
        s.close()
        st = run.states[ci]
        for got_t, ref_t in zip((st.x, st.y, st.z, st.px, st.py, st.pz), (a.x, a.y, a.z, a.px, a.py, a.pz)):
            assert torch.equal(got_t[j * n:(j + 1) * n], ref_t), (ci, j)
    run.close()
