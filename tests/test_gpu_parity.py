"""GPU parity: every stage of the HIP PM step against the CPU oracle, through the
C ABI (libhyowon_pm).  Integer outputs (bucket / x-plane permutations) and the
binary64 deposit and gather must be bit-exact; FFT-based outputs agree to
binary64 roundoff (FP64) or to the stated binary32 tolerances (FP32)."""
import ctypes as C
import math

import numpy as np
import pytest

from conftest import clumpy_particles, face_particles, uniform_particles

pytestmark = pytest.mark.gpu

G_REL_FP64 = 1e-10     # binary64 force vs oracle (FFT roundoff, plan-dependent in the reference too)
G_RMS_FP32 = 1e-4      # binary32 force RMS relative error vs binary64 (target of the build)


def _dev(a, torch_device, dtype=None):
    import torch
    t = torch.as_tensor(np.ascontiguousarray(a), device=torch_device)
    return t.to(dtype) if dtype is not None else t


def _solver(L, N, prec, n, deconv=False):
    import hyowon_amd as H
    return H.PMSolver(L, N, deconvolve_cic=deconv, precision=prec, max_particles=n)


CASES = [
    ("uniform", 16, 32.0, 4096),
    ("uniform", 32, 32.0, 20000),
    ("clumpy", 48, 48.0, 30000),
    ("uniform", 12, 24.0, 3000),
    ("clumpy", 64, 64.0, 40000),
]


def _positions(kind, N, L, n, seed):
    if kind == "uniform":
        return uniform_particles(n, L, seed)
    return clumpy_particles(n, L, seed, n_clumps=max(4, n // 800), sigma=0.5 * L / N)


@pytest.mark.parametrize("kind,N,L,n", CASES)
def test_bucket_order_bit_exact(gpu, oracle_mod, kind, N, L, n):
    pos = _positions(kind, N, L, n, seed=11)
    s = _solver(L, N, "fp32", n)
    x, y, z = (_dev(p, gpu) for p in pos)
    order, offsets = s.bucket_order(x, y, z)
    T = int(s.info().tile_edge)
    off_ref, ord_ref = oracle_mod.bucket_order(pos[0], pos[1], pos[2], L, N, T)
    assert np.array_equal(offsets.cpu().numpy().astype(np.uint64), off_ref)
    assert np.array_equal(order.cpu().numpy().astype(np.uint32), ord_ref)


@pytest.mark.parametrize("kind,N,L,n", CASES + [("faces", 16, 32.0, 0)])
def test_xplane_order_bit_exact(gpu, oracle_mod, kind, N, L, n):
    pos = face_particles(N, L, 5) if kind == "faces" else _positions(kind, N, L, n, seed=12)
    s = _solver(L, N, "fp64", pos.shape[1])
    x = _dev(pos[0], gpu)
    order, offsets = s.xplane_order(x)
    off_ref, ord_ref = oracle_mod.xplane_order(pos[0], L, N)
    assert np.array_equal(offsets.cpu().numpy().astype(np.uint64), off_ref)
    assert np.array_equal(order.cpu().numpy().astype(np.uint32), ord_ref)


@pytest.mark.parametrize("kind,N,L,n", CASES)
def test_density_fp64_bit_exact(gpu, oracle_mod, kind, N, L, n):
    pos = _positions(kind, N, L, n, seed=13)
    m = 0.75
    s = _solver(L, N, "fp64", n)
    x, y, z = (_dev(p, gpu) for p in pos)
    mesh = s.density(x, y, z, None, m).cpu().numpy()
    dep = oracle_mod.deposit(pos[0], pos[1], pos[2], L, N, uniform_mass=m)
    ref = oracle_mod.density_contrast(dep, L, N, m * n)
    assert np.array_equal(mesh, ref), np.abs(mesh - ref).max()


def test_density_fp64_faces_and_explicit_masses(gpu, oracle_mod):
    N, L = 16, 32.0
    pos = face_particles(N, L, 7)
    n = pos.shape[1]
    masses = np.random.default_rng(1).uniform(0.5, 2.0, n)
    s = _solver(L, N, "fp64", n)
    x, y, z = (_dev(p, gpu) for p in pos)
    mesh = s.density(x, y, z, _dev(masses, gpu), 0.0).cpu().numpy()
    dep = oracle_mod.deposit(pos[0], pos[1], pos[2], L, N, masses=masses)
    total = float(oracle_mod.lib().or_exact_sum(masses.ctypes.data_as(oracle_mod._pd), n))
    ref = oracle_mod.density_contrast(dep, L, N, total)
    assert np.array_equal(mesh, ref)


@pytest.mark.parametrize("kind,N,L,n", CASES[:3])
def test_density_fp32(gpu, oracle_mod, kind, N, L, n):
    pos = _positions(kind, N, L, n, seed=14)
    m = 1.3
    s = _solver(L, N, "fp32", n)
    x, y, z = (_dev(p, gpu) for p in pos)
    mesh = s.density(x, y, z, None, m).cpu().numpy().astype(np.float64)
    dep = oracle_mod.deposit(pos[0], pos[1], pos[2], L, N, uniform_mass=m)
    ref = oracle_mod.density_contrast(dep, L, N, m * n)
    scale = np.abs(ref).max()
    assert np.abs(mesh - ref).max() <= 1e-5 * scale


@pytest.mark.parametrize("N", [8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384])
@pytest.mark.parametrize("deconv", [False, True])
def test_potential_fp64(gpu, oracle_mod, N, deconv):
    L = 2.0 * N
    rng = np.random.default_rng(N)
    src = rng.standard_normal((N, N, N))
    src -= src.mean()
    s = _solver(L, N, "fp64", 1, deconv)
    phi = s.potential(_dev(src, gpu)).cpu().numpy()
    spec = oracle_mod.rfftn(src)
    spec = oracle_mod.green_apply(spec, L, deconv)
    ref = oracle_mod.irfftn(spec)
    err = np.abs(phi - ref).max() / np.abs(ref).max()
    assert err < 1e-12, err


@pytest.mark.parametrize("N", [16, 48, 64, 128, 192, 256])
def test_potential_fp32(gpu, oracle_mod, N):
    L = 2.0 * N
    rng = np.random.default_rng(N + 1)
    src = rng.standard_normal((N, N, N))
    s = _solver(L, N, "fp32", 1)
    phi = s.potential(_dev(src, gpu)).cpu().numpy().astype(np.float64)
    ref = oracle_mod.irfftn(oracle_mod.green_apply(oracle_mod.rfftn(src.astype(np.float32).astype(np.float64)), L))
    err = np.sqrt(np.mean((phi - ref) ** 2) / np.mean(ref ** 2))
    assert err < 2e-6, err


@pytest.mark.parametrize("kind,N,L,n", CASES)
def test_gather_fp64_bit_exact(gpu, oracle_mod, kind, N, L, n):
    pos = _positions(kind, N, L, n, seed=15)
    rng = np.random.default_rng(3)
    phi = rng.standard_normal((N, N, N))
    acc0 = rng.standard_normal((3, n))
    s = _solver(L, N, "fp64", n)
    import torch
    acc = [_dev(a, gpu) for a in acc0]
    x, y, z = (_dev(p, gpu) for p in pos)
    s.gather(_dev(phi, gpu), x, y, z, *acc)
    got = np.stack([a.cpu().numpy() for a in acc])
    ref = oracle_mod.gather(phi, pos[0], pos[1], pos[2], L, acc=acc0)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("kind,N,L,n", CASES)
def test_force_fp64(gpu, oracle_mod, kind, N, L, n):
    pos = _positions(kind, N, L, n, seed=16)
    m = 2.5
    s = _solver(L, N, "fp64", n)
    import torch
    acc = [torch.zeros(n, dtype=torch.float64, device=gpu) for _ in range(3)]
    x, y, z = (_dev(p, gpu) for p in pos)
    s.compute_forces(x, y, z, None, m, *acc)
    got = np.stack([a.cpu().numpy() for a in acc])
    ref = oracle_mod.pm_force(pos[0], pos[1], pos[2], L, N, uniform_mass=m)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < G_REL_FP64, err


@pytest.mark.parametrize("kind,N,L,n", CASES)
def test_force_fp32_rms(gpu, oracle_mod, kind, N, L, n):
    pos = _positions(kind, N, L, n, seed=17)
    m = 2.5
    s = _solver(L, N, "fp32", n)
    import torch
    acc = [torch.zeros(n, dtype=torch.float64, device=gpu) for _ in range(3)]
    x, y, z = (_dev(p, gpu) for p in pos)
    s.compute_forces(x, y, z, None, m, *acc)
    got = np.stack([a.cpu().numpy() for a in acc])
    ref = oracle_mod.pm_force(pos[0], pos[1], pos[2], L, N, uniform_mass=m)
    rms = math.sqrt(((got - ref) ** 2).sum() / (ref ** 2).sum())
    assert rms < G_RMS_FP32, rms


def test_force_host_pointers_and_accumulate(gpu, oracle_mod):
    N, L, n = 32, 32.0, 10000
    pos = uniform_particles(n, L, 21)
    s = _solver(L, N, "fp64", n)
    acc = np.random.default_rng(2).standard_normal((3, n))
    a0 = acc.copy()
    s.compute_forces(pos[0].copy(), pos[1].copy(), pos[2].copy(), None, 1.0, acc[0], acc[1], acc[2])
    ref = oracle_mod.pm_force(pos[0], pos[1], pos[2], L, N, uniform_mass=1.0)
    assert np.abs((acc - a0) - ref).max() <= G_REL_FP64 * np.abs(ref).max()


def test_force_deterministic(gpu):
    import torch
    N, L, n = 64, 64.0, 60000
    pos = clumpy_particles(n, L, 31, n_clumps=20, sigma=0.5)
    for prec in ("fp32", "fp64"):
        s = _solver(L, N, prec, n)
        x, y, z = (_dev(p, gpu) for p in pos)
        outs = []
        for _ in range(2):
            acc = [torch.zeros(n, dtype=torch.float64, device=gpu) for _ in range(3)]
            s.compute_forces(x, y, z, None, 1.0, *acc)
            outs.append(torch.stack(acc).cpu().numpy())
        assert np.array_equal(outs[0], outs[1])


def test_errors_map_to_reference_exceptions(gpu):
    import torch
    import hyowon_amd as H
    from hyowon_amd._lib import HyInvalidArgument, HyRuntimeError
    N, L, n = 16, 16.0, 100
    s = _solver(L, N, "fp64", n)
    pos = uniform_particles(n, L, 1)
    pos[1, 7] = np.nan
    x, y, z = (_dev(p, gpu) for p in pos)
    acc = [torch.zeros(n, dtype=torch.float64, device=gpu) for _ in range(3)]
    with pytest.raises(HyInvalidArgument, match="finite"):
        s.compute_forces(x, y, z, None, 1.0, *acc)
    assert all(float(a.abs().sum()) == 0.0 for a in acc)  # transactional
    pos = uniform_particles(n, L, 1)
    x, y, z = (_dev(p, gpu) for p in pos)
    with pytest.raises(HyInvalidArgument):
        s.compute_forces(x, y, z, None, -1.0, *acc)
    masses = np.ones(n)
    masses[3] = -2.0
    with pytest.raises(HyInvalidArgument):
        s.compute_forces(x, y, z, _dev(masses, gpu), 0.0, *acc)
    with pytest.raises(HyInvalidArgument):
        H.PMSolver(L, 10, precision="fp64", max_particles=4)  # unsupported mesh size
    with pytest.raises(HyInvalidArgument):
        big = _solver(L, N, "fp64", 10)
        big.compute_forces(x, y, z, None, 1.0, *acc)  # exceeds capacity
    with pytest.raises(HyRuntimeError):
        s.compute_forces(x[:0], y[:0], z[:0], None, 1.0, acc[0][:0], acc[1][:0], acc[2][:0])


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("fp32", 2e-3)])
def test_kdk_run_matches_oracle(gpu, oracle_mod, prec, tol):
    """Short evolution z=9->4 through the fused KDK path vs the oracle step loop."""
    import hyowon_amd as H
    N, L = 32, 64.0
    ng = 16
    q = (np.arange(ng) + 0.5) * (L / ng)
    X, Y, Z = np.meshgrid(q, q, q, indexing="ij")
    rng = np.random.default_rng(4)
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()]) + rng.normal(0, 0.3, (3, ng ** 3))
    pos = np.mod(pos, L)
    mom = rng.normal(0, 5.0, (3, ng ** 3))
    n = ng ** 3
    m = H.rho_crit0() * 0.3 * (L / ng) ** 3
    a0, a1, nsteps = 0.1, 0.2, 4
    stepper = H.TimeStepper(a0, a1, math.log(a1 / a0) / nsteps, [a1])
    b = stepper.boundaries
    s = _solver(L, N, prec, n)
    st = H.ParticleState.from_numpy(pos, mom, precision=prec)
    integ = H.LeapfrogIntegrator(s, H.DriftKickIntegrator(0.3, 0.7), m)
    st = integ.run(st, stepper)
    gpos, gmom = st.to_numpy()
    (rx, ry, rz, rpx, rpy, rpz), _ = oracle_mod.run(pos[0], pos[1], pos[2], mom[0], mom[1], mom[2],
                                                    L, N, b, m)
    rpos = np.stack([rx, ry, rz])
    rmom = np.stack([rpx, rpy, rpz])
    d = np.abs(gpos - rpos)
    d = np.minimum(d, L - d)
    assert d.max() / (L / N) < tol
    assert np.abs(gmom - rmom).max() / np.abs(rmom).max() < tol


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_reused_bucket_keys_match_fresh_keys(gpu, prec):
    """A fused step makes the next step's bucket keys; reusing them
    (HY_REUSE_KEYS) must give results bit-identical to recomputing them."""
    import hyowon_amd as H
    N, L, n = 32, 32.0, 20000
    pos = clumpy_particles(n, L, 41, n_clumps=12, sigma=0.5)
    mom = np.random.default_rng(5).normal(0, 20.0, (3, n))
    K, D = 0.02, 0.05

    def run(reuse):
        s = _solver(L, N, prec, n)
        a = H.ParticleState.from_numpy(pos, mom, precision=prec)
        b = a.empty_like()
        for i in range(4):
            s.kdk(a, b, 1.0, K if i else None, K, D, reuse_keys=reuse)
            a, b = b, a
        return [t.cpu().numpy() for t in (a.x, a.y, a.z, a.px, a.py, a.pz)]

    for u, v in zip(run(True), run(False)):
        assert np.array_equal(u, v)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_reused_keys_after_positions_change(gpu, prec):
    """Positions rewritten in place between fused steps: a synchronous call
    with HY_REUSE_KEYS detects the stale keys and routes afresh, an
    asynchronous one reports logic_error through hy_pm_check."""
    import hyowon_amd as H
    from hyowon_amd._lib import HyLogicError
    N, L, n = 32, 32.0, 5000
    pos = uniform_particles(n, L, 42)
    mom = np.zeros((3, n))

    def first_step():
        s = _solver(L, N, prec, n)
        a = H.ParticleState.from_numpy(pos, mom, precision=prec)
        b = a.empty_like()
        s.kdk(a, b, 1.0, None, 0.01, 0.01)
        b.x.add_(8.0).remainder_(L)  # every particle moves to another tile, same buffers
        return s, a, b

    s, a, b = first_step()
    s.kdk(b, a, 1.0, 0.01, 0.01, 0.01, reuse_keys=True)
    got = [t.cpu().numpy() for t in (a.x, a.y, a.z, a.px, a.py, a.pz)]
    s, a, b = first_step()
    s.kdk(b, a, 1.0, 0.01, 0.01, 0.01, reuse_keys=False)
    ref = [t.cpu().numpy() for t in (a.x, a.y, a.z, a.px, a.py, a.pz)]
    for u, v in zip(got, ref):
        assert np.array_equal(u, v)

    s, a, b = first_step()
    s.kdk(b, a, 1.0, 0.01, 0.01, 0.01, reuse_keys=True, asynchronous=True)
    with pytest.raises(HyLogicError):
        s.check()
    # the context recovers: a fresh synchronous step afterwards is correct
    s.kdk(b, a, 1.0, 0.01, 0.01, 0.01)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("N", [192, 256])
def test_fft_launch_variants_bit_identical(gpu, prec, N):
    """Every compiled launch variant and plane grouping of the FFT passes, and
    the autotuned plan, give the same potential bit for bit."""
    L = float(N)
    rng = np.random.default_rng(N + 7)
    src = rng.standard_normal((N, N, N))
    src -= src.mean()
    s = _solver(L, N, prec, 1)
    t = s.tuning()
    nv = len(t["rows"]["z_r2c"])
    assert nv == 4
    base = s.potential(_dev(src, gpu)).cpu().numpy()
    for v in range(nv):
        for group in (0, 16):
            s.set_tuning([v] * 5, plane_group=group)
            got = s.potential(_dev(src, gpu)).cpu().numpy()
            assert np.array_equal(got, base), (v, group)
    plan = s.autotune(reps=2)
    assert all(0 <= plan["variant"][p] < nv for p in plan["variant"])
    got = s.potential(_dev(src, gpu)).cpu().numpy()
    assert np.array_equal(got, base)


def test_fft_tuning_rejects_uncompiled_variant(gpu):
    from hyowon_amd._lib import HyInvalidArgument
    s = _solver(32.0, 32, "fp32", 1)
    assert len(s.tuning()["rows"]["z_r2c"]) == 1
    with pytest.raises(HyInvalidArgument):
        s.set_tuning([1, 0, 0, 0, 0])


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_lattice_block_schedule_bit_identical(gpu, prec):
    """256^3 particles in lattice order take the Lagrangian-tiled gather
    schedule; results equal, bit for bit, under every tile size the gather
    autotuner can choose (and the particle-order schedule), and the
    autotuner's choice is one of them."""
    import hyowon_amd as H
    S, N, L = 256, 64, 64.0
    n = S ** 3
    rng = np.random.default_rng(9)
    q = (np.arange(S) + 0.5) * (L / S)
    X, Y, Z = np.meshgrid(q, q, q, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()]) + rng.normal(0, 0.05, (3, n))
    pos = np.mod(pos, L)
    mom = rng.normal(0, 5.0, (3, n))

    def run(rows, tune=False):
        s = _solver(L, N, prec, n)
        a = H.ParticleState.from_numpy(pos, mom, precision=prec)
        if tune:
            t = s.autotune_gather(a.x, a.y, a.z, 1.0, reps=1)
            assert t["rows"] in t["ms"] and t["ms"][t["rows"]] > 0.0
        else:
            s.set_gather_rows(rows)
        b = a.empty_like()
        for i in range(2):
            s.kdk(a, b, 1.0, 0.01 if i else None, 0.01, 0.02, reuse_keys=True)
            a, b = b, a
        out = [t.cpu().numpy() for t in (a.x, a.y, a.z, a.px, a.py, a.pz)]
        s.close()
        return out

    tiled = run(None)
    for other in [run(0), run(16), run(1024), run(None, tune=True)]:
        for u, v in zip(tiled, other):
            assert np.array_equal(u, v)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
@pytest.mark.parametrize("explicit", [False, True])
def test_potential_energy_diagnostic(gpu, oracle_mod, prec, tol, explicit):
    """compute_forces(..., collect_potential_energy=True) returns
    PMForceDiagnostics with 0.5 sum m phi (src/gravity/pm_solver.cpp:817-841)
    equal to the oracle's (bit-exact against the reference's blocked sum) on
    the oracle's potential, to binary64 FFT roundoff; forces are unchanged."""
    import torch
    import hyowon_amd as H
    from hyowon_amd._lib import HyLogicError
    N, L, n = 48, 48.0, 30000
    pos = clumpy_particles(n, L, 21, n_clumps=30, sigma=0.5)
    masses = np.random.default_rng(3).uniform(0.2, 2.0, n) if explicit else None
    m = 1.3
    ref, phi = oracle_mod.pm_force(pos[0], pos[1], pos[2], L, N, masses=masses, uniform_mass=m, return_phi=True)
    e_ref = oracle_mod.potential_energy(phi, pos[0], pos[1], pos[2], L, masses=masses, uniform_mass=m)
    from hyowon_amd._lib import check, lib
    s = _solver(L, N, prec, n)
    assert s.compute_forces(*(_dev(p, gpu) for p in pos), None if masses is None else _dev(masses, gpu), m,
                            *(torch.zeros(n, dtype=torch.float64, device=gpu) for _ in range(3))) is None
    with pytest.raises(HyLogicError):  # that call did not collect it
        check(lib().hy_pm_potential_energy(s.handle, C.byref(C.c_double())))
    acc = [torch.zeros(n, dtype=torch.float64, device=gpu) for _ in range(3)]
    d = s.compute_forces(*(_dev(p, gpu) for p in pos), None if masses is None else _dev(masses, gpu), m, *acc,
                         collect_potential_energy=True)
    got = np.stack([a.cpu().numpy() for a in acc])
    assert math.sqrt(((got - ref) ** 2).sum() / (ref ** 2).sum()) < (1e-10 if prec == "fp64" else 1e-4)
    print(f"{prec}: potential energy {d.potential_energy_comoving:.17g} vs oracle {e_ref:.17g}")
    assert abs(d.potential_energy_comoving - e_ref) <= tol * abs(e_ref)
    s.close()


def test_potential_energy_bit_exact_on_a_given_potential(gpu, oracle_mod):
    """With the same binary64 potential on both sides (the GPU's own, copied
    out), the diagnostic's per-particle arithmetic and blocked Neumaier order
    reproduce the oracle bit for bit."""
    import torch
    import hyowon_amd as H
    N, L, n = 32, 40.0, 70000
    pos = uniform_particles(n, L, 31)
    m = 0.8
    s = _solver(L, N, "fp64", n)
    x, y, z = (_dev(p, gpu) for p in pos)
    src = s.density(x, y, z, None, m)
    phi = s.potential(src).cpu().numpy()
    acc = [torch.zeros(n, dtype=torch.float64, device=gpu) for _ in range(3)]
    d = s.compute_forces(x, y, z, None, m, *acc, collect_potential_energy=True)
    e_ref = oracle_mod.potential_energy(phi, pos[0], pos[1], pos[2], L, uniform_mass=m)
    assert d.potential_energy_comoving == e_ref
    s.close()


@pytest.mark.parametrize("explicit", [False, True])
def test_deposited_mass_exact_fp64(gpu, oracle_mod, explicit):
    """The deposited mass the mass-conservation check uses is the exact sum of
    every cell's mass, rounded once (ExactPositiveDoubleSum,
    src/gravity/pm_solver.cpp:132-180): equal to the oracle's exact sum of its
    own (bit-identical) cell masses."""
    import ctypes as C
    N, L, n = 48, 48.0, 30000
    pos = clumpy_particles(n, L, 41, n_clumps=20, sigma=0.4)
    masses = np.random.default_rng(5).uniform(0.01, 3.0, n) if explicit else None
    m = 0.37
    s = _solver(L, N, "fp64", n)
    s.density(*(_dev(p, gpu) for p in pos), _dev(masses, gpu) if explicit else None, m)
    got = s.info().last_deposited_mass
    mesh = oracle_mod.deposit(pos[0], pos[1], pos[2], L, N, masses=masses, uniform_mass=m)
    flat = np.ascontiguousarray(mesh.reshape(-1))
    ref = oracle_mod.lib().or_exact_sum(flat.ctypes.data_as(C.POINTER(C.c_double)), flat.size)
    assert got == ref, (got, ref, got - ref)


def test_deposited_mass_exact_fp32(gpu, oracle_mod):
    """Binary32 contexts deposit 32.32 fixed-point cell masses; the deposited
    mass is their exact sum times the mass unit, rounded once.  Restated here
    from the particles: per corner q = round(((m wx) wy) wz 2^32) in binary32
    with the reference's CIC fractions (oracle cic_coordinate), summed exactly
    with Python integers, scaled and rounded with fractions.Fraction."""
    from fractions import Fraction
    N, L, n = 32, 32.0, 6000
    pos = clumpy_particles(n, L, 42, n_clumps=10, sigma=0.5)
    m = 0.7
    s = _solver(L, N, "fp32", n)
    s.density(*(_dev(p, gpu) for p in pos), None, m)
    got = s.info().last_deposited_mass
    fr = np.array([[oracle_mod.cic_coordinate(float(v), L, N)[1] for v in pos[a]] for a in range(3)])
    f = fr.astype(np.float32)
    one = np.float32(1.0)
    Q = 0
    for ox in range(2):
        wx = (one - f[0]) if ox == 0 else f[0]
        for oy in range(2):
            wy = (one - f[1]) if oy == 0 else f[1]
            mxy = (one * wx).astype(np.float32) * wy  # uniform mass: the records carry m = 1 in mass units
            mxy = mxy.astype(np.float32)
            for oz in range(2):
                wz = (one - f[2]) if oz == 0 else f[2]
                w = (mxy * wz).astype(np.float32)
                scaled = (w * np.float32(4294967296.0)).astype(np.float32)
                Q += int(np.rint(scaled.astype(np.float64)).astype(np.int64).sum())
    exact = Fraction(Q) * Fraction(m) / (1 << 32)
    assert got == float(exact), (got, float(exact))
