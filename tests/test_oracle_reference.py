"""The CPU oracle against the reference build (oracle/_ref, compiled from the
reference sources) on fresh seeded inputs: bit-exact for every deterministic
function; the FFT-based full force agrees to binary64 roundoff because the
reference's FFT (FFTW) is absent and both sides use this repository's FFT."""
import ctypes as C
import math

import numpy as np
import pytest

from conftest import clumpy_particles, face_particles, uniform_particles

pd = C.POINTER(C.c_double)


def p(a):
    return a.ctypes.data_as(pd)


def test_wrap_minimum_image_and_cic_coordinate(oracle_mod, ref_lib):
    L_, N = 32.0, 16
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-100, 100, 2000), np.arange(-40, 80) * (L_ / N),
                         [0.0, -0.0, L_, np.nextafter(L_, 0), np.nextafter(L_, 99), -1e-300, 1e-300,
                          3 * L_, -3 * L_, -np.nextafter(L_, 0), 1e15, -1e15]])
    lib = oracle_mod.lib()
    for x in xs:
        assert lib.or_wrap(x, L_) == ref_lib.ref_wrap(x, L_) or (math.isnan(lib.or_wrap(x, L_)))
        assert lib.or_minimum_image(x, L_) == ref_lib.ref_minimum_image(x, L_)
        b1, f1 = C.c_int64(), C.c_double()
        b2, f2 = C.c_int64(), C.c_double()
        r1 = lib.or_cic_coordinate(x, L_, L_ / N, N, C.byref(b1), C.byref(f1))
        r2 = ref_lib.ref_cic_coordinate(x, L_, N, C.byref(b2), C.byref(f2))
        assert r1 == r2
        if r1 == 0:
            assert (b1.value, f1.value) == (b2.value, f2.value)


@pytest.mark.parametrize("N,L,n,kind", [(16, 32.0, 5000, "u"), (24, 48.0, 8000, "c"), (12, 10.0, 2000, "u")])
def test_deposit_density_green_gather(oracle_mod, ref_lib, N, L, n, kind):
    pos = uniform_particles(n, L, N) if kind == "u" else clumpy_particles(n, L, N, n_clumps=9, sigma=0.4)
    x, y, z = (np.ascontiguousarray(a) for a in pos)
    m = 1.7
    mesh_ref = np.zeros(N ** 3)
    assert ref_lib.ref_deposit(n, p(x), p(y), p(z), None, m, L, N, p(mesh_ref)) == 0
    mesh = oracle_mod.deposit(x, y, z, L, N, uniform_mass=m).reshape(-1)
    assert np.array_equal(mesh, mesh_ref)
    total = oracle_mod.lib().or_exact_repeated(m, n)
    assert total == ref_lib.ref_exact_repeated(m, n)
    con_ref = mesh_ref.copy()
    assert ref_lib.ref_density_contrast(p(con_ref), N, L / N, total) == 0
    con = oracle_mod.density_contrast(mesh.reshape(N, N, N), L, N, total).reshape(-1)
    assert np.array_equal(con, con_ref)
    spec = oracle_mod.rfftn(con.reshape(N, N, N))
    for dec in (0, 1):
        g_ref = spec.copy()
        assert ref_lib.ref_green_apply(N, L, dec, g_ref.ctypes.data_as(pd)) == 0
        assert np.array_equal(oracle_mod.green_apply(spec, L, dec), g_ref)
    phi = oracle_mod.irfftn(oracle_mod.green_apply(spec, L, 0)).reshape(-1)
    acc_ref = np.zeros((3, n))
    assert ref_lib.ref_gather(n, p(phi), p(x), p(y), p(z), L, N, p(acc_ref[0]), p(acc_ref[1]), p(acc_ref[2])) == 0
    acc = oracle_mod.gather(phi.reshape(N, N, N), x, y, z, L)
    assert np.array_equal(acc, acc_ref)
    # the oracle's composed force equals the reference stages composed with the same FFT
    assert np.array_equal(oracle_mod.pm_force(x, y, z, L, N, uniform_mass=m), acc_ref)


def test_density_normalize_path_for_non_dyadic_mesh(ref_lib, oracle_mod):
    # 12^3 cells: M/N_c is not an exact quotient, the fma-centred path is used
    direct = C.c_int()
    assert ref_lib.ref_density_direct_available(12, 1.0, 3000 * 2.0, C.byref(direct)) == 0
    assert direct.value == 0


def test_explicit_masses_and_faces(oracle_mod, ref_lib):
    N, L = 16, 32.0
    pos = face_particles(N, L, 9)
    n = pos.shape[1]
    masses = np.random.default_rng(3).uniform(0.0, 2.0, n)
    x, y, z = (np.ascontiguousarray(a) for a in pos)
    mesh_ref = np.zeros(N ** 3)
    assert ref_lib.ref_deposit(n, p(x), p(y), p(z), p(masses), 0.0, L, N, p(mesh_ref)) == 0
    assert np.array_equal(oracle_mod.deposit(x, y, z, L, N, masses=masses).reshape(-1), mesh_ref)
    assert oracle_mod.lib().or_exact_sum(p(masses), n) == ref_lib.ref_exact_sum(p(masses), n)


def test_mass_conservation_bound(oracle_mod, ref_lib):
    lib = oracle_mod.lib()
    for (pm, dm, cnt) in [(1.0, 1.0, 10), (1.0, 1.0 + 1e-13, 1000), (1.0, 1.0 + 1e-9, 10),
                          (5.0, 5.0 * (1 + 7e-12), 100000), (1.0, 1.0, 0), (-1.0, 1.0, 4)]:
        assert lib.or_require_cic_mass_conservation(pm, dm, cnt) == ref_lib.ref_mass_conservation(pm, dm, cnt)


def test_leapfrog_step_loop(oracle_mod, ref_lib):
    """or_run (the oracle's step loop) against the reference LeapfrogIntegrator +
    TimeStepper + DriftKickIntegrator driving the oracle force through a callback."""
    N, L, ng = 16, 32.0, 8
    q = (np.arange(ng) + 0.5) * (L / ng)
    X, Y, Z = np.meshgrid(q, q, q, indexing="ij")
    rng = np.random.default_rng(1)
    pos = np.mod(np.stack([X.ravel(), Y.ravel(), Z.ravel()]) + rng.normal(0, 0.4, (3, ng ** 3)), L)
    mom = rng.normal(0, 3.0, (3, ng ** 3))
    n = ng ** 3
    m = oracle_mod.lib().or_rho_crit0() * 0.3 * (L / ng) ** 3
    a0, a1, steps = 0.1, 0.5, 5
    dl = math.log(a1 / a0) / steps

    @oracle_mod.FORCE_CB
    def force(nn, fx, fy, fz, ax, ay, az, a, user):
        xs = np.ctypeslib.as_array(fx, (nn,)).copy()
        ys = np.ctypeslib.as_array(fy, (nn,)).copy()
        zs = np.ctypeslib.as_array(fz, (nn,)).copy()
        g = oracle_mod.pm_force(xs, ys, zs, L, N, uniform_mass=m)
        for k, ptr in enumerate((ax, ay, az)):
            np.ctypeslib.as_array(ptr, (nn,))[:] += g[k]
        return 0

    st = [np.ascontiguousarray(a).copy() for a in (*pos, *mom)]
    acc = [np.zeros(n) for _ in range(3)]
    nsteps = C.c_uint64()
    rc = ref_lib.ref_leapfrog_run(n, *[p(a) for a in st], *[p(a) for a in acc], m, L, 0.3, 0.7, a0, a1, dl,
                                  force, None, C.byref(nsteps))
    b = oracle_mod.time_boundaries(a0, a1, dl, [a1])
    assert rc == 0 and nsteps.value == len(b) - 1
    (rx, ry, rz, rpx, rpy, rpz), racc = oracle_mod.run(*pos, *mom, L, N, b, m)
    for got, want in zip(st, (rx, ry, rz, rpx, rpy, rpz)):
        assert np.array_equal(got, want)
    for got, want in zip(acc, racc):
        assert np.array_equal(got, want)


def _underflow_cases():
    """Inputs that lose a non-zero product in the centered-gradient gather
    (src/mesh/mass_assignment_gradient.cpp:157-196), one per condition, plus a
    control with normal magnitudes."""
    N, L = 8, 16.0  # dx = 2
    rng = np.random.default_rng(11)
    cases = {}
    # (a) a weight whose three factors are non-zero rounds to zero: fx = fy = 1e-200
    phi = rng.normal(size=N ** 3)
    cases["weight"] = (phi, np.array([2e-200]), np.array([2e-200]), np.array([1.0]))
    # (b) a non-zero potential difference whose centred quotient rounds to zero
    phi = np.zeros(N ** 3)
    phi.reshape(N, N, N)[2, 0, 0] = 5e-324  # cell x = 2 against x = 0 around x = 1
    cases["gradient"] = (phi, np.array([2.5]), np.array([0.5]), np.array([0.5]))
    # (c) a weighted contribution rounds to zero: w ~ 1e-10, g ~ 1e-316
    phi = rng.normal(size=N ** 3) * 1e-315
    cases["contribution"] = (phi, np.array([2e-10]), np.array([0.5]), np.array([0.5]))
    cases["control"] = (rng.normal(size=N ** 3), np.array([2e-10]), np.array([3.3]), np.array([7.9]))
    return N, L, cases


@pytest.mark.parametrize("which", ["weight", "gradient", "contribution", "control"])
def test_gather_underflow_errors_match_reference(oracle_mod, ref_lib, which):
    N, L, cases = _underflow_cases()
    phi, x, y, z = cases[which]
    outs = []
    codes = []
    for fn in (ref_lib.ref_gather, oracle_mod.lib().or_gather_centered_gradient_add):
        acc = np.zeros((3, len(x)))
        codes.append(fn(len(x), p(phi), p(x), p(y), p(z), L, N, p(acc[0]), p(acc[1]), p(acc[2])))
        outs.append(acc)
    assert codes[0] == codes[1] == (0 if which == "control" else 4)
    if which == "control":
        assert np.array_equal(outs[0], outs[1])


def test_potential_energy_bit_exact_vs_reference(oracle_mod, ref_lib):
    """or_potential_energy = 0.5 * CICMassAssignment::interpolate_mass_weighted_sum_full_mesh
    (fixed-block Neumaier over 4096-particle blocks), bit for bit, incl.
    wrapped coordinates, explicit masses and several blocks."""
    import ctypes as C
    pd = C.POINTER(C.c_double)
    rng = np.random.default_rng(17)
    for N, L, n in ((16, 32.0, 5000), (12, 24.0, 9000), (32, 10.0, 20000)):
        phi = rng.standard_normal((N, N, N))
        pos = rng.uniform(-L, 2 * L, (3, n))
        m = rng.uniform(0.1, 2.0, n)
        for masses in (None, m):
            got = oracle_mod.potential_energy(phi, *pos, L, masses=masses, uniform_mass=1.7)
            out = C.c_double()
            mm = np.ascontiguousarray(masses) if masses is not None else None
            rc = ref_lib.ref_potential_energy(n, phi.ctypes.data_as(pd),
                                              *[np.ascontiguousarray(p).ctypes.data_as(pd) for p in pos],
                                              mm.ctypes.data_as(pd) if mm is not None else None, 1.7, L, N,
                                              C.byref(out))
            assert rc == 0
            assert got == out.value
