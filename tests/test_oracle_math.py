"""Checks of the oracle against what the mathematics fixes: textbook DFTs,
closed forms of the discrete Green function and centered gradient, CIC
weights, conservation laws and Einstein-de Sitter time integrals.  Each is
chosen so that a plausible slip (sign, index, transposed axis, dropped term)
fails it."""
import math

import numpy as np
import pytest
from scipy import integrate

from conftest import uniform_particles


def _dft3_long(real):
    """Brute-force separable DFT in long double (tiny meshes)."""
    N = real.shape[0]
    k = np.arange(N)
    W = np.exp(-2j * np.pi * np.outer(k, k).astype(np.longdouble) / N).astype(np.clongdouble)
    a = real.astype(np.clongdouble)
    a = np.tensordot(W, a, axes=([1], 0))
    a = np.tensordot(W, a, axes=([1], 1)).transpose(1, 0, 2)
    a = np.tensordot(W, a, axes=([1], 2)).transpose(1, 2, 0)
    return a[:, :, : N // 2 + 1]


@pytest.mark.parametrize("N", [4, 6, 8, 9, 10, 12, 15])
def test_fft_matches_brute_force_dft(oracle_mod, N):
    r = np.random.default_rng(N).standard_normal((N, N, N))
    got = oracle_mod.rfftn(r)
    want = _dft3_long(r)
    assert np.abs(got - want.astype(np.complex128)).max() < 1e-12 * np.abs(want).max()
    assert np.abs(oracle_mod.irfftn(got) - r).max() < 1e-13


@pytest.mark.parametrize("N", [16, 24, 32, 48, 64, 96, 128])
def test_fft_matches_pocketfft(oracle_mod, N):
    r = np.random.default_rng(N).standard_normal((N, N, N))
    spec = oracle_mod.rfftn(r)
    ref = np.fft.rfftn(r)
    assert np.abs(spec - ref).max() < 1e-12 * np.abs(ref).max()
    # backward: c2r including the 1/N^3 of the reference convention
    assert np.abs(oracle_mod.irfftn(ref) - np.fft.irfftn(ref, s=r.shape, axes=(0, 1, 2))).max() < 1e-13


def test_green_single_mode_closed_form(oracle_mod):
    """delta = cos(2 pi m.x/N) gives phi = 4 pi G delta / lambda(m) with the
    discrete Laplacian lambda = -(4/dx^2) sum sin^2(pi m_j/N)."""
    N, L = 32, 64.0
    dx = L / N
    G = oracle_mod.lib().or_gravitational_constant()
    i = np.arange(N)
    for m in [(1, 0, 0), (0, 3, 0), (2, 5, 7), (16, 0, 1)]:
        ph = 2 * np.pi * (m[0] * i[:, None, None] + m[1] * i[None, :, None] + m[2] * i[None, None, :]) / N
        delta = np.cos(ph)
        lam = sum(-4.0 / dx ** 2 * math.sin(math.pi * mj / N) ** 2 for mj in m)
        phi = oracle_mod.irfftn(oracle_mod.green_apply(oracle_mod.rfftn(delta), L))
        assert np.abs(phi - 4 * math.pi * G * delta / lam).max() < 1e-12 * abs(4 * math.pi * G / lam)


def test_green_zero_mode_and_deconvolution_window(oracle_mod):
    N, L = 16, 16.0
    lib = oracle_mod.lib()
    assert lib.or_green_kernel(N, L, 0, 0, 0, 0) == 0.0
    k0 = lib.or_green_kernel(N, L, 0, 3, 1, 2)
    k1 = lib.or_green_kernel(N, L, 1, 3, 1, 2)
    w = 1.0
    for mm in (3, 1, 2):
        x = math.pi * mm / N
        w *= (math.sin(x) / x) ** 2
    assert k1 == pytest.approx(k0 / w ** 2, rel=1e-14)


def test_gather_plane_wave_closed_form(oracle_mod):
    """phi = cos(2 pi m x/L) along x: the centered difference at a mesh node is
    sin(k x) sin(k dx)/dx and CIC interpolates it linearly."""
    N, L = 32, 32.0
    dx = L / N
    k = 2 * math.pi * 3 / L
    i = np.arange(N)
    phi = np.broadcast_to(np.cos(k * i * dx)[:, None, None], (N, N, N)).copy()
    pos = uniform_particles(2000, L, 2)
    acc = oracle_mod.gather(phi, *pos, L)
    ix = np.floor(pos[0] / dx).astype(int)
    f = pos[0] / dx - ix
    node = lambda j: np.sin(k * (j % N) * dx) * math.sin(k * dx) / dx
    want = (1 - f) * node(ix) + f * node(ix + 1)
    assert np.abs(acc[0] - want).max() < 1e-12
    assert np.abs(acc[1]).max() < 1e-13 and np.abs(acc[2]).max() < 1e-13


def test_single_particle_deposit_weights(oracle_mod):
    N, L = 8, 8.0
    x, y, z = 2.25, 5.5, 7.75  # base (2,5,7), fractions (0.25, 0.5, 0.75), z wraps to 0
    mesh = oracle_mod.deposit([x], [y], [z], L, N, uniform_mass=2.0)
    want = np.zeros((N, N, N))
    for ox, wx in ((0, 0.75), (1, 0.25)):
        for oy, wy in ((0, 0.5), (1, 0.5)):
            for oz, wz in ((0, 0.25), (1, 0.75)):
                want[(2 + ox) % N, (5 + oy) % N, (7 + oz) % N] += 2.0 * wx * wy * wz
    assert np.array_equal(mesh, want)


def test_force_lone_particle_and_momentum_balance(oracle_mod):
    N, L = 16, 16.0
    # the self-force cancels exactly in exact arithmetic (same CIC kernel for
    # assignment and interpolation, antisymmetric difference, even Green
    # function); what remains is roundoff of the potential scale
    acc, phi = oracle_mod.pm_force([3.3], [7.1], [12.9], L, N, uniform_mass=1.0, return_phi=True)
    assert np.abs(acc).max() < 1e-14 * np.ptp(phi) / (L / N)
    pos = uniform_particles(3000, L, 4)
    acc = oracle_mod.pm_force(*pos, L, N, uniform_mass=1.0)
    # sum m g vanishes to roundoff for the momentum-conserving CIC scheme
    assert np.abs(acc.sum(axis=1)).max() < 1e-9 * np.abs(acc).sum(axis=1).max()


def test_two_particle_force_is_newtonian(oracle_mod):
    N, L = 64, 64.0
    G = oracle_mod.lib().or_gravitational_constant()
    m = 1.0
    for r in (8.0, 12.0):
        x = np.array([20.0, 20.0 + r])
        acc = oracle_mod.pm_force(x, np.array([30.0, 30.0]), np.array([10.0, 10.0]), L, N, uniform_mass=m)
        assert acc[0][0] > 0 and acc[0][1] < 0  # attraction
        assert acc[0][0] == pytest.approx(G * m / r ** 2, rel=0.1)


def test_mass_conservation_of_deposit(oracle_mod):
    N, L = 16, 16.0
    pos = uniform_particles(10000, L, 6)
    mesh = oracle_mod.deposit(*pos, L, N, uniform_mass=0.5)
    assert mesh.sum() == pytest.approx(5000.0, rel=1e-13)
    assert (mesh >= 0).all()


def test_sorts_are_stable_argsorts(oracle_mod):
    N, L = 16, 16.0
    pos = uniform_particles(5000, L, 7)
    pos[0, :100] = np.floor(pos[0, :100])  # exact faces
    off, order = oracle_mod.xplane_order(pos[0], L, N)
    bx = np.floor(np.mod(pos[0], L) / (L / N)).astype(int)
    assert np.array_equal(order, np.argsort(bx, kind="stable"))
    assert np.array_equal(off, np.concatenate([[0], np.cumsum(np.bincount(bx, minlength=N))]))
    T = oracle_mod.bucket_tile_edge(N)
    off, order = oracle_mod.bucket_order(*pos, L, N, T)
    b = [np.floor(np.mod(c, L) / (L / N)).astype(int) for c in pos]
    t = [bb // T for bb in b]
    cls = ((b[0] % T == T - 1) * 4 + (b[1] % T == T - 1) * 2 + (b[2] % T == T - 1))
    key = ((t[0] * (N // T) + t[1]) * (N // T) + t[2]) * 8 + cls
    assert np.array_equal(order, np.argsort(key, kind="stable"))


def _eds(a1, a2):
    # Omega_m = 1: D = (2/H0)(a1^-1/2 - a2^-1/2), K = (2/H0)(a2^1/2 - a1^1/2)
    return (2 / 100.0) * (a1 ** -0.5 - a2 ** -0.5), (2 / 100.0) * (a2 ** 0.5 - a1 ** 0.5)


@pytest.mark.parametrize("a1,a2", [(0.02, 0.03), (0.1, 0.9), (0.5, 0.25)])
def test_drift_kick_einstein_de_sitter(oracle_mod, a1, a2):
    D, K = _eds(a1, a2)
    assert oracle_mod.drift_factor(a1, a2, 1.0, 0.0) == pytest.approx(D, rel=1e-14)
    assert oracle_mod.kick_factor(a1, a2, 1.0, 0.0) == pytest.approx(K, rel=1e-14)


def test_drift_kick_lcdm_quadrature(oracle_mod):
    H = lambda a: 100.0 * math.sqrt(0.3 / a ** 3 + 0.7)
    for a1, a2 in [(0.02, 0.05), (0.3, 1.0)]:
        D, _ = integrate.quad(lambda a: 1 / (a ** 3 * H(a)), a1, a2, epsabs=0, epsrel=1e-13)
        K, _ = integrate.quad(lambda a: 1 / (a ** 2 * H(a)), a1, a2, epsabs=0, epsrel=1e-13)
        assert oracle_mod.drift_factor(a1, a2) == pytest.approx(D, rel=1e-11)
        assert oracle_mod.kick_factor(a1, a2) == pytest.approx(K, rel=1e-11)


def test_kdk_constant_force_is_a_parabola(oracle_mod):
    """With g constant, one K1+D step gives p1 = p0 + g K1, x1 = x0 + p1 D."""
    L = 100.0
    x, y, z, px, py, pz = [np.array([v]) for v in (10.0, 20.0, 30.0, 1.0, -2.0, 0.5)]
    out = oracle_mod.kick_drift(x, y, z, px, py, pz, [3.0], [0.0], [-1.0], 0.25, 0.5, L)
    assert out[3][0] == 1.0 + 3.0 * 0.25 and out[0][0] == 10.0 + (1.75 * 0.5)
    assert out[5][0] == 0.5 - 0.25 and out[2][0] == 30.0 + 0.25 * 0.5
