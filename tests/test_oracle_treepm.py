"""The oracle's TreePM long-range PM force (continuum Green with the Gaussian
split, spectral gradient, CIC interpolation) pinned against the reference's
GreenFunction and CICMassAssignment::interpolate_add (built into oracle/_ref,
bit-exact), and by physics the reference's pieces do not fix on their own:
the Gaussian-filtered Newtonian pair force and momentum conservation."""
import math

import numpy as np
import pytest


@pytest.mark.parametrize("N,L,rs", [(8, 16.0, 1.5), (16, 40.0, 3.1), (12, 7.0, 0.9)])
def test_treepm_green_matches_reference(oracle_mod, ref_lib, N, L, rs):
    import ctypes as C
    rng = np.random.default_rng(N)
    spec = rng.normal(size=(N, N, N // 2 + 1)) + 1j * rng.normal(size=(N, N, N // 2 + 1))
    mine = oracle_mod.green_apply_treepm(spec, L, rs)
    ref = np.ascontiguousarray(spec, np.complex128).copy()
    assert ref_lib.ref_green_apply_treepm(N, L, rs, ref.ctypes.data_as(C.POINTER(C.c_double))) == 0
    assert np.array_equal(mine, ref)


def test_interpolate_add_matches_reference(oracle_mod, ref_lib):
    import ctypes as C
    pd = C.POINTER(C.c_double)
    rng = np.random.default_rng(3)
    for N, L in ((8, 10.0), (16, 32.0), (12, 5.5)):
        mesh = rng.normal(size=(N, N, N))
        n = 2000
        pos = rng.uniform(-L, 2 * L, (3, n))
        pos[:, :7] = np.array([[0.0, L, np.nextafter(L, 0), -1e-300, 3 * L, L / N, L - L / N]] * 3)
        base = rng.normal(size=n)
        mine = oracle_mod.interpolate_add(mesh, *pos, L, out=base)
        ref = base.copy()
        x, y, z = (np.ascontiguousarray(p) for p in pos)
        m = np.ascontiguousarray(mesh.reshape(-1))
        assert ref_lib.ref_interpolate_add(n, m.ctypes.data_as(pd), x.ctypes.data_as(pd),
                                           y.ctypes.data_as(pd), z.ctypes.data_as(pd), L, N,
                                           ref.ctypes.data_as(pd)) == 0
        assert np.array_equal(mine, ref)


def _long_range_pair(r, rs, G, m):
    """|F| of the Gaussian long-range Newtonian pair force: G m / r^2 (1 - M_S(r/rs)),
    M_S(q) = erfc(q/2) + q/sqrt(pi) exp(-q^2/4)."""
    q = r / rs
    ms = math.erfc(q / 2) + q / math.sqrt(math.pi) * math.exp(-q * q / 4)
    return G * m / r ** 2 * (1.0 - ms)


@pytest.mark.parametrize("sep_cells", [5.0, 7.5, 10.0])
def test_pair_force_is_the_gaussian_filtered_newtonian(oracle_mod, sep_cells):
    """A heavy particle and a light test particle along a diagonal direction
    of a 64^3 mesh: the light particle's acceleration matches the Gaussian
    long-range pair force (r_s = 1.25 cells) to 2%, beyond a few r_s where the
    mesh resolves the filtered force (periodic images contribute < 1%)."""
    N, L = 64, 64.0
    dx = L / N
    rs = 1.25 * dx
    G = oracle_mod.lib().or_gravitational_constant()
    M = 1.0e6
    c = np.array([20.3, 21.7, 19.1])
    d = np.array([1.0, 2.0, 2.0]) / 3.0
    p2 = c + sep_cells * dx * d
    x, y, z = (np.array([c[i], p2[i]]) for i in range(3))
    masses = np.array([M, 1e-6])
    ax, ay, az = oracle_mod.pm_force_treepm(x, y, z, L, N, rs, masses=masses)
    a = np.array([ax[1], ay[1], az[1]])
    r = sep_cells * dx
    want = _long_range_pair(r, rs, G, M)
    assert np.dot(a, d) < 0  # attractive
    assert abs(np.linalg.norm(a) / want - 1.0) < 0.02, (np.linalg.norm(a), want)
    # the transverse part is small
    assert np.linalg.norm(a - np.dot(a, d) * d) < 0.02 * want


def test_momentum_is_conserved(oracle_mod):
    """CIC deposit and CIC interpolation with a real, even Green function and
    an odd spectral gradient: sum_i m_i a_i vanishes to roundoff."""
    rng = np.random.default_rng(9)
    N, L, n = 32, 50.0, 3000
    pos = rng.uniform(0, L, (3, n))
    m = rng.uniform(0.5, 2.0, n)
    ax, ay, az = oracle_mod.pm_force_treepm(*pos, L, N, 2.0, masses=m)
    for a in (ax, ay, az):
        assert abs((m * a).sum()) < 1e-10 * (m * np.abs(a)).sum()
