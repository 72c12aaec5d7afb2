"""The oracle's matter power-spectrum estimator (MatterPowerSpectrum::compute_pk
over FourierDensityBuilder::build) pinned three ways:

* its helpers bit-exact against the reference's own header-only helpers
  (shell edges and lookup, CIC window, |v|^2 L^3 with and without the
  shot-noise subtraction, L^3 scaling) and the reference's unit-normalised
  CIC deposit, all built by oracle/Makefile;
* closed forms: particle patterns on cell corners whose contrast has a
  hand-computed spectrum (Nyquist plane, period-4 pattern with its
  interlacing partner), the number and mean |k| of the modes in each shell
  from a brute-force count over the full (not half) grid;
* invariants: mass-scale and integer-cell translation invariance, and the
  Poisson expectation E[P_raw] = S prod(1 - 2/3 sin^2(k_i dx/2)) for random
  particles (which a wrong alias model or deconvolution fails).

The reference estimator itself runs its FFT through FFTW loaded at run time,
which this image does not have, so the estimator loop is pinned through the
above rather than through a reference run."""
import ctypes as C
import math

import numpy as np
import pytest

pd = C.POINTER(C.c_double)


def _full_grid_shells(N, L, edges, kmax):
    """Mode count and mean |k| per shell over the full N^3 grid (no half-space)."""
    m = np.arange(N)
    m = np.where(m <= N // 2, m, m - N).astype(np.float64)
    k1 = 2.0 * math.pi * m / L
    kx, ky, kz = np.meshgrid(k1, k1, k1, indexing="ij")
    # the half-complex grid has kz >= 0 only; mirror-equivalent modes of the
    # full grid are (kx, ky, kz) and (-kx, -ky, -kz), i.e. multiplicity 2 off
    # the kz = 0 and Nyquist planes.  Enumerate the full grid with the half
    # grid's representative (kz index <= N/2) of each pair.
    kmag = np.hypot(np.hypot(kx, ky), kz)
    sel = (kmag > 0) & (kmag <= kmax)
    counts = np.zeros(len(edges) - 1, np.int64)
    ksum = np.zeros(len(edges) - 1)
    for b in range(len(edges) - 1):
        hi_closed = b == len(edges) - 2
        inb = sel & (kmag >= edges[b]) & ((kmag <= edges[b + 1]) if hi_closed else (kmag < edges[b + 1]))
        counts[b] = int(inb.sum())
        ksum[b] = kmag[inb].sum()
    return counts, ksum


def _corner_particles(counts3, L):
    """Particles exactly on cell corners: counts3[i, j, k] particles at (i, j, k) dx."""
    N = counts3.shape[0]
    dx = L / N
    idx = np.repeat(np.arange(N ** 3), counts3.reshape(-1))
    i, j, k = np.unravel_index(idx, counts3.shape)
    return i * dx, j * dx, k * dx


# ---------------------------------------------------------------------------
# helpers against the reference build
# ---------------------------------------------------------------------------
def test_edges_and_lookup_match_reference(oracle_mod, ref_lib):
    rng = np.random.default_rng(11)
    O = oracle_mod.lib()
    cases = [(2 * math.pi / 512, math.pi, 32), (0.1, 0.2, 7), (1e-3, 50.0, 100), (0.05, 0.051, 3),
             (2 * math.pi / 128, 0.5 * math.pi, 1)]
    for _ in range(40):
        lo = 10 ** rng.uniform(-4, 1)
        cases.append((lo, lo * 10 ** rng.uniform(0.01, 4), int(rng.integers(1, 200))))
    for lo, hi, nb in cases:
        a, b = np.zeros(nb + 1), np.zeros(nb + 1)
        assert O.or_pk_edges(lo, hi, nb, a.ctypes.data_as(pd)) == 0
        assert ref_lib.ref_pk_edges(lo, hi, nb, b.ctypes.data_as(pd)) == 0
        assert np.array_equal(a, b), (lo, hi, nb)
        probes = np.concatenate([a, np.nextafter(a, 0), np.nextafter(a, np.inf),
                                 rng.uniform(lo * 0.9, hi * 1.1, 50)])
        for v in probes:
            r = C.c_int()
            assert ref_lib.ref_pk_bin_index(lo, hi, nb, float(v), C.byref(r)) == 0
            assert O.or_pk_bin_index(a.ctypes.data_as(pd), nb, float(v)) == r.value, (lo, hi, nb, v)


def test_window_and_power_helpers_match_reference(oracle_mod, ref_lib):
    rng = np.random.default_rng(12)
    O = oracle_mod.lib()
    for N, L in ((16, 64.0), (192, 384.0), (1024, 512.0), (12, 7.3)):
        dx = L / N
        for i in range(N):
            m = i if i <= N // 2 else i - N
            k = 2.0 * math.pi * m / L
            w = C.c_longdouble()
            assert ref_lib.ref_pk_window_axis(k, dx, C.byref(w)) == 0
            assert O.or_pk_window_axis(k, dx) == w.value
    out = C.c_double()
    for _ in range(4000):
        L = 10 ** rng.uniform(-1, 3.5)
        re, im = rng.normal(size=2) * 10 ** rng.uniform(-8, 2, 2)
        if rng.random() < 0.1:
            re = 0.0
        assert ref_lib.ref_pk_raw_power(L, re, im, C.byref(out)) == 0
        assert O.or_pk_raw_power(L, re, im) == out.value
        raw = out.value
        # subtraction: generic, near-cancelling and exactly cancelling values
        for s in (abs(raw) * rng.uniform(0, 3), raw * (1 + 1e-17 * rng.normal()), raw,
                  np.nextafter(raw, np.inf), 0.0):
            if not s >= 0.0:
                continue
            assert ref_lib.ref_pk_raw_power_minus(L, re, im, s, C.byref(out)) == 0
            got = O.or_pk_raw_power_minus(L, re, im, s)
            assert got == out.value, (L, re, im, s, got, out.value)
    # the shot-noise reference L^3 * (1/n) of the estimator output
    for n in (1, 7, 3000):
        for L in (64.0, 100.0, 3.3):
            assert ref_lib.ref_pk_volume_times(L, 1.0 / n, C.byref(out)) == 0
            x = rng.uniform(0, L, (3, n))
            o = oracle_mod.power_spectrum(x[0], x[1], x[2], L, 8, 1, 0, 0, 1.0)
            assert o[0, 6] == out.value


def test_unit_normalised_deposit_matches_reference(oracle_mod, ref_lib):
    """The oracle deposits weights m / max(m) (uniform: weight 1), as the
    reference's CICMassAssignment with a deposition mass unit."""
    rng = np.random.default_rng(13)
    n, N, L = 3000, 16, 40.0
    x, y, z = rng.uniform(-L, 2 * L, (3, n))
    m = rng.uniform(0.2, 3.0, n)
    unit = m.max()
    mesh_ref = np.zeros(N ** 3)
    assert ref_lib.ref_deposit_unit(n, x.ctypes.data_as(pd), y.ctypes.data_as(pd), z.ctypes.data_as(pd),
                                    m.ctypes.data_as(pd), 0.0, unit, L, N, mesh_ref.ctypes.data_as(pd)) == 0
    mine = oracle_mod.deposit(x, y, z, L, N, masses=m / unit).reshape(-1)
    assert np.array_equal(mine, mesh_ref)
    u = 0.7371
    assert ref_lib.ref_deposit_unit(n, x.ctypes.data_as(pd), y.ctypes.data_as(pd), z.ctypes.data_as(pd),
                                    None, u, u, L, N, mesh_ref.ctypes.data_as(pd)) == 0
    mine = oracle_mod.deposit(x, y, z, L, N, uniform_mass=1.0).reshape(-1)
    assert np.array_equal(mine, mesh_ref)


# ---------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("N", [8, 16])
def test_alternating_planes_nyquist(oracle_mod, N):
    """Counts 2 / 1 on alternating x planes: contrast +-1/3, so the only
    nonzero mode is (N/2, 0, 0) with delta_k = 1/3, P_raw = L^3 / 9 and a CIC
    power deconvolution (pi/2)^4 = 1 / sinc^4(pi/2)."""
    L = 50.0
    c = np.ones((N, N, N), np.int64)
    c[0::2] = 2
    x, y, z = _corner_particles(c, L)
    nb = 4
    out = oracle_mod.power_spectrum(x, y, z, L, N, nb, 0, 0, 1.0, uniform_mass=2.5)
    praw_sum = out[:, 3] * out[:, 7]
    pc_sum = out[:, 4] * out[:, 7]
    assert praw_sum[-1] == pytest.approx(L ** 3 / 9.0, rel=1e-12)
    assert np.all(np.abs(praw_sum[:-1]) < 1e-20 * L ** 3)
    assert pc_sum[-1] == pytest.approx(L ** 3 / 9.0 * (math.pi / 2) ** 4, rel=1e-12)
    # interlaced: the half-cell-shifted partner is uniform, so delta_k halves
    out_i = oracle_mod.power_spectrum(x, y, z, L, N, nb, 0, 1, 1.0, uniform_mass=2.5)
    assert (out_i[:, 3] * out_i[:, 7])[-1] == pytest.approx(L ** 3 / 36.0, rel=1e-12)


@pytest.mark.parametrize("N", [8, 16, 32])
def test_period4_interlacing_phase(oracle_mod, N):
    """Counts [2, 1, 1, 1] repeating along x.  At k = 2 pi h / (4 dx) the
    contrast has delta_1 = DFT of (c / 1.25 - 1) over one period; the
    partner deposit of the particles shifted by -dx/2 in x, y and z is
    c2(j) = (c(j) + c(j+1)) / 2 along x.  By the shift theorem the partner's
    mode carries e^{i k dx/2}; the interlaced mode is 0.5 (delta_1 +
    delta_2 e^{-i k dx/2}).  The opposite phase sign gives a different
    power at h = 1, 3."""
    L = 32.0
    dx = L / N
    pat = np.array([2, 1, 1, 1])
    c = np.broadcast_to(np.tile(pat, N // 4)[:, None, None], (N, N, N)).astype(np.int64)
    x, y, z = _corner_particles(c, L)
    mean = pat.mean()
    d1 = pat / mean - 1.0
    d2 = (pat + np.roll(pat, -1)) / 2.0 / mean - 1.0
    j = np.arange(4)
    expect = {}
    for h in (1, 2, 3):
        kx_idx = h * N // 4
        mnum = kx_idx if kx_idx <= N // 2 else kx_idx - N
        k = 2.0 * math.pi * mnum / L
        e = np.exp(-2j * math.pi * h * j / 4)
        m1 = (d1 * e).mean()
        m2 = (d2 * e).mean()
        mi = 0.5 * (m1 + m2 * np.exp(-1j * k * dx / 2))
        expect[abs(k)] = expect.get(abs(k), 0.0) + L ** 3 * abs(mi) ** 2
    nb = 3
    out = oracle_mod.power_spectrum(x, y, z, L, N, nb, 0, 1, 1.0, uniform_mass=1.0)
    edges = np.concatenate([out[:, 0], out[-1:, 2]])
    got = np.zeros(nb)
    want = np.zeros(nb)
    for k, p in expect.items():
        b = oracle_mod.lib().or_pk_bin_index(edges.ctypes.data_as(pd), nb, k)
        want[b] += p
    got = out[:, 3] * out[:, 7]
    assert np.allclose(got, want, rtol=1e-11, atol=1e-20 * L ** 3)


@pytest.mark.parametrize("N,frac,nb", [(8, 1.0, 5), (12, 0.5, 4), (16, 0.5, 7), (16, 1.0, 9), (24, 0.75, 6)])
def test_shell_counts_and_mean_k_bruteforce(oracle_mod, N, frac, nb):
    rng = np.random.default_rng(N)
    L = 77.0
    x, y, z = rng.uniform(0, L, (3, 500))
    out = oracle_mod.power_spectrum(x, y, z, L, N, nb, 0, 0, frac, uniform_mass=1.0)
    edges = np.concatenate([out[:, 0], out[-1:, 2]])
    kmax = frac * (math.pi * N / L)
    counts, ksum = _full_grid_shells(N, L, edges, kmax)
    assert np.array_equal(out[:, 7].astype(np.int64), counts)
    nz = counts > 0
    assert np.allclose(out[nz, 1], ksum[nz] / counts[nz], rtol=1e-14)
    assert edges[0] == 2.0 * math.pi / L and edges[-1] == kmax


def test_shot_noise_reference(oracle_mod):
    rng = np.random.default_rng(5)
    L, N, n = 20.0, 8, 777
    x, y, z = rng.uniform(0, L, (3, n))
    out = oracle_mod.power_spectrum(x, y, z, L, N, 2, 1, 0, 1.0, uniform_mass=3.0)
    assert out[0, 6] == pytest.approx(L ** 3 / n, rel=1e-15)
    m = rng.uniform(0.5, 2.0, n)
    out = oracle_mod.power_spectrum(x, y, z, L, N, 2, 1, 0, 1.0, masses=m)
    assert out[0, 6] == pytest.approx(L ** 3 * (m ** 2).sum() / m.sum() ** 2, rel=1e-14)


# ---------------------------------------------------------------------------
# invariants
# ---------------------------------------------------------------------------
def test_mass_scale_and_translation_invariance(oracle_mod):
    rng = np.random.default_rng(6)
    L, N, n = 32.0, 16, 6000
    x, y, z = rng.uniform(0, L, (3, n))
    m = rng.uniform(0.5, 2.0, n)
    a = oracle_mod.power_spectrum(x, y, z, L, N, 6, 1, 0, 1.0, masses=m)
    b = oracle_mod.power_spectrum(x, y, z, L, N, 6, 1, 0, 1.0, masses=m * 1024.0)
    assert np.array_equal(a, b)  # a power-of-two scale is exact
    c = oracle_mod.power_spectrum(x, y, z, L, N, 6, 1, 0, 1.0, masses=m * 3.7)
    assert np.allclose(a, c, rtol=1e-12, atol=1e-12)
    dx = L / N
    t = oracle_mod.power_spectrum(x + 3 * dx, y - 5 * dx, z + 7 * dx, L, N, 6, 0, 0, 1.0, masses=m)
    u = oracle_mod.power_spectrum(x, y, z, L, N, 6, 0, 0, 1.0, masses=m)
    assert np.allclose(t[:, 3], u[:, 3], rtol=1e-9)


def test_poisson_expectation_of_aliased_shot_noise(oracle_mod):
    """Uniform random particles: E[P_raw(k)] = S prod_i (1 - 2/3 sin^2(k_i dx/2)),
    so the shot-noise-subtracted power averages to zero over many modes."""
    L, N, n = 64.0, 32, 20000
    acc, wsum = 0.0, 0.0
    for seed in range(6):
        rng = np.random.default_rng(100 + seed)
        x, y, z = rng.uniform(0, L, (3, n))
        out = oracle_mod.power_spectrum(x, y, z, L, N, 4, 1, 0, 0.5, uniform_mass=1.0)
        shot = out[0, 6]
        acc += (out[:, 4] * out[:, 7]).sum() / shot
        wsum += out[:, 7].sum()
    mean = acc / wsum
    # per-mode scatter of P_raw / S is ~1 (times the deconvolution, < 1.6 here)
    assert abs(mean) < 4.0 * 1.6 / math.sqrt(wsum / 2), mean


def test_errors(oracle_mod):
    x = np.array([1.0, 2.0])
    with pytest.raises(oracle_mod.OracleError):
        oracle_mod.power_spectrum(x, x, x, 10.0, 8, 0, 0, 0, 1.0)
    with pytest.raises(oracle_mod.OracleError):
        oracle_mod.power_spectrum(x, x, x, 10.0, 8, 4, 1, 1, 1.0)
    with pytest.raises(oracle_mod.OracleError):
        oracle_mod.power_spectrum(x, x, x, 10.0, 8, 4, 0, 0, 0.0)
    with pytest.raises(oracle_mod.OracleError):
        oracle_mod.power_spectrum(x, x, np.array([1.0, np.nan]), 10.0, 8, 4, 0, 0, 1.0)
    with pytest.raises(oracle_mod.OracleError):
        oracle_mod.power_spectrum(x, x, x, 10.0, 8, 4, 0, 0, 1.0, masses=np.array([1.0, 0.0]))
    with pytest.raises(oracle_mod.OracleError):  # k_max = 0.2 k_Nyq < k_f at N = 8
        oracle_mod.power_spectrum(x, x, x, 10.0, 8, 4, 0, 0, 0.2)
