"""The oracle's mode-keyed Gaussian initial field (Philox4x32-10, open
53-bit uniforms, Box-Muller, Hermitian mode assembly) pinned against the
Random123 known-answer vectors, the reference's own philox_rng.cpp (built by
oracle/Makefile), and the invariants the field must satisfy."""
import numpy as np
import pytest

# Random123 philox4x32_10 known-answer vectors (counter, key) -> output
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("counter,key,expect", KAT)
def test_philox_known_answers(oracle_mod, counter, key, expect):
    assert tuple(int(v) for v in oracle_mod.philox(counter, key)) == expect


def test_philox_uniform_normal_bit_exact_vs_reference(oracle_mod, ref_lib):
    import ctypes as C
    rng = np.random.default_rng(3)
    O, R = oracle_mod.lib(), ref_lib
    pd = C.POINTER(C.c_double)
    pu = C.POINTER(C.c_uint32)
    for _ in range(3000):
        ctr = rng.integers(0, 2 ** 32, 4, dtype=np.uint64).astype(np.uint32)
        key = rng.integers(0, 2 ** 32, 2, dtype=np.uint64).astype(np.uint32)
        a, b = np.zeros(4, np.uint32), np.zeros(4, np.uint32)
        O.or_philox4x32_10(ctr.ctypes.data_as(pu), key.ctypes.data_as(pu), a.ctypes.data_as(pu))
        R.ref_philox4x32_10(ctr.ctypes.data_as(pu), key.ctypes.data_as(pu), b.ctypes.data_as(pu))
        assert np.array_equal(a, b)
    for _ in range(3000):
        seed = int(rng.integers(0, 2 ** 63))
        kx, ky, kz = (int(v) for v in rng.integers(-600, 600, 3))
        stream = int(rng.integers(0, 2))
        for fo, fr in ((O.or_uniform_pair, R.ref_uniform_pair), (O.or_normal_pair, R.ref_normal_pair)):
            a, b = np.zeros(2), np.zeros(2)
            fo(seed, kx, ky, kz, stream, a.ctypes.data_as(pd))
            fr(seed, kx, ky, kz, stream, b.ctypes.data_as(pd))
            assert np.array_equal(a, b), (seed, kx, ky, kz, stream)


def test_open_uniform_edges(oracle_mod):
    """Uniforms stay strictly inside (0, 1) on the 53-bit grid."""
    import ctypes as C
    pd = C.POINTER(C.c_double)
    u = np.zeros(2)
    lo, hi = 1.0, 0.0
    for kz in range(20000):
        oracle_mod.lib().or_uniform_pair(7, 3, -2, kz, 0, u.ctypes.data_as(pd))
        assert 0.0 < u[0] < 1.0 and 0.0 < u[1] < 1.0
        assert (u * 2.0 ** 53 == np.round(u * 2.0 ** 53)).all()
        lo, hi = min(lo, u.min()), max(hi, u.max())
    assert lo < 1e-3 and hi > 1 - 1e-3


@pytest.mark.parametrize("N", [8, 12, 16])
@pytest.mark.parametrize("flags", [0, 1, 2])
def test_gaussian_modes_hermitian_and_real_inverse(oracle_mod, N, flags):
    """On the stored conjugate planes X(-m) = conj X(m); self-conjugate modes
    are real; the zero mode is zero; the c2r inverse reproduces the field, so
    the half spectrum is that of a real field."""
    NH = N // 2 + 1
    var = np.linspace(1.0, 2.0, 3 * (N // 2) ** 2 + 1)
    X = oracle_mod.gaussian_modes(N, 12345, N // 2, var, flags)
    assert X[0, 0, 0] == 0
    planes = [0] + ([N // 2] if N % 2 == 0 else [])
    for k in planes:
        for i in range(N):
            for j in range(N):
                ci, cj = (-i) % N, (-j) % N
                assert X[i, j, k] == np.conj(X[ci, cj, k])
                if (ci, cj) == (i, j):
                    assert X[i, j, k].imag == 0.0
    full = np.fft.irfftn(X, s=(N, N, N), axes=(0, 1, 2))
    assert np.allclose(np.fft.rfftn(full, axes=(0, 1, 2)), X, atol=1e-12)


def test_gaussian_modes_statistics_and_band_limit(oracle_mod):
    """Unit-variance draws: the mean of |X|^2 / var over modes is 1; modes past
    max_mode on any axis are zero; the reversed pair is the negated field; the
    fixed-amplitude field has |X|^2 = var exactly (up to rounding)."""
    N, K = 32, 10
    var = np.full(3 * (N // 2) ** 2 + 1, 3.0)
    X = oracle_mod.gaussian_modes(N, 99, K, var, 0)
    m = np.fft.fftfreq(N, 1.0 / N).astype(int)
    mz = np.arange(N // 2 + 1)
    inside = (np.abs(m)[:, None, None] <= K) & (np.abs(m)[None, :, None] <= K) & (mz[None, None, :] <= K)
    assert np.all(X[~inside] == 0)
    p = np.abs(X[inside]) ** 2 / 3.0
    assert abs(p[1:].mean() - 1.0) < 0.03
    assert np.array_equal(oracle_mod.gaussian_modes(N, 99, K, var, 2), -X)
    F = oracle_mod.gaussian_modes(N, 99, K, var, 1)
    nz = inside.copy()
    nz[0, 0, 0] = False
    assert np.allclose(np.abs(F[nz]) ** 2, 3.0, rtol=1e-14)


def _single_mode_spectrum(N, m, amp, theta):
    """Half spectrum whose c2r (with 1/N^3) is amp cos(2 pi m.i / N + theta); m_z in (0, N/2)."""
    dk = np.zeros((N, N, N // 2 + 1), np.complex128)
    dk[m[0] % N, m[1] % N, m[2]] = 0.5 * amp * N ** 3 * np.exp(1j * theta)
    return dk


@pytest.mark.parametrize("N,n_side", [(16, 16), (16, 8), (12, 12)])
def test_lpt_single_mode_closed_form(oracle_mod, N, n_side):
    """delta = A cos(k.x + t)  =>  psi = -A (k/k^2) sin(k.x + t) (psi_k = i k/k^2
    delta_k), sampled on the lattice; x = wrap(q + psi), p = p_factor psi."""
    L, A, th, pf = 2.0 * N, 0.37, 0.6, 1.7
    m = (3, -2, 1)
    dk = _single_mode_spectrum(N, m, A, th)
    pos, mom = oracle_mod.lpt_first_order(dk, L, n_side, pf)
    s = N // n_side
    idx = np.indices((n_side, n_side, n_side)).reshape(3, -1) * s
    kv = np.array(m) * 2 * np.pi / L
    k2 = (kv ** 2).sum()
    phase = 2 * np.pi * (np.array(m)[:, None] * idx).sum(0) / N + th
    for d in range(3):
        psi = -A * kv[d] / k2 * np.sin(phase)
        q = idx[d] // s * (L / n_side)
        x = np.mod(q + psi, L)
        dd = pos[d] - x
        dd -= L * np.round(dd / L)
        assert np.abs(dd).max() < 1e-13 * L
        assert np.allclose(mom[d], pf * psi, rtol=0, atol=1e-13)


def test_lpt_nyquist_derivative_is_zero(oracle_mod):
    """A mode on the x Nyquist index has no x displacement (odd derivative
    zeroed there) while y and z keep -A (k_d/k^2) sin with k^2 including k_x."""
    N, L, A = 16, 16.0, 0.5
    m = (8, 3, 2)
    dk = _single_mode_spectrum(N, m, A, 0.0)
    pos, mom = oracle_mod.lpt_first_order(dk, L, N, 1.0)
    idx = np.indices((N, N, N)).reshape(3, -1)
    kv = np.array(m) * 2 * np.pi / L
    k2 = (kv ** 2).sum()
    phase = 2 * np.pi * (np.array(m)[:, None] * idx).sum(0) / N
    assert np.abs(mom[0]).max() < 1e-14
    for d in (1, 2):
        assert np.allclose(mom[d], -A * kv[d] / k2 * np.sin(phase), rtol=0, atol=1e-13)


def test_lpt_discrete_divergence_matches_numpy(oracle_mod):
    """For a random field, the spectral divergence of the oracle's psi (numpy's
    own FFT) is -delta_k on every mode with no Nyquist index."""
    N, L = 16, 20.0
    rng = np.random.default_rng(4)
    dk = np.fft.rfftn(rng.standard_normal((N, N, N)))
    dk[0, 0, 0] = 0.0
    _, mom = oracle_mod.lpt_first_order(dk, L, N, 1.0)
    m = np.fft.fftfreq(N, 1.0 / N)
    mz = np.arange(N // 2 + 1)
    ks = [m[:, None, None], m[None, :, None], mz[None, None, :]]
    div = np.zeros_like(dk)
    for d in range(3):
        div += 1j * (ks[d] * 2 * np.pi / L) * np.fft.rfftn(mom[d].reshape(N, N, N))
    ok = (np.abs(ks[0]) != N // 2) & (np.abs(ks[1]) != N // 2) & (ks[2] != N // 2)
    ok = np.broadcast_to(ok, dk.shape)
    assert np.abs(div[ok] + dk[ok]).max() < 1e-10 * np.abs(dk).max()
