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
