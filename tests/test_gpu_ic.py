"""GPU mode-keyed Gaussian initial field (hy_ic_gaussian_modes) against the
oracle's restatement, which is itself pinned bit-exact to the reference's
philox_rng.cpp.  The Philox words and uniforms are integer/exact; Box-Muller
uses the GPU's log/sin/cos, so values agree to a few ulp."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _gpu_modes(N, seed, max_mode, var, flags):
    import torch
    from hyowon_amd._lib import lib
    v = torch.as_tensor(var, device="cuda:0")
    out = torch.empty((N, N, N // 2 + 1), dtype=torch.complex128, device="cuda:0")
    rc = lib().hy_ic_gaussian_modes(N, seed, max_mode, v.data_ptr(), out.data_ptr(), flags, 0, None)
    assert rc == 0, lib().hy_ic_last_error()
    return out.cpu().numpy()


@pytest.mark.parametrize("N", [8, 12, 32, 48])
@pytest.mark.parametrize("flags", [0, 1, 2])
def test_gpu_gaussian_modes_match_oracle(gpu, oracle_mod, N, flags):
    rng = np.random.default_rng(N)
    var = rng.uniform(0.5, 2.0, 3 * (N // 2) ** 2 + 1)
    K = N // 2 if N != 32 else 11
    seed = 0x1234_5678_9ABC_DEF0 + N
    got = _gpu_modes(N, seed, K, var, flags)
    ref = oracle_mod.gaussian_modes(N, seed, K, var, flags)
    assert np.array_equal(got == 0, ref == 0)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 1e-14, err


def test_gpu_modes_are_keyed_by_integer_mode(gpu):
    """A mode's draw depends only on (seed, integer mode): the shared modes of
    a 16^3 and a 32^3 IC mesh are identical when the per-mode variance is."""
    K = 7
    var16 = np.linspace(1.0, 2.0, 3 * 8 ** 2 + 1)
    var32 = np.concatenate([var16, np.ones(3 * 16 ** 2 + 1 - var16.size)])
    a = _gpu_modes(16, 42, K, var16, 0)
    b = _gpu_modes(32, 42, K, var32, 0)
    for mx in range(-K, K + 1):
        for my in range(-K, K + 1):
            for mz in range(0, K + 1):
                assert a[mx % 16, my % 16, mz] == b[mx % 32, my % 32, mz]


def test_philox_lpt_displacements_match_numpy(gpu):
    """1LPT ICs from the Philox field: the lattice displacements equal
    numpy's inverse r2c transform of i k_x / k^2 delta_k (Nyquist plane of the
    derivative axis zeroed), and the momenta are a^2 H f times them."""
    import torch
    from hyowon_amd import ic
    n, L, a = 32, 64.0, 1.0 / 50.0
    dk = ic.philox_modes(n, L, a, seed=5)
    pos, mom = ic.lpt_from_delta_k(dk, L, a)
    D = dk.cpu().numpy()
    m = np.fft.fftfreq(n, 1.0 / n)
    kx = (m * 2 * np.pi / L)[:, None, None]
    mz = np.arange(n // 2 + 1)
    k2 = ((m[:, None, None] ** 2 + m[None, :, None] ** 2 + mz[None, None, :] ** 2) * (2 * np.pi / L) ** 2)
    k2[0, 0, 0] = 1.0
    kxn = np.where(np.abs(m)[:, None, None] == n // 2, 0.0, kx)
    psi = np.fft.irfftn(1j * kxn / k2 * D, s=(n, n, n), axes=(0, 1, 2))
    q = np.arange(n) * (L / n)
    x_ref = np.mod(q[:, None, None] + psi, L).reshape(-1)
    d = pos[0].cpu().numpy() - x_ref
    d -= L * np.round(d / L)
    assert np.abs(d).max() < 1e-9
    vel = ic.momentum_factor(a)
    assert np.allclose(mom[0].cpu().numpy(), vel * psi.reshape(-1), rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("n", [16, 32, 48, 128])
def test_hip_lpt_matches_numpy(gpu, n):
    """hy_ic_lpt on the library's FFT passes: lattice displacements equal
    numpy's inverse r2c transform of i k_x / k^2 delta_k (Nyquist index of the
    derivative axis zeroed) on every axis, positions wrap(q + psi), momenta
    a^2 H f psi."""
    from hyowon_amd import ic
    L, a = 2.0 * n, 1.0 / 50.0
    dk = ic.philox_modes(n, L, a, seed=11)
    pos, mom = ic.lpt_from_delta_k_hip(dk, L, a)
    D = dk.cpu().numpy()
    m = np.fft.fftfreq(n, 1.0 / n)
    mz = np.arange(n // 2 + 1)
    k1 = m * 2 * np.pi / L
    kzz = mz * 2 * np.pi / L
    k2 = k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + kzz[None, None, :] ** 2
    k2[0, 0, 0] = 1.0
    vel = ic.momentum_factor(a)
    q = np.arange(n) * (L / n)
    for axis in range(3):
        kd = [k1[:, None, None], k1[None, :, None], kzz[None, None, :]][axis]
        kd = np.where(np.abs([m[:, None, None], m[None, :, None], mz[None, None, :]][axis]) == n // 2, 0.0, kd)
        psi = np.fft.irfftn(1j * kd / k2 * D, s=(n, n, n), axes=(0, 1, 2))
        qa = [q[:, None, None], q[None, :, None], q[None, None, :]][axis]
        x_ref = np.mod(qa + psi, L).reshape(-1)
        d = pos[axis].cpu().numpy() - x_ref
        d -= L * np.round(d / L)
        assert np.abs(d).max() < 1e-10 * L
        assert np.allclose(mom[axis].cpu().numpy(), vel * psi.reshape(-1), rtol=1e-9, atol=1e-12 * np.abs(vel * psi).max())
        assert (pos[axis] >= 0).all() and (pos[axis] < L).all()


def test_hip_lpt_sub_sampled_lattice(gpu):
    """A 16^3 lattice on a 32^3 IC mesh takes psi at every second cell."""
    from hyowon_amd import ic
    n, N, L, a = 16, 32, 40.0, 1.0 / 20.0
    dk = ic.philox_modes(N, L, a, seed=4)
    pos_full, _ = ic.lpt_from_delta_k_hip(dk, L, a)
    pos_sub, _ = ic.lpt_from_delta_k_hip(dk, L, a, n_side=n)
    for axis in range(3):
        full = pos_full[axis].cpu().numpy().reshape(N, N, N)[::2, ::2, ::2].reshape(-1)
        assert np.array_equal(pos_sub[axis].cpu().numpy(), full)


@pytest.mark.parametrize("n,n_side", [(16, 16), (32, 32), (48, 48), (64, 32), (128, 128)])
def test_hip_lpt_matches_oracle(gpu, oracle_mod, n, n_side):
    """hy_ic_lpt against the oracle's restatement of the reference's first-order
    LPT (src/ic/lpt_displacement.cpp:202-384) on the same Philox modes, incl. a
    lattice coarser than the IC mesh."""
    from hyowon_amd import ic
    L, a = 3.0 * n, 1.0 / 50.0
    dk = ic.philox_modes(n, L, a, seed=5)
    pf = ic.momentum_factor(a)
    pos, mom = ic.lpt_from_delta_k_hip(dk, L, a, n_side=n_side)
    rpos, rmom = oracle_mod.lpt_first_order(dk.cpu().numpy(), L, n_side, pf)
    for ax in range(3):
        d = pos[ax].cpu().numpy() - rpos[ax]
        d -= L * np.round(d / L)
        disp = np.abs(rmom[ax]).max() / abs(pf)
        assert np.abs(d).max() <= 1e-12 * disp + 1e-13 * L, ax
        assert np.abs(mom[ax].cpu().numpy() - rmom[ax]).max() <= 1e-12 * np.abs(rmom[ax]).max(), ax
