"""Single PM force evaluation at full size: 256^3 particles (half uniform,
half in 10^4 Gaussian clumps of sigma 0.5 Mpc/h), 512^3 CIC mesh, box
256 Mpc/h.  Binary64 GPU forces against the oracle, binary32 GPU forces by
RMS relative error against the oracle, and the cell-sort permutations (the
deposit buckets and the reference x-plane bins) bit-exact."""
import math

import numpy as np
import pytest

from conftest import clumpy_particles

pytestmark = pytest.mark.gpu

N, L, NSIDE = 512, 256.0, 256


@pytest.fixture(scope="module")
def case(oracle_mod):
    n = NSIDE ** 3
    pos = clumpy_particles(n, L, 1, n_clumps=10000, sigma=0.5)
    m = 0.3 * 27.75 * (L / NSIDE) ** 3  # Omega_m rho_crit0 dx^3 in 1e10 Msun/h (scale only)
    ref = oracle_mod.pm_force(pos[0], pos[1], pos[2], L, N, uniform_mass=m)
    return pos, m, ref


def _forces(gpu, pos, m, prec):
    import torch
    import hyowon_amd as H
    n = pos.shape[1]
    s = H.PMSolver(L, N, precision=prec, max_particles=n)
    x, y, z = (torch.as_tensor(p, device=gpu) for p in pos)
    acc = [torch.zeros(n, dtype=torch.float64, device=gpu) for _ in range(3)]
    s.compute_forces(x, y, z, None, m, *acc)
    out = np.stack([a.cpu().numpy() for a in acc])
    s.close()
    return out


def test_config1_force_fp64(gpu, case):
    pos, m, ref = case
    got = _forces(gpu, pos, m, "fp64")
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 1e-10, err


def test_config1_force_fp32_rms(gpu, case):
    pos, m, ref = case
    got = _forces(gpu, pos, m, "fp32")
    d2 = ((got - ref) ** 2).sum(0)
    g2 = (ref ** 2).sum(0)
    rms = math.sqrt(d2.sum() / g2.sum())
    p99 = np.quantile(np.sqrt(d2 / np.maximum(g2, 1e-300)), 0.99)
    print(f"config-1 FP32 force: RMS rel {rms:.3e}, p99 per-particle rel {p99:.3e}")
    assert rms < 1e-4, rms


def test_config1_permutations_bit_exact(gpu, oracle_mod, case):
    import torch
    import hyowon_amd as H
    pos, m, _ = case
    n = pos.shape[1]
    s = H.PMSolver(L, N, precision="fp32", max_particles=n)
    x, y, z = (torch.as_tensor(p, device=gpu) for p in pos)
    order, offsets = s.bucket_order(x, y, z)
    T = int(s.info().tile_edge)
    off_ref, ord_ref = oracle_mod.bucket_order(pos[0], pos[1], pos[2], L, N, T)
    assert np.array_equal(offsets.cpu().numpy().astype(np.uint64), off_ref)
    assert np.array_equal(order.cpu().numpy().astype(np.uint32), ord_ref)
    order, offsets = s.xplane_order(x)
    off_ref, ord_ref = oracle_mod.xplane_order(pos[0], L, N)
    assert np.array_equal(offsets.cpu().numpy().astype(np.uint64), off_ref)
    assert np.array_equal(order.cpu().numpy().astype(np.uint32), ord_ref)
    s.close()
