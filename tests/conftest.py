import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libhyowon_pm on the GPU)")
    config.addinivalue_line("markers", "slow: larger GPU parity cases")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build(ref=False)
    return oracle


@pytest.fixture(scope="session")
def ref_lib(oracle_mod):
    if not oracle_mod.ref_available():
        if os.path.isdir(oracle_mod.REFERENCE_TREE):
            oracle_mod.build(ref=True)
        else:
            pytest.skip("reference build (oracle/_ref) not available")
    return oracle_mod.ref()


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import hyowon_amd
    hyowon_amd.lib()  # fails loudly if the HIP library is missing
    return torch.device("cuda:0")


def uniform_particles(n, L, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, L, (3, n))


def clumpy_particles(n, L, seed, n_clumps=16, sigma=0.5, frac=0.5):
    """Uniform background plus Gaussian clumps (the config-1 recipe, scaled)."""
    rng = np.random.default_rng(seed)
    nu = n - int(n * frac)
    pos = np.empty((3, n))
    pos[:, :nu] = rng.uniform(0.0, L, (3, nu))
    centres = rng.uniform(0.0, L, (3, n_clumps))
    k = np.arange(n - nu) % n_clumps
    pos[:, nu:] = centres[:, k] + rng.normal(0.0, sigma, (3, n - nu))
    return np.mod(pos, L)


def face_particles(N, L, seed):
    """Particles exactly on cell faces, at 0, just below L and outside [0,L)."""
    rng = np.random.default_rng(seed)
    dx = L / N
    idx = rng.integers(0, N, (3, 512))
    pos = idx * dx
    extra = np.array([[0.0, np.nextafter(L, 0), -dx, L, L + 0.25 * dx, -1e-300, 3 * L + dx],
                      [np.nextafter(L, 0), 0.0, L, -dx, 0.5 * dx, 1e-300, -2 * L],
                      [L, -dx, 0.0, np.nextafter(L, 0), 7 * dx, 0.0, dx]])
    return np.concatenate([pos, extra], axis=1)
