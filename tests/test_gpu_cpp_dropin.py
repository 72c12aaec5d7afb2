"""The C++20 drop-in header (include/hyowon/pm_solver.hpp) used from a compiled
program the way a caller of the reference's PMSolver / TimeStepper /
DriftKickIntegrator would use it; forces checked against the oracle and the
schedule factors against the reference's golden vectors."""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import clumpy_particles

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "dropin_force")


@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 1e-4)])
def test_cpp_dropin_force_and_schedule(gpu, oracle_mod, tmp_path, prec, tol):
    if not os.path.exists(EXE):
        pytest.skip("C++ drop-in example not built (python -m hyowon_amd.build)")
    N, L, n = 32, 32.0, 20000
    pos = clumpy_particles(n, L, 8, n_clumps=10, sigma=0.5)
    m = 1.25
    src = tmp_path / "in.bin"
    out = tmp_path / "out.bin"
    with open(src, "wb") as f:
        np.array([n], np.uint64).tofile(f)
        pos.astype(np.float64).tofile(f)
    subprocess.run([EXE, str(src), str(out), str(L), str(N), str(prec), str(m)], check=True,
                   timeout=120)
    raw = np.fromfile(out, dtype=np.float64)
    acc = raw[: 3 * n].reshape(3, n)
    steps = int(raw[3 * n:3 * n + 1].view(np.uint64)[0])
    fac = raw[3 * n + 1:3 * n + 1 + 3 * steps].reshape(steps, 3)
    rest = raw[3 * n + 1 + 3 * steps:]
    acc2 = rest[: 3 * n].reshape(3, n)
    energy = rest[3 * n]
    ref, phi = oracle_mod.pm_force(pos[0], pos[1], pos[2], L, N, uniform_mass=m, return_phi=True)
    err = math.sqrt(((acc - ref) ** 2).sum() / (ref ** 2).sum())
    assert err < tol, err
    err2 = math.sqrt(((acc2 - ref) ** 2).sum() / (ref ** 2).sum())
    assert err2 < tol, err2
    e_ref = oracle_mod.potential_energy(phi, pos[0], pos[1], pos[2], L, uniform_mass=m)
    assert abs(energy - e_ref) <= (1e-12 if prec == 0 else 1e-5) * abs(e_ref), (energy, e_ref)
    g = np.load(os.path.join(ROOT, "tests", "golden", "time_integrals.npz"), allow_pickle=False)
    b = g["boundaries_c0"]
    assert steps == len(b) - 1 == 32
    for i in range(steps):
        a0, a1 = b[i], b[i + 1]
        ah = oracle_mod.lib().or_geometric_mean(a0, a1)
        assert fac[i, 0] == oracle_mod.kick_factor(a0, ah)
        assert fac[i, 1] == oracle_mod.drift_factor(a0, a1)
        assert fac[i, 2] == oracle_mod.kick_factor(ah, a1)
