"""The CPU oracle against golden vectors written by the reference's own classes
(tests/golden/make_golden.py over oracle/_ref).  Bit-exact where the reference
is deterministic (deposit, density source, Green kernel, gather, KDK update,
time integrals, schedules)."""
import ctypes as C
import glob
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.mark.parametrize("name", sorted(os.path.basename(p) for p in glob.glob(os.path.join(GOLDEN, "deposit_*.npz"))))
def test_deposit_and_density_source(oracle_mod, name):
    g = _load(name)
    pos, L, N = g["pos"], float(g["L"]), int(g["N"])
    masses = g["masses"] if "masses" in g else None
    mesh = oracle_mod.deposit(pos[0], pos[1], pos[2], L, N, masses=masses, uniform_mass=float(g["uniform_mass"]))
    assert np.array_equal(mesh, g["mesh"])
    if masses is not None:
        total = oracle_mod.lib().or_exact_sum(np.ascontiguousarray(masses).ctypes.data_as(oracle_mod._pd), len(masses))
    else:
        total = oracle_mod.lib().or_exact_repeated(float(g["uniform_mass"]), pos.shape[1])
    assert total == float(g["total_mass"])
    assert np.array_equal(oracle_mod.density_contrast(mesh, L, N, total), g["contrast"])


@pytest.mark.parametrize("name", sorted(os.path.basename(p) for p in glob.glob(os.path.join(GOLDEN, "green_*.npz"))))
def test_green_kernel(oracle_mod, name):
    g = _load(name)
    N, L = int(g["N"]), float(g["L"])
    dec = int(name.split("_")[2].split(".")[0])
    idx = g["index"]
    got = np.array([oracle_mod.lib().or_green_kernel(N, L, dec, int(i), int(j), int(k))
                    for i, j, k in idx.T])
    assert np.array_equal(got, g["kernel"])


def test_gather(oracle_mod):
    g = _load("gather_16.npz")
    pos = g["pos"]
    acc = oracle_mod.gather(g["phi"], pos[0], pos[1], pos[2], float(g["L"]))
    assert np.array_equal(acc, g["acc"])


def test_time_integrals_and_schedules(oracle_mod):
    g = _load("time_integrals.npz")
    L = oracle_mod.lib()
    for a1, a2, d, k, h, gm in zip(g["a1"], g["a2"], g["drift"], g["kick"], g["hubble"], g["geometric_mean"]):
        assert oracle_mod.drift_factor(a1, a2) == d
        assert oracle_mod.kick_factor(a1, a2) == k
        assert L.or_hubble(a1, 0.3, 0.7) == h
        if a1 != a2:
            assert L.or_geometric_mean(min(a1, a2), max(a1, a2)) == gm
    assert L.or_gravitational_constant() == float(g["G"])
    assert L.or_rho_crit0() == float(g["rho_crit0"])
    for name, steps in (("c0", 32), ("c2", 64), ("c3", 32), ("c4", 16)):
        a0, a1, dl = g[f"args_{name}"]
        b = oracle_mod.time_boundaries(a0, a1, dl, [1.0])
        assert np.array_equal(b, g[f"boundaries_{name}"])
        assert len(b) - 1 == steps


def test_kick_drift_one_particle(oracle_mod):
    g = _load("kick_drift.npz")
    L = float(g["L"])
    for row in g["cases"]:
        pos, mom, acc = row[0:3], row[3:6], row[6:9]
        K, D = row[9], row[10]
        out = oracle_mod.kick_drift(*[[v] for v in pos], *[[v] for v in mom], *[[v] for v in acc], K, D, L)
        got = np.array([o[0] for o in out])
        assert np.array_equal(got[:3], row[11:14])
        assert np.array_equal(got[3:], row[14:17])
