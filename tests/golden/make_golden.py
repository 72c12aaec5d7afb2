"""Write golden vectors from the reference build (oracle/_ref/libhyref.so).

The reference's own classes (CICMassAssignment, GreenFunction, DensityNormalization,
LeapfrogIntegrator, TimeStepper, DriftKickIntegrator, CosmologyModel) are run on
small seeded inputs; inputs and outputs are stored as .npz fixtures next to this
script.  Run in a tree where the reference build exists:

    make -C oracle ref && python tests/golden/make_golden.py

The IC fixtures also read the reference's example CAMB table
(examples/quickstart_camb/linear_cb_z0.dat) as data.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import oracle  # noqa: E402
from conftest import clumpy_particles, face_particles, uniform_particles  # noqa: E402

pd = C.POINTER(C.c_double)


def p(a):
    return a.ctypes.data_as(pd)


def deposit_case(R, name, pos, L, N, masses=None, m=1.0):
    n = pos.shape[1]
    x, y, z = (np.ascontiguousarray(a) for a in pos)
    mesh = np.zeros(N ** 3)
    mm = np.ascontiguousarray(masses) if masses is not None else None
    rc = R.ref_deposit(n, p(x), p(y), p(z), p(mm) if mm is not None else None, m, L, N, p(mesh))
    assert rc == 0
    total = R.ref_exact_sum(p(mm), n) if mm is not None else R.ref_exact_repeated(m, n)
    contrast = mesh.copy()
    assert R.ref_density_contrast(p(contrast), N, L / N, total) == 0
    out = dict(pos=pos, L=L, N=N, mesh=mesh.reshape(N, N, N), contrast=contrast.reshape(N, N, N),
               total_mass=total, uniform_mass=m)
    if masses is not None:
        out["masses"] = masses
    np.savez_compressed(os.path.join(HERE, f"deposit_{name}.npz"), **out)


def main():
    oracle.build(ref=True)
    R = oracle.ref()
    rng = np.random.default_rng(2024)

    deposit_case(R, "uniform_16", uniform_particles(4096, 32.0, 1), 32.0, 16, m=0.7)
    deposit_case(R, "clumpy_32", clumpy_particles(32 ** 3, 32.0, 2, n_clumps=12, sigma=0.5), 32.0, 32, m=1.5)
    fp = face_particles(16, 32.0, 3)
    deposit_case(R, "faces_16", fp, 32.0, 16, masses=rng.uniform(0.1, 3.0, fp.shape[1]))
    deposit_case(R, "uniform_12", uniform_particles(3000, 24.0, 4), 24.0, 12, m=2.0)

    # Green kernel over the half spectrum
    for N in (8, 16, 64, 192):
        L = 2.0 * N
        for dec in (0, 1):
            c = np.ones((N, N, N // 2 + 1), np.complex128)
            assert R.ref_green_apply(N, L, dec, c.ctypes.data_as(pd)) == 0
            if N <= 16:
                idx = np.indices(c.shape).reshape(3, -1)
            else:  # a seeded sample of modes keeps the fixture small
                idx = np.stack([rng.integers(0, s_, 4096) for s_ in c.shape])
                idx[:, :4] = [[0, N // 2, 0, N - 1], [0, N // 2, 1, 0], [0, N // 2, 0, N // 2]]
            np.savez_compressed(os.path.join(HERE, f"green_{N}_{dec}.npz"), index=idx,
                                kernel=c.real[idx[0], idx[1], idx[2]], L=L, N=N)

    # centered-gradient gather of a random potential
    N, L, n = 16, 32.0, 4096
    pos = uniform_particles(n, L, 5)
    phi = rng.standard_normal((N, N, N))
    acc = np.zeros((3, n))
    x, y, z = (np.ascontiguousarray(a) for a in pos)
    assert R.ref_gather(n, p(phi.reshape(-1)), p(x), p(y), p(z), L, N, p(acc[0]), p(acc[1]), p(acc[2])) == 0
    np.savez_compressed(os.path.join(HERE, "gather_16.npz"), pos=pos, phi=phi, acc=acc, L=L, N=N)

    # time integrals, H(a), geometric mean, boundary schedules of the five configs
    a = np.concatenate([np.geomspace(1 / 50, 1.0, 40), rng.uniform(0.01, 1.0, 40)])
    b = np.concatenate([np.geomspace(1 / 50, 1.0, 40)[1:], [1.0], rng.uniform(0.01, 1.0, 40)])
    drift, kick, hub, gm = [], [], [], []
    for a1, a2 in zip(a, b):
        d = C.c_double(); k = C.c_double()
        assert R.ref_drift_kick(a1, a2, 0.3, 0.7, C.byref(d), C.byref(k)) == 0
        drift.append(d.value); kick.append(k.value)
        hub.append(R.ref_hubble(a1, 0.3, 0.7))
        gm.append(R.ref_geometric_mean(min(a1, a2), max(a1, a2)) if a1 != a2 else a1)
    sched = {}
    for name, (a0, a1, ns) in {"c0": (1 / 50, 1.0, 32), "c2": (1 / 50, 1.0, 64), "c3": (1 / 50, 1.0, 32),
                               "c4": (1 / 20, 1.0, 16)}.items():
        dl = math.log(a1 / a0) / ns
        out = np.zeros(1000)
        cnt = C.c_uint64()
        s = np.array([1.0])
        assert R.ref_time_boundaries(a0, a1, dl, p(s), 1, p(out), 1000, C.byref(cnt)) == 0
        sched[f"boundaries_{name}"] = out[: cnt.value]
        sched[f"args_{name}"] = np.array([a0, a1, dl])
    np.savez_compressed(os.path.join(HERE, "time_integrals.npz"), a1=a, a2=b, drift=np.array(drift),
                        kick=np.array(kick), hubble=np.array(hub), geometric_mean=np.array(gm),
                        G=R.ref_gravitational_constant(), rho_crit0=R.ref_rho_crit0(), **sched)

    # one-particle K1 + periodic drift through particle_update.hpp, incl. wraps
    L = 10.0
    cases = []
    for _ in range(200):
        pos = rng.uniform(-2 * L, 3 * L, 3)
        mom = rng.normal(0, 50, 3)
        acc = rng.normal(0, 20, 3)
        K, D = rng.uniform(0.001, 0.1), rng.uniform(0.0, 0.5)
        pp, mm = pos.copy(), mom.copy()
        assert R.ref_kick_drift_one(p(pp), p(mm), p(acc), K, D, L) == 0
        cases.append(np.concatenate([pos, mom, acc, [K, D], pp, mm]))
    np.savez_compressed(os.path.join(HERE, "kick_drift.npz"), cases=np.array(cases), L=L)
    ic_cases(R)
    print("golden vectors written to", HERE)


REF_CAMB_TABLE = os.path.join(oracle.REFERENCE_TREE, "examples", "quickstart_camb", "linear_cb_z0.dat")


def ic_cases(R):
    """CosmologyModel growth, CubicSpline, LinearPowerSpectrum::evaluate and the
    RandomField mode variance, from the reference's classes."""
    import tempfile
    from hyowon_amd import ic
    rng = np.random.default_rng(77)
    a = np.concatenate([np.geomspace(1e-7, 1.0, 120), rng.uniform(1e-3, 1.0, 60), [1 / 50, 1 / 20, 1.0]])
    out = {"a": a}
    for tag, (om, ol) in {"lcdm": (0.3, 0.7), "low": (0.25, 0.75), "eds": (1.0, 0.0)}.items():
        g = [np.zeros(len(a)) for _ in range(4)]
        assert R.ref_growth(om, ol, len(a), p(a), *[p(v) for v in g]) == 0
        out.update({f"{tag}_{k}": v for k, v in zip(("D1", "f1", "D2", "f2"), g)})
        out[f"{tag}_omega"] = np.array([om, ol])
    np.savez_compressed(os.path.join(HERE, "ic_growth.npz"), **out)

    x = np.sort(rng.uniform(-3.0, 5.0, 40))
    y = rng.normal(0.0, 1.0, 40)
    q = np.concatenate([x, rng.uniform(-4.0, 6.0, 300)])
    xs = 1e-200 * np.arange(1.0, 41.0)  # knot spacing whose square underflows: normalised build
    qs = np.concatenate([xs, rng.uniform(0.5e-200, 41e-200, 100)])
    o, os_ = np.zeros(len(q)), np.zeros(len(qs))
    assert R.ref_cubic_spline(40, p(x), p(y), len(q), p(q), p(o)) == 0
    assert R.ref_cubic_spline(40, p(xs), p(y), len(qs), p(qs), p(os_)) == 0
    np.savez_compressed(os.path.join(HERE, "ic_spline.npz"), x=x, y=y, q=q, out=o, xs=xs, qs=qs, out_s=os_)

    def power_case(name, cosmo6, text, z_start, L, n_side, K, k):
        with tempfile.NamedTemporaryFile("w", suffix=".dat", delete=False) as f:
            f.write(text)
            path = f.name
        c6 = np.array(cosmo6, np.float64)
        pk = np.zeros(len(k))
        assert R.ref_linear_power(p(c6), path.encode(), 0.0, z_start, L, n_side, n_side, K, len(k), p(k), p(pk)) == 0
        half = n_side // 2
        var = np.zeros(3 * half * half + 1)
        assert R.ref_mode_variance(p(c6), path.encode(), 0.0, z_start, L, n_side, n_side, K, p(var)) == 0
        os.unlink(path)
        np.savez_compressed(os.path.join(HERE, f"ic_power_{name}.npz"), cosmology=c6, table=np.array(text),
                            z_start=z_start, L=L, n_side=n_side, K=K, k=k, power=pk, variance=var)

    c = ic.Cosmology()
    power_case("bbks", [c.h, c.omega_m, c.omega_b, c.omega_lambda, c.sigma8, c.n_s], c.table(), 49.0, 128.0, 64, 31,
               np.concatenate([np.geomspace(1e-4, 1e2, 400), rng.uniform(1e-4, 1e2, 100)]))
    # the reference's CAMB example table (data), rows rewritten with 17 digits
    head, rows = [], []
    for line in open(REF_CAMB_TABLE):
        (head if line.lstrip().startswith("#") else rows).append(line.rstrip("\n"))
    kp = np.array([[float(v) for v in r.split()] for r in rows if r.strip()])
    text = "\n".join(head + [f"{kv:.17e} {pv:.17e}" for kv, pv in kp]) + "\n"
    power_case("camb", [0.7, 0.3, 0.05, 0.7, 0.8, 1.0], text, 9.0, 64.0, 32, 15,
               np.concatenate([np.geomspace(kp[0, 0], kp[-1, 0], 300), kp[::50, 0]]))


if __name__ == "__main__":
    main()
