"""GPU TreePM long-range PM force (HY_METHOD_TREEPM_LONG_RANGE) against the
oracle's restatement, itself pinned bit-exact to the reference's
GreenFunction (TreePM branch) and CICMassAssignment::interpolate_add.

CIC interpolation is bit-exact in binary64 (same loads, same order).  The
force goes through the FFT, so binary64 forces agree to roundoff (1e-9 of the
RMS force) and binary32 forces to 1e-4 RMS relative."""
import numpy as np
import pytest

from conftest import clumpy_particles, uniform_particles

pytestmark = pytest.mark.gpu


def _solver(L, N, prec, cap, rs):
    from hyowon_amd import PMSolver
    return PMSolver(L, N, precision=prec, max_particles=cap, method="treepm_long_range", split_scale=rs)


def _dev(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda:0")


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("N,L", [(16, 32.0), (48, 71.0), (128, 128.0)])
def test_interpolate_matches_oracle(gpu, oracle_mod, prec, N, L):
    import torch
    rng = np.random.default_rng(N)
    mesh = rng.normal(size=(N, N, N))
    n = 20000
    pos = rng.uniform(-L, 2 * L, (3, n))
    pos[:, :6] = np.array([[0.0, L, np.nextafter(L, 0), L - L / N, L / N * 3, -1e-300]] * 3)
    s = _solver(L, N, prec, n, 2.0 * L / N)
    base = rng.normal(size=n)
    out = _dev(base)
    mdev = torch.as_tensor(mesh, device="cuda:0")
    print(f"interp prec={prec} N={N}", flush=True)
    s.interpolate(mdev, *(_dev(p) for p in pos), out)
    ref = oracle_mod.interpolate_add(mesh if prec == "fp64" else mesh.astype(np.float32).astype(np.float64),
                                     *pos, L, out=base)
    got = out.cpu().numpy()
    if prec == "fp64":
        assert np.array_equal(got, ref)
    else:
        assert np.abs(got - ref).max() < 1e-5 * np.abs(mesh).max()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("fp32", 1e-4)])
@pytest.mark.parametrize("N,L,rs_cells,kind", [(16, 32.0, 1.25, "uniform"), (48, 96.0, 1.5, "clumpy"),
                                                (128, 128.0, 1.25, "clumpy"), (192, 192.0, 2.0, "uniform")])
def test_treepm_force_matches_oracle(gpu, oracle_mod, prec, tol, N, L, rs_cells, kind):
    import torch
    n = 4 * N ** 2 if N > 16 else 3000
    pos = uniform_particles(n, L, N) if kind == "uniform" else clumpy_particles(n, L, N, sigma=0.04 * L)
    rs = rs_cells * L / N
    m = 2.5
    print(f"treepm prec={prec} N={N} n={n}", flush=True)
    s = _solver(L, N, prec, n, rs)
    acc = [torch.zeros(n, dtype=torch.float64, device="cuda:0") for _ in range(3)]
    s.compute_forces(*(_dev(p) for p in pos), None, m, *acc)
    got = np.stack([a.cpu().numpy() for a in acc])
    ref = np.stack(oracle_mod.pm_force_treepm(*pos, L, N, rs, uniform_mass=m))
    rms = np.sqrt((ref ** 2).sum(0).mean())
    err = np.sqrt(((got - ref) ** 2).sum(0).mean()) / rms
    assert err < tol, err
    if prec == "fp64":
        # momentum conservation of the CIC deposit / spectral gradient / CIC interpolation
        for a in got:
            assert abs(a.sum()) < 1e-9 * np.abs(a).sum()


# fp64 at N = 32 runs the LDS Stockham passes; fp32 at N = 128 runs the
# two-level dual passes (both x outputs, and the -i k_y and plain y^-1 outputs,
# from one read each), so each of F_x, F_y, F_z is checked on its own.
@pytest.mark.parametrize("prec,N,L,tol", [("fp64", 32, 40.0, 1e-10), ("fp32", 128, 160.0, 1e-4),
                                          ("fp32", 256, 256.0, 1e-4)])
def test_force_meshes_match_oracle(gpu, oracle_mod, prec, N, L, tol):
    import torch
    n = {32: 6000, 128: 40000}.get(N, 200000)
    pos = clumpy_particles(n, L, 3, sigma=L / 40)
    rs = 1.7 * L / 40
    print(f"force meshes prec={prec} N={N} n={n}", flush=True)
    s = _solver(L, N, prec, n, rs)
    x, y, z = (_dev(p) for p in pos)
    src = s.density(x, y, z, None, 1.3)
    fx, fy, fz = s.force_meshes(src)
    *_, meshes = oracle_mod.pm_force_treepm(*pos, L, N, rs, uniform_mass=1.3, meshes=True)
    for got, ref in zip((fx, fy, fz), meshes):
        g = got.double().cpu().numpy()
        err = np.abs(g - ref).max() / np.abs(ref).max()
        assert err < tol, err


def test_treepm_fp32_matches_fp64_at_1024(gpu):
    """The benchmarked size, 1024^3: the binary32 dual passes (one read, two
    outputs in x and in y^-1) against the binary64 single-output passes on
    the same particles, two independent kernel paths."""
    import torch
    N, L, n = 1024, 512.0, 2_000_000
    pos = clumpy_particles(n, L, 11, n_clumps=64, sigma=4.0)
    rs = 1.25 * L / N
    x, y, z = (_dev(p) for p in pos)
    got = {}
    for prec in ("fp32", "fp64"):
        print(f"treepm 1024 prec={prec} n={n}", flush=True)
        s = _solver(L, N, prec, n, rs)
        acc = [torch.zeros(n, dtype=torch.float64, device="cuda:0") for _ in range(3)]
        s.compute_forces(x, y, z, None, 1.0, *acc)
        got[prec] = torch.stack(acc).cpu().numpy()
        s.close()
        del acc
        torch.cuda.empty_cache()
    ref = got["fp64"]
    rms = np.sqrt((ref ** 2).sum(0).mean())
    for axis in range(3):
        err = np.sqrt(((got["fp32"][axis] - ref[axis]) ** 2).mean()) / rms
        assert err < 1e-4, (axis, err)


def test_treepm_context_rules(gpu):
    import torch
    from hyowon_amd import PMSolver, ParticleState
    from hyowon_amd._lib import HyInvalidArgument
    with pytest.raises(HyInvalidArgument, match="split scale"):
        PMSolver(32.0, 16, precision="fp64", max_particles=10, method="treepm_long_range", split_scale=0.0)
    with pytest.raises(HyInvalidArgument, match="single box"):
        PMSolver(32.0, 16, precision="fp64", max_particles=20, method="treepm_long_range", split_scale=1.0,
                 boxes=2)
    s = _solver(32.0, 16, "fp64", 100, 2.0)
    pos = uniform_particles(100, 32.0, 2)
    st = ParticleState(*(_dev(p) for p in pos), *(torch.zeros(100, dtype=torch.float64, device="cuda:0")
                                                 for _ in range(3)))
    with pytest.raises(HyInvalidArgument, match="pure-PM"):
        s.kdk(st, st.empty_like(), 1.0, None, 0.1, 0.1)
    pm = PMSolver(32.0, 16, precision="fp64", max_particles=100)
    with pytest.raises(HyInvalidArgument, match="TreePM long-range context"):
        pm.force_meshes(torch.zeros((16, 16, 16), dtype=torch.float64, device="cuda:0"))
    # the power spectrum and the density work on TreePM contexts too
    got = s.power_spectrum(*(_dev(p) for p in pos), 4, False, False, 1.0)
    assert sum(b.mode_count for b in got) > 0
