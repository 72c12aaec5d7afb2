"""Fail-closed behaviour of the asynchronous and graph-launched paths.

Calls made with HY_ASYNC (every timed step and every ensemble replay) return
before anything is checked on the host.  The deposit's routing-count identity
and mass-conservation bound (src/math/cic_mass_conservation.cpp:25-63) are
therefore evaluated on the device and reported by hy_pm_check; the fused
gather raises the reference's lost-product underflow errors
(src/mesh/mass_assignment_gradient.cpp:157-211); an ensemble whose positions
are rewritten outside its graphs either reports stale keys or, after
hy_ensemble_reset, reproduces a fresh run bit for bit.  Every condition is
forced from the host: all kernels exit normally with error bits set."""
import numpy as np
import pytest

from conftest import uniform_particles

pytestmark = pytest.mark.gpu


def _dev(a, gpu):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device=gpu)


def _lost_mass_load(n_tiny=2_000_000, seed=3):
    """One unit mass plus n_tiny masses of 1e-10: in the binary32 fixed-point
    deposit (2^-32 of the largest mass per contribution) every light
    contribution rounds to zero, so 2e-4 of the input mass is not deposited,
    above the binary32 conservation bound (1e-4 relative)."""
    L = 32.0
    n = n_tiny + 1
    pos = uniform_particles(n, L, seed)
    m = np.full(n, 1e-10)
    m[0] = 1.0
    return L, pos, m


def test_async_step_with_lost_mass_raises_on_check(gpu):
    import torch
    import hyowon_amd as H
    from hyowon_amd._lib import HyRuntimeError
    L, pos, m = _lost_mass_load()
    n = pos.shape[1]
    x, y, z = (_dev(p, gpu) for p in pos)
    md = _dev(m, gpu)
    s = H.PMSolver(L, 32, precision="fp32", max_particles=n)
    acc = [torch.zeros(n, dtype=torch.float64, device=gpu) for _ in range(3)]
    with pytest.raises(HyRuntimeError, match="mass"):
        s.compute_forces(x, y, z, md, 0.0, *acc)  # synchronous: raised by the call
    assert all(float(a.abs().sum()) == 0.0 for a in acc)
    st = H.ParticleState(x, y, z, *(torch.zeros(n, dtype=torch.float32, device=gpu) for _ in range(3)))
    out = st.empty_like()
    s.kdk(st, out, 0.0, None, 0.01, 0.01, masses=md, asynchronous=True)  # returns without checking
    with pytest.raises(HyRuntimeError, match="mass"):
        s.check()
    s.check()  # reported once
    # the binary64 deposit keeps every contribution: no error
    s64 = H.PMSolver(L, 32, precision="fp64", max_particles=n)
    st64 = H.ParticleState(x, y, z, *(torch.zeros(n, dtype=torch.float64, device=gpu) for _ in range(3)))
    s64.kdk(st64, st64.empty_like(), 0.0, None, 0.01, 0.01, masses=md, asynchronous=True)
    s64.check()


def _underflow_inputs():
    N, L = 8, 16.0
    rng = np.random.default_rng(11)
    cases = {}
    phi = rng.normal(size=(N, N, N))
    cases["weight"] = (phi, [2e-200], [2e-200], [1.0])
    phi = np.zeros((N, N, N))
    phi[2, 0, 0] = 5e-324
    cases["gradient"] = (phi, [2.5], [0.5], [0.5])
    cases["contribution"] = (rng.normal(size=(N, N, N)) * 1e-315, [2e-10], [0.5], [0.5])
    cases["control"] = (rng.normal(size=(N, N, N)), [2e-10], [3.3], [7.9])
    return N, L, cases


@pytest.mark.parametrize("which", ["weight", "gradient", "contribution", "control"])
def test_gather_lost_products_raise_underflow(gpu, oracle_mod, which):
    """binary64 centered-gradient gather of a given potential: the three
    lost-product conditions raise underflow_error like the oracle (pinned to
    the reference's gather); the control case is bit-exact."""
    import torch
    import hyowon_amd as H
    from hyowon_amd._lib import HyUnderflowError
    N, L, cases = _underflow_inputs()
    phi, x, y, z = cases[which]
    x, y, z = (np.array(v, dtype=np.float64) for v in (x, y, z))
    s = H.PMSolver(L, N, precision="fp64", max_particles=4)
    acc = [torch.zeros(1, dtype=torch.float64, device=gpu) for _ in range(3)]
    args = (_dev(phi, gpu), _dev(x, gpu), _dev(y, gpu), _dev(z, gpu), *acc)
    if which == "control":
        s.gather(*args)
        ref = oracle_mod.gather(phi, x, y, z, L)
        assert np.array_equal(np.stack([a.cpu().numpy() for a in acc]), ref)
    else:
        with pytest.raises(HyUnderflowError, match="centered-gradient"):
            s.gather(*args)
        with pytest.raises(Exception):
            oracle_mod.gather(phi, x, y, z, L)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_ensemble_reset_after_positions_rewritten(gpu, prec):
    """New initial conditions written into an ensemble's buffers: binary64
    members' next recorded step reports the stale keys (logic_error); binary32
    members carry deposit records, which are not re-checked against the
    positions (the documented contract: reset after rewriting).  After
    hy_ensemble_reset the run reproduces a fresh ensemble run bit for bit."""
    import torch
    from hyowon_amd._lib import HyLogicError
    from hyowon_amd.ensemble import BoxClass, EnsembleRun
    classes = [BoxClass(8, 16, 32.0, [1, 2]), BoxClass(12, 24, 48.0, [3])]
    a0, steps = 1.0 / 20.0, 6
    run = EnsembleRun(classes, a0, 1.0, steps, precision=prec)
    ics = [[t.clone() for t in (st.x, st.y, st.z, st.px, st.py, st.pz)] for st in run.states]

    def rewrite():
        torch.cuda.synchronize()
        for st, ic0 in zip(run.states, ics):
            for t, v in zip((st.x, st.y, st.z, st.px, st.py, st.pz), ic0):
                t.copy_(v)
        torch.cuda.synchronize()

    for _ in range(3):
        run.next_step()
    run.ensemble.check()
    rewrite()
    run.ensemble.step(run.K2[0], run.K1[1], run.D[1])  # recorded step on stale routing
    if prec == "fp64":
        with pytest.raises(HyLogicError, match="stale|reused"):
            run.ensemble.check()
    else:
        run.ensemble.check()
    rewrite()
    run.ensemble.reset()
    run.ensemble.step(None, run.K1[0], run.D[0])  # eager re-keying step
    for i in (1, 2):
        run.ensemble.step(run.K2[i - 1], run.K1[i], run.D[i])
    run.ensemble.check()
    torch.cuda.synchronize()
    fresh = EnsembleRun(classes, a0, 1.0, steps, precision=prec)
    for _ in range(3):
        fresh.next_step()
    fresh.ensemble.check()
    torch.cuda.synchronize()
    for a, b in zip(run.states, fresh.states):
        for u, v in zip((a.x, a.y, a.z, a.px, a.py, a.pz), (b.x, b.y, b.z, b.px, b.py, b.pz)):
            assert torch.equal(u, v)
    run.close()
    fresh.close()
