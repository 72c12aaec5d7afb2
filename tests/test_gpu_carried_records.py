"""Deposit records carried between asynchronous binary32 steps.

An HY_ASYNC hy_pm_kdk call on a uniform-mass binary32 context writes, in its
gather, the deposit record of every drifted particle straight into a
capacity layout (each bucket's count at this step plus an eighth plus a
slack), with an overflow list for records past a bucket's capacity; the next
HY_ASYNC | HY_REUSE_KEYS call deposits those records without re-routing the
particles.  The binary32 deposit sums integers, so a run that carries
records must equal, bit for bit, the same run made of synchronous calls,
which route the particles from their positions (checked keys) every step."""
import numpy as np
import pytest

from conftest import uniform_particles

pytestmark = pytest.mark.gpu


def _run(s, st0, m, K1, D, K2, asynchronous, sync_from=None):
    """S steps as S + 1 fused refreshes; calls from index sync_from on are synchronous."""
    import hyowon_amd as H
    a = H.ParticleState(*(t.clone() for t in (st0.x, st0.y, st0.z, st0.px, st0.py, st0.pz)))
    b = a.empty_like()
    S = len(K1)
    for i in range(S + 1):
        is_async = asynchronous and (sync_from is None or i < sync_from)
        if i < S:
            s.kdk(a, b, m, K2[i - 1] if i else None, K1[i], D[i], asynchronous=is_async, reuse_keys=i > 0)
            a, b = b, a
        else:
            s.kdk(a, b, m, K2[i - 1], None, None, asynchronous=is_async, reuse_keys=True)
            a = H.ParticleState(a.x, a.y, a.z, b.px, b.py, b.pz)
    s.check()
    return a


def _equal(u, v):
    import torch
    return all(torch.equal(p, q) for p, q in zip((u.x, u.y, u.z, u.px, u.py, u.pz), (v.x, v.y, v.z, v.px, v.py, v.pz)))


def test_carried_records_equal_fresh_routing(gpu):
    """1LPT lattice (Philox seed 3), 32^3 particles on a 64^3 mesh, 8 steps."""
    import hyowon_amd as H
    from hyowon_amd import ic
    from hyowon_amd.ensemble import schedule
    n_side, N, L, steps = 32, 64, 64.0, 8
    a0 = 1.0 / 50.0
    pos, mom = ic.lpt_ics(n_side, L, a0, seed=3, device=gpu, rng="philox")
    st0 = H.ParticleState(pos[0], pos[1], pos[2], *(t.float() for t in mom))
    m = ic.particle_mass(n_side, L)
    K1, D, K2 = schedule(a0, 1.0, steps)
    s = H.PMSolver(L, N, precision="fp32", max_particles=n_side ** 3)
    ref = _run(s, st0, m, K1, D, K2, asynchronous=False)
    got = _run(s, st0, m, K1, D, K2, asynchronous=True)
    assert _equal(got, ref)
    # carried records, then synchronous calls (which route afresh)
    mixed = _run(s, st0, m, K1, D, K2, asynchronous=True, sync_from=5)
    assert _equal(mixed, ref)
    s.close()


def test_carried_records_overflow(gpu):
    """Every particle of a uniform set is sent into one 6-cell cube by the
    first drift and out again by the second, so the carried records of the
    second and third deposits overflow their buckets' capacities (made from
    the previous step's counts) by orders of magnitude."""
    import torch
    import hyowon_amd as H
    from hyowon_amd.ensemble import schedule
    n, N, L = 32768, 64, 64.0
    pos = uniform_particles(n, L, 21)
    target = L / 2 + np.random.default_rng(22).uniform(-3.0, 3.0, (3, n))
    K1, D, K2 = schedule(1.0 / 50.0, 1.0, 3)
    d = target - pos
    d -= L * np.round(d / L)
    mom = d / D[0]
    m = 1e-6  # forces far below the momenta's binary32 resolution: the drifts land as aimed
    st0 = H.ParticleState.from_numpy(pos, mom, precision="fp32", device=gpu)
    s = H.PMSolver(L, N, precision="fp32", max_particles=n)
    ref = _run(s, st0, m, K1, D, K2, asynchronous=False)
    got = _run(s, st0, m, K1, D, K2, asynchronous=True)
    # the first drift did gather the particles: positions after one step
    one = H.ParticleState(*(t.clone() for t in (st0.x, st0.y, st0.z, st0.px, st0.py, st0.pz)))
    nxt = one.empty_like()
    s.kdk(one, nxt, m, None, K1[0], D[0])
    c = torch.stack([nxt.x, nxt.y, nxt.z]).cpu().numpy()
    assert np.abs(c - L / 2).max() < 3.5
    assert _equal(got, ref)
    s.close()
