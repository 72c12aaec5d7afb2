"""hy_xplane_order_static (the allocation-free x-plane binning) against
hy_pm_xplane_order through the C ABI: bit-identical order and offsets on two
calls in a row (the second smaller, so the shared scratch holds the first
call's counters), at a size where the chunk count reaches its cap, and its
error paths."""
import ctypes as C

import numpy as np
import pytest

from conftest import clumpy_particles, uniform_particles

pytestmark = pytest.mark.gpu


def _static_fn():
    from hyowon_amd._lib import lib
    L = lib()
    f = L.hy_xplane_order_static
    f.restype = C.c_int
    f.argtypes = [C.c_double, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    e = L.hy_static_last_error
    e.restype = C.c_char_p
    return f, e


def _static(gpu, L, N, x):
    import torch
    f, err = _static_fn()
    n = x.numel()
    order = torch.full((max(n, 1),), -1, dtype=torch.int32, device=gpu)
    offsets = torch.full((N + 1,), -1, dtype=torch.int32, device=gpu)
    torch.cuda.synchronize()
    rc = f(L, N, n, x.data_ptr() if n else None, order.data_ptr() if n else None, offsets.data_ptr(), 0, None)
    if rc != 0:
        raise RuntimeError(f"rc={rc}: {err().decode()}")
    return (order[:n].cpu().numpy().view(np.uint32).copy(), offsets.cpu().numpy().view(np.uint32).copy())


def _context(gpu, L, N, x):
    import hyowon_amd as H
    s = H.PMSolver(L, N, precision="fp32", max_particles=max(x.numel(), 1))
    order, offsets = s.xplane_order(x)
    out = (order.cpu().numpy().view(np.uint32).copy(), offsets.cpu().numpy().view(np.uint32).copy())
    s.close()
    return out


@pytest.mark.parametrize("N,L,n1,n2", [(64, 100.0, 300_000, 70_001), (1024, 512.0, 1 << 24, 1 << 20),
                                       (192, 384.0, 262_144, 1000)])
def test_static_xplane_order_equals_context_entry_point(gpu, N, L, n1, n2):
    import torch
    first = clumpy_particles(n1, L, 3, n_clumps=64, sigma=0.5 * L / N)
    second = uniform_particles(n2, L, 4)
    for pos in (first, second):
        x = torch.as_tensor(pos[0], device=gpu)
        got = _static(gpu, L, N, x)
        ref = _context(gpu, L, N, x)
        print(f"N={N} n={x.numel()}: {int(ref[1][-1])} particles binned")
        assert np.array_equal(got[1], ref[1])
        assert np.array_equal(got[0], ref[0])


def test_static_xplane_order_matches_oracle_with_faces_and_wraps(gpu, oracle_mod):
    import torch
    N, L = 16, 32.0
    dx = L / N
    rng = np.random.default_rng(9)
    x = np.concatenate([rng.integers(0, N, 4000) * dx, [0.0, np.nextafter(L, 0), -dx, L, 3 * L + dx, -1e-300],
                        rng.uniform(-2 * L, 3 * L, 3000)])
    got = _static(gpu, L, N, torch.as_tensor(x, device=gpu))
    off_ref, ord_ref = oracle_mod.xplane_order(x, L, N)
    assert np.array_equal(got[1].astype(np.uint64), off_ref)
    assert np.array_equal(got[0], ord_ref)


def test_static_xplane_order_empty_and_errors(gpu):
    import torch
    f, err = _static_fn()
    order, offsets = _static(gpu, 32.0, 16, torch.zeros(0, dtype=torch.float64, device=gpu))
    assert offsets.tolist() == [0] * 17
    x = torch.as_tensor(np.array([1.0, float("nan"), 3.0]), device=gpu)
    with pytest.raises(RuntimeError, match="rc=1: CIC positions must be finite"):
        _static(gpu, 32.0, 16, x)
    # the same scratch serves a valid call right after the failed one
    order, offsets = _static(gpu, 32.0, 16, torch.as_tensor(np.array([31.0, 1.0, 3.0, 0.5]), device=gpu))
    assert order.tolist() == [1, 3, 2, 0] and offsets[-1] == 4
    with pytest.raises(RuntimeError, match="rc=1: mesh size"):
        _static(gpu, 32.0, 20, torch.as_tensor(np.array([1.0]), device=gpu))
