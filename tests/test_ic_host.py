"""The library's initial-condition host scalars without a GPU: the growth table
(CosmologyModel::D1/f1/D2/f2), the natural cubic spline (math::CubicSpline),
the tabulated linear spectrum (LinearPowerSpectrum::evaluate) and the
RandomField per-mode variance, bit for bit against golden vectors written by
the reference's own classes (tests/golden/make_golden.py), and against the
closed forms the mathematics fixes."""
import ctypes as C
import math
import os

import numpy as np
import pytest
from scipy import integrate

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
pd = C.POINTER(C.c_double)


def P(a):
    return a.ctypes.data_as(pd)


@pytest.fixture(scope="module")
def L():
    import hyowon_amd
    if not os.path.exists(hyowon_amd.LIB_PATH):
        from hyowon_amd import build
        build.build()
    from hyowon_amd._lib import lib
    return lib()


def _cosmo(h=0.7, om=0.3, ob=0.05, ol=0.7, s8=0.8, ns=0.96):
    from hyowon_amd._lib import hy_cosmology
    return hy_cosmology(h, om, ob, ol, s8, ns)


def _growth(L, om, ol, a):
    from hyowon_amd._lib import check
    out = [np.zeros(len(a)) for _ in range(4)]
    check(L.hy_growth(C.byref(_cosmo(om=om, ol=ol)), len(a), P(a), *[P(v) for v in out]))
    return out


@pytest.mark.parametrize("tag", ["lcdm", "low", "eds"])
def test_growth_matches_reference_golden(L, tag):
    g = np.load(os.path.join(GOLDEN, "ic_growth.npz"), allow_pickle=False)
    om, ol = g[f"{tag}_omega"]
    got = _growth(L, om, ol, g["a"])
    for name, v in zip(("D1", "f1", "D2", "f2"), got):
        assert np.array_equal(v, g[f"{tag}_{name}"]), name


def test_growth_closed_forms(L):
    """EdS: D1 = a, f1 = 1, D2 = -3/7 a^2, f2 = 2 exactly.  LCDM: D1 equals the
    growing mode H(a) int_0^a da'/(a' H)^3 normalised at a = 1, and f1 its
    logarithmic derivative, to the RK4 table's accuracy."""
    a = np.geomspace(1e-4, 1.0, 50)
    d1, f1, d2, f2 = _growth(L, 1.0, 0.0, a)
    assert np.array_equal(d1, a) and np.all(f1 == 1.0) and np.all(f2 == 2.0)
    assert np.allclose(d2, -3.0 / 7.0 * a * a, rtol=1e-15, atol=0)
    om, ol = 0.3, 0.7
    E = lambda x: math.sqrt(om / x ** 3 + ol)  # noqa: E731

    def unnorm(x):
        v, _ = integrate.quad(lambda t: 1.0 / (t * E(t)) ** 3, 0.0, x, epsabs=0, epsrel=1e-13, limit=400)
        return E(x) * v

    aa = np.array([0.02, 0.05, 0.1, 0.3, 0.6, 1.0])
    d1, f1, _, _ = _growth(L, om, ol, aa)
    ref = np.array([unnorm(x) for x in aa]) / unnorm(1.0)
    assert np.allclose(d1, ref, rtol=1e-10, atol=0)
    h = 1e-5
    fref = [(math.log(unnorm(x * (1 + h))) - math.log(unnorm(x * (1 - h)))) / (math.log1p(h) - math.log1p(-h))
            for x in aa[:-1]]
    assert np.allclose(f1[:-1], fref, rtol=1e-8, atol=0)


def test_cubic_spline_matches_reference_golden(L):
    from hyowon_amd._lib import check
    g = np.load(os.path.join(GOLDEN, "ic_spline.npz"), allow_pickle=False)
    for x, q, ref in ((g["x"], g["q"], g["out"]), (g["xs"], g["qs"], g["out_s"])):
        out = np.zeros(len(q))
        check(L.hy_cubic_spline(len(x), P(x), P(g["y"]), len(q), P(q), P(out)))
        assert np.array_equal(out, ref)


def test_cubic_spline_properties(L):
    """Knots are reproduced exactly, a straight line is reproduced to roundoff
    (natural end conditions hold for it), and the second derivative vanishes
    at both ends (natural spline)."""
    from hyowon_amd._lib import HyInvalidArgument, check
    x = np.sort(np.random.default_rng(2).uniform(0.0, 10.0, 25))
    y = 3.0 * x - 2.0
    q = np.linspace(x[0], x[-1], 301)
    out = np.zeros(len(q))
    check(L.hy_cubic_spline(len(x), P(x), P(y), len(q), P(q), P(out)))
    assert np.allclose(out, 3.0 * q - 2.0, rtol=0, atol=1e-12)
    yr = np.random.default_rng(3).normal(0, 1, 25)
    kn = np.zeros(25)
    check(L.hy_cubic_spline(25, P(x), P(yr), 25, P(x), P(kn)))
    assert np.array_equal(kn, yr)
    e = 1e-4 * (x[1] - x[0])
    q3 = np.array([x[0], x[0] + e, x[0] + 2 * e])
    o3 = np.zeros(3)
    check(L.hy_cubic_spline(25, P(x), P(yr), 3, P(q3), P(o3)))
    second = (o3[2] - 2 * o3[1] + o3[0]) / (e * e)
    assert abs(second) < 1e-3 * np.abs(np.diff(yr)).max() / (x[1] - x[0]) ** 2
    with pytest.raises(HyInvalidArgument):
        check(L.hy_cubic_spline(2, P(x), P(yr), 1, P(q3), P(o3)))


@pytest.mark.parametrize("name", ["bbks", "camb"])
def test_linear_power_and_mode_variance_match_reference_golden(L, name):
    from hyowon_amd._lib import check, hy_cosmology
    g = np.load(os.path.join(GOLDEN, f"ic_power_{name}.npz"), allow_pickle=False)
    c = hy_cosmology(*g["cosmology"])
    text = str(g["table"]).encode()
    k = g["k"]
    out = np.zeros(len(k))
    check(L.hy_linear_power(C.byref(c), text, 0.0, float(g["z_start"]), len(k), P(k), P(out)))
    assert np.array_equal(out, g["power"])
    n_side, K = int(g["n_side"]), int(g["K"])
    var = np.zeros(3 * (n_side // 2) ** 2 + 1)
    check(L.hy_ic_mode_variance(C.byref(c), text, 0.0, float(g["z_start"]), n_side, float(g["L"]), K, P(var)))
    ref = g["variance"]
    used = ref >= 0.0  # m^2 values that are sums of three squares within the band limit
    assert used.sum() > 0.6 * (3 * K * K)
    assert np.array_equal(var[used], ref[used])


def test_product_bbks_table_is_the_golden_one():
    """The synthetic spectrum table ic.Cosmology().table() writes is the one
    the fixture was generated from (so the GPU ICs use pinned variances)."""
    from hyowon_amd import ic
    g = np.load(os.path.join(GOLDEN, "ic_power_bbks.npz"), allow_pickle=False)
    assert ic.Cosmology().table() == str(g["table"])


def test_linear_power_growth_rescaling_and_header_checks(L):
    """P(k, z) / P(k, 0) = (D1(a)/D1(1))^2 up to the spline's rounding, at the
    knots exactly interpolated; a header that disagrees with the cosmology is
    rejected as the reference does."""
    from hyowon_amd import ic
    from hyowon_amd._lib import HyRuntimeError, check
    c = ic.Cosmology()
    text = c.table().encode()
    k = np.geomspace(1e-3, 10.0, 40)
    p0, p9 = np.zeros(40), np.zeros(40)
    check(L.hy_linear_power(C.byref(c.c_struct()), text, 0.0, 0.0, 40, P(k), P(p0)))
    check(L.hy_linear_power(C.byref(c.c_struct()), text, 0.0, 9.0, 40, P(k), P(p9)))
    d1 = _growth(L, 0.3, 0.7, np.array([0.1]))[0][0]
    assert np.allclose(p9 / p0, d1 * d1, rtol=1e-12, atol=0)
    assert np.allclose(p0, c.bbks_power(k), rtol=1e-6, atol=0)  # spline of the closed form
    bad = ic.Cosmology(omega_m=0.31, omega_lambda=0.69)
    with pytest.raises(HyRuntimeError, match="omega_m"):
        check(L.hy_linear_power(C.byref(bad.c_struct()), text, 0.0, 9.0, 40, P(k), P(p9)))
    with pytest.raises(HyRuntimeError, match="missing 'h'"):
        check(L.hy_linear_power(C.byref(c.c_struct()), text.replace(b"# h=0.7\n", b""), 0.0, 9.0, 40, P(k), P(p9)))


def test_reference_direct(L, ref_lib):
    """Fresh random inputs through both the library and the reference build."""
    from hyowon_amd._lib import check
    rng = np.random.default_rng(12)
    a = rng.uniform(1e-6, 1.0, 500)
    got = _growth(L, 0.27, 0.73, a)
    ref = [np.zeros(500) for _ in range(4)]
    assert ref_lib.ref_growth(0.27, 0.73, 500, P(a), *[P(v) for v in ref]) == 0
    for u, v in zip(got, ref):
        assert np.array_equal(u, v)
    for _ in range(5):
        n = int(rng.integers(3, 60))
        x = np.cumsum(rng.uniform(0.01, 2.0, n))
        y = rng.normal(0, 3, n)
        q = rng.uniform(x[0] - 1, x[-1] + 1, 200)
        o1, o2 = np.zeros(200), np.zeros(200)
        check(L.hy_cubic_spline(n, P(x), P(y), 200, P(q), P(o1)))
        assert ref_lib.ref_cubic_spline(n, P(x), P(y), 200, P(q), P(o2)) == 0
        assert np.array_equal(o1, o2)
