"""1LPT initial conditions for benchmarks and tests.

Builds the workload the PM step is measured on: a corner lattice of particles
(index (i*N+j)*N+k, q = (i,j,k)*L/N; src/ic/initial_conditions.cpp:296-312)
displaced by the first-order LPT field of a Gaussian density
(src/ic/lpt_displacement.cpp:202-384), momenta p = a^2 H(a) f1(a) psi.

The linear spectrum follows the reference's input path: a two-column table
"k P(k)" at z = 0 with its semantic header (src/ic/linear_power_spectrum.cpp),
rescaled by the growth table's (D1(a_start)/D1(1))^2 and interpolated as a
natural cubic spline in ln k -- all evaluated by libhyowon_pm (hy_linear_power,
hy_ic_mode_variance, hy_growth).  The table written here is the synthetic BBKS
spectrum P(k) = A k^n_s T(k)^2 (Bardeen et al. 1986, q = k/Gamma) over
1e-4..1e2 h/Mpc, A normalised to sigma_8 at z = 0.

rng="philox" draws the reference's mode-keyed Philox field
(src/ic/random_field.cpp); rng="torch" draws real-space white noise with
torch's generator (the differentiable runs' noise -> IC map).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from functools import lru_cache

import numpy as np
from scipy import integrate


@dataclass(frozen=True)
class Cosmology:
    omega_m: float = 0.3
    omega_lambda: float = 0.7
    h: float = 0.7
    omega_b: float = 0.05
    sigma8: float = 0.8
    n_s: float = 0.96
    gamma: float = 0.21

    def transfer(self, k):
        q = np.asarray(k, dtype=np.float64) / self.gamma
        out = np.ones_like(q)
        nz = q > 0
        qq = q[nz]
        out[nz] = (np.log1p(2.34 * qq) / (2.34 * qq)
                   * (1 + 3.89 * qq + (16.1 * qq) ** 2 + (5.46 * qq) ** 3 + (6.71 * qq) ** 4) ** -0.25)
        return out

    def _unnormalised(self, k):
        return np.asarray(k, np.float64) ** self.n_s * self.transfer(k) ** 2

    @property
    def amplitude(self) -> float:
        """A with sigma_8(z=0) = sigma8 for the closed-form BBKS shape."""
        return _bbks_amplitude(self.n_s, self.gamma, self.sigma8)

    def bbks_power(self, k):
        """Closed-form BBKS P(k) at z = 0 (the table's content)."""
        return self.amplitude * self._unnormalised(k)

    def table(self) -> str:
        """The spectrum table (text) with the semantic header the reference requires."""
        return _bbks_table(self)

    def c_struct(self):
        from ._lib import hy_cosmology
        return hy_cosmology(self.h, self.omega_m, self.omega_b, self.omega_lambda, self.sigma8, self.n_s)

    def growth(self, a):
        """(D1(a), f1(a)) of the reference's growth table (hy_growth)."""
        from ._lib import check, lib
        av = np.array([a], np.float64)
        d1 = np.zeros(1)
        f1 = np.zeros(1)
        pd = C.POINTER(C.c_double)
        check(lib().hy_growth(C.byref(self.c_struct()), 1, av.ctypes.data_as(pd), d1.ctypes.data_as(pd),
                              f1.ctypes.data_as(pd), None, None))
        return float(d1[0]), float(f1[0])

    def hubble(self, a: float) -> float:
        from ._lib import lib
        return float(lib().hy_hubble(a, self.omega_m, self.omega_lambda))

    def power(self, k, a_start: float):
        """P(k) at a_start: the table rescaled by the growth and spline-interpolated (hy_linear_power)."""
        from ._lib import check, lib
        kk = np.ascontiguousarray(k, np.float64).reshape(-1)
        out = np.zeros_like(kk)
        pd = C.POINTER(C.c_double)
        check(lib().hy_linear_power(C.byref(self.c_struct()), self.table().encode(), 0.0, start_redshift(a_start),
                                    kk.size, kk.ctypes.data_as(pd), out.ctypes.data_as(pd)))
        return out.reshape(np.shape(k))


def start_redshift(a_start: float) -> float:
    """z with 1/(1+z) = a_start (the reference derives a_start from z_start)."""
    return 1.0 / a_start - 1.0


@lru_cache(maxsize=8)
def _bbks_amplitude(n_s: float, gamma: float, sigma8: float) -> float:
    shape = Cosmology(n_s=n_s, gamma=gamma, sigma8=sigma8)
    R = 8.0

    def integrand(lnk):
        k = math.exp(lnk)
        x = k * R
        w = 3.0 * (math.sin(x) - x * math.cos(x)) / x ** 3 if x > 1e-4 else 1.0 - x * x / 10.0
        return k ** 3 * float(shape._unnormalised(np.array([k]))[0]) * w * w / (2 * math.pi ** 2)

    s2, _ = integrate.quad(integrand, math.log(1e-5), math.log(1e3), limit=400)
    return sigma8 ** 2 / s2


@lru_cache(maxsize=8)
def _bbks_table(c: Cosmology, k_min: float = 1e-4, k_max: float = 1e2, per_decade: int = 100) -> str:
    decades = math.log10(k_max / k_min)
    n = int(round(decades * per_decade)) + 1
    k = k_min * 10.0 ** (np.arange(n) * (decades / (n - 1)))
    k[-1] = k_max
    p = c.bbks_power(k)
    head = [
        "k_unit=h/Mpc", "power_unit=(Mpc/h)^3", "redshift=0", "species=cold_plus_baryon",
        "gauge=synchronous", "fidelity=precision_boltzmann", "normalization=sigma8_z0",
        f"normalization_value={c.sigma8!r}", "nonlinear=false", "massive_neutrino_sum_eV=0",
        f"h={c.h!r}", f"omega_m={c.omega_m!r}", f"omega_b={c.omega_b!r}",
        f"omega_lambda={c.omega_lambda!r}", f"n_s={c.n_s!r}",
    ]
    lines = [f"# {h}" for h in head]
    lines.append(f"# BBKS transfer, Gamma={c.gamma!r}, A={c.amplitude!r}")
    lines += [f"{kv:.17e} {pv:.17e}" for kv, pv in zip(k, p)]
    return "\n".join(lines) + "\n"


def white_noise(n_side: int, seed: int, device=None):
    """Seeded unit Gaussian white noise on the n_side^3 IC mesh (binary32)."""
    import torch  # data generation only
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    return torch.randn((n_side, n_side, n_side), generator=gen, device=device, dtype=torch.float32)


def variance_by_m2(n_side: int, box: float, a_start: float, cosmo: Cosmology | None = None,
                   max_mode: int | None = None):
    """Per-mode variance N^6/L^3 P(|m| 2 pi/L) at a_start of the r2c modes of an
    n_side^3 IC mesh, indexed by m^2 = mx^2 + my^2 + mz^2 (float64, length
    3 (n/2)^2 + 1; zero at m^2 = 0 and beyond 3 K^2), from hy_ic_mode_variance:
    the reference's table + spline + growth rescaling and its scaled product
    (src/ic/random_field.cpp:108-235)."""
    from ._lib import check, lib
    cosmo = cosmo or Cosmology()
    K = _band_limit(n_side, max_mode)
    half = n_side // 2
    out = np.zeros(3 * half * half + 1, np.float64)
    check(lib().hy_ic_mode_variance(C.byref(cosmo.c_struct()), cosmo.table().encode(), 0.0, start_redshift(a_start),
                                    n_side, float(box), K, out.ctypes.data_as(C.POINTER(C.c_double))))
    return out


def _band_limit(n_side: int, max_mode: int | None) -> int:
    """The particle-lattice band limit (src/ic/particle_lattice_bandlimit.cpp:39-80):
    |m_axis| <= K, by default and at most (n - 1) // 2 for an n^3 lattice."""
    K = (n_side - 1) // 2 if max_mode is None else int(max_mode)
    if K > (n_side - 1) // 2 or K < 1:
        raise ValueError("IC Fourier support exceeds the safe particle-lattice support")
    return K


def philox_modes(n_side: int, box: float, a_start: float, seed: int, cosmo: Cosmology | None = None,
                 max_mode: int | None = None, device=0):
    """Mode-keyed Gaussian initial density modes delta_k (complex128 r2c half
    spectrum [n][n][n/2+1] on the GPU) drawn by libhyowon_pm's
    hy_ic_gaussian_modes: the reference's Philox4x32-10 mode-counter
    convention, so a mode's draw depends only on (seed, integer mode)."""
    import torch
    from ._lib import check, lib
    K = _band_limit(n_side, max_mode)
    var = torch.as_tensor(variance_by_m2(n_side, box, a_start, cosmo, K), device=f"cuda:{device}")
    out = torch.empty((n_side, n_side, n_side // 2 + 1), dtype=torch.complex128, device=f"cuda:{device}")
    torch.cuda.synchronize(device)
    rc = lib().hy_ic_gaussian_modes(n_side, int(seed), K, var.data_ptr(), out.data_ptr(), 0, device, None)
    if rc != 0:
        raise RuntimeError(lib().hy_ic_last_error().decode())
    return out


def momentum_factor(a_start: float, cosmo: Cosmology | None = None) -> float:
    """a^2 H(a) f1(a), the first-order momentum factor (src/ic/lpt_displacement.cpp:202-221)."""
    cosmo = cosmo or Cosmology()
    a = 1.0 / (1.0 + start_redshift(a_start))
    _, f1 = cosmo.growth(a)
    return a * a * cosmo.hubble(a) * f1


def lpt_from_noise(noise, box: float, a_start: float, cosmo: Cosmology | None = None):
    """1LPT lattice ICs from a white-noise field (torch ops only, so the map
    noise -> (positions, momenta) is differentiable with torch.autograd; it is
    linear in the noise apart from the periodic wrap).  Returns (pos, mom) as
    lists of three float64 tensors of n^3 particles."""
    import torch

    cosmo = cosmo or Cosmology()
    device = noise.device
    n = noise.shape[0]
    kf = 2.0 * math.pi / box
    m = torch.fft.fftfreq(n, d=1.0 / n, device=device, dtype=torch.float64)
    mz = torch.fft.rfftfreq(n, d=1.0 / n, device=device, dtype=torch.float64)
    kx = (m * kf).view(n, 1, 1)
    ky = (m * kf).view(1, n, 1)
    kz = (mz * kf).view(1, 1, n // 2 + 1)
    k2 = kx * kx + ky * ky + kz * kz
    kmag = torch.sqrt(k2).cpu().numpy()
    pk = cosmo.power(kmag, a_start)  # already at a_start (growth-rescaled table)
    vcell = (box / n) ** 3
    amp = torch.as_tensor(np.sqrt(pk / vcell), device=device)
    amp[0, 0, 0] = 0.0
    delta_k = torch.fft.rfftn(noise).to(torch.complex128) * amp
    return lpt_from_delta_k(delta_k, box, a_start, cosmo)


def lpt_from_delta_k(delta_k, box: float, a_start: float, cosmo: Cosmology | None = None):
    """1LPT lattice ICs from the initial density modes delta_k (r2c half
    spectrum of the n^3 lattice mesh, at a_start): psi = IFFT(i k / k^2
    delta_k), x = q + psi, p = a^2 H f psi.  Torch ops (differentiable)."""
    import torch

    cosmo = cosmo or Cosmology()
    device = delta_k.device
    n = delta_k.shape[0]
    vel_fac = momentum_factor(a_start, cosmo)
    kf = 2.0 * math.pi / box
    m = torch.fft.fftfreq(n, d=1.0 / n, device=device, dtype=torch.float64)
    mz = torch.fft.rfftfreq(n, d=1.0 / n, device=device, dtype=torch.float64)
    kx = (m * kf).view(n, 1, 1)
    ky = (m * kf).view(1, n, 1)
    kz = (mz * kf).view(1, 1, n // 2 + 1)
    k2 = kx * kx + ky * ky + kz * kz
    k2 = k2.clone()
    k2[0, 0, 0] = 1.0
    nyq = n // 2
    q = torch.arange(n, device=device, dtype=torch.float64) * (box / n)
    qs = (q.view(n, 1, 1), q.view(1, n, 1), q.view(1, 1, n))
    pos, mom = [], []
    for axis, kd in enumerate((kx, ky, kz)):
        kd = kd.clone()
        if n % 2 == 0:  # odd derivative vanishes on the Nyquist plane
            if axis < 2:
                kd[(nyq,) if axis == 0 else (slice(None), nyq)] = 0.0
            else:
                kd[..., nyq] = 0.0
        psi_k = 1j * kd / k2 * delta_k
        psi = torch.fft.irfftn(psi_k, s=(n, n, n)).to(torch.float64)
        del psi_k
        qa = qs[axis].expand(n, n, n)
        x = torch.remainder(qa + psi, box)
        x = torch.where(x >= box, torch.zeros_like(x), x)
        pos.append(x.reshape(-1).contiguous())
        mom.append((psi * vel_fac).reshape(-1).contiguous())
        del psi
    return pos, mom


def lpt_from_delta_k_hip(delta_k, box: float, a_start: float, cosmo: Cosmology | None = None,
                         n_side: int | None = None):
    """1LPT lattice ICs from delta_k on the GPU kernels (hy_ic_lpt): the
    reference's first-order displacement (src/ic/lpt_displacement.cpp) on
    libhyowon_pm's own FFT passes.  delta_k: complex128 CUDA tensor
    [N][N][N/2+1]; n_side (default N) must divide N.  Returns (pos, mom) as
    lists of three float64 tensors of n_side^3 particles."""
    import torch
    from ._lib import check, lib
    from .pm import PMSolver

    cosmo = cosmo or Cosmology()
    N = delta_k.shape[0]
    n = N if n_side is None else int(n_side)
    p_factor = momentum_factor(a_start, cosmo)
    dev = delta_k.device
    dk = delta_k.to(torch.complex128).contiguous()
    out = [torch.empty(n ** 3, dtype=torch.float64, device=dev) for _ in range(6)]
    s = PMSolver(box, N, precision="fp64", max_particles=0, device=dev.index or 0)
    try:
        torch.cuda.synchronize(dev)
        check(lib().hy_ic_lpt(s.handle, n, dk.data_ptr(), p_factor, *(o.data_ptr() for o in out)))
    finally:
        s.close()
    return out[:3], out[3:]


def lpt_ics(n_side: int, box: float, a_start: float, seed: int = 1,
            cosmo: Cosmology | None = None, device=None, rng: str = "torch",
            max_mode: int | None = None):
    """Return (pos, mom) as lists of three tensors (float64) of n_side^3 particles.

    rng="torch": seeded torch white noise (fast synthetic ICs);
    rng="philox": the reference's mode-keyed Philox Gaussian field, drawn on
    the GPU by hy_ic_gaussian_modes, with the Cartesian band limit max_mode
    (default (n_side - 1) // 2, the reference's), displaced by hy_ic_lpt."""
    import torch  # data generation only

    device = device or ("cuda" if torch.cuda.is_available() else "cpu")
    if rng == "philox":
        dev = torch.device(device)
        dk = philox_modes(n_side, box, a_start, seed, cosmo, max_mode=max_mode, device=dev.index or 0)
        return lpt_from_delta_k_hip(dk, box, a_start, cosmo)
    noise = white_noise(n_side, seed, device)
    return lpt_from_noise(noise, box, a_start, cosmo)


def particle_mass(n_side: int, box: float, omega_m: float = 0.3) -> float:
    """m_p = rho_crit0 * Omega_m * (L/N)^3 (src/config/simulation_parameters.cpp:403-416)."""
    from . import rho_crit0
    return rho_crit0() * omega_m * (box / n_side) ** 3
