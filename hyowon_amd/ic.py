"""Synthetic 1LPT initial conditions for benchmarks and tests.

Builds the workload the PM step is measured on: a corner lattice of particles
(index (i*N+j)*N+k, q = (i,j,k)*L/N; src/ic/initial_conditions.cpp:296-312)
displaced by the first-order LPT field of a Gaussian density with the BBKS
spectrum P(k) = A k^n_s T(k)^2, T from Bardeen et al. (1986) with q = k/Gamma,
normalised to sigma_8 at z=0 and scaled by the linear growth D1(a).
Momenta p = a^2 H(a) f(a) psi (src/ic/lpt_displacement.cpp:301-380).

The white noise comes from torch's seeded Gaussian generator, not from the
reference's Philox mode-keyed generator, so the realisations differ from
HYOWON's for the same seed; the statistics and the path being timed do not.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
from scipy import integrate


@dataclass
class Cosmology:
    omega_m: float = 0.3
    omega_lambda: float = 0.7
    h: float = 0.7
    sigma8: float = 0.8
    n_s: float = 0.96
    gamma: float = 0.21

    def E(self, a):
        return np.sqrt(self.omega_m / a ** 3 + self.omega_lambda)

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
        R = 8.0

        def integrand(lnk):
            k = math.exp(lnk)
            x = k * R
            w = 3.0 * (math.sin(x) - x * math.cos(x)) / x ** 3 if x > 1e-4 else 1.0 - x * x / 10.0
            return k ** 3 * float(self._unnormalised(np.array([k]))[0]) * w * w / (2 * math.pi ** 2)

        s2, _ = integrate.quad(integrand, math.log(1e-5), math.log(1e3), limit=400)
        return self.sigma8 ** 2 / s2

    def power(self, k):
        return self.amplitude * self._unnormalised(k)

    def growth(self, a):
        """Linear growth D1(a), D1(1) = 1, and f = dlnD/dlna (flat matter + Lambda)."""
        def unnorm(a_):
            val, _ = integrate.quad(lambda x: 1.0 / (x * self.E(x)) ** 3, 0.0, a_, limit=200)
            return 2.5 * self.omega_m * self.E(a_) * val

        d = unnorm(a) / unnorm(1.0)
        h = 1e-4
        f = (math.log(unnorm(a * (1 + h))) - math.log(unnorm(a * (1 - h)))) / (math.log1p(h) - math.log1p(-h))
        return d, f


def white_noise(n_side: int, seed: int, device=None):
    """Seeded unit Gaussian white noise on the n_side^3 IC mesh (binary32)."""
    import torch  # data generation only
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    return torch.randn((n_side, n_side, n_side), generator=gen, device=device, dtype=torch.float32)


def variance_by_m2(n_side: int, box: float, a_start: float, cosmo: Cosmology | None = None,
                   max_mode: int | None = None):
    """Per-mode variance N^6/L^3 P(k) D1(a)^2 of the r2c modes of an n_side^3
    IC mesh, indexed by m^2 = mx^2 + my^2 + mz^2 (float64, length 3 K^2 + 1
    with K = n_side // 2); entry 0 is unused."""
    cosmo = cosmo or Cosmology()
    D, _ = cosmo.growth(a_start)
    K = n_side // 2
    m2 = np.arange(3 * K * K + 1, dtype=np.float64)
    k = np.sqrt(m2) * (2.0 * math.pi / box)
    pk = np.zeros_like(k)
    pk[1:] = cosmo.power(k[1:])
    return float(n_side) ** 6 / box ** 3 * pk * D * D


def philox_modes(n_side: int, box: float, a_start: float, seed: int, cosmo: Cosmology | None = None,
                 max_mode: int | None = None, device=0):
    """Mode-keyed Gaussian initial density modes delta_k (complex128 r2c half
    spectrum [n][n][n/2+1] on the GPU) drawn by libhyowon_pm's
    hy_ic_gaussian_modes: the reference's Philox4x32-10 mode-counter
    convention, so a mode's draw depends only on (seed, integer mode)."""
    import torch
    from ._lib import check, lib
    var = torch.as_tensor(variance_by_m2(n_side, box, a_start, cosmo), device=f"cuda:{device}")
    out = torch.empty((n_side, n_side, n_side // 2 + 1), dtype=torch.complex128, device=f"cuda:{device}")
    # the particle-lattice band limit (src/ic/particle_lattice_bandlimit.cpp:39-80):
    # |m_axis| <= K, by default and at most (n - 1) // 2 for an n^3 lattice on this mesh
    K = (n_side - 1) // 2 if max_mode is None else int(max_mode)
    if K > (n_side - 1) // 2:
        raise ValueError("IC Fourier support exceeds the safe particle-lattice support")
    torch.cuda.synchronize(device)
    rc = lib().hy_ic_gaussian_modes(n_side, int(seed), K, var.data_ptr(), out.data_ptr(), 0, device, None)
    if rc != 0:
        raise RuntimeError(lib().hy_ic_last_error().decode())
    return out


def lpt_from_noise(noise, box: float, a_start: float, cosmo: Cosmology | None = None):
    """1LPT lattice ICs from a white-noise field (torch ops only, so the map
    noise -> (positions, momenta) is differentiable with torch.autograd; it is
    linear in the noise apart from the periodic wrap).  Returns (pos, mom) as
    lists of three float64 tensors of n^3 particles."""
    import torch

    cosmo = cosmo or Cosmology()
    device = noise.device
    n = noise.shape[0]
    D, f = cosmo.growth(a_start)
    H = 100.0 * float(cosmo.E(a_start))
    vel_fac = a_start * a_start * H * f

    kf = 2.0 * math.pi / box
    m = torch.fft.fftfreq(n, d=1.0 / n, device=device, dtype=torch.float64)
    mz = torch.fft.rfftfreq(n, d=1.0 / n, device=device, dtype=torch.float64)
    kx = (m * kf).view(n, 1, 1)
    ky = (m * kf).view(1, n, 1)
    kz = (mz * kf).view(1, 1, n // 2 + 1)
    k2 = kx * kx + ky * ky + kz * kz
    kmag = torch.sqrt(k2).cpu().numpy()
    pk = cosmo.power(kmag.ravel()).reshape(kmag.shape)
    vcell = (box / n) ** 3
    amp = torch.as_tensor(np.sqrt(pk / vcell), device=device) * D
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
    _, f = cosmo.growth(a_start)
    H = 100.0 * float(cosmo.E(a_start))
    vel_fac = a_start * a_start * H * f
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
    _, f = cosmo.growth(a_start)
    H = 100.0 * float(cosmo.E(a_start))
    p_factor = a_start * a_start * H * f
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
