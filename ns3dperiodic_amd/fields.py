"""Synthetic inputs in the reference layout, generated on the host in FP64.

Spectral arrays follow the reference normalisation (unnormalised forward DFT,
Source/math_support.cu:301-312), shape [Nx, Ny, Nz//2+1].  Generation is
seeded (numpy PCG64) so CPU and GPU runs start from identical bits.
"""
import math

import numpy as np


def wavenumbers(Nx, Ny, Nz, Lx=2 * math.pi, Ly=2 * math.pi, Lz=2 * math.pi):
    """k vectors as in build_Nabla_Wavenumbers (Source/math_support.cu:393-414)."""
    m = np.arange(Nx); m = np.where(m >= Nx // 2, m - Nx, m)
    n = np.arange(Ny); n = np.where(n >= Ny // 2, n - Ny, n)
    q = np.arange(Nz // 2 + 1)
    kx = 2.0 * math.pi / Lx * m
    ky = 2.0 * math.pi / Ly * n
    kz = 2.0 * math.pi / Lz * q
    return np.meshgrid(kx, ky, kz, indexing="ij")


def grid(Nx, Ny, Nz, Lx=2 * math.pi, Ly=2 * math.pi, Lz=2 * math.pi):
    x = np.arange(Nx) * (Lx / Nx)
    y = np.arange(Ny) * (Ly / Ny)
    z = np.arange(Nz) * (Lz / Nz)
    return np.meshgrid(x, y, z, indexing="ij")


def _spectrum(kk, kind, k0):
    if kind == "k4exp":     # E(k) ~ k^4 exp(-2 (k/k0)^2)
        return kk ** 4 * np.exp(-2.0 * (kk / k0) ** 2)
    if kind == "k2exp":     # E(k) ~ k^2 exp(-(k/k0)^2)
        return kk ** 2 * np.exp(-((kk / k0) ** 2))
    raise ValueError(kind)


def random_solenoidal(Nx, Ny=None, Nz=None, seed=0, kind="k4exp", k0=4.0, energy=0.5,
                      Lx=2 * math.pi, Ly=2 * math.pi, Lz=2 * math.pi):
    """Seeded divergence-free field: Gaussian complex coefficients with amplitude
    sqrt(E(|k|)/(4 pi |k|^2)), projected solenoidal, made real (Hermitian) by a
    C2R/R2C round trip, k=0 and the Nyquist planes zeroed, scaled so that
    0.5 <|u|^2> = energy.  Returns three complex128 spectral arrays."""
    Ny = Nx if Ny is None else Ny
    Nz = Nx if Nz is None else Nz
    rng = np.random.default_rng(seed)
    KX, KY, KZ = wavenumbers(Nx, Ny, Nz, Lx, Ly, Lz)
    kk = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    kk_safe = np.where(kk == 0, 1.0, kk)
    amp = np.sqrt(_spectrum(kk_safe, kind, k0) / (4 * math.pi * kk_safe ** 2))
    amp[kk == 0] = 0.0
    shape = KX.shape
    u = [amp * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) for _ in range(3)]
    # solenoidal projection (I - k k^T/|k|^2)
    kd = (KX * u[0] + KY * u[1] + KZ * u[2]) / kk_safe ** 2
    u = [u[0] - KX * kd, u[1] - KY * kd, u[2] - KZ * kd]
    # zero the Nyquist planes of every axis (keeps the field exactly solenoidal after the round trip)
    for c in u:
        c[Nx // 2, :, :] = 0
        c[:, Ny // 2, :] = 0
        c[:, :, Nz // 2] = 0
        c[0, 0, 0] = 0
    phys = [np.fft.irfftn(c, s=(Nx, Ny, Nz), axes=(0, 1, 2)) for c in u]
    e = 0.5 * np.mean(phys[0] ** 2 + phys[1] ** 2 + phys[2] ** 2)
    s = math.sqrt(energy / e) if e > 0 else 0.0
    return [np.fft.rfftn(p * s, axes=(0, 1, 2)) for p in phys]


def random_scalar(Nx, Ny=None, Nz=None, seed=0, k0=4.0):
    """Seeded Gaussian random scalar field, E(k) ~ k^2 exp(-(k/k0)^2), unit variance (physical space)."""
    Ny = Nx if Ny is None else Ny
    Nz = Nx if Nz is None else Nz
    rng = np.random.default_rng(seed)
    KX, KY, KZ = wavenumbers(Nx, Ny, Nz)
    kk = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    kk_safe = np.where(kk == 0, 1.0, kk)
    amp = np.sqrt(_spectrum(kk_safe, "k2exp", k0) / (4 * math.pi * kk_safe ** 2))
    amp[kk == 0] = 0.0
    c = amp * (rng.standard_normal(KX.shape) + 1j * rng.standard_normal(KX.shape))
    th = np.fft.irfftn(c, s=(Nx, Ny, Nz), axes=(0, 1, 2))
    th -= th.mean()
    return th / th.std()


def shapiro(Nx, Ny=None, Nz=None, t=0.0, Re=20.0):
    """Shapiro Beltrami solution (Source/NS3D_periodic.cu:810-829, Source/Shapiro_test.cu:26-31)."""
    Ny = Nx if Ny is None else Ny
    Nz = Nx if Nz is None else Nz
    X, Y, Z = grid(Nx, Ny, Nz)
    f = math.sqrt(3.0)
    dec = math.exp(-(f * f) * t / Re)
    ux = -0.5 * (f * np.cos(X) * np.sin(Y) * np.sin(Z) + np.sin(X) * np.cos(Y) * np.cos(Z)) * dec
    uy = 0.5 * (f * np.sin(X) * np.cos(Y) * np.sin(Z) - np.cos(X) * np.sin(Y) * np.cos(Z)) * dec
    uz = np.cos(X) * np.cos(Y) * np.sin(Z) * dec
    return [ux, uy, uz]


def kolmogorov_forcing_sin(Nx, Ny=None, Nz=None, n=4, amp=1.0):
    """f_x = amp*sin(n y): f^_x(0, n, 0) = -i amp N/2, f^_x(0, Ny-n, 0) = +i amp N/2 (N = NxNyNz)."""
    Ny = Nx if Ny is None else Ny
    Nz = Nx if Nz is None else Nz
    Mz = Nz // 2 + 1
    f = [np.zeros((Nx, Ny, Mz), np.complex128) for _ in range(3)]
    N3 = Nx * Ny * Nz
    f[0][0, n, 0] = -0.5j * amp * N3
    f[0][0, Ny - n, 0] = 0.5j * amp * N3
    return f


def reference_kolmogorov(Nx, Ny=None, Nz=None, n=1, Re=10.0):
    """The reference's Fourier initial condition and forcing (Source/NS3D_periodic.cu:725-764,
    Source/math_support.cu:42-88).  Both are called with Nz := Mz (NS3D_periodic.cu:735,1271),
    so the amplitudes carry Nx*Ny*Mz: f_x = (Mz/Nz) cos(n y) sin(n z) and
    u_x = (Re/2) f_x.  For n = 1 this is a steady state of the discrete equations."""
    Ny = Nx if Ny is None else Ny
    Nz = Nx if Nz is None else Nz
    Mz = Nz // 2 + 1
    N3 = Nx * Ny * Mz
    u = [np.zeros((Nx, Ny, Mz), np.complex128) for _ in range(3)]
    f = [np.zeros((Nx, Ny, Mz), np.complex128) for _ in range(3)]
    u[0][0, n, n] = -1j * N3 * 0.25 * 0.5 * Re
    u[0][0, Ny - n, n] = -1j * N3 * 0.25 * 0.5 * Re
    f[0][0, n, n] = -1j * N3 * 0.25
    f[0][0, Ny - n, n] = -1j * N3 * 0.25
    return u, f
