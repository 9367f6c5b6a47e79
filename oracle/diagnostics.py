"""numpy restatement of the reference driver's diagnostics (test infrastructure
only: the checker for the HIP kernels ns3d_curl / ns3d_helmholtz_filter /
ns3d_cutoff_filter; nothing in the product path imports it).

Pinned in tests/test_oracle_diagnostics.py by closed forms: the central
difference of sin(a x) is sin(a h)/h cos(a x) exactly, and each filter acts
on a single Fourier mode by the factor of its formula.
"""
import numpy as np


def curl(ux, uy, uz, dx, dy, dz, fix=False):
    """get_curl_device (Source/math_support.cu:680-718): central differences,
    periodic neighbours (I3, Source/Macros.h:19).  The reference's y-component
    differences uz(j+1) against ux(j-1); fix=True uses uz(j-1)."""
    def d(a, axis, h):
        return 0.5 * (np.roll(a, -1, axis) - np.roll(a, 1, axis)) / h
    rx = d(uz, 1, dy) - d(uy, 2, dz)
    xm = uz if fix else ux
    ry = d(ux, 2, dz) - 0.5 * (np.roll(uz, -1, 0) - np.roll(xm, 1, 0)) / dx
    rz = d(uy, 0, dx) - d(ux, 1, dy)
    return rx, ry, rz


def _mnq(Nx, Ny, Mz):
    j, k, l = np.arange(Nx), np.arange(Ny), np.arange(Mz)
    m = np.where(j >= Nx // 2, j - Nx, j)
    n = np.where(k >= Ny // 2, k - Ny, k)
    return np.meshgrid(m, n, l, indexing="ij")


def helmholtz_filter(u_hat, eps, Lx, Ly, Lz):
    """Helmholz_Fourier_Filter_Device (Source/math_support.cu:860-905)."""
    Nx, Ny, Mz = u_hat[0].shape
    M, N, Q = _mnq(Nx, Ny, Mz)
    a, b, c = 2 * np.pi / Lx * M, 2 * np.pi / Ly * N, 2 * np.pi / Lz * Q
    f = 1.0 / (1.0 + eps * eps * (a * a + b * b + c * c))
    return [f * x for x in u_hat]


def cutoff_filter(u_hat, radius):
    """CutOff_Fourier_Filter_Device (Source/math_support.cu:931-970), called
    with Nz := Mz (:978)."""
    Nx, Ny, Mz = u_hat[0].shape
    M, N, Q = _mnq(Nx, Ny, Mz)
    keep = (M * M / (Nx * Nx) + N * N / (Ny * Ny) + Q * Q / (Mz * Mz)) <= radius * radius
    return [np.where(keep, x, 0.0) for x in u_hat]
