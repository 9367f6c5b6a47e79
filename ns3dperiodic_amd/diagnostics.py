"""Host-side diagnostics of the reference driver.

calculate_energy_spectrum (Source/math_support.cu:819-866): the spectrum over
"taxicab" shells n = j + k + l of the first octant j < Nx/2, k < Ny/2,
l < Mz (negative wavenumbers are not visited), E(n) = 1/2 sum |u^|^2 / norm^2
with norm = Nx*Ny*Mz, written as "%i %.16le %.16le" lines of n, E(n) and
the reference slope 0.25 n^(-5/3).
"""
import numpy as np


def energy_spectrum(ux_hat, uy_hat, uz_hat):
    """Returns (n, E(n), 0.25 n^(-5/3)) as arrays of length Nx/2 + Ny/2 + Mz."""
    Nx, Ny, Mz = ux_hat.shape
    size = Nx // 2 + Ny // 2 + Mz
    norm = float(Nx * Ny * Mz)
    j, k, l = np.meshgrid(np.arange(Nx // 2), np.arange(Ny // 2), np.arange(Mz), indexing="ij")
    shell = (j + k + l).ravel()
    e = sum((np.abs(np.asarray(c)[: Nx // 2, : Ny // 2, :]) ** 2).ravel() / norm / norm for c in (ux_hat, uy_hat, uz_hat))
    E = 0.5 * np.bincount(shell, weights=e, minlength=size)[:size]
    n = np.arange(size)
    with np.errstate(divide="ignore"):
        slope = 0.25 * np.power(n.astype(float), -5.0 / 3.0)
    # shells never visited keep their index 0 in the reference's first column
    visited = np.zeros(size, bool)
    visited[np.unique(shell)] = True
    label = np.where(visited, n, 0)
    slope = np.where(visited, slope, 0.0)
    return label, E, slope


def write_energy_spectrum(path, ux_hat, uy_hat, uz_hat):
    n, E, s = energy_spectrum(ux_hat, uy_hat, uz_hat)
    with open(path, "w") as fh:
        for a, b, c in zip(n, E, s):
            fh.write("%i %.16e %.16e\n" % (a, b, c))
    return n, E, s
