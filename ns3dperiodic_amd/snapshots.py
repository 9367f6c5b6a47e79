"""Snapshot files of the reference (checkpoint / resume, interchange).

Two text formats, both in the reference's index order IN(j,k,l) (x slowest,
z fastest; Source/Macros.h:11):

* ``control_file.dat`` (Source/file_operations.cu:4-44): physical velocity,
  one line "ux uy uz" per grid point, each "%.16le".
* ``control_file_Fourier.dat`` (Source/file_operations.cu:107-222): spectral
  velocity [Nx][Ny][Mz], one line "Re(ux) Im(ux) Re(uy) Im(uy) Re(uz) Im(uz)"
  per mode.

Reading a Fourier file written on a coarser grid embeds it in the finer
spectrum (Source/file_operations.cu:47-176): the file's size is guessed from
its line count with the aspect ratios of the target grid, modes with
j >= Nx_file/2 (k >= Ny_file/2) move to j + Nx - Nx_file (k + Ny - Ny_file),
l is kept, and every value is multiplied by
Nx*Ny*(2*Mz-1) / (Nx_file*Ny_file*(2*Mz_file-1)) evaluated in integer
arithmetic (:153; e.g. 7 instead of 7.88 for 32^3 -> 64^3).  ``fix_scale``
uses the exact ratio of the unnormalised spectra, NxNyNz / (NxNyNz)_file.

Host-side file I/O: arrays are numpy (copy device tensors with .cpu()).
"""
import numpy as np


def _fmt_lines(cols):
    return "\n".join(" ".join("%.16e" % v for v in row) for row in zip(*cols)) + "\n"


def write_control_file(path, ux, uy, uz):
    """write_control_file (Source/file_operations.cu:27-44): physical [Nx][Ny][Nz] fields."""
    cols = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (ux, uy, uz)]
    with open(path, "w") as fh:
        np.savetxt(fh, np.stack(cols, axis=1), fmt="%.16e")


def read_control_file(path, Nx, Ny, Nz):
    """read_control_file (Source/file_operations.cu:4-24)."""
    d = np.loadtxt(path, dtype=np.float64, ndmin=2)
    if d.shape != (Nx * Ny * Nz, 3):
        raise ValueError(f"{path}: {d.shape[0]} points, expected {Nx * Ny * Nz}")
    return [np.ascontiguousarray(d[:, c].reshape(Nx, Ny, Nz)) for c in range(3)]


def write_control_file_fourier(path, ux_hat, uy_hat, uz_hat):
    """write_control_file_fourier (Source/file_operations.cu:180-222): spectral [Nx][Ny][Mz]
    fields.  Returns the largest |Re| over the three components (the value the
    reference prints)."""
    u = [np.ascontiguousarray(a, dtype=np.complex128).ravel() for a in (ux_hat, uy_hat, uz_hat)]
    cols = np.stack([u[0].real, u[0].imag, u[1].real, u[1].imag, u[2].real, u[2].imag], axis=1)
    with open(path, "w") as fh:
        np.savetxt(fh, cols, fmt="%.16e")
    return float(max(np.abs(c.real).max() for c in u)) if u[0].size else 0.0


def fourier_file_size(count, Nx, Ny, Mz):
    """get_control_fourier_size (Source/file_operations.cu:47-103) on a line
    count: None if the file matches the target (Nx, Ny, Mz), else the guessed
    (Nx_file, Ny_file, Mz_file).  Guesses keep the target's integer aspect
    ratios Nx/Ny and (2Mz-1)/Ny, with Ny_file = 4, 6, 8, ..."""
    if count == Nx * Ny * Mz:
        return None
    rxy = Nx // Ny
    rzy = (2 * Mz - 1) // Ny
    ny = 2
    while True:
        ny += 2
        nx, nz = rxy * ny, rzy * ny
        mz = nz // 2 + 1
        if nx * ny * mz == count:
            return nx, ny, mz
        if nx * ny * mz > count and ny > 4 * max(Nx, Ny, 2 * Mz):
            raise ValueError(f"no grid with the target's aspect ratios has {count} modes")


def read_control_fourier(path, Nx, Ny, Nz, fix_scale=False):
    """read_control_fourier (Source/file_operations.cu:107-176) for a target
    grid Nx x Ny x Nz (Mz = Nz/2+1).  Returns three complex128 [Nx][Ny][Mz]
    arrays; modes absent from a coarser file are zero."""
    Mz = Nz // 2 + 1
    d = np.loadtxt(path, dtype=np.float64, ndmin=2)
    if d.shape[1] != 6:
        raise ValueError(f"{path}: expected 6 columns")
    size = fourier_file_size(d.shape[0], Nx, Ny, Mz)
    u = [np.zeros((Nx, Ny, Mz), np.complex128) for _ in range(3)]
    if size is None:
        for c in range(3):
            u[c][...] = (d[:, 2 * c] + 1j * d[:, 2 * c + 1]).reshape(Nx, Ny, Mz)
        return u
    nxr, nyr, mzr = size
    if nxr > Nx or nyr > Ny or mzr > Mz:
        raise ValueError(f"{path}: file grid {nxr}x{nyr}x(Mz={mzr}) is larger than the target")
    if fix_scale:
        scale = (Nx * Ny * Nz) / (nxr * nyr * 2 * (mzr - 1))
    else:
        scale = float((Nx * Ny * (2 * Mz - 1)) // (nxr * nyr * (2 * mzr - 1)))
    j = np.arange(nxr)
    k = np.arange(nyr)
    jj = np.where(j >= nxr // 2, j + (Nx - nxr), j)
    kk = np.where(k >= nyr // 2, k + (Ny - nyr), k)
    for c in range(3):
        src = (d[:, 2 * c] + 1j * d[:, 2 * c + 1]).reshape(nxr, nyr, mzr)
        u[c][np.ix_(jj, kk, np.arange(mzr))] = scale * src
    return u
