"""Pin the oracle's FFT layer and spectral tables (CPU only).

FFT: brute-force DFT sums, numpy.fft, single-mode closed forms, round trip,
Parseval, and the C2R convention for non-Hermitian input.  Tables: closed
forms, projector identities, and the 2/3 ellipsoid decided in exact integer
arithmetic for every supported grid size.
"""
import itertools
import math

import numpy as np
import pytest

from tests.conftest import rel_l2

SUPPORTED = [8, 16, 32, 48, 64, 80, 96, 128, 160, 192, 256, 320, 384, 512]


def brute_r2c(x):
    Nx, Ny, Nz = x.shape
    Mz = Nz // 2 + 1
    j = np.arange(Nx)[:, None, None]
    k = np.arange(Ny)[None, :, None]
    l = np.arange(Nz)[None, None, :]
    out = np.zeros((Nx, Ny, Mz), np.complex128)
    for a, b, c in itertools.product(range(Nx), range(Ny), range(Mz)):
        ph = np.exp(-2j * math.pi * (a * j / Nx + b * k / Ny + c * l / Nz))
        out[a, b, c] = np.sum(x * ph)
    return out


@pytest.mark.parametrize("shape", [(4, 4, 4), (6, 4, 8), (8, 6, 4)])
def test_r2c_matches_brute_force(oracle_mod, shape):
    o = oracle_mod.Oracle(*shape)
    x = np.random.default_rng(0).standard_normal(shape)
    assert rel_l2(o.fft_r2c(x), brute_r2c(x)) < 1e-14


@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 12, 10), (32, 48, 24), (48, 16, 96), (6, 10, 14), (80, 16, 160)])
def test_fft_vs_numpy_and_round_trip(oracle_mod, shape):
    o = oracle_mod.Oracle(*shape)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(shape)
    X = o.fft_r2c(x)
    assert rel_l2(X, np.fft.rfftn(x)) < 1e-14
    assert rel_l2(o.fft_c2r(X), x) < 1e-14
    # Parseval for the unnormalised transform
    w = np.ones(X.shape[2])
    w[1:(shape[2] + 1) // 2] = 2.0
    e_spec = np.sum(w * np.abs(X) ** 2) / np.prod(shape)
    assert abs(e_spec - np.sum(x ** 2)) < 1e-12 * np.sum(x ** 2)
    # unscaled C2R is N times the scaled one (iFFTN_Device without scale_double)
    assert rel_l2(o.fft_c2r(X, scale=False), x * np.prod(shape)) < 1e-14


@pytest.mark.parametrize("shape", [(8, 8, 8), (12, 16, 10)])
def test_c2r_convention_non_hermitian(oracle_mod, shape):
    """x,y inverse C2C first, then z C2R using Re of the l=0 and l=Nz/2 planes (numpy irfftn)."""
    o = oracle_mod.Oracle(*shape)
    rng = np.random.default_rng(2)
    Z = rng.standard_normal(o.cshape) + 1j * rng.standard_normal(o.cshape)
    ref = np.fft.irfftn(Z, s=shape, axes=(0, 1, 2))
    assert rel_l2(o.fft_c2r(Z), ref) < 1e-14
    # changing Im of the l=0 plane of a mode that is its own partner (x,y Nyquist) changes nothing
    Z2 = Z.copy()
    Z2[shape[0] // 2, shape[1] // 2, 0] += 5j
    assert rel_l2(o.fft_c2r(Z2), o.fft_c2r(Z)) < 1e-15


def test_single_mode_closed_form(oracle_mod):
    N = 16
    o = oracle_mod.Oracle(N)
    x = np.arange(N) * 2 * math.pi / N
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    f = np.cos(3 * X + 2 * Y + 5 * Z)
    F = o.fft_r2c(f)
    expect = np.zeros_like(F)
    expect[3, 2, 5] = N ** 3 / 2
    expect[N - 3, N - 2, 0] = 0  # only l >= 0 is stored; (-3,-2,-5) is not in the half spectrum
    assert np.abs(F - expect).max() < 1e-9


@pytest.mark.parametrize("n", [2, 3, 4, 5, 7, 8, 9, 11, 12, 16, 48, 96])
def test_c2c_1d(oracle_mod, n):
    x = np.random.default_rng(n).standard_normal(n) + 1j * np.random.default_rng(n + 1).standard_normal(n)
    o = oracle_mod.Oracle
    assert rel_l2(o.fft_c2c_1d(x, -1), np.fft.fft(x)) < 1e-14
    assert rel_l2(o.fft_c2c_1d(x, +1), np.fft.ifft(x) * n) < 1e-14


@pytest.mark.parametrize("shape,L", [((16, 16, 16), (2 * math.pi,) * 3), ((8, 12, 16), (2 * math.pi / 0.7, 2 * math.pi, 3.0))])
def test_tables_closed_forms(oracle_mod, shape, L):
    o = oracle_mod.Oracle(*shape, Lx=L[0], Ly=L[1], Lz=L[2])
    t = o.tables()
    Nx, Ny, Nz = shape
    m = np.fft.fftfreq(Nx, 1.0 / Nx)
    n = np.fft.fftfreq(Ny, 1.0 / Ny)
    q = np.arange(Nz // 2 + 1)
    # build_Nabla_Wavenumbers: m wraps at N/2 (Nyquist index carries -N/2), z does not wrap
    np.testing.assert_allclose(t["kx"], 2 * math.pi / L[0] * m, rtol=0, atol=1e-13)
    np.testing.assert_allclose(t["ky"], 2 * math.pi / L[1] * n, rtol=0, atol=1e-13)
    np.testing.assert_allclose(t["kz"], 2 * math.pi / L[2] * q, rtol=0, atol=1e-13)
    KX, KY, KZ = np.meshgrid(t["kx"], t["ky"], t["kz"], indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    np.testing.assert_allclose(t["din_diffusion"], k2, rtol=1e-14)
    assert t["din_poisson"][0, 0, 0] == 1.0
    # AM is the projector k k^T/|k|^2: idempotent and AM k = k (except k = 0 where AM = 0)
    A = np.empty(KX.shape + (3, 3))
    A[..., 0, 0], A[..., 1, 1], A[..., 2, 2] = t["AM11"], t["AM22"], t["AM33"]
    A[..., 0, 1] = A[..., 1, 0] = t["AM12"]
    A[..., 0, 2] = A[..., 2, 0] = t["AM13"]
    A[..., 1, 2] = A[..., 2, 1] = t["AM23"]
    assert np.abs(A @ A - A).max() < 1e-14
    kv = np.stack([KX, KY, KZ], -1)
    Ak = np.einsum("...ab,...b->...a", A, kv)
    nz = k2 > 0
    assert np.abs(Ak[nz] - kv[nz]).max() < 1e-12
    assert np.all(A[0, 0, 0] == 0)


@pytest.mark.parametrize("Nx,Ny,Nz", [(n, n, n) for n in SUPPORTED if n <= 192] +
                         [(64, 32, 32), (128, 64, 64), (256, 128, 128), (256, 256, 48), (512, 32, 64), (32, 512, 16), (16, 16, 512)])
def test_mask_is_the_ellipsoid_with_Mz(oracle_mod, Nx, Ny, Nz):
    """Q2: keep iff (m/Nx)^2+(n/Ny)^2+(q/Mz)^2 < 4/9 with Mz = Nz/2+1, decided exactly; mask(0)=0."""
    o = oracle_mod.Oracle(Nx, Ny, Nz)
    mask = o.tables()["mask"]
    Mz = Nz // 2 + 1
    m = np.fft.fftfreq(Nx, 1.0 / Nx).astype(np.int64)
    n = np.fft.fftfreq(Ny, 1.0 / Ny).astype(np.int64)
    q = np.arange(Mz, dtype=np.int64)
    M, Nn, Q = np.meshgrid(m, n, q, indexing="ij")
    lhs = 9 * (M * M * Ny * Ny * Mz * Mz + Nn * Nn * Nx * Nx * Mz * Mz + Q * Q * Nx * Nx * Ny * Ny)
    rhs = 4 * Nx * Nx * Ny * Ny * Mz * Mz
    exact = (lhs < rhs).astype(float)
    exact[0, 0, 0] = 0.0
    tie = lhs == rhs
    np.testing.assert_array_equal(mask[~tie], exact[~tie])
    # exact ties (e.g. q/Mz = 22/33 at 64^3) are decided by the reference's floating-point
    # expression sqrt(kx^2/kxMax^2 + ...) < 2/3 (math_support.cu:498-504), evaluated here in numpy
    if tie.any():
        f = lambda k, N2: (k * 2 * math.pi / (2 * math.pi)) ** 2 / (2 * N2 * math.pi / (2 * math.pi)) ** 2
        s2 = np.sqrt(f(M[tie], Nx) + f(Nn[tie], Ny) + f(Q[tie], Mz))
        np.testing.assert_array_equal(mask[tie], (s2 < 2.0 / 3.0).astype(float))
    # x Nyquist (m = -Nx/2) survives the mask on the k=l=0 line (Q5)
    assert mask[Nx // 2, 0, 0] == 1.0
