"""Snapshot formats (Source/file_operations.cu): exact text round trips, the
reference's line format, and reading a coarser Fourier file into a finer grid
(index embedding and the reference's integer scale factor, plus the exact
scale with fix_scale, pinned by physical-space agreement)."""
import numpy as np
import pytest

from ns3dperiodic_amd import fields, snapshots


def test_physical_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    u = [rng.standard_normal((4, 6, 8)) * 10.0 ** rng.integers(-5, 5) for _ in range(3)]
    p = tmp_path / "control_file.dat"
    snapshots.write_control_file(p, *u)
    lines = p.read_text().splitlines()
    assert len(lines) == 4 * 6 * 8
    # "%.16le %.16le %.16le" per point, IN(j,k,l) order
    assert lines[1] == "%.16e %.16e %.16e" % (u[0][0, 0, 1], u[1][0, 0, 1], u[2][0, 0, 1])
    back = snapshots.read_control_file(p, 4, 6, 8)
    for a, b in zip(u, back):
        assert np.array_equal(a, b)
    with pytest.raises(ValueError):
        snapshots.read_control_file(p, 4, 6, 4)


def test_fourier_round_trip_and_max_real(tmp_path):
    u = fields.random_solenoidal(8, 8, 12, seed=2)
    p = tmp_path / "control_file_Fourier.dat"
    mre = snapshots.write_control_file_fourier(p, *u)
    assert mre == max(np.abs(c.real).max() for c in u)
    line = p.read_text().splitlines()[3]
    vals = [u[0].ravel()[3], u[1].ravel()[3], u[2].ravel()[3]]
    assert line == " ".join("%.16e %.16e" % (v.real, v.imag) for v in vals)
    back = snapshots.read_control_fourier(p, 8, 8, 12)
    for a, b in zip(u, back):
        assert np.array_equal(a, b)


def test_fourier_size_guess():
    assert snapshots.fourier_file_size(16 * 16 * 9, 16, 16, 9) is None
    assert snapshots.fourier_file_size(8 * 8 * 5, 16, 16, 9) == (8, 8, 5)
    assert snapshots.fourier_file_size(32 * 16 * 9, 64, 32, 17) == (32, 16, 9)


@pytest.mark.parametrize("fix", [False, True])
def test_fourier_coarse_to_fine(tmp_path, fix):
    N0, N1 = 8, 16
    u0 = fields.random_solenoidal(N0, seed=5)
    p = tmp_path / "control_file_Fourier.dat"
    snapshots.write_control_file_fourier(p, *u0)
    u1 = snapshots.read_control_fourier(p, N1, N1, N1, fix_scale=fix)
    scale = (N1 ** 3 / N0 ** 3) if fix else float((N1 * N1 * (2 * (N1 // 2 + 1) - 1)) // (N0 * N0 * (2 * (N0 // 2 + 1) - 1)))
    assert fix or scale == 7.0          # the reference's integer ratio 4352 // 576 (exact: 8)
    m0 = np.arange(N0)
    jj = np.where(m0 >= N0 // 2, m0 + N1 - N0, m0)
    for a, b in zip(u0, u1):
        assert np.array_equal(b[np.ix_(jj, jj, np.arange(N0 // 2 + 1))], scale * a)
        mask = np.ones(b.shape, bool)
        mask[np.ix_(jj, jj, np.arange(N0 // 2 + 1))] = False
        assert not np.any(b[mask])
    if fix:
        # same physical field: the fine grid's values at the coarse points equal the coarse field
        # (the coarse field has zero Nyquist planes, so the embedding is exact)
        for a, b in zip(u0, u1):
            x0 = np.fft.irfftn(a, s=(N0,) * 3, axes=(0, 1, 2))
            x1 = np.fft.irfftn(b, s=(N1,) * 3, axes=(0, 1, 2))
            assert np.allclose(x1[::2, ::2, ::2], x0, rtol=0, atol=1e-13 * np.abs(x0).max())
