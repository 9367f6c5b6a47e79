"""Error paths of the C ABI that return before any device work (run on the CPU):
argument and size validation of ns3d_create, null contexts and pointers, the
status strings, and the supported-size query."""
import ctypes

import pytest

from ns3dperiodic_amd import _lib


@pytest.fixture(scope="module")
def lib():
    try:
        return _lib.load()
    except _lib.NS3DError:
        pytest.skip("libns3d_hip.so is not built")


def _cfg(**kw):
    d = dict(Nx=16, Ny=16, Nz=16, Lx=1.0, Ly=1.0, Lz=1.0, Re=10.0, dt=1e-3, precision=_lib.NS3D_FP64, flags=0,
             stream=None, device=-1)
    d.update(kw)
    return _lib.Config(d["Nx"], d["Ny"], d["Nz"], d["Lx"], d["Ly"], d["Lz"], d["Re"], d["dt"], d["precision"], d["flags"],
                       ctypes.c_void_p(d["stream"]), d["device"])


@pytest.mark.parametrize("kw,code", [
    (dict(Nx=100), -2), (dict(Nz=20), -2), (dict(Ny=1024), -2),
    (dict(Re=0.0), -1), (dict(Lx=-1.0), -1), (dict(precision=7), -1),
])
def test_create_rejects(lib, kw, code):
    h = ctypes.c_void_p()
    assert lib.ns3d_create(ctypes.byref(_cfg(**kw)), ctypes.byref(h)) == code
    assert not h.value


def test_null_arguments(lib):
    assert lib.ns3d_create(None, None) == -1
    n = None
    assert lib.ns3d_rk3_ssp(n, n, n, n, n, n, n) == -1
    assert lib.ns3d_rk3_ssp_steps(n, 3, n, n, n, n, n, n) == -1
    assert lib.ns3d_rk3_ssp_graph(n, 3, n, n, n, n, n, n) == -1
    assert lib.ns3d_nonlinear(n, n, n, n, n, n, n) == -1
    assert lib.ns3d_fft_r2c(n, n, n) == -1
    assert lib.ns3d_probe(n, n, n, n, 0, 0, 0, None) == -1
    assert lib.ns3d_vec_axpby(n, 10, 1.0, n, 1.0, n) == -1
    assert lib.ns3d_last_error(n) == -1
    assert lib.ns3d_destroy(n) == -1


def test_status_strings_and_sizes(lib):
    for code in (0, -1, -2, -3, -4, -5, -6, -7):
        s = lib.ns3d_status_string(code).decode()
        assert s and s != lib.ns3d_status_string(12345).decode()
    for n in (8, 16, 32, 48, 64, 80, 96, 128, 160, 192, 256, 320, 384, 512):
        assert lib.ns3d_supported_size(n) == 1
    for n in (0, 4, 12, 24, 40, 100, 240, 1024):
        assert lib.ns3d_supported_size(n) == 0
