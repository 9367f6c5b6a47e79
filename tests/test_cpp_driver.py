"""The drop-in from host C++: tests/cpp/shapiro_driver.cpp calls the
reference-named wrappers of include/ns3d_compat.hpp on HIP device arrays.
It is compiled on the CPU (g++, host code only) and run on the GPU, where
its Shapiro error must equal the Python path's to the last digits."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "shapiro_driver.cpp")
SRC_GRAPH = os.path.join(ROOT, "tests", "cpp", "multi_graph_driver.cpp")


def _compile(out, src=SRC):
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I" + os.path.join(ROOT, "include"), src, "-o", str(out), "-L" + os.path.join(ROOT, "ns3dperiodic_amd"),
           "-lns3d_hip", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath," + os.path.join(ROOT, "ns3dperiodic_amd")]
    subprocess.run(cmd, check=True)


def test_driver_compiles(tmp_path):
    if not os.path.exists(os.path.join(ROOT, "ns3dperiodic_amd", "libns3d_hip.so")):
        pytest.skip("libns3d_hip.so is not built")
    _compile(tmp_path / "shapiro_driver")
    _compile(tmp_path / "multi_graph_driver", SRC_GRAPH)


@pytest.mark.gpu
def test_driver_runs_shapiro_32(tmp_path):
    from ns3dperiodic_amd import NS3D, fields
    import torch
    print("\n[case] C++ driver, Shapiro 32^3", flush=True)
    exe = tmp_path / "shapiro_driver"
    _compile(exe)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    err_cpp = float(r.stdout.split()[-1])
    g = NS3D(32, Re=20.0, dt=1e-3)
    u = [g.FFTN(torch.from_numpy(c).cuda()) for c in fields.shapiro(32)]
    g.RK3_SSP(u, None, steps=100)
    err_py = g.Shapiro_test_case(0.1, [g.iFFTN(c) for c in u])
    print(f"[cpp] shapiro L-inf: C++ {err_cpp:.17e}, Python {err_py:.17e}", flush=True)
    assert err_cpp < 1e-12
    assert abs(err_cpp - err_py) <= 1e-15


@pytest.mark.gpu
def test_driver_multi_size_graph(tmp_path):
    """tests/cpp/multi_graph_driver.cpp: a 32^3 and a 16^3 context captured
    from two forked streams into one HIP graph (plain C ABI), five replays of
    10 steps bitwise equal to 50 eager steps, and Shapiro-exact."""
    print("\n[case] C++ driver, 32^3 + 16^3 contexts in one graph", flush=True)
    exe = tmp_path / "multi_graph_driver"
    _compile(exe, SRC_GRAPH)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stderr
    lines = [ln.split() for ln in r.stdout.splitlines()]
    assert ["bitwise_equal", "1"] in lines
    errs = {int(p[1]): float(p[2]) for p in lines if p[0] == "shapiro_linf"}
    assert set(errs) == {16, 32}
    assert all(e < 1e-12 for e in errs.values()), errs
