"""GPU parity of every hot-path operator against the CPU oracle (FP64).

Each test calls the HIP library through the C ABI (ns3dperiodic_amd.NS3D) and
compares with the oracle on the same seeded inputs.  Tolerance: 1e-12
relative L2 for FP64 (the per-step bar), exact for the mask, min/max and
Krylov packing.
"""
import numpy as np
import pytest

from ns3dperiodic_amd import fields
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-12
SIZES = [(16, 16, 16), (32, 32, 32), (32, 16, 48), (64, 48, 32)]


def _make(shape, **kw):
    from ns3dperiodic_amd import NS3D
    from oracle.oracle import Oracle
    Nx, Ny, Nz = shape
    kw.setdefault("Re", 100.0)
    kw.setdefault("dt", 2e-3)
    return NS3D(Nx, Ny, Nz, **kw), Oracle(Nx, Ny, Nz, **{k: v for k, v in kw.items() if k in ("Re", "dt", "Lx", "Ly", "Lz", "ipq", "weno_zfix")})


def _field(shape, seed=0, energy=0.5):
    return fields.random_solenoidal(*shape, seed=seed, energy=energy)


@pytest.fixture(scope="module", params=SIZES, ids=lambda s: "x".join(map(str, s)))
def pair(request):
    print(f"\n[case] grid {request.param}", flush=True)
    return _make(request.param)


def test_fft_r2c_c2r(pair):
    from tests.gpu_util import to_dev, to_host
    g, o = pair
    rng = np.random.default_rng(1)
    x = rng.standard_normal(o.rshape)
    X = to_host([g.FFTN(to_dev([x])[0])])[0]
    assert rel_l2(X, o.fft_r2c(x)) < 1e-14
    assert rel_l2(X, np.fft.rfftn(x)) < 1e-14
    # C2R of a non-Hermitian spectrum follows the documented convention (oracle == numpy irfftn)
    Z = rng.standard_normal(o.cshape) + 1j * rng.standard_normal(o.cshape)
    y = to_host([g.iFFTN(to_dev([Z])[0])])[0]
    assert rel_l2(y, o.fft_c2r(Z)) < 1e-14
    yu = to_host([g.iFFTN(to_dev([Z])[0], scale=False)])[0]
    assert rel_l2(yu, o.fft_c2r(Z, scale=False)) < 1e-14


def test_tables(pair):
    from tests.gpu_util import to_host
    g, o = pair
    tg = g.tables()
    to = o.tables()
    for k in ("kx", "ky", "kz", "mask"):
        np.testing.assert_array_equal(to_host([tg[k]])[0], to[k])
    for k in ("din_diffusion", "din_poisson", "AM11", "AM22", "AM33", "AM12", "AM13", "AM23"):
        assert rel_l2(to_host([tg[k]])[0], to[k]) < 1e-15, k


def test_nonlinear(pair):
    from tests.gpu_util import to_dev, to_host
    g, o = pair
    u = _field((o.Nx, o.Ny, o.Nz), seed=3)
    Qg = to_host(g.calculate_convolution_2p3(to_dev(u)))
    Qo = o.nonlinear(u)
    assert rel_l2(Qg, Qo) < TOL


def test_bilinear(pair):
    from tests.gpu_util import to_dev, to_host
    g, o = pair
    u = _field((o.Nx, o.Ny, o.Nz), seed=4)
    v = _field((o.Nx, o.Ny, o.Nz), seed=5, energy=0.1)
    Qg = to_host(g.calculate_convolution_2p3_UV(to_dev(u), to_dev(v)))
    Qo = o.bilinear(u, v)
    assert rel_l2(Qg, Qo) < TOL


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("fp32", 2e-6)])
def test_bilinear_small_perturbation(precision, tol):
    """B(u, s v) = s B(u, v) for a perturbation 1e-6 times smaller than u (the
    Krylov regime): the z pass must not share a transform between u- and
    v-derived fields, or v inherits u's absolute rounding."""
    from ns3dperiodic_amd import NS3D
    from oracle.oracle import Oracle
    from tests.gpu_util import to_dev, to_host
    sh = (32, 32, 32)
    print(f"\n[case] bilinear small perturbation {precision}", flush=True)
    g = NS3D(*sh, Re=100.0, dt=2e-3, precision=precision)
    o = Oracle(*sh, Re=100.0, dt=2e-3)
    u = _field(sh, seed=4)
    v = _field(sh, seed=5, energy=1e-12)
    Qg = to_host(g.calculate_convolution_2p3_UV(to_dev(u), to_dev(v)))
    Qo = o.bilinear(u, v)
    assert rel_l2(Qg, Qo) < tol


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
def test_project_update(pair, mode):
    from tests.gpu_util import to_dev, to_host
    g, o = pair
    sh = (o.Nx, o.Ny, o.Nz)
    u = _field(sh, seed=6)
    Q = _field(sh, seed=7)
    f = _field(sh, seed=8, energy=0.01)
    ud = to_dev(u)
    out = g.project_update(mode, ud, to_dev(Q), to_dev(f))
    Rg = to_host(out)
    Ro = o.update(mode, u, Q, f)
    assert rel_l2(Rg, Ro) < 1e-14


def test_stage_rhs_rk3(pair):
    from tests.gpu_util import to_dev, to_host
    g, o = pair
    sh = (o.Nx, o.Ny, o.Nz)
    u = _field(sh, seed=9)
    f = fields.kolmogorov_forcing_sin(*sh, n=2)
    # single_time_step
    ud = to_dev(u)
    g.single_time_step(ud, to_dev(f))
    assert rel_l2(to_host(ud), o.single_time_step(u, f)) < TOL
    # return_RHS
    assert rel_l2(to_host(g.return_RHS(to_dev(u), to_dev(f))), o.return_rhs(u, f)) < TOL
    # RK3_SSP, one and three steps
    ud = to_dev(u)
    g.RK3_SSP(ud, to_dev(f))
    assert rel_l2(to_host(ud), o.rk3(u, f)) < TOL
    ud = to_dev(u)
    g.RK3_SSP(ud, None, steps=3)
    zero = [np.zeros(o.cshape, np.complex128)] * 3
    assert rel_l2(to_host(ud), o.rk3(u, zero, steps=3)) < 3 * TOL


def test_tangent_family(pair):
    from tests.gpu_util import to_dev, to_host
    g, o = pair
    sh = (o.Nx, o.Ny, o.Nz)
    u = _field(sh, seed=10)
    v = _field(sh, seed=11, energy=1e-3)
    f = fields.kolmogorov_forcing_sin(*sh, n=1)
    assert rel_l2(to_host(g.RK3_SSP_UV_RHS(to_dev(u), to_dev(v))), o.rk3_uv_rhs(u, v)) < TOL
    assert rel_l2(to_host(g.single_time_step_iUV_RHS(to_dev(u), to_dev(v))), o.iuv_rhs(u, v)) < TOL
    ud, vd = to_dev(u), to_dev(v)
    g.RK3_SSP_UV(ud, vd, to_dev(f))
    uo, vo = o.rk3_uv(u, v, f)
    assert rel_l2(to_host(ud), uo) < TOL
    assert rel_l2(to_host(vd), vo) < TOL


def test_divergence(pair):
    from tests.gpu_util import to_dev, to_host
    g, o = pair
    rng = np.random.default_rng(12)
    u = [rng.standard_normal(o.cshape) + 1j * rng.standard_normal(o.cshape) for _ in range(3)]
    dg = to_host([g.divergence(to_dev(u))])[0]
    assert rel_l2(dg, o.divergence(u)) < 1e-15


@pytest.mark.parametrize("scheme", [3, 5, 1])
def test_weno(pair, scheme):
    from tests.gpu_util import to_dev, to_host
    g, o = pair
    sh = (o.Nx, o.Ny, o.Nz)
    u = [o.fft_c2r(c) for c in _field(sh, seed=13)]
    th = fields.random_scalar(*sh, seed=14)
    og = to_host([g.weno_advect(scheme, to_dev(u), to_dev([th])[0])])[0]
    assert rel_l2(og, o.weno_advect(scheme, u, th)) < 1e-13


def test_weno_spectral(pair):
    from tests.gpu_util import to_dev, to_host
    g, o = pair
    u = _field((o.Nx, o.Ny, o.Nz), seed=15)
    Qg = to_host(g.WenoAdvection(5, to_dev(u)))
    assert rel_l2(Qg, o.weno_advection_spectral(5, u)) < 1e-12


def test_reductions_and_diagnostics(pair):
    from tests.gpu_util import to_dev
    g, o = pair
    sh = (o.Nx, o.Ny, o.Nz)
    rng = np.random.default_rng(16)
    x = rng.standard_normal(sh) * 3.0
    xd = to_dev([x])[0]
    assert g.compute_reduction(xd) == tuple(o.reduce_minmax(x))
    # the integer seed of the reference (Q8) shows through when all values exceed it
    big = np.full(sh, 3.0e9)
    assert g.compute_reduction(to_dev([big])[0]) == (2139095040.0, 3.0e9)
    assert abs(g.reduction_sum(xd) - o.reduce_sum(x)) < 1e-12 * np.abs(x).sum()
    u = [o.fft_c2r(c) for c in _field(sh, seed=17)]
    ud = to_dev(u)
    assert abs(g.get_kinetic_energy(ud) - o.kinetic_energy(u)) < 1e-13 * abs(o.kinetic_energy(u))
    assert abs(g.get_dissipation(ud) - o.dissipation(u)) < 1e-13 * abs(o.dissipation(u))
    assert g.calc_dt(0.5, ud) == pytest.approx(o.cfl_dt(0.5, u), rel=1e-15)
    assert g.Shapiro_test_case(0.3, ud) == pytest.approx(o.shapiro_error(0.3, u), rel=1e-14)


def test_spectral_energy_and_dissipation(pair):
    """Physical diagnostics from the spectral fields against physical-space
    numpy sums (spectral derivatives), and tied to the reference formulas:
    with u_x = 0 the reference's transversal energy times dz is the spectral
    energy, and its central-difference dissipation converges to the spectral
    one at second order."""
    from tests.gpu_util import to_dev
    g, o = pair
    Nx, Ny, Nz = o.Nx, o.Ny, o.Nz
    L = 2 * np.pi
    rng = np.random.default_rng(23)
    uh = [np.fft.rfftn(rng.standard_normal((Nx, Ny, Nz))) for _ in range(3)]
    for c in uh:                              # no Nyquist modes: their derivative is not real
        c[Nx // 2] = 0
        c[:, Ny // 2] = 0
        c[:, :, Nz // 2] = 0
    u = [np.fft.irfftn(c, s=(Nx, Ny, Nz), axes=(0, 1, 2)) for c in uh]
    dV = L ** 3 / (Nx * Ny * Nz)
    e_ref = 0.5 * dV * sum(float(np.sum(x * x)) for x in u)
    kx = np.fft.fftfreq(Nx, 1.0 / Nx)[:, None, None]
    ky = np.fft.fftfreq(Ny, 1.0 / Ny)[None, :, None]
    kz = np.arange(Nz // 2 + 1)[None, None, :]
    ks = [kx, ky, kz]
    grad = [[np.fft.irfftn(1j * ks[i] * uh[j], s=(Nx, Ny, Nz), axes=(0, 1, 2)) for j in range(3)] for i in range(3)]
    d_ref = dV * sum(float(np.sum(0.5 * (grad[i][j] + grad[j][i]) ** 2)) for i in range(3) for j in range(3))
    uhd = to_dev(uh)
    assert g.kinetic_energy_spectral(uhd) == pytest.approx(e_ref, rel=1e-12)
    assert g.dissipation_spectral(uhd) == pytest.approx(d_ref, rel=1e-12)
    # the reference's formulas on the same flow
    z = [np.zeros_like(uh[0]), uh[1], uh[2]]
    ez = g.kinetic_energy_spectral(to_dev(z))
    assert g.get_kinetic_energy(to_dev([np.fft.irfftn(c, s=(Nx, Ny, Nz), axes=(0, 1, 2)) for c in z])) * (L / Nz) == \
        pytest.approx(ez, rel=1e-12)


def test_dissipation_formulas_converge():
    """The reference's central-difference dissipation tends to 4x the
    spectral value as (k dx)^2 for a smooth field: its difference quotients
    (f[j+1] - f[j-1])/dx (Source/math_support.cu:777-785) carry no 1/2, so
    every squared derivative is four times the derivative squared."""
    from tests.gpu_util import to_dev
    from ns3dperiodic_amd import NS3D
    errs = []
    for N in (16, 32, 64):
        g = NS3D(N)
        X, Y, Z = fields.grid(N, N, N)
        u = [np.sin(Y) * np.cos(2 * Z), np.sin(2 * Z) * np.cos(X), np.sin(X) * np.cos(Y)]
        spec = g.dissipation_spectral(to_dev([np.fft.rfftn(c) for c in u]))
        fd = g.get_dissipation(to_dev(u))
        errs.append(abs(0.25 * fd - spec) / spec)
        g.close()
    print(f"[diag] FD vs spectral dissipation: {errs}", flush=True)
    assert errs[0] > errs[1] > errs[2] > 0
    assert 3.5 < errs[1] / errs[2] < 4.5        # second order


def test_krylov_packing_and_jv(pair):
    from tests.gpu_util import to_dev, to_host
    import torch
    g, o = pair
    sh = (o.Nx, o.Ny, o.Nz)
    n = g.vector_length_reduced()
    rng = np.random.default_rng(18)
    vec = rng.standard_normal(n)
    vd = torch.from_numpy(vec).cuda()
    vg = to_host(g.unpack_reduced(vd))
    vo = o.unpack_reduced(vec)
    for a, b in zip(vg, vo):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(to_host([g.pack_reduced(to_dev(vo))])[0], o.pack_reduced(vo))
    u = _field(sh, seed=19)
    ud = to_dev(u)
    g.set_base_state(ud)
    out = to_host([g.jv_reduced(vd)])[0]
    assert rel_l2(out, o.jv_reduced(u, vec)) < TOL


def test_shapiro_32():
    """Shapiro Beltrami flow, N=32, nu=0.05, dt=1e-3, 100 RK3 steps."""
    from tests.gpu_util import to_dev, to_host
    from ns3dperiodic_amd import NS3D
    g = NS3D(32, Re=20.0, dt=1e-3)
    u0 = fields.shapiro(32)
    u = [g.FFTN(c) for c in to_dev(u0)]
    g.RK3_SSP(u, None, steps=100)
    ur = [g.iFFTN(c) for c in u]
    err = g.Shapiro_test_case(0.1, ur)
    print("[shapiro] L-inf error after 100 steps:", err)
    assert err < 1e-12


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5), ("fp32s", 1e-5)])
def test_shapiro_512(precision, tol):
    """The largest grid: the Shapiro flow at 512^3 (an exact solution whose
    nonlinear term is a pure gradient) stays on its analytic decay over 10
    RK3 steps, through every 512-point plan of the fused passes.  Re = 500:
    the explicit diffusion (Source/diffusion.cu:96-142) is stable only for
    dt*|k|max^2/Re below ~2.5 (RK3's real-axis limit), and |k|max^2 = 3*256^2
    here, so Re = 20 with dt = 1e-3 would amplify round-off in the highest
    modes by ~100 per step."""
    from tests.gpu_util import to_dev
    from ns3dperiodic_amd import NS3D
    import torch
    print(f"\n[case] shapiro 512^3 {precision}", flush=True)
    N, Re, dt, steps = 512, 500.0, 1e-3, 10
    g = NS3D(N, Re=Re, dt=dt, precision=precision)
    u = [g.FFTN(c) for c in to_dev(fields.shapiro(N))]
    g.RK3_SSP(u, None, steps=steps)
    ur = [g.iFFTN(c) for c in u]
    del u
    err = g.Shapiro_test_case(steps * dt, ur)
    print(f"[shapiro] 512^3 {precision}: L-inf error after {steps} steps: {err:.3e}", flush=True)
    g.close()
    del ur
    torch.cuda.empty_cache()
    assert err < tol


@pytest.mark.parametrize("shape", [(512, 16, 16), (16, 512, 16), (16, 16, 512), (16, 192, 96),
                                   (80, 16, 16), (16, 160, 16), (16, 16, 320), (320, 16, 16), (16, 80, 160),
                                   (384, 16, 16), (16, 16, 384)],
                         ids=lambda s: "x".join(map(str, s)))
def test_large_axes(shape):
    """Largest, radix-3 and radix-5 plans on each axis: FFT, nonlinear term and one RK3 step."""
    from tests.gpu_util import to_dev, to_host
    print(f"\n[case] grid {shape}", flush=True)
    g, o = _make(shape)
    rng = np.random.default_rng(21)
    x = rng.standard_normal(o.rshape)
    assert rel_l2(to_host([g.FFTN(to_dev([x])[0])])[0], np.fft.rfftn(x)) < 1e-14
    u = _field(shape, seed=22)
    assert rel_l2(to_host(g.calculate_convolution_2p3(to_dev(u))), o.nonlinear(u)) < TOL
    ud = to_dev(u)
    g.RK3_SSP(ud, None)
    zero = [np.zeros(o.cshape, np.complex128)] * 3
    assert rel_l2(to_host(ud), o.rk3(u, zero)) < TOL


@pytest.mark.parametrize("shape", [(80, 16, 32), (16, 32, 80), (32, 160, 16), (16, 16, 320), (48, 16, 96)],
                         ids=lambda s: "x".join(map(str, s)))
def test_bilinear_tangent_and_jv_on_radix5_and_radix3_axes(shape):
    """The bilinear term B(u,v), the tangent-map family and the reduced J.v
    (their own z-pass and epilogue kernels) on radix-5 and radix-3 axes
    against the oracle."""
    import torch
    from tests.gpu_util import to_dev, to_host
    print(f"\n[case] bilinear / tangent / J.v grid {shape}", flush=True)
    g, o = _make(shape)
    u = _field(shape, seed=41)
    v = _field(shape, seed=42, energy=1e-3)
    f = fields.kolmogorov_forcing_sin(*shape, n=1)
    assert rel_l2(to_host(g.calculate_convolution_2p3_UV(to_dev(u), to_dev(v))), o.bilinear(u, v)) < TOL
    assert rel_l2(to_host(g.RK3_SSP_UV_RHS(to_dev(u), to_dev(v))), o.rk3_uv_rhs(u, v)) < TOL
    assert rel_l2(to_host(g.single_time_step_iUV_RHS(to_dev(u), to_dev(v))), o.iuv_rhs(u, v)) < TOL
    ud, vd = to_dev(u), to_dev(v)
    g.RK3_SSP_UV(ud, vd, to_dev(f))
    uo, vo = o.rk3_uv(u, v, f)
    assert rel_l2(to_host(ud), uo) < TOL
    assert rel_l2(to_host(vd), vo) < TOL
    n = g.vector_length_reduced()
    vec = np.random.default_rng(43).standard_normal(n)
    g.set_base_state(to_dev(u))
    out = to_host([g.jv_reduced(torch.from_numpy(vec).cuda())])[0]
    assert rel_l2(out, o.jv_reduced(u, vec)) < TOL


@pytest.mark.parametrize("shape", [(32, 32, 32), (32, 16, 48)], ids=lambda s: "x".join(map(str, s)))
def test_rk3_with_passive_scalar(shape):
    """Velocity RK3 plus a WENO5-advected scalar (the decaying-turbulence passive scalar) against the oracle."""
    from tests.gpu_util import to_dev, to_host
    print(f"\n[case] scalar grid {shape}", flush=True)
    g, o = _make(shape, Re=500.0, dt=2e-3)
    u = _field(shape, seed=23)
    th = fields.random_scalar(*shape, seed=24)
    f = fields.kolmogorov_forcing_sin(*shape, n=1)
    ud, thd = to_dev(u), to_dev([th])[0]
    g.RK3_SSP_scalar(ud, thd, to_dev(f), steps=2, scheme=5)
    uo, tho = o.rk3_scalar(u, f, th, scheme=5, steps=2)
    assert rel_l2(to_host(ud), uo) < TOL
    assert rel_l2(to_host([thd])[0], tho) < TOL


@pytest.mark.parametrize("precision,tol", [("fp32", 2e-6), ("fp32s", 2e-6), ("fp32i", 2e-6)])
@pytest.mark.parametrize("shape", [(32, 32, 32), (48, 64, 32), (80, 32, 160)], ids=lambda s: "x".join(map(str, s)))
def test_rk3_fp32_paths(shape, precision, tol):
    """FP32 paths: five forced RK3 steps in one call against the FP64 oracle
    (the FP32-state path converts the state once on entry and once on exit)."""
    from tests.gpu_util import to_dev, to_host
    print(f"\n[case] rk3 {precision} grid {shape}", flush=True)
    g, o = _make(shape, precision=precision)
    u = _field(shape, seed=31)
    f = fields.kolmogorov_forcing_sin(*shape, n=2)
    ud = to_dev(u)
    g.RK3_SSP(ud, to_dev(f), steps=5)
    assert rel_l2(to_host(ud), o.rk3(u, f, steps=5)) < tol


@pytest.mark.parametrize("shape", [(32, 32, 32), (48, 64, 32)], ids=lambda s: "x".join(map(str, s)))
def test_physical_mode(shape):
    """NS3D_FLAG_MASK_NYQUIST with the IPQ rule off (the request server's mode):
    mask, nonlinear term and RK3 steps against the oracle in the same mode."""
    from ns3dperiodic_amd import NS3D
    from oracle.oracle import Oracle
    from tests.gpu_util import to_dev, to_host
    print(f"\n[case] physical mode grid {shape}", flush=True)
    g = NS3D(*shape, Re=300.0, dt=2e-3, ipq=False, mask_nyquist=True)
    o = Oracle(*shape, Re=300.0, dt=2e-3, ipq=False, mask_nyq=True)
    np.testing.assert_array_equal(to_host([g.tables()["mask"]])[0], o.tables()["mask"])
    u = _field(shape, seed=41)
    f = fields.kolmogorov_forcing_sin(*shape, n=4)
    assert rel_l2(to_host(g.calculate_convolution_2p3(to_dev(u))), o.nonlinear(u)) < TOL
    ud = to_dev(u)
    g.RK3_SSP(ud, to_dev(f), steps=3)
    assert rel_l2(to_host(ud), o.rk3(u, f, steps=3)) < 3 * TOL


@pytest.mark.parametrize("fix", [False, True])
def test_diagnostics_curl_and_filters(pair, fix):
    """f4 diagnostics: get_curl (both y-component variants), Helmholtz and cut-off filters."""
    from oracle import diagnostics as od
    from tests.gpu_util import to_dev, to_host
    g, o = pair
    sh = (o.Nx, o.Ny, o.Nz)
    u_hat = _field(sh, seed=51)
    u = [o.fft_c2r(c) for c in u_hat]
    r = to_host(g.get_curl(to_dev(u), fix=fix))
    ro = od.curl(*u, 2 * np.pi / o.Nx, 2 * np.pi / o.Ny, 2 * np.pi / o.Nz, fix=fix)
    assert rel_l2(r, list(ro)) < 1e-14
    fh = to_host(g.Helmholz_Fourier_Filter(to_dev(u_hat), 0.2))
    assert rel_l2(fh, od.helmholtz_filter(u_hat, 0.2, 2 * np.pi, 2 * np.pi, 2 * np.pi)) < 1e-15
    fc = to_host(g.CutOff_Fourier_Filter(to_dev(u_hat), 0.1))
    for a, b in zip(fc, od.cutoff_filter(u_hat, 0.1)):
        np.testing.assert_array_equal(a, b)
