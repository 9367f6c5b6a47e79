"""Host-side mirror of NS3DPeriodic's hot-path operators over the HIP C ABI.

Method names follow the reference functions they replace (RK_time_step.h,
advection_2_3.h, advection_WENO.h, diffusion.h, divergence.h,
min_max_reduction.h, math_support.h, the J.v callbacks of NS3D_periodic.cu).
Fields are torch tensors on the GPU in the reference layout:
  spectral  complex128 [Nx, Ny, Nz//2+1]  (interleaved, = cufftDoubleComplex)
  physical  float64    [Nx, Ny, Nz]
torch provides device memory and the stream; all arithmetic runs in the
hand-written gfx950 kernels of libns3d_hip.so.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib
from ._lib import check


def _ptr(t):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("ns3d fields must live on the GPU")
    if not t.is_contiguous():
        raise ValueError("ns3d fields must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


class NS3D:
    """One grid/box configuration; owns an ns3d_ctx (struct_NS3D_RK3_Call analogue)."""

    def __init__(self, Nx, Ny=None, Nz=None, Lx=2 * math.pi, Ly=2 * math.pi, Lz=2 * math.pi,
                 Re=100.0, dt=1e-3, precision="fp64", ipq=True, weno_zfix=False, device=0, stream=None,
                 mask_nyquist=False, check_finite=False):
        self.lib = _lib.load()
        Ny = Nx if Ny is None else Ny
        Nz = Nx if Nz is None else Nz
        self.Nx, self.Ny, self.Nz, self.Mz = Nx, Ny, Nz, Nz // 2 + 1
        self.Lx, self.Ly, self.Lz = Lx, Ly, Lz
        self.Re, self.dt = Re, dt
        self.device = torch.device("cuda", device)
        flags = ((0 if ipq else _lib.NS3D_FLAG_NO_IPQ) | (_lib.NS3D_FLAG_WENO_ZFIX if weno_zfix else 0)
                 | (_lib.NS3D_FLAG_MASK_NYQUIST if mask_nyquist else 0)
                 | (_lib.NS3D_FLAG_CHECK_FINITE if check_finite else 0))
        # "fp32": FP32 transforms, FP64 state; "fp32s": also FP32 state inside the RK3 loops;
        # "fp32i": FP32 transforms and FP32 stage increments over the FP64 state
        prec = {"fp64": _lib.NS3D_FP64, "fp32": _lib.NS3D_FP32_BULK, "fp32s": _lib.NS3D_FP32_STATE,
                "fp32i": _lib.NS3D_FP32_INCR}[precision]
        self.precision = precision
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        self.stream = stream
        cfg = _lib.Config(Nx, Ny, Nz, Lx, Ly, Lz, Re, dt, prec, flags, ctypes.c_void_p(stream.cuda_stream), device)
        h = ctypes.c_void_p()
        check(self.lib.ns3d_create(ctypes.byref(cfg), ctypes.byref(h)), "ns3d_create")
        self.ctx = h

    def set_stream(self, stream):
        """Move the context's launches to another torch stream (ns3d_set_stream)."""
        check(self.lib.ns3d_set_stream(self.ctx, ctypes.c_void_p(stream.cuda_stream)), "ns3d_set_stream")
        self.stream = stream

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.ns3d_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ helpers
    @property
    def cshape(self):
        return (self.Nx, self.Ny, self.Mz)

    @property
    def rshape(self):
        return (self.Nx, self.Ny, self.Nz)

    @property
    def n_complex(self):
        return self.Nx * self.Ny * self.Mz

    def cfield(self, n=1):
        z = [torch.zeros(self.cshape, dtype=torch.complex128, device=self.device) for _ in range(n)]
        return z if n > 1 else z[0]

    def rfield(self, n=1):
        z = [torch.zeros(self.rshape, dtype=torch.float64, device=self.device) for _ in range(n)]
        return z if n > 1 else z[0]

    def _c3(self, u):
        assert len(u) == 3
        for c in u:
            assert c.dtype == torch.complex128 and tuple(c.shape) == self.cshape, (c.dtype, c.shape)
        return [_ptr(c) for c in u]

    def _r3(self, u):
        assert len(u) == 3
        for c in u:
            assert c.dtype == torch.float64 and tuple(c.shape) == self.rshape, (c.dtype, c.shape)
        return [_ptr(c) for c in u]

    def _f3(self, f):
        """Forcing: None, or three components of which any may be None (zero)."""
        if f is None:
            return [None, None, None]
        assert len(f) == 3
        out = []
        for c in f:
            if c is not None:
                assert c.dtype == torch.complex128 and tuple(c.shape) == self.cshape, (c.dtype, c.shape)
            out.append(_ptr(c))
        return out

    def set_params(self, dt=None, Re=None):
        self.dt = self.dt if dt is None else dt
        self.Re = self.Re if Re is None else Re
        check(self.lib.ns3d_set_params(self.ctx, self.dt, self.Re), "ns3d_set_params")

    def synchronize(self):
        check(self.lib.ns3d_synchronize(self.ctx), "ns3d_synchronize")

    # ------------------------------------------------------------ a4 FFT layer
    def FFTN(self, x):
        """FFTN_Device: D2Z, unnormalised."""
        out = torch.empty(self.cshape, dtype=torch.complex128, device=self.device)
        check(self.lib.ns3d_fft_r2c(self.ctx, _ptr(x.contiguous()), _ptr(out)), "ns3d_fft_r2c")
        return out

    def iFFTN(self, X, scale=True):
        """iFFTN_Device: Z2D (scaled by 1/(NxNyNz) as with scale_double)."""
        out = torch.empty(self.rshape, dtype=torch.float64, device=self.device)
        fn = self.lib.ns3d_fft_c2r if scale else self.lib.ns3d_fft_c2r_unscaled
        check(fn(self.ctx, _ptr(X.contiguous()), _ptr(out)), "ns3d_fft_c2r")
        return out

    # ------------------------------------------------------------ a5 tables
    def tables(self):
        t = {k: torch.empty(self.cshape, dtype=torch.float64, device=self.device) for k in
             ("din_diffusion", "din_poisson", "AM11", "AM22", "AM33", "AM12", "AM13", "AM23", "mask")}
        t["kx"] = torch.empty(self.Nx, dtype=torch.float64, device=self.device)
        t["ky"] = torch.empty(self.Ny, dtype=torch.float64, device=self.device)
        t["kz"] = torch.empty(self.Mz, dtype=torch.float64, device=self.device)
        check(self.lib.ns3d_tables(self.ctx, *[_ptr(t[k]) for k in (
            "kx", "ky", "kz", "din_diffusion", "din_poisson", "AM11", "AM22", "AM33", "AM12", "AM13", "AM23",
            "mask")]), "ns3d_tables")
        return t

    # ------------------------------------------------------------ a3 nonlinear
    def calculate_convolution_2p3(self, u):
        Q = self.cfield(3)
        check(self.lib.ns3d_nonlinear(self.ctx, *self._c3(u), *self._c3(Q)), "ns3d_nonlinear")
        return Q

    def calculate_convolution_2p3_UV(self, u, v):
        Q = self.cfield(3)
        check(self.lib.ns3d_bilinear(self.ctx, *self._c3(u), *self._c3(v), *self._c3(Q)), "ns3d_bilinear")
        return Q

    # ------------------------------------------------------------ a6 update family (in place on u)
    def project_update(self, mode, u, Q, f=None):
        R = self.cfield(3) if mode == _lib.UPD_RHS else [None, None, None]
        Rp = self._c3(R) if mode == _lib.UPD_RHS else R
        check(self.lib.ns3d_project_update(self.ctx, mode, *self._c3(u), *self._f3(f), *self._c3(Q), *Rp),
              "ns3d_project_update")
        return R if mode == _lib.UPD_RHS else u

    # ------------------------------------------------------------ a7
    def copy_arrays(self, src, dst):
        check(self.lib.ns3d_copy3(self.ctx, *self._c3(src), *self._c3(dst)), "ns3d_copy3")

    def single_RK_step(self, a, wa, b, wb, dst):
        check(self.lib.ns3d_axpby3(self.ctx, *self._c3(a), wa, *self._c3(b), wb, *self._c3(dst)), "ns3d_axpby3")

    # ------------------------------------------------------------ a2 / a1 (in place on u)
    def single_time_step(self, u, f=None):
        check(self.lib.ns3d_stage(self.ctx, *self._c3(u), *self._f3(f)), "ns3d_stage")
        return u

    def return_RHS(self, u, f=None):
        R = self.cfield(3)
        check(self.lib.ns3d_rhs(self.ctx, *self._c3(u), *self._f3(f), *self._c3(R)), "ns3d_rhs")
        return R

    def RK3_SSP(self, u, f=None, steps=1):
        check(self.lib.ns3d_rk3_ssp_steps(self.ctx, steps, *self._c3(u), *self._f3(f)), "ns3d_rk3_ssp_steps")
        return u

    def RK3_SSP_graph(self, u, f=None, steps=1):
        """steps RK3 steps replayed from a captured HIP graph (needs a non-default stream)."""
        check(self.lib.ns3d_rk3_ssp_graph(self.ctx, steps, *self._c3(u), *self._f3(f)), "ns3d_rk3_ssp_graph")
        return u

    def RK3_SSP_batch(self, ub, f=None, steps=1):
        """steps RK3 steps of a batch of fields advanced together: ub is a
        complex128 tensor [B, 3, Nx, Ny, Mz] (member b, component k)."""
        assert ub.dtype == torch.complex128 and ub.is_contiguous() and tuple(ub.shape[1:]) == (3,) + self.cshape
        check(self.lib.ns3d_rk3_ssp_batch(self.ctx, int(ub.shape[0]), steps, _ptr(ub), *self._f3(f)),
              "ns3d_rk3_ssp_batch")
        return ub

    def RK3_SSP_batch_members(self, ub, f, dt, Re, steps, nsteps, graph=False):
        """nsteps RK3 steps of the members of ub ([B, 3, Nx, Ny, Mz] complex128) with per-member
        dt, Re (float64 device tensors [B]) and step counts (int32 device tensor [B]: member b takes
        the first steps[b] steps, 0 = idle); ns3d_rk3_ssp_batch_members, optionally graph-replayed."""
        assert ub.dtype == torch.complex128 and tuple(ub.shape[1:]) == (3,) + self.cshape and ub.is_contiguous()
        B = int(ub.shape[0])
        for t, dty in ((dt, torch.float64), (Re, torch.float64), (steps, torch.int32)):
            assert t.dtype == dty and t.numel() >= B and t.is_cuda and t.is_contiguous()
        check(self.lib.ns3d_rk3_ssp_batch_members(self.ctx, B, nsteps, _ptr(ub), *self._f3(f), _ptr(dt), _ptr(Re),
                                                   _ptr(steps), int(graph)), "ns3d_rk3_ssp_batch_members")
        return ub

    def RK3_SSP_scalar(self, u, theta, f=None, steps=1, scheme=5):
        """RK3 of the velocity plus a WENO-advected passive scalar theta (physical float64)."""
        assert theta.dtype == torch.float64 and tuple(theta.shape) == self.rshape
        check(self.lib.ns3d_rk3_ssp_scalar(self.ctx, steps, *self._c3(u), *self._f3(f), _ptr(theta), scheme),
              "ns3d_rk3_ssp_scalar")
        return u, theta

    # ------------------------------------------------------------ a8
    def divergence(self, u):
        d = self.cfield(1)
        check(self.lib.ns3d_divergence(self.ctx, *self._c3(u), _ptr(d)), "ns3d_divergence")
        return d

    def divergence_physical(self, u):
        d = self.cfield(1)
        r = self.rfield(1)
        check(self.lib.ns3d_divergence_physical(self.ctx, *self._c3(u), _ptr(d), _ptr(r)), "ns3d_divergence_physical")
        return d, r

    # ------------------------------------------------------------ a9
    def weno_advect(self, scheme, u, x):
        out = self.rfield(1)
        check(self.lib.ns3d_weno_advect(self.ctx, scheme, *self._r3(u), _ptr(x), _ptr(out)), "ns3d_weno_advect")
        return out

    def WenoAdvection(self, scheme, u_hat):
        Q = self.cfield(3)
        check(self.lib.ns3d_weno_advection(self.ctx, scheme, *self._c3(u_hat), *self._c3(Q)), "ns3d_weno_advection")
        return Q

    # ------------------------------------------------------------ a10 (in place on v; RK3_SSP_UV also on u)
    def RK3_SSP_UV_RHS(self, u, v):
        check(self.lib.ns3d_rk3_ssp_uv_rhs(self.ctx, *self._c3(u), *self._c3(v)), "ns3d_rk3_ssp_uv_rhs")
        return v

    def single_time_step_UV(self, u, v):
        check(self.lib.ns3d_stage_uv(self.ctx, *self._c3(u), *self._c3(v)), "ns3d_stage_uv")
        return v

    def single_time_step_iUV_RHS(self, u, v):
        check(self.lib.ns3d_stage_iuv_rhs(self.ctx, *self._c3(u), *self._c3(v)), "ns3d_stage_iuv_rhs")
        return v

    def RK3_SSP_UV(self, u, v, f=None):
        check(self.lib.ns3d_rk3_ssp_uv(self.ctx, *self._c3(u), *self._c3(v), *self._f3(f)), "ns3d_rk3_ssp_uv")
        return u, v

    def vector_length_reduced(self):
        return 3 * (self.n_complex - 1)

    def unpack_reduced(self, vec):
        v = self.cfield(3)
        check(self.lib.ns3d_unpack_reduced(self.ctx, _ptr(vec), *self._c3(v)), "ns3d_unpack_reduced")
        return v

    def pack_reduced(self, v):
        vec = torch.empty(self.vector_length_reduced(), dtype=torch.float64, device=self.device)
        check(self.lib.ns3d_pack_reduced(self.ctx, *self._c3(v), _ptr(vec)), "ns3d_pack_reduced")
        return vec

    def set_base_state(self, u, f=None, timesteps_period=1, tau=1.0):
        self._base = (u, f)  # keep the tensors alive while registered
        check(self.lib.ns3d_set_base_state(self.ctx, *self._c3(u), *self._f3(f), timesteps_period, tau),
              "ns3d_set_base_state")

    def jv_reduced(self, vec_in, newton=False):
        out = torch.empty_like(vec_in)
        fn = self.lib.ns3d_jv_reduced_newton if newton else self.lib.ns3d_jv_reduced
        fn(self.ctx, _ptr(vec_in), _ptr(out))
        check(self.lib.ns3d_last_error(self.ctx), "ns3d_jv_reduced")
        return out

    def jv_exponential(self, vec_in, vec_out):
        self.lib.ns3d_jv_exponential(self.ctx, _ptr(vec_in), _ptr(vec_out))
        check(self.lib.ns3d_last_error(self.ctx), "ns3d_jv_exponential")
        return vec_out

    # ------------------------------------------------------------ f4 diagnostics
    def get_curl(self, u, fix=False):
        """get_curl: vorticity of physical u by central differences (fix: see ns3d_curl)."""
        r = self.rfield(3)
        check(self.lib.ns3d_curl(self.ctx, *self._r3(u), *self._r3(r), int(fix)), "ns3d_curl")
        return r

    def Helmholz_Fourier_Filter(self, u_hat, eps):
        out = self.cfield(3)
        check(self.lib.ns3d_helmholtz_filter(self.ctx, eps, *self._c3(u_hat), *self._c3(out)), "ns3d_helmholtz_filter")
        return out

    def CutOff_Fourier_Filter(self, u_hat, radius):
        out = self.cfield(3)
        check(self.lib.ns3d_cutoff_filter(self.ctx, radius, *self._c3(u_hat), *self._c3(out)), "ns3d_cutoff_filter")
        return out

    # ------------------------------------------------------------ f3 return map probe
    def probe(self, u_hat, j, k, l):
        """(ux, uy, uz) at grid point (j, k, l) of the spectral field u_hat (host floats)."""
        out = (ctypes.c_double * 3)()
        check(self.lib.ns3d_probe(self.ctx, *self._c3(u_hat), j, k, l, out), "ns3d_probe")
        return np.array(out[:])

    def probe_batch(self, ub, j, k, l):
        """probe for every member of a batch ub ([B, 3, Nx, Ny, Mz] complex128): [B, 3] host array."""
        assert ub.dtype == torch.complex128 and tuple(ub.shape[1:]) == (3,) + self.cshape and ub.is_contiguous()
        B = int(ub.shape[0])
        out = (ctypes.c_double * (3 * B))()
        check(self.lib.ns3d_probe_batch(self.ctx, B, _ptr(ub), j, k, l, out), "ns3d_probe_batch")
        return np.array(out[:]).reshape(B, 3)

    # ------------------------------------------------------------ a11 / a12 (host results)
    def compute_reduction(self, x):
        out = (ctypes.c_double * 2)()
        check(self.lib.ns3d_reduce_minmax(self.ctx, x.numel(), _ptr(x), out), "ns3d_reduce_minmax")
        return out[0], out[1]

    def reduction_sum(self, x):
        out = ctypes.c_double()
        check(self.lib.ns3d_reduce_sum(self.ctx, x.numel(), _ptr(x), ctypes.byref(out)), "ns3d_reduce_sum")
        return out.value

    def check_nans(self, x):
        """check_nans_kernel: True if every value of x (any float64/complex128 device tensor) is finite."""
        x = x.contiguous()
        n = x.numel() * (2 if x.is_complex() else 1)
        st = self.lib.ns3d_check_finite(self.ctx, n, _ptr(x))
        if st == _lib.NS3D_ERR_NONFINITE:
            return False
        check(st, "ns3d_check_finite")
        return True

    def finite_status(self, clear=True):
        """False if a fused stage wrote a NaN or Inf since the flag was last cleared (synchronises)."""
        st = self.lib.ns3d_finite_status(self.ctx, int(clear))
        if st == _lib.NS3D_ERR_NONFINITE:
            return False
        check(st, "ns3d_finite_status")
        return True

    def finite_flag_async(self, host_int32):
        """Enqueue a copy of the non-finite flag (0/1) into a pinned int32 host tensor."""
        assert host_int32.dtype == torch.int32 and not host_int32.is_cuda
        check(self.lib.ns3d_finite_flag_async(self.ctx, ctypes.c_void_p(host_int32.data_ptr())),
              "ns3d_finite_flag_async")

    def kinetic_energy_spectral(self, u_hat):
        """0.5 * integral of |u|^2 over the box, exact from the spectral fields."""
        out = ctypes.c_double()
        check(self.lib.ns3d_kinetic_energy_spectral(self.ctx, *self._c3(u_hat), ctypes.byref(out)),
              "ns3d_kinetic_energy_spectral")
        return out.value

    def dissipation_spectral(self, u_hat):
        """Integral of 2 S:S over the box (|grad u|^2 + (div u)^2), exact from the spectral fields."""
        out = ctypes.c_double()
        check(self.lib.ns3d_dissipation_spectral(self.ctx, *self._c3(u_hat), ctypes.byref(out)),
              "ns3d_dissipation_spectral")
        return out.value

    def get_kinetic_energy(self, u):
        out = ctypes.c_double()
        check(self.lib.ns3d_kinetic_energy(self.ctx, *self._r3(u), ctypes.byref(out)), "ns3d_kinetic_energy")
        return out.value

    def get_dissipation(self, u):
        out = ctypes.c_double()
        check(self.lib.ns3d_dissipation(self.ctx, *self._r3(u), ctypes.byref(out)), "ns3d_dissipation")
        return out.value

    def calc_dt(self, cfl, u):
        out = ctypes.c_double()
        check(self.lib.ns3d_cfl_dt(self.ctx, cfl, *self._r3(u), ctypes.byref(out)), "ns3d_cfl_dt")
        return out.value

    def Shapiro_test_case(self, t, u):
        out = ctypes.c_double()
        check(self.lib.ns3d_shapiro_error(self.ctx, t, *self._r3(u), ctypes.byref(out)), "ns3d_shapiro_error")
        return out.value


class StepGraph:
    """K RK3 steps of several contexts -- any grid sizes and precisions --
    captured into one HIP graph: ``replay()`` advances every registered state
    by K steps with a single graph launch (no reference counterpart; for many
    small fields, whose kernels are too small to fill the GPU one at a time).

    items: (NS3D, u, f) triples; u (three spectral fields) is advanced in
    place by every replay, f may be None.  Items are dealt to ``branches``
    streams (largest grid first, each to the branch with the least work so
    far) forked from the capture stream and joined back, so
    the graph holds independent branches that the GPU runs concurrently;
    items on one branch run in order.  (Measured on MI355X with 24 fields of
    16^3..64^3 and 12 branches: 2.5-2.8x the throughput of eager steps on one
    stream, 1.03-1.18x that of eager steps on 12 streams.)  Each context is moved
    to its branch's stream (its later eager calls run there too).  Before
    capturing, each context takes one step on a copy of its state, so every
    buffer the step uses exists (capture allows no allocation) and the states
    are left untouched.  The step launches no memset, so the graph holds
    kernels only.  Each item's result is bitwise the one its context gives
    eagerly.  The graph holds the contexts' buffer addresses and dt/Re as they
    were at capture: ``replay()`` raises (and launches nothing) once a
    context was closed, reallocated a buffer the step uses
    (ns3d_workspace_generation changed: a bilinear B(u,v) call or a larger
    batch grows the scratch) or changed dt or Re; build a new StepGraph then.
    Buffers the step does not use (J.v unpacking, WENO scratch) may grow.

    Ordering: ``replay()`` runs the graph on its own stream after all work
    already queued on the caller's current stream and on every branch stream
    (where the contexts' eager calls run), and makes both wait for the replay,
    so eager calls before and after it need no host synchronisation.
    """

    def __init__(self, items, steps=1, branches=12, stream=None):
        self.items = list(items)
        self.steps = int(steps)
        self.stream = stream if stream is not None else torch.cuda.Stream()
        nb = max(1, min(int(branches), len(self.items)))
        self.branch_streams = [torch.cuda.Stream() for _ in range(nb)]
        self.branch_of = self.assign([self._cost(g) for g, _, _ in self.items], nb)
        torch.cuda.synchronize()
        for i, (g, u, f) in enumerate(self.items):
            s = self.branch_streams[self.branch_of[i]]
            g.set_stream(s)
            with torch.cuda.stream(s):
                g.RK3_SSP([c.clone() for c in u], f, steps=1)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=self.stream):
            for s in self.branch_streams:
                s.wait_stream(self.stream)
            for g, u, f in self.items:
                g.RK3_SSP(u, f, steps=self.steps)
            for s in self.branch_streams:
                self.stream.wait_stream(s)
        torch.cuda.synchronize()
        self._keys = [self._key(g) for g, _, _ in self.items]

    @staticmethod
    def _cost(g):
        """A step's cost: grid volume plus a fixed floor (a small grid's kernels cost their latency)."""
        return g.Nx * g.Ny * g.Nz + 32 ** 3

    @staticmethod
    def assign(costs, nb):
        """Branch of each item: largest cost first, each onto the branch with
        the least total cost so far (lowest index on ties)."""
        load = [0] * nb
        out = [0] * len(costs)
        for i in sorted(range(len(costs)), key=lambda i: (-costs[i], i)):
            b = min(range(nb), key=lambda b: (load[b], b))
            out[i] = b
            load[b] += costs[i]
        return out

    @staticmethod
    def _key(g):
        if not getattr(g, "ctx", None):
            return None
        return (g.ctx.value, int(g.lib.ns3d_workspace_generation(g.ctx)), g.dt, g.Re)

    def replay(self):
        for i, ((g, _, _), k) in enumerate(zip(self.items, self._keys)):
            if self._key(g) != k:
                raise RuntimeError(f"StepGraph: context {i} ({g.Nx}x{g.Ny}x{g.Nz}) was closed, reallocated a "
                                   "buffer of its step or changed dt/Re since the capture; build a new StepGraph")
        cur = torch.cuda.current_stream()
        self.stream.wait_stream(cur)
        for s in self.branch_streams:
            self.stream.wait_stream(s)
        with torch.cuda.stream(self.stream):
            self.graph.replay()
        cur.wait_stream(self.stream)
        for s in self.branch_streams:
            s.wait_stream(self.stream)
