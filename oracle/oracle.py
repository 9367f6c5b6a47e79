"""ctypes front-end of the CPU oracle (test infrastructure only).

The oracle is a from-source FP64 restatement of NS3DPeriodic's hot path
(see ns3d_oracle.c for the reference citations and what pins each function).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product path (ns3dperiodic_amd) never does.

Arrays: spectral fields are complex128 [Nx, Ny, Nz//2+1]; physical fields are
float64 [Nx, Ny, Nz]; both C-contiguous (reference layout, Source/Macros.h:11,15).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libns3d_oracle.so")

UPD_STAGE, UPD_UV_RHS, UPD_UV, UPD_IUV, UPD_RHS = 0, 1, 2, 3, 4


class Params(ctypes.Structure):
    _fields_ = [("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("Nz", ctypes.c_int),
                ("Lx", ctypes.c_double), ("Ly", ctypes.c_double), ("Lz", ctypes.c_double),
                ("Re", ctypes.c_double), ("dt", ctypes.c_double),
                ("ipq", ctypes.c_int), ("weno_zfix", ctypes.c_int), ("nthreads", ctypes.c_int),
                ("mask_nyq", ctypes.c_int)]


def build(force=False):
    """Compile libns3d_oracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.or_reduce_sum.restype = ctypes.c_double
        for fn in ("or_kinetic_energy", "or_dissipation", "or_cfl_dt", "or_shapiro_error"):
            getattr(_lib, fn).restype = ctypes.c_double
    return _lib


class mutation:
    """Context manager planting mistake m in the oracle (test-only fault
    injection, see or_set_mutation in ns3d_oracle.c); restores 0 on exit."""

    def __init__(self, m):
        self.m = int(m)

    def __enter__(self):
        lib().or_set_mutation(self.m)
        return self

    def __exit__(self, *exc):
        lib().or_set_mutation(0)
        return False


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """One grid/box configuration of the CPU oracle."""

    def __init__(self, Nx, Ny=None, Nz=None, Lx=2 * np.pi, Ly=2 * np.pi, Lz=2 * np.pi,
                 Re=100.0, dt=1e-3, ipq=True, weno_zfix=False, nthreads=0, mask_nyq=False):
        Ny = Nx if Ny is None else Ny
        Nz = Nx if Nz is None else Nz
        self.Nx, self.Ny, self.Nz, self.Mz = Nx, Ny, Nz, Nz // 2 + 1
        self.Lx, self.Ly, self.Lz = Lx, Ly, Lz
        self.Re, self.dt = Re, dt
        self.ipq, self.weno_zfix, self.nthreads, self.mask_nyq = ipq, weno_zfix, nthreads, mask_nyq

    @property
    def params(self):
        return Params(self.Nx, self.Ny, self.Nz, self.Lx, self.Ly, self.Lz, self.Re, self.dt,
                      int(self.ipq), int(self.weno_zfix), int(self.nthreads), int(self.mask_nyq))

    @property
    def cshape(self):
        return (self.Nx, self.Ny, self.Mz)

    @property
    def rshape(self):
        return (self.Nx, self.Ny, self.Nz)

    def _c(self, a):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        assert a.shape == self.cshape, (a.shape, self.cshape)
        return a

    def _r(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == self.rshape, (a.shape, self.rshape)
        return a

    # FFT layer (Source/math_support.cu:301-334)
    def fft_r2c(self, x):
        x = self._r(x)
        out = np.empty(self.cshape, np.complex128)
        lib().or_fft_r2c(ctypes.byref(self.params), _p(x), _p(out))
        return out

    def fft_c2r(self, X, scale=True):
        X = self._c(X)
        out = np.empty(self.rshape, np.float64)
        lib().or_fft_c2r(ctypes.byref(self.params), _p(X), _p(out), int(scale))
        return out

    @staticmethod
    def fft_c2c_1d(x, sign):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        out = np.empty_like(x)
        lib().or_fft_c2c_1d(len(x), int(sign), _p(x), _p(out))
        return out

    def tables(self):
        n = self.cshape
        t = {k: np.empty(n, np.float64) for k in
             ("din_diffusion", "din_poisson", "AM11", "AM22", "AM33", "AM12", "AM13", "AM23", "mask")}
        kx, ky, kz = np.empty(self.Nx), np.empty(self.Ny), np.empty(self.Mz)
        lib().or_tables(ctypes.byref(self.params), _p(kx), _p(ky), _p(kz),
                        _p(t["din_diffusion"]), _p(t["din_poisson"]), _p(t["AM11"]), _p(t["AM22"]),
                        _p(t["AM33"]), _p(t["AM12"]), _p(t["AM13"]), _p(t["AM23"]), _p(t["mask"]))
        t.update(kx=kx, ky=ky, kz=kz)
        return t

    def nonlinear(self, u):
        u = [self._c(c) for c in u]
        Q = [np.empty(self.cshape, np.complex128) for _ in range(3)]
        lib().or_nonlinear(ctypes.byref(self.params), *map(_p, u), *map(_p, Q))
        return Q

    def bilinear(self, u, v):
        u = [self._c(c) for c in u]
        v = [self._c(c) for c in v]
        Q = [np.empty(self.cshape, np.complex128) for _ in range(3)]
        lib().or_bilinear(ctypes.byref(self.params), *map(_p, u), *map(_p, v), *map(_p, Q))
        return Q

    def update(self, mode, w, Q, f=None):
        w = [self._c(c).copy() for c in w]
        Q = [self._c(c) for c in Q]
        f = [self._c(c) for c in f] if f is not None else [None, None, None]
        R = [np.empty(self.cshape, np.complex128) for _ in range(3)]
        lib().or_update(ctypes.byref(self.params), int(mode), *map(_p, w), *map(_p, f), *map(_p, Q), *map(_p, R))
        return R if mode == UPD_RHS else w

    def single_time_step(self, u, f):
        u = [self._c(c).copy() for c in u]
        f = [self._c(c) for c in f]
        lib().or_single_time_step(ctypes.byref(self.params), *map(_p, u), *map(_p, f))
        return u

    def return_rhs(self, u, f):
        u = [self._c(c) for c in u]
        f = [self._c(c) for c in f]
        R = [np.empty(self.cshape, np.complex128) for _ in range(3)]
        lib().or_return_rhs(ctypes.byref(self.params), *map(_p, u), *map(_p, f), *map(_p, R))
        return R

    def rk3(self, u, f, steps=1):
        u = [self._c(c).copy() for c in u]
        f = [self._c(c) for c in f]
        for _ in range(steps):
            lib().or_rk3_ssp(ctypes.byref(self.params), *map(_p, u), *map(_p, f))
        return u

    def rk3_uv_rhs(self, u, v):
        u = [self._c(c) for c in u]
        v = [self._c(c).copy() for c in v]
        lib().or_rk3_ssp_uv_rhs(ctypes.byref(self.params), *map(_p, u), *map(_p, v))
        return v

    def iuv_rhs(self, u, v):
        u = [self._c(c) for c in u]
        v = [self._c(c).copy() for c in v]
        lib().or_single_time_step_iuv_rhs(ctypes.byref(self.params), *map(_p, u), *map(_p, v))
        return v

    def rk3_uv(self, u, v, f):
        u = [self._c(c).copy() for c in u]
        v = [self._c(c).copy() for c in v]
        f = [self._c(c) for c in f]
        lib().or_rk3_ssp_uv(ctypes.byref(self.params), *map(_p, u), *map(_p, v), *map(_p, f))
        return u, v

    def rk3_scalar(self, u, f, theta, scheme=5, steps=1):
        u = [self._c(c).copy() for c in u]
        f = [self._c(c) for c in f]
        th = self._r(theta).copy()
        for _ in range(steps):
            lib().or_rk3_ssp_scalar(ctypes.byref(self.params), *map(_p, u), *map(_p, f), _p(th), int(scheme))
        return u, th

    def divergence(self, u):
        u = [self._c(c) for c in u]
        d = np.empty(self.cshape, np.complex128)
        lib().or_divergence(ctypes.byref(self.params), *map(_p, u), _p(d))
        return d

    def weno_advect(self, scheme, u, x):
        u = [self._r(c) for c in u]
        x = self._r(x)
        out = np.empty(self.rshape, np.float64)
        lib().or_weno_advect(ctypes.byref(self.params), int(scheme), *map(_p, u), _p(x), _p(out))
        return out

    def weno_advection_spectral(self, scheme, u_hat):
        u = [self._c(c) for c in u_hat]
        Q = [np.empty(self.cshape, np.complex128) for _ in range(3)]
        lib().or_weno_advection_spectral(ctypes.byref(self.params), int(scheme), *map(_p, u), *map(_p, Q))
        return Q

    @staticmethod
    def reduce_minmax(x):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        out = np.empty(2)
        lib().or_reduce_minmax(ctypes.c_long(x.size), _p(x), _p(out))
        return out

    @staticmethod
    def reduce_sum(x):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        return lib().or_reduce_sum(ctypes.c_long(x.size), _p(x))

    def kinetic_energy(self, u):
        u = [self._r(c) for c in u]
        return lib().or_kinetic_energy(ctypes.byref(self.params), *map(_p, u))

    def dissipation(self, u):
        u = [self._r(c) for c in u]
        return lib().or_dissipation(ctypes.byref(self.params), *map(_p, u))

    def cfl_dt(self, cfl, u):
        u = [self._r(c) for c in u]
        return lib().or_cfl_dt(ctypes.byref(self.params), ctypes.c_double(cfl), *map(_p, u))

    def shapiro_error(self, t, u):
        u = [self._r(c) for c in u]
        return lib().or_shapiro_error(ctypes.byref(self.params), ctypes.c_double(t), *map(_p, u))

    def pack_reduced(self, v):
        v = [self._c(c) for c in v]
        vec = np.empty(3 * (self.Nx * self.Ny * self.Mz - 1))
        lib().or_pack_reduced(ctypes.byref(self.params), *map(_p, v), _p(vec))
        return vec

    def unpack_reduced(self, vec):
        vec = np.ascontiguousarray(vec, dtype=np.float64)
        assert vec.size == 3 * (self.Nx * self.Ny * self.Mz - 1)
        v = [np.empty(self.cshape, np.complex128) for _ in range(3)]
        lib().or_unpack_reduced(ctypes.byref(self.params), _p(vec), *map(_p, v))
        return v

    def jv_reduced(self, u, vec):
        u = [self._c(c) for c in u]
        vec = np.ascontiguousarray(vec, dtype=np.float64)
        out = np.empty_like(vec)
        lib().or_jv_reduced(ctypes.byref(self.params), *map(_p, u), _p(vec), _p(out))
        return out

    def pack_full(self, v, vec=None):
        """A_vector_from_velocities_kernel: writes vec[c*N + i] for 0 < i < N into vec (a copy
        of the given one, or zeros); the i = 0 entries keep their value."""
        v = [self._c(c) for c in v]
        N = self.Nx * self.Ny * self.Mz
        out = np.zeros(3 * N) if vec is None else np.array(vec, dtype=np.float64, copy=True)
        lib().or_pack_full(ctypes.byref(self.params), *map(_p, v), _p(out))
        return out

    def unpack_full(self, vec):
        vec = np.ascontiguousarray(vec, dtype=np.float64)
        assert vec.size == 3 * self.Nx * self.Ny * self.Mz
        v = [np.empty(self.cshape, np.complex128) for _ in range(3)]
        lib().or_unpack_full(ctypes.byref(self.params), _p(vec), *map(_p, v))
        return v

    def jv_reduced_newton(self, u, vec, tau):
        u = [self._c(c) for c in u]
        vec = np.ascontiguousarray(vec, dtype=np.float64)
        out = np.empty_like(vec)
        lib().or_jv_reduced_newton(ctypes.byref(self.params), *map(_p, u), ctypes.c_double(tau), _p(vec), _p(out))
        return out

    def jv_exponential(self, u, f, timesteps_period, vec, vec_out):
        """Returns (vec_out after the call, the base flow advanced by timesteps_period steps)."""
        u = [self._c(c).copy() for c in u]
        f = [self._c(c) for c in f]
        vec = np.ascontiguousarray(vec, dtype=np.float64)
        out = np.array(vec_out, dtype=np.float64, copy=True)
        lib().or_jv_exponential(ctypes.byref(self.params), *map(_p, u), *map(_p, f), int(timesteps_period),
                                _p(vec), _p(out))
        return out, u
