"""ctypes binding of libns3d_hip.so (the C ABI declared in include/ns3d.h).

The library is the only compute path: if it is missing or cannot be loaded
this module raises, there is no fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libns3d_hip.so")

NS3D_OK = 0
NS3D_FP64 = 0
NS3D_FP32_BULK = 1
NS3D_FP32_STATE = 2
NS3D_FP32_INCR = 3
NS3D_FLAG_NO_IPQ = 0x1
NS3D_FLAG_WENO_ZFIX = 0x2
NS3D_FLAG_MASK_NYQUIST = 0x4
NS3D_FLAG_CHECK_FINITE = 0x8
NS3D_ERR_NONFINITE = -5
NS3D_ERR_CAPACITY = -7
UPD_STAGE, UPD_UV_RHS, UPD_UV, UPD_IUV, UPD_RHS = 0, 1, 2, 3, 4


class Config(ctypes.Structure):
    _fields_ = [("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("Nz", ctypes.c_int),
                ("Lx", ctypes.c_double), ("Ly", ctypes.c_double), ("Lz", ctypes.c_double),
                ("Re", ctypes.c_double), ("dt", ctypes.c_double),
                ("precision", ctypes.c_int), ("flags", ctypes.c_uint),
                ("stream", ctypes.c_void_p), ("device", ctypes.c_int)]


P = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double
LL = ctypes.c_longlong
DP = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes); every symbol include/ns3d.h declares
SIGNATURES = {
    "ns3d_create": (I, [ctypes.POINTER(Config), ctypes.POINTER(P)]),
    "ns3d_destroy": (I, [P]),
    "ns3d_set_params": (I, [P, D, D]),
    "ns3d_set_stream": (I, [P, P]),
    "ns3d_synchronize": (I, [P]),
    "ns3d_last_error": (I, [P]),
    "ns3d_status_string": (ctypes.c_char_p, [I]),
    "ns3d_supported_size": (I, [I]),
    "ns3d_workspace_bytes": (LL, [P]),
    "ns3d_workspace_generation": (LL, [P]),
    "ns3d_fft_r2c": (I, [P, P, P]),
    "ns3d_fft_c2r": (I, [P, P, P]),
    "ns3d_fft_c2r_unscaled": (I, [P, P, P]),
    "ns3d_tables": (I, [P] + [P] * 12),
    "ns3d_nonlinear": (I, [P] + [P] * 6),
    "ns3d_bilinear": (I, [P] + [P] * 9),
    "ns3d_project_update": (I, [P, I] + [P] * 12),
    "ns3d_copy3": (I, [P] + [P] * 6),
    "ns3d_axpby3": (I, [P, P, P, P, D, P, P, P, D, P, P, P]),
    "ns3d_stage": (I, [P] + [P] * 6),
    "ns3d_rhs": (I, [P] + [P] * 9),
    "ns3d_rk3_ssp": (I, [P] + [P] * 6),
    "ns3d_rk3_ssp_steps": (I, [P, I] + [P] * 6),
    "ns3d_rk3_ssp_steps_noalloc": (I, [P, I] + [P] * 6),
    "ns3d_rk3_ssp_scalar": (I, [P, I] + [P] * 7 + [I]),
    "ns3d_rk3_ssp_graph": (I, [P, I] + [P] * 6),
    "ns3d_rk3_ssp_batch": (I, [P, I, I] + [P] * 4),
    "ns3d_rk3_ssp_batch_members": (I, [P, I, I] + [P] * 7 + [I]),
    "ns3d_kinetic_energy_spectral": (I, [P] + [P] * 4),
    "ns3d_dissipation_spectral": (I, [P] + [P] * 4),
    "ns3d_divergence": (I, [P] + [P] * 4),
    "ns3d_divergence_physical": (I, [P] + [P] * 5),
    "ns3d_weno_advect": (I, [P, I] + [P] * 5),
    "ns3d_weno_advection": (I, [P, I] + [P] * 6),
    "ns3d_rk3_ssp_uv_rhs": (I, [P] + [P] * 6),
    "ns3d_stage_uv": (I, [P] + [P] * 6),
    "ns3d_stage_iuv_rhs": (I, [P] + [P] * 6),
    "ns3d_rk3_ssp_uv": (I, [P] + [P] * 9),
    "ns3d_unpack_reduced": (I, [P] + [P] * 4),
    "ns3d_pack_reduced": (I, [P] + [P] * 4),
    "ns3d_unpack_full": (I, [P] + [P] * 4),
    "ns3d_pack_full": (I, [P] + [P] * 4),
    "ns3d_set_base_state": (I, [P] + [P] * 6 + [I, D]),
    "ns3d_jv_reduced": (None, [P, P, P]),
    "ns3d_jv_reduced_newton": (None, [P, P, P]),
    "ns3d_jv_exponential": (None, [P, P, P]),
    "ns3d_reduce_minmax": (I, [P, LL, P, DP]),
    "ns3d_reduce_sum": (I, [P, LL, P, DP]),
    "ns3d_check_finite": (I, [P, LL, P]),
    "ns3d_finite_status": (I, [P, I]),
    "ns3d_finite_flag_async": (I, [P, P]),
    "ns3d_vec_dot": (I, [P, LL, P, P, DP]),
    "ns3d_vec_gemv_t": (I, [P, LL, I, P, LL, P, DP]),
    "ns3d_vec_gemv": (I, [P, LL, I, P, LL, DP, D, D, P]),
    "ns3d_vec_gemm": (I, [P, LL, I, I, P, LL, DP, P, LL]),
    "ns3d_vec_axpby": (I, [P, LL, D, P, D, P]),
    "ns3d_curl": (I, [P] + [P] * 6 + [I]),
    "ns3d_helmholtz_filter": (I, [P, D] + [P] * 6),
    "ns3d_cutoff_filter": (I, [P, D] + [P] * 6),
    "ns3d_probe": (I, [P, P, P, P, I, I, I, DP]),
    "ns3d_probe_batch": (I, [P, I, P, I, I, I, DP]),
    "ns3d_kinetic_energy": (I, [P, P, P, P, DP]),
    "ns3d_dissipation": (I, [P, P, P, P, DP]),
    "ns3d_cfl_dt": (I, [P, D, P, P, P, DP]),
    "ns3d_shapiro_error": (I, [P, D, P, P, P, DP]),
    "ns3d_profile_enable": (I, [P, I]),
    "ns3d_profile_read": (I, [P, DP, ctypes.POINTER(LL)]),
}

_lib = None


class NS3DError(RuntimeError):
    pass


class NonFiniteError(NS3DError):
    """NS3D_ERR_NONFINITE: a checking call found a NaN or Inf."""


def load():
    """Load libns3d_hip.so and declare every entry point; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NS3DError(f"{LIB_PATH} is missing: build it with `python -m ns3dperiodic_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status, what=""):
    if status != NS3D_OK:
        msg = load().ns3d_status_string(status).decode()
        cls = NonFiniteError if status == NS3D_ERR_NONFINITE else NS3DError
        raise cls(f"{what}: status {status} ({msg})")
