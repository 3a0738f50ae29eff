"""ctypes binding of libchemsim_hip.so (include/chemsim.h).

The library is the product: a C ABI over hand-written gfx950 HIP kernels.
This module only declares its signatures.  There is no fallback: if the
shared library is missing or does not load, importing raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libchemsim_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "chemsim.h")

# enums (include/chemsim.h)
CS_OK, CS_ERR_ARG, CS_ERR_HIP, CS_ERR_UNSUPPORTED, CS_ERR_STATE, CS_ERR_NOMEM, CS_ERR_CALLBACK = range(7)
CS_RECORD_MAX_DEPTH = 16
CS_STEP_VECTOR = -1
CS_STEP_VECTOR_WIDE = -2
CS_TEMPORAL_AUTO = -1
CS_NOALLOC_MAX_VALUES = 3 << 20
CS_F32, CS_F64, CS_BF16DEV, CS_F16DEV, CS_F32DEV = range(5)
CS_EDGE_ZERO, CS_EDGE_PERIODIC = range(2)
CS_BB_REFERENCE, CS_BB_HALFWAY = range(2)
CS_FLUID, CS_WALL, CS_LID, CS_INLET, CS_OUTLET = range(5)
CS_ROUND_NEAREST, CS_ROUND_STOCHASTIC = range(2)
CS_ORDER_LARGEST_FIRST, CS_ORDER_GIVEN, CS_ORDER_SMALLEST_FIRST = range(3)
CS_LAUNCH_PERSISTENT, CS_LAUNCH_TILES, CS_LAUNCH_TEMPORAL, CS_LAUNCH_TEMPORAL_PERSISTENT = range(4)
CS_COLL_BGK, CS_COLL_TRT, CS_COLL_REGULARIZED, CS_COLL_KBC = range(4)
CS_RENDER_DENSITY, CS_RENDER_SPEED, CS_RENDER_VELOCITY, CS_RENDER_MOMENTUM = range(4)


class cs_params(C.Structure):
    _fields_ = [
        ("tau", C.c_double),
        ("delta_x", C.c_double),
        ("delta_t", C.c_double),
        ("edge", C.c_int32),
        ("boundary", C.c_int32),
        ("lid_ux", C.c_double),
        ("lid_uy", C.c_double),
        ("inlet_ux", C.c_double),
        ("inlet_uy", C.c_double),
        ("outlet_rho", C.c_double),
        ("rho0", C.c_double),
        ("rounding", C.c_int32),
        ("seed", C.c_uint32),
        ("collision", C.c_int32),
        ("tau_minus", C.c_double),
        ("ks_viscosity", C.c_double),
    ]


class cs_record_stats(C.Structure):
    _fields_ = [
        ("wall_ms", C.c_double),
        ("gpu_ms", C.c_double),
        ("wait_ms", C.c_double),
        ("callback_ms", C.c_double),
        ("frames", C.c_int64),
    ]


# int (*cs_frame_fn)(void* user, int64_t frame, const uint8_t* rgba, int w, int h)
FRAME_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int64, C.POINTER(C.c_uint8), C.c_int, C.c_int)


class cs_batch_tuning(C.Structure):
    _fields_ = [
        ("block_threads", C.c_int32),
        ("workgroups_per_cu", C.c_int32),
        ("cells_per_thread", C.c_int32),
        ("order", C.c_int32),
        ("launch", C.c_int32),
    ]


class ChemsimError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where} failed with status {code}: {msg}")
        self.code = code


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    return C.CDLL(LIB_PATH)


lib = _load()

_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_F = C.c_float
_I64 = C.c_int64
_SZ = C.c_size_t
_FP = C.POINTER(C.c_float)
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)
_I64P = C.POINTER(C.c_int64)

# name -> argtypes; every entry returns int status (except the two noted).
SIGNATURES = {
    "cs_ctx_create": [_I, C.POINTER(_P)],
    "cs_ctx_destroy": [_P],
    "cs_ctx_device_name": [_P, C.c_char_p, _SZ],
    "cs_lattice_create": [_P, _I, _I, _I, C.POINTER(_P)],
    "cs_lattice_destroy": [_P],
    "cs_lattice_shape": [_P, _IP, _IP, _IP],
    "cs_lattice_set_params": [_P, C.POINTER(cs_params)],
    "cs_lattice_set_flags": [_P, _P],
    "cs_lattice_upload_f32": [_P, _P],
    "cs_lattice_download_f32": [_P, _P],
    "cs_lattice_upload_f64": [_P, _P],
    "cs_lattice_download_f64": [_P, _P],
    "cs_lattice_upload_raw": [_P, _P],
    "cs_lattice_download_raw": [_P, _P],
    "cs_lattice_init_equilibrium": [_P, _P, _P, _P],
    "cs_lattice_init_equilibrium_f64": [_P, _P, _P, _P],
    "cs_lattice_init_nonequilibrium": [_P, _P, _P, _P],
    "cs_bgk_evaluate": [_P, C.c_int64, _P, _P, C.c_float, C.c_float, _P],
    "cs_trt_new": [C.c_float, C.c_float, C.c_float, C.c_float, _P, _P],
    "cs_lattice_render": [_P, _I, _I, _P, _FP],
    "cs_lattice_record": [_P, _I, _I, _I, _I, _I, FRAME_FN, _P, C.POINTER(cs_record_stats)],
    "cs_lattice_step": [_P, _I],
    "cs_lattice_step_timed": [_P, _I, _FP],
    "cs_lattice_temporal_steps": [_P, _IP],
    "cs_lattice_stream": [_P],
    "cs_lattice_bounce_back": [_P],
    "cs_lattice_collide": [_P],
    "cs_lattice_time": [_P, _DP],
    "cs_lattice_sync": [_P],
    "cs_lattice_device_ptr": [_P, C.POINTER(_P), _IP, _I64P],
    "cs_lattice_stream_handle": [_P, C.POINTER(_P)],
    "cs_lattice_density": [_P, _P],
    "cs_lattice_momentum_density": [_P, _P, _P],
    "cs_lattice_velocity": [_P, _P, _P],
    "cs_lattice_speed": [_P, _P],
    "cs_lattice_pressure": [_P, _P],
    "cs_lattice_equilibrium": [_P, _P],
    "cs_lattice_non_equilibrium": [_P, _P],
    "cs_lattice_is_unstable": [_P, _IP],
    "cs_lattice_moments_f64": [_P, _P, _P, _P],
    "cs_lattice_moments_f64_noalloc": [_P, _P, _P, _P],
    "cs_lattice_sums_f64": [_P, _DP],
    "cs_lattice_snapshot_velocity": [_P],
    "cs_lattice_residual_f64": [_P, _DP],
    "cs_lattice_run_to_residual": [_P, _D, _I, _I64, _I64P, _DP],
    "cs_batch_create": [_P, _I, _P, _P, _I, C.POINTER(_P)],
    "cs_batch_destroy": [_P],
    "cs_batch_count": [_P, _IP],
    "cs_batch_set_params": [_P, _I, C.POINTER(cs_params)],
    "cs_batch_set_flags": [_P, _I, _P],
    "cs_batch_init_equilibrium": [_P, _I, _P, _P, _P],
    "cs_batch_upload_f32": [_P, _I, _P],
    "cs_batch_download_f32": [_P, _I, _P],
    "cs_batch_download_raw": [_P, _I, _P],
    "cs_batch_velocity": [_P, _I, _P, _P],
    "cs_batch_set_tuning": [_P, C.POINTER(cs_batch_tuning)],
    "cs_batch_get_tuning": [_P, C.POINTER(cs_batch_tuning)],
    "cs_batch_step": [_P, _I, _FP],
    "cs_batch_step_per_lattice": [_P, _I, _FP],
    "cs_batch_autotune": [_P, _I, C.POINTER(cs_batch_tuning), _FP],
    "cs_batch_set_active": [_P, _I, _I],
    "cs_batch_snapshot_velocity": [_P],
    "cs_batch_residuals": [_P, _DP],
    "cs_batch_moments": [_P, _I, _P, _P, _P],
    "cs_cnn_create": [_P, _I, _I, _I, _I, C.c_uint32, C.POINTER(_P)],
    "cs_cnn_destroy": [_P],
    "cs_cnn_param_count": [_P, _I64P],
    "cs_cnn_get_params": [_P, _P],
    "cs_cnn_set_params": [_P, _P],
    "cs_cnn_get_grads": [_P, _P],
    "cs_cnn_forward": [_P, _I, _P, _P],
    "cs_cnn_loss_grad": [_P, _I, _P, _P, _P, _FP],
    "cs_cnn_adam_step": [_P, _F, _F, _F, _F],
    "cs_cnn_set_dataset": [_P, _I, _P, _P, _P],
    "cs_cnn_train_epoch": [_P, _P, _I, _F, _FP],
    "cs_cnn_time_step": [_P, _I, _FP, _DP],
    "cs_gemm_bf16": [_P, _I, _I, _I, _P, _P, _P, _FP],
    "cs_bench_copy": [_P, _I64, _I, _FP],
    "cs_check_reciprocal": [_P, _P],
    "cs_set_cells_per_thread": [_I],
    "cs_set_step_graphs": [_I],
    "cs_set_temporal_steps": [_I],
}

OPTIONAL = set()


def _declare():
    lib.cs_last_error.restype = C.c_char_p
    lib.cs_last_error.argtypes = []
    lib.cs_abi_version.restype = C.c_int
    lib.cs_abi_version.argtypes = []
    for name, args in SIGNATURES.items():
        try:
            fn = getattr(lib, name)
        except AttributeError:
            if name in OPTIONAL:
                continue
            raise
        fn.argtypes = args
        fn.restype = C.c_int


_declare()


def last_error() -> str:
    msg = lib.cs_last_error()
    return msg.decode() if msg else ""


def check(code: int, where: str) -> None:
    if code != CS_OK:
        raise ChemsimError(code, where, last_error())


def call(name: str, *args) -> None:
    check(getattr(lib, name)(*args), name)
