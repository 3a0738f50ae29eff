"""ctypes binding of the CPU oracle (oracle/lbm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the parity reference.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "liblbm_oracle.so")

EDGE_ZERO, EDGE_PERIODIC = 0, 1
BB_REFERENCE, BB_HALFWAY = 0, 1
FLUID, WALL, LID, INLET, OUTLET = range(5)
BF16, F16 = 0, 1


class Params(C.Structure):
    _fields_ = [("w", C.c_int), ("h", C.c_int), ("edge", C.c_int), ("boundary", C.c_int),
                ("tau", C.c_double), ("delta_x", C.c_double), ("delta_t", C.c_double),
                ("lid_ux", C.c_double), ("lid_uy", C.c_double),
                ("inlet_ux", C.c_double), ("inlet_uy", C.c_double),
                ("outlet_rho", C.c_double), ("rho0", C.c_double),
                ("collision", C.c_int), ("tau_minus", C.c_double), ("ks_viscosity", C.c_double)]


class Consts32(C.Structure):
    _fields_ = [("K", C.c_float), ("A", C.c_float), ("B", C.c_float), ("C", C.c_float), ("cs2", C.c_float),
                ("w", C.c_float * 9), ("lid", C.c_float * 9), ("base", C.c_float * 9),
                ("wp", C.c_float), ("wm", C.c_float), ("hdt", C.c_float),
                ("dx", C.c_float), ("beta", C.c_float), ("sf", C.c_float * 9)]


class Consts64(C.Structure):
    _fields_ = [("K", C.c_double), ("A", C.c_double), ("B", C.c_double), ("C", C.c_double), ("cs2", C.c_double),
                ("w", C.c_double * 9), ("lid", C.c_double * 9), ("base", C.c_double * 9),
                ("wp", C.c_double), ("wm", C.c_double), ("hdt", C.c_double),
                ("dx", C.c_double), ("beta", C.c_double), ("sf", C.c_double * 9)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not built (make -C oracle)")
    return C.CDLL(LIB_PATH)


lib = _load()
_P = C.c_void_p
_PP = C.POINTER(Params)
for name, args, res in [
    ("orc_consts_f32_make", [_PP, C.POINTER(Consts32)], None),
    ("orc_consts_f64_make", [_PP, C.POINTER(Consts64)], None),
    ("orc_consts_devf32_make", [_PP, C.POINTER(Consts32)], None),
    ("orc_stream_f32", [C.c_int, C.c_int, C.c_int, _P, _P], None),
    ("orc_stream_f64", [C.c_int, C.c_int, C.c_int, _P, _P], None),
    ("orc_bounce_back_f32", [C.c_int, C.c_int, _P, _P, _P], None),
    ("orc_bounce_back_f64", [C.c_int, C.c_int, _P, _P, _P], None),
    ("orc_density_f32", [C.c_int, C.c_int, _P, _P], None),
    ("orc_density_f64", [C.c_int, C.c_int, _P, _P], None),
    ("orc_momentum_f32", [C.c_int, C.c_int, _P, _P, _P], None),
    ("orc_momentum_f64", [C.c_int, C.c_int, _P, _P, _P], None),
    ("orc_velocity_f32", [C.c_int, C.c_int, _P, _P, _P], None),
    ("orc_velocity_f64", [C.c_int, C.c_int, _P, _P, _P], None),
    ("orc_equilibrium_f32", [_PP, C.c_int, _P, _P, _P, _P], None),
    ("orc_equilibrium_f64", [_PP, C.c_int, _P, _P, _P, _P], None),
    ("orc_bgk_f32", [_PP, C.c_int, _P, _P, _P], None),
    ("orc_bgk_f64", [_PP, C.c_int, _P, _P, _P], None),
    ("orc_step_unfused_f32", [_PP, _P, _P, _P, _P], None),
    ("orc_step_unfused_f64", [_PP, _P, _P, _P, _P], None),
    ("orc_step_f32", [_PP, _P, _P, _P], None),
    ("orc_step_f64", [_PP, _P, _P, _P], None),
    ("orc_run_f32", [_PP, _P, _P, C.c_int], C.c_int),
    ("orc_run_f64", [_PP, _P, _P, C.c_int], C.c_int),
    ("orc_run_f32_omp", [_PP, _P, _P, C.c_int, C.c_int], C.c_int),
    ("orc_run_unfused_f32", [_PP, _P, _P, C.c_int], C.c_int),
    ("orc_step_devf32", [_PP, _P, _P, _P], None),
    ("orc_run_devf32", [_PP, _P, _P, C.c_int], C.c_int),
    ("orc_equilibrium_devf32", [_PP, C.c_int, _P, _P, _P, _P], None),
    ("orc_add_neq_f32", [_PP, _P, _P, _P, _P, _P], None),
    ("orc_trt_f32", [_PP, C.c_int, _P, _P], None),
    ("orc_trt_f64", [_PP, C.c_int, _P, _P], None),
    ("orc_regularized_f32", [_PP, C.c_int, _P, _P], None),
    ("orc_regularized_f64", [_PP, C.c_int, _P, _P], None),
    ("orc_kbc_f32", [_PP, C.c_int, _P, _P], None),
    ("orc_kbc_f64", [_PP, C.c_int, _P, _P], None),
    ("orc_trt_new", [C.c_float, C.c_float, C.c_float, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)], None),
    ("orc_add_neq_f64", [_PP, _P, _P, _P, _P, _P], None),
    ("orc_f32_to_bf16_n", [_P, _P, C.c_long], None),
    ("orc_f32_to_f16_n", [_P, _P, C.c_long], None),
    ("orc_step_dev16", [_PP, C.c_int, _P, _P, _P], None),
    ("orc_run_dev16", [_PP, C.c_int, _P, _P, C.c_int], C.c_int),
    ("orc_run_dev16_ex", [_PP, C.c_int, C.c_int, C.c_uint32, C.c_uint32, _P, _P, C.c_int], C.c_int),
    ("orc_f32_to_bf16_sr", [C.c_float, C.c_uint32], C.c_uint16),
    ("orc_f32_to_f16_sr", [C.c_float, C.c_uint32], C.c_uint16),
    ("orc_sr_key", [C.c_uint32, C.c_uint32], C.c_uint32),
    ("orc_sr_bits", [C.c_uint32, C.c_uint32, _P], None),
    ("orc_encode_dev16", [_PP, C.c_int, _P, _P], None),
    ("orc_decode_dev16", [_PP, C.c_int, _P, _P], None),
    ("orc_sums_f32", [C.c_int, C.c_int, _P, _P, _P], None),
    ("orc_sums_f64", [C.c_int, C.c_int, _P, _P, _P], None),
    ("orc_residual", [C.c_int, C.c_int, _P, _P, _P, _P, _P], C.c_double),
]:
    fn = getattr(lib, name)
    fn.argtypes = args
    fn.restype = res


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


COLL_BGK, COLL_TRT, COLL_REGULARIZED, COLL_KBC = 0, 1, 2, 3


def params(w, h, tau=1.0, edge=EDGE_ZERO, boundary=BB_REFERENCE, lid=(0.0, 0.0), inlet=(0.0, 0.0),
           outlet_rho=1.0, rho0=1.0, delta_x=1.0, delta_t=1.0, trt=None, regularized=False,
           kbc=None) -> Params:
    """trt = (tau_plus, tau_minus) selects the TRT operator (tau is then tau_plus);
    regularized=True the Regularized operator; kbc = ks_viscosity the KBC operator."""
    coll, tau_m, nu = COLL_BGK, float(tau), 0.0
    if trt is not None:
        coll, tau, tau_m = COLL_TRT, float(trt[0]), float(trt[1])
    elif regularized:
        coll = COLL_REGULARIZED
    elif kbc is not None:
        coll, nu = COLL_KBC, float(kbc)
    return Params(int(w), int(h), int(edge), int(boundary), float(tau), float(delta_x), float(delta_t),
                  float(lid[0]), float(lid[1]), float(inlet[0]), float(inlet[1]), float(outlet_rho), float(rho0),
                  int(coll), float(tau_m), nu)


def trt_new(lam, ks_viscosity, delta_x=1.0, delta_t=1.0):
    """TRT::new (lbm.rs:380-390) in f32: (tau_plus, tau_minus)."""
    tp, tm = C.c_float(), C.c_float()
    lib.orc_trt_new(float(lam), float(ks_viscosity), float(delta_x), float(delta_t), C.byref(tp), C.byref(tm))
    return tp.value, tm.value


def trt(p: Params, f):
    """TRT::evaluate (lbm.rs:401-444) of populations f [9][h][w]."""
    f = np.ascontiguousarray(f)
    out = np.empty_like(f)
    getattr(lib, "orc_trt_" + _sfx(f.dtype))(C.byref(p), int(f[0].size), _p(f), _p(out))
    return out


def regularized(p: Params, f):
    """Regularized::evaluate (lbm.rs:606-661) of populations f [9][h][w]."""
    f = np.ascontiguousarray(f)
    out = np.empty_like(f)
    getattr(lib, "orc_regularized_" + _sfx(f.dtype))(C.byref(p), int(f[0].size), _p(f), _p(out))
    return out


def kbc(p: Params, f):
    """KBC::evaluate (lbm.rs:468-572) of populations f [9][h][w] (p.ks_viscosity)."""
    f = np.ascontiguousarray(f)
    out = np.empty_like(f)
    getattr(lib, "orc_kbc_" + _sfx(f.dtype))(C.byref(p), int(f[0].size), _p(f), _p(out))
    return out


def consts_f32(p: Params) -> Consts32:
    k = Consts32()
    lib.orc_consts_f32_make(C.byref(p), C.byref(k))
    return k


def consts_devf32(p: Params) -> Consts32:
    k = Consts32()
    lib.orc_consts_devf32_make(C.byref(p), C.byref(k))
    return k


def consts_f64(p: Params) -> Consts64:
    k = Consts64()
    lib.orc_consts_f64_make(C.byref(p), C.byref(k))
    return k


def _flags(flags, p):
    if flags is None:
        return None
    f = np.ascontiguousarray(flags, dtype=np.uint8)
    assert f.shape == (p.h, p.w)
    return f


def _real(a):
    return np.float64 if a.dtype == np.float64 else np.float32


def _sfx(dt):
    return "f64" if dt == np.float64 else "f32"


def stream(f, edge=EDGE_ZERO):
    f = np.ascontiguousarray(f)
    out = np.empty_like(f)
    getattr(lib, "orc_stream_" + _sfx(f.dtype))(f.shape[2], f.shape[1], edge, _p(f), _p(out))
    return out


def bounce_back(f, flags):
    f = np.ascontiguousarray(f)
    fl = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint8)
    out = np.empty_like(f)
    getattr(lib, "orc_bounce_back_" + _sfx(f.dtype))(f.shape[2], f.shape[1], _p(fl), _p(f), _p(out))
    return out


def density(f):
    f = np.ascontiguousarray(f)
    out = np.empty(f.shape[1:], f.dtype)
    getattr(lib, "orc_density_" + _sfx(f.dtype))(f.shape[2], f.shape[1], _p(f), _p(out))
    return out


def momentum(f):
    f = np.ascontiguousarray(f)
    mx, my = np.empty(f.shape[1:], f.dtype), np.empty(f.shape[1:], f.dtype)
    getattr(lib, "orc_momentum_" + _sfx(f.dtype))(f.shape[2], f.shape[1], _p(f), _p(mx), _p(my))
    return mx, my


def velocity(f):
    f = np.ascontiguousarray(f)
    ux, uy = np.empty(f.shape[1:], f.dtype), np.empty(f.shape[1:], f.dtype)
    getattr(lib, "orc_velocity_" + _sfx(f.dtype))(f.shape[2], f.shape[1], _p(f), _p(ux), _p(uy))
    return ux, uy


def equilibrium(p: Params, rho, ux, uy):
    dt = _real(np.asarray(rho))
    rho, ux, uy = (np.ascontiguousarray(a, dtype=dt) for a in (rho, ux, uy))
    out = np.empty((9,) + rho.shape, dt)
    getattr(lib, "orc_equilibrium_" + _sfx(dt))(C.byref(p), int(rho.size), _p(rho), _p(ux), _p(uy), _p(out))
    return out


def bgk(p: Params, f, feq):
    f = np.ascontiguousarray(f)
    feq = np.ascontiguousarray(feq, dtype=f.dtype)
    out = np.empty_like(f)
    getattr(lib, "orc_bgk_" + _sfx(f.dtype))(C.byref(p), int(f[0].size), _p(f), _p(feq), _p(out))
    return out


def step_unfused(p: Params, f, flags=None):
    f = np.ascontiguousarray(f)
    out = np.empty_like(f)
    scratch = np.empty(21 * f[0].size, f.dtype)
    getattr(lib, "orc_step_unfused_" + _sfx(f.dtype))(C.byref(p), _p(_flags(flags, p)), _p(f), _p(out),
                                                       _p(scratch))
    return out


def step(p: Params, f, flags=None):
    f = np.ascontiguousarray(f)
    out = np.empty_like(f)
    getattr(lib, "orc_step_" + _sfx(f.dtype))(C.byref(p), _p(_flags(flags, p)), _p(f), _p(out))
    return out


def run(p: Params, f, nsteps, flags=None):
    f = np.array(f, copy=True, order="C")
    rc = getattr(lib, "orc_run_" + _sfx(f.dtype))(C.byref(p), _p(_flags(flags, p)), _p(f), int(nsteps))
    assert rc == 0
    return f


def run_omp(p: Params, f, nsteps, flags=None, threads=0):
    f = np.array(f, copy=True, order="C", dtype=np.float32)
    used = lib.orc_run_f32_omp(C.byref(p), _p(_flags(flags, p)), _p(f), int(nsteps), int(threads))
    assert used > 0
    return f, used


def run_unfused(p: Params, f, nsteps, flags=None):
    f = np.array(f, copy=True, order="C", dtype=np.float32)
    assert lib.orc_run_unfused_f32(C.byref(p), _p(_flags(flags, p)), _p(f), int(nsteps)) == 0
    return f


def to_bf16(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.shape, np.uint16)
    lib.orc_f32_to_bf16_n(_p(x), _p(out), x.size)
    return out


def to_f16(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.shape, np.uint16)
    lib.orc_f32_to_f16_n(_p(x), _p(out), x.size)
    return out


def encode_dev16(p: Params, kind, f):
    f = np.ascontiguousarray(f, dtype=np.float32)
    out = np.empty(f.shape, np.uint16)
    lib.orc_encode_dev16(C.byref(p), int(kind), _p(f), _p(out))
    return out


def decode_dev16(p: Params, kind, d):
    d = np.ascontiguousarray(d, dtype=np.uint16)
    out = np.empty(d.shape, np.float32)
    lib.orc_decode_dev16(C.byref(p), int(kind), _p(d), _p(out))
    return out


def step_devf32(p: Params, g, flags=None):
    g = np.ascontiguousarray(g, dtype=np.float32)
    out = np.empty_like(g)
    lib.orc_step_devf32(C.byref(p), _p(_flags(flags, p)), _p(g), _p(out))
    return out


def run_devf32(p: Params, g, nsteps, flags=None):
    g = np.array(g, copy=True, order="C", dtype=np.float32)
    assert lib.orc_run_devf32(C.byref(p), _p(_flags(flags, p)), _p(g), int(nsteps)) == 0
    return g


def equilibrium_dev(p: Params, rho, ux, uy):
    """Deviation-form equilibrium g_eq = f_eq - w_i rho0 in f32."""
    rho, ux, uy = (np.ascontiguousarray(a, dtype=np.float32) for a in (rho, ux, uy))
    out = np.empty((9,) + rho.shape, np.float32)
    lib.orc_equilibrium_devf32(C.byref(p), int(rho.size), _p(rho), _p(ux), _p(uy), _p(out))
    return out


def nonequilibrium_init(p: Params, rho, ux, uy, flags=None, dev=False):
    """f_eq(rho, u) (deviation form with dev=True) plus the first-order
    Chapman-Enskog non-equilibrium part at fluid cells."""
    dt = np.float32 if dev else _real(np.asarray(rho))
    rho, ux, uy = (np.ascontiguousarray(a, dtype=dt) for a in (rho, ux, uy))
    F = equilibrium_dev(p, rho, ux, uy) if dev else equilibrium(p, rho, ux, uy)
    getattr(lib, "orc_add_neq_" + _sfx(dt))(C.byref(p), _p(_flags(flags, p)), _p(rho), _p(ux), _p(uy), _p(F))
    return F


def to_dev(p: Params, f):
    """f32 deviations g_i = f_i - w_i rho0 of full populations."""
    k = consts_f32(p)
    return (np.asarray(f, np.float32) - np.array(k.base, np.float32)[:, None, None]).astype(np.float32)


def from_dev(p: Params, g):
    k = consts_f32(p)
    return (np.asarray(g, np.float32) + np.array(k.base, np.float32)[:, None, None]).astype(np.float32)


def run_dev16(p: Params, kind, d, nsteps, flags=None, rounding=0, seed=0, step0=0):
    d = np.array(d, copy=True, order="C", dtype=np.uint16)
    assert lib.orc_run_dev16_ex(C.byref(p), int(kind), int(rounding), int(seed) & 0xffffffff,
                                int(step0) & 0xffffffff, _p(_flags(flags, p)), _p(d), int(nsteps)) == 0
    return d


def sr_bits(key, cell):
    out = np.zeros(9, np.uint32)
    lib.orc_sr_bits(int(key) & 0xffffffff, int(cell) & 0xffffffff, _p(out))
    return out


def sums(f, flags=None):
    f = np.ascontiguousarray(f)
    out = np.zeros(3, np.float64)
    fl = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint8)
    getattr(lib, "orc_sums_" + _sfx(f.dtype))(f.shape[2], f.shape[1], _p(fl), _p(f), _p(out))
    return out


def residual(ux, uy, ux_prev, uy_prev, flags=None):
    arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (ux, uy, ux_prev, uy_prev)]
    fl = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint8)
    h, w = arrs[0].shape
    return lib.orc_residual(w, h, _p(fl), *[_p(a) for a in arrs])
