"""Python handle over the C ABI: one GPU context and D2Q9 lattices.

Arrays follow the reference host layout (src/matrix.rs:24-30,120-126):
populations are (9, h, w) in direction order 0..8 (src/lbm.rs:209-231),
scalar fields are (h, w), row r = image y, column c = image x.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import capi
from .capi import call

STORAGE = {"f32": capi.CS_F32, "f64": capi.CS_F64, "bf16": capi.CS_BF16DEV, "f16": capi.CS_F16DEV,
           "f32dev": capi.CS_F32DEV}
RAW_DTYPE = {capi.CS_F32: np.float32, capi.CS_F64: np.float64,
             capi.CS_BF16DEV: np.uint16, capi.CS_F16DEV: np.uint16, capi.CS_F32DEV: np.float32}


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


@dataclass
class Params:
    """Collision operator + discretization + boundary choices (src/lbm.rs:75-86,
    345-347; TRT: collision = CS_COLL_TRT, tau = tau_plus, lbm.rs:374-377;
    Regularized: collision = CS_COLL_REGULARIZED, lbm.rs:596-666; KBC:
    collision = CS_COLL_KBC with ks_viscosity, lbm.rs:458-460)."""
    tau: float = 1.0
    delta_x: float = 1.0
    delta_t: float = 1.0
    edge: int = capi.CS_EDGE_ZERO
    boundary: int = capi.CS_BB_REFERENCE
    lid_ux: float = 0.0
    lid_uy: float = 0.0
    inlet_ux: float = 0.0
    inlet_uy: float = 0.0
    outlet_rho: float = 1.0
    rho0: float = 1.0
    rounding: int = capi.CS_ROUND_NEAREST
    seed: int = 0
    collision: int = capi.CS_COLL_BGK
    tau_minus: float = 1.0
    ks_viscosity: float = 0.0

    def to_c(self) -> capi.cs_params:
        return capi.cs_params(self.tau, self.delta_x, self.delta_t, int(self.edge), int(self.boundary),
                              self.lid_ux, self.lid_uy, self.inlet_ux, self.inlet_uy, self.outlet_rho,
                              self.rho0, int(self.rounding), int(self.seed) & 0xffffffff, int(self.collision),
                              float(self.tau_minus), float(self.ks_viscosity))


def trt_new(lam: float, ks_viscosity: float, delta_x: float = 1.0, delta_t: float = 1.0):
    """TRT::new (src/lbm.rs:380-390): (tau_plus, tau_minus) in f32."""
    tp, tm = C.c_float(), C.c_float()
    call("cs_trt_new", float(lam), float(ks_viscosity), float(delta_x), float(delta_t), C.byref(tp), C.byref(tm))
    return tp.value, tm.value


class Context:
    """One bound GPU (af::init + af::set_backend, src/main.rs:331-332)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        call("cs_ctx_create", device, C.byref(h))
        self.handle = h

    def device_name(self) -> str:
        buf = C.create_string_buffer(256)
        call("cs_ctx_device_name", self.handle, buf, 256)
        return buf.value.decode()

    def bench_copy(self, nbytes: int, iters: int = 20) -> float:
        ms = C.c_float()
        call("cs_bench_copy", self.handle, int(nbytes), int(iters), C.byref(ms))
        return ms.value

    def check_reciprocal(self) -> int:
        """Inputs (of all 2^32) where the step kernels' reciprocal differs from IEEE 1/b."""
        n = C.c_uint64()
        call("cs_check_reciprocal", self.handle, C.byref(n))
        return n.value

    def close(self):
        if self.handle:
            capi.lib.cs_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Lattice:
    """A D2Q9 lattice on the GPU: State<D2Q9> with a BGK operator (src/lbm.rs:670-819)."""

    def __init__(self, ctx: Context, w: int, h: int, storage: str | int = "f32",
                 params: Params | None = None, flags: np.ndarray | None = None):
        self.ctx = ctx
        self.w, self.h = int(w), int(h)
        self.storage = STORAGE[storage] if isinstance(storage, str) else int(storage)
        hdl = C.c_void_p()
        call("cs_lattice_create", ctx.handle, self.w, self.h, self.storage, C.byref(hdl))
        self.handle = hdl
        self.params = params or Params()
        self.set_params(self.params)
        if flags is not None:
            self.set_flags(flags)

    # -- configuration -----------------------------------------------------
    def set_params(self, params: Params) -> None:
        self.params = params
        p = params.to_c()
        call("cs_lattice_set_params", self.handle, C.byref(p))

    def set_flags(self, flags: np.ndarray | None) -> None:
        if flags is None:
            call("cs_lattice_set_flags", self.handle, None)
            return
        f = np.ascontiguousarray(flags, dtype=np.uint8)
        assert f.shape == (self.h, self.w), f.shape
        call("cs_lattice_set_flags", self.handle, _ptr(f))

    # -- populations -------------------------------------------------------
    def upload(self, f: np.ndarray) -> None:
        if f.dtype == np.float64:
            a = np.ascontiguousarray(f)
            assert a.shape == (9, self.h, self.w)
            call("cs_lattice_upload_f64", self.handle, _ptr(a))
        else:
            a = np.ascontiguousarray(f, dtype=np.float32)
            assert a.shape == (9, self.h, self.w)
            call("cs_lattice_upload_f32", self.handle, _ptr(a))

    def download(self, dtype=np.float32) -> np.ndarray:
        out = np.empty((9, self.h, self.w), dtype=dtype)
        name = "cs_lattice_download_f64" if dtype == np.float64 else "cs_lattice_download_f32"
        call(name, self.handle, _ptr(out))
        return out

    def upload_raw(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=RAW_DTYPE[self.storage])
        assert a.shape == (9, self.h, self.w)
        call("cs_lattice_upload_raw", self.handle, _ptr(a))

    def download_raw(self) -> np.ndarray:
        out = np.empty((9, self.h, self.w), dtype=RAW_DTYPE[self.storage])
        call("cs_lattice_download_raw", self.handle, _ptr(out))
        return out

    def init_equilibrium(self, rho, ux, uy) -> None:
        """f = compute_equilibrium(rho, (ux, uy)) (src/lbm.rs:43-71)."""
        shape = (self.h, self.w)
        if any(np.asarray(a).dtype == np.float64 for a in (rho, ux, uy)):
            arrs = [np.ascontiguousarray(np.broadcast_to(a, shape), dtype=np.float64) for a in (rho, ux, uy)]
            call("cs_lattice_init_equilibrium_f64", self.handle, *[_ptr(a) for a in arrs])
        else:
            arrs = [np.ascontiguousarray(np.broadcast_to(a, shape), dtype=np.float32) for a in (rho, ux, uy)]
            call("cs_lattice_init_equilibrium", self.handle, *[_ptr(a) for a in arrs])

    def render(self, mode: int = capi.CS_RENDER_DENSITY, overlay: bool = True):
        """RGB image [h][w][3] of a display mode (src/render.rs:7-178) and the
        GPU milliseconds of the render passes."""
        out = np.empty((self.h, self.w, 3), np.uint8)
        ms = C.c_float()
        call("cs_lattice_render", self.handle, int(mode), int(bool(overlay)), _ptr(out), C.byref(ms))
        return out, ms.value

    def record(self, frames: int, steps_per_frame: int = 1, mode: int = capi.CS_RENDER_DENSITY,
               overlay: bool = True, depth: int = 3, on_frame=None):
        """record (src/display.rs:157-185): `frames` times steps_per_frame steps
        then the render pass into RGBA.  on_frame(k, rgba[h][w][4]) receives
        each frame in order (the array is valid only during the call; return
        False to stop); without it the frames are returned as one array
        [frames][h][w][4].  Returns (frames or None, stats dict)."""
        out = None if on_frame is not None else np.empty((frames, self.h, self.w, 4), np.uint8)
        err = []

        def cb(_user, k, ptr, w, h):
            try:
                view = np.ctypeslib.as_array(ptr, shape=(h, w, 4))
                if out is not None:
                    out[k] = view
                    return 0
                return 0 if on_frame(int(k), view) is not False else 1
            except BaseException as e:  # re-raised after the call
                err.append(e)
                return 1

        fn = capi.FRAME_FN(cb)
        st = capi.cs_record_stats()
        code = capi.lib.cs_lattice_record(self.handle, int(frames), int(steps_per_frame), int(mode),
                                          int(bool(overlay)), int(depth), fn, None, C.byref(st))
        if err:
            raise err[0]
        if code != capi.CS_ERR_CALLBACK:
            capi.check(code, "cs_lattice_record")
        stats = {k: getattr(st, k) for k, _ in capi.cs_record_stats._fields_}
        stats["stopped"] = code == capi.CS_ERR_CALLBACK
        return out, stats

    def init_nonequilibrium(self, rho, ux, uy) -> None:
        """f = f_eq(rho, u) + first-order Chapman-Enskog f_neq from grad u."""
        shape = (self.h, self.w)
        arrs = [np.ascontiguousarray(np.broadcast_to(a, shape), dtype=np.float32) for a in (rho, ux, uy)]
        call("cs_lattice_init_nonequilibrium", self.handle, *[_ptr(a) for a in arrs])

    # -- stepping ------------------------------------------------------------
    def step(self, n: int = 1) -> None:
        call("cs_lattice_step", self.handle, int(n))

    def step_timed(self, n: int) -> float:
        ms = C.c_float()
        call("cs_lattice_step_timed", self.handle, int(n), C.byref(ms))
        return ms.value

    def temporal_steps(self) -> int:
        """Steps per launch cs_lattice_step takes now (1: one-step kernels)."""
        k = C.c_int()
        call("cs_lattice_temporal_steps", self.handle, C.byref(k))
        return k.value

    def stream(self) -> None:
        call("cs_lattice_stream", self.handle)

    def bounce_back(self) -> None:
        call("cs_lattice_bounce_back", self.handle)

    def collide(self) -> None:
        call("cs_lattice_collide", self.handle)

    def sync(self) -> None:
        call("cs_lattice_sync", self.handle)

    @property
    def time(self) -> float:
        t = C.c_double()
        call("cs_lattice_time", self.handle, C.byref(t))
        return t.value

    # -- moments -------------------------------------------------------------
    def _planes(self, name: str, k: int):
        outs = [np.empty((self.h, self.w), np.float32) for _ in range(k)]
        call(name, self.handle, *[_ptr(o) for o in outs])
        return outs[0] if k == 1 else tuple(outs)

    def density(self):
        return self._planes("cs_lattice_density", 1)

    def momentum_density(self):
        return self._planes("cs_lattice_momentum_density", 2)

    def velocity(self):
        return self._planes("cs_lattice_velocity", 2)

    def speed(self):
        return self._planes("cs_lattice_speed", 1)

    def pressure(self):
        return self._planes("cs_lattice_pressure", 1)

    def equilibrium(self):
        out = np.empty((9, self.h, self.w), np.float32)
        call("cs_lattice_equilibrium", self.handle, _ptr(out))
        return out

    def non_equilibrium(self):
        out = np.empty((9, self.h, self.w), np.float32)
        call("cs_lattice_non_equilibrium", self.handle, _ptr(out))
        return out

    def is_unstable(self) -> bool:
        v = C.c_int()
        call("cs_lattice_is_unstable", self.handle, C.byref(v))
        return bool(v.value)

    def moments_f64(self):
        outs = [np.empty((self.h, self.w), np.float64) for _ in range(3)]
        call("cs_lattice_moments_f64", self.handle, *[_ptr(o) for o in outs])
        return tuple(outs)

    def moments_f64_noalloc(self, planes=(True, True, True)):
        """(rho, ux, uy) through the entry point that allocates nothing; a
        False in `planes` passes a null pointer for that plane (None back)."""
        outs = [np.empty((self.h, self.w), np.float64) if want else None for want in planes]
        call("cs_lattice_moments_f64_noalloc", self.handle, *[None if o is None else _ptr(o) for o in outs])
        return tuple(outs)

    def sums_f64(self) -> np.ndarray:
        out = (C.c_double * 3)()
        call("cs_lattice_sums_f64", self.handle, out)
        return np.array(out[:], dtype=np.float64)

    def snapshot_velocity(self) -> None:
        call("cs_lattice_snapshot_velocity", self.handle)

    def residual_f64(self) -> float:
        r = C.c_double()
        call("cs_lattice_residual_f64", self.handle, C.byref(r))
        return r.value

    def run_to_residual(self, tol: float, interval: int, max_steps: int):
        steps = C.c_int64()
        res = C.c_double()
        call("cs_lattice_run_to_residual", self.handle, float(tol), int(interval), int(max_steps),
             C.byref(steps), C.byref(res))
        return steps.value, res.value

    def close(self):
        if getattr(self, "handle", None):
            capi.lib.cs_lattice_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
