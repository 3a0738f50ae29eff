"""Python handle over the batch C ABI: many independent lattices stepped by one
persistent launch (include/chemsim.h, cs_batch_*)."""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

from . import capi
from .capi import call
from .lattice import RAW_DTYPE, STORAGE, Context, Params, _ptr


class Batch:
    def __init__(self, ctx: Context, sizes, storage: str | int = "f32"):
        """sizes: sequence of (w, h)."""
        self.ctx = ctx
        self.sizes = [(int(w), int(h)) for w, h in sizes]
        self.storage = STORAGE[storage] if isinstance(storage, str) else int(storage)
        ws = np.array([s[0] for s in self.sizes], np.int32)
        hs = np.array([s[1] for s in self.sizes], np.int32)
        hdl = C.c_void_p()
        call("cs_batch_create", ctx.handle, len(self.sizes), _ptr(ws), _ptr(hs), self.storage, C.byref(hdl))
        self.handle = hdl

    def __len__(self):
        return len(self.sizes)

    @property
    def cells(self) -> int:
        return int(sum(w * h for w, h in self.sizes))

    def set_params(self, i: int, params: Params) -> None:
        p = params.to_c()
        call("cs_batch_set_params", self.handle, int(i), C.byref(p))

    def set_flags(self, i: int, flags) -> None:
        if flags is None:
            call("cs_batch_set_flags", self.handle, int(i), None)
            return
        w, h = self.sizes[i]
        f = np.ascontiguousarray(flags, dtype=np.uint8)
        assert f.shape == (h, w)
        call("cs_batch_set_flags", self.handle, int(i), _ptr(f))

    def init_equilibrium(self, i: int, rho, ux, uy) -> None:
        w, h = self.sizes[i]
        arrs = [np.ascontiguousarray(np.broadcast_to(a, (h, w)), dtype=np.float32) for a in (rho, ux, uy)]
        call("cs_batch_init_equilibrium", self.handle, int(i), *[_ptr(a) for a in arrs])

    def upload(self, i: int, f) -> None:
        w, h = self.sizes[i]
        a = np.ascontiguousarray(f, dtype=np.float32)
        assert a.shape == (9, h, w)
        call("cs_batch_upload_f32", self.handle, int(i), _ptr(a))

    def download(self, i: int) -> np.ndarray:
        w, h = self.sizes[i]
        out = np.empty((9, h, w), np.float32)
        call("cs_batch_download_f32", self.handle, int(i), _ptr(out))
        return out

    def download_raw(self, i: int) -> np.ndarray:
        w, h = self.sizes[i]
        out = np.empty((9, h, w), RAW_DTYPE[self.storage])
        call("cs_batch_download_raw", self.handle, int(i), _ptr(out))
        return out

    def velocity(self, i: int):
        w, h = self.sizes[i]
        ux, uy = np.empty((h, w), np.float32), np.empty((h, w), np.float32)
        call("cs_batch_velocity", self.handle, int(i), _ptr(ux), _ptr(uy))
        return ux, uy

    def set_active(self, i: int, active: bool) -> None:
        call("cs_batch_set_active", self.handle, int(i), 1 if active else 0)

    def snapshot_velocity(self) -> None:
        call("cs_batch_snapshot_velocity", self.handle)

    def residuals(self) -> np.ndarray:
        out = np.empty(len(self.sizes), np.float64)
        call("cs_batch_residuals", self.handle, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def moments(self, i: int):
        w, h = self.sizes[i]
        outs = [np.empty((h, w), np.float32) for _ in range(3)]
        call("cs_batch_moments", self.handle, int(i), *[_ptr(o) for o in outs])
        return tuple(outs)

    @property
    def tuning(self) -> dict:
        t = capi.cs_batch_tuning()
        call("cs_batch_get_tuning", self.handle, C.byref(t))
        return {"block_threads": t.block_threads, "workgroups_per_cu": t.workgroups_per_cu,
                "cells_per_thread": t.cells_per_thread, "order": t.order, "launch": t.launch}

    @tuning.setter
    def tuning(self, d: dict) -> None:
        t = capi.cs_batch_tuning(int(d["block_threads"]), int(d["workgroups_per_cu"]),
                                 int(d["cells_per_thread"]), int(d["order"]),
                                 int(d.get("launch", capi.CS_LAUNCH_PERSISTENT)))
        call("cs_batch_set_tuning", self.handle, C.byref(t))

    def step(self, n: int) -> float:
        """n steps of every active lattice (persistent or tile-list launches,
        by the tuning); returns milliseconds."""
        ms = C.c_float()
        call("cs_batch_step", self.handle, int(n), C.byref(ms))
        return ms.value

    def step_per_lattice(self, n: int) -> float:
        ms = C.c_float()
        call("cs_batch_step_per_lattice", self.handle, int(n), C.byref(ms))
        return ms.value

    def autotune(self, trial_steps: int = 20):
        t = capi.cs_batch_tuning()
        ms = C.c_float()
        call("cs_batch_autotune", self.handle, int(trial_steps), C.byref(t), C.byref(ms))
        return self.tuning, ms.value

    def close(self):
        if getattr(self, "handle", None):
            capi.lib.cs_batch_destroy(self.handle)
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


def signature(ctx: Context, batch: Batch) -> str:
    """Key of a tuning record: device, storage and the size histogram."""
    sizes = sorted(batch.sizes)
    h = hash(tuple(sizes)) & 0xffffffff
    return f"{ctx.device_name()}|storage={batch.storage}|n={len(sizes)}|cells={batch.cells}|sizes={h:08x}"


def tune(ctx: Context, batch: Batch, cache_path: str, trial_steps: int = 20) -> dict:
    """Autotuned launch shape for this batch, persisted in a JSON cache."""
    key = signature(ctx, batch)
    cache = {}
    if os.path.exists(cache_path):
        try:
            with open(cache_path) as fh:
                cache = json.load(fh)
        except (OSError, ValueError):
            cache = {}
    if key in cache:
        batch.tuning = cache[key]["tuning"]
        return cache[key]
    tuning, ms = batch.autotune(trial_steps)
    rec = {"tuning": tuning, "trial_ms": ms, "trial_steps": trial_steps}
    cache[key] = rec
    os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
    with open(cache_path, "w") as fh:
        json.dump(cache, fh, indent=1, sort_keys=True)
    return rec
