"""Python view of one device context (the C ABI of include/fluidsim_amd.h).

`DeviceSim` mirrors the reference's DeviceCalculation stage functions
(include/calculation.cuh:12-20) by name; every method is a thin call into
libfluidsim_amd.so and raises FsaError on a non-zero status.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native
from ._native import FLUID, SOLID, VIRT, FsaParams, check

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)


def _f32(a, n_cols=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a


def _ptr(a, t=_fp):
    return None if a is None else a.ctypes.data_as(t)


class DeviceSim:
    def __init__(self, params: FsaParams, stream: int | None = None):
        self.lib = _native.device_lib()
        self.params = params
        self.nf, self.ns, self.nv = params.n_fluid, params.n_solid, params.n_virt
        self.ncell = params.cell_number_total
        h = C.c_void_p()
        check(self.lib.fsa_create(C.byref(params), C.c_void_p(stream) if stream else None, C.byref(h)))
        self.handle = h

    # ---- lifetime ----------------------------------------------------------
    def close(self):
        if getattr(self, "handle", None):
            self.lib.fsa_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- uploads -------------------------------------------------------------
    def upload(self, a: dict):
        L = self.lib
        if self.nf:
            x, u = _f32(a["fluid_position"]), _f32(a["fluid_velocity"])
            r, p = _f32(a["fluid_density"]), _f32(a["fluid_pressure"])
            check(L.fsa_upload_fluid(self.handle, _ptr(x), _ptr(u), _ptr(r), _ptr(p)))
        if self.ns:
            x, u, X = _f32(a["solid_position"]), _f32(a["solid_velocity"]), _f32(a["solid_reference_position"])
            t = np.ascontiguousarray(a["solid_type"], dtype=np.int32)
            check(L.fsa_upload_solid(self.handle, _ptr(x), _ptr(u), _ptr(X), _ptr(t, _ip)))
        if self.nv:
            x = _f32(a["virt_position"])
            t = np.ascontiguousarray(a["virt_type"], dtype=np.int32)
            check(L.fsa_upload_virt(self.handle, _ptr(x), _ptr(t, _ip)))

    # ---- stages (DeviceCalculation) -------------------------------------------
    def particle_cell_update(self):
        check(self.lib.fsa_particle_cell_update(self.handle))

    def fill_cell(self):
        check(self.lib.fsa_fill_cell(self.handle))

    def couple_calculation(self):
        check(self.lib.fsa_couple_calculation(self.handle))

    def virt_update(self):
        check(self.lib.fsa_virt_update(self.handle))

    def fluid_calculation(self):
        check(self.lib.fsa_fluid_calculation(self.handle))

    def solid_calculation(self):
        check(self.lib.fsa_solid_calculation(self.handle))

    def step(self, n: int = 1):
        check(self.lib.fsa_step(self.handle, int(n)))

    def synchronize(self):
        check(self.lib.fsa_synchronize(self.handle))

    # ---- downloads (.dat order) ------------------------------------------------
    def fluid(self) -> dict:
        n = self.nf
        x, u = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32)
        r, p = np.zeros(n, np.float32), np.zeros(n, np.float32)
        check(self.lib.fsa_download_fluid(self.handle, _ptr(x), _ptr(u), _ptr(r), _ptr(p)))
        return {"position": x, "velocity": u, "density": r, "pressure": p}

    def fluid_rates(self) -> dict:
        n = self.nf
        drho, a = np.zeros(n, np.float32), np.zeros((n, 3), np.float32)
        check(self.lib.fsa_download_fluid_rates(self.handle, _ptr(drho), _ptr(a)))
        return {"drho_dt": drho, "du_dt": a}

    def pair_counts(self) -> np.ndarray:
        """(N, 3) partners per fluid particle and kind found by the last fluid stage (debug runs)."""
        out = np.zeros((self.nf, 3), np.int32)
        check(self.lib.fsa_download_pair_counts(self.handle, out.ctypes.data_as(C.POINTER(C.c_int32))))
        return out

    def solid(self) -> dict:
        n = self.ns
        out = {k: np.zeros((n, 3), np.float32) for k in ("position", "velocity", "normal", "du_dt", "couple_dudt")}
        out["pressure"] = np.zeros(n, np.float32)
        for k in ("deformation", "stress_1st_piola", "correction"):
            out[k] = np.zeros((n, 9), np.float32)
        check(self.lib.fsa_download_solid(self.handle, _ptr(out["position"]), _ptr(out["velocity"]),
                                          _ptr(out["pressure"]), _ptr(out["normal"]), _ptr(out["du_dt"]),
                                          _ptr(out["couple_dudt"]), _ptr(out["deformation"]),
                                          _ptr(out["stress_1st_piola"]), _ptr(out["correction"])))
        return out

    def virt(self) -> dict:
        n = self.nv
        out = {k: np.zeros((n, 3), np.float32) for k in ("position", "normal", "velocity")}
        out["density"] = np.zeros(n, np.float32)
        out["pressure"] = np.zeros(n, np.float32)
        check(self.lib.fsa_download_virt(self.handle, _ptr(out["position"]), _ptr(out["normal"]),
                                         _ptr(out["velocity"]), _ptr(out["density"]), _ptr(out["pressure"])))
        return out

    # ---- introspection ---------------------------------------------------------
    def _count(self, kind):
        return {FLUID: self.nf, SOLID: self.ns, VIRT: self.nv}[kind]

    def cell_ids(self, kind: int) -> np.ndarray:
        out = np.zeros(self._count(kind), np.int32)
        check(self.lib.fsa_get_cell_ids(self.handle, kind, _ptr(out, _ip)))
        return out

    def cell_start(self, kind: int) -> np.ndarray:
        out = np.zeros(self.ncell + 1, np.int32)
        check(self.lib.fsa_get_cell_start(self.handle, kind, _ptr(out, _ip)))
        return out

    def sorted_ids(self, kind: int) -> np.ndarray:
        out = np.zeros(self._count(kind), np.int32)
        check(self.lib.fsa_get_sorted_ids(self.handle, kind, _ptr(out, _ip)))
        return out

    def jump_count(self) -> int:
        v = C.c_int64()
        check(self.lib.fsa_get_jump_count(self.handle, C.byref(v)))
        return int(v.value)

    def list_stats(self) -> dict:
        """Neighbour-list builds since the last fluid upload; the particles the
        last build could not list, the candidates it listed, and the steps
        walked without lists (fsa_get_list_stats)."""
        v = [C.c_int64(0) for _ in range(4)]
        check(self.lib.fsa_get_list_stats(self.handle, *[C.byref(x) for x in v]))
        return dict(zip(("builds", "unlisted", "entries", "walks"), (x.value for x in v)))

    def health(self) -> int:
        """FSA_HEALTH_* bits: a non-finite velocity was written since the last upload."""
        v = C.c_uint()
        check(self.lib.fsa_get_health(self.handle, C.byref(v)))
        return int(v.value)

    def observe(self) -> dict:
        o = _native.FsaObservables()
        check(self.lib.fsa_observe(self.handle, C.byref(o)))
        return o.as_dict()

    def probe_pressure(self, points) -> np.ndarray:
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        out = np.zeros(len(pts), np.float32)
        check(self.lib.fsa_probe_pressure(self.handle, _ptr(pts), len(pts), _ptr(out)))
        return out

    def pair_census(self) -> dict:
        c = (C.c_int64 * 6)()
        check(self.lib.fsa_pair_census(self.handle, c))
        kinds = ("fluid", "solid", "virt")
        return {"candidates": dict(zip(kinds, c[:3])), "pairs": dict(zip(kinds, c[3:]))}

    def set_debug(self, flags: int):
        check(self.lib.fsa_set_debug(self.handle, flags))

    def set_timing(self, on: bool, stages=None):
        """Stage timing: every stage, or only the named ones
        (e.g. ("fluid",)), whose events cost the step less."""
        if stages is None:
            check(self.lib.fsa_set_timing(self.handle, 1 if on else 0))
            return
        mask = 0
        for name in stages:
            mask |= 1 << _native.STAGES.index(name)
        check(self.lib.fsa_set_timing_stages(self.handle, mask if on else 0))

    def stage_times(self) -> dict:
        ms = (C.c_double * 5)()
        n = (C.c_int64 * 5)()
        check(self.lib.fsa_get_stage_times(self.handle, ms, n))
        return {name: (ms[i], n[i]) for i, name in enumerate(_native.STAGES)}

    def reset_stage_times(self):
        check(self.lib.fsa_reset_stage_times(self.handle))

    def set_fluid_threads(self, threads: int):
        check(self.lib.fsa_set_fluid_threads(self.handle, int(threads)))

    def fluid_threads(self) -> int:
        v = C.c_int()
        check(self.lib.fsa_get_fluid_threads(self.handle, C.byref(v)))
        return int(v.value)

    def autotune_fluid(self, repeats: int = 3) -> int:
        v = C.c_int()
        check(self.lib.fsa_autotune_fluid(self.handle, int(repeats), C.byref(v)))
        return int(v.value)

    def memory_used(self) -> int:
        v = C.c_int64()
        check(self.lib.fsa_memory_used(self.handle, C.byref(v)))
        return int(v.value)


class DeviceBatch:
    """Executor that advances many contexts with one launch per stage
    (fsa_batch_* of include/fluidsim_amd.h).  The list of contexts is given on
    every call and may change between calls."""

    def __init__(self, stream: int | None = None, graphs: bool = True, autotune: bool = False,
                 streams: int | None = None):
        self.lib = _native.device_lib()
        h = C.c_void_p()
        flags = (0 if graphs else _native.FSA_BATCH_NO_GRAPH) | (_native.FSA_BATCH_AUTOTUNE if autotune else 0)
        check(self.lib.fsa_batch_create(C.c_void_p(stream) if stream else None, flags, C.byref(h)))
        self.handle = h
        if streams is not None:
            check(self.lib.fsa_batch_set_streams(self.handle, streams))

    def close(self):
        if getattr(self, "handle", None):
            self.lib.fsa_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def step(self, scenes, n: int = 1):
        arr = (C.c_void_p * max(1, len(scenes)))(*[s.handle for s in scenes])
        check(self.lib.fsa_batch_step(self.handle, arr, len(scenes), n))

    def set_streams(self, n: int):
        """Concurrent launch chains the scenes are split over (1..8)."""
        check(self.lib.fsa_batch_set_streams(self.handle, n))

    def synchronize(self):
        check(self.lib.fsa_batch_synchronize(self.handle))

    def stats(self) -> dict:
        plans, steps, threads = C.c_int64(), C.c_int64(), C.c_int()
        check(self.lib.fsa_batch_stats(self.handle, C.byref(plans), C.byref(steps), C.byref(threads)))
        tunes, tune_ms = C.c_int64(), C.c_double()
        check(self.lib.fsa_batch_tuning(self.handle, C.byref(tunes), C.byref(tune_ms)))
        return {"plans_built": int(plans.value), "steps_run": int(steps.value), "fluid_threads": int(threads.value),
                "tunes": int(tunes.value), "tune_ms": float(tune_ms.value)}


def footprint(params: FsaParams) -> tuple:
    """(resident, upload_peak) device bytes of a context for these parameters;
    allocates nothing (fsa_footprint)."""
    lib = _native.device_lib()
    r, p = C.c_int64(), C.c_int64()
    check(lib.fsa_footprint(C.byref(params), C.byref(r), C.byref(p)))
    return int(r.value), int(p.value)


def device_memory() -> tuple:
    """(free, total) bytes of the current device."""
    lib = _native.device_lib()
    f, t = C.c_int64(), C.c_int64()
    check(lib.fsa_device_memory(C.byref(f), C.byref(t)))
    return int(f.value), int(t.value)
