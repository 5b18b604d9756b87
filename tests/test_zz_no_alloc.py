"""fsa_probe_pressure_noalloc: the probe entry point that allocates nothing.

It stages points and pressures in the library's static per-device area
instead of buffers it allocates, so its results must be the bits of
fsa_probe_pressure on the same state, on two calls in a row (the second
with fewer points, so a stale tail of the first call's area is never
read), from two contexts sharing the area, and it must neither grow the
context's device memory nor accept more points than the area holds.
"""
import ctypes as C

import numpy as np
import pytest

from fluidsim_amd import _native, scenes
from fluidsim_amd.sim import DeviceSim

pytestmark = pytest.mark.gpu


def _noalloc(d, pts):
    pts = np.ascontiguousarray(pts, np.float32)
    out = np.full(len(pts), np.nan, np.float32)
    fp = C.POINTER(C.c_float)
    rc = _native.device_lib().fsa_probe_pressure_noalloc(d.handle, pts.ctypes.data_as(fp), len(pts),
                                                          out.ctypes.data_as(fp))
    return rc, out


def _sim(sc, steps):
    a = sc.arrays()
    d = DeviceSim(sc.params((len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"]))))
    d.upload(a)
    d.step(steps)
    return d


def _points(sc, n, seed):
    rng = np.random.default_rng(seed)
    hi = np.array(sc.tank, np.float64) * sc.dp
    return (rng.uniform(0.0, 1.0, (n, 3)) * hi).astype(np.float32)


def test_noalloc_probe_matches_the_allocating_entry_point(gpu_available):
    sc = scenes.tiny_scene(seed=41, fluid=(16, 10, 6))
    d = _sim(sc, 6)
    print("case noalloc probe, two calls and a second context", flush=True)
    before = d.memory_used()
    first, second = _points(sc, 3000, 1), _points(sc, 217, 2)
    for pts in (first, second):
        rc, got = _noalloc(d, pts)
        assert rc == _native.FSA_OK, _native.device_lib().fsa_last_error()
        want = d.probe_pressure(pts)
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
        assert (got != 0).any()
    # the allocating entry point grew its own buffers; the noalloc one added nothing
    d2 = _sim(scenes.tiny_scene(seed=42), 3)
    m2 = d2.memory_used()
    pts = _points(scenes.tiny_scene(seed=42), 500, 3)
    rc, got = _noalloc(d2, pts)
    assert rc == _native.FSA_OK
    assert d2.memory_used() == m2
    np.testing.assert_array_equal(got.view(np.uint32), d2.probe_pressure(pts).view(np.uint32))
    # and the first context again, after the second one used the area
    rc, got = _noalloc(d, second)
    assert rc == _native.FSA_OK
    np.testing.assert_array_equal(got.view(np.uint32), d.probe_pressure(second).view(np.uint32))
    assert before <= d.memory_used()
    d.close()
    d2.close()


def test_noalloc_probe_rejects_more_points_than_its_area(gpu_available):
    sc = scenes.tiny_scene(seed=43)
    d = _sim(sc, 1)
    pts = np.tile(_points(sc, 1, 4), (_native.PROBE_NOALLOC_MAX + 1, 1))
    rc, _ = _noalloc(d, pts)
    assert rc == _native.FSA_ERR_INVALID
    assert b"at most" in _native.device_lib().fsa_last_error()
    rc, _ = _noalloc(d, pts[:_native.PROBE_NOALLOC_MAX])
    assert rc == _native.FSA_OK
    d.close()
