"""Batched stepping: many contexts advanced with one launch per stage.

Every context's state after a batched step must be bit-identical to the same
context stepped on its own with fsa_step (integer and float fields alike:
both run the same per-scene operation code, only the grid differs).  The
scenes differ in size, particle kinds present, stiffness and solid
sub-step count; the batch list changes between calls, a scene is stepped on
its own in between (which must force new recorded launches), and the step
counts are odd so both recorded steps of a plan are exercised.
"""
import numpy as np
import pytest

from fluidsim_amd import scenes
from fluidsim_amd._native import FLUID, SOLID, VIRT, FsaError
from fluidsim_amd.sim import DeviceBatch, DeviceSim

pytestmark = pytest.mark.gpu


def _variants():
    base = scenes.tiny_scene(seed=11)
    return [
        ("base", base),
        ("stiff_s6", base.with_(seed=12, youngs_modulus=8.0e6, solid_sub_step=6)),
        ("soft_s2", base.with_(seed=13, youngs_modulus=1.0e6, solid_sub_step=2, fluid=(10, 8, 6))),
        ("no_gate", base.with_(seed=14, gate=None)),
        ("no_walls", base.with_(seed=15, wall_layers=0)),
        ("no_fluid", base.with_(seed=16, fluid=(0, 0, 0))),
        ("bigger", scenes.tiny_scene(seed=17, fluid=(12, 12, 8), tank=(22, 16, 8), gate=(2, 10, 8), gate_x0=16,
                                     solid_sub_step=3)),
    ]


def _make(sc):
    a = sc.arrays()
    p = sc.params((len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"])))
    d = DeviceSim(p)
    d.upload(a)
    return d


def _state(d: DeviceSim) -> dict:
    out = {}
    if d.nf:
        out.update({"f_" + k: v for k, v in d.fluid().items()})
    if d.ns:
        out.update({"s_" + k: v for k, v in d.solid().items()})
    if d.nv:
        out.update({"v_" + k: v for k, v in d.virt().items()})
    for kind, name in ((FLUID, "fluid"), (SOLID, "solid"), (VIRT, "virt")):
        out["cells_" + name] = d.cell_ids(kind)
        out["order_" + name] = d.sorted_ids(kind)
        out["start_" + name] = d.cell_start(kind)
    out["jumps"] = np.array([d.jump_count()])
    return out


def _assert_same(name, got, want):
    assert got.keys() == want.keys()
    for k in want:
        g, w = got[k], want[k]
        assert g.shape == w.shape, (name, k)
        # bit-identical, NaN patterns included
        assert np.array_equal(g.view(np.uint32) if g.dtype == np.float32 else g,
                              w.view(np.uint32) if w.dtype == np.float32 else w), f"{name}: {k} differs"


@pytest.mark.parametrize("graphs,streams", [(True, 1), (True, 2), (False, 2), (True, 3)],
                         ids=["graph-1", "graph-2", "direct-2", "graph-3"])
def test_batched_steps_are_bit_identical_to_single_steps(gpu_available, graphs, streams):
    """The scenes are split over `streams` launch chains, each recording its
    own steps; plans_built counts the chains' recordings."""
    cases = _variants()
    alone = {n: _make(sc) for n, sc in cases}
    mixed = {n: _make(sc) for n, sc in cases}
    names = [n for n, _ in cases]
    print("case batch", "graphs" if graphs else "direct", streams, names, flush=True)
    with DeviceBatch(graphs=graphs, streams=streams) as b:
        # 1) all scenes, odd count: both recorded steps of every chain's plan run
        b.step([mixed[n] for n in names], 3)
        for n in names:
            alone[n].step(3)
        assert b.stats()["plans_built"] == streams and b.stats()["steps_run"] == 3
        # 2) same list again: the plans continue at the right recorded step
        b.step([mixed[n] for n in names], 2)
        for n in names:
            alone[n].step(2)
        assert b.stats()["plans_built"] == streams
        # 3) one scene stepped on its own: its buffers moved, its chain's plan is stale
        mixed["base"].step(1)
        alone["base"].step(1)
        b.step([mixed[n] for n in names], 1)
        for n in names:
            alone[n].step(1)
        assert b.stats()["plans_built"] == streams + 1
        # 4) continuous batching: a scene leaves, the order changes
        sub = ["bigger", "no_walls", "stiff_s6", "base", "no_fluid"]
        b.step([mixed[n] for n in sub], 3)
        for n in sub:
            alone[n].step(3)
        assert streams + 2 <= b.stats()["plans_built"] <= 2 * streams + 1
        b.synchronize()
    for n in names:
        _assert_same(n, _state(mixed[n]), _state(alone[n]))


def test_batch_of_one_matches_and_reuses_its_plan(gpu_available):
    sc = scenes.tiny_scene(seed=21)
    one, ref = _make(sc), _make(sc)
    with DeviceBatch() as b:
        for _ in range(4):
            b.step([one], 1)
        ref.step(4)
        assert b.stats()["plans_built"] == 1 and b.stats()["steps_run"] == 4
    _assert_same("single", _state(one), _state(ref))


def test_batch_stream_count_is_checked(gpu_available):
    with DeviceBatch() as b:
        for bad in (0, 9):
            with pytest.raises(FsaError, match="1..8"):
                b.set_streams(bad)


def test_batch_errors_are_status_codes(gpu_available):
    sc = scenes.tiny_scene(seed=22)
    d = _make(sc)
    a = sc.arrays()
    p = sc.params((len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"])))
    not_uploaded = DeviceSim(p)
    with DeviceBatch() as b:
        with pytest.raises(FsaError, match="listed twice"):
            b.step([d, d], 1)
        with pytest.raises(FsaError, match="uploaded"):
            b.step([d, not_uploaded], 1)
        b.step([], 1)  # nothing to do
        assert b.stats()["steps_run"] == 0


def test_fluid_workgroup_sizes_give_identical_bits(gpu_available):
    """The fluid pass at 64, 128 and 256 threads per workgroup: same bits."""
    sc = scenes.tiny_scene(seed=23, fluid=(12, 12, 8), tank=(22, 16, 8), gate=(2, 10, 8), gate_x0=16)
    states = []
    for t in (64, 128, 256):
        d = _make(sc)
        d.set_fluid_threads(t)
        assert d.fluid_threads() == t
        d.step(5)
        states.append(_state(d))
        d.close()
    _assert_same("128 vs 64", states[1], states[0])
    _assert_same("256 vs 64", states[2], states[0])
    d = _make(sc)
    with pytest.raises(FsaError, match="64, 128 or 256"):
        d.set_fluid_threads(96)


def test_autotune_keeps_the_scene_where_it_was(gpu_available):
    sc = scenes.tiny_scene(seed=24)
    tuned, ref = _make(sc), _make(sc)
    tuned.step(2)
    ref.step(2)
    before = _state(tuned)
    t = tuned.autotune_fluid(repeats=2)
    assert t in (64, 128, 256) and tuned.fluid_threads() == t
    _assert_same("autotune", _state(tuned), before)
    tuned.step(3)
    ref.step(3)
    _assert_same("after autotune", _state(tuned), _state(ref))


def test_batch_autotune_matches_single_steps(gpu_available):
    cases = _variants()[:4]
    alone = [_make(sc) for _, sc in cases]
    mixed = [_make(sc) for _, sc in cases]
    with DeviceBatch(autotune=True) as b:
        b.step(mixed, 3)
        st = b.stats()
        assert st["plans_built"] == 3 and st["fluid_threads"] in (64, 128, 256)  # three chains (the default)
    for d in alone:
        d.step(3)
    for (n, _), x, y in zip(cases, mixed, alone):
        _assert_same(n, _state(x), _state(y))
