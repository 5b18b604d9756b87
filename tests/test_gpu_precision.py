"""The precision forms on the GPU (FSA_FORM_COMPENSATED, FSA_FORM_PACKED).

Compensated form: the oracle restates the same Kahan updates (pinned on the
CPU against exact sums), so the GPU is held to the plain form's parity bar
against it (1e-5 relative, the scales of tests/helpers.py).

Packed form: the fluid pass gathers its fluid partners from FP16 records
(position relative to the partner's cell centre, rho/rho0 - 1, velocity,
p/B), so there is no FP32 restatement to match bit for bit.  Its cut-off
test still reads FP32 positions: cell ids, the sort and the pair census
must equal the plain form's exactly.  Its state is held against the FP64
oracle to bounds that follow from FP16 rounding: a partner's position
moves by up to 2^-13 cell (~1e-4 dp at 2.6 dp cells) and its velocity,
density and pressure by 2^-11 relative, i.e. ~1e-4 of each pair term, so
after a few steps of these small scenes the fields are held to 1e-4 of
their scales (10x the plain bar).  The measured errors are printed.

The batched step must give the same bits as single steps in every form, and
a batch refuses to mix packed and unpacked scenes.
"""
import numpy as np
import pytest

import oracle as orc
from fluidsim_amd import scenes
from fluidsim_amd._native import FLUID, FORM_COMPENSATED, FORM_PACKED, SOLID, FsaError
from fluidsim_amd.sim import DeviceBatch, DeviceSim

from helpers import REL, assert_within, compare_fluid, max_abs

pytestmark = pytest.mark.gpu


def _pair(sc, oracle_precision=32):
    a = sc.arrays()
    p = sc.params((len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"])))
    dev = DeviceSim(p)
    dev.upload(a)
    o = orc.Oracle(p, oracle_precision)
    o.upload(a)
    return a, p, dev, o


def _errors(sc, got, ref):
    rep = compare_fluid(sc, got, ref, rel=1.0)
    return {k: v[0] for k, v in rep.items()}


def test_compensated_form_matches_the_compensated_oracle(gpu_available):
    sc = scenes.tiny_scene(seed=21, precision_form=FORM_COMPENSATED)
    a, p, dev, o = _pair(sc)
    assert p.precision_form == FORM_COMPENSATED
    print("case compensated tiny", p.n_fluid, p.n_solid, p.n_virt, flush=True)
    for steps in (1, 20):
        dev.step(steps)
        o.step(steps)
        assert_within(compare_fluid(sc, dev.fluid(), o.fluid()), f"compensated form, +{steps} steps")
        L = max(sc.tank) * sc.dp
        assert max_abs(dev.solid()["position"], o.solid()["position"]) <= REL * L


@pytest.mark.parametrize("form", [FORM_PACKED, FORM_PACKED | FORM_COMPENSATED], ids=["packed", "packed+comp"])
def test_packed_form_keeps_neighbour_sets_and_fp16_accuracy(gpu_available, form):
    base = scenes.tiny_scene(seed=23)
    sc = base.with_(precision_form=form)
    a = sc.arrays()
    counts = (len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"]))
    print(f"case packed form {form}", counts, flush=True)
    plain = DeviceSim(base.params(counts))
    packed = DeviceSim(sc.params(counts))
    for d in (plain, packed):
        d.upload(a)
    # same cells, sort and cut-off hits as the plain form on the same state
    for kind in (FLUID, SOLID):
        np.testing.assert_array_equal(packed.cell_ids(kind), plain.cell_ids(kind))
        np.testing.assert_array_equal(packed.sorted_ids(kind), plain.sorted_ids(kind))
    assert packed.pair_census() == plain.pair_census()
    o64 = orc.Oracle(base.params(counts), 64)
    o64.upload(a)
    for steps in (1, 10):
        packed.step(steps)
        plain.step(steps)
        o64.step(steps)
        ref = o64.fluid()
        e_packed, e_plain = _errors(sc, packed.fluid(), ref), _errors(sc, plain.fluid(), ref)
        print(f"+{steps} steps: packed {e_packed} plain {e_plain}", flush=True)
        for key in ("position", "velocity", "density", "pressure"):
            assert e_packed[key] <= 1e-4, (steps, key, e_packed)
        L = max(sc.tank) * sc.dp
        assert max_abs(packed.solid()["position"], o64.solid()["position"]) <= 1e-4 * L
    for d in (plain, packed):
        d.close()


@pytest.mark.parametrize("form", [FORM_COMPENSATED, FORM_PACKED | FORM_COMPENSATED], ids=["comp", "packed+comp"])
def test_batched_forms_are_bit_identical_to_single_steps(gpu_available, form):
    from test_gpu_batch import _assert_same, _make, _state
    cases = [scenes.tiny_scene(seed=24, precision_form=form),
             scenes.tiny_scene(seed=25, precision_form=form, fluid=(10, 8, 6), solid_sub_step=3)]
    alone = [_make(sc) for sc in cases]
    mixed = [_make(sc) for sc in cases]
    with DeviceBatch() as b:
        b.step(mixed, 5)
    for d in alone:
        d.step(5)
    for i, (g, w) in enumerate(zip(mixed, alone)):
        _assert_same(f"form {form} scene {i}", _state(g), _state(w))
    for d in alone + mixed:
        d.close()


def test_family_scene_gate_tip_within_one_percent_of_fp64(gpu_available):
    """The smallest scene of the batched family (compensated form, as the batch
    and serving measurements run it), with its gate moved to 2 dp in front of
    the column so that the water loads it from the start, keeps its gate-tip
    displacement within 1 % of the FP64 oracle's over 1500 steps once the
    tip has moved 1e-3 dp; the same scene in plain FP32 is printed beside it."""
    fam = scenes.batch_scenes(64, seed=0)
    sc = min(fam, key=lambda s: s.counts()[0])
    sc = sc.with_(gate_x0=sc.fluid[0] + 2)
    a = sc.arrays()
    counts = (len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"]))
    print("case family scene", counts, "S", sc.solid_sub_step, "form", sc.precision_form, flush=True)
    assert sc.precision_form == FORM_COMPENSATED
    orc.set_threads(16)
    o64 = orc.Oracle(sc.with_(precision_form=0).params(counts), 64)
    o64.upload(a)
    devs = {}
    for form in (FORM_COMPENSATED, 0):
        d = DeviceSim(sc.with_(precision_form=form).params(counts))
        d.upload(a)
        devs[form] = d
    X = np.asarray(a["solid_reference_position"], np.float64)
    free = np.asarray(a["solid_type"]) == 0
    row = free & (X[:, 1] >= X[free, 1].max() - 0.5 * sc.dp)
    checked = 0
    for k in range(6):
        for d in devs.values():
            d.step(250)
        o64.step(250)
        ref = (o64.solid()["position"][row] - X[row]).mean(axis=0)
        got = {f: np.asarray(d.observe()["gate_tip_displacement"]) for f, d in devs.items()}
        err = {f: float(np.linalg.norm(g - ref) / np.linalg.norm(ref)) for f, g in got.items()}
        print(f"step {250 * (k + 1)}: tip fp64 {np.linalg.norm(ref):.3e} m, rel error compensated "
              f"{err[FORM_COMPENSATED]:.2e}, plain {err[0]:.2e}", flush=True)
        if np.linalg.norm(ref) > 1e-3 * sc.dp:
            assert err[FORM_COMPENSATED] <= 0.01
            checked += 1
    assert checked >= 2, "the gate tip must move during the run"
    for d in devs.values():
        d.close()


def test_batch_refuses_mixed_packing(gpu_available):
    from test_gpu_batch import _make
    ds = [_make(scenes.tiny_scene(seed=26)), _make(scenes.tiny_scene(seed=27, precision_form=FORM_PACKED))]
    with DeviceBatch() as b:
        with pytest.raises(FsaError, match="packed form"):
            b.step(ds, 1)
    for d in ds:
        d.close()


def test_cfl_bounds_from_fp64_speed_maxima(gpu_available):
    sc = scenes.tiny_scene(seed=28)
    a, p, dev, o = _pair(sc)
    dev.step(30)
    obs = dev.observe()
    f, s = dev.fluid(), dev.solid()
    umax = float(np.sqrt((f["velocity"].astype(np.float64) ** 2).sum(axis=1)).max())
    smax = float(np.sqrt((s["velocity"].astype(np.float64) ** 2).sum(axis=1)).max())
    assert umax > 0
    assert obs["max_fluid_speed"] == pytest.approx(umax, rel=1e-12)
    assert obs["max_solid_speed"] == pytest.approx(smax, rel=1e-12, abs=1e-300)
    h = float(p.smoothing_length)
    assert obs["cfl_dt"] == pytest.approx(0.25 * h / (float(p.fluid_sound_speed) + umax), rel=1e-12)
    assert obs["solid_cfl_dt"] == pytest.approx(0.3 * h / (float(p.solid_sound_speed) + smax), rel=1e-12)
    # the scene's own step is the bound at rest, 0.25 h / c0
    assert float(p.dt) >= obs["cfl_dt"] * (1 - 1e-6)
