"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Integer work (cell ids, sort permutation, per-cell starts) must be
bit-exact.  Fluid state must agree per particle within the 1e-5 relative
tolerances of tests/helpers.py.  Solid stress-derived fields are
ill-conditioned in FP32 at small strain (the Green strain is a difference of
numbers close to 1), so they are held to "as accurate as the FP32 oracle":
|gpu - fp64| <= 4 |fp32 oracle - fp64| + 1e-5 * scale.
"""
import numpy as np
import pytest

import oracle as orc
from fluidsim_amd import scenes
from fluidsim_amd._native import FLUID, SOLID, VIRT
from fluidsim_amd.sim import DeviceBatch, DeviceSim

from helpers import REL, assert_within, compare_fluid, max_abs, same_class, velocity_scale

pytestmark = pytest.mark.gpu


def _setup(sc, precision=32, mode=orc.CANONICAL):
    a = sc.arrays()
    counts = (len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"]))
    p = sc.params(counts)
    dev = DeviceSim(p)
    dev.upload(a)
    o = orc.Oracle(p, precision, mode)
    o.upload(a)
    return a, p, dev, o


@pytest.fixture(scope="module")
def tiny():
    return scenes.tiny_scene(seed=3)


def test_cell_ids_and_sort_are_bit_exact(gpu_available, tiny):
    a, p, dev, o = _setup(tiny)
    print("case tiny seed=3", p.n_fluid, p.n_solid, p.n_virt, flush=True)
    dev.particle_cell_update()
    dev.fill_cell()
    o.particle_cell_update()
    o.fill_cell()
    for kind in (FLUID, SOLID, VIRT):
        np.testing.assert_array_equal(dev.cell_ids(kind), o.cell_ids(kind))
        np.testing.assert_array_equal(dev.cell_start(kind), o.cell_start(kind))
        np.testing.assert_array_equal(dev.sorted_ids(kind), o.sorted_ids(kind))


def test_one_step_stage_by_stage(gpu_available, tiny):
    a, p, dev, o = _setup(tiny)
    dev.set_debug(1)
    o64 = orc.Oracle(p, 64)
    o64.upload(a)
    for s in (dev, o, o64):
        s.particle_cell_update()
        s.fill_cell()
        s.couple_calculation()
        s.virt_update()
    # wall state; normals are unit vectors normalised from sums whose terms
    # cancel, so they are held to 1e-5 on each component
    gv, ov = dev.virt(), o.virt()
    np.testing.assert_allclose(gv["normal"], ov["normal"], atol=REL)
    np.testing.assert_array_equal(gv["density"], ov["density"].astype(np.float32))
    assert max_abs(gv["velocity"], ov["velocity"]) <= REL * velocity_scale(tiny, ov["velocity"])
    # solid coupling fields
    gs, os_ = dev.solid(), o.solid()
    np.testing.assert_allclose(gs["normal"], os_["normal"], atol=REL)
    np.testing.assert_allclose(gs["pressure"], os_["pressure"], atol=1e-3, rtol=REL)
    # fluid pass
    for s in (dev, o, o64):
        s.fluid_calculation()
    gf, of = dev.fluid(), o.fluid()
    assert_within(compare_fluid(tiny, gf, of), "fluid after one fluid stage")
    rates = dev.fluid_rates()
    # pair sums: errors relative to the largest sum (the terms cancel)
    amax = np.abs(of["du_dt"]).max()
    assert max_abs(rates["du_dt"], of["du_dt"]) <= 1e-4 * amax
    dmax = np.abs(of["drho_dt"]).max()
    assert max_abs(rates["drho_dt"], of["drho_dt"]) <= 1e-4 * max(dmax, 1e-6)
    # solid sub-steps
    for s in (dev, o, o64):
        s.solid_calculation()
    gs, os_, os64 = dev.solid(), o.solid(), o64.solid()
    L = max(tiny.tank) * tiny.dp
    assert max_abs(gs["position"], os_["position"]) <= REL * L
    for key in ("deformation", "correction"):
        e_g, e_32, ok = same_class(gs[key], os_[key], os64[key], 1.0)
        assert ok, (key, e_g, e_32)
    amax = np.abs(os64["du_dt"]).max()
    e_g, e_32, ok = same_class(gs["du_dt"], os_["du_dt"], os64["du_dt"], amax)
    assert ok, ("du_dt", e_g, e_32)


@pytest.mark.parametrize("h_ratio", [1.3, 1.9, 2.6])
def test_solid_partner_records(gpu_available, h_ratio):
    """The first solid sub-step of a step records each lane's reference
    partners (at most 16 per lane) and the later deformation and force passes
    replay them; a lane with more partners walks its candidates again.  A
    thick gate at growing smoothing lengths covers both paths: ~5 partners per
    lane at h = 1.3 dp (at most 11), a few lanes over the cap at 1.9 dp, a
    tenth of them at 2.6 dp."""
    sc = scenes.tiny_scene(seed=5, gate=(6, 8, 6), gate_x0=9, h_ratio=h_ratio, solid_sub_step=3)
    a, p, dev, o = _setup(sc)
    o64 = orc.Oracle(p, 64)
    o64.upload(a)
    print(f"case thick gate h_ratio={h_ratio}", p.n_fluid, p.n_solid, p.n_virt, flush=True)
    for st in range(3):
        for s in (dev, o, o64):
            s.step(1)
        gs, os_, os64 = dev.solid(), o.solid(), o64.solid()
        L = max(sc.tank) * sc.dp
        assert max_abs(gs["position"], os_["position"]) <= REL * L, st
        for key in ("deformation", "correction"):
            e_g, e_32, ok = same_class(gs[key], os_[key], os64[key], 1.0)
            assert ok, (st, key, e_g, e_32)
        amax = np.abs(os64["du_dt"]).max()
        e_g, e_32, ok = same_class(gs["du_dt"], os_["du_dt"], os64["du_dt"], amax)
        assert ok, (st, "du_dt", e_g, e_32)


def test_replayed_partners_give_the_walk_bit_for_bit(gpu_available):
    """Replaying the recorded solid and fluid partners of a step gives the
    same bits as walking the candidate cells in every solid pass
    (FSA_DEBUG_WALK), over steps in which the fluid surrounds the gate."""
    from fluidsim_amd._native import DEBUG_WALK
    sc = scenes.tiny_scene(seed=9, fluid=(16, 10, 6), gate=(4, 8, 6), gate_x0=9, solid_sub_step=4)
    a = sc.arrays()
    p = sc.params((len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"])))
    print("case walk vs replay", p.n_fluid, p.n_solid, p.n_virt, flush=True)
    devs = []
    for flags in (0, DEBUG_WALK):
        d = DeviceSim(p)
        d.upload(a)
        d.set_debug(flags)
        devs.append(d)
    for st in range(4):
        for d in devs:
            d.step(2)
        s0, s1 = devs[0].solid(), devs[1].solid()
        for key in ("position", "velocity", "du_dt", "couple_dudt", "deformation"):
            np.testing.assert_array_equal(s0[key], s1[key], err_msg=f"{key} after {2 * (st + 1)} steps")
        f0, f1 = devs[0].fluid(), devs[1].fluid()
        np.testing.assert_array_equal(f0["velocity"], f1["velocity"])
    assert np.abs(devs[0].solid()["couple_dudt"]).max() > 0, "the fluid must press on the gate"
    for d in devs:
        d.close()


def test_replay_holds_for_fast_fluid_and_a_moving_gate(gpu_available):
    """The coupling pass records fluid candidates from the positions before
    the fluid update; the force passes replay them after it.  Fluid driven
    at the gate faster than the speed cap (first step, uploaded 0.5 c0) and
    at the cap afterwards moves up to |u| dt in between, and a gate pushed
    at 0.45 c0 leaves the 0.04 R replay zone within a step, so its later
    sub-steps walk.  Replay must still give the walk's bits."""
    from fluidsim_amd._native import DEBUG_WALK
    sc = scenes.tiny_scene(seed=13, fluid=(16, 10, 6), gate=(4, 8, 6), gate_x0=11, solid_sub_step=4)
    a = sc.arrays()
    c0 = sc.sound_speed
    a["fluid_velocity"][:, 0] = 0.5 * c0
    a["fluid_velocity"][:, 1] = -0.05 * c0
    free = a["solid_type"] == 0
    a["solid_velocity"][free, 0] = -0.45 * c0
    p = sc.params((len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"])))
    print("case fast fluid and moving gate, walk vs replay", p.n_fluid, p.n_solid, p.n_virt, flush=True)
    devs = []
    for flags in (0, DEBUG_WALK):
        d = DeviceSim(p)
        d.upload(a)
        d.set_debug(flags)
        devs.append(d)
    for st in range(6):
        for d in devs:
            d.step(1)
        s0, s1 = devs[0].solid(), devs[1].solid()
        for key in ("position", "velocity", "du_dt", "couple_dudt", "deformation"):
            np.testing.assert_array_equal(s0[key], s1[key], err_msg=f"{key} after {st + 1} steps")
        np.testing.assert_array_equal(devs[0].fluid()["velocity"], devs[1].fluid()["velocity"])
    assert np.abs(devs[0].solid()["couple_dudt"]).max() > 0, "the fluid must press on the gate"
    for d in devs:
        d.close()


@pytest.mark.parametrize("steps", [1, 20])
def test_multi_step_state(gpu_available, tiny, steps):
    a, p, dev, o = _setup(tiny)
    print(f"case tiny steps={steps}", flush=True)
    dev.step(steps)
    o.step(steps)
    assert_within(compare_fluid(tiny, dev.fluid(), o.fluid()), f"fluid after {steps} steps")
    gs, os_ = dev.solid(), o.solid()
    L = max(tiny.tank) * tiny.dp
    assert max_abs(gs["position"], os_["position"]) <= REL * L
    assert dev.jump_count() == 0


# ---- golden fixtures ---------------------------------------------------------

@pytest.mark.parametrize("name", ["tiny_seed7", "tiny_inviscid_seed8"])
def test_gpu_matches_golden_fixtures(gpu_available, name):
    from test_golden import inputs, load, state
    sc, d = load(name)
    a = inputs(d)
    p = sc.params((len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"])))
    dev = DeviceSim(p)
    dev.upload(a)
    done = 0
    for s in (1, 10):
        print(f"case golden {name} step {s}", flush=True)
        dev.step(s - done)
        done = s
        assert_within(compare_fluid(sc, dev.fluid(), state(d, s, "fluid")), f"{name} step {s}")
        L = max(sc.tank) * sc.dp
        assert max_abs(dev.solid()["position"], state(d, s, "solid")["position"]) <= REL * L


# ---- the CPU-scale scene: 43,750 fluid + 2,000 solid --------------------------

@pytest.fixture(scope="module")
def c1_pair():
    sc = scenes.c1_cpu_scene(seed=2)
    a = sc.arrays()
    p = sc.params((len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"])))
    orc.set_threads(16)
    return sc, a, p


@pytest.mark.parametrize("steps", [1, 100])
def test_c1_scene_parity(gpu_available, c1_pair, steps):
    sc, a, p = c1_pair
    print(f"case c1 steps={steps} fluid={p.n_fluid} solid={p.n_solid} walls={p.n_virt}", flush=True)
    dev = DeviceSim(p)
    dev.upload(a)
    o = orc.Oracle(p, 32)
    o.upload(a)
    o64 = orc.Oracle(p, 64)
    o64.upload(a)
    dev.step(steps)
    o.step(steps)
    o64.step(steps)
    rep = compare_fluid(sc, dev.fluid(), o.fluid())
    print({k: v[0] for k, v in rep.items()}, flush=True)
    assert_within(rep, f"C1 fluid after {steps} steps")
    # gate positions: TLSPH in FP32 amplifies rounding (the Green strain is a
    # difference of numbers near 1), so the GPU is held to the accuracy of the
    # FP32 oracle against FP64, plus 1e-5 of the tank size
    L = max(sc.tank) * sc.dp
    e_gpu, e_32, ok = same_class(dev.solid()["position"], o.solid()["position"], o64.solid()["position"], L)
    print(f"gate position error vs FP64: gpu {e_gpu:.3g}, fp32 oracle {e_32:.3g}", flush=True)
    assert ok, (e_gpu, e_32)
    for kind in (FLUID, SOLID):
        np.testing.assert_array_equal(dev.cell_ids(kind), o.cell_ids(kind))


# ---- full size: 4,000,000 fluid + 27,000 solid, size-independent checks ------

def test_showcase_scale_integer_invariants(gpu_available):
    sc = scenes.c2_showcase_scene()
    a = sc.arrays()
    p = sc.params((len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"])))
    print(f"case c2 fluid={p.n_fluid} solid={p.n_solid} walls={p.n_virt}", flush=True)
    dev = DeviceSim(p)
    dev.upload(a)
    dev.step(10)
    dev.particle_cell_update()  # hash the post-step positions, as the next step would
    dev.fill_cell()
    f = dev.fluid()
    cmin = np.array(p.cell_min, np.float32)
    inv = np.array(p.interval_inv, np.float32)
    n = np.array(p.cell_number)
    for kind, x in ((FLUID, f["position"]), (SOLID, dev.solid()["position"])):
        idx = ((x.astype(np.float32) - cmin) * inv).astype(np.int64)
        expect = (idx[:, 0] * n[1] + idx[:, 1]) * n[2] + idx[:, 2]
        cells = dev.cell_ids(kind)
        np.testing.assert_array_equal(cells, expect)
        # slots: by cell, then (fluid and walls) eighth-cell z index inside the cell, then .dat index
        tz = (x[:, 2].astype(np.float32) - cmin[2]) * inv[2]
        sub = np.clip((tz * np.float32(8)).astype(np.int64) - 8 * idx[:, 2], 0, 7) if kind == FLUID else 0 * idx[:, 2]
        order = dev.sorted_ids(kind)
        np.testing.assert_array_equal(order, np.lexsort((np.arange(len(cells)), sub, cells)))
        start = dev.cell_start(kind)
        np.testing.assert_array_equal(np.diff(start), np.bincount(cells, minlength=p.cell_number_total))
    assert np.isfinite(f["velocity"]).all() and np.isfinite(f["pressure"]).all()
    assert f["density"].min() >= p.fluid_min_density and f["density"].max() <= p.fluid_max_density
    lo, hi = np.array(p.domain_min), np.array(p.domain_max)
    assert (f["position"] >= lo).all() and (f["position"] <= hi).all()
    assert dev.jump_count() == 0


# ---- edge cases --------------------------------------------------------------

@pytest.mark.parametrize("variant", ["no_gate", "no_walls", "no_fluid", "single_fluid"])
def test_missing_kinds(gpu_available, variant):
    from helpers import custom_case
    base = scenes.tiny_scene(seed=4)
    if variant == "no_gate":
        sc = base.with_(gate=None)
        a = sc.arrays()
    elif variant == "no_walls":
        sc = base.with_(wall_layers=0)
        a = sc.arrays()
    elif variant == "no_fluid":
        sc = base.with_(fluid=(0, 0, 0))
        a = sc.arrays()
    else:
        sc = base
        _, a = custom_case(sc, fluid={"position": [[0.05, 0.05, 0.03]], "velocity": [[0.1, 0.0, 0.0]]})
    p = sc.params((len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"])))
    print(f"case {variant}: {p.n_fluid} {p.n_solid} {p.n_virt}", flush=True)
    dev = DeviceSim(p)
    dev.upload(a)
    o = orc.Oracle(p, 32)
    o.upload(a)
    dev.step(3)
    o.step(3)
    if p.n_fluid:
        assert_within(compare_fluid(sc, dev.fluid(), o.fluid()), variant)
    if p.n_solid:
        L = max(sc.tank) * sc.dp
        assert max_abs(dev.solid()["position"], o.solid()["position"]) <= REL * L


def test_dense_cells_long_runs(gpu_available):
    """h = 7 dp: cells 18 dp wide (2 per axis), ~6,500 fluid slots in a
    column's z-window and ~11,000 partners per particle.  The fluid pass then
    splits a run at 4096 slots (the reach of a 16-bit list entry) and
    evaluates a full list hundreds of times per particle; solid particles
    overflow their recorded-partner lists.  Fluid: the usual tolerances; the
    gate: as accurate as the FP32 oracle against FP64 (this coarse, stiff
    gate is badly conditioned in FP32)."""
    sc = scenes.DamBreak(dp=0.01, fluid=(20, 20, 20), tank=(20, 20, 20), gate=(2, 8, 6), gate_x0=12,
                         gate_clamp_rows=2, wall_layers=1, h_ratio=7.0, sound_speed=20.0, youngs_modulus=2.0e6,
                         seed=11)
    a, p, dev, o = _setup(sc)
    assert tuple(p.cell_number) == (2, 2, 2)
    print(f"dense cells: {p.n_fluid} fluid, {p.n_solid} solid, {p.n_virt} walls", flush=True)
    o64 = orc.Oracle(p, 64)
    o64.upload(a)
    dev.step(2)
    o.step(2)
    o64.step(2)
    assert_within(compare_fluid(sc, dev.fluid(), o.fluid()), "dense cells")
    L = max(sc.tank) * sc.dp
    e_gpu, e_32, ok = same_class(dev.solid()["position"], o.solid()["position"], o64.solid()["position"], L)
    print(f"gate position error vs FP64: gpu {e_gpu:.3g}, fp32 oracle {e_32:.3g}", flush=True)
    assert ok, (e_gpu, e_32)


@pytest.mark.parametrize("scene,steps,form", [("tiny", 12, 0), ("c1", 20, 0), ("c1", 20, 3)])
def test_overlapped_steps_match_the_stage_sequence(gpu_available, scene, steps, form):
    """fsa_step runs the next step's fluid cell build and wall pass beside the
    solid sub-steps; the six stage calls made one by one give the same bits."""
    sc = (scenes.tiny_scene(seed=5) if scene == "tiny" else scenes.c1_cpu_scene(seed=2)).with_(precision_form=form)
    a = sc.arrays()
    p = sc.params((len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"])))
    print(f"case {scene} form {form}: {p.n_fluid} fluid, {p.n_solid} solid, {p.n_virt} walls, {steps} steps",
          flush=True)
    fused = DeviceSim(p)
    fused.upload(a)
    fused.step(steps)
    staged = DeviceSim(p)
    staged.upload(a)
    for _ in range(steps):
        staged.particle_cell_update()
        staged.fill_cell()
        staged.couple_calculation()
        staged.virt_update()
        staged.fluid_calculation()
        staged.solid_calculation()
    for got, want in ((fused.fluid(), staged.fluid()), (fused.solid(), staged.solid()), (fused.virt(), staged.virt())):
        for key in want:
            np.testing.assert_array_equal(got[key], want[key], err_msg=key)
    for kind in (FLUID, SOLID):
        np.testing.assert_array_equal(fused.sorted_ids(kind), staged.sorted_ids(kind))
        np.testing.assert_array_equal(fused.cell_start(kind), staged.cell_start(kind))
    assert fused.jump_count() == staged.jump_count()


def _same_state(got: DeviceSim, want: DeviceSim):
    for g, w in ((got.fluid(), want.fluid()), (got.solid(), want.solid()), (got.virt(), want.virt())):
        for key in w:
            np.testing.assert_array_equal(g[key], w[key], err_msg=key)
    for kind in (FLUID, SOLID):
        np.testing.assert_array_equal(got.sorted_ids(kind), want.sorted_ids(kind))
        np.testing.assert_array_equal(got.cell_start(kind), want.cell_start(kind))


@pytest.mark.parametrize("then", ["hash_again", "upload", "batch"])
def test_pending_hash_leaves_no_counts_behind(gpu_available, tiny, then):
    """The per-key counts are zero between cell builds (every sort's scan
    clears them).  A stage-call hash without its sort leaves counts behind:
    hashing again, re-uploading or stepping in a batch must still give the
    bits of a clean run."""
    a = tiny.arrays()
    p = tiny.params((len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"])))
    clean = DeviceSim(p)
    clean.upload(a)
    clean.step(4)
    dev = DeviceSim(p)
    dev.upload(a)
    dev.particle_cell_update()
    if then == "hash_again":
        dev.particle_cell_update()
        dev.fill_cell()
        dev.couple_calculation()
        dev.virt_update()
        dev.fluid_calculation()
        dev.solid_calculation()
        dev.step(3)
    elif then == "upload":
        dev.upload(a)
        dev.step(4)
    else:
        b = DeviceBatch()
        b.step([dev], 4)
        b.synchronize()
    _same_state(dev, clean)


def test_teleported_particle_is_kept_and_counted(gpu_available, tiny):
    a = tiny.arrays()
    p = tiny.params((len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"])))
    a["fluid_velocity"][0] = [3.0 * p.interval[0] / p.dt, 0.0, 0.0]
    dev = DeviceSim(p)
    dev.upload(a)
    o = orc.Oracle(p, 32)
    o.upload(a)
    dev.step(2)
    o.step(2)
    assert dev.jump_count() == 1
    assert_within(compare_fluid(tiny, dev.fluid(), o.fluid()), "after a jump")
    ids = dev.sorted_ids(FLUID)
    assert sorted(ids.tolist()) == list(range(p.n_fluid))


def test_nan_pressure_is_scrubbed_like_the_oracle(gpu_available, tiny):
    a = tiny.arrays()
    p = tiny.params((len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"])))
    a["fluid_pressure"][5] = np.nan
    dev = DeviceSim(p)
    dev.upload(a)
    o = orc.Oracle(p, 32)
    o.upload(a)
    dev.particle_cell_update(); dev.fill_cell(); dev.couple_calculation(); dev.virt_update()
    dev.fluid_calculation()
    o.particle_cell_update(); o.fill_cell(); o.couple_calculation(); o.virt_update()
    o.fluid_calculation()
    g, r = dev.fluid(), o.fluid()
    assert np.isfinite(g["position"]).all() and np.isfinite(g["density"]).all()
    assert_within(compare_fluid(tiny, g, r), "NaN neighbour")


def test_errors_are_status_codes(gpu_available, tiny):
    from fluidsim_amd._native import FsaError, FSA_ERR_STATE, FSA_ERR_INVALID
    a = tiny.arrays()
    p = tiny.params((len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"])))
    dev = DeviceSim(p)
    with pytest.raises(FsaError) as e:
        dev.fluid_calculation()  # nothing uploaded
    assert e.value.code == FSA_ERR_STATE
    dev.upload(a)
    with pytest.raises(FsaError) as e:
        dev.fill_cell()  # no hash of the current order
    assert e.value.code == FSA_ERR_STATE
    bad = {k: v.copy() for k, v in a.items()}
    bad["fluid_position"][3] = [-1.0, 0.0, 0.0]
    with pytest.raises(FsaError) as e:
        dev.upload(bad)
    assert e.value.code == FSA_ERR_INVALID
    dev.upload(a)  # a good upload resets the context
    dev.step(1)
    dev.set_timing(True)
    dev.step(2)
    t = dev.stage_times()
    assert t["fluid"][1] == 2 and t["fluid"][0] > 0
    # timing the fluid pass alone: only its samples grow
    dev.set_timing(True, stages=("fluid",))
    dev.reset_stage_times()
    dev.step(3)
    t = dev.stage_times()
    assert t["fluid"][1] == 3 and t["fluid"][0] > 0
    assert all(n == 0 for k, (ms, n) in t.items() if k != "fluid"), t
    dev.set_timing(False)


def test_wall_pass_follows_fluid_coming_and_going(gpu_available):
    """The wall pass leaves a wall with no fluid in its 27 cells alone once its
    far-field values (rho0, p = 0, u = 0) stand.  With the fluid thrown in all
    directions walls gain and lose fluid in reach from step to step; after
    every step's wall pass the walls must be the oracle's wall pass over the
    device's own fluid state: density bit for bit, velocity and pressure
    within the tolerances, and walls must change sides both ways."""
    sc = scenes.tiny_scene(seed=41)
    a = sc.arrays()
    rng = np.random.default_rng(41)
    n = len(a["fluid_position"])
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    speed = rng.uniform(0.0, 0.9 * 0.3 * sc.sound_speed, size=(n, 1))
    a["fluid_velocity"] = np.ascontiguousarray((d * speed).astype(np.float32))
    p = sc.params((n, len(a["solid_position"]), len(a["virt_position"])))
    print(f"case thrown fluid, walls: {p.n_fluid} fluid, {p.n_virt} walls", flush=True)
    dev = DeviceSim(p)
    dev.upload(a)
    rho0 = np.float32(p.fluid_reference_density)
    prev, gained, lost = None, 0, 0
    for k in range(24):
        dev.particle_cell_update()
        dev.fill_cell()
        dev.couple_calculation()
        dev.virt_update()
        gv = dev.virt()
        b = dict(a)
        f = dev.fluid()
        b["fluid_position"], b["fluid_velocity"] = f["position"], f["velocity"]
        b["fluid_density"], b["fluid_pressure"] = f["density"], f["pressure"]
        s = dev.solid()
        b["solid_position"], b["solid_velocity"] = s["position"], s["velocity"]
        o = orc.Oracle(p, 32)
        o.upload(b)
        o.particle_cell_update()
        o.fill_cell()
        o.virt_update()
        ov = o.virt()
        o.close()
        np.testing.assert_array_equal(gv["density"], ov["density"].astype(np.float32), err_msg=f"step {k}")
        assert max_abs(gv["velocity"], ov["velocity"]) <= REL * velocity_scale(sc, ov["velocity"]), k
        pmax = max(float(np.abs(ov["pressure"]).max()), 1.0)
        assert max_abs(gv["pressure"], ov["pressure"]) <= REL * pmax, k
        far = (gv["density"] == rho0) & np.all(gv["velocity"] == 0.0, axis=1) & (gv["pressure"] == 0.0)
        if prev is not None:
            gained += int(np.sum(prev & ~far))
            lost += int(np.sum(~prev & far))
        prev = far
        dev.fluid_calculation()
        dev.solid_calculation()
    print(f"walls that gained fluid in reach {gained}, lost it {lost}", flush=True)
    assert gained > 0 and lost > 0
    dev.close()
