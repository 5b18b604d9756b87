#!/usr/bin/env python3
"""Throughput of the FluidSim step pipeline on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--scene c2|c1|tiny]
    python bench.py --workload batch [--batch-scenes 64] [--steps K]
    python bench.py --workload serve [--requests 64] [--load 0.8] [--sim-time 0.5]
    python bench.py --workload precision [--steps 1000]

One step is one pass of the reference's step loop body (src/calculation.cu:23-33:
cell hash, cell lists, solid coupling fields, walls, fluid pair+update,
solid_sub_step solid sub-steps) over the resident scene.  The default scene
is the 4,000,000 fluid + 27,000 solid dam break (dp = 1 mm, 3 wall layers,
h = 1.3 dp, dt = 0.25 h / c0, 4 solid sub-steps).  Inputs are uploaded once
before the timed region.  The default line also carries the 500-step run of
the 43,750 + 2,000 particle scene (`c1_gpu`), the 64-scene batched step
from rest (`batched`) and after 3000 batched steps (`batched_mid_solve`),
64 served requests of 0.05 s simulated time each (`serving`, p50/p99
time-to-result) and the CPU baseline.  Prints one JSON line on stdout.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from fluidsim_amd import scenes  # noqa: E402
from fluidsim_amd.metrics import field_errors, tip_displacement  # noqa: E402
from fluidsim_amd.sim import DeviceSim  # noqa: E402

HBM_PEAK_GBS = 8000.0     # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def scene_for(name: str) -> scenes.DamBreak:
    if name == "c2":
        return scenes.c2_showcase_scene()
    if name == "c1":
        return scenes.c1_cpu_scene()
    if name == "tiny":
        return scenes.tiny_scene()
    raise SystemExit(f"unknown scene {name}")


def active_walls(params, fluid_pos: np.ndarray, virt_pos: np.ndarray) -> int:
    """Wall particles inside the 27-cell neighbourhood of some fluid cell (the walls the fluid pass reads)."""
    cmin = np.array(params.cell_min, np.float32)
    inv = np.array(params.interval_inv, np.float32)
    n = np.array(params.cell_number)
    fc = np.clip(((fluid_pos - cmin) * inv).astype(np.int64), 0, n - 1)
    occ = np.zeros(tuple(n), dtype=bool)
    occ[fc[:, 0], fc[:, 1], fc[:, 2]] = True
    dil = occ.copy()
    for ax in range(3):
        d = dil.copy()
        sl_lo = [slice(None)] * 3
        sl_hi = [slice(None)] * 3
        sl_lo[ax], sl_hi[ax] = slice(0, -1), slice(1, None)
        d[tuple(sl_lo)] |= dil[tuple(sl_hi)]
        d[tuple(sl_hi)] |= dil[tuple(sl_lo)]
        dil = d
    vc = np.clip(((virt_pos - cmin) * inv).astype(np.int64), 0, n - 1)
    return int(dil[vc[:, 0], vc[:, 1], vc[:, 2]].sum())


def measured_traffic(scene: str):
    """HBM bytes per k_fluid launch from the committed rocprofv3 --pmc summary of this
    workload (FETCH_SIZE and WRITE_SIZE in KiB per dispatch, each scaled by what the
    counter reads for a plain copy of known size); None when there is none."""
    path = ROOT / "profiles" / f"k_fluid_{scene}_pmc.md"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("traffic:"):
            kv = dict(tok.split("=") for tok in line.split()[1:])
            b = (float(kv["fetch_kib"]) * float(kv["fetch_scale"]) +
                 float(kv["write_kib"]) * float(kv["write_scale"])) * 1024
            return int(b)
    return None


def sync_all(dev):
    dev.synchronize()
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def host_threads() -> int:
    """Threads for the CPU baseline: OMP_NUM_THREADS if set (the host share of
    one GPU on a shared box), else every core this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_baseline(params, arrays, budget_s: float, threads: int) -> dict:
    """The FP32 oracle (OpenMP over particles) on the same scene: three timed
    samples of whole steps after one untimed step (median reported), each
    about a third of `budget_s`; then the full 500-step run of the
    43,750 + 2,000 particle scene."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    used = oracle.set_threads(threads)
    o = oracle.Oracle(params, 32, oracle.CANONICAL)
    o.upload(arrays)
    t0 = time.perf_counter()
    o.step(1)
    first = time.perf_counter() - t0
    per = max(1, min(20, int(budget_s / 3 / max(first, 1e-3))))
    n = params.n_fluid + params.n_solid
    samples = []
    for _ in range(3):
        t1 = time.perf_counter()
        o.step(per)
        samples.append(n * per / (time.perf_counter() - t1))
    o.close()
    # the CPU-scale scene, run in full
    c1 = scenes.c1_cpu_scene(seed=0)
    a1 = c1.arrays()
    p1 = c1.params((len(a1["fluid_position"]), len(a1["solid_position"]), len(a1["virt_position"])))
    o = oracle.Oracle(p1, 32, oracle.CANONICAL)
    o.upload(a1)
    t1 = time.perf_counter()
    o.step(c1.iend)
    c1_s = time.perf_counter() - t1
    o.close()
    n1 = p1.n_fluid + p1.n_solid
    return {
        "value": float(np.median(samples)),
        "unit": "particle-steps/s",
        "cores": used,
        "kind": "port",
        "sample": f"median of 3 samples of {per} step(s) of the same scene after 1 untimed step "
                  f"({n} particles, FP32 canonical traversal, OpenMP over particles)",
        "samples": [round(v, 1) for v in samples],
        "omp_max_threads": used,
        "nproc": os.cpu_count(),
        "affinity_cpus": len(os.sched_getaffinity(0)),
        "c1_full_run": {"workload": f"{p1.n_fluid} fluid + {p1.n_solid} solid + {p1.n_virt} walls, "
                                    f"{c1.iend} steps", "seconds": round(c1_s, 3),
                        "particle_steps_per_s": n1 * c1.iend / c1_s},
    }


def batch_bench(n_scenes: int, steps: int, warmup: int, seed: int, sim_time: float, streams: int = 3,
                advance: int = 0) -> dict:
    """Many scenes of uneven size and stiffness: one batched step of all of them
    (one launch per stage, replayed as a HIP graph) against stepping them one
    after another, each on its own (the plain per-scene solve).  Scenes/hour
    is for solves of `sim_time` seconds of simulated time, from the measured
    per-step times.  advance: untimed batched steps first (both forms are then
    timed from that state of the solve)."""
    from fluidsim_amd.sim import DeviceBatch

    t = time.perf_counter()
    family = scenes.batch_scenes(n_scenes, seed=seed)
    sims, counts = [], []
    for sc in family:
        a = sc.arrays()
        c = (len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"]))
        d = DeviceSim(sc.params(c))
        d.upload(a)
        sims.append(d)
        counts.append(c)
    nf = sum(c[0] for c in counts)
    ns = sum(c[1] for c in counts)
    nv = sum(c[2] for c in counts)
    mem = sum(d.memory_used() for d in sims)
    subs = [sc.solid_sub_step for sc in family]
    log(f"batch: {n_scenes} scenes, {nf} fluid + {ns} solid + {nv} walls, S {min(subs)}..{max(subs)}, "
        f"{mem / 2**30:.2f} GiB (built in {time.perf_counter() - t:.1f}s)")
    dt = family[0].time_step
    solve_steps = math.ceil(sim_time / dt)

    b = DeviceBatch(autotune=True, streams=streams)
    if advance:
        t = time.perf_counter()
        b.step(sims, advance)
        b.synchronize()
        log(f"batch: advanced {advance} steps ({advance * family[0].time_step:.4f} s simulated) in "
            f"{time.perf_counter() - t:.1f}s")
    lists0 = [d.list_stats() for d in sims]
    b.step(sims, warmup)
    b.synchronize()
    t0 = time.perf_counter()
    b.step(sims, steps)
    b.synchronize()
    batched = (time.perf_counter() - t0) / steps
    lists1 = [d.list_stats() for d in sims]
    builds = sum(l1["builds"] - l0["builds"] for l0, l1 in zip(lists0, lists1))
    walks = sum(l1["walks"] - l0["walks"] for l0, l1 in zip(lists0, lists1))
    log(f"batch: batched {1e3 * batched:.3f} ms/step for all {n_scenes} scenes on {streams} launch chains, "
        f"fluid workgroup {b.stats()['fluid_threads']}; list builds {builds}, walking scene-steps {walks} "
        f"over {warmup + steps} steps")

    for d in sims:
        d.step(warmup)
        d.autotune_fluid(2)
    for d in sims:
        d.synchronize()
    t0 = time.perf_counter()
    for d in sims:
        d.step(steps)
        d.synchronize()  # one plain solve after another
    sequential = (time.perf_counter() - t0) / steps
    log(f"batch: sequential {1e3 * sequential:.3f} ms/step summed over the scenes")
    plans, b_threads = b.stats()["plans_built"], b.stats()["fluid_threads"]
    b.close()
    for d in sims:
        d.close()
    per_hour = lambda step_s: n_scenes * 3600.0 / (solve_steps * step_s)  # noqa: E731
    return {
        "workload": f"{n_scenes} dam breaks (seed {seed}): fluid log-uniform in [20k, 500k] on columns of "
                    "aspect ratio U[0.5, 2], gate 1-3 % of the fluid count, gate E U[5, 50] MPa, S from the solid CFL, "
                    "dp 4 mm, c0 30 m/s, compensated form (batched and sequential alike)",
        "scenes": n_scenes, "fluid": nf, "solid": ns, "walls": nv,
        "solid_sub_steps": [min(subs), max(subs)],
        "solve": f"{sim_time} s simulated = {solve_steps} steps of dt {dt:.4g} s",
        "steps_timed": steps, "warmup": warmup, "advanced_steps": advance,
        "list_builds": builds, "walking_scene_steps": walks,
        "batched_ms_per_step": round(1e3 * batched, 4),
        "sequential_ms_per_step": round(1e3 * sequential, 4),
        "scenes_per_hour_batched": round(per_hour(batched), 1),
        "scenes_per_hour_sequential": round(per_hour(sequential), 1),
        "speedup": round(sequential / batched, 3),
        "particle_steps_per_s_batched": (nf + ns) / batched,
        "plans_built": plans,
        "launch_chains": streams,
        "fluid_threads_batched": b_threads,
        "bit_identical_to_single_scene": "tests/test_gpu_batch.py",
    }


def serve_bench(n_requests: int, seed: int, sim_time: float, load: float, chunk: int,
                saturation: int | None = None) -> dict:
    """Scene-solve requests with Poisson arrivals served by continuous batching
    (fluidsim_amd.serving.Server), against the same requests solved one at a
    time in arrival order (each scene's solve time measured alone, the FIFO
    queue then computed for the same arrival times)."""
    from fluidsim_amd.serving import Request, Server, percentile
    from fluidsim_amd.sim import DeviceBatch

    rng = np.random.default_rng(seed + 7)
    t = time.perf_counter()
    family = scenes.serving_scenes(n_requests, seed=seed)
    reqs, work = [], []
    for i, sc in enumerate(family):
        a = sc.arrays()
        nf, ns = len(a["fluid_position"]), len(a["solid_position"])
        steps = math.ceil(sim_time / sc.time_step)
        reqs.append(Request(i, sc, steps, 0.0, a))
        work.append((nf + ns) * steps)
    log(f"serve: {n_requests} requests, fluid {min(len(r.arrays['fluid_position']) for r in reqs)}.."
        f"{max(len(r.arrays['fluid_position']) for r in reqs)}, {sum(work) / 1e9:.2f} G particle-steps "
        f"(built in {time.perf_counter() - t:.1f}s)")

    # each request alone: upload + per-step time over a few steps
    alone = []
    for r in reqs:
        t0 = time.perf_counter()
        d = DeviceSim(r.scene.params(tuple(len(r.arrays[k]) for k in
                                          ("fluid_position", "solid_position", "virt_position"))))
        d.upload(r.arrays)
        d.synchronize()
        up = time.perf_counter() - t0
        d.step(2)
        d.autotune_fluid(2)
        d.synchronize()
        t1 = time.perf_counter()
        d.step(8)
        d.synchronize()
        per_step = (time.perf_counter() - t1) / 8
        alone.append(up + per_step * r.steps)
        d.close()
        log(f"serve: request {r.rid} alone: upload {up:.3f}s + {r.steps} x {1e3 * per_step:.3f} ms")
    # batched throughput on the whole set, to set the arrival rate for the target load
    sims = []
    for r in reqs[: min(len(reqs), 12)]:
        d = DeviceSim(r.scene.params(tuple(len(r.arrays[k]) for k in
                                          ("fluid_position", "solid_position", "virt_position"))))
        d.upload(r.arrays)
        sims.append(d)
    b = DeviceBatch(autotune=True)
    b.step(sims, 2)
    b.synchronize()
    t1 = time.perf_counter()
    b.step(sims, 8)
    b.synchronize()
    rate = sum((d.nf + d.ns) for d in sims) * 8 / (time.perf_counter() - t1)
    b.close()
    for d in sims:
        d.close()
    mean_service = float(np.mean(work)) / rate
    lam = load / mean_service
    gaps = rng.exponential(1.0 / lam, size=n_requests)
    arrivals = np.cumsum(gaps) - gaps[0]
    for r, a in zip(reqs, arrivals):
        r.arrival = float(a)
    log(f"serve: batched rate {rate:.3g} particle-steps/s, mean service {mean_service:.2f}s, "
        f"lambda {lam:.3f}/s for load {load}")

    srv = Server(chunk_steps=chunk, saturation=saturation,
                 on_event=lambda kind, rid, t: log(f"serve: t={t:7.2f}s {kind} {rid}"))
    t0 = time.perf_counter()
    res = srv.run(reqs)
    wall = time.perf_counter() - t0
    lat = [r.latency for r in res if r.status == "done"]
    busy = sum(e - s0 for s0, e, *_ in srv.stats.chunk_log)
    done_work = sum(n * k for *_, k, n in srv.stats.chunk_log)
    log(f"serve: chunks busy {busy:.1f}s of {wall:.1f}s, {done_work / max(busy, 1e-9):.3g} particle-steps/s in chunks")
    for s0, e, na, k, n in srv.stats.chunk_log[:: max(1, len(srv.stats.chunk_log) // 25)]:
        log(f"serve: chunk t={s0:7.2f} active {na:2d} particles {n:8d} steps {k:3d}: "
            f"{1e3 * (e - s0) / k:.3f} ms/step")
    makespan = max(r.finished for r in res) - min(r.arrival for r in res)
    # one at a time in arrival order, same arrivals
    free_at, seq_lat, seq_end = 0.0, [], 0.0
    for r, s_alone in sorted(zip(reqs, alone), key=lambda x: x[0].arrival):
        start = max(free_at, r.arrival)
        free_at = start + s_alone
        seq_lat.append(free_at - r.arrival)
        seq_end = free_at
    seq_makespan = seq_end - min(r.arrival for r in reqs)
    log(f"serve: done in {wall:.1f}s; p50 {percentile(lat, 50):.2f}s p99 {percentile(lat, 99):.2f}s; "
        f"sequential p50 {percentile(seq_lat, 50):.2f}s p99 {percentile(seq_lat, 99):.2f}s")
    return {
        "workload": f"{n_requests} scene-solve requests (seed {seed}): fluid log-uniform in [10k, 2M], solid "
                    f"1-3 %, gate E U[5, 50] MPa, c0 = 10 sqrt(2 g H), {sim_time} s simulated each, "
                    f"Poisson arrivals at {load:.0%} of the batched device rate, compensated form",
        "requests": n_requests,
        "done": sum(r.status == "done" for r in res),
        "particle_steps": int(sum(work)),
        "arrival_rate_per_s": lam,
        "p50_s": percentile(lat, 50), "p99_s": percentile(lat, 99),
        "scenes_per_hour": 3600.0 * len(lat) / makespan,
        "makespan_s": makespan,
        "sequential_p50_s": percentile(seq_lat, 50), "sequential_p99_s": percentile(seq_lat, 99),
        "sequential_scenes_per_hour": 3600.0 * n_requests / seq_makespan,
        "sequential_capacity_scenes_per_hour": 3600.0 * n_requests / sum(alone),
        "max_resident_scenes": srv.stats.max_active,
        "max_resident_gib": srv.stats.max_resident_bytes / 2**30,
        "chunks": srv.stats.chunks, "chunk_steps": chunk, "plans_built": srv.stats.plans_built,
        "same_alone_and_in_mix": "tests/test_gpu_serving.py (bit-identical)",
    }


def precision_bench(steps: int, seed: int, threads: int) -> dict:
    """The precision forms on the 1,000,000 fluid + 8,000 solid dam break:
    each form advanced on the GPU, the FP64 oracle (the same method in double
    precision, OpenMP on the host) advanced over the same steps, per-field
    errors and the gate-tip displacement error of each form against it."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    base = scenes.c3_precision_scene(seed)
    a = base.arrays()
    counts = (len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"]))
    checkpoints = sorted({1, min(100, steps), steps})
    forms = [(0, "plain"), (2, "compensated"), (1, "packed"), (3, "packed+compensated")]
    gpu = {}
    for form, name in forms:
        d = DeviceSim(base.with_(precision_form=form).params(counts))
        d.upload(a)
        d.synchronize()
        states, done, secs = {}, 0, 0.0
        for c in checkpoints:
            t0 = time.perf_counter()
            d.step(c - done)
            d.synchronize()
            secs += time.perf_counter() - t0
            done = c
            states[c] = (d.fluid(), d.solid(), d.observe())
        gpu[name] = (states, 1e3 * secs / steps)
        d.close()
        log(f"precision: {name} form on the GPU, {steps} steps, {1e3 * secs / steps:.3f} ms/step")
    used = oracle.set_threads(threads)
    o = oracle.Oracle(base.params(counts), 64, oracle.CANONICAL)
    o.upload(a)
    ref, done = {}, 0
    t0 = time.perf_counter()
    for c in checkpoints:
        while done < c:
            k = min(10, c - done)
            o.step(k)
            done += k
            log(f"precision: FP64 oracle step {done}/{steps} ({time.perf_counter() - t0:.0f}s, {used} threads)")
        ref[c] = (o.fluid(), o.solid())
    o.close()
    types = np.asarray(a["solid_type"])
    X = a["solid_reference_position"]
    out = {}
    for name, (states, ms) in gpu.items():
        per = {}
        for c in checkpoints:
            f, s, obs = states[c]
            rf, rs = ref[c]
            tip_ref = tip_displacement(rs["position"], X, types, base.dp)
            tip_gpu = np.asarray(obs["gate_tip_displacement"], np.float64)
            per[str(c)] = {
                "fluid": field_errors(base, f, rf),
                "solid_position_max": float(np.abs(np.asarray(s["position"], np.float64) - rs["position"]).max()) /
                                      (max(base.tank) * base.dp),
                "gate_tip_fp64_m": [float(v) for v in tip_ref],
                "gate_tip_abs_error_m": float(np.linalg.norm(tip_gpu - tip_ref)),
                "gate_tip_rel_error": float(np.linalg.norm(tip_gpu - tip_ref) / max(np.linalg.norm(tip_ref), 1e-30)),
            }
        out[name] = {"ms_per_step": round(ms, 4), "errors": per}
        last = per[str(checkpoints[-1])]
        log(f"precision: {name}: step {checkpoints[-1]} position {last['fluid']['position']['max']:.3g} "
            f"velocity {last['fluid']['velocity']['max']:.3g} density {last['fluid']['density']['max']:.3g} "
            f"pressure {last['fluid']['pressure']['max']:.3g} gate tip {last['gate_tip_rel_error']:.3g}")
    return {
        "workload": f"dam break, {counts[0]} fluid (100^3, dp 2 mm) + {counts[1]} solid (4x50x40) + "
                    f"{counts[2]} wall particles, gate 2 dp in front of the column, {steps} steps",
        "steps": steps, "checkpoints": checkpoints, "oracle": f"FP64 CANONICAL oracle, {used} host threads",
        "forms": out,
    }


def c1_gpu_run() -> dict:
    """The CPU-scale scene's full 500-step run on the GPU, beside the CPU's."""
    c1 = scenes.c1_cpu_scene(seed=0)
    a1 = c1.arrays()
    d1 = DeviceSim(c1.params((len(a1["fluid_position"]), len(a1["solid_position"]), len(a1["virt_position"]))))
    d1.upload(a1)
    d1.step(1)
    sync_all(d1)
    t0 = time.perf_counter()
    d1.step(c1.iend)
    sync_all(d1)
    c1_s = time.perf_counter() - t0
    out = {"workload": f"{d1.nf} fluid + {d1.ns} solid + {d1.nv} walls, {c1.iend} steps (after 1 untimed step)",
           "seconds": round(c1_s, 4), "particle_steps_per_s": (d1.nf + d1.ns) * c1.iend / c1_s}
    d1.close()
    log(f"c1 scene on the GPU: {c1.iend} steps in {c1_s:.3f}s")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 1000; batch workload 30)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps first (default 20; batch workload 3)")
    ap.add_argument("--scene", default="c2")
    ap.add_argument("--advance", type=int, default=0,
                    help="step workload: untimed steps before the warm-up, to time a flow in motion "
                         "(the default times the first steps from rest, as the configuration says)")
    ap.add_argument("--stage-steps", type=int, default=100,
                    help="step workload: the first steps of the timed run, run again on a fresh upload with "
                         "every stage timed (the stage_ms breakdown; the timed steps time the fluid pass only)")
    ap.add_argument("--skin", type=float, default=None,
                    help="neighbour_skin of the step workload's scene (default: the library's; 0: walk every step)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0,
                    help="seconds of CPU oracle work on the headline scene (3 samples), before the 500-step CPU scene")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="CPU baseline threads (0: OMP_NUM_THREADS, else every core of the affinity mask)")
    ap.add_argument("--workload", choices=["step", "batch", "serve", "precision"], default="step",
                    help="step: one scene's step throughput (the headline line, plus a short batched "
                         "measurement); batch: only the many-scene batched measurement")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched measurement of the step workload")
    ap.add_argument("--no-c1", action="store_true", help="skip the 500-step run of the 43,750-particle scene")
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=30)
    ap.add_argument("--batch-seed", type=int, default=0)
    ap.add_argument("--batch-streams", type=int, default=3, help="concurrent launch chains of the batch (1..8)")
    ap.add_argument("--batch-advance", type=int, default=0,
                    help="untimed batched steps before the batch measurement (time the scenes mid-solve)")
    ap.add_argument("--batch-mid-advance", type=int, default=3000,
                    help="step workload: a second batched measurement after this many untimed batched steps "
                         "(the scenes mid-solve; 0: skip)")
    ap.add_argument("--sim-time", type=float, default=0.5, help="simulated seconds per solve (scenes/hour)")
    ap.add_argument("--requests", type=int, default=64, help="serve workload: number of requests")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement of the step workload")
    ap.add_argument("--serve-requests", type=int, default=64,
                    help="requests of the step workload's serving measurement")
    ap.add_argument("--serve-sim-time", type=float, default=0.05,
                    help="simulated seconds per request of the step workload's serving measurement "
                         "(the serve workload uses --sim-time)")
    ap.add_argument("--load", type=float, default=0.8, help="serve workload: offered load")
    ap.add_argument("--chunk", type=int, default=64, help="serve workload: steps between admission points")
    ap.add_argument("--saturation", type=int, default=0,
                    help="serve workload: particles a chunk steps together, oldest scenes first (0: all)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES (3 times): the reference kernel for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": n * 4, "dispatches": 3}))
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; --gpus must be 1 "
                                   "(a scene is one GPU's share; batched and served scenes spread over GPUs as whole scenes)"}))
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))
    if args.workload == "step":
        args.steps = 1000 if args.steps is None else args.steps
        args.warmup = 20 if args.warmup is None else args.warmup

    if args.workload == "serve":
        r = serve_bench(args.requests, args.batch_seed, args.sim_time, args.load, args.chunk,
                        args.saturation or None)
        out = {
            "metric": "p99 latency",
            "value": r["p99_s"],
            "unit": "s",
            "n_gpus": 1,
            "steps": None,
            "warmup": None,
            "ms_per_step": None,
            "higher_is_better": False,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f32",
            "data": "synthetic (seeded dam-break requests, Poisson arrivals)",
            "config": {"workload": r["workload"]},
            "serving": r,
        }
        print(json.dumps(out), flush=True)
        return

    if args.workload == "precision":
        steps = args.steps or 100
        r = precision_bench(steps, args.batch_seed, args.cpu_threads or host_threads())
        pc = r["forms"]["packed+compensated"]["errors"][str(steps)]
        out = {
            "metric": "gate-tip displacement error vs FP64 oracle (packed+compensated form)",
            "value": pc["gate_tip_rel_error"],
            "unit": "relative",
            "n_gpus": 1,
            "steps": steps,
            "warmup": 0,
            "ms_per_step": r["forms"]["packed+compensated"]["ms_per_step"],
            "higher_is_better": False,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f32 (f16 partner records)",
            "data": "synthetic (seeded dam-break lattice)",
            "config": {"workload": r["workload"]},
            "precision": r,
        }
        print(json.dumps(out), flush=True)
        return

    if args.workload == "batch":
        r = batch_bench(args.batch_scenes, args.steps or args.batch_steps,
                        3 if args.warmup is None else args.warmup, args.batch_seed, args.sim_time, args.batch_streams,
                        args.batch_advance)
        out = {
            "metric": "batched scenes/h",
            "value": r["scenes_per_hour_batched"],
            "unit": "scenes/h",
            "n_gpus": 1,
            "steps": r["steps_timed"],
            "warmup": r["warmup"],
            "ms_per_step": r["batched_ms_per_step"],
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f32",
            "data": "synthetic (seeded dam-break family, uploaded once)",
            "config": {"workload": r["workload"]},
            "batched": r,
        }
        print(json.dumps(out), flush=True)
        return

    sc = scene_for(args.scene)
    if args.skin is not None:
        sc = sc.with_(neighbour_skin=args.skin)
    t = time.perf_counter()
    arrays = sc.arrays()
    nf, ns, nv = (len(arrays["fluid_position"]), len(arrays["solid_position"]), len(arrays["virt_position"]))
    params = sc.params((nf, ns, nv))
    log(f"scene {args.scene}: {nf} fluid + {ns} solid + {nv} wall particles, grid "
        f"{list(params.cell_number)} = {params.cell_number_total} cells (built in {time.perf_counter() - t:.1f}s)")

    dev = DeviceSim(params)
    dev.upload(arrays)
    sync_all(dev)
    log(f"uploaded; device memory {dev.memory_used() / 2**20:.0f} MiB")

    if args.advance:
        t = time.perf_counter()
        dev.step(args.advance)
        sync_all(dev)
        log(f"advanced {args.advance} steps ({args.advance * sc.time_step:.4f} s simulated) in "
            f"{time.perf_counter() - t:.1f}s; max fluid speed {dev.observe()['max_fluid_speed']:.3f} m/s")
    dev.step(args.warmup)
    sync_all(dev)
    threads = dev.autotune_fluid(3)
    log(f"warmup {args.warmup} steps done; fluid workgroup size {threads} (autotuned)")

    # the timed steps bracket only the fluid pass with events (on its
    # stream): events around every stage cost the step ~2.5 %
    dev.set_timing(True, stages=("fluid",))
    dev.reset_stage_times()
    lists0 = dev.list_stats()
    sync_all(dev)
    t0 = time.perf_counter()
    dev.step(args.steps)
    sync_all(dev)
    t1 = time.perf_counter()
    fluid_ms, fluid_n = dev.stage_times()["fluid"]
    dev.set_timing(False)
    lists1 = dev.list_stats()
    secs = t1 - t0
    ms_per_step = 1e3 * secs / args.steps
    value = (nf + ns) * args.steps / secs
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step, {value:.4g} particle-steps/s; "
        f"neighbour lists {lists1}, {lists1['builds'] - lists0['builds']} builds in the timed steps; "
        f"fluid pass {fluid_ms / max(fluid_n, 1):.4f} ms avg over {fluid_n}")

    # roofline of the dominant kernel: the fused fluid pair+update pass
    fl = dev.fluid()
    n_active = active_walls(params, fl["position"], arrays["virt_position"])
    fluid_avg_s = fluid_ms / max(fluid_n, 1) / 1e3
    # the pass is VALU-bound: price it in flops too (8 per candidate test,
    # 90 per in-range Riemann pair) from an exact census of this state
    census = dev.pair_census()
    pairs = sum(census["pairs"].values())
    lists_on = params.neighbour_skin > 0 and lists1["builds"] > 0
    if lists_on:
        # with neighbour lists: per fluid particle its home state in (32 B),
        # the update out to the home and the sorted buffers (64 B), its sorted
        # slot, entry count and build position (24 B), its list (4 B per
        # candidate); fluid candidates are the listed ones
        entries = lists1["entries"]
        bytes_per_launch = (120 * nf + 4 * entries + 32 * n_active + 40 * ns
                            + 8 * params.cell_number_total)
        cand = entries + census["candidates"]["solid"] + census["candidates"]["virt"]
    else:
        # per fluid particle its state in and out (64 B) and its cell (4 B)
        bytes_per_launch = 68 * nf + 32 * n_active + 40 * ns + 8 * params.cell_number_total
        cand = sum(census["candidates"].values())
    achieved = bytes_per_launch / fluid_avg_s / 1e9
    flops = 8 * cand + 90 * pairs
    tflops = flops / fluid_avg_s / 1e12
    traffic = measured_traffic(args.scene)
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": traffic,
        "kernel": "k_fluid (fused pair sums + update)",
        "kernel_avg_ms": round(fluid_avg_s * 1e3, 5),
        "algorithmic_bytes_per_launch": int(bytes_per_launch),
        "active_walls": n_active,
        "valu": {
            "achieved_tflops": round(tflops, 3),
            "peak_tflops": FP32_PEAK_TFLOPS,
            "frac": round(tflops / FP32_PEAK_TFLOPS, 4),
            "flops_per_launch": int(flops),
            "candidates_per_fluid_particle": round(cand / nf, 2),
            "pairs_per_fluid_particle": round(pairs / nf, 2),
        },
    }

    # the stage breakdown: the same scene uploaded again and stepped from the
    # same start (warm-up, then the first steps of the timed run) with every
    # stage bracketed
    stage = {}
    if args.stage_steps > 0:
        dev.close()
        dev = DeviceSim(params)
        dev.upload(arrays)
        dev.set_fluid_threads(threads)
        dev.step(args.advance + args.warmup)
        sync_all(dev)
        dev.set_timing(True)
        dev.reset_stage_times()
        t2 = time.perf_counter()
        dev.step(args.stage_steps)
        sync_all(dev)
        t3 = time.perf_counter()
        stage = dev.stage_times()
        dev.set_timing(False)
        log(f"stage breakdown over the first {args.stage_steps} of the timed steps again, every stage timed: "
            f"{1e3 * (t3 - t2) / args.stage_steps:.3f} ms/step")
        for k, (ms, n) in stage.items():
            if n:
                log(f"  stage {k:6s}: {ms / n:8.4f} ms avg over {n}")

    result = {
        "metric": "particle-steps/s @4M+27k FSI",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded dam-break lattice, uploaded once)",
        "config": {
            "workload": f"dam break onto elastic gate, scene {args.scene}",
            "fluid": nf, "solid": ns, "walls": nv,
            "dp": sc.dp, "h_over_dp": sc.h_ratio, "solid_sub_step": sc.solid_sub_step,
            "cells": params.cell_number_total,
            "fluid_threads": threads,
            "advanced_steps": args.advance,
        },
        "stage_ms": {k: (ms / n if n else None) for k, (ms, n) in stage.items()},
        "stage_ms_steps": args.stage_steps,
        "roofline": roofline,
        "neighbour_lists": {
            "skin": round(float(params.neighbour_skin), 4),
            "builds_in_timed_steps": lists1["builds"] - lists0["builds"],
            "candidates_per_fluid_particle": round(lists1["entries"] / nf, 2),
            "unlisted": lists1["unlisted"],
        } if lists_on else None,
    }
    dev.close()
    if not args.no_c1:
        result["c1_gpu"] = c1_gpu_run()
    if not args.no_batch:
        result["batched"] = batch_bench(args.batch_scenes, args.batch_steps, 3, args.batch_seed, args.sim_time,
                                        args.batch_streams, args.batch_advance)
        if args.batch_mid_advance > 0:
            # the same family timed mid-solve (columns collapsing onto the
            # gates, every scene walking its cells), beside the from-rest line
            m = batch_bench(args.batch_scenes, args.batch_steps, 3, args.batch_seed, args.sim_time,
                            args.batch_streams, args.batch_mid_advance)
            result["batched_mid_solve"] = {k: m[k] for k in (
                "advanced_steps", "steps_timed", "batched_ms_per_step", "sequential_ms_per_step", "speedup",
                "list_builds", "walking_scene_steps", "scenes_per_hour_batched", "scenes_per_hour_sequential")}
    if not args.no_serve:
        result["serving"] = serve_bench(args.serve_requests, args.batch_seed, args.serve_sim_time, args.load,
                                        args.chunk, args.saturation or None)
    if rank == 0 and not args.no_cpu_baseline:
        threads = args.cpu_threads or host_threads()
        log(f"cpu baseline: oracle FP32 on {threads} threads (nproc {os.cpu_count()}), "
            f"~{args.cpu_budget:.0f}s budget + the 500-step CPU scene")
        cb = cpu_baseline(params, arrays, args.cpu_budget, threads)
        result["cpu_baseline"] = cb
        log(f"cpu baseline: {cb['value']:.4g} particle-steps/s (samples {cb['samples']}); "
            f"500-step CPU scene {cb['c1_full_run']['seconds']:.1f}s")
    if rank == 0:
        print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
