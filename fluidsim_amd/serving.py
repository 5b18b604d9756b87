"""Scene-solve serving with continuous batching.

Requests (a scene and a number of steps) arrive over time.  A request is
admitted when its device footprint fits the memory budget and a scene slot
is free; its state then stays resident on the device until it is done.  All
admitted scenes advance together, a chunk of steps at a time, through one
batch executor (one launch per stage for all of them, fsa_batch_step), and
the list of scenes changes between chunks: finished scenes leave, newly
admitted ones join.  Admission work (building inputs, creating the context,
uploading) overlaps the chunk in flight on the device, since each context
uploads on its own stream.

A scene's result does not depend on which other scenes shared its chunks:
a batched step is bit-identical to stepping the scene on its own.  A scene
whose state diverges (a non-finite velocity, fsa_get_health) is evicted
after the chunk in which it happened, with status "diverged", so that it
does not slow its neighbours (a torn-apart gate crowds its particles into a
few cells, where the per-cell work grows with the square of the count).

The reference has no counterpart (it runs one scene per process,
src/main.cu); the step each chunk repeats is its step loop body
(src/calculation.cu:23-33).
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

from .scenes import DamBreak
from .sim import DeviceBatch, DeviceSim, device_memory, footprint


@dataclass
class Request:
    rid: int
    scene: DamBreak
    steps: int
    arrival: float = 0.0        # clock time at which the request arrives
    arrays: dict | None = None  # prepared inputs; built on admission when None


@dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    finished: float
    steps: int
    counts: tuple
    observables: dict
    state: dict | None = None
    status: str = "done"        # or "rejected" (larger than the whole budget), or "diverged"
    health: int = 0             # FSA_HEALTH_* bits of a diverged scene

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


@dataclass
class _Active:
    req: Request
    sim: DeviceSim
    remaining: int
    resident: int
    admitted: float
    counts: tuple


@dataclass
class ServerStats:
    chunks: int = 0
    steps: int = 0
    scene_steps: int = 0
    max_resident_bytes: int = 0
    max_active: int = 0
    plans_built: int = 0
    tunes: int = 0          # autotuner timings (FSA_BATCH_AUTOTUNE)
    tune_ms: float = 0.0    # their host wall time
    diverged: int = 0
    events: list = field(default_factory=list)
    chunk_log: list = field(default_factory=list)  # (start, end, active scenes, steps, particles)


class Server:
    """Admission by footprint + continuous batching over fsa_batch_step.

    memory_budget: device bytes the resident scenes may hold (default 90 % of
    the free memory when the server starts).  max_scenes: scene slots.
    chunk_steps: most steps between two admission points.  saturation: the
    particles (fluid + solid) a chunk steps together: the resident scenes in
    arrival order, each taken while it fits (the oldest always), so that a
    chunk fills the device and the oldest requests are not slowed by newer
    ones beyond that (None: every resident scene).  The clock of run() is a
    callable giving the current time with a wait_until(t) method (default
    WallClock); StepClock makes a run reproducible."""

    def __init__(self, memory_budget: int | None = None, max_scenes: int = 64, chunk_steps: int = 32,
                 graphs: bool = True, autotune: bool = True, keep_state: bool = False, on_event=None,
                 saturation: int | None = None):
        if memory_budget is None:
            free, _ = device_memory()
            memory_budget = int(0.9 * free)
        self.memory_budget = memory_budget
        self.max_scenes = max_scenes
        self.chunk_steps = chunk_steps
        self.graphs = graphs
        self.autotune = autotune  # time the fluid workgroup sizes whenever the batch is re-recorded
        self.keep_state = keep_state
        self.saturation = saturation
        self.on_event = on_event  # on_event(kind, request id, time), kind "admit", "finish" or "diverged"
        self.stats = ServerStats()

    def _event(self, kind: str, rid: int, t: float):
        self.stats.events.append((kind, rid, t))
        if self.on_event is not None:
            self.on_event(kind, rid, t)

    # ---- one request ------------------------------------------------------------
    @staticmethod
    def _inputs(req: Request):
        a = req.arrays if req.arrays is not None else req.scene.arrays()
        counts = (len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"]))
        return a, counts, req.scene.params(counts)

    def _finish(self, act: _Active, now: float) -> Result:
        obs = act.sim.observe()
        state = None
        if self.keep_state:
            state = {"fluid": act.sim.fluid(), "solid": act.sim.solid() if act.counts[1] else None}
        act.sim.close()
        return Result(act.req.rid, act.req.arrival, act.admitted, now, act.req.steps, act.counts, obs, state)

    def _chunk_scenes(self, active: list) -> list:
        """The resident scenes one chunk steps: in arrival order, each while
        the particles fit `saturation` (the oldest always)."""
        if self.saturation is None:
            return list(active)
        chosen, total = [], 0
        for x in active:
            n = x.counts[0] + x.counts[1]
            if chosen and total + n > self.saturation:
                continue
            chosen.append(x)
            total += n
        return chosen

    # ---- the loop -------------------------------------------------------------------
    def run(self, requests, clock=None) -> list:
        if clock is None:
            clock = WallClock()
        pending = deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        queue: deque = deque()   # arrived, waiting for memory or a slot: (req, arrays, counts, params, fp)
        active: list = []
        results: list = []
        used = 0
        st = self.stats
        batch = DeviceBatch(graphs=self.graphs, autotune=self.autotune)
        try:
            while pending or queue or active:
                now = clock()
                while pending and pending[0].arrival <= now:
                    req = pending.popleft()
                    a, counts, params = self._inputs(req)
                    fp = footprint(params)
                    queue.append((req, a, counts, params, fp))
                # admission, first come first served
                while queue and len(active) < self.max_scenes:
                    req, a, counts, params, (resident, peak) = queue[0]
                    if peak > self.memory_budget:
                        queue.popleft()
                        results.append(Result(req.rid, req.arrival, now, now, 0, counts, {}, None, "rejected"))
                        continue
                    if used + peak > self.memory_budget:
                        break
                    queue.popleft()
                    sim = DeviceSim(params)
                    sim.upload(a)
                    used += resident
                    active.append(_Active(req, sim, req.steps, resident, clock(), counts))
                    self._event("admit", req.rid, clock())
                st.max_resident_bytes = max(st.max_resident_bytes, used)
                st.max_active = max(st.max_active, len(active))
                if not active:
                    if queue:
                        raise RuntimeError("a queued request fits the budget but nothing is active")
                    if pending:
                        clock.wait_until(pending[0].arrival)
                    continue
                stepped = self._chunk_scenes(active)
                k = min(self.chunk_steps, min(x.remaining for x in stepped))
                sims = [x.sim for x in stepped]
                t_chunk = clock()
                batch.step(sims, k)
                # arrivals during the chunk are prepared while the device works
                now = clock()
                while pending and pending[0].arrival <= now:
                    req = pending.popleft()
                    a, counts, params = self._inputs(req)
                    queue.append((req, a, counts, params, footprint(params)))
                batch.synchronize()
                st.chunk_log.append((t_chunk, clock(), len(stepped), k,
                                     sum(x.counts[0] + x.counts[1] for x in stepped)))
                st.chunks += 1
                st.steps += k
                st.scene_steps += k * len(stepped)
                now = clock()
                still = []
                for x in active:
                    if any(x is y for y in stepped):
                        x.remaining -= k
                        bad = x.sim.health()
                        if bad:
                            x.sim.close()
                            results.append(Result(x.req.rid, x.req.arrival, x.admitted, now, x.req.steps - x.remaining,
                                                  x.counts, {}, None, "diverged", bad))
                            used -= x.resident
                            st.diverged += 1
                            self._event("diverged", x.req.rid, now)
                            continue
                    if x.remaining == 0:
                        results.append(self._finish(x, now))
                        used -= x.resident
                        self._event("finish", x.req.rid, now)
                    else:
                        still.append(x)
                active = still
            bs = batch.stats()
            st.plans_built, st.tunes, st.tune_ms = bs["plans_built"], bs["tunes"], bs["tune_ms"]
        finally:
            for x in active:
                x.sim.close()
            batch.close()
        return sorted(results, key=lambda r: r.rid)


class WallClock:
    """Seconds since the clock was made."""

    def __init__(self):
        self.t0 = time.perf_counter()

    def __call__(self) -> float:
        return time.perf_counter() - self.t0

    def wait_until(self, t: float):
        time.sleep(max(0.0, min(t - self(), 0.05)))


class StepClock:
    """A reproducible clock for Server.run: time is `tick` per batched step
    the server has run; with nothing to run it jumps to the next arrival."""

    def __init__(self, server: Server, tick: float = 1.0):
        self.server = server
        self.tick = tick
        self.offset = 0.0

    def __call__(self) -> float:
        return self.offset + self.server.stats.steps * self.tick

    def wait_until(self, t: float):
        self.offset += max(0.0, t - self())


def percentile(values, q: float) -> float:
    """Nearest-rank percentile (q in [0, 100])."""
    v = sorted(values)
    if not v:
        return float("nan")
    k = max(0, min(len(v) - 1, int(-(-q * len(v) // 100)) - 1))
    return v[k]
