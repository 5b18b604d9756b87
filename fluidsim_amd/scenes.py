"""Seeded dam-break-onto-an-elastic-gate scenes in the reference's formats.

The reference ships one case generator (cases/dam_break/particle_generate.cpp)
that fills a lattice at spacing dp, marks sites outside the tank as virtual
wall particles, sites in the gate box as solid (FIXED below the floor), and
sites in the water column as fluid, then writes the three .dat files
(particle_generate.cpp:93-137).  This module builds the same kind of scene for
any size, with gravity along -y as the solver hard-wires it
(src/physics.cpp:12-15): x is the flow direction, y is up, z is the width.

A scene gives
  * `ini_text()`            -- config/input.ini with the reference's keys
  * `arrays()`              -- float32/int32 particle arrays in .dat order
  * `params()`              -- the derived fsa_params, from the C++ host layer
  * `write_case(root)`      -- <root>/config/input.ini + <root>/data/*.dat
"""
from __future__ import annotations

import math
import os
import tempfile
from dataclasses import dataclass, field, replace
from pathlib import Path

import numpy as np

from . import _native


@dataclass(frozen=True)
class DamBreak:
    """Water column in a closed tank, flowing onto a gate clamped in the floor.

    Sizes are lattice counts (sites of spacing dp) as (x, y, z) = (flow, up, width).
    """

    dp: float = 0.004
    fluid: tuple = (50, 35, 25)       # water column, starting at the tank corner
    tank: tuple = (100, 70, 25)       # inner tank size
    gate: tuple | None = (4, 20, 25)  # gate size including its clamped rows
    gate_x0: int = 75                 # first gate column
    gate_z0: int | None = None        # first gate layer in z (None: centred)
    gate_clamp_rows: int = 3          # bottom rows below the floor, type FIXED
    wall_layers: int = 3
    jitter: float = 0.01              # fluid position jitter, fraction of dp (uniform +-)
    seed: int = 0
    # physics (config/input.ini keys)
    h_ratio: float = 1.3              # smoothing_length / particle_diameter
    impact_ratio: float = 2.0         # impact_length_hsml_ratio
    fluid_density: float = 1000.0
    sound_speed: float = 30.0
    gamma: float = 7.0
    viscosity: float = 1.0e-3         # dynamic viscosity (the key's meaning, include/properties.h:13)
    is_viscid: bool = True
    gravity: float = 9.81
    solid_density: float = 1100.0
    youngs_modulus: float = 12.0e6
    poisson_ratio: float = 0.4
    arti_vis_alpha: float = 2.5
    arti_vis_beta: float = 2.5
    dt: float | None = None           # default 0.25 h / c0
    solid_sub_step: int = 4
    iend: int = 1
    result_interval: int = 1000000
    precision_form: int = 0           # FSA_FORM_* bits (optional ini key)
    neighbour_skin: float | None = None  # fluid neighbour-list skin (optional ini key; None: the library default)

    # ---- derived --------------------------------------------------------
    @property
    def h(self) -> float:
        return self.h_ratio * self.dp

    @property
    def time_step(self) -> float:
        return self.dt if self.dt is not None else 0.25 * self.h / self.sound_speed

    def domain(self):
        lo = (0.0, 0.0, 0.0)
        hi = tuple(n * self.dp for n in self.tank)
        return lo, hi

    def ini_text(self, extra: dict | None = None) -> str:
        lo, hi = self.domain()
        keys = {
            "gravity": self.gravity,
            "is_viscid": "true" if self.is_viscid else "false",
            "fluid_density": self.fluid_density,
            "fluid_sound_speed": self.sound_speed,
            "gamma": self.gamma,
            "viscosity": self.viscosity,
            "solid_density": self.solid_density,
            "youngs_modulus": self.youngs_modulus,
            "poisson_ratio": self.poisson_ratio,
            "arti_vis_alpha": self.arti_vis_alpha,
            "arti_vis_beta": self.arti_vis_beta,
            "xmin": lo[0], "xmax": hi[0], "ymin": lo[1], "ymax": hi[1], "zmin": lo[2], "zmax": hi[2],
            "dt": self.time_step,
            "istart": 0,
            "iend": self.iend,
            "result_interval": self.result_interval,
            "sort_every_step": 1,
            "solid_sub_step": self.solid_sub_step,
            "smoothing_length": self.h,
            "particle_diameter": self.dp,
            "impact_length_hsml_ratio": self.impact_ratio,
            "thread_num": 256,
        }
        if self.precision_form:
            keys["precision_form"] = self.precision_form
        if self.neighbour_skin is not None:
            keys["neighbour_skin"] = self.neighbour_skin
        if extra:
            keys.update(extra)
        sections = {
            "physics": ["gravity"],
            "fluid": ["is_viscid", "fluid_density", "fluid_sound_speed", "gamma", "viscosity"],
            "solid": ["solid_density", "youngs_modulus", "poisson_ratio", "arti_vis_alpha", "arti_vis_beta"],
            "domain": ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"],
            "time": ["dt", "istart", "iend", "result_interval", "sort_every_step", "solid_sub_step"],
            "kernel": ["smoothing_length", "particle_diameter", "impact_length_hsml_ratio", "thread_num"],
        }
        lines = []
        used = set()
        for sec, names in sections.items():
            lines.append(f"[{sec}]")
            for n in names:
                lines.append(f"{n} = {_fmt(keys[n])}")
                used.add(n)
            lines.append("")
        rest = [k for k in keys if k not in used]
        if rest:
            lines.append("[extra]")
            lines += [f"{k} = {_fmt(keys[k])}" for k in rest]
            lines.append("")
        return "\n".join(lines)

    # ---- particles --------------------------------------------------------
    def _sites(self, lo, hi):
        ax = [np.arange(lo[a], hi[a], dtype=np.int64) for a in range(3)]
        g = np.meshgrid(*ax, indexing="ij")
        return np.stack([x.ravel() for x in g], axis=1)

    def _gate_box(self):
        if self.gate is None:
            return None
        gx, gy, gz = self.gate
        z0 = self.gate_z0 if self.gate_z0 is not None else (self.tank[2] - gz) // 2
        lo = (self.gate_x0, -self.gate_clamp_rows, z0)
        hi = (self.gate_x0 + gx, -self.gate_clamp_rows + gy, z0 + gz)
        return lo, hi

    def arrays(self) -> dict:
        dp = np.float64(self.dp)
        W = self.wall_layers
        L = self.tank
        gate = self._gate_box()

        def in_box(s, box):
            if box is None:
                return np.zeros(len(s), dtype=bool)
            lo, hi = box
            return np.all((s >= np.array(lo)) & (s < np.array(hi)), axis=1)

        def pos(s):
            return ((s.astype(np.float64) + 0.5) * dp).astype(np.float32)

        # solid gate
        if gate is not None:
            ss = self._sites(*gate)
            s_pos = pos(ss)
            s_type = np.where(ss[:, 1] < 0, _native.FIXED_SOLID, _native.UNCONSTRAINED_SOLID).astype(np.int32)
        else:
            ss = np.zeros((0, 3), dtype=np.int64)
            s_pos = np.zeros((0, 3), dtype=np.float32)
            s_type = np.zeros(0, dtype=np.int32)

        # walls: every lattice site of the padded box outside the tank, minus gate sites
        vs = self._sites((-W, -W, -W), (L[0] + W, L[1] + W, L[2] + W))
        inside = np.all((vs >= 0) & (vs < np.array(L)), axis=1)
        vs = vs[~inside & ~in_box(vs, gate)]
        v_pos = pos(vs)
        v_type = np.zeros(len(vs), dtype=np.int32)  # FIXED_VIRT

        # fluid column with seeded jitter
        fs = self._sites((0, 0, 0), tuple(min(self.fluid[a], L[a]) for a in range(3)))
        fs = fs[~in_box(fs, gate)]
        rng = np.random.default_rng(self.seed)
        jit = rng.uniform(-self.jitter, self.jitter, size=fs.shape) * dp
        f_pos = ((fs.astype(np.float64) + 0.5) * dp + jit).astype(np.float32)
        nf = len(f_pos)
        return {
            "fluid_position": f_pos,
            "fluid_velocity": np.zeros((nf, 3), dtype=np.float32),
            "fluid_density": np.full(nf, self.fluid_density, dtype=np.float32),
            "fluid_pressure": np.zeros(nf, dtype=np.float32),
            "solid_position": s_pos,
            "solid_velocity": np.zeros_like(s_pos),
            "solid_reference_position": s_pos.copy(),
            "solid_type": s_type,
            "virt_position": v_pos,
            "virt_type": v_type,
        }

    def counts(self) -> tuple:
        a = self.arrays()
        return len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"])

    def params(self, counts: tuple | None = None, extra: dict | None = None) -> _native.FsaParams:
        """Derived parameters from the host layer's Parameter::initiate."""
        nf, ns, nv = counts if counts is not None else self.counts()
        return params_from_ini_text(self.ini_text(extra), nf, ns, nv)

    def write_case(self, root, arrays: dict | None = None, extra: dict | None = None) -> Path:
        root = Path(root)
        a = arrays if arrays is not None else self.arrays()
        (root / "config").mkdir(parents=True, exist_ok=True)
        (root / "data").mkdir(parents=True, exist_ok=True)
        (root / "config" / "input.ini").write_text(self.ini_text(extra))
        write_dat(root / "data", a)
        return root

    def with_(self, **kw) -> "DamBreak":
        return replace(self, **kw)


def _fmt(v) -> str:
    if isinstance(v, float):
        return repr(v)
    return str(v)


def params_from_ini_text(text: str, nf: int, ns: int, nv: int) -> _native.FsaParams:
    lib = _native.host_lib()
    fd, path = tempfile.mkstemp(suffix=".ini")
    try:
        with os.fdopen(fd, "w") as f:
            f.write(text)
        p = _native.FsaParams()
        _native.check_host(lib.fsh_parameter_from_ini(path.encode(), nf, ns, nv, p))
        return p
    finally:
        os.unlink(path)


def write_dat(data_dir, a: dict) -> None:
    """The three .dat files in the reference's text layout (include/host_data.cuh:25, 84, 152)."""
    data_dir = Path(data_dir)
    fp, fu, fr, fpr = a["fluid_position"], a["fluid_velocity"], a["fluid_density"], a["fluid_pressure"]
    rows = np.concatenate([fp, fu, fr[:, None], fpr[:, None]], axis=1).astype(np.float32)
    _write_rows(data_dir / "fluid_particle.dat", rows, None)
    sp, su, sX, st = (a["solid_position"], a["solid_velocity"], a["solid_reference_position"], a["solid_type"])
    rows = np.concatenate([sp, su, sX], axis=1).astype(np.float32)
    _write_rows(data_dir / "solid_particle.dat", rows, st)
    _write_rows(data_dir / "virt_particle.dat", a["virt_position"].astype(np.float32), a["virt_type"])


def _write_rows(path, rows: np.ndarray, ints) -> None:
    # %.9g round-trips every float32 exactly through the reader's istream >> float
    with open(path, "w") as f:
        f.write(f"{len(rows)}\n")
        for r in range(len(rows)):
            vals = " ".join(f"{float(v):.9g}" for v in rows[r])
            f.write(vals + (f" {int(ints[r])}" if ints is not None else "") + "\n")


# The benchmark scenes: dam break onto an elastic gate at CPU, showcase and
# precision scale (y-up mapping of the z-up specifications)
def c1_cpu_scene(seed: int = 0) -> DamBreak:
    """43,750 fluid (50x35x25, dp 4 mm) + 2,000 solid (4x20x25) + 3 wall layers, S = 4."""
    return DamBreak(dp=0.004, fluid=(50, 35, 25), tank=(100, 70, 25), gate=(4, 20, 25), gate_x0=75,
                    gate_clamp_rows=3, wall_layers=3, seed=seed, iend=500)


def c2_showcase_scene(seed: int = 0) -> DamBreak:
    """4,000,000 fluid (200x200x100, dp 1 mm) + 27,000 solid (6x60x75) + 3 wall layers, S = 4."""
    return DamBreak(dp=0.001, fluid=(200, 200, 100), tank=(400, 400, 100), gate=(6, 60, 75), gate_x0=300,
                    gate_clamp_rows=3, wall_layers=3, seed=seed, iend=1000)


def c3_precision_scene(seed: int = 0, precision_form: int = 0) -> DamBreak:
    """1,000,000 fluid (100^3, dp 2 mm) + 8,000 solid (4x50x40) + 3 wall layers: the
    precision-form comparison scene.  The gate stands 2 dp in front of the column,
    so the water loads it within the first few hundred steps."""
    return DamBreak(dp=0.002, fluid=(100, 100, 100), tank=(200, 160, 100), gate=(4, 50, 40), gate_x0=102,
                    gate_clamp_rows=3, wall_layers=3, seed=seed, iend=1000, precision_form=precision_form)


def tiny_scene(seed: int = 0, **kw) -> DamBreak:
    """A few thousand particles: every stage active, fits an oracle run in well under a second."""
    base = dict(dp=0.01, fluid=(8, 10, 6), tank=(16, 14, 6), gate=(2, 8, 6), gate_x0=11, gate_clamp_rows=2,
                wall_layers=2, seed=seed, sound_speed=20.0, youngs_modulus=2.0e6, solid_density=1100.0)
    base.update(kw)
    return DamBreak(**base)


# ---- scene families: many scenes of uneven size and stiffness ----------------

def solid_sound_speed(youngs_modulus: float, poisson_ratio: float, density: float) -> float:
    """c_s = sqrt(K / rho) with K = E / (3 (1 - 2 nu)), as the host layer derives it
    (src/properties.cpp:59-66)."""
    return math.sqrt(youngs_modulus / (3.0 * (1.0 - 2.0 * poisson_ratio)) / density)


def solid_sub_steps_for(sc: DamBreak, cfl: float = 0.3) -> int:
    """Smallest sub-step count with dt / S <= cfl * h / c_s.

    At cfl = 0.3 this gives the 4 sub-steps of the 12 MPa gate of the 43,750-particle
    dam break (dt c_s / h = 1.12); a 50 MPa gate needs 8, a 5 MPa gate 3."""
    cs = solid_sound_speed(sc.youngs_modulus, sc.poisson_ratio, sc.solid_density)
    return max(1, math.ceil(sc.time_step * cs / (cfl * sc.h) - 1e-9))


def column_scene(n_fluid: float, flow_ratio: float, width_ratio: float, youngs_modulus: float, seed: int,
                 dp: float = 0.004, sound_speed: float | None = None,
                 solid_fraction: float | None = None, precision_form: int = 0) -> DamBreak:
    """A dam break with about `n_fluid` fluid particles in a column of aspect
    ratios flow/height = flow_ratio and width/height = width_ratio, a gate of a
    given stiffness at 1.5 column lengths, and the sub-step count its solid CFL
    needs.  sound_speed None: 10x the dam-break velocity bound sqrt(2 g H).
    solid_fraction: gate particles as a fraction of the fluid count, in a gate
    of 3/4 the column width whose height is 4x its thickness (a slender plate
    under a large column deflects until it tears apart); None: 4 layers thick,
    half the column height."""
    ny = max(4, round((n_fluid / (flow_ratio * width_ratio)) ** (1.0 / 3.0)))
    nx = max(4, round(flow_ratio * ny))
    nz = max(4, round(width_ratio * ny))
    tank = (2 * nx + 10, ny + max(6, ny // 2), nz)
    if solid_fraction is None:
        gate_t = 4
        gate_h = max(8, ny // 2)
        gate_w = max(2, nz - 2 * max(1, nz // 8))
    else:
        gate_w = max(2, min(nz, round(0.75 * nz)))
        gate_t = max(2, round(math.sqrt(solid_fraction * nx * ny * nz / (4.0 * gate_w))))
        gate_h = 4 * gate_t
    c0 = sound_speed if sound_speed is not None else 10.0 * math.sqrt(2.0 * 9.81 * ny * dp)
    sc = DamBreak(dp=dp, fluid=(nx, ny, nz), tank=tank, gate=(gate_t, gate_h, gate_w), gate_x0=(3 * nx) // 2,
                  gate_clamp_rows=3, wall_layers=3, seed=seed, sound_speed=float(c0),
                  youngs_modulus=float(youngs_modulus), precision_form=precision_form)
    return sc.with_(solid_sub_step=solid_sub_steps_for(sc))


# Scene families default to the compensated form: in plain FP32 the gate tip
# of a dam break drifts far from the FP64 solution (much of every solid
# sub-step's displacement is below the ulp of the position: 64 % off after
# 1000 steps of the 1M-particle dam break), compensated updates kept it
# within 0.014 % there, at 2 % more step time.
FAMILY_FORM = 2  # FSA_FORM_COMPENSATED


def batch_scenes(n: int = 64, seed: int = 0, fluid_range=(20_000, 500_000), e_range=(5.0e6, 50.0e6),
                 dp: float = 0.004, sound_speed: float = 30.0, precision_form: int = FAMILY_FORM) -> list:
    """`n` independent dam breaks: fluid counts log-uniform in `fluid_range`,
    column aspect ratios uniform in [0.5, 2], gate particles 1-3 % of the
    fluid count, gate E uniform in `e_range` (nu 0.4, rho 1100), one sound
    speed (so one dt) for all, S per scene from the solid CFL."""
    rng = np.random.default_rng(seed)
    lo, hi = math.log(fluid_range[0]), math.log(fluid_range[1])
    out = []
    for i in range(n):
        nf = math.exp(rng.uniform(lo, hi))
        ax, az = rng.uniform(0.5, 2.0, size=2)
        E = rng.uniform(*e_range)
        frac = rng.uniform(0.01, 0.03)
        out.append(column_scene(nf, ax, az, E, seed=seed * 1000 + i, dp=dp, sound_speed=sound_speed,
                                solid_fraction=frac, precision_form=precision_form))
    return out


def serving_scenes(n: int, seed: int = 0, fluid_range=(10_000, 2_000_000), e_range=(5.0e6, 50.0e6),
                   dp: float = 0.004, precision_form: int = FAMILY_FORM) -> list:
    """Scene-solve requests: fluid counts log-uniform in `fluid_range`, gate
    particles 1-3 % of the fluid count, gate E uniform in `e_range`,
    c0 = 10 sqrt(2 g H) per scene (so dt differs)."""
    rng = np.random.default_rng(seed)
    lo, hi = math.log(fluid_range[0]), math.log(fluid_range[1])
    out = []
    for i in range(n):
        nf = math.exp(rng.uniform(lo, hi))
        ax, az = rng.uniform(0.5, 2.0, size=2)
        E = rng.uniform(*e_range)
        frac = rng.uniform(0.01, 0.03)
        out.append(column_scene(nf, ax, az, E, seed=seed * 1000 + i, dp=dp, solid_fraction=frac,
                                precision_form=precision_form))
    return out
