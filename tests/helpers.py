"""Shared helpers of the test suite: scenes, run drivers and tolerance checks."""
from __future__ import annotations

import math

import numpy as np

from fluidsim_amd import scenes

# ---- tolerances (relative, stated once for every parity test) --------------
# Per-particle fluid density, velocity and position: 1e-5 relative.
#   position  |dx|   <= 1e-5 * L          (L = largest domain extent)
#   velocity  |du|   <= 1e-5 * U          (U = max(max|u|, sqrt(2 g H)), H = water depth)
#   density   |drho| <= 1e-5 * rho
#   pressure  |dp|   <= 1e-5 * rho0 * c0^2 * 1.1^(gamma-1)   (density tolerance through the EOS slope)
REL = 1.0e-5


def velocity_scale(sc: scenes.DamBreak, *us) -> float:
    H = sc.fluid[1] * sc.dp
    umax = max([float(np.abs(u).max()) if len(u) else 0.0 for u in us] + [0.0])
    return max(umax, math.sqrt(2.0 * sc.gravity * H))


def length_scale(sc: scenes.DamBreak) -> float:
    return max(sc.tank) * sc.dp


def pressure_tolerance(sc: scenes.DamBreak, rel: float = REL) -> float:
    return rel * sc.fluid_density * sc.sound_speed ** 2 * 1.1 ** (sc.gamma - 1.0)


def max_abs(a, b) -> float:
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max()) if a.size else 0.0


def compare_fluid(sc, got: dict, ref: dict, rel: float = REL) -> dict:
    """Errors of a fluid state against a reference state, and whether each is within tolerance."""
    L = length_scale(sc)
    U = velocity_scale(sc, ref["velocity"])
    dx = max_abs(got["position"], ref["position"])
    du = max_abs(got["velocity"], ref["velocity"])
    drho = float(np.max(np.abs(np.asarray(got["density"], np.float64) - ref["density"]) / np.abs(ref["density"])))
    dp = max_abs(got["pressure"], ref["pressure"])
    ptol = pressure_tolerance(sc, rel)
    return {
        "position": (dx / L, dx <= rel * L),
        "velocity": (du / U, du <= rel * U),
        "density": (drho, drho <= rel),
        "pressure": (dp / ptol * rel, dp <= ptol),
    }


def assert_within(report: dict, what: str = ""):
    bad = {k: v[0] for k, v in report.items() if not v[1]}
    assert not bad, f"{what}: relative errors over tolerance {bad}; all: { {k: v[0] for k, v in report.items()} }"


def custom_case(sc: scenes.DamBreak, fluid=None, solid=None, virt=None) -> tuple:
    """Parameters of scene `sc` with hand-placed particles.

    fluid: dict(position, velocity=0, density=rho0, pressure=0)
    solid: dict(position, velocity=0, reference_position=position, type=0)
    virt:  dict(position)
    """
    def arr(x, n, cols=3, fill=0.0, dt=np.float32):
        if x is None:
            return np.full((n, cols) if cols > 1 else n, fill, dtype=dt)
        return np.ascontiguousarray(np.asarray(x, dtype=dt).reshape((n, cols) if cols > 1 else n))

    fluid = fluid or {}
    solid = solid or {}
    virt = virt or {}
    fx = np.asarray(fluid.get("position", np.zeros((0, 3))), np.float32).reshape(-1, 3)
    sx = np.asarray(solid.get("position", np.zeros((0, 3))), np.float32).reshape(-1, 3)
    vx = np.asarray(virt.get("position", np.zeros((0, 3))), np.float32).reshape(-1, 3)
    nf, ns, nv = len(fx), len(sx), len(vx)
    a = {
        "fluid_position": fx,
        "fluid_velocity": arr(fluid.get("velocity"), nf),
        "fluid_density": arr(fluid.get("density"), nf, 1, sc.fluid_density),
        "fluid_pressure": arr(fluid.get("pressure"), nf, 1, 0.0),
        "solid_position": sx,
        "solid_velocity": arr(solid.get("velocity"), ns),
        "solid_reference_position": arr(solid.get("reference_position", sx), ns),
        "solid_type": arr(solid.get("type"), ns, 1, 0, np.int32),
        "virt_position": vx,
        "virt_type": np.zeros(nv, np.int32),
    }
    return sc.params((nf, ns, nv)), a


def same_class(got, o32, o64, scale: float, factor: float = 4.0, rel: float = REL) -> tuple:
    """FP32 result `got` is as accurate as the FP32 oracle: |got-o64| <= factor*|o32-o64| + rel*scale."""
    e_got = max_abs(got, o64)
    e_32 = max_abs(o32, o64)
    return e_got, e_32, e_got <= factor * e_32 + rel * scale


def observables_from_state(params, fluid: dict, solid: dict, solid_ref, solid_type) -> dict:
    """The quantities of fsa_observe, from a downloaded state, in float64."""
    dp = float(params.particle_diameter)
    V = dp ** 3
    rho = np.asarray(fluid["density"], np.float64)
    u = np.asarray(fluid["velocity"], np.float64)
    x = np.asarray(fluid["position"], np.float64)
    g = np.array(params.gravity, np.float64)
    out = {
        "fluid_mass": rho.sum() * V,
        "fluid_momentum": list((rho[:, None] * u).sum(axis=0) * V),
        "fluid_kinetic_energy": 0.5 * (rho * (u * u).sum(axis=1)).sum() * V,
        "fluid_potential_energy": -(rho * (x @ g)).sum() * V,
        "front_x": float(x[:, 0].max()) if len(x) else 0.0,
    }
    ns = len(solid_type)
    if ns:
        us = np.asarray(solid["velocity"], np.float64)
        out["solid_momentum"] = list(us.sum(axis=0) * float(params.solid_reference_density) * V)
        X = np.asarray(solid_ref, np.float64)
        free = np.asarray(solid_type) == 0
        ytop = X[free, 1].max()
        tip = free & (X[:, 1].astype(np.float32) >= np.float32(np.float32(ytop) - np.float32(0.5) * np.float32(dp)))
        xs = np.asarray(solid["position"], np.float64)
        out["gate_tip_displacement"] = list((xs[tip] - X[tip]).mean(axis=0))
        out["gate_tip_count"] = int(tip.sum())
    return out


def shepard_probe(params, fluid: dict, points) -> np.ndarray:
    """Brute-force Shepard pressure over all fluid particles (solid.cu:275-291 weights)."""
    x = np.asarray(fluid["position"], np.float64)
    p = np.asarray(fluid["pressure"], np.float64)
    out = []
    for q in np.asarray(points, np.float64):
        d2 = ((x - q) ** 2).sum(axis=1)
        m = d2 < params.impact_distance_square
        w = (np.exp(-d2[m] * params.smoothing_length_inverse_square) - params.kernel_coefficient_1) \
            * params.kernel_coefficient_2
        out.append((p[m] * w).sum() / w.sum() if w.sum() > 1e-6 else 0.0)
    return np.array(out)
