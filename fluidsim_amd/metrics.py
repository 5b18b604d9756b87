"""Error measures of a scene state against a reference state (float64).

Used by the precision-form measurement of bench.py and by the tests that
hold each form to its bounds against the FP64 oracle.
"""
from __future__ import annotations

import math

import numpy as np

from .scenes import DamBreak


def tip_displacement(solid_pos, ref_pos, types, dp) -> np.ndarray:
    """Mean x - X over the gate tip row (unconstrained solids whose reference y is
    within dp/2 of the largest one), in float64."""
    types = np.asarray(types)
    free = types == 0
    if not free.any():
        return np.zeros(3)
    X = np.asarray(ref_pos, np.float64)
    row = free & (X[:, 1] >= X[free, 1].max() - 0.5 * dp)
    return (np.asarray(solid_pos, np.float64)[row] - X[row]).mean(axis=0)


def field_errors(sc: DamBreak, got: dict, ref: dict) -> dict:
    """Max and RMS errors of a fluid state against a reference state, each divided
    by its scale: L (largest tank extent), U = max(max |u|, sqrt(2 g H)), rho0,
    and rho0 g H (hydrostatic pressure at the column foot)."""
    H = sc.fluid[1] * sc.dp
    L = max(sc.tank) * sc.dp
    uref = np.asarray(ref["velocity"], np.float64)
    U = max(float(np.sqrt((uref ** 2).sum(axis=1)).max()), math.sqrt(2.0 * sc.gravity * H))
    scales = {"position": L, "velocity": U, "density": sc.fluid_density,
              "pressure": sc.fluid_density * sc.gravity * H}
    out = {}
    for k, s in scales.items():
        d = np.abs(np.asarray(got[k], np.float64) - np.asarray(ref[k], np.float64))
        out[k] = {"max": float(d.max()) / s, "rms": float(np.sqrt((d ** 2).mean())) / s}
    return out
