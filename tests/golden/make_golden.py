"""Write the golden fixtures of tests/golden/ from the CPU oracle.

    python tests/golden/make_golden.py

Each fixture holds the input arrays of a small seeded dam-break scene and
the FP64 oracle state after 1 and 10 steps.  The oracle is the CPU
restatement of the reference method (oracle/fsa_oracle.cpp); the reference
itself cannot be built here (CUDA-only, glm and Windows headers absent), so
these vectors are written by the oracle, not by the reference.
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

import oracle  # noqa: E402
from fluidsim_amd import scenes  # noqa: E402

CASES = {
    "tiny_seed7": dict(seed=7),
    "tiny_inviscid_seed8": dict(seed=8, is_viscid=False),
}
STEPS = (1, 10)


def scene(name):
    return scenes.tiny_scene(**CASES[name])


def make(name):
    sc = scene(name)
    a = sc.arrays()
    counts = (len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"]))
    p = sc.params(counts)
    o = oracle.Oracle(p, 64)
    o.upload(a)
    out = {f"in_{k}": v for k, v in a.items()}
    done = 0
    for s in STEPS:
        o.step(s - done)
        done = s
        f, so, v = o.fluid(), o.solid(), o.virt()
        for k in ("position", "velocity", "density", "pressure"):
            out[f"s{s}_fluid_{k}"] = f[k]
        for k in ("position", "velocity", "pressure", "normal"):
            out[f"s{s}_solid_{k}"] = so[k]
        for k in ("pressure", "density", "velocity"):
            out[f"s{s}_virt_{k}"] = v[k]
    np.savez_compressed(HERE / f"{name}.npz", **out)
    return HERE / f"{name}.npz"


if __name__ == "__main__":
    for n in CASES:
        print(make(n))
