"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the scene generator reproduces the stored inputs bit for bit, the FP64
oracle reproduces the stored states, and the FP32 oracle stays within the
parity tolerances of them.  The GPU side of the same fixtures is in
test_gpu_parity.py.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle as orc
from helpers import compare_fluid, assert_within

GOLDEN = Path(__file__).resolve().parent / "golden"
sys_path_added = False


def load(name):
    import sys
    sys.path.insert(0, str(GOLDEN))
    import make_golden
    d = dict(np.load(GOLDEN / f"{name}.npz"))
    return make_golden.scene(name), d


def inputs(d):
    return {k[3:]: v for k, v in d.items() if k.startswith("in_")}


def state(d, s, kind):
    pre = f"s{s}_{kind}_"
    return {k[len(pre):]: v for k, v in d.items() if k.startswith(pre)}


NAMES = ["tiny_seed7", "tiny_inviscid_seed8"]


@pytest.mark.parametrize("name", NAMES)
def test_generator_reproduces_fixture_inputs(name):
    sc, d = load(name)
    a = sc.arrays()
    for k, v in inputs(d).items():
        np.testing.assert_array_equal(a[k], v, err_msg=k)


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("precision", [64, 32])
def test_oracle_reproduces_fixture_states(name, precision):
    sc, d = load(name)
    a = inputs(d)
    p = sc.params((len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"])))
    o = orc.Oracle(p, precision)
    o.upload(a)
    done = 0
    for s in (1, 10):
        o.step(s - done)
        done = s
        ref = state(d, s, "fluid")
        got = o.fluid()
        if precision == 64:
            for k in ("position", "velocity", "density", "pressure"):
                np.testing.assert_allclose(got[k], ref[k], rtol=1e-12, atol=1e-12, err_msg=f"{k} step {s}")
        else:
            assert_within(compare_fluid(sc, got, ref), f"FP32 oracle vs FP64 fixture, step {s}")
