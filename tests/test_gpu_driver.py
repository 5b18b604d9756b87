"""The drop-in host driver on the GPU: a case directory in, .vtk files out.

Runs the C++ host layer (Parameter, Sim, Output over the C ABI) exactly as
the `fluidsim` executable does, and checks the reference's output set
(src/output.cu:299-335): nine ASCII POLYDATA files per output step, named
<field>-<step/result_interval>.vtk, the state at the start of the step,
plus the final state after the loop.
"""
import re

import numpy as np
import pytest

import oracle as orc
from fluidsim_amd import _native, scenes

pytestmark = pytest.mark.gpu


def read_vtk(path):
    text = path.read_text().splitlines()
    assert text[0] == "# vtk DataFile Version 2.0"
    assert text[2] == "ASCII" and text[3] == "DATASET POLYDATA"
    n = int(re.match(r"POINTS (\d+) float", text[4]).group(1))
    pts = np.array([[float(v) for v in l.split()] for l in text[5:5 + n]])
    assert text[5 + n] == f"POINT_DATA {n}"
    head = text[6 + n]
    if head.startswith("VECTORS"):
        vals = np.array([[float(v) for v in l.split()] for l in text[7 + n:7 + 2 * n]])
    else:
        assert text[7 + n] == "LOOKUP_TABLE default"
        vals = np.array([float(l) for l in text[8 + n:8 + 2 * n]])
    return head, pts, vals


def test_driver_writes_reference_output_set(gpu_available, tmp_path):
    sc = scenes.tiny_scene(seed=9).with_(iend=4, result_interval=2)
    a = sc.arrays()
    root = sc.write_case(tmp_path / "case", a)
    host = _native.host_lib()
    rc = host.fsh_run(str(root).encode(), 0)
    assert rc == 0, host.fsh_last_error()
    out = root / "out"
    # outputs before steps 1 and 3 (file 0, 1) and after the loop (i = 5 -> file 2)
    names = {
        "fluid": ["fluid_velocity", "fluid_density", "fluid_pressure"],
        "solid": ["solid_velocity", "solid_pressure", "solid_density", "solid_normal"],
        "virtual": ["virt_normal", "virt_pressure"],
    }
    for sub, fields in names.items():
        for f in fields:
            for k in (0, 1):
                assert (out / sub / f"{f}-{k}.vtk").exists(), (sub, f, k)
    # file 2 is written only if iend is a multiple of result_interval (iend = 4)
    assert (out / "fluid" / "fluid_velocity-2.vtk").exists()
    head, pts, vel = read_vtk(out / "fluid" / "fluid_velocity-0.vtk")
    assert head == "VECTORS velocity float"
    np.testing.assert_allclose(pts, a["fluid_position"], atol=1e-6)  # %f keeps 6 decimals
    head, _, p = read_vtk(out / "fluid" / "fluid_pressure-1.vtk")
    assert head == "SCALARS pressure float 1" and np.all(p == 0)  # the reference writes `color` (= 0)
    head, _, rho = read_vtk(out / "solid" / "solid_density-1.vtk")
    assert np.all(rho == sc.solid_density)
    # final state against the oracle after 4 steps
    params = sc.params((len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"])))
    o = orc.Oracle(params, 32)
    o.upload(a)
    o.step(4)
    _, pts, vel = read_vtk(out / "fluid" / "fluid_velocity-2.vtk")
    ref = o.fluid()
    np.testing.assert_allclose(pts, ref["position"], atol=2e-6)
    np.testing.assert_allclose(vel, ref["velocity"], atol=2e-6)


def test_driver_true_pressure_key(gpu_available, tmp_path):
    sc = scenes.tiny_scene(seed=9).with_(iend=1, result_interval=1)
    root = sc.write_case(tmp_path / "case", extra={"output_fluid_pressure": "true"})
    assert _native.host_lib().fsh_run(str(root).encode(), 0) == 0
    _, _, p0 = read_vtk(root / "out" / "fluid" / "fluid_pressure-1.vtk")  # start of step 1: p = 0 as uploaded
    _, _, p = read_vtk(root / "out" / "fluid" / "fluid_pressure-2.vtk")   # after the step
    assert np.all(p0 == 0) and np.abs(p).max() > 0
