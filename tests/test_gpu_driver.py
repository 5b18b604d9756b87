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


# ---- byte-exact .vtk and checkpoint / resume ----------------------------------------

def _f(v) -> str:
    """std::to_string(float): printf("%f") of the value widened to double."""
    return "%f" % float(v)


def _vtk_vectors(title, name, pts, vec) -> str:
    n = len(pts)
    s = f"# vtk DataFile Version 2.0\n{title}\nASCII\nDATASET POLYDATA\nPOINTS {n} float\n"
    s += "".join(f"{_f(p[0])} {_f(p[1])} {_f(p[2])}\n" for p in pts)
    s += f"POINT_DATA {n}\nVECTORS {name} float\n"
    s += "".join(f"{_f(v[0])} {_f(v[1])} {_f(v[2])}\n" for v in vec)
    return s


def _vtk_scalars(name, pts, val) -> str:
    n = len(pts)
    s = f"# vtk DataFile Version 2.0\nscalar vector file\nASCII\nDATASET POLYDATA\nPOINTS {n} float\n"
    s += "".join(f"{_f(p[0])} {_f(p[1])} {_f(p[2])}\n" for p in pts)
    s += f"POINT_DATA {n}\nSCALARS {name} float 1\nLOOKUP_TABLE default\n"
    s += "".join(f"{_f(v)}\n" for v in val)
    return s


def _expected_files(sc, f, s, v, vpos):
    """The nine files of src/output.cu:6-297 for a state, keyed by path under out/."""
    zero = np.zeros(len(f["position"]), np.float32)  # the `color` field the reference writes as pressure
    return {
        "fluid/fluid_velocity": _vtk_vectors("Velocity vector file", "velocity", f["position"], f["velocity"]),
        "fluid/fluid_density": _vtk_scalars("density", f["position"], f["density"]),
        "fluid/fluid_pressure": _vtk_scalars("pressure", f["position"], zero),
        "solid/solid_velocity": _vtk_vectors("Velocity vector file", "velocity", s["position"], s["velocity"]),
        "solid/solid_pressure": _vtk_scalars("pressure", s["position"], s["pressure"]),
        "solid/solid_density": _vtk_scalars("density", s["position"],
                                            np.full(len(s["position"]), sc.solid_density, np.float32)),
        "solid/solid_normal": _vtk_vectors("Velocity vector file", "normal", s["position"], s["normal"]),
        "virtual/virt_normal": _vtk_vectors("Velocity vector file", "normal", vpos, v["normal"]),
        "virtual/virt_pressure": _vtk_scalars("pressure", vpos, v["pressure"]),
    }


@pytest.mark.parametrize("crlf", [False, True])
def test_vtk_files_are_byte_exact(gpu_available, tmp_path, crlf):
    """Every byte of the nine files: headers, %f numbers, .dat order, the
    state at the start of step 1 (file 0) and after the loop (file 1); with
    vtk_crlf the CR LF line ends of the reference's Windows text streams."""
    from fluidsim_amd.sim import DeviceSim
    sc = scenes.tiny_scene(seed=19).with_(iend=3, result_interval=3)
    a = sc.arrays()
    extra = {"vtk_crlf": "true"} if crlf else None
    root = sc.write_case(tmp_path / "case", a, extra=extra)
    host = _native.host_lib()
    assert host.fsh_run(str(root).encode(), 0) == 0, host.fsh_last_error()
    p = sc.params((len(a["fluid_position"]), len(a["solid_position"]), len(a["virt_position"])))
    dev = DeviceSim(p)
    dev.upload(a)
    states = {0: (dev.fluid(), dev.solid(), dev.virt())}
    dev.step(3)
    states[1] = (dev.fluid(), dev.solid(), dev.virt())
    dev.close()
    for k, (f, s, v) in states.items():
        for name, text in _expected_files(sc, f, s, v, a["virt_position"]).items():
            want = text.replace("\n", "\r\n") if crlf else text
            got = (root / "out" / f"{name}-{k}.vtk").read_bytes()
            assert got == want.encode(), f"{name}-{k}.vtk differs"


def test_checkpoint_resumes_bit_for_bit(gpu_available, tmp_path):
    """checkpoint_interval writes the state after step 10 as a case with
    istart = 10; running that case repeats steps 11..20 of the uninterrupted
    run exactly: the same .vtk bytes for files 2..4 and the same step-20
    checkpoint."""
    sc = scenes.tiny_scene(seed=20, fluid=(12, 10, 6)).with_(iend=20, result_interval=5)
    root = sc.write_case(tmp_path / "case", extra={"checkpoint_interval": 10})
    host = _native.host_lib()
    assert host.fsh_run(str(root).encode(), 0) == 0, host.fsh_last_error()
    ck10, ck20 = root / "checkpoint" / "step-10", root / "checkpoint" / "step-20"
    assert (ck10 / "config" / "input.ini").exists() and (ck20 / "data" / "fluid_particle.dat").exists()
    assert "istart = 10" in (ck10 / "config" / "input.ini").read_text()
    assert host.fsh_run(str(ck10).encode(), 0) == 0, host.fsh_last_error()
    state = [("fluid", "fluid_velocity"), ("fluid", "fluid_density"), ("fluid", "fluid_pressure"),
             ("solid", "solid_velocity"), ("solid", "solid_density")]
    derived = [("solid", "solid_pressure"), ("solid", "solid_normal"), ("virtual", "virt_normal"),
               ("virtual", "virt_pressure")]
    # file 2 is the state at the start of step 11: the resumed run has just
    # uploaded it, so the fields the coupling and wall passes derive anew in
    # every step still hold their upload defaults there (as the reference's
    # first output does); from file 3 on every field matches
    for k, fields in ((2, state), (3, state + derived), (4, state + derived)):
        for sub, f in fields:
            want = (root / "out" / sub / f"{f}-{k}.vtk").read_bytes()
            got = (ck10 / "out" / sub / f"{f}-{k}.vtk").read_bytes()
            assert got == want, (sub, f, k)
    for dat in ("fluid_particle.dat", "solid_particle.dat", "virt_particle.dat"):
        assert (ck10 / "checkpoint" / "step-20" / "data" / dat).read_bytes() == (ck20 / "data" / dat).read_bytes()
