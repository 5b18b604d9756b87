"""Host layer: the ini reader and parameter derivation (CPU only).

Config semantics follow src/config.cpp:59-129 and include/config.h:140-159;
the derived constants follow src/{properties,kernel,domain,physics,sim_time}.cpp.
Every expected value below comes from the closed form of the quantity, not
from the code under test.
"""
import ctypes as C
import math

import numpy as np
import pytest

from fluidsim_amd import _native, scenes
from fluidsim_amd._native import FsaParams


@pytest.fixture(scope="module")
def host():
    return _native.host_lib()


def _ini(tmp_path, text, name="input.ini"):
    p = tmp_path / name
    p.write_text(text)
    return str(p).encode()


def read_str(host, path, key):
    buf = C.create_string_buffer(256)
    found = C.c_int(0)
    _native.check_host(host.fsh_config_read_string(path, key.encode(), buf, 256, C.byref(found)))
    return buf.value.decode(), bool(found.value)


def test_comments_sections_and_trimming(host, tmp_path):
    path = _ini(tmp_path, "[physics]\n  gravity = 1.5   ## down\n\n[fluid]\n# is_viscid = true\nfluid_density=998.0\n")
    assert read_str(host, path, "gravity") == ("1.5", True)
    assert read_str(host, path, "fluid_density") == ("998.0", True)
    assert read_str(host, path, "is_viscid") == ("", False)  # commented out
    assert read_str(host, path, "[fluid]") == ("", False)


def test_repeated_key_overwrites_and_defaults(host, tmp_path):
    path = _ini(tmp_path, "dt = 1e-4\ndt = 2e-4\n")
    v = C.c_float()
    _native.check_host(host.fsh_config_read_float(path, b"dt", 0.0, C.byref(v)))
    assert v.value == np.float32(2e-4)
    _native.check_host(host.fsh_config_read_float(path, b"missing", 3.25, C.byref(v)))
    assert v.value == 3.25


def test_continuation_lines(host, tmp_path):
    # a non-blank line without '=' directly after a key continues its value
    path = _ini(tmp_path, "a = one\ntwo\n\nb = 3\n")
    assert read_str(host, path, "a") == ("one\ntwo", True)
    assert read_str(host, path, "b") == ("3", True)


def test_integer_reads_stop_at_non_digits(host, tmp_path):
    # istream >> int reads "1e5" as 1 (the reference warns about this, cases/dam_break/config/input.ini:29)
    path = _ini(tmp_path, "iend = 1e5\nistart = 7\n")
    v = C.c_int()
    _native.check_host(host.fsh_config_read_int(path, b"iend", 0, C.byref(v)))
    assert v.value == 1
    _native.check_host(host.fsh_config_read_int(path, b"istart", 0, C.byref(v)))
    assert v.value == 7


@pytest.mark.parametrize("text,expected", [("false", 0), ("F", 0), ("no", 0), ("N", 0), ("0", 0), ("none", 0),
                                           ("true", 1), ("yes", 1), ("1", 1), ("-1", 1), ("anything", 1)])
def test_bool_spelling(host, tmp_path, text, expected):
    path = _ini(tmp_path, f"is_viscid = {text}\n")
    v = C.c_int()
    _native.check_host(host.fsh_config_read_bool(path, b"is_viscid", 0, C.byref(v)))
    assert v.value == expected


def test_missing_file_is_an_error(host, tmp_path):
    v = C.c_float()
    rc = host.fsh_config_read_float(str(tmp_path / "nope.ini").encode(), b"dt", 0.0, C.byref(v))
    assert rc == _native.FSA_ERR_INVALID
    assert b"not found" in host.fsh_last_error()


def _params(tmp_path, text, counts=(10, 2, 3)):
    path = _ini(tmp_path, text)
    p = FsaParams()
    rc = _native.host_lib().fsh_parameter_from_ini(path, *counts, C.byref(p))
    return rc, p


BASE = """gravity = 9.81
is_viscid = true
fluid_density = 1000
fluid_sound_speed = 30
gamma = 7
viscosity = 1e-3
solid_density = 1100
youngs_modulus = 12e6
poisson_ratio = 0.4
xmin = 0
xmax = 0.4
ymin = 0
ymax = 0.28
zmin = 0
zmax = 0.1
dt = 4.3333e-5
iend = 10
solid_sub_step = 4
smoothing_length = 0.0052
particle_diameter = 0.004
impact_length_hsml_ratio = 2.0
"""


def test_derived_fluid_and_solid_constants(tmp_path):
    rc, p = _params(tmp_path, BASE)
    assert rc == 0
    rho0, c0, gam = 1000.0, 30.0, 7.0
    # Tait: p = B((rho/rho0)^gamma - 1), B = rho0 c0^2 / gamma
    assert math.isclose(p.fluid_coefficient_rho2p, rho0 * c0 ** 2 / gam, rel_tol=1e-6)
    assert math.isclose(p.fluid_min_density, 0.9 * rho0, rel_tol=1e-7)
    assert math.isclose(p.fluid_max_density, 1.1 * rho0, rel_tol=1e-7)
    assert p.fluid_min_pressure == np.float32(-1.013e5)  # default key value
    E, nu, rs = 12e6, 0.4, 1100.0
    assert math.isclose(p.solid_nu, E / (2 * (1 + nu)), rel_tol=1e-6)  # shear modulus
    assert math.isclose(p.solid_lambda, E * nu / ((1 + nu) * (1 - 2 * nu)), rel_tol=1e-6)
    assert math.isclose(p.solid_sound_speed, math.sqrt(E / (3 * (1 - 2 * nu)) / rs), rel_tol=1e-6)
    assert list(p.gravity) == [0.0, np.float32(-9.81), 0.0]  # the key is a magnitude along -y
    assert list(p.solid_artificial_viscosity) == [2.5, 2.5]
    assert p.fluid_is_viscid == 1


def test_grid_covers_the_domain_with_cells_at_least_one_cutoff_wide(tmp_path):
    rc, p = _params(tmp_path, BASE)
    assert rc == 0
    K, h, dp = 2.0, 0.0052, 0.004
    ext = [0.4, 0.28, 0.1]
    for a in range(3):
        pad = 8 * dp  # VIRT_LAYER spacings, include/global.h:10
        assert math.isclose(p.cell_min[a], -pad, abs_tol=1e-7)
        n = int((ext[a] + 2 * pad) / (K * h))
        assert p.cell_number[a] == n
        assert p.interval[a] >= K * h * (1 - 1e-6)
        assert math.isclose(p.interval[a] * n, ext[a] + 2 * pad, rel_tol=1e-6)
    assert p.cell_number_total == p.cell_number[0] * p.cell_number[1] * p.cell_number[2]
    assert math.isclose(p.impact_distance_square, (K * h) ** 2, rel_tol=1e-6)


def test_kernel_normalisation_matches_closed_form(tmp_path):
    """coef = 1 / (4 pi int_0^K x^2 e^-x^2 dx - e^-K^2 4/3 pi K^3).

    The integral is sqrt(pi)/4 erf(K) - K/2 e^-K^2.  The solver takes a left
    Riemann sum of 10^4 FP32 terms instead (src/kernel.cpp:25-41), which sits
    below the integral by ~dx*K^2 e^-K^2/2; the test pins that gap, so a
    changed sum rule, term or constant shows up."""
    host = _native.host_lib()
    for K in (2.0, 3.0):
        integral = math.sqrt(math.pi) / 4 * math.erf(K) - K / 2 * math.exp(-K * K)
        denom = 4 * math.pi * integral - math.exp(-K * K) * 4 / 3 * math.pi * K ** 3
        exact = 1.0 / denom
        got = host.fsh_kernel_coefficient(K)
        # left-endpoint rule error: -(f(K)-f(0)) dx / 2 with f = x^2 e^-x^2, dx = K/1e4
        dx = K / 1e4
        riemann = 4 * math.pi * (integral - K * K * math.exp(-K * K) * dx / 2)
        expect = 1.0 / (riemann - math.exp(-K * K) * 4 / 3 * math.pi * K ** 3)
        # plus FP32 accumulation of 1e4 terms (~N eps/4 relative)
        assert abs(got / expect - 1) < 5e-6, (K, got, expect)
        gap = got / exact - 1
        assert 1e-6 < abs(gap) < 1e-4, gap


def test_kernel_constants(tmp_path):
    rc, p = _params(tmp_path, BASE)
    h, dp, K = 0.0052, 0.004, 2.0
    coef = _native.host_lib().fsh_kernel_coefficient(K)
    assert math.isclose(p.kernel_coefficient_1, math.exp(-K * K), rel_tol=1e-6)
    assert math.isclose(p.kernel_coefficient_2, coef * (h / dp) ** -3, rel_tol=1e-6)
    assert math.isclose(p.kernel_differential, 2 * coef * (h / dp) ** -3 / h ** 2, rel_tol=1e-6)
    assert math.isclose(p.smoothing_length_inverse_square, 1 / h ** 2, rel_tol=1e-6)


@pytest.mark.parametrize("drop", ["dt", "smoothing_length", "xmax"])
def test_missing_required_keys_fail_like_the_reference(tmp_path, drop):
    text = "\n".join(l for l in BASE.splitlines() if not l.startswith(drop + " "))
    rc, _ = _params(tmp_path, text)
    assert rc == _native.FSA_ERR_INVALID
    assert _native.host_lib().fsh_last_error() == b"FAILED TO READ INPUT.INI"


def test_case_directory_counts_and_mismatch(tmp_path):
    sc = scenes.tiny_scene()
    root = sc.write_case(tmp_path / "case")
    a = sc.arrays()
    p = FsaParams()
    host = _native.host_lib()
    _native.check_host(host.fsh_parameter_initiate(str(root).encode(), C.byref(p)))
    assert (p.n_fluid, p.n_solid, p.n_virt) == (len(a["fluid_position"]), len(a["solid_position"]),
                                                len(a["virt_position"]))
    # header count that disagrees with the number of lines
    f = root / "data" / "solid_particle.dat"
    lines = f.read_text().splitlines()
    f.write_text("\n".join([str(int(lines[0]) + 1)] + lines[1:]) + "\n")
    rc = host.fsh_parameter_initiate(str(root).encode(), C.byref(p))
    assert rc == _native.FSA_ERR_INVALID
    assert host.fsh_last_error() == b"FAILED TO READ PARTICLES PARAMETER"


def test_dat_files_round_trip_float32(tmp_path):
    sc = scenes.tiny_scene(seed=5)
    a = sc.arrays()
    root = sc.write_case(tmp_path / "case", a)
    rows = np.loadtxt(root / "data" / "fluid_particle.dat", skiprows=1, dtype=np.float64)
    assert rows.shape == (len(a["fluid_position"]), 8)
    np.testing.assert_array_equal(rows[:, :3].astype(np.float32), a["fluid_position"])
    srows = np.loadtxt(root / "data" / "solid_particle.dat", skiprows=1)
    np.testing.assert_array_equal(srows[:, 9].astype(np.int32), a["solid_type"])
