"""The C ABI libraries load without a GPU and export every declared symbol (CPU only)."""
import ctypes as C
import re
from pathlib import Path

import pytest

from fluidsim_amd import _native

ROOT = Path(__file__).resolve().parents[1]


def declared(header: str, prefix: str) -> set:
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(" + prefix + r"\w+)\s*\(", text))


def test_device_library_exports_every_declared_symbol():
    names = declared("fluidsim_amd.h", "fsa_")
    assert len(names) >= 25
    lib = C.CDLL(str(_native.DEVICE_LIB))
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes table covers the whole header
    assert names == set(_native.DEVICE_API), names ^ set(_native.DEVICE_API)


def test_host_library_exports_every_declared_symbol():
    names = declared("fluidsim_host.h", "fsh_")
    lib = C.CDLL(str(_native.HOST_LIB))
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert names == set(_native.HOST_API)


def test_params_struct_layout_matches_the_header():
    """ctypes mirror of fsa_params has the fields of the header in order."""
    text = (ROOT / "include" / "fluidsim_amd.h").read_text()
    body = text[text.index("typedef struct fsa_params"):text.index("} fsa_params;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in re.findall(r"(?:float|int32_t)\s+([^;]+);", body):
        for name in decl.split(","):
            fields.append(re.sub(r"\[.*\]", "", name).strip())
    assert fields == [f for f, _ in _native.FsaParams._fields_]


def test_abi_version_and_error_paths_without_a_device():
    lib = _native.device_lib()
    assert lib.fsa_abi_version() == 3
    n = C.c_int(-1)
    assert lib.fsa_device_count(C.byref(n)) == 0
    if n.value > 0:
        pytest.skip("a GPU is visible; the no-device path is not reachable")
    p = _native.FsaParams()
    h = C.c_void_p()
    # an invalid parameter block is rejected before any device call
    assert lib.fsa_create(C.byref(p), None, C.byref(h)) == _native.FSA_ERR_INVALID
    assert b"cell_number" in lib.fsa_last_error()
    assert lib.fsa_create(None, None, C.byref(h)) == _native.FSA_ERR_INVALID
    assert lib.fsa_step(None, 1) == _native.FSA_ERR_INVALID


def test_host_driver_reports_missing_device(tmp_path):
    """fsh_run on a valid case fails cleanly (no abort) where no GPU is visible."""
    from fluidsim_amd import scenes
    n = C.c_int(0)
    _native.device_lib().fsa_device_count(C.byref(n))
    if n.value > 0:
        pytest.skip("a GPU is visible")
    root = scenes.tiny_scene().with_(iend=2).write_case(tmp_path / "case")
    host = _native.host_lib()
    rc = host.fsh_run(str(root).encode(), 0)
    assert rc != 0
    assert b"fsa_create" in host.fsh_last_error()


def test_footprint_matches_a_hand_count_of_the_buffers():
    """fsa_footprint needs no device: per fluid particle the context holds
    position + velocity double buffers (4 x 16 B), id and cell double buffers
    (4 x 4 B) and four sort scratch ints (16 B) = 96 B; per solid particle
    additionally the reference position double buffer (32 B), np/acc/couple
    (48 B), four 3x3 planes (144 B), the batched deformation pass's pair
    sums (18 floats: 72 B), the force pass's partner record (X, x, u and P L
    as three float4 rows: 96 B), the step's recorded solid and fluid
    partners (2 x (16 lanes x 16 slots x 4 B + 16 counts x 4 B) = 2176 B)
    and the step-start position (16 B) = 2680 B;
    per wall particle one
    position/velocity buffer (32 B), id and cell double buffers (16 B), sort
    scratch (16 B), normal and constant position (32 B) = 96 B.  (Without
    neighbour lists; their buffers are counted in the next test.)"""
    from fluidsim_amd import scenes
    from fluidsim_amd.sim import footprint
    sc = scenes.tiny_scene(seed=1).with_(neighbour_skin=0.0)
    base = sc.params((1000, 100, 2000))
    r0, p0 = footprint(base)
    r1, _ = footprint(sc.params((1500, 100, 2000)))
    r2, _ = footprint(sc.params((1000, 150, 2000)))
    r3, p3 = footprint(sc.params((1000, 100, 2400)))
    assert r1 - r0 == 500 * 96
    assert r2 - r0 == 50 * 2680
    assert r3 - r0 == 400 * 96
    assert p0 - r0 == 2000 * 32 and p3 - r3 == 2400 * 32


def test_footprint_counts_the_neighbour_lists():
    """With neighbour lists a fluid particle also holds the home-order
    position + velocity double buffers (64 B), home index double buffer and
    sorted slot (12 B), its list -- (rows + 18) 32-bit entries and the
    entry count (4 B) -- its wall list -- (wall rows + 18) entries and a
    count -- and its build position (16 B); the lists' few words of state
    do not grow with the particles.  rows is 1.5 x the lattice sites within
    the list reach (1 + skin) R, rounded up to a multiple of 8, at most 256;
    wall rows half of that, rounded down to a multiple of 8, at least 32."""
    import math
    from fluidsim_amd import scenes
    from fluidsim_amd.sim import footprint
    skin = 0.15
    sc = scenes.tiny_scene(seed=1).with_(neighbour_skin=skin)
    p0, p1 = sc.params((1000, 100, 2000)), sc.params((1640, 100, 2000))
    r0, _ = footprint(p0)
    r1, _ = footprint(p1)
    reach = (1 + skin) * math.sqrt(p0.impact_distance_square) / p0.particle_diameter
    rows = min(max(math.ceil(1.5 * 4 / 3 * math.pi * reach ** 3 / 8) * 8, 32), 256)
    wrows = max(rows // 2 // 8 * 8, 32)
    per = 96 + 64 + 12 + 4 * (rows + 18) + 4 + 4 * (wrows + 18) + 4 + 16
    assert r1 - r0 == 640 * per
    # without a skin, or with the packed form, there are no lists
    assert footprint(sc.with_(neighbour_skin=0.0).params((1000, 100, 2000)))[0] < r0
    assert footprint(sc.with_(precision_form=_native.FORM_PACKED).params((1000, 100, 2000)))[0] < r0
