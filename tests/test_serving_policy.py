"""Host logic of the serving loop that needs no GPU: which resident scenes a
chunk steps."""
from types import SimpleNamespace

from fluidsim_amd.serving import Server


def _active(*sizes):
    # resident scenes in arrival order; counts = (fluid, solid, walls)
    return [SimpleNamespace(counts=(n, n // 100, 0), remaining=10, name=i) for i, n in enumerate(sizes)]


def _server(saturation):
    s = Server.__new__(Server)  # the policy only reads `saturation`: no device needed
    s.saturation = saturation
    return s


def test_every_resident_scene_without_a_budget():
    act = _active(3_000_000, 500_000, 20_000)
    assert [x.name for x in _server(None)._chunk_scenes(act)] == [0, 1, 2]


def test_oldest_first_within_the_budget_and_later_small_scenes_fill_in():
    act = _active(600_000, 800_000, 300_000, 50_000)
    # 606k + 303k + 50.5k fit 1M; the 808k scene does not and waits
    assert [x.name for x in _server(1_000_000)._chunk_scenes(act)] == [0, 2, 3]


def test_the_oldest_scene_is_stepped_even_past_the_budget():
    act = _active(2_000_000, 10_000)
    assert [x.name for x in _server(1_000_000)._chunk_scenes(act)] == [0]
