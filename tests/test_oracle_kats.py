"""The CPU oracle against every known answer the reference's own tests hold for this path
(test_Surface.cpp; fixtures in tests/golden/surface_kats.json)."""
import pytest

import kat_replay
import oracle_lib

KATS = kat_replay.load_kats()


@pytest.mark.parametrize("kat", KATS, ids=[k["name"] for k in KATS])
def test_oracle_matches_reference_kat(kat):
    fails = kat_replay.check_kat(oracle_lib.surface, kat)
    assert not fails, f"{kat['name']} ({kat['cite']}): {fails}"


def test_kat_detects_planted_mistake():
    """A grid with one voxel changed must fail the element-wise check."""
    kat = next(k for k in KATS if k["name"] == "expand_1")
    s = oracle_lib.surface(kat["D"], kat["L"], kat["size"], kat["part"])
    kat_replay.replay(s, kat["ops"])
    bad = list(kat["grid"])
    bad[40] += 1.0
    bad[41] -= 1.0  # keeps the reference's summed check at zero
    ok, _ = kat_replay.grid_matches(s.snapshot(), bad)
    assert not ok
