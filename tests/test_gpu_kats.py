"""The HIP surface against every known answer of the reference's own tests."""
import sys

import pytest

import kat_replay

pytestmark = pytest.mark.gpu
KATS = kat_replay.load_kats()


def _gpu_surface(D, L, size, part):
    import felt_amd
    return felt_amd.Surface(D, L, size, part)


@pytest.mark.parametrize("kat", KATS, ids=[k["name"] for k in KATS])
def test_gpu_matches_reference_kat(kat):
    print(f"[kat] {kat['name']}", flush=True)
    fails = kat_replay.check_kat(_gpu_surface, kat)
    assert not fails, f"{kat['name']} ({kat['cite']}): {fails}"
