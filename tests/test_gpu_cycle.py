"""The device on the band whose round loop cycles (tests/cycle_case.py): an update stops with
FELT_E_STATE after the handle's round limit, in one of the two states the reference's algorithm
alternates between, and the handle stays usable (its state reads back, a later update with a
larger limit fails the same way)."""
import numpy as np
import pytest

import cycle_case
from felt_amd._binding import FeltError, speed_constant

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "launches"])
def test_round_limit_stops_a_cycling_band(graph):
    import felt_amd
    g = felt_amd.Surface(3, cycle_case.L, cycle_case.SIZE, cycle_case.PART)
    g.set_graph(graph)
    g.load_dense(cycle_case.initial())
    states = (cycle_case.initial(), cycle_case.flipped())
    for limit in (9, 40):
        g.set_round_limit(limit)
        with pytest.raises(FeltError, match="did not converge"):
            g.update(speed_constant(0.0))
        snap = g.snapshot()
        assert any(np.array_equal(snap, s) for s in states), f"limit {limit}"
    with pytest.raises(FeltError):
        g.set_round_limit(0)
    g.close()


def test_bounded_update_without_zero_layer_leaves_the_blob():
    """update(lo, hi, fn) starts its affected set from the zero layer in the box
    (Surface.hpp:310-386, 780-913); with none, nothing is recomputed and the blob stays."""
    import felt_amd
    g = felt_amd.Surface(3, cycle_case.L, cycle_case.SIZE, cycle_case.PART)
    g.load_dense(cycle_case.initial())
    g.set_round_limit(12)
    g.update_box([-4, -4, -4], [4, 4, 4], speed_constant(0.0))
    assert np.array_equal(g.snapshot(), cycle_case.initial())
    g.close()
