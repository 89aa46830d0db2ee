"""The device on tests/background_case.py: a partition emptied by one flush takes the target of
its last from-layer's untrack as background (the reference's ascending flush order), the same as
the oracle, for a flush with untracks out of both sides and for one out of the inside only."""
import numpy as np
import pytest

import background_case as bc
from felt_amd._binding import speed_constant

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("with_plus", [True, False], ids=["both_sides", "inside_only"])
@pytest.mark.parametrize("graph", [True, False], ids=["graph", "launches"])
def test_background_after_emptying_flush(with_plus, graph):
    import felt_amd
    g = felt_amd.Surface(2, bc.L, bc.SIZE, bc.PART)
    g.set_graph(graph)
    g.load_dense(bc.field(with_plus))
    g.update(speed_constant(0.0))
    assert [g.layer_size(k) for k in range(-3, 4)] == [0] * 7
    assert np.array_equal(g.snapshot(), bc.expected(with_plus))
    assert not g.partition_active().any()
    g.close()
