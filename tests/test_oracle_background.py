"""The oracle on tests/background_case.py: the background of a partition emptied by untracks
out of both sides of the band in one flush is the target of the last from-layer's untrack, as the
reference's ascending flush order makes it."""
import numpy as np
import pytest

import background_case as bc
import oracle_lib
from felt_amd._binding import speed_constant


@pytest.mark.parametrize("with_plus", [True, False], ids=["both_sides", "inside_only"])
def test_oracle_background_after_emptying_flush(with_plus):
    s = oracle_lib.surface(2, bc.L, bc.SIZE, bc.PART)
    s.load_dense(bc.field(with_plus))
    assert [s.layer_size(k) for k in range(-3, 4)] == [1, 0, 0, 0, 0, 0, 1 if with_plus else 0]
    s.update(speed_constant(0.0))
    assert [s.layer_size(k) for k in range(-3, 4)] == [0] * 7
    assert np.array_equal(s.snapshot(), bc.expected(with_plus))
