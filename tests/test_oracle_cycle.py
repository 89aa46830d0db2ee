"""The oracle on a band whose round loop cycles (tests/cycle_case.py): stopped after an odd or
an even number of rounds, it holds exactly the two states the reference's algorithm alternates
between, derived by hand in cycle_case."""
import numpy as np
import pytest

import cycle_case
import oracle_lib
from felt_amd._binding import FeltError, speed_constant


@pytest.mark.parametrize("cap,want", [(1, "initial"), (2, "flipped"), (7, "initial"),
                                      (8, "flipped")])
def test_oracle_round_loop_cycles(cap, want):
    lib = oracle_lib.binding().lib
    s = oracle_lib.surface(3, cycle_case.L, cycle_case.SIZE, cycle_case.PART)
    s.load_dense(cycle_case.initial())
    assert [s.layer_size(k) for k in range(-3, 4)] == [18, 5, 1, 0, 0, 0, 0]
    lib.felt_oracle_set_round_diag(0, cap, 0)
    try:
        with pytest.raises(FeltError, match="round cap"):
            s.update(speed_constant(0.0))
    finally:
        lib.felt_oracle_set_round_diag(0, 0, 0)
    # the loop stops after round cap + 1: odd rounds leave the flipped state, even the initial
    expect = cycle_case.initial() if want == "initial" else cycle_case.flipped()
    assert np.array_equal(s.snapshot(), expect)
