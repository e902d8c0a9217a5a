"""Host logic of the tile-width autotuner: the choice among timed candidates."""
from mindquantum_amd import autotune


def test_pick_keeps_start_within_tolerance():
    times = {(12, 11): 0.2428, (12, 12): 0.2459, (13, 12): 0.2497}
    assert autotune._pick(times, (12, 12), 0.02) == (12, 12)   # 1.3% apart: noise, keep
    assert autotune._pick(times, (12, 12), 0.0) == (12, 11)


def test_pick_moves_when_clearly_faster():
    assert autotune._pick({10: 0.72, 11: 0.65, 12: 0.57}, 11, 0.02) == 12
    assert autotune._pick({10: 0.72, 12: 0.57}, 11, 0.02) == 12   # start not timed
