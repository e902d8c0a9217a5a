"""Host logic of sampling: measurement keys, bit-string order and the measure-ending rule
(mindquantum/core/gates/measurement.py:254-298, vector_state.tpp:1326-1347)."""
import numpy as np
import pytest

import mindquantum_amd as mq


def test_keys_are_stored_in_reverse_circuit_order():
    c = mq.Circuit().h(0).x(1, 0).measure(0).measure("b", 1)
    res = mq.MeasureResult()
    res.add_measure(c.all_measures)
    assert res.keys == ["b", "q0"] and res.keys_map == {"b": 0, "q0": 1}
    res.collect_data(np.array([[1, 0], [1, 0], [0, 0]]))
    assert res.data == {"00": 1, "10": 2} and res.shots == 3
    with pytest.raises(ValueError):
        res.add_measure(mq.Measure("b").on(3))


def test_measure_ending_rule():
    assert mq.Circuit().h(0).x(1, 0).measure(0).measure(1).is_measure_end
    assert not mq.Circuit().measure(0).h(0).is_measure_end
    assert not mq.Circuit().h(1).measure(1).x(0, 1).is_measure_end  # a control after the measurement
    assert mq.Circuit().measure(0).h(1).is_measure_end
    assert len(mq.Circuit().h(0).measure(0).remove_measure()) == 1
    with pytest.raises(ValueError):
        mq.Measure().on(0, 1)
