"""Host logic of expectation_with_grad_batch: native circuits are reused across calls."""
import mindquantum_amd as mq
from mindquantum_amd.simulator import batch


def test_native_circuit_cache_reuses_and_refreshes():
    circ = mq.Circuit().rx("a", 0).ry("b", 1)
    names = list(circ.params_name)
    first = batch._native_circuit(circ, names)
    assert batch._native_circuit(circ, names) is first
    circ.rz("c", 1)  # the circuit grew: a new native circuit
    grown = batch._native_circuit(circ, list(circ.params_name))
    assert grown is not first
    other = mq.Circuit().rx("a", 0).ry("b", 1)
    assert batch._native_circuit(other, list(other.params_name)) is not grown
