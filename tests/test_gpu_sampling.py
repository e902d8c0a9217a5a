"""GPU sampling of measurement-ending circuits.

Exact check: the uniforms come from numpy's generator seeded with ``seed``, so the lower-bound
index of every shot is recomputed on the host from the GPU's own final state (global cumulative
sum in double) and must agree shot by shot (bar: all but 1e-4 of the shots, for rounding at
interval boundaries).  Statistical checks against the CPU oracle's state: basis states, a Bell
pair, and marginals of random circuits (total variation distance bars stated per test).
"""
import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd import ir
from mindquantum_amd.algorithm import layered_random_circuit
from oracle import cpu

pytestmark = pytest.mark.gpu

DTYPES = [mq.complex64, mq.complex128]


def full_index(res, n):
    idx = np.zeros(res.shots, dtype=np.int64)
    for q in range(n):
        idx |= res.samples[:, res.keys_map[f"q{q}"]].astype(np.int64) << q
    return idx


@pytest.mark.parametrize("dtype", DTYPES)
def test_basis_state(dtype):
    sim = mq.Simulator("mqvector_amd", 5, dtype=dtype)
    c = mq.Circuit().x(0).x(2)
    for q in range(5):
        c.measure(q)
    res = sim.sampling(c, shots=1000, seed=3)
    assert res.data == {"00101": 1000}
    z = np.zeros(32)
    z[0] = 1
    assert np.allclose(sim.get_qs(), z)  # the simulator's state is untouched


@pytest.mark.parametrize("dtype", DTYPES)
def test_bell_pair(dtype):
    sim = mq.Simulator("mqvector_amd", 2, dtype=dtype)
    c = mq.Circuit().h(0).x(1, 0).measure(0).measure(1)
    shots = 100000
    res = sim.sampling(c, shots=shots, seed=11)
    assert set(res.data) <= {"00", "11"}
    assert abs(res.data.get("00", 0) - shots / 2) < 5 * np.sqrt(shots / 4)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n", [8, 15, 20])
def test_indices_match_lower_bound_of_the_state(n, dtype):
    circ = layered_random_circuit(n, 6, n)
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    sim.apply_circuit(circ)
    qs = sim.get_qs().astype(np.complex128)
    meas = mq.Circuit()
    for q in range(n):
        meas.measure(q)
    shots, seed = 20000, 5
    res = sim.sampling(meas, shots=shots, seed=seed)
    got = np.sort(full_index(res, n))
    cum = np.cumsum(np.abs(qs) ** 2)
    u = np.sort(np.random.default_rng(seed).random(shots))
    want = np.sort(np.searchsorted(cum, u * cum[-1], side="left"))
    assert np.mean(got != want) <= 1e-4


@pytest.mark.parametrize("dtype", DTYPES)
def test_marginals_against_oracle(dtype):
    """22 qubits, random circuit (parameters resolved), 3 measured qubits, 400k shots: total
    variation distance to the oracle's exact marginal below 0.01."""
    n = 22
    circ = layered_random_circuit(n, 5, 9)
    rng = np.random.default_rng(1)
    pc = mq.Circuit([mq.RY("a").on(3), mq.RZ("b").on(17, 3)])
    full = circ + pc
    vals = {"a": 0.7, "b": -1.3}
    for q in (3, 11, 17):
        full.measure(q)
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    res = sim.sampling(full, pr=vals, shots=400000, seed=int(rng.integers(1 << 20)))
    body = full.remove_measure()
    names = list(body.params_name)
    ref = cpu.apply_circuit(cpu.zero_state(n), ir.lower(body, names), np.array([vals[k] for k in names]))
    p = np.abs(ref) ** 2
    idx = np.arange(1 << n)
    marg = {}
    for bits in range(8):
        m = np.ones(1 << n, dtype=bool)
        for k, q in enumerate((17, 11, 3)):  # key order: last measurement first
            m &= ((idx >> q) & 1) == ((bits >> (2 - k)) & 1)
        marg[format(bits, "03b")] = float(p[m].sum())
    tv = 0.5 * sum(abs(res.data.get(k, 0) / res.shots - v) for k, v in marg.items())
    assert tv < 0.01


def test_errors():
    sim = mq.Simulator("mqvector_amd", 3)
    with pytest.raises(ValueError):
        sim.sampling(mq.Circuit().h(0), shots=10)  # no measurement
    with pytest.raises(ValueError):
        sim.sampling(mq.Circuit().measure(0).h(0).measure(1), shots=10)  # not measurement-ending
    with pytest.raises(ValueError):
        sim.sampling(mq.Circuit().h(0).measure(0), shots=0)
    with pytest.raises(ValueError):
        sim.apply_circuit(mq.Circuit().h(0).measure(0))
