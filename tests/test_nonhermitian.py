"""Pins of the dense non-hermitian gradient (oracle/dense.py nonhermitian_gradient), the checker
of the GPU's circ_left / simulator_left gradient: the reference's closed form
(test_simulator.py:199-224) and complex central finite differences on random circuit pairs."""
import numpy as np

import mindquantum_amd as mq
from oracle import dense
from tests.test_gate_families import family_circuit


def test_dense_matches_reference_closed_form():
    cr = mq.Circuit().ry("b", 0)
    cl = mq.Circuit().rz("a", 0)
    f, g = dense.nonhermitian_gradient(cr, cl, {"b": 1.2, "a": 2.3}, ["b", "a"], 1, mq.QubitOperator(""))
    f_exp = np.exp(1j * 2.3 / 2) * np.cos(1.2 / 2)
    g_exp = [-0.5 * np.exp(1j * 2.3 / 2) * np.sin(1.2 / 2), 1j / 2 * np.exp(1j * 2.3 / 2) * np.cos(1.2 / 2)]
    assert np.allclose(f, f_exp, atol=1e-14) and np.allclose(g, g_exp, atol=1e-14)


def test_dense_matches_central_differences():
    n = 4
    cr = family_circuit(n, 14, 1)
    cl = family_circuit(n, 10, 2)
    names = list(cr.params_name) + [k for k in cl.params_name if k not in cr.params_name]
    vals = np.random.default_rng(3).uniform(0, 2 * np.pi, size=len(names))
    op = mq.QubitOperator("Z0 X1") + 0.5 * mq.QubitOperator("Y2")
    rng = np.random.default_rng(4)
    init_l = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    init_l /= np.linalg.norm(init_l)
    f, g = dense.nonhermitian_gradient(cr, cl, dict(zip(names, vals)), names, n, op, init_l=init_l)
    h = 1e-6
    for j in range(len(names)):
        vp, vm = vals.copy(), vals.copy()
        vp[j] += h
        vm[j] -= h
        fp, _ = dense.nonhermitian_gradient(cr, cl, dict(zip(names, vp)), [], n, op, init_l=init_l)
        fm, _ = dense.nonhermitian_gradient(cr, cl, dict(zip(names, vm)), [], n, op, init_l=init_l)
        assert abs((fp - fm) / (2 * h) - g[j]) < 1e-7
