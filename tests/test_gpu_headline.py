"""The benchmarked 20-qubit workload at the size it is timed, against the complex128 oracle.

Encoder RX(x_i) on 20 qubits + HEA-8 (RY, RZ on every qubit, CNOT ladder; 320 ansatz parameters),
H = sum_i w_i Z_i Z_{i+1 mod 20} + sum_i v_i X_i (40 terms), batch 1024, seed 2024 - exactly
bench.py's workload (mindquantum_amd.algorithm.batched_adjoint) - through the torch.autograd operator
(MQOps) on compiled passes, at the forward / back-sweep tile widths the bench's tuner takes (13 / 12),
at 12 / 12, and at 13 / 13 (a width the library's tuner may take).  Sixteen sampled rows of f, g_enc and g_ans are compared with the complex128 oracle
(gradient.h:100-197 restated; framework/operations.py:132-146 for the operator):

* complex64: |f - f_ref| <= 1e-5 * sum |c|, max |g - g_ref| / max |g_ref| <= 1e-5;
* complex128: the same with 1e-10.

The ansatz gradient that MQOps hands to torch (sum over the batch of g_ans weighted by dout) is
checked against the per-sample gradients of the same call.
"""
import numpy as np
import pytest
import torch

import mindquantum_amd as mq
from mindquantum_amd import _native, ir
from mindquantum_amd.algorithm import batched_adjoint
from mindquantum_amd.framework import MQOps
from oracle import cpu
from tests.bars import assert_f, assert_g, coeff_sum

pytestmark = pytest.mark.gpu

ROWS = np.array([0, 1, 2, 3, 97, 255, 256, 400, 511, 512, 640, 777, 900, 1000, 1022, 1023])


@pytest.fixture(scope="module")
def workload():
    circ, op, enc, ans = batched_adjoint()
    names = list(circ.encoder_params_name) + list(circ.ansatz_params_name)
    f_ref, g_ref = cpu.grad_batch(20, np.complex128, ir.lower(circ, names), [mq.Hamiltonian(op).terms_ir()],
                                  enc[ROWS], ans, batch_threads=len(ROWS))
    return circ, op, enc, ans, f_ref[:, 0], g_ref[:, 0]


@pytest.fixture
def compiled():
    saved = {k: _native.get_option(k) for k in ("jit",)}
    _native.set_option("jit", 2)
    yield
    for k, v in saved.items():
        _native.set_option(k, v)


@pytest.mark.parametrize("dtype,tiles", [(mq.complex64, (13, 13)), (mq.complex64, (13, 12)), (mq.complex64, (12, 12)),
                                         (mq.complex128, (12, 12))])
def test_headline_workload_against_c128_oracle(workload, compiled, dtype, tiles):
    circ, op, enc, ans, f_ref, g_ref = workload
    n_enc = enc.shape[1]
    sim = mq.Simulator("mqvector_amd", 20, dtype=dtype)
    grad_ops = sim.get_expectation_with_grad(mq.Hamiltonian(op, dtype), circ)
    code = mq.dtype.native_code(dtype)
    nc = grad_ops.native_circuits[0]
    nc.set_tile_qubits(code, False, tiles[0])
    nc.set_tile_qubits(code, True, tiles[1])
    before = _native.get_option("jit_launches")

    ops = MQOps(grad_ops)
    x = torch.tensor(enc, dtype=torch.float64, requires_grad=True)
    a = torch.tensor(ans, dtype=torch.float64, requires_grad=True)
    out = ops(x, a)
    out.sum().backward()
    assert _native.get_option("jit_launches") > before, "compiled passes did not run"
    assert out.shape == (1024, 1) and x.grad.shape == (1024, n_enc) and a.grad.shape == (len(ans),)

    # per-sample gradients of the same operator (the values MQOps caches)
    f, g_enc, g_ans = grad_ops(enc, ans)
    csum = coeff_sum(op)
    err_f = assert_f(f[ROWS, 0], f_ref, dtype, csum)
    err_g = assert_g(np.concatenate([g_enc[ROWS, 0], g_ans[ROWS, 0]], axis=1).real, g_ref.real, dtype)
    print(f"{np.dtype(dtype).name} tiles {tiles}: |f - f_ref| = {err_f:.2e}, max|g - g_ref| = {err_g:.2e} "
          f"(max|g_ref| = {np.max(np.abs(g_ref)):.3f})")
    # what MQOps returned and handed to torch (float32 values, dout = 1)
    assert np.max(np.abs(out.detach().numpy()[:, 0] - f[:, 0].real)) <= 1e-6 * max(1.0, csum)
    assert np.max(np.abs(x.grad.numpy() - g_enc[:, 0].real)) <= 1e-6 * max(1.0, float(np.max(np.abs(g_enc))))
    ga_sum = np.sum(g_ans[:, 0].real.astype(np.float64), axis=0)
    assert np.max(np.abs(a.grad.numpy() - ga_sum)) <= 1e-4 * max(1.0, float(np.max(np.abs(ga_sum))))
