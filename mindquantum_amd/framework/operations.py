"""torch.autograd wrappers of ``Simulator.get_expectation_with_grad``.

Same semantics as the reference MindSpore cells (framework/operations.py:67-147, 237-307):
the forward pass runs the whole adjoint sweep on the GPU (f and the Jacobian come back
together), caches Re(g), and returns Re(f) as float32; the backward pass contracts the cached
Jacobian with the incoming gradient: enc_grad = einsum('smp,sm->sp'), ans_grad =
einsum('smp,sm->p').  PyTorch is only the autograd plumbing here.
"""
import numpy as np
import torch
from torch import nn


class _MQOpsFunction(torch.autograd.Function):  # pylint: disable=abstract-method
    @staticmethod
    def forward(ctx, enc, ans, grad_ops):  # pylint: disable=arguments-differ
        f, g_enc, g_ans = grad_ops(enc.detach().cpu().numpy().astype(np.float64),
                                   ans.detach().cpu().numpy().astype(np.float64))
        ctx.save_for_backward(torch.from_numpy(np.real(g_enc).astype(np.float32)),
                              torch.from_numpy(np.real(g_ans).astype(np.float32)))
        ctx.devices = (enc.device, ans.device)
        return torch.from_numpy(np.real(f).astype(np.float32)).to(enc.device)

    @staticmethod
    def backward(ctx, dout):  # pylint: disable=arguments-differ
        g_enc, g_ans = ctx.saved_tensors
        dout = dout.detach().float().cpu()
        enc_grad = torch.einsum("smp,sm->sp", g_enc, dout).to(ctx.devices[0])
        ans_grad = torch.einsum("smp,sm->p", g_ans, dout).to(ctx.devices[1])
        return enc_grad, ans_grad, None


class _MQAnsatzOnlyFunction(torch.autograd.Function):  # pylint: disable=abstract-method
    @staticmethod
    def forward(ctx, ans, grad_ops):  # pylint: disable=arguments-differ
        f, g = grad_ops(ans.detach().cpu().numpy().astype(np.float64))
        ctx.save_for_backward(torch.from_numpy(np.real(g[0]).astype(np.float32)))
        ctx.device = ans.device
        return torch.from_numpy(np.real(f[0]).astype(np.float32)).to(ans.device)

    @staticmethod
    def backward(ctx, dout):  # pylint: disable=arguments-differ
        (g,) = ctx.saved_tensors
        return (dout.detach().float().cpu() @ g).to(ctx.device), None


class MQOps(nn.Module):
    """f(enc[b, E], ans[A]) -> [b, n_hams] with gradients to both inputs."""

    def __init__(self, expectation_with_grad):
        super().__init__()
        self.expectation_with_grad = expectation_with_grad

    def forward(self, enc, ans):
        return _MQOpsFunction.apply(enc, ans, self.expectation_with_grad)


class MQAnsatzOnlyOps(nn.Module):
    """f(ans[A]) -> [n_hams] for circuits without encoder parameters."""

    def __init__(self, expectation_with_grad):
        super().__init__()
        self.expectation_with_grad = expectation_with_grad

    def forward(self, ans):
        return _MQAnsatzOnlyFunction.apply(ans, self.expectation_with_grad)


class MQLayer(nn.Module):
    """MQOps with the ansatz weights as a trainable parameter (framework/layer.py)."""

    def __init__(self, expectation_with_grad, weight=None):
        super().__init__()
        n = len(expectation_with_grad.ansatz_params_name)
        w = torch.zeros(n) if weight is None else torch.as_tensor(weight, dtype=torch.float32)
        self.weight = nn.Parameter(w)
        self.ops = MQOps(expectation_with_grad)

    def forward(self, enc):
        return self.ops(enc, self.weight)


class MQAnsatzOnlyLayer(nn.Module):
    def __init__(self, expectation_with_grad, weight=None):
        super().__init__()
        n = len(expectation_with_grad.ansatz_params_name)
        w = torch.zeros(n) if weight is None else torch.as_tensor(weight, dtype=torch.float32)
        self.weight = nn.Parameter(w)
        self.ops = MQAnsatzOnlyOps(expectation_with_grad)

    def forward(self):
        return self.ops(self.weight)
