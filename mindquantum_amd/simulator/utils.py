"""Helpers of the simulator API (mindquantum/simulator/utils.py:18-76)."""
import numbers


def check_parallel_worker(parallel_worker):
    """``parallel_worker`` of get_expectation_with_grad: None or a positive int.  The reference
    splits that many host threads between samples and Hamiltonians; here every sample of a batch
    runs inside the same device launches, so the value is validated and otherwise unused."""
    if parallel_worker is None:
        return
    if isinstance(parallel_worker, bool) or not isinstance(parallel_worker, numbers.Integral):
        raise TypeError(f"parallel_worker requires a int, but get {type(parallel_worker)}")
    if parallel_worker <= 0:
        raise ValueError("parallel_worker must be positive")


class GradOpsWrapper:  # pylint: disable=too-many-instance-attributes
    """Gradient operator plus the information that generated it (utils.py:35-76)."""

    def __init__(self, grad_ops, hams, circ_right, circ_left, encoder_params_name, ansatz_params_name,
                 parallel_worker, sim=None):
        self.grad_ops = grad_ops
        self.hams = hams
        self.circ_right = circ_right
        self.circ_left = circ_left
        self.encoder_params_name = encoder_params_name
        self.ansatz_params_name = ansatz_params_name
        self.parallel_worker = parallel_worker
        self.str = ""
        self.sim = sim

    def __call__(self, *args):
        return self.grad_ops(*args)

    def set_str(self, grad_str):
        self.str = grad_str
