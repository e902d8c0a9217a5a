"""Helpers of the simulator API (mindquantum/simulator/utils.py:18-76)."""


def _thread_balance(n_prs, n_meas, parallel_worker):
    """Split of ``parallel_worker`` threads between samples and Hamiltonians.

    Same arithmetic as the reference (utils.py:18-32).  The HIP backend runs every sample of a
    batch inside one launch, so the split only validates the argument and is reported back.
    """
    if parallel_worker is None:
        parallel_worker = n_meas * n_prs
    if parallel_worker <= 0:
        raise ValueError("parallel_worker must be positive")
    if n_meas * n_prs <= parallel_worker:
        batch_threads = n_prs
        mea_threads = n_meas
    else:
        if n_meas < n_prs:
            batch_threads = min(n_prs, parallel_worker)
            mea_threads = min(n_meas, max(1, parallel_worker // batch_threads))
        else:
            mea_threads = min(n_meas, parallel_worker)
            batch_threads = min(n_prs, max(1, parallel_worker // mea_threads))
    return batch_threads, mea_threads


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
