"""torch.autograd operators over the adjoint gradient (mindquantum/framework/operations.py)."""
from .operations import MQAnsatzOnlyLayer, MQAnsatzOnlyOps, MQLayer, MQOps

__all__ = ["MQOps", "MQAnsatzOnlyOps", "MQLayer", "MQAnsatzOnlyLayer"]
