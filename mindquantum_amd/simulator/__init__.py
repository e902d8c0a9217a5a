"""Simulator API (mindquantum.simulator) backed by the MI355X state-vector engine."""
from .backend import MQVectorAMD, inner_product
from .batch import expectation_with_grad_batch
from .simulator import SUPPORTED_SIMULATOR, Simulator
from .utils import GradOpsWrapper

__all__ = ["Simulator", "MQVectorAMD", "inner_product", "GradOpsWrapper", "SUPPORTED_SIMULATOR",
           "expectation_with_grad_batch"]
