"""Circuit, gate, operator and parameter-resolver objects (mindquantum.core)."""
from . import gates
from .circuit import UN, Circuit
from .measurement import Measure, MeasureResult
from .gates import *  # noqa: F401,F403
from .operators import Hamiltonian, QubitOperator
from .parameterresolver import PR, ParameterResolver
