"""ParameterResolver: an affine form ``const + sum_i coeff_i * name_i`` over named parameters.

Host-side mirror of mindquantum/core/parameterresolver/parameterresolver.py (the Python class
over the C++ resolver, ccsrc/lib/math/pr/parameter_resolver.cpp:481-489 Combination,
:546-572 jacobian).  Each name carries two flags, as in the reference: encoder vs ansatz
(as_encoder / as_ansatz, :677-700) and requires-grad vs no-grad (:993-1110).
"""
import numbers

import numpy as np

from .mutation import bump


class ParameterResolver:
    """Affine combination of named real parameters."""

    def __init__(self, data=None, const=None, dtype=None):
        self._data = {}
        self._const = 0.0
        self._encoder = set()
        self._no_grad = set()
        if isinstance(data, ParameterResolver):
            self._data = dict(data._data)
            self._const = data._const
            self._encoder = set(data._encoder)
            self._no_grad = set(data._no_grad)
        elif isinstance(data, str):
            self._data[data] = 1.0
        elif isinstance(data, dict):
            for k, v in data.items():
                if not isinstance(k, str):
                    raise TypeError(f"parameter name must be a string, but get {type(k)}")
                self._data[k] = float(np.real(v)) if np.isreal(v) else complex(v)
        elif isinstance(data, numbers.Number):
            self._const = data
        elif data is not None:
            raise TypeError(f"Unsupported parameter resolver data type {type(data)}")
        if const is not None:
            self._const = const

    # ---- properties
    @property
    def const(self):
        return self._const

    @const.setter
    def const(self, value):
        self._const = value
        bump()

    @property
    def params_name(self):
        return list(self._data)

    @property
    def encoder_parameters(self):
        return [k for k in self._data if k in self._encoder]

    @property
    def ansatz_parameters(self):
        return [k for k in self._data if k not in self._encoder]

    @property
    def requires_grad_parameters(self):
        return [k for k in self._data if k not in self._no_grad]

    @property
    def no_grad_parameters(self):
        return [k for k in self._data if k in self._no_grad]

    def keys(self):
        yield from self._data.keys()

    def values(self):
        yield from self._data.values()

    def items(self):
        yield from self._data.items()

    def __contains__(self, key):
        return key in self._data

    def __getitem__(self, key):
        return self._data[key]

    def __len__(self):
        return len(self._data)

    def is_const(self):
        return not self._data

    @property
    def is_complex(self):
        return isinstance(self._const, complex) or any(isinstance(v, complex) for v in self._data.values())

    # ---- flags (return self, like the reference)
    def as_encoder(self):
        self._encoder |= set(self._data)
        bump()
        return self

    def as_ansatz(self):
        self._encoder -= set(self._data)
        bump()
        return self

    def no_grad(self):
        self._no_grad |= set(self._data)
        bump()
        return self

    def requires_grad(self):
        self._no_grad -= set(self._data)
        bump()
        return self

    def no_grad_part(self, *names):
        for n in names:
            if n not in self._data:
                raise KeyError(f"Parameter {n} not in this parameter resolver!")
            self._no_grad.add(n)
        bump()
        return self

    def requires_grad_part(self, *names):
        for n in names:
            if n not in self._data:
                raise KeyError(f"Parameter {n} not in this parameter resolver!")
            self._no_grad.discard(n)
        bump()
        return self

    def encoder_part(self, *names):
        for n in names:
            self._encoder.add(n)
        bump()
        return self

    def ansatz_part(self, *names):
        for n in names:
            self._encoder.discard(n)
        bump()
        return self

    # ---- evaluation
    def combination(self, other):
        """Value for parameter values ``other`` (dict or ParameterResolver): const + sum c_i * v_i."""
        vals = other._data if isinstance(other, ParameterResolver) else dict(other)
        out = self._const
        for k, c in self._data.items():
            if k not in vals:
                raise ValueError(f"Parameter {k} not in given parameter resolver.")
            out = out + c * vals[k]
        return out

    def subs(self, other):
        return ParameterResolver(self.combination(other))

    # ---- arithmetic
    def _merge_flags(self, other, out):
        if isinstance(other, ParameterResolver):
            out._encoder |= other._encoder
            out._no_grad |= other._no_grad
        return out

    def __add__(self, other):
        out = ParameterResolver(self)
        if isinstance(other, ParameterResolver):
            for k, v in other._data.items():
                out._data[k] = out._data.get(k, 0.0) + v
            out._const = out._const + other._const
            return self._merge_flags(other, out)
        if isinstance(other, str):
            return self + ParameterResolver(other)
        out._const = out._const + other
        return out

    __radd__ = __add__

    def __neg__(self):
        out = ParameterResolver(self)
        out._data = {k: -v for k, v in self._data.items()}
        out._const = -self._const
        return out

    def __sub__(self, other):
        return self + (-ParameterResolver(other) if not isinstance(other, ParameterResolver) else -other)

    def __rsub__(self, other):
        return (-self) + other

    def __mul__(self, other):
        if isinstance(other, ParameterResolver):
            if other.is_const():
                other = other.const
            elif self.is_const():
                return other * self.const
            else:
                raise ValueError("product of two non-constant parameter resolvers is not affine")
        out = ParameterResolver(self)
        out._data = {k: v * other for k, v in self._data.items()}
        out._const = self._const * other
        return out

    __rmul__ = __mul__

    def __truediv__(self, other):
        return self * (1.0 / other)

    def __eq__(self, other):
        if not isinstance(other, ParameterResolver):
            return self.is_const() and self._const == other
        return self._data == other._data and self._const == other._const

    def __hash__(self):
        return id(self)

    def expression(self):
        parts = [f"{v}*{k}" if v != 1 else k for k, v in self._data.items()]
        if self._const or not parts:
            parts.append(str(self._const))
        return " + ".join(parts)

    def __repr__(self):
        return f"ParameterResolver({self.expression()})"


PR = ParameterResolver
