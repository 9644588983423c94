"""Circuit parameters.

Mirrors quafu/elements/parameters.py:25-286: a `ParameterExpression` is a pivot `Parameter`
followed by a chain of (function, operand) steps; `_variables` orders the pivot first, then the
operands' variables in order of appearance; `grad()` returns d(value)/d(variable) in that order.
The reference differentiates with the `autograd` package (parameters.py:92-96); here the same
chain is evaluated with forward-mode dual numbers, which gives the same exact derivative.
"""

from __future__ import annotations

import copy
import math
import operator
from typing import Union

import numpy as np


class _Dual:
    """value + sum_k d[k] * eps_k (first-order forward mode)."""

    __slots__ = ("v", "d")

    def __init__(self, v, d):
        self.v = float(v)
        self.d = np.asarray(d, dtype=float)

    @staticmethod
    def lift(x, n):
        return x if isinstance(x, _Dual) else _Dual(x, np.zeros(n))

    def __add__(self, o):
        o = _Dual.lift(o, len(self.d))
        return _Dual(self.v + o.v, self.d + o.d)

    __radd__ = __add__

    def __sub__(self, o):
        o = _Dual.lift(o, len(self.d))
        return _Dual(self.v - o.v, self.d - o.d)

    def __rsub__(self, o):
        return _Dual.lift(o, len(self.d)) - self

    def __mul__(self, o):
        o = _Dual.lift(o, len(self.d))
        return _Dual(self.v * o.v, self.d * o.v + o.d * self.v)

    __rmul__ = __mul__

    def __truediv__(self, o):
        o = _Dual.lift(o, len(self.d))
        return _Dual(self.v / o.v, (self.d * o.v - o.d * self.v) / (o.v * o.v))

    def __rtruediv__(self, o):
        return _Dual.lift(o, len(self.d)) / self

    def __pow__(self, k):
        if isinstance(k, _Dual):
            raise NotImplementedError
        return _Dual(self.v**k, self.d * (k * self.v ** (k - 1)))

    def __neg__(self):
        return _Dual(-self.v, -self.d)


def _unary(f, df):
    def g(x):
        if isinstance(x, _Dual):
            return _Dual(f(x.v), x.d * df(x.v))
        return f(x)

    g.__name__ = f.__name__
    return g


sin = _unary(math.sin, math.cos)
cos = _unary(math.cos, lambda x: -math.sin(x))
tan = _unary(math.tan, lambda x: 1.0 / math.cos(x) ** 2)
arcsin = _unary(math.asin, lambda x: 1.0 / math.sqrt(1.0 - x * x))
arccos = _unary(math.acos, lambda x: -1.0 / math.sqrt(1.0 - x * x))
arctan = _unary(math.atan, lambda x: 1.0 / (1.0 + x * x))
exp = _unary(math.exp, math.exp)
log = _unary(math.log, lambda x: 1.0 / x)


class ParameterExpression:
    def __init__(self, pivot: "Parameter", value=0.0, operands=None, funcs=None, latex=""):
        self.pivot = pivot
        self.value = value
        self.operands = [] if operands is None else operands
        self.funcs = [] if funcs is None else funcs
        self.latex = latex

    @property
    def _variables(self):
        variables = {self.pivot: 0}
        vi = 1
        for o in self.operands:
            if isinstance(o, Parameter):
                if o not in variables:
                    variables[o] = vi
                    vi += 1
            elif isinstance(o, ParameterExpression):
                for v in o._variables:
                    if v not in variables:
                        variables[v] = vi
                        vi += 1
        return variables

    def _undo(self, step):
        for _ in range(step):
            if len(self.operands) > 0:
                self.operands.pop()
                self.funcs.pop()
            else:
                return self.pivot
        return self

    @property
    def _func(self):
        variables = self._variables

        def __func(x):
            z = x[0]
            for i, f in enumerate(self.funcs):
                op = self.operands[i]
                if op is None:
                    z = f(z)
                elif isinstance(op, (float, int)):
                    z = f(z, op)
                elif isinstance(op, Parameter):
                    z = f(z, x[variables[op]])
                elif isinstance(op, ParameterExpression):
                    opvars = op._variables.keys()
                    z = f(z, op._func([x[variables[ov]] for ov in opvars]))
                else:
                    raise NotImplementedError
            return z

        return __func

    def __repr__(self):
        return str(self.get_value())

    def grad(self, input_value=None):
        variables = list(self._variables)
        if input_value is None or len(input_value) == 0:
            input_value = [v.value for v in variables]
        n = len(variables)
        x = [_Dual(input_value[i], np.eye(n)[i]) for i in range(n)]
        z = self._func(x)
        if not isinstance(z, _Dual):
            return np.zeros(n)
        return z.d

    def get_value(self):
        value = self._func([v.value for v in self._variables])
        self.value = value
        return value

    def _chain(self, func, operand, value):
        operands = list(self.operands)
        funcs = copy.copy(self.funcs)
        operands.append(operand)
        funcs.append(func)
        return ParameterExpression(self.pivot, value, operands, funcs)

    @staticmethod
    def _val(r):
        if isinstance(r, (Parameter, ParameterExpression)):
            return r.value
        if isinstance(r, (float, int)):
            return r
        raise NotImplementedError

    def __add__(self, r):
        return self._chain(operator.add, r, self.value + self._val(r))

    def __radd__(self, r):
        return self + r

    def __mul__(self, r):
        return self._chain(operator.mul, r, self.value * self._val(r))

    def __rmul__(self, r):
        return self * r

    def __neg__(self):
        return -1.0 * self

    def __sub__(self, r):
        return self._chain(operator.sub, r, self.value - self._val(r))

    def __rsub__(self, r):
        # r - self. (The reference's `return r - self` recurses without end; this is the
        # intended value.)
        return (-1.0 * self) + r

    def __truediv__(self, r):
        return self._chain(operator.truediv, r, self.value / self._val(r))

    def __rtruediv__(self, r):
        # r / self, written as self**-1 * r (the reference's version recurses, as __rsub__).
        return self._chain(operator.pow, -1, self.value**-1) * r

    def __pow__(self, n):
        if not isinstance(n, (float, int)):
            raise NotImplementedError
        return self._chain(operator.pow, n, self.value**n)

    def sin(self):
        return self._chain(sin, None, math.sin(self.value))

    def cos(self):
        return self._chain(cos, None, math.cos(self.value))

    def tan(self):
        return self._chain(tan, None, math.tan(self.value))

    def arcsin(self):
        return self._chain(arcsin, None, math.asin(self.value))

    def arccos(self):
        return self._chain(arccos, None, math.acos(self.value))

    def arctan(self):
        return self._chain(arctan, None, math.atan(self.value))

    def exp(self):
        return self._chain(exp, None, math.exp(self.value))

    def log(self):
        return self._chain(log, None, math.log(self.value))


class Parameter(ParameterExpression):
    # pylint: disable=super-init-not-called
    def __init__(self, name, value: float = 0.0, tunable: bool = True):
        self.name = name
        self.value = float(value)
        self.operands = []
        self.funcs = []
        self.latex = self.name
        self.tunable = tunable

    @property
    def pivot(self):
        return self

    def grad(self, input_value=None):  # pylint: disable=arguments-differ
        return np.array([1.0])

    def get_value(self):
        return self.value

    def __hash__(self):
        return hash(self.name)

    def __eq__(self, other):
        return isinstance(other, Parameter) and self.name == other.name

    def __repr__(self):
        return f"{self.name}({self.value})"


ParameterType = Union[float, Parameter, ParameterExpression]


def extract_float(paras):
    """Numeric values of a parameter list (quafu/elements/utils.py:32-41)."""
    if not isinstance(paras, (list, tuple)):
        paras = [paras]
    out = []
    for para in paras:
        if isinstance(para, (float, int, np.floating, np.integer)):
            out.append(para)
        elif isinstance(para, (Parameter, ParameterExpression)):
            out.append(para.get_value())
    return out
