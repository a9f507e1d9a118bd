"""Named constants/variables of a model (mirrors opesci/variable.py:6-21)."""
from sympy import Symbol

__all__ = ['Variable']


class Variable(Symbol):
    """A SymPy symbol that also carries the C type, value and constness the
    generator needs (dims, spacing, dt, ntsteps, media constants)."""

    def __new__(cls, name, value=0, type='int', constant=False):
        obj = Symbol.__xnew__(cls, name)
        obj.type = type
        obj.constant = constant
        obj.value = value
        return obj

    def __getnewargs__(self):
        return (self.name, self.value, self.type, self.constant)
