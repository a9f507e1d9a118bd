"""Placeholder derivative symbols (mirrors opesci/derivative.py:6-26)."""
from sympy import Symbol

__all__ = ['DDerivative']


class DDerivative(Symbol):
    """Symbol standing for d^order F / d var^order inside a user's PDE.

    ``fd[accuracy]`` holds the finite-difference expression of that accuracy;
    ``StaggeredGrid.solve_fd`` substitutes ``fd[max_accuracy]``."""

    def __new__(cls, name, var=None, order=1, max_accuracy=2):
        obj = Symbol.__xnew__(cls, name)
        obj.var = var
        obj.order = order
        obj.max_accuracy = max_accuracy
        obj.fd = [None] * (max_accuracy + 1)
        return obj

    def __getnewargs__(self):
        return (self.name, self.var, self.order, self.max_accuracy)
