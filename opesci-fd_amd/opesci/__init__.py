"""opesci-fd, MI355X backend: the StaggeredGrid elastic model description
(SField/VField/Media, DDerivative, Variable) and its generate / compile /
execute / convergence entry points, running on hand-written gfx950 kernels
in libopesci_hip.so (mirrors the reference package's opesci/__init__.py)."""
from sympy import Eq, Float, cos, pi, sin, sqrt, symbols  # noqa: F401  (reference re-exports)

from .derivative import *  # noqa: F401,F403
from .fields import *  # noqa: F401,F403
from .regulargrid import *  # noqa: F401,F403
from .staggeredgrid import *  # noqa: F401,F403
from .util import *  # noqa: F401,F403
from .variable import *  # noqa: F401,F403
from .emit import ccode  # noqa: F401
from .eigenwave import eigenwave3d  # noqa: F401

__version__ = '0.1.0'
