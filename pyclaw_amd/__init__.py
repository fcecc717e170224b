"""pyclaw_amd: PyClaw's 2D time step on AMD MI355X (gfx950).

Drop-in for the reference's 2D API (Controller, Dimension, Patch, Domain,
Solution, State, CFL, ClawSolver2D, SharpClawSolver2D, BC, limiters, and a
riemann registry), with q/aux on the GPU and the hot path in hand-written HIP
kernels (libclawhip.so, C ABI in include/clawhip.h).
"""
from . import limiters
from . import riemann
from . import sources
from .cfl import CFL
from .classic import ClawSolver2D
from .controller import Controller
from .geometry import Dimension, Domain, Grid, Patch
from .solution import Solution
from .solver import BC, CFLError
from .state import State

try:
    from .sharpclaw import SharpClawSolver2D
except ImportError:  # pragma: no cover
    SharpClawSolver2D = None

__all__ = ['Controller', 'Dimension', 'Patch', 'Domain', 'Grid', 'Solution', 'State', 'CFL',
           'ClawSolver2D', 'SharpClawSolver2D', 'BC', 'CFLError', 'limiters', 'riemann', 'sources']
