"""Riemann-solver registry (replaces clawpack.riemann's f2py modules).

The reference selects a solver by assigning a clawpack.riemann module to
solver.rp (e.g. apps/acoustics_2d_homogeneous/acoustics.py:25-26) and reads
num_eqn/num_waves from riemann.static (src/pyclaw/solver.py:206-211); the
solver's scalar parameters come from state.problem_data through the /cparam/
common block (state.py:162-182).  Here each solver is a small descriptor whose
`kind` selects the device implementation in libclawhip.
"""
from . import _lib


class RiemannSolver(object):
    def __init__(self, name, kind, num_eqn, num_waves, num_aux, cparam):
        self.__name__ = name
        self.kind = kind
        self.num_eqn = num_eqn
        self.num_waves = num_waves
        self.num_aux = num_aux
        self.cparam = tuple(cparam)

    def params(self, problem_data):
        """Pack problem_data into the C ABI's ClawRiemann.p array."""
        missing = [k for k in self.cparam if k not in problem_data]
        if missing:
            raise Exception("Some required value(s) in the cparam common block in the Riemann solver "
                            "have not been set in problem_data: %s" % ', '.join(missing))
        return [float(problem_data[k]) for k in self.cparam]

    def __repr__(self):
        return "<Riemann solver %s>" % self.__name__


rp2_acoustics = RiemannSolver('rp2_acoustics', _lib.RP_ACOUSTICS, 3, 2, 0, ('rho', 'bulk', 'cc', 'zz'))
rp2_vc_acoustics = RiemannSolver('rp2_vc_acoustics', _lib.RP_VC_ACOUSTICS, 3, 2, 2, ())
rp2_euler_5wave = RiemannSolver('rp2_euler_5wave', _lib.RP_EULER, 5, 5, 0, ('gamma', 'gamma1'))
rp2_shallow_roe_with_efix = RiemannSolver('rp2_shallow_roe_with_efix', _lib.RP_SHALLOW, 3, 3, 0, ('grav',))


class static(object):
    """Counterpart of clawpack.riemann.static."""
    num_eqn = {s.__name__: s.num_eqn for s in (rp2_acoustics, rp2_vc_acoustics, rp2_euler_5wave,
                                                 rp2_shallow_roe_with_efix)}
    num_waves = {s.__name__: s.num_waves for s in (rp2_acoustics, rp2_vc_acoustics, rp2_euler_5wave,
                                                     rp2_shallow_roe_with_efix)}
