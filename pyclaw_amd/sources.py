"""Source terms as device operations.

The Euler equations with cylindrical (radial) symmetry: the geometric source
terms of the shock-bubble application (apps/euler_2d/shockbubble.py:89-153),
as HIP kernels on the device-resident state instead of numpy on a host copy
(claw2d_euler_radial_step / claw2d_euler_radial_dq).  Drop-in replacements for
the application's routines:

    solver.step_source = pyclaw_amd.sources.euler_radial_step   # ClawSolver2D
    solver.dq_src = pyclaw_amd.sources.euler_radial_dq           # SharpClawSolver2D

The radius is aux component 0 (the cell centre's y), gamma - 1 comes from the
state's problem_data ('gamma1', else 'gamma' - 1).  FP64 results equal the
reference's numpy arithmetic bit for bit.
"""
import torch

from . import _lib
from .state import NG_STORE


def _grid(state):
    g = _lib.ClawGrid()
    g.mx, g.my = state.patch.num_cells_global
    g.ng = NG_STORE
    buf = state._q.buf
    g.ldq = buf.shape[2]
    g.plane = buf.shape[1] * buf.shape[2]
    g.dx, g.dy = state.grid.delta
    return g


def _gamma1(state):
    pd = state.problem_data
    if 'gamma1' in pd:
        return float(pd['gamma1'])
    if 'gamma' in pd:
        return float(pd['gamma']) - 1.0
    raise ValueError("problem_data needs 'gamma1' or 'gamma' for the Euler source terms")


def _prepare(state):
    if state.num_aux < 1:
        raise ValueError("the radial source terms read the radius from aux component 0")
    if state.num_eqn < 4:
        raise ValueError("the radial source terms need the four Euler components")
    state.sync_to_device()
    L = _lib.lib()
    for name in ("claw2d_euler_radial_step", "claw2d_euler_radial_dq"):
        if not hasattr(L, name):
            raise _lib.ClawError("libclawhip.so lacks %s: rebuild it" % name)
    buf = state._q.buf
    prec = _lib.F32 if buf.dtype == torch.float32 else _lib.F64
    return L, prec, torch.cuda.current_stream(buf.device).cuda_stream


def euler_radial_step(solver, state, dt):
    """step_Euler_radial (apps/euler_2d/shockbubble.py:89-124): one 2-stage
    Runge-Kutta step of the radial source terms, in place on the device."""
    L, prec, st = _prepare(state)
    rc = L.claw2d_euler_radial_step(_grid(state), prec, state.num_eqn, state._q.buf.data_ptr(),
                                    state._aux.buf.data_ptr(), float(dt), _gamma1(state), st)
    _lib.check(rc, "claw2d_euler_radial_step")
    state._device_updated()


def euler_radial_dq(solver, state, dt):
    """dq_Euler_radial (apps/euler_2d/shockbubble.py:127-153): dt times the
    radial source terms, returned as a device tensor (num_eqn, mx, my)."""
    L, prec, st = _prepare(state)
    dq = torch.empty_like(state._q.buf)
    rc = L.claw2d_euler_radial_dq(_grid(state), prec, state.num_eqn, state._q.buf.data_ptr(),
                                  state._aux.buf.data_ptr(), dq.data_ptr(), float(dt), _gamma1(state), st)
    _lib.check(rc, "claw2d_euler_radial_dq")
    return state._q.interior(dq)
