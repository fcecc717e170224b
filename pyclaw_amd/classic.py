"""Classic Clawpack solver in 2D on the GPU.

API of the reference's ClawSolver / ClawSolver2D
(src/pyclaw/classic/solver.py:19-547): limiters, order, source_split, fwave,
step_source, before_step, cfl_max, cfl_desired, dimensional_split,
transverse_waves (no_trans / trans_inc / trans_cor), step(), setup(),
teardown().  step_hyperbolic calls the fused HIP kernel through
claw2d_classic_step instead of the f2py step2/step2ds.
"""
import ctypes
import warnings

import torch

from . import _lib
from .limiters import tvd
from .solver import Solver


class ClawSolver(Solver):
    def __init__(self, riemann_solver=None, claw_package=None):
        self.num_ghost = 2
        self.limiters = tvd.minmod
        self.order = 2
        self.source_split = 1
        self.fwave = False
        self.step_source = None
        self.before_step = None
        self.kernel_language = 'Fortran'
        self.verbosity = 0
        self.cfl_max = 1.0
        self.cfl_desired = 0.9
        self._mthlim = self.limiters
        self._method = None
        super(ClawSolver, self).__init__(riemann_solver, claw_package)

    def step(self, solution):
        """classic/solver.py:100-151."""
        if self.before_step is not None:
            self.before_step(self, solution.states[0])
        if self.source_split == 2 and self.step_source is not None:
            self.step_source(self, solution.states[0], self.dt / 2.0)
        self.step_hyperbolic(solution)
        if self.cfl.get_cached_max() >= self.cfl_max:
            return False
        if self.step_source is not None:
            if self.source_split == 2:
                self.step_source(self, solution.states[0], self.dt / 2.0)
            if self.source_split == 1:
                self.step_source(self, solution.states[0], self.dt)
        return True

    def _hooks_active(self):
        return self.before_step is not None or self.step_source is not None

    def check_cfl_settings(self):
        pass

    def set_mthlim(self):
        """classic/solver.py:167-176."""
        self._mthlim = self.limiters
        if not isinstance(self.limiters, list):
            self._mthlim = [self._mthlim]
        if len(self._mthlim) == 1:
            self._mthlim = self._mthlim * self.num_waves
        if len(self._mthlim) != self.num_waves:
            raise Exception('Length of solver.limiters is not equal to 1 or to solver.num_waves')

    def step_hyperbolic(self, solution):
        raise Exception("Dummy routine, please override!")


class ClawSolver2D(ClawSolver):
    no_trans = 0
    trans_inc = 1
    trans_cor = 2

    def __init__(self, riemann_solver=None, claw_package=None):
        self.dimensional_split = True
        self.transverse_waves = self.trans_inc
        self.num_dim = 2
        self.aux1 = None
        self.aux2 = None
        self.aux3 = None
        self.work = None
        self._method_struct = None
        self._rp_struct = None
        # FP64 states only: use the kernels built with fused multiply-add
        # (CLAW_F64_FMA).  Faster; each step agrees with the reference
        # arithmetic to rounding (not bit for bit).
        self.fma = False
        super(ClawSolver2D, self).__init__(riemann_solver, claw_package)

    def check_cfl_settings(self):
        if (not self.dimensional_split) and (self.transverse_waves == 0):
            cfl_recommended = 0.5
        else:
            cfl_recommended = 1.0
        if self.cfl_max > cfl_recommended:
            warnings.warn('cfl_max is set higher than the recommended value of %s' % cfl_recommended)
            warnings.warn(str(self.cfl_desired))

    def setup(self, solution):
        """classic/solver.py:200-224: validate options and move the state to
        the GPU."""
        state = solution.states[0]
        state.set_num_ghost(self.num_ghost)
        self.check_cfl_settings()
        if self.rp is None:
            raise Exception("solver.rp (the Riemann solver) has not been set")
        if self.num_waves is None:
            self.num_waves = self.rp.num_waves
        if self.num_waves != self.rp.num_waves:
            raise Exception("solver.num_waves (%d) does not match %s (%d waves)"
                            % (self.num_waves, self.rp.__name__, self.rp.num_waves))
        if state.num_eqn != self.rp.num_eqn:
            raise Exception("state has %d equations, %s needs %d" % (state.num_eqn, self.rp.__name__,
                                                                   self.rp.num_eqn))
        if self.rp.num_aux > state.num_aux:
            raise Exception("%s needs %d aux fields" % (self.rp.__name__, self.rp.num_aux))
        if self.fwave:
            raise NotImplementedError("f-wave splitting needs an f-wave Riemann solver; none is provided")
        if state.index_capa >= 0:
            raise NotImplementedError("capacity functions (index_capa) are not supported")
        if self.order not in (1, 2):
            raise Exception("order must be 1 or 2")
        self.set_mthlim()
        m = _lib.ClawMethod()
        m.order = self.order
        m.transverse = 0 if self.dimensional_split else int(self.transverse_waves)
        m.split = 1 if self.dimensional_split else 0
        m.fwave = 0
        m.num_waves = self.num_waves
        for k, v in enumerate(self._mthlim):
            m.mthlim[k] = int(v)
        self._method_struct = m
        r = _lib.ClawRiemann()
        r.kind = self.rp.kind
        r.efix = 1
        for k, v in enumerate(self.rp.params(state.problem_data)):
            r.p[k] = v
        self._rp_struct = r
        self._device_setup(state)
        self._is_set_up = True

    def teardown(self):
        self._ctx = None

    def _precision(self, state):
        return _lib.F32 if state.dtype == torch.float32 else (_lib.F64_FMA if self.fma else _lib.F64)

    def _autotune_once(self, state, grid, bc, prec, aux, stream):
        key = (grid.mx, grid.my, prec, self._method_struct.order, self._method_struct.transverse,
               self._method_struct.split, self._rp_struct.kind)
        if self.autotune and key not in self._tuned:
            band = ctypes.c_int(0)
            rc = _lib.lib().claw2d_classic_autotune(self._ctx.handle, grid, self._method_struct, self._rp_struct,
                                                    bc, prec, state._q.buf.data_ptr(), self._qalt.data_ptr(), aux,
                                                    self._dt_dev.data_ptr(), self._cfl_dev.data_ptr(), stream,
                                                    ctypes.byref(band))
            _lib.check(rc, "claw2d_classic_autotune")
            self._tuned.add(key)

    def _loop_launcher(self, state):
        """Iterations of the device step loop: the fused kernel(s) of
        step_hyperbolic reading dt from the device and accumulating the Courant
        number there, each followed by the loop's bookkeeping, all launched by
        claw2d_classic_loop."""
        if not self._is_set_up or self._hooks_active():
            return None
        state.sync_to_device()
        bc = self._kernel_bcs(state)
        grid = self._grid_struct(state)
        prec = self._precision(state)
        aux = state._aux.buf.data_ptr() if state._aux is not None else None
        if self.autotune:
            self._dt_dev.fill_(float(self.dt))
            self._autotune_once(state, grid, bc, prec, aux, torch.cuda.current_stream().cuda_stream)
        fn = _lib.lib().claw2d_classic_loop
        ctx, m, r = self._ctx.handle, self._method_struct, self._rp_struct

        def launch(k0, niter, qa, qb, rule, states, cfl, nbytes, stream, gauges=None):
            rc = fn(ctx, grid, m, r, bc, prec, qa, qb, aux, rule, states, cfl, nbytes, k0, niter, gauges, stream)
            _lib.check(rc, "claw2d_classic_loop")
        return launch

    def step_hyperbolic(self, solution):
        """classic/solver.py:498-547 on the GPU: ghost handling, one fused
        kernel (unsplit) or two passes (split), Courant number back to the host."""
        state = solution.states[0]
        if not self._is_set_up:
            self.setup(solution)
        state.sync_to_device()
        bc = self._kernel_bcs(state)
        grid = self._grid_struct(state)
        prec = self._precision(state)
        self._dt_dev.fill_(float(self.dt))
        aux = state._aux.buf.data_ptr() if state._aux is not None else None
        stream = torch.cuda.current_stream().cuda_stream
        self._autotune_once(state, grid, bc, prec, aux, stream)
        rc = _lib.lib().claw2d_classic_step(self._ctx.handle, grid, self._method_struct, self._rp_struct, bc,
                                            prec, state._q.buf.data_ptr(), self._qalt.data_ptr(), aux,
                                            self._dt_dev.data_ptr(), self._cfl_dev.data_ptr(), stream)
        _lib.check(rc, "claw2d_classic_step")
        self._swap_in(state)
        self.cfl.update_global_max(float(self._cfl_dev.item()))
