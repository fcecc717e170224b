"""SharpClaw solver in 2D on the GPU (WENO5 + SSP Runge-Kutta).

API of the reference's SharpClawSolver / SharpClawSolver2D
(src/pyclaw/sharpclaw/solver.py:29-582): weno_order, lim_type, char_decomp,
time_integrator ('Euler', 'SSP33', 'SSP104'), limiters, cfl_max=2.5,
cfl_desired=2.45, dq_src, before_step, call_before_step_each_stage, dq(),
dqdt(), step().

Each Runge-Kutta stage is one launch of the fused HIP kernel (WENO5,
Riemann solves and the stage combination); a whole SSP104 step without user
hooks is a single C call (claw2d_sharpclaw_ssp104_step).  Courant numbers of
all stages stay on the device until the step ends; the reference's rule is
then applied on the host: the first stage above cfl_max aborts the step
(CFLError, sharpclaw/solver.py:276-277) and q is left unchanged, otherwise the
last stage's number is the step's (cfl.update_global_max per dq call).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .solver import CFLError, Solver
from .state import NG_STORE, State


def before_step(solver, solution):
    """Default hook: does nothing (sharpclaw/solver.py:20-26)."""
    pass


_SCHEDULES = {
    'Euler': [(_lib.SM_EULER, 'q', 'qnew')],
    'SSP33': [(_lib.SM_EULER, 'q', 's1a'), (_lib.SM_SSP33_2, 's1a', 's1b'), (_lib.SM_SSP33_3, 's1b', 'qnew')],
    'SSP104': [(_lib.SM_SSP104_A, 'q', 's1a'), (_lib.SM_SSP104_A, 's1a', 's1b'), (_lib.SM_SSP104_A, 's1b', 's1a'),
               (_lib.SM_SSP104_A, 's1a', 's1b'), (_lib.SM_SSP104_MID, 's1b', 's1a'),
               (_lib.SM_SSP104_A, 's1a', 's1b'), (_lib.SM_SSP104_A, 's1b', 's1a'),
               (_lib.SM_SSP104_A, 's1a', 's1b'), (_lib.SM_SSP104_A, 's1b', 's1a'),
               (_lib.SM_SSP104_END, 's1a', 'qnew')],
}
# stage time offsets in units of dt (s1.t updates, sharpclaw/solver.py:226-246)
_STAGE_T = {
    'Euler': [0.0],
    'SSP33': [0.0, 1.0, 0.5],
    'SSP104': [0.0, 1 / 6., 2 / 6., 3 / 6., 4 / 6., 1 / 3., 1 / 3. + 1 / 6., 1 / 3. + 2 / 6., 1 / 3. + 3 / 6.,
               1 / 3. + 4 / 6.],
}


class SharpClawSolver(Solver):
    def __init__(self, riemann_solver=None, claw_package=None):
        self.limiters = [1]
        self.before_step = before_step
        self.lim_type = 2
        self.weno_order = 5
        self.time_integrator = 'SSP104'
        self.char_decomp = 0
        self.tfluct_solver = False
        self.aux_time_dep = False
        self.kernel_language = 'Fortran'
        self.num_ghost = 3
        self.fwave = False
        self.cfl_desired = 2.45
        self.cfl_max = 2.5
        self.dq_src = None
        self.call_before_step_each_stage = False
        # FP64 states only: use the kernels built with fused multiply-add and
        # shared reciprocals (CLAW_F64_FMA); each stage agrees with the
        # reference arithmetic to rounding (not bit for bit).
        self.fma = False
        self._mthlim = self.limiters
        self._method = None
        self._rk_stages = None
        self._regs = None
        self._rp_struct = None
        self._cfl_stages = None
        super(SharpClawSolver, self).__init__(riemann_solver, claw_package)

    def set_mthlim(self):
        self._mthlim = self.limiters
        if not isinstance(self.limiters, list):
            self._mthlim = [self._mthlim]
        if len(self._mthlim) == 1:
            self._mthlim = self._mthlim * self.num_waves
        if len(self._mthlim) != self.num_waves:
            raise Exception('Length of solver.limiters is not equal to 1 or to solver.num_waves')

    def setup(self, solution):
        """sharpclaw/solver.py:155-182."""
        self.num_ghost = (self.weno_order + 1) // 2
        state = solution.states[0]
        state.set_num_ghost(self.num_ghost)
        if self.rp is None:
            raise Exception("solver.rp (the Riemann solver) has not been set")
        if self.num_waves is None:
            self.num_waves = self.rp.num_waves
        if self.lim_type != 2 or self.weno_order != 5:
            raise NotImplementedError("the GPU SharpClaw path implements WENO5 (lim_type=2, weno_order=5)")
        if self.char_decomp != 0:
            raise NotImplementedError("characteristic-wise WENO (char_decomp>0) is not implemented; the "
                                      "reference's eigenvector routine is a stub (sharpclaw/evec.f90)")
        if self.tfluct_solver or self.fwave:
            raise NotImplementedError("tfluct_solver / fwave are not supported")
        if state.index_capa >= 0:
            raise NotImplementedError("capacity functions (index_capa) are not supported")
        if self.time_integrator not in _SCHEDULES:
            raise Exception('Unrecognized time integrator')
        if state.num_eqn != self.rp.num_eqn or self.num_waves != self.rp.num_waves:
            raise Exception("Riemann solver %s does not match num_eqn/num_waves" % self.rp.__name__)
        self.set_mthlim()
        r = _lib.ClawRiemann()
        r.kind = self.rp.kind
        r.efix = 1
        for k, v in enumerate(self.rp.params(state.problem_data)):
            r.p[k] = v
        self._rp_struct = r
        self._device_setup(state)
        buf = state._q.buf
        self._regs = {k: torch.zeros_like(buf) for k in ('s1a', 's1b', 's2')}
        self._cfl_stages = torch.zeros(10, dtype=torch.float64, device=buf.device)
        self.allocate_rk_stages(solution)
        self._is_set_up = True

    def allocate_rk_stages(self, solution):
        """Stage States for user hooks (sharpclaw/solver.py:330-355)."""
        state = solution.states[0]
        self._rk_stages = []
        for _ in range(2):
            s = State.__new__(State)
            s.__dict__.update({k: v for k, v in state.__dict__.items() if k not in ('_q',)})
            s._q = type(state._q).__new__(type(state._q))
            s._q.__dict__.update(state._q.__dict__)
            s._q.host_valid = False
            self._rk_stages.append(s)

    def teardown(self):
        self._ctx = None

    def _static_targets(self, state):
        return [state._q.buf, self._qalt] + list(self._regs.values())

    # ------------------------------------------------------------------------
    def _stage_bcs(self, state, buf):
        """Kernel boundary codes; user boundary functions fill `buf`'s ghosts."""
        saved = state._q.buf
        state._q.buf = buf
        try:
            return self._kernel_bcs(state)
        finally:
            state._q.buf = saved

    def step(self, solution):
        """sharpclaw/solver.py:185-255 on the GPU."""
        state = solution.states[0]
        if not self._is_set_up:
            self.setup(solution)
        self.before_step(self, state)
        state.sync_to_device()
        has_custom = any(b == 0 for b in self.bc_lower + self.bc_upper) and not self.custom_bc_static
        simple = (self.dq_src is None and not self.call_before_step_each_stage and not has_custom)
        if self.autotune:
            self._tune(state)
        if simple:
            ok = self._step_fused(state)
        else:
            ok = self._step_staged(state)
        return ok

    def _loop_launcher(self, state):
        """Iterations of the device step loop: whole SSP104 steps
        (claw2d_sharpclaw_ssp104_step), each followed by the loop's bookkeeping
        with the stage rule of _apply_cfl_rule, launched by
        claw2d_sharpclaw_loop.  None where a step needs the host (another
        integrator, dq_src, a before_step hook, user boundary functions)."""
        has_custom = any(b == 0 for b in self.bc_lower + self.bc_upper) and not self.custom_bc_static
        # fixed dt stays on the host: an aborted step leaves q as it was, where
        # the device loop with fixed dt keeps the classic step's result
        if (not self._is_set_up or not self.dt_variable or self.time_integrator != 'SSP104' or self.dq_src is not None
                or self.call_before_step_each_stage or self.before_step is not before_step or has_custom):
            return None
        state.sync_to_device()
        if self.autotune:
            self._tune(state)
        bc = self._kernel_bcs(state)
        grid = self._grid_struct(state)
        prec = self._precision(state)
        aux = state._aux.buf.data_ptr() if state._aux is not None else None
        fn = _lib.lib().claw2d_sharpclaw_loop
        ctx, r = self._ctx.handle, self._rp_struct
        regs = [self._regs[k].data_ptr() for k in ('s1a', 's1b', 's2')]

        def launch(k0, niter, qa, qb, rule, states, cfl, nbytes, stream, gauges=None):
            rc = fn(ctx, grid, r, bc, prec, qa, qb, regs[0], regs[1], regs[2], aux, rule, states, cfl, nbytes,
                    k0, niter, gauges, stream)
            _lib.check(rc, "claw2d_sharpclaw_loop")
        launch.ncfl = 10
        return launch

    def _precision(self, state):
        if state.dtype == torch.float32:
            return _lib.F32
        return _lib.F64_FMA if self.fma else _lib.F64

    def _tune(self, state):
        """Time the stage kernel's launch shapes once per grid shape
        (claw2d_sharpclaw_autotune); the choice never changes results."""
        grid = self._grid_struct(state)
        prec = self._precision(state)
        key = (grid.mx, grid.my, prec, self._rp_struct.kind)
        if key in self._tuned:
            return
        bc = self._kernel_bcs(state)
        self._dt_dev.fill_(float(self.dt))
        aux = state._aux.buf.data_ptr() if state._aux is not None else None
        band = ctypes.c_int(0)
        rc = _lib.lib().claw2d_sharpclaw_autotune(self._ctx.handle, grid, self._rp_struct, bc, prec,
                                                  state._q.buf.data_ptr(), self._qalt.data_ptr(), aux,
                                                  self._dt_dev.data_ptr(), self._cfl_stages.data_ptr(),
                                                  torch.cuda.current_stream().cuda_stream, ctypes.byref(band))
        _lib.check(rc, "claw2d_sharpclaw_autotune")
        self._tuned.add(key)

    def _launch_stage(self, state, mode, qin, q0, out, cfl_ptr, bc):
        grid = self._grid_struct(state)
        prec = self._precision(state)
        aux = state._aux.buf.data_ptr() if state._aux is not None else None
        rc = _lib.lib().claw2d_sharpclaw_stage(self._ctx.handle, grid, self._rp_struct, bc, prec, mode,
                                               qin.data_ptr(), q0.data_ptr(), self._regs['s2'].data_ptr(),
                                               out.data_ptr(), aux, self._dt_dev.data_ptr(), cfl_ptr,
                                               torch.cuda.current_stream().cuda_stream)
        _lib.check(rc, "claw2d_sharpclaw_stage")

    def _apply_cfl_rule(self, cfls):
        for c in cfls:
            if c > self.cfl_max:
                self.cfl.update_global_max(c)
                return False
        self.cfl.update_global_max(cfls[-1])
        return True

    def _step_fused(self, state):
        bc = self._kernel_bcs(state)
        self._dt_dev.fill_(float(self.dt))
        q = state._q.buf
        qnew = self._qalt
        sched = _SCHEDULES[self.time_integrator]
        if self.time_integrator == 'SSP104':
            grid = self._grid_struct(state)
            prec = self._precision(state)
            aux = state._aux.buf.data_ptr() if state._aux is not None else None
            rc = _lib.lib().claw2d_sharpclaw_ssp104_step(
                self._ctx.handle, grid, self._rp_struct, bc, prec, q.data_ptr(), qnew.data_ptr(),
                self._regs['s1a'].data_ptr(), self._regs['s1b'].data_ptr(), self._regs['s2'].data_ptr(), aux,
                self._dt_dev.data_ptr(), self._cfl_stages.data_ptr(), torch.cuda.current_stream().cuda_stream)
            _lib.check(rc, "claw2d_sharpclaw_ssp104_step")
        else:
            bufs = dict(self._regs, q=q, qnew=qnew)
            for k, (mode, a, b) in enumerate(sched):
                self._launch_stage(state, mode, bufs[a], q, bufs[b], self._cfl_stages.data_ptr() + 8 * k, bc)
        cfls = self._cfl_stages[:len(sched)].tolist()
        ok = self._apply_cfl_rule(cfls)
        if ok:
            self._swap_in(state)
        return ok

    def _stage_state(self, state, buf, t):
        s = self._rk_stages[0]
        s._q.buf = buf
        s._q.host_valid = False
        s._q.host_dirty = False
        s.t = t
        return s

    def _step_staged(self, state):
        """Stage by stage: user boundary functions, per-stage hooks and dq_src
        (evaluated on the host between launches)."""
        self._dt_dev.fill_(float(self.dt))
        q = state._q.buf
        qnew = self._qalt
        bufs = dict(self._regs, q=q, qnew=qnew)
        sched = _SCHEDULES[self.time_integrator]
        tstage = _STAGE_T[self.time_integrator]
        dqbuf = None
        for k, (mode, a, b) in enumerate(sched):
            sst = self._stage_state(state, bufs[a], state.t + tstage[k] * self.dt) if a != 'q' else state
            if k > 0 and self.call_before_step_each_stage:
                self.before_step(self, sst)
                sst.sync_to_device()
            bc = self._stage_bcs(sst, bufs[a])
            cptr = self._cfl_stages.data_ptr() + 8 * k
            if self.dq_src is None:
                self._launch_stage(state, mode, bufs[a], q, bufs[b], cptr, bc)
                c = float(self._cfl_stages[k].item())
                if c > self.cfl_max:
                    self.cfl.update_global_max(c)
                    return False
            else:
                if dqbuf is None:
                    dqbuf = torch.zeros_like(q)
                self._launch_stage(state, _lib.SM_DQ, bufs[a], q, dqbuf, cptr, bc)
                c = float(self._cfl_stages[k].item())
                if c > self.cfl_max:
                    self.cfl.update_global_max(c)
                    return False
                src = self.dq_src(self, sst, self.dt)
                dqi = state._q.interior(dqbuf)
                if torch.is_tensor(src):   # a device source (pyclaw_amd.sources): no host copy
                    dqi += src.to(dqi.device, dqi.dtype)
                else:
                    src = np.asarray(src, dtype=np.float64)
                    dqi += torch.from_numpy(np.ascontiguousarray(src)).to(dqi.device, dqi.dtype)
                self._combine(mode, state._q.interior(bufs[a]), state._q.interior(q), dqi,
                              state._q.interior(self._regs['s2']), state._q.interior(bufs[b]))
            self.cfl.update_global_max(c)
        self._swap_in(state)
        return True

    @staticmethod
    def _combine(mode, qin, q0, dq, s2, out):
        """The stage combinations of sharpclaw/solver.py:201-251, elementwise in
        the same operation order as the kernel epilogue."""
        if mode == _lib.SM_EULER:
            out.copy_(qin + dq)
        elif mode == _lib.SM_SSP104_A:
            out.copy_(qin + dq / 6.)
        elif mode == _lib.SM_SSP104_MID:
            t = qin + dq / 6.
            s2.copy_(q0 / 25. + (9. / 25) * t)
            out.copy_(15. * s2 - 5. * t)
        elif mode == _lib.SM_SSP104_END:
            out.copy_((s2 + 0.6 * qin) + 0.1 * dq)
        elif mode == _lib.SM_SSP33_2:
            out.copy_(0.75 * q0 + 0.25 * (qin + dq))
        elif mode == _lib.SM_SSP33_3:
            out.copy_((1. / 3.) * q0 + (2. / 3.) * (qin + dq))
        else:
            out.copy_(dq)

    # ---- reference API: dq and dq/dt on the host -------------------------------
    def dq(self, state):
        """dt * dq/dt (plus dq_src), as a host array (sharpclaw/solver.py:266-282)."""
        if not self._is_set_up:
            raise Exception("call solver.setup(solution) first")
        state.sync_to_device()
        bc = self._kernel_bcs(state)
        self._dt_dev.fill_(float(self.dt))
        dqbuf = torch.zeros_like(state._q.buf)
        self._launch_stage(state, _lib.SM_DQ, state._q.buf, state._q.buf, dqbuf, self._cfl_stages.data_ptr(), bc)
        c = float(self._cfl_stages[0].item())
        self.cfl.update_global_max(c)
        if c > self.cfl_max:
            raise CFLError('cfl_max exceeded')
        out = state._q.interior(dqbuf).to('cpu', torch.float64).numpy().copy(order='F')
        if self.dq_src is not None:
            src = self.dq_src(self, state, self.dt)
            out += src.to('cpu', torch.float64).numpy() if torch.is_tensor(src) else src
        return out

    def dqdt(self, state):
        self.dt = 1
        d = self.dq(state)
        return d.flatten('F')


class SharpClawSolver2D(SharpClawSolver):
    def __init__(self, riemann_solver=None, claw_package=None):
        self.num_dim = 2
        super(SharpClawSolver2D, self).__init__(riemann_solver, claw_package)
