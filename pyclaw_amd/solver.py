"""Solver base class: boundary conditions, the time-step loop, gauges.

Behaviour follows the reference's Solver (src/pyclaw/solver.py:25-742):
  * the BC enumeration and the ghost-fill rules (extrap / periodic / wall /
    custom callback; x sides then y sides, each fill spanning the full padded
    extent of the other dimension, solver.py:299-434 and aux :438-576);
  * evolve_to_time (solver.py:582-699): the same end-time clipping, accept /
    reject on cfl <= cfl_max, dt = min(dt_max, dt*cfl_desired/cfl), status
    dictionary, and the same exceptions;
  * unknown attribute assignment after construction raises TypeError
    (solver.py:121-126).
On the GPU the step writes into a second buffer, so the reference's q_backup
copy is replaced by keeping the previous buffer until the step is accepted.
"""
import ctypes
import logging
import math

import numpy as np
import torch

from . import _lib
from .cfl import CFL
from .state import NG_STORE


class CFLError(Exception):
    def __init__(self, msg):
        super(CFLError, self).__init__(msg)


class BC(object):
    """Boundary condition codes (solver.py:12-18)."""
    custom = 0
    extrap = 1
    periodic = 2
    wall = 3


def default_compute_gauge_values(q, aux):
    return q


def fill_ghosts(field, ng, bc_lower, bc_upper, negate, state=None, t=0.0, user_lower=None,
                user_upper=None, user_args=None):
    """Fill the ng ghost layers around a padded field view, reference order.

    field: tensor (ncomp, mx+2ng, my+2ng) indexed [m, i, j] (the reference's qbc).
    negate: True for q (wall negates the normal component), False for aux."""
    for idim in range(2):
        for side in (0, 1):
            code = (bc_lower if side == 0 else bc_upper)[idim]
            qb = field if idim == 0 else field.permute(0, 2, 1)   # dimension idim moved to axis 1
            if code == BC.custom:
                fn = user_lower if side == 0 else user_upper
                if fn is None:
                    raise Exception("Custom BC function has not been set.")
                dim = state.patch.dimensions[idim] if state is not None else idim
                fn(state, dim, t, field, ng)
            elif code == BC.extrap:
                for i in range(ng):
                    if side == 0:
                        qb[:, i, ...] = qb[:, ng, ...]
                    else:
                        qb[:, -i - 1, ...] = qb[:, -ng - 1, ...]
            elif code == BC.periodic:
                if side == 0:
                    qb[:, :ng, ...] = qb[:, -2 * ng:-ng, ...].clone()
                else:
                    qb[:, -ng:, ...] = qb[:, ng:2 * ng, ...].clone()
            elif code == BC.wall:
                for i in range(ng):
                    if side == 0:
                        qb[:, i, ...] = qb[:, 2 * ng - 1 - i, ...]
                        if negate:
                            qb[idim + 1, i, ...] = -qb[idim + 1, 2 * ng - 1 - i, ...]
                    else:
                        qb[:, -i - 1, ...] = qb[:, -2 * ng + i, ...]
                        if negate:
                            qb[idim + 1, -i - 1, ...] = -qb[idim + 1, -2 * ng + i, ...]
            elif code is None:
                raise Exception("One or more of the boundary conditions has not been specified.")
            else:
                raise NotImplementedError("Boundary condition %s not implemented" % code)


class Solver(object):
    """Base class of the 2D solvers."""

    def __setattr__(self, key, value):
        if not hasattr(self, '_isinitialized'):
            self.__dict__['_isinitialized'] = False
        if self._isinitialized and not hasattr(self, key):
            raise TypeError("%s has no attribute %s" % (self.__class__, key))
        object.__setattr__(self, key, value)

    def __init__(self, riemann_solver=None, claw_package=None):
        self.logger = logging.getLogger('evolve')
        self.dt_initial = 0.1
        self.dt_max = 1e99
        self.max_steps = 10000
        self.dt_variable = True
        self.num_waves = None
        self.num_eqn = None
        self.rp = None
        self._is_set_up = False
        self.dt = self.dt_initial
        self.cfl = CFL(self.cfl_desired)
        self.status = {'cflmax': self.cfl.get_cached_max(), 'dtmin': self.dt, 'dtmax': self.dt,
                       'numsteps': 0}
        self.bc_lower = [None] * self.num_dim
        self.bc_upper = [None] * self.num_dim
        self.aux_bc_lower = [None] * self.num_dim
        self.aux_bc_upper = [None] * self.num_dim
        self.user_bc_lower = None
        self.user_bc_upper = None
        self.user_aux_bc_lower = None
        self.user_aux_bc_upper = None
        self.custom_bc_static = False
        self.compute_gauge_values = default_compute_gauge_values
        self.qbc = None
        self.auxbc = None
        self._ctx = None
        self._qalt = None
        self._gauge_rec = None    # gauge values recorded on the device, written out by flush_gauge_values
        self._dt_dev = None
        self._cfl_dev = None
        self._static_filled = set()
        self._backup_buf = None
        # When True, the first step on each grid shape times the kernel's launch
        # shapes and keeps the fastest for the solver's context; results do not change.
        self.autotune = False
        self._tuned = set()
        # evolve_to_time queues its steps with the accept/reject bookkeeping on
        # the device where no step needs the host; results do not change.
        self.device_loop = True
        self._loop_ws = None
        if riemann_solver is not None:
            self.rp = riemann_solver
            self.num_eqn = riemann_solver.num_eqn
            self.num_waves = riemann_solver.num_waves
        self._isinitialized = True

    @property
    def all_bcs(self):
        return self.bc_lower, self.bc_upper

    @all_bcs.setter
    def all_bcs(self, all_bcs):
        for i in range(self.num_dim):
            self.bc_lower[i] = all_bcs
            self.bc_upper[i] = all_bcs

    def is_valid(self):
        valid = True
        if any(b == BC.custom for b in self.bc_lower) and self.user_bc_lower is None:
            self.logger.debug('Lower custom BC function has not been set.')
            valid = False
        if any(b == BC.custom for b in self.bc_upper) and self.user_bc_upper is None:
            self.logger.debug('Upper custom BC function has not been set.')
            valid = False
        return valid

    def setup(self, solution):
        pass

    def teardown(self):
        pass

    def __str__(self):
        return "Solver Status:\n" + "\n".join("%s = %s" % (k.rjust(25), v) for k, v in self.status.items())

    # ---- device resources -------------------------------------------------------
    def _device_setup(self, state):
        """Move the state to the GPU and allocate the step buffers."""
        if not torch.cuda.is_available():
            raise _lib.ClawError("pyclaw_amd solvers run on the GPU; no GPU is visible")
        _lib.lib()
        dev = torch.device('cuda', torch.cuda.current_device())
        if state.device != dev:
            state.to(device=dev)
        state.sync_to_device()
        if self._ctx is None:
            self._ctx = _lib.Context(dev.index)
            self._tuned = set()
        buf = state._q.buf
        if self._qalt is None or self._qalt.shape != buf.shape or self._qalt.dtype != buf.dtype \
                or self._qalt.device != buf.device:
            self._qalt = torch.zeros_like(buf)
        if self._dt_dev is None:
            self._dt_dev = torch.zeros((), dtype=torch.float64, device=dev)
            self._cfl_dev = torch.zeros((), dtype=torch.float64, device=dev)
        self._static_filled = set()

    def _grid_struct(self, state):
        g = _lib.ClawGrid()
        g.mx, g.my = state.patch.num_cells_global
        g.ng = NG_STORE
        buf = state._q.buf
        g.ldq = buf.shape[2]
        g.plane = buf.shape[1] * buf.shape[2]
        g.dx, g.dy = state.grid.delta
        return g

    def _kernel_bcs(self, state):
        """Boundary codes for the kernel; fills ghost cells on the host where the
        reference calls a user function."""
        q_custom = any(b == BC.custom for b in self.bc_lower + self.bc_upper)
        a_custom = state.num_aux > 0 and any(b == BC.custom for b in self.aux_bc_lower + self.aux_bc_upper)
        bc = _lib.ClawBC()
        ng = self.num_ghost
        if q_custom:
            if self.custom_bc_static:
                if any(b == BC.custom for b in (self.bc_lower[1], self.bc_upper[1])):
                    raise ValueError("custom_bc_static supports custom x boundaries only")
                for b in self._static_targets(state):
                    if id(b) not in self._static_filled:
                        self._fill_custom_sides(state, b)
                        self._static_filled.add(id(b))
                for s in range(2):
                    bc.lower[s] = _lib.BC_STORED if self.bc_lower[s] == BC.custom else self.bc_lower[s]
                    bc.upper[s] = _lib.BC_STORED if self.bc_upper[s] == BC.custom else self.bc_upper[s]
            else:
                self._fill_all_q(state)
                for s in range(2):
                    bc.lower[s] = bc.upper[s] = _lib.BC_STORED
        else:
            for s in range(2):
                if self.bc_lower[s] is None or self.bc_upper[s] is None:
                    raise Exception("One or more of the boundary conditions has not been specified.")
                bc.lower[s], bc.upper[s] = int(self.bc_lower[s]), int(self.bc_upper[s])
        if state.num_aux > 0:
            if a_custom:
                field = state._aux.padded()[:, NG_STORE - ng:field_end(state, 0, ng), NG_STORE - ng:field_end(state, 1, ng)]
                fill_ghosts(field, ng, self.aux_bc_lower, self.aux_bc_upper, False, state, state.t,
                            self.user_aux_bc_lower, self.user_aux_bc_upper)
                for s in range(2):
                    bc.aux_lower[s] = bc.aux_upper[s] = _lib.BC_STORED
            else:
                for s in range(2):
                    if self.aux_bc_lower[s] is None or self.aux_bc_upper[s] is None:
                        raise Exception("One or more of the aux boundary conditions has not been specified.")
                    bc.aux_lower[s], bc.aux_upper[s] = int(self.aux_bc_lower[s]), int(self.aux_bc_upper[s])
        else:
            for s in range(2):
                bc.aux_lower[s] = bc.aux_upper[s] = _lib.BC_EXTRAP
        return bc

    def _static_targets(self, state):
        """Buffers whose ghost cells hold time-independent user boundary values."""
        return [state._q.buf, self._qalt]

    def _qbc_view(self, state, buf):
        ng = self.num_ghost
        return state._q.padded(buf)[:, NG_STORE - ng:field_end(state, 0, ng), NG_STORE - ng:field_end(state, 1, ng)]

    def _fill_all_q(self, state):
        fill_ghosts(self._qbc_view(state, state._q.buf), self.num_ghost, self.bc_lower, self.bc_upper, True,
                    state, state.t, self.user_bc_lower, self.user_bc_upper)

    def _fill_custom_sides(self, state, buf):
        view = self._qbc_view(state, buf)
        for side, fn in ((0, self.user_bc_lower), (1, self.user_bc_upper)):
            if (self.bc_lower if side == 0 else self.bc_upper)[0] == BC.custom:
                fn(state, state.patch.dimensions[0], state.t, view, self.num_ghost)

    def _swap_in(self, state):
        """Make the step's output buffer current; keep the old one as backup."""
        old = state._q.buf
        state._q.buf = self._qalt
        self._qalt = old
        state._device_updated()

    # ---- time stepping ------------------------------------------------------------
    def evolve_to_time(self, solution, tend=None):
        """solver.py:582-699."""
        if not self._is_set_up:
            self.setup(solution)
        take_one_step = tend is None
        tstart = solution.t
        self.status['cflmax'] = self.cfl.get_cached_max()
        self.status['dtmin'] = self.dt
        self.status['dtmax'] = self.dt
        self.status['numsteps'] = 0
        if not self.dt_variable:
            if take_one_step:
                self.max_steps = 1
            else:
                self.max_steps = int((tend - tstart + 1e-10) / self.dt)
                if abs(self.max_steps * self.dt - (tend - tstart)) > 1e-5 * (tend - tstart):
                    raise Exception('dt does not divide (tend-tstart) and dt is fixed!')
        if self.dt_variable == 1 and self.cfl_desired > self.cfl_max:
            raise Exception('Variable time-stepping and desired CFL > maximum CFL')
        if not take_one_step and tend <= tstart:
            self.logger.info("Already at or beyond end time: no evolution required.")
            self.max_steps = 0

        try:
            launch = self._device_loop_launcher(solution, take_one_step)
            if launch is not None:
                self._device_step_loop(solution, tend, tstart, launch)
            else:
                self._step_loop(solution, tend, take_one_step, tstart)
        finally:
            # gauge values recorded on the device reach their files and
            # gauge_data at the end of every evolve_to_time call
            self.flush_gauge_values(solution)
        if self.dt_variable and tend is not None and solution.t < tend \
                and self.status['numsteps'] == self.max_steps:
            raise Exception("Maximum number of timesteps have been taken")
        return self.status

    def _step_loop(self, solution, tend, take_one_step, tstart):
        """The step loop of evolve_to_time (solver.py:629-696)."""
        for n in range(self.max_steps):
            state = solution.state
            if not take_one_step:
                if solution.t + self.dt > tend and tstart < tend:
                    self.dt = tend - solution.t
                if tend - solution.t - self.dt < 1.e-14 * solution.t:
                    self.dt = tend - solution.t
            if self.dt_variable:
                backup = self._take_backup(state)
                told = solution.t
            self.step(solution)
            cfl = self.cfl.get_cached_max()
            if cfl <= self.cfl_max:
                self.status['cflmax'] = max(cfl, self.status['cflmax'])
                if self.dt_variable:
                    solution.t += self.dt
                else:
                    solution.t = tstart + (n + 1) * self.dt
                self.logger.debug("Step %i  CFL = %f   dt = %f   t = %f" % (n, cfl, self.dt, solution.t))
                self.write_gauge_values(solution)
                self.status['numsteps'] += 1
            else:
                self.logger.debug("Rejecting time step, CFL number too large")
                if self.dt_variable:
                    self._restore_backup(state, backup)
                    solution.t = told
                else:
                    self.status['cflmax'] = max(cfl, self.status['cflmax'])
                    raise Exception('CFL too large, giving up!')
            if self.dt_variable:
                if cfl > 0.0:
                    self.dt = min(self.dt_max, self.dt * self.cfl_desired / cfl)
                    self.status['dtmin'] = min(self.dt, self.status['dtmin'])
                    self.status['dtmax'] = max(self.dt, self.status['dtmax'])
                else:
                    self.dt = self.dt_max
            if solution.t >= tend if tend is not None else False:
                break
            if take_one_step:
                break

    def _device_loop_launcher(self, solution, take_one_step):
        """The device step of _loop_launcher, or None for the host loop.  The
        step loop runs on the device when no step needs the host: gauges only
        with the default compute_gauge_values on an FP64 state (recorded on the
        device), no per-step debug log, no user hooks or ghost cells filled by
        user functions each step, and a solver with a device step
        (_loop_launcher)."""
        if not self.device_loop or take_one_step or self.max_steps <= 0:
            return None
        state = solution.state
        if self._hooks_active() or self.logger.isEnabledFor(logging.DEBUG):
            return None
        if state.grid.gauges and not (self.compute_gauge_values is default_compute_gauge_values
                                      and state._q.buf.dtype == torch.float64):
            return None
        if state.grid.gauges and not gauge_cells_inside(state):
            # the host loop raises the reference's IndexError after the first step
            return None
        if not hasattr(_lib.lib(), 'claw2d_classic_loop'):
            return None
        q_custom = any(b == BC.custom for b in self.bc_lower + self.bc_upper)
        a_custom = state.num_aux > 0 and any(b == BC.custom for b in self.aux_bc_lower + self.aux_bc_upper)
        if (q_custom and not self.custom_bc_static) or a_custom:
            return None
        return self._loop_launcher(state)

    def _loop_launcher(self, state):
        """A function launch(k0, niter, qa, qb, rule, states, cfl, nbytes, stream)
        that queues niter iterations of the device step loop (claw2d_loop_advance
        after each step; iteration k0 + i reads q[(k0+i)%2], states[(k0+i)%2] and
        the launch.ncfl (default 1) Courant numbers at cfl[ncfl*((k0+i)%2)]), or
        None where the solver has none."""
        return None

    def _device_step_loop(self, solution, tend, tstart, launch):
        """_step_loop with the bookkeeping of solver.py:629-696 on the device
        (claw2d_loop_advance after each step): steps are queued in chunks of the
        number left at the current dt, with one host synchronisation per chunk.
        Solution, status, dt and Courant number end as the host loop leaves them."""
        state = solution.state
        dev = state._q.buf.device
        t = solution.t
        dt = self.dt
        # the first step's clip to tend (solver.py:637-641); the device clips the rest
        if t + dt > tend and tstart < tend:
            dt = tend - t
        if tend - t - dt < 1.e-14 * t:
            dt = tend - t
        st = _lib.ClawLoopState(dt, t, self.status['cflmax'], self.status['dtmin'], self.status['dtmax'],
                                self.cfl.get_cached_max(), 0, 0, 0, 0)
        ncfl = getattr(launch, 'ncfl', 1)
        rule = _lib.ClawLoopRule(self.cfl_max, self.cfl_desired, self.dt_max, tend, tstart, self.max_steps, ncfl,
                                 1 if self.dt_variable else 0, 0)
        size = ctypes.sizeof(_lib.ClawLoopState)
        ws = self._loop_ws
        if ws is None or ws['dev'] != dev or ws['ncfl'] != ncfl:
            # kept between calls: the loop state, the Courant slots and a pinned upload buffer
            ws = dict(dev=dev, ncfl=ncfl, sbuf=torch.zeros(2 * size, dtype=torch.uint8, device=dev),
                      cfl=torch.zeros(2 * ncfl, dtype=torch.float64, device=dev),
                      host=torch.empty(size, dtype=torch.uint8, pin_memory=True), gauges=None)
            self._loop_ws = ws
        sbuf, cfl = ws['sbuf'], ws['cfl']
        ctypes.memmove(ws['host'].data_ptr(), ctypes.addressof(st), size)
        sbuf[:size].copy_(ws['host'], non_blocking=True)
        cfl.zero_()
        bufs = [state._q.buf, self._qalt]
        qp = [b.data_ptr() for b in bufs]
        nbytes = bufs[0].numel() * bufs[0].element_size()
        stream = torch.cuda.current_stream(dev).cuda_stream
        rref = ctypes.byref(rule)
        gauges = state.grid.gauges
        G = None
        if gauges:
            # recorded after each accepted step on the device, written out per chunk
            self.flush_gauge_values(solution)
            ne, rows, ldq = state.num_eqn, bufs[0].shape[1], bufs[0].shape[2]
            gkey = (tuple(map(tuple, gauges)), tuple(bufs[0].shape))
            if ws['gauges'] is None or ws['gauges'][0] != gkey:
                offs = torch.tensor([m * rows * ldq + (iy + NG_STORE) * ldq + (ix + NG_STORE)
                                     for (ix, iy) in gauges for m in range(ne)], dtype=torch.int64, device=dev)
                ws['gauges'] = (gkey, offs, torch.empty((self._GAUGE_CAP, offs.numel()), dtype=torch.float64,
                                                        device=dev),
                                torch.empty(self._GAUGE_CAP, dtype=torch.float64, device=dev))
            _, offs, vals, times = ws['gauges']
            G = _lib.ClawLoopGauges(offs.data_ptr(), vals.data_ptr(), times.data_ptr(), offs.numel(), 0,
                                    self._GAUGE_CAP, 0)
        k = 0
        while True:
            left = self.max_steps - st.niter
            x = (tend - st.t) / st.dt if st.dt > 0 else 1.0
            est = int(math.ceil(x)) if x < 1024 else 1024      # steps left at the current dt
            n = max(1, min(left, est, self._GAUGE_CAP))
            base = st.naccepted
            if G is not None:
                G.base = base
            launch(k, n, qp[0], qp[1], rref, sbuf.data_ptr(), cfl.data_ptr(), nbytes, stream,
                   None if G is None else ctypes.byref(G))
            k += n
            a = k & 1
            st = _lib.ClawLoopState.from_buffer_copy(sbuf[a * size:(a + 1) * size].cpu().numpy().tobytes())
            if G is not None and st.naccepted > base:
                m = st.naccepted - base
                tv = times[:m].cpu().numpy()
                vv = vals[:m].cpu().numpy()
                for r in range(m):
                    for i in range(len(gauges)):
                        self._emit_gauge(state, i, float(tv[r]), vv[r, i * ne:(i + 1) * ne])
            if st.done:
                break
        if k & 1:
            self._swap_in(state)
        else:
            state._device_updated()
        solution.t = st.t
        self.dt = st.dt
        self.cfl.update_global_max(st.cfl_last)
        self.status['cflmax'] = st.cflmax
        self.status['dtmin'] = st.dtmin
        self.status['dtmax'] = st.dtmax
        self.status['numsteps'] += st.naccepted
        if st.failed:
            raise Exception('CFL too large, giving up!')

    def _hooks_active(self):
        return False

    def _take_backup(self, state):
        """Keep q as it was before the step.  Without user hooks the step's input
        buffer is untouched by the kernel, so holding a reference suffices."""
        state.sync_to_device()
        if self._hooks_active():
            return ('copy', state._q.buf.clone())
        return ('buf', state._q.buf)

    def _restore_backup(self, state, backup):
        kind, buf = backup
        if kind == 'buf':
            if state._q.buf is not buf:
                self._qalt = state._q.buf
                state._q.buf = buf
        else:
            state._q.buf.copy_(buf)
        state._device_updated()

    def step(self, solution):
        raise NotImplementedError("No stepping routine has been defined!")

    # ---- gauges (solver.py:713-742) ------------------------------------------------------
    def write_gauge_values(self, solution):
        """Record the gauge values of this step.  With the default
        compute_gauge_values (q itself) and an FP64 device state, the values are
        gathered into a device buffer (claw2d_gather, no host copy per step)
        and written to the gauge files and gauge_data by flush_gauge_values, at
        the end of every evolve_to_time call or when the buffer fills; a user
        compute_gauge_values is called per step on host values, as in the
        reference."""
        gauges = solution.state.grid.gauges
        if not gauges:
            return
        state = solution.state
        check_gauge_cells(state)
        buf = state._q.buf
        if self.compute_gauge_values is default_compute_gauge_values and buf.is_cuda \
                and buf.dtype == torch.float64:
            self._record_gauges(solution)
            return
        self.flush_gauge_values(solution)
        idx = torch.tensor(gauges, dtype=torch.long)
        q = state._q.interior()
        qg = q[:, idx[:, 0].to(q.device), idx[:, 1].to(q.device)].to('cpu', torch.float64).numpy()
        aux = None
        if state._aux is not None:
            a = state._aux.interior()
            aux = a[:, idx[:, 0].to(a.device), idx[:, 1].to(a.device)].to('cpu', torch.float64).numpy()
        for i in range(len(gauges)):
            p = self.compute_gauge_values(qg[:, i], None if aux is None else aux[:, i])
            self._emit_gauge(state, i, solution.t, p)

    def _emit_gauge(self, state, i, t, p):
        """solver.py:727-737: keep_gauges data and one file line per gauge."""
        if state.keep_gauges:
            gd = state.gauge_data
            if len(gd) == len(state.grid.gauges):
                gd[i] = np.vstack((gd[i], np.append(t, p)))
            else:
                gd.append(np.append(t, p))
        files = state.grid.gauge_files
        if i < len(files):
            files[i].write(py2_str(t) + ' ' + ' '.join(py2_str(j) for j in p) + '\n')

    _GAUGE_CAP = 1024   # steps recorded on the device between flushes

    def _record_gauges(self, solution):
        state = solution.state
        gauges = state.grid.gauges
        buf = state._q.buf
        ne = state.num_eqn
        key = (tuple(map(tuple, gauges)), tuple(buf.shape), buf.device)
        rec = self._gauge_rec
        if rec is None or rec['key'] != key:
            self.flush_gauge_values(solution)
            rows, ldq = buf.shape[1], buf.shape[2]
            plane = rows * ldq
            # gauge-major: value g * num_eqn + m is component m at gauge g
            offs = [m * plane + (iy + NG_STORE) * ldq + (ix + NG_STORE) for (ix, iy) in gauges for m in range(ne)]
            rec = dict(key=key, offsets=torch.tensor(offs, dtype=torch.int64, device=buf.device),
                       vals=torch.empty((self._GAUGE_CAP, len(offs)), dtype=torch.float64, device=buf.device),
                       times=[], state=state)
            self._gauge_rec = rec
        if len(rec['times']) == self._GAUGE_CAP:
            self.flush_gauge_values(solution)
        k = len(rec['times'])
        n = rec['offsets'].numel()
        _lib.check(_lib.lib().claw2d_gather(buf.data_ptr(), rec['offsets'].data_ptr(), n,
                                            rec['vals'][k].data_ptr(),
                                            torch.cuda.current_stream(buf.device).cuda_stream), "claw2d_gather")
        rec['times'].append(solution.t)
        rec['state'] = state

    def flush_gauge_values(self, solution=None):
        """Write the gauge values recorded on the device (one host copy for all
        recorded steps), in step order, exactly as the per-step path would."""
        rec = self._gauge_rec
        if rec is None or not rec['times']:
            return
        state = rec['state']
        ne = state.num_eqn
        vals = rec['vals'][:len(rec['times'])].to('cpu').numpy()
        for r, t in enumerate(rec['times']):
            for i in range(len(state.grid.gauges)):
                self._emit_gauge(state, i, t, vals[r, i * ne:(i + 1) * ne])
        rec['times'] = []


def py2_str(x):
    """str(x) as Python 2 writes a float (12 significant digits, '.0' added to
    integral values) -- the reference writes gauge lines with str() under
    Python 2 (solver.py:736-737), and so did numpy's float64 there.  Other
    values use str()."""
    if isinstance(x, (float, np.floating)) and not isinstance(x, np.float16):
        s = '%.12g' % float(x)
        if all(ch in '-0123456789' for ch in s):
            s += '.0'
        return s
    return str(x)


def gauge_cells_inside(state):
    mx, my = state.patch.num_cells_global[0], state.patch.num_cells_global[1]
    return all(0 <= ix < mx and 0 <= iy < my for (ix, iy) in state.grid.gauges)


def check_gauge_cells(state):
    """The reference reads q[:, ix, iy] for each gauge (solver.py:722-723), so a
    gauge index outside the grid raises numpy's IndexError; the device buffer
    is padded, so the same check is made here before any read."""
    shape = state.patch.num_cells_global
    for g in state.grid.gauges:
        for axis, (i, n) in enumerate(zip(g, shape)):
            if not 0 <= i < n:
                raise IndexError("index %d is out of bounds for axis %d with size %d" % (i, axis + 1, n))


def field_end(state, idim, ng):
    n = state.patch.num_cells_global[idim]
    return NG_STORE + n + ng
