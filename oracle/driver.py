"""CPU time-stepping driver around the oracle kernels (TEST INFRASTRUCTURE ONLY).

Restates the reference's host logic so that whole runs can be compared with
the reference's stored fixtures:
  * Solver.evolve_to_time      src/pyclaw/solver.py:582-699
  * Controller.run output loop src/pyclaw/controller.py:202-316 (output_style 1)
  * ClawSolver.step            src/pyclaw/classic/solver.py:100-151
  * ClawSolver2D.step_hyperbolic classic/solver.py:498-547 (step2 / two step2ds calls)
  * SharpClawSolver.step SSP104 src/pyclaw/sharpclaw/solver.py:185-255, dq :266-282
  * ghost fill                 solver.py:299-434 (q) and :438-576 (aux)
Arrays are the reference's: q (meqn, mx, my) float64.
"""
import numpy as np

import oracle as O


class Problem(object):
    """Everything a run needs (mirrors the solver attributes of the reference)."""

    def __init__(self, rp, num_waves, problem_data, dx, dy, bc_lower, bc_upper, aux=None,
                 aux_bc_lower=(1, 1), aux_bc_upper=(1, 1), user_bc_lower=None, user_bc_upper=None,
                 kind='classic', dimensional_split=True, transverse=1, order=2, limiters=1,
                 cfl_max=1.0, cfl_desired=0.9, dt_initial=0.1, dt_max=1e99, dt_variable=True,
                 max_steps=10000, step_source=None, source_split=1, time_integrator='SSP104'):
        self.__dict__.update(locals())
        del self.__dict__['self']


def _padded(q, ng):
    ne, mx, my = q.shape
    qbc = np.zeros((ne, mx + 2 * ng, my + 2 * ng), order='F')
    qbc[:, ng:-ng, ng:-ng] = q
    return qbc


def fill_q_ghosts(P, qbc, ng, t):
    """apply_q_bcs: x lower, x upper, y lower, y upper; custom (0) via callbacks."""
    for idim in range(2):
        for side in (0, 1):
            code = (P.bc_lower if side == 0 else P.bc_upper)[idim]
            if code == 0:
                fn = P.user_bc_lower if side == 0 else P.user_bc_upper
                fn(None, idim, t, qbc, ng)
            else:
                lo = [99, 99]
                hi = [99, 99]
                (lo if side == 0 else hi)[idim] = code
                O.fill_bc(qbc, ng, lo, hi, True)


def fill_aux_ghosts(P, auxbc, ng):
    O.fill_bc(auxbc, ng, list(P.aux_bc_lower), list(P.aux_bc_upper), False)


def classic_step_hyperbolic(P, q, dt, t):
    ng = 2
    qbc = _padded(q, ng)
    fill_q_ghosts(P, qbc, ng, t)
    auxbc = None
    if P.aux is not None:
        auxbc = _padded(P.aux, ng)
        fill_aux_ghosts(P, auxbc, ng)
    O.set_params(P.rp, P.problem_data)
    meth = O.method_array(P.order, P.transverse, P.dimensional_split, P.dt_variable)
    lim = O.mthlim_array(P.limiters, P.num_waves)
    if P.dimensional_split:
        qold = qbc.copy(order='F')
        c1 = O.step2ds(P.rp, qold, qbc, auxbc, P.dx, P.dy, dt, meth, lim, P.num_waves, 1)
        c2 = O.step2ds(P.rp, qbc, qbc, auxbc, P.dx, P.dy, dt, meth, lim, P.num_waves, 2)
        cfl = max(c1, c2)
        qn = qbc
    else:
        qn, cfl = O.step2(P.rp, qbc, auxbc, P.dx, P.dy, dt, meth, lim, P.num_waves)
    return np.asfortranarray(qn[:, ng:-ng, ng:-ng]), cfl


def classic_step(P, q, dt, t):
    """ClawSolver.step: returns (q, cfl)."""
    if P.source_split == 2 and P.step_source is not None:
        q = P.step_source(q, dt / 2.0)
    q, cfl = classic_step_hyperbolic(P, q, dt, t)
    if cfl >= P.cfl_max:
        return q, cfl
    if P.step_source is not None:
        if P.source_split == 2:
            q = P.step_source(q, dt / 2.0)
        if P.source_split == 1:
            q = P.step_source(q, dt)
    return q, cfl


def sharp_dq(P, q, dt, t):
    ng = 3
    qbc = _padded(q, ng)
    fill_q_ghosts(P, qbc, ng, t)
    auxbc = None
    if P.aux is not None:
        auxbc = _padded(P.aux, ng)
        fill_aux_ghosts(P, auxbc, ng)
    O.set_params(P.rp, P.problem_data)
    return O.sharp_dq(P.rp, qbc, auxbc, P.dx, P.dy, dt, P.num_waves)


class _CFLError(Exception):
    pass


def sharp_step(P, q, dt, t):
    """SharpClawSolver.step (sharpclaw/solver.py:185-255): Euler (:199-201),
    SSP33 (:203-218) or SSP104 (:221-251).  Returns (q, cfl) where cfl is the
    last dq's Courant number, or the one that exceeded cfl_max (the reference's
    dq raises CFLError there, :272-275, and q is left as it was)."""
    state = {'cfl': None}

    def dq(qs, ts):
        d, c = sharp_dq(P, qs, dt, ts)
        state['cfl'] = c
        if c > P.cfl_max:
            raise _CFLError()
        return d

    ti = getattr(P, 'time_integrator', 'SSP104')
    if ti == 'Euler':
        try:
            q = q + dq(q, t)
        except _CFLError:
            pass
        return np.asfortranarray(q), state['cfl']
    if ti == 'SSP33':
        try:
            s1 = q + dq(q, t)
            s1 = 0.75 * q + 0.25 * (s1 + dq(s1, t + dt))
            q = 1. / 3. * q + 2. / 3. * (s1 + dq(s1, t + 0.5 * dt))
        except _CFLError:
            pass
        return np.asfortranarray(q), state['cfl']
    if ti != 'SSP104':
        raise Exception('Unrecognized time integrator')
    try:
        s1 = q + dq(q, t) / 6.
        t1 = t + dt / 6.
        for _ in range(4):
            s1 = s1 + dq(s1, t1) / 6.
            t1 = t1 + dt / 6.
        s2 = q / 25. + 9. / 25 * s1
        s1 = 15. * s2 - 5. * s1
        t1 = t + dt / 3.
        for _ in range(4):
            s1 = s1 + dq(s1, t1) / 6.
            t1 = t1 + dt / 6.
        d = dq(s1, t1)
        q = s2 + 0.6 * s1 + 0.1 * d
    except _CFLError:
        pass
    return np.asfortranarray(q), state['cfl']


class Run(object):
    """Controller.run + Solver.evolve_to_time on the CPU."""

    def __init__(self, P, q0, t0=0.0):
        self.P = P
        self.q = np.asfortranarray(q0, dtype=np.float64)
        self.t = t0
        self.dt = P.dt_initial
        self.cfl = P.cfl_desired
        self.status = {}
        self.nsteps = 0
        self.nrejected = 0

    def step(self, q, dt, t):
        if self.P.kind == 'classic':
            return classic_step(self.P, q, dt, t)
        return sharp_step(self.P, q, dt, t)

    def evolve_to_time(self, tend):
        P = self.P
        tstart = self.t
        max_steps = P.max_steps
        if not P.dt_variable:
            max_steps = int((tend - tstart + 1e-10) / self.dt)
        numsteps = 0
        for n in range(max_steps):
            if self.t + self.dt > tend and tstart < tend:
                self.dt = tend - self.t
            if tend - self.t - self.dt < 1.e-14 * self.t:
                self.dt = tend - self.t
            qnew, cfl = self.step(self.q, self.dt, self.t)
            self.cfl = cfl
            if cfl <= P.cfl_max:
                self.q = qnew
                if P.dt_variable:
                    self.t += self.dt
                else:
                    self.t = tstart + (n + 1) * self.dt
                numsteps += 1
                self.nsteps += 1
            else:
                if not P.dt_variable:
                    raise Exception('CFL too large, giving up!')
                self.nrejected += 1
            if P.dt_variable:
                if cfl > 0.0:
                    self.dt = min(P.dt_max, self.dt * P.cfl_desired / cfl)
                else:
                    self.dt = P.dt_max
            if self.t >= tend:
                break
        if P.dt_variable and self.t < tend and numsteps == max_steps:
            raise Exception("Maximum number of timesteps have been taken")

    def run(self, tfinal, num_output_times):
        """Controller.run with output_style 1; returns the list of frames (q copies)."""
        frames = [self.q.copy(order='F')]
        self.dt = self.P.dt_initial
        for tout in np.linspace(self.t, tfinal, num_output_times + 1)[1:]:
            self.evolve_to_time(tout)
            frames.append(self.q.copy(order='F'))
        return frames


def check_diff(expected, test, abstol=None, reltol=None):
    """src/pyclaw/util.py:149-167: L2 norm of the difference, absolute or relative."""
    err = np.linalg.norm(expected - test)
    if abstol is not None:
        return err < abstol, err
    return err / np.linalg.norm(expected) < reltol, err / np.linalg.norm(expected)
