"""The shock-bubble application's own host routines, restated for the checker.

TEST INFRASTRUCTURE ONLY.  tests/ use these as the reference behaviour that the
device source kernels (pyclaw_amd/csrc/sources.hip) and the drop-in API's
user boundary callbacks are compared with; the product never imports them.

  shockbubble_inflow_numpy -- shockbc          apps/euler_2d/shockbubble.py:68-87
  step_euler_radial_numpy  -- step_Euler_radial apps/euler_2d/shockbubble.py:89-124
  dq_euler_radial_numpy    -- dq_Euler_radial   apps/euler_2d/shockbubble.py:127-153

Each expression keeps numpy's left-to-right evaluation order of the
application's code, so results can be compared bit for bit.
"""
import numpy as np

GAMMA = 1.4
GAMMA1 = GAMMA - 1.


def _shock_state(pinf=5.):
    """Post-shock state from Rankine-Hugoniot at pressure ratio pinf (shockbubble.py:26-28)."""
    rinf = (GAMMA1 + pinf * (GAMMA + 1.)) / ((GAMMA + 1.) + GAMMA1 * pinf)
    vinf = 1. / np.sqrt(GAMMA) * (pinf - 1.) / np.sqrt(0.5 * ((GAMMA + 1.) / GAMMA) * pinf + 0.5 * GAMMA1 / GAMMA)
    einf = 0.5 * rinf * vinf ** 2 + pinf / GAMMA1
    return rinf, vinf, einf


def shockbubble_inflow_numpy(state, dim, t, qbc, num_ghost):
    """shockbc (shockbubble.py:68-87): constant post-shock state in the x-lower ghosts."""
    rinf, vinf, einf = _shock_state()
    for i in range(num_ghost):
        qbc[0, i, ...] = rinf
        qbc[1, i, ...] = rinf * vinf
        qbc[2, i, ...] = 0.
        qbc[3, i, ...] = einf
        qbc[4, i, ...] = 0.


def step_euler_radial_numpy(q, dt, rad):
    """step_Euler_radial (shockbubble.py:89-124): geometric source, 2-stage RK."""
    dt2 = dt / 2.
    ndim = 2
    rho = q[0, :, :]
    u = q[1, :, :] / rho
    v = q[2, :, :] / rho
    press = GAMMA1 * (q[3, :, :] - 0.5 * rho * (u ** 2 + v ** 2))
    qstar = np.empty(q.shape)
    qstar[0, :, :] = q[0, :, :] - dt2 * (ndim - 1) / rad * q[2, :, :]
    qstar[1, :, :] = q[1, :, :] - dt2 * (ndim - 1) / rad * rho * u * v
    qstar[2, :, :] = q[2, :, :] - dt2 * (ndim - 1) / rad * rho * v * v
    qstar[3, :, :] = q[3, :, :] - dt2 * (ndim - 1) / rad * v * (q[3, :, :] + press)
    rho = qstar[0, :, :]
    u = qstar[1, :, :] / rho
    v = qstar[2, :, :] / rho
    press = GAMMA1 * (qstar[3, :, :] - 0.5 * rho * (u ** 2 + v ** 2))
    q[0, :, :] = q[0, :, :] - dt * (ndim - 1) / rad * qstar[2, :, :]
    q[1, :, :] = q[1, :, :] - dt * (ndim - 1) / rad * rho * u * v
    q[2, :, :] = q[2, :, :] - dt * (ndim - 1) / rad * rho * v * v
    q[3, :, :] = q[3, :, :] - dt * (ndim - 1) / rad * v * (qstar[3, :, :] + press)
    return q


def dq_euler_radial_numpy(q, dt, rad):
    """dq_Euler_radial (shockbubble.py:127-153): dt times the geometric source."""
    ndim = 2
    rho = q[0, :, :]
    u = q[1, :, :] / rho
    v = q[2, :, :] / rho
    press = GAMMA1 * (q[3, :, :] - 0.5 * rho * (u ** 2 + v ** 2))
    dq = np.empty(q.shape)
    dq[0, :, :] = -dt * (ndim - 1) / rad * q[2, :, :]
    dq[1, :, :] = -dt * (ndim - 1) / rad * rho * u * v
    dq[2, :, :] = -dt * (ndim - 1) / rad * rho * v * v
    dq[3, :, :] = -dt * (ndim - 1) / rad * v * (q[3, :, :] + press)
    dq[4, :, :] = 0
    return dq
