"""The reference's 2D regression apps, restated as data (initial conditions,
boundary codes, solver options), so the same set-up drives the CPU oracle
(oracle.driver) and the GPU path (pyclaw_amd).

  acoustics_homogeneous -- apps/acoustics_2d_homogeneous/acoustics.py:6-92
  acoustics_variable    -- apps/acoustics_2d_variable/acoustics.py:8-88
  shockbubble           -- apps/euler_2d/shockbubble.py:17-235 (classic variant)
"""
import numpy as np

from pyclaw_amd.problems import GAMMA, GAMMA1, acoustics_pulse, centers, shock_state, shockbubble as sb_init


def acoustics_homogeneous(solver_type='classic'):
    mx = my = 100
    rho, bulk = 1.0, 4.0
    cc = np.sqrt(bulk / rho)
    zz = rho * cc
    dx = dy = 2.0 / 100
    q0 = acoustics_pulse(mx, my, center=(0., 0.), width=0.2)
    opts = dict(rp='acoustics', num_waves=2, problem_data=dict(rho=rho, bulk=bulk, zz=zz, cc=cc),
                dx=dx, dy=dy, lower=(-1., -1.), upper=(1., 1.), mx=mx, my=my,
                bc_lower=(1, 1), bc_upper=(1, 1), kind=solver_type, dimensional_split=True,
                limiters=4, cfl_max=0.5, cfl_desired=0.45,
                dt_initial=min(dx, dy) / cc * 0.45, tfinal=0.12, num_output_times=10)
    if solver_type == 'sharpclaw':
        opts.update(limiters=1)
    return q0, None, opts


def acoustics_variable(solver_type='classic'):
    mx = my = 200
    x, dx = centers(-1.0, 1.0, mx)
    y, dy = centers(-1.0, 1.0, my)
    Y, X = np.meshgrid(y, x)
    rhol, rhor, bulkl, bulkr = 4.0, 1.0, 4.0, 4.0
    cl, cr = np.sqrt(bulkl / rhol), np.sqrt(bulkr / rhor)
    aux = np.zeros((2, mx, my), order='F')
    aux[0] = rhol * (X < 0.) + rhor * (X >= 0.)
    aux[1] = cl * (X < 0.) + cr * (X >= 0.)
    x0, y0 = -0.5, 0.
    r = np.sqrt((X - x0) ** 2 + (Y - y0) ** 2)
    width, rad = 0.1, 0.25
    q0 = np.zeros((3, mx, my), order='F')
    q0[0] = (np.abs(r - rad) <= width) * (1. + np.cos(np.pi * (r - rad) / width))
    opts = dict(rp='vc_acoustics', num_waves=2, problem_data={}, dx=dx, dy=dy, lower=(-1., -1.),
                upper=(1., 1.), mx=mx, my=my, bc_lower=(3, 3), bc_upper=(1, 1), aux_bc_lower=(3, 3),
                aux_bc_upper=(1, 1), kind=solver_type, tfinal=0.6, num_output_times=20)
    if solver_type == 'classic':
        opts.update(dimensional_split=False, transverse=1, limiters=4, cfl_max=1.0, cfl_desired=0.9,
                    dt_initial=0.1)
    else:
        opts.update(limiters=1, cfl_max=2.5, cfl_desired=2.45, dt_initial=0.1)
    return q0, aux, opts


def shockbubble_inflow_numpy(state, dim, t, qbc, num_ghost):
    """shockbc (shockbubble.py:68-87): constant post-shock state in the x-lower ghosts."""
    rinf, vinf, einf = shock_state()
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


def shockbubble():
    mx, my = 160, 40
    q0, dx, dy = sb_init(mx, my)
    y, _ = centers(0.0, 0.5, my)
    aux = np.zeros((1, mx, my), order='F')
    aux[0, :, :] = y[None, :]
    # gamma1 = gamma - 1. as the application sets it (shockbubble.py:6-7, 191-192)
    opts = dict(rp='euler', num_waves=5, problem_data=dict(gamma=GAMMA, gamma1=GAMMA1), dx=dx, dy=dy,
                lower=(0., 0.), upper=(2., 0.5), mx=mx, my=my, bc_lower=(0, 3), bc_upper=(1, 1),
                aux_bc_lower=(1, 1), aux_bc_upper=(1, 1), kind='classic', dimensional_split=True,
                limiters=[4, 4, 4, 4, 2], cfl_max=0.5, cfl_desired=0.45, dt_initial=0.005, tfinal=0.2,
                num_output_times=1, source_split=1)
    return q0, aux, opts
