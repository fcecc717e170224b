"""The reference's 2D regression apps, restated as data (initial conditions,
boundary codes, solver options), so the same set-up drives the CPU oracle
(oracle.driver) and the GPU path (pyclaw_amd).

  acoustics_homogeneous -- apps/acoustics_2d_homogeneous/acoustics.py:6-92
  acoustics_variable    -- apps/acoustics_2d_variable/acoustics.py:8-88
  shockbubble           -- apps/euler_2d/shockbubble.py:17-235 (classic variant)
"""
import numpy as np

from pyclaw_amd.problems import GAMMA, GAMMA1, acoustics_pulse, centers, shockbubble as sb_init
# the application's host routines live with the oracle (checker code)
from oracle.shockbubble_ref import (dq_euler_radial_numpy, shockbubble_inflow_numpy,  # noqa: F401
                                    step_euler_radial_numpy)


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


# ---------------------------------------------------------------------------
# The benchmark configurations as applications
# ---------------------------------------------------------------------------
def config1_acoustics(seed=1, transverse=2):
    """Homogeneous acoustics 200x200 on [-1,1]^2: rho=1, K=4 (c=2); ring pulse
    p = (|r-0.5|<=0.2)(1+cos(pi(r-0.5)/0.2)) (acoustics.py:82-92) about a centre
    drawn uniformly from [-0.1,0.1]^2; walls at x- and y-lower, extrapolation
    at the upper sides; unsplit, order 2, MC, CFL 0.9 (cfl_max 1.0), first dt
    0.9 dx / c, to t = 0.6 in 10 frames."""
    mx = my = 200
    rng = np.random.default_rng(seed)
    c0 = tuple(-0.1 + 0.2 * rng.random(2))
    q0 = acoustics_pulse(mx, my, center=c0)
    rho, bulk = 1.0, 4.0
    cc = np.sqrt(bulk / rho)
    dx = dy = 2.0 / mx
    opts = dict(rp='acoustics', num_waves=2, problem_data=dict(rho=rho, bulk=bulk, zz=rho * cc, cc=cc),
                dx=dx, dy=dy, lower=(-1., -1.), upper=(1., 1.), mx=mx, my=my, bc_lower=(3, 3), bc_upper=(1, 1),
                kind='classic', dimensional_split=False, transverse=transverse, limiters=4, cfl_max=1.0,
                cfl_desired=0.9, dt_initial=0.9 * dx / cc, tfinal=0.6, num_output_times=10)
    return q0, None, opts


def config2_shockbubble(mx=640, my=160, tfinal=0.2, num_output_times=4):
    """The plain shock-bubble (no radial source): domain [0,2]x[0,0.5], Roe with
    entropy fix, gamma 1.4, inflow of the post-shock state at x-lower (the
    application's shockbc, a static custom boundary), wall at y-lower,
    extrapolation at the upper sides, unsplit, MC on all five waves, transverse
    correction waves, CFL 0.8 (cfl_max 1.0), first dt at CFL 0.8 of the initial
    data, to t = 0.2."""
    from pyclaw_amd.problems import euler_max_speed
    q0, dx, dy = sb_init(mx, my)
    dt0 = 0.8 * min(dx, dy) / euler_max_speed(q0)
    opts = dict(rp='euler', num_waves=5, problem_data=dict(gamma=GAMMA, gamma1=GAMMA1), dx=dx, dy=dy,
                lower=(0., 0.), upper=(2., 0.5), mx=mx, my=my, bc_lower=(0, 3), bc_upper=(1, 1),
                kind='classic', dimensional_split=False, transverse=2, limiters=4, cfl_max=1.0, cfl_desired=0.8,
                dt_initial=dt0, tfinal=tfinal, num_output_times=num_output_times, source_split=1)
    return q0, None, opts


def config3_shallow(mx=1024, seed=7, nsteps=500, num_output_times=1):
    """SharpClaw shallow water on [0,1]^2, periodic, g=1: h = 1 + 0.5 exp(-|x-x0|^2/0.01)
    with x0 uniform in [0.25,0.75]^2, hu = hv = 0; WENO5 component-wise +
    SSP104, Roe with entropy fix, CFL 0.5 (cfl_max 0.6); run for about nsteps
    steps (tfinal = nsteps steps at the first dt, 0.5 dx / max sqrt(g h))."""
    from pyclaw_amd.problems import shallow_hump
    rng = np.random.default_rng(seed)
    x0 = 0.25 + 0.5 * rng.random(2)
    q0 = shallow_hump(mx, mx, x0)
    dx = 1.0 / mx
    smax = float(np.max(np.sqrt(q0[0])))
    dt0 = 0.5 * dx / smax
    opts = dict(rp='shallow', num_waves=3, problem_data=dict(grav=1.0), dx=dx, dy=dx, lower=(0., 0.),
                upper=(1., 1.), mx=mx, my=mx, bc_lower=(2, 2), bc_upper=(2, 2), kind='sharpclaw', limiters=1,
                cfl_max=0.6, cfl_desired=0.5, dt_initial=dt0, tfinal=nsteps * dt0,
                num_output_times=num_output_times)
    return q0, None, opts
