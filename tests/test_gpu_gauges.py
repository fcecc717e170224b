"""Gauges (solver.py:713-742, geometry.py:336-374): values at gauge cells
recorded every accepted step.  With the default compute_gauge_values they are
gathered on the device and written out in batches; a user function runs per
step on host values.  Both must give the same gauge files, byte for byte, and
the same keep_gauges data."""
import os

import numpy as np
import pytest

import apps
from test_gpu_apps import RP

pytestmark = pytest.mark.gpu

GAUGES = [(0.0, 0.0), (0.31, -0.42), (-0.77, 0.5), (0.5, 0.9)]


def _run(outdir, custom, device_loop=True, kind='classic'):
    import pyclaw_amd as pc
    q0, aux, opts = apps.acoustics_variable(kind)
    x = pc.Dimension('x', opts['lower'][0], opts['upper'][0], opts['mx'])
    y = pc.Dimension('y', opts['lower'][1], opts['upper'][1], opts['my'])
    domain = pc.Domain([x, y])
    state = pc.State(domain, q0.shape[0], aux.shape[0])
    state.problem_data.update(opts['problem_data'])
    state.q[...] = q0
    state.aux[...] = aux
    state.keep_gauges = True
    domain.grid.add_gauges(GAUGES)
    if kind == 'classic':
        solver = pc.ClawSolver2D(getattr(pc.riemann, RP[opts['rp']]))
        solver.dimensional_split = opts.get('dimensional_split', True)
        solver.transverse_waves = opts.get('transverse', 1)
    else:
        solver = pc.SharpClawSolver2D(getattr(pc.riemann, RP[opts['rp']]))
    solver.limiters = opts['limiters']
    solver.cfl_max, solver.cfl_desired = opts['cfl_max'], opts['cfl_desired']
    solver.dt_initial = opts['dt_initial']
    solver.bc_lower, solver.bc_upper = list(opts['bc_lower']), list(opts['bc_upper'])
    solver.aux_bc_lower = list(opts.get('aux_bc_lower', (1, 1)))
    solver.aux_bc_upper = list(opts.get('aux_bc_upper', (1, 1)))
    solver.device_loop = device_loop
    if custom:
        solver.compute_gauge_values = lambda q, a: q      # a user function: the per-step host path
    claw = pc.Controller()
    claw.solution = pc.Solution(state, domain)
    claw.solver = solver
    claw.output_format = None
    claw.outdir = str(outdir)
    claw.keep_copy = True
    claw.tfinal = 0.2
    claw.num_output_times = 4
    claw.run()
    files = {}
    gdir = os.path.join(str(outdir), '_gauges')
    for name in sorted(os.listdir(gdir)):
        with open(os.path.join(gdir, name), 'rb') as f:
            files[name] = f.read()
    return files, [np.array(g) for g in state.gauge_data], claw


@pytest.mark.parametrize("kind,device_loop", [('classic', True), ('classic', False), ('sharpclaw', True)])
def test_device_recorded_gauges_equal_the_per_step_path(tmp_path, kind, device_loop):
    """Default compute_gauge_values: recorded inside the device step loop
    (device_loop) or gathered on the device after each host-loop step; a user
    compute_gauge_values: the per-step host path."""
    print("gauges: vc acoustics", kind, "device loop", device_loop, flush=True)
    dev_files, dev_data, claw = _run(tmp_path / 'dev', custom=False, device_loop=device_loop, kind=kind)
    host_files, host_data, _ = _run(tmp_path / 'host', custom=True, kind=kind)
    assert len(dev_files) == len(GAUGES)
    assert dev_files == host_files
    assert len(dev_data) == len(host_data) == len(GAUGES)
    for a, b in zip(dev_data, host_data):
        assert np.array_equal(a, b)
    # one line per accepted step plus the initial values; the last line holds the final frame's values
    nsteps = sum(1 for _ in dev_files[sorted(dev_files)[0]].splitlines())
    assert nsteps == dev_data[0].shape[0] > 10
    final = claw.frames[-1].state.q
    for (gx, gy), data in zip(claw.solution.state.grid.gauges, dev_data):
        assert np.array_equal(data[-1, 1:], final[:, gx, gy])
        assert data[-1, 0] == claw.frames[-1].t
