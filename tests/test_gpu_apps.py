"""The reference's 2D regression apps run end to end through the drop-in API
(pyclaw_amd Controller / State / ClawSolver2D / SharpClawSolver2D) on the GPU,
checked two ways:
  * against the reference's stored data with the reference's own tolerances;
  * against the CPU oracle's run of the same set-up: every output frame must
    be identical (FP64 kernels keep the reference's operation order).
"""
import os

import numpy as np
import pytest

import apps
from test_fixtures_oracle import GOLD, oracle_run
from oracle import driver

pytestmark = pytest.mark.gpu

RP = {'acoustics': 'rp2_acoustics', 'vc_acoustics': 'rp2_vc_acoustics', 'euler': 'rp2_euler_5wave',
      'shallow': 'rp2_shallow_roe_with_efix'}


def amd_run(q0, aux, opts, step_source=None, user_bc_lower=None, autotune=False, fma=False, dq_src=None):
    claw = amd_controller(q0, aux, opts, step_source, user_bc_lower, autotune, fma, dq_src)
    status = claw.run()
    return [f.state.q.copy() for f in claw.frames], status


def amd_controller(q0, aux, opts, step_source=None, user_bc_lower=None, autotune=False, fma=False, dq_src=None,
                   **solver_attrs):
    """The application's Controller, set up and not yet run; solver_attrs are
    set on the solver last."""
    import pyclaw_amd as pc
    x = pc.Dimension('x', opts['lower'][0], opts['upper'][0], opts['mx'])
    y = pc.Dimension('y', opts['lower'][1], opts['upper'][1], opts['my'])
    domain = pc.Domain([x, y])
    state = pc.State(domain, q0.shape[0], 0 if aux is None else aux.shape[0])
    state.problem_data.update(opts['problem_data'])
    state.q[...] = q0
    if aux is not None:
        state.aux[...] = aux
    rp = getattr(pc.riemann, RP[opts['rp']])
    if opts['kind'] == 'classic':
        solver = pc.ClawSolver2D(rp)
        solver.dimensional_split = opts.get('dimensional_split', True)
        solver.transverse_waves = opts.get('transverse', 1)
        solver.source_split = opts.get('source_split', 1)
        if step_source is not None:
            solver.step_source = step_source
    else:
        solver = pc.SharpClawSolver2D(rp)
        solver.time_integrator = opts.get('time_integrator', 'SSP104')
        if dq_src is not None:
            solver.dq_src = dq_src
    solver.limiters = opts['limiters']
    solver.autotune = autotune
    solver.fma = fma
    solver.cfl_max = opts['cfl_max']
    solver.cfl_desired = opts['cfl_desired']
    solver.dt_initial = opts['dt_initial']
    solver.bc_lower = list(opts['bc_lower'])
    solver.bc_upper = list(opts['bc_upper'])
    solver.aux_bc_lower = list(opts.get('aux_bc_lower', (1, 1)))
    solver.aux_bc_upper = list(opts.get('aux_bc_upper', (1, 1)))
    if user_bc_lower is not None:
        solver.user_bc_lower = user_bc_lower
    claw = pc.Controller()
    claw.solution = pc.Solution(state, domain)
    claw.solver = solver
    claw.output_format = None
    claw.keep_copy = True
    claw.tfinal = opts['tfinal']
    claw.num_output_times = opts['num_output_times']
    for k, v in solver_attrs.items():
        setattr(solver, k, v)
    return claw


def _compare_frames(got, ref):
    assert len(got) == len(ref)
    for k, (a, b) in enumerate(zip(got, ref)):
        assert np.array_equal(a, b), "frame %d differs: max %g" % (k, np.abs(a - b).max())


def test_homogeneous_acoustics_classic():
    print("app: acoustics_2d_homogeneous classic", flush=True)
    q0, aux, opts = apps.acoustics_homogeneous('classic')
    got, status = amd_run(q0, aux, opts)
    ok, err = driver.check_diff(GOLD['acoustics_homog_classic_p'], got[-1][0], reltol=1e-3)
    assert ok, err
    ref, _ = oracle_run(q0, aux, opts)
    _compare_frames(got, ref)


def test_variable_acoustics_classic():
    print("app: acoustics_2d_variable classic", flush=True)
    q0, aux, opts = apps.acoustics_variable('classic')
    got, status = amd_run(q0, aux, opts)
    assert np.abs(GOLD['acoustics_vc_pressure_classic'] - got[20][0]).max() < 1e-1
    ok, err = driver.check_diff(GOLD['acoustics_vc_fortq0020'], got[20], reltol=1e-4)
    assert ok, err
    ref, _ = oracle_run(q0, aux, opts)
    _compare_frames(got, ref)


def test_shockbubble_classic():
    print("app: euler_2d shockbubble classic", flush=True)
    q0, aux, opts = apps.shockbubble()

    def src(solver, state, dt):
        apps.step_euler_radial_numpy(state.q, dt, state.aux[0])

    got, status = amd_run(q0, aux, opts, step_source=src, user_bc_lower=apps.shockbubble_inflow_numpy)
    err = np.linalg.norm(GOLD['shockbubble_density'] - got[-1][0])
    print("shockbubble L2 err", err)
    assert err < 1e-12


def test_homogeneous_acoustics_sharpclaw():
    print("app: acoustics_2d_homogeneous sharpclaw", flush=True)
    q0, aux, opts = apps.acoustics_homogeneous('sharpclaw')
    got, status = amd_run(q0, aux, opts)
    ok, err = driver.check_diff(GOLD['acoustics_homog_sharpclaw_p'], got[-1][0], reltol=1e-3)
    assert ok, err
    ref, _ = oracle_run(q0, aux, opts)
    _compare_frames(got, ref)


def test_variable_acoustics_sharpclaw():
    print("app: acoustics_2d_variable sharpclaw", flush=True)
    q0, aux, opts = apps.acoustics_variable('sharpclaw')
    got, status = amd_run(q0, aux, opts)
    assert np.abs(GOLD['acoustics_vc_pressure_classic'] - got[20][0]).max() < 1e-1
    ref, _ = oracle_run(q0, aux, opts)
    _compare_frames(got, ref)


@pytest.mark.parametrize("kind", ['classic', 'sharpclaw'])
def test_autotune_keeps_frames(kind):
    """solver.autotune = True times the kernel's launch shapes on the first
    step; every frame must match the untuned run bit for bit."""
    print("app: acoustics_2d_variable %s autotune" % kind, flush=True)
    q0, aux, opts = apps.acoustics_variable(kind)
    ref, _ = amd_run(q0, aux, opts)
    got, _ = amd_run(q0, aux, opts, autotune=True)
    _compare_frames(got, ref)


@pytest.mark.parametrize("kind,tol", [('classic', 1e-11), ('sharpclaw', 1e-6)])
def test_fma_app_frames_close(kind, tol):
    """solver.fma = True runs the fused-multiply-add kernels through the drop-in
    API: every frame stays close to the bit-exact run (classic: rounding level;
    SharpClaw: WENO5 amplifies rounding differences on this smooth pulse) and
    differs from it somewhere (the fused build really ran)."""
    print("app: acoustics_2d_variable %s fma" % kind, flush=True)
    q0, aux, opts = apps.acoustics_variable(kind)
    ref, _ = amd_run(q0, aux, opts)
    got, _ = amd_run(q0, aux, opts, fma=True)
    assert len(got) == len(ref)
    worst = max(float(np.abs(a - b).max() / np.abs(b).max()) for a, b in zip(got, ref))
    print("fma app worst frame rel Linf", worst, flush=True)
    assert worst <= tol
    assert any(not np.array_equal(a, b) for a, b in zip(got, ref))


@pytest.mark.parametrize("integrator", ['Euler', 'SSP33'])
@pytest.mark.parametrize("app", ['acoustics_homogeneous', 'shallow_hump', 'shallow_hump_rejected_first_step'])
def test_sharpclaw_euler_and_ssp33_frames_equal_oracle(integrator, app):
    """SharpClawSolver2D.time_integrator 'Euler' and 'SSP33'
    (sharpclaw/solver.py:199-218) through the drop-in API: every frame equal to
    the oracle's run, fused-stage path and the stage-by-stage path (taken when
    a hook runs before each stage)."""
    print("app: sharpclaw %s %s" % (integrator, app), flush=True)
    if app == 'acoustics_homogeneous':
        q0, aux, opts = apps.acoustics_homogeneous('sharpclaw')
        cfl = dict(cfl_max=0.5, cfl_desired=0.45) if integrator == 'Euler' else {}
    else:
        q0, aux, opts = apps.config3_shallow(96, nsteps=24, num_output_times=3)
        cfl = dict(cfl_max=0.45, cfl_desired=0.4) if integrator == 'Euler' else dict(cfl_max=1.0, cfl_desired=0.9)
    opts = dict(opts, time_integrator=integrator, **cfl)
    if app.endswith('rejected_first_step'):
        opts['dt_initial'] = 10 * opts['dt_initial']
    ref, run = oracle_run(q0, aux, opts)
    assert run.nsteps >= 3
    got, _ = amd_run(q0, aux, opts)
    _compare_frames(got, ref)
    staged = amd_controller(q0, aux, opts, call_before_step_each_stage=True)
    staged.run()
    _compare_frames([f.state.q.copy() for f in staged.frames], ref)
