"""evolve_to_time's step loop with its bookkeeping on the device
(claw2d_loop_advance) against the host loop (src/pyclaw/solver.py:629-696):
the same frames bit for bit, the same status dict, final dt, time and Courant
number, including rejected steps (a first dt far above the stable one), the
clip of the last step to each output time and the max_steps limit."""
import numpy as np
import pytest

import apps
from test_gpu_apps import amd_controller

pytestmark = pytest.mark.gpu


def _run(case, device_loop, **kw):
    q0, aux, opts = case
    claw = amd_controller(q0, aux, opts, device_loop=device_loop, **kw)
    err = None
    try:
        claw.run()
    except Exception as e:   # the max_steps limit: both loops must raise it at the same step
        err = str(e)
    s = claw.solver
    frames = [f.state.q.copy() for f in claw.frames] + [claw.solution.state.q.copy()]
    return frames, dict(s.status), s.dt, claw.solution.t, s.cfl.get_cached_max(), err


def _same(a, b):
    fa, sa, dta, ta, ca, ea = a
    fb, sb, dtb, tb, cb, eb = b
    assert len(fa) == len(fb)
    for k, (x, y) in enumerate(zip(fa, fb)):
        assert np.array_equal(x, y), "frame %d differs: max %g" % (k, np.abs(x - y).max())
    assert sa == sb and dta == dtb and ta == tb and ca == cb and ea == eb, (sa, sb, dta, dtb, ta, tb, ca, cb, ea, eb)


def _cases():
    out = {}
    q0, aux, opts = apps.acoustics_homogeneous('classic')
    out['acoustics'] = (q0, aux, opts)
    big = dict(opts, dt_initial=0.1)          # the first step is rejected, its dt cut to the stable one
    out['acoustics_rejected_first_step'] = (q0, aux, big)
    out['vc_acoustics'] = apps.acoustics_variable('classic')
    out['acoustics_unsplit'] = (q0, aux, dict(opts, dimensional_split=False, transverse=2))
    q0, aux, opts = apps.acoustics_homogeneous('sharpclaw')
    out['sharpclaw_acoustics'] = (q0, aux, opts)
    out['sharpclaw_acoustics_rejected_first_step'] = (q0, aux, dict(opts, dt_initial=0.1))
    out['sharpclaw_vc_acoustics'] = apps.acoustics_variable('sharpclaw')
    return out


@pytest.mark.parametrize("name", ['acoustics', 'acoustics_rejected_first_step', 'vc_acoustics',
                                  'acoustics_unsplit', 'sharpclaw_acoustics',
                                  'sharpclaw_acoustics_rejected_first_step', 'sharpclaw_vc_acoustics'])
@pytest.mark.parametrize("fma", [False, True])
def test_device_loop_equals_host_loop(name, fma):
    print("loop:", name, "fma", fma, flush=True)
    case = _cases()[name]
    dev = _run(case, True, fma=fma)
    host = _run(case, False, fma=fma)
    assert dev[5] is None and dev[1]['numsteps'] >= 1
    _same(dev, host)


def test_device_loop_with_static_custom_inflow():
    """Euler shock-bubble without its source term: the inflow ghost cells are
    filled once by the user function (custom_bc_static) and the steps run on
    the device."""
    print("loop: shockbubble static inflow", flush=True)
    q0, aux, opts = apps.shockbubble()
    opts = dict(opts, tfinal=0.05)
    kw = dict(user_bc_lower=apps.shockbubble_inflow_numpy, custom_bc_static=True)
    _same(_run((q0, aux, opts), True, **kw), _run((q0, aux, opts), False, **kw))


@pytest.mark.parametrize("name", ['acoustics', 'acoustics_rejected_first_step', 'sharpclaw_acoustics'])
def test_device_loop_stops_at_max_steps(name):
    """max_steps iterations end the loop short of tend.  solver.py:697-699
    raises only when every one of them was accepted; with a rejected step the
    run ends early without an error, in both loops."""
    print("loop: max_steps", name, flush=True)
    q0, aux, opts = _cases()[name]
    case = (q0, aux, dict(opts, num_output_times=1))       # one evolve_to_time call to t=0.12
    dev = _run(case, True, max_steps=3)
    host = _run(case, False, max_steps=3)
    assert dev[3] < opts['tfinal']
    if name in ('acoustics', 'sharpclaw_acoustics'):
        assert dev[5] is not None and 'Maximum number of timesteps' in dev[5]
    else:
        assert dev[5] is None
    _same(dev, host)


@pytest.mark.parametrize("kind", ['classic', 'sharpclaw'])
def test_device_loop_is_used(kind):
    """The device loop runs (no host fallback) for the homogeneous acoustics app."""
    import pyclaw_amd.solver as S
    calls = []
    orig = S.Solver._device_step_loop

    def spy(self, *a, **k):
        calls.append(1)
        return orig(self, *a, **k)
    S.Solver._device_step_loop = spy
    try:
        q0, aux, opts = apps.acoustics_homogeneous(kind)
        amd_controller(q0, aux, opts).run()
    finally:
        S.Solver._device_step_loop = orig
    assert len(calls) == opts['num_output_times']


def _fixed(opts, steps_per_frame=None, dt=None):
    """Fixed-dt options: dt divides the output interval."""
    interval = opts['tfinal'] / opts['num_output_times']
    if dt is None:
        dt = interval / steps_per_frame
    return dict(opts, dt_initial=dt, dt_variable=False)


@pytest.mark.parametrize("fma", [False, True])
def test_device_loop_with_fixed_dt(fma):
    """dt_variable = False: t = tstart + n dt (solver.py:655-659), the clip of
    the last step to each output time, no dt rule."""
    print("loop: fixed dt fma", fma, flush=True)
    q0, aux, opts = apps.acoustics_homogeneous('classic')
    interval = opts['tfinal'] / opts['num_output_times']
    k = int(np.ceil(interval / opts['dt_initial']))
    case = (q0, aux, _fixed(opts, steps_per_frame=k))
    kw = dict(dt_variable=False, fma=fma)
    dev = _run(case, True, **kw)
    host = _run(case, False, **kw)
    assert dev[5] is None
    _same(dev, host)


def test_device_loop_fixed_dt_above_cfl_max_fails_like_the_host_loop():
    """A fixed dt above the stable one: the first step's Courant number is above
    cfl_max, both loops raise 'CFL too large' with that step's result kept."""
    print("loop: fixed dt CFL failure", flush=True)
    q0, aux, opts = apps.acoustics_homogeneous('classic')
    big = 2.0 * opts['dt_initial'] / 0.45          # Courant number about 2 (dt_initial is 0.45 dx / c)
    case = (q0, aux, dict(opts, dt_initial=big, tfinal=4 * big, num_output_times=1))
    dev = _run(case, True, dt_variable=False)
    host = _run(case, False, dt_variable=False)
    assert dev[5] is not None and 'CFL too large' in dev[5]
    _same(dev, host)
