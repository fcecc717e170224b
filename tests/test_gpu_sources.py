"""Source terms as device operations (pyclaw_amd.sources): the Euler radial
source of the shock-bubble application (apps/euler_2d/shockbubble.py:89-153)
as HIP kernels, against the application's numpy routines restated in
tests/apps.py, bit for bit in FP64, and end to end through the drop-in API in
place of the numpy hooks."""
import numpy as np
import pytest

import apps
from test_fixtures_oracle import GOLD
from test_gpu_apps import _compare_frames, amd_run

pytestmark = pytest.mark.gpu


def _state(mx, my, seed, dtype=None):
    import torch
    import pyclaw_amd as pc
    rng = np.random.default_rng(seed)
    x = pc.Dimension('x', 0.0, 2.0, mx)
    y = pc.Dimension('y', 0.0, 0.5, my)
    domain = pc.Domain([x, y])
    state = pc.State(domain, 5, 1, dtype=dtype or torch.float64)
    state.problem_data.update(dict(gamma=apps.GAMMA, gamma1=apps.GAMMA1))
    rho = 0.5 + rng.random((mx, my))
    u = rng.standard_normal((mx, my))
    v = rng.standard_normal((mx, my))
    p = 0.5 + rng.random((mx, my))
    q = np.empty((5, mx, my))
    q[0] = rho
    q[1] = rho * u
    q[2] = rho * v
    q[3] = p / apps.GAMMA1 + 0.5 * rho * (u * u + v * v)
    q[4] = rng.random((mx, my))
    state.q[...] = q
    yc, _ = apps.centers(0.0, 0.5, my)
    state.aux[0, :, :] = yc[None, :]
    return state, q, np.broadcast_to(yc[None, :], (mx, my))


@pytest.mark.parametrize("shape", [(160, 40), (67, 5), (1, 1)])
def test_radial_step_matches_the_application_routine(shape):
    from pyclaw_amd.sources import euler_radial_step
    mx, my = shape
    print("radial step", shape, flush=True)
    state, q, rad = _state(mx, my, mx + my)
    dt = 0.0037
    euler_radial_step(None, state, dt)
    ref = apps.step_euler_radial_numpy(q.copy(), dt, rad)
    assert np.array_equal(state.q, ref)


@pytest.mark.parametrize("shape", [(160, 40), (67, 5)])
def test_radial_dq_matches_the_application_routine(shape):
    from pyclaw_amd.sources import euler_radial_dq
    mx, my = shape
    state, q, rad = _state(mx, my, 3 * mx + my)
    dt = 0.0021
    dq = euler_radial_dq(None, state, dt).cpu().numpy()
    ref = apps.dq_euler_radial_numpy(q, dt, rad)
    assert np.array_equal(dq, ref)
    assert np.array_equal(state.q, q)   # q unchanged


def test_radial_step_f32_close_to_f64():
    import torch
    from pyclaw_amd.sources import euler_radial_step
    state, q, rad = _state(96, 24, 7, dtype=torch.float32)
    euler_radial_step(None, state, 0.003)
    ref = apps.step_euler_radial_numpy(q.copy(), 0.003, rad)
    assert np.abs(state.q - ref).max() <= 1e-5 * np.abs(ref).max()


def test_shockbubble_classic_with_device_source():
    """The shock-bubble application with solver.step_source =
    pyclaw_amd.sources.euler_radial_step: every frame equals the run with the
    application's numpy routine, and the reference's stored density."""
    import pyclaw_amd as pc
    print("app: shockbubble classic, device source", flush=True)
    q0, aux, opts = apps.shockbubble()

    def host_src(solver, state, dt):
        apps.step_euler_radial_numpy(state.q, dt, state.aux[0])

    host, _ = amd_run(q0, aux, opts, step_source=host_src, user_bc_lower=apps.shockbubble_inflow_numpy)
    dev, _ = amd_run(q0, aux, opts, step_source=pc.sources.euler_radial_step,
                     user_bc_lower=apps.shockbubble_inflow_numpy)
    _compare_frames(dev, host)
    err = np.linalg.norm(GOLD['shockbubble_density'] - dev[-1][0])
    print("shockbubble L2 err", err)
    assert err < 1e-12


def test_shockbubble_sharpclaw_with_device_dq_src():
    """SharpClaw shock-bubble (shockbubble.py:171-176: dq_src =
    dq_Euler_radial) with the device dq_src: the same frames as the numpy
    dq_src, bit for bit."""
    import pyclaw_amd as pc
    print("app: shockbubble sharpclaw, device dq_src", flush=True)
    q0, aux, opts = apps.shockbubble()
    opts = dict(opts, kind='sharpclaw', tfinal=0.02, num_output_times=2)

    def host_dq(solver, state, dt):
        return apps.dq_euler_radial_numpy(state.q, dt, state.aux[0])

    host, _ = amd_run(q0, aux, opts, user_bc_lower=apps.shockbubble_inflow_numpy, dq_src=host_dq)
    dev, _ = amd_run(q0, aux, opts, user_bc_lower=apps.shockbubble_inflow_numpy, dq_src=pc.sources.euler_radial_dq)
    assert len(dev) == 3
    _compare_frames(dev, host)
    assert np.isfinite(dev[-1]).all()
