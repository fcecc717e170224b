"""The benchmark configurations on their own workloads, through the drop-in API
(Controller / ClawSolver2D / SharpClawSolver2D) on the GPU, against the oracle:

  C1  homogeneous acoustics 200x200, ring pulse about a seeded centre, walls
      low / extrapolation high, unsplit MC, CFL 0.9 to t=0.6
      (apps/acoustics_2d_homogeneous/acoustics.py:82-92): every frame equal.
  C2  plain shock-bubble (apps/euler_2d/shockbubble.py:17-87, no source term),
      unsplit MC on 5 waves, transverse correction waves, CFL 0.8 to t=0.2 at
      640x160: reference arithmetic frame for frame equal, fused build within
      1e-12; 4096x1024: three steps of the default (large-grid) launch shape and
      of a pinned shape equal to the oracle.
  C3  SharpClaw shallow-water hump, periodic, WENO5 + SSP104 (sharpclaw/solver.py:221-251):
      128x128 frames equal to the oracle; 1024x1024 for about 500 steps with
      total h, hu, hv kept to 1e-13 relative to total h (compensated sums), and
      the large-grid default launch shape's dq equal to a pinned shape's and the
      oracle's.
The oracle runs on all the host's cores (bit-identical to one thread).
"""
import ctypes
import math

import numpy as np
import pytest

import apps
import oracle as O
from helpers import NG, _bc_struct, _dev_grid, _rp_struct, from_device_interior, gpu_classic_step, \
    oracle_classic_step, padded_with_bc, RP_CASES, to_device_padded
from test_fixtures_oracle import oracle_run
from test_gpu_apps import amd_controller

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _oracle_threads():
    old = O.set_threads(O.host_threads())
    yield
    O.set_threads(old)


def _frames(claw):
    claw.run()
    return [f.state.q.copy() for f in claw.frames]


def _compare(got, ref):
    assert len(got) == len(ref)
    for k, (a, b) in enumerate(zip(got, ref)):
        assert np.array_equal(a, b), "frame %d differs: max %g" % (k, np.abs(a - b).max())


def _rel_per_field(a, b):
    return [float(np.abs(a[m] - b[m]).max() / max(np.abs(b[m]).max(), 1e-300)) for m in range(a.shape[0])]


@pytest.mark.parametrize("trans", [1, 2])
@pytest.mark.parametrize("device_loop", [True, False])
def test_config1_acoustics_200_to_t06(trans, device_loop):
    print("config 1: acoustics 200x200 to t=0.6, transverse", trans, "device loop", device_loop, flush=True)
    q0, aux, opts = apps.config1_acoustics(seed=1, transverse=trans)
    claw = amd_controller(q0, aux, opts, device_loop=device_loop)
    got = _frames(claw)
    ref, run = oracle_run(q0, aux, opts)
    assert run.nsteps > 100
    _compare(got, ref)
    assert claw.solution.t == run.t


def test_config2_shockbubble_640x160_to_t02():
    print("config 2: shock-bubble 640x160 to t=0.2 (exact and fused builds)", flush=True)
    q0, aux, opts = apps.config2_shockbubble(640, 160)
    kw = dict(user_bc_lower=apps.shockbubble_inflow_numpy)
    ref, run = oracle_run(q0, aux, opts, **kw)
    print("oracle steps", run.nsteps, flush=True)
    claw = amd_controller(q0, aux, opts, custom_bc_static=True, **kw)
    got = _frames(claw)
    _compare(got, ref)
    assert claw.solution.t == run.t
    claw = amd_controller(q0, aux, opts, fma=True, custom_bc_static=True, **kw)
    fused = _frames(claw)
    assert len(fused) == len(ref)
    worst = max(max(_rel_per_field(a, b)) for a, b in zip(fused, ref))
    print("fused build vs oracle, worst frame relative L-inf per field: %.3g" % worst, flush=True)
    assert worst <= 1e-12
    assert any(not np.array_equal(a, b) for a, b in zip(fused, ref))


def _launch(ctx, mx, my, case, prec):
    from pyclaw_amd import _lib
    band, w, wpb = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    m = _lib.ClawMethod()
    m.order, m.transverse, m.split, m.num_waves = 2, 2, 0, case['nw']
    for k in range(case['nw']):
        m.mthlim[k] = 4
    _lib.check(_lib.lib().claw2d_classic_get_launch(ctx.handle, _dev_grid(mx, my, 1.0, 1.0), m, _rp_struct(case),
                                                   prec, ctypes.byref(band), ctypes.byref(w), ctypes.byref(wpb)),
               'claw2d_classic_get_launch')
    return band.value, w.value, wpb.value


def test_config2_shockbubble_4096x1024_steps_default_and_pinned_shape():
    """Three steps at the headline size with the default launch shape (the
    large-grid default: the paired fill decomposition, one-wavefront blocks
    of the two-waves build) and one step with each of two pinned shapes (18-row
    bands in about two rounds of wavefronts, the previous default; 36-row
    bands of the one-wave build in 4-wavefront blocks), each equal to the
    oracle."""
    from pyclaw_amd import _lib
    from pyclaw_amd.problems import euler_max_speed, shock_state
    print("config 2: shock-bubble 4096x1024, 3 exact steps", flush=True)
    mx, my = 4096, 1024
    q, dx, dy = apps.sb_init(mx, my)
    dt = 0.8 * min(dx, dy) / euler_max_speed(q)
    case = RP_CASES['euler']
    ctx = _lib.Context(0)
    assert _launch(ctx, mx, my, case, _lib.F64) == (0, 2, 1)
    inflow = np.array(shock_state())
    inflow = np.array([inflow[0], inflow[0] * inflow[1], 0.0, inflow[2], 0.0])

    def padded(qi):
        qbc = np.zeros((5, mx + 4, my + 4), order='F')
        qbc[:, 2:-2, 2:-2] = qi
        qbc[:, :2, :] = inflow[:, None, None]
        O.fill_bc(qbc, 2, [99, 3], [1, 1], True)
        return qbc

    O.set_params(O.RP_EULER, case['pd'])   # gamma1 = 0.4 as the device step takes it
    pinned = None
    for k in range(3):
        qbc = padded(q)
        ref, cref = O.step2(O.RP_EULER, qbc, None, dx, dy, dt, O.method_array(2, 2, False), O.mthlim_array(4, 5), 5)
        ref = ref[:, 2:-2, 2:-2]
        got, cgot = gpu_classic_step(qbc, None, case, dt, 2, 2, bc_lower=(0, 3), bc_upper=(1, 1), dx=dx, dy=dy)
        assert np.array_equal(got, ref) and cgot == cref, "step %d max %g" % (k, np.abs(got - ref).max())
        if k == 0:
            for shape in ((18, 2, 1), (36, 1, 4)):
                pinned = _pinned_step(qbc, case, dt, dx, dy, shape)
                assert np.array_equal(pinned[0], ref) and pinned[1] == cref, shape
        print("step", k, "equal, cfl", cgot, flush=True)
        q = ref


def _pinned_step(qbc, case, dt, dx, dy, shape):
    """gpu_classic_step with a launch shape set on the context first."""
    import torch
    from pyclaw_amd import _lib
    ne, nxp, nyp = qbc.shape
    mx, my = nxp - 4, nyp - 4
    pad = np.zeros((ne, mx + 2 * NG, my + 2 * NG), order='F')
    pad[:, NG - 2:NG - 2 + nxp, NG - 2:NG - 2 + nyp] = qbc
    qd = to_device_padded(pad)
    qn = torch.zeros_like(qd)
    m = _lib.ClawMethod()
    m.order, m.transverse, m.split, m.num_waves = 2, 2, 0, case['nw']
    for k in range(case['nw']):
        m.mthlim[k] = 4
    grid = _dev_grid(mx, my, dx, dy)
    ctx = _lib.Context(0)
    _lib.check(_lib.lib().claw2d_classic_set_launch_ex(ctx.handle, grid, m, _rp_struct(case), _lib.F64, *shape),
               'claw2d_classic_set_launch_ex')
    assert _launch(ctx, mx, my, case, _lib.F64) == shape
    dtd = torch.tensor(dt, dtype=torch.float64, device='cuda')
    cfl = torch.zeros((), dtype=torch.float64, device='cuda')
    _lib.check(_lib.lib().claw2d_classic_step(ctx.handle, grid, m, _rp_struct(case), _bc_struct((0, 3), (1, 1)),
                                              _lib.F64, qd.data_ptr(), qn.data_ptr(), None, dtd.data_ptr(),
                                              cfl.data_ptr(), torch.cuda.current_stream().cuda_stream),
               'claw2d_classic_step')
    torch.cuda.synchronize()
    return from_device_interior(qn, mx, my), float(cfl.item())


@pytest.mark.parametrize("device_loop", [True, False])
def test_config3_shallow_128_frames_equal_oracle(device_loop):
    print("config 3: shallow water 128x128 SSP104, device loop", device_loop, flush=True)
    q0, aux, opts = apps.config3_shallow(128, nsteps=20, num_output_times=2)
    claw = amd_controller(q0, aux, opts, device_loop=device_loop)
    got = _frames(claw)
    ref, run = oracle_run(q0, aux, opts)
    assert run.nsteps >= 15
    _compare(got, ref)
    assert claw.solution.t == run.t


def test_config3_shallow_1024_conservation_500_steps():
    print("config 3: shallow water 1024x1024, about 500 SSP104 steps", flush=True)
    q0, aux, opts = apps.config3_shallow(1024, nsteps=500)
    claw = amd_controller(q0, aux, opts)
    got = _frames(claw)
    n = claw.solver.status['numsteps']
    qf = got[-1]
    assert n >= 450 and np.all(np.isfinite(qf))
    tot0 = [math.fsum(q0[m].ravel()) for m in range(3)]
    tot1 = [math.fsum(qf[m].ravel()) for m in range(3)]
    rel = [abs(tot1[m] - tot0[m]) / tot0[0] for m in range(3)]
    print("steps", n, "relative change of total h, hu, hv (over total h):", rel, flush=True)
    assert max(rel) <= 1e-13


def test_config3_sharpclaw_large_grid_default_shape():
    """claw2d_sharpclaw_dq at 1024x1024: the default launch shape for large grids
    (the paired fill decomposition of the default build, one-wavefront blocks)
    against pinned shapes (16-row bands of the one-wave build, 7-row bands of
    the three-waves build, the equal-piece fill of the three-waves build) and
    the oracle's dq, bit for bit."""
    import torch
    from pyclaw_amd import _lib
    print("config 3: sharpclaw dq 1024x1024 default vs pinned shape vs oracle", flush=True)
    q0, aux, opts = apps.config3_shallow(1024, nsteps=1)
    case = RP_CASES['shallow']
    mx = my = 1024
    dx = 1.0 / mx
    pad = np.zeros((3, mx + 2 * NG, my + 2 * NG), order='F')
    pad[:, NG:-NG, NG:-NG] = q0
    O.fill_bc(pad, NG, [2, 2], [2, 2], True)
    O.set_params(O.RP_SHALLOW, dict(grav=1.0))
    dt = opts['dt_initial']
    ref, cref = O.sharp_dq(O.RP_SHALLOW, pad, None, dx, dx, dt, 3)
    grid, rp, bc = _dev_grid(mx, my, dx, dx), _rp_struct(case), _bc_struct((2, 2), (2, 2))
    qd = to_device_padded(pad)
    dtd = torch.tensor(dt, dtype=torch.float64, device='cuda')
    out = []
    for shape in (None, (16, 1, 4), (7, 3, 4), (0, 3, 1)):
        ctx = _lib.Context(0)
        band, w, wpb = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        if shape is not None:
            _lib.check(_lib.lib().claw2d_sharpclaw_set_launch_ex(ctx.handle, grid, rp, _lib.F64, *shape),
                       'claw2d_sharpclaw_set_launch_ex')
        _lib.check(_lib.lib().claw2d_sharpclaw_get_launch(ctx.handle, grid, rp, _lib.F64, ctypes.byref(band),
                                                          ctypes.byref(w), ctypes.byref(wpb)),
                   'claw2d_sharpclaw_get_launch')
        assert (band.value, w.value, wpb.value) == (shape if shape is not None else (0, 1, 1))
        dq = torch.zeros_like(qd)
        cfl = torch.zeros((), dtype=torch.float64, device='cuda')
        _lib.check(_lib.lib().claw2d_sharpclaw_dq(ctx.handle, grid, rp, bc, _lib.F64, qd.data_ptr(), dq.data_ptr(),
                                                  None, dtd.data_ptr(), cfl.data_ptr(),
                                                  torch.cuda.current_stream().cuda_stream), 'claw2d_sharpclaw_dq')
        torch.cuda.synchronize()
        out.append((from_device_interior(dq, mx, my), float(cfl.item())))
    for got, cgot in out:
        assert np.array_equal(got, ref) and cgot == cref


@pytest.mark.parametrize("name,trans", [("vc_acoustics", 1), ("acoustics", 2), ("shallow", 1)])
def test_midsize_default_band_rule_equals_oracle(name, trans):
    """Mid-size grids (512x512, 9 strips): the default launch shape sizes its
    row bands for about one wavefront per SIMD (5 rows here, where two per
    SIMD would give 4-row bands); two steps of the light solvers' one-block
    row march at that shape (its odd and even row counts, walls low,
    extrapolation high, MC, transverse increment or correction waves) equal
    the oracle bit for bit."""
    from pyclaw_amd import _lib
    print("mid-size 512x512 default band rule:", name, "transverse", trans, flush=True)
    mx = my = 512
    case = RP_CASES[name]
    rng = np.random.default_rng(11)
    ctx = _lib.Context(0)
    band, w, wpb = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    m = _lib.ClawMethod()
    m.order, m.transverse, m.split, m.num_waves = 2, trans, 0, case['nw']
    for k in range(case['nw']):
        m.mthlim[k] = 4
    _lib.check(_lib.lib().claw2d_classic_get_launch(ctx.handle, _dev_grid(mx, my, 1.0, 1.0), m, _rp_struct(case),
                                                   _lib.F64, ctypes.byref(band), ctypes.byref(w), ctypes.byref(wpb)),
               'claw2d_classic_get_launch')
    assert 5 <= band.value <= 7, band.value
    q = case['make'](rng, mx, my)
    aux = None
    if case['naux']:
        from helpers import vc_aux
        aux = vc_aux(rng, mx, my)
    dx, dy = 1.0 / mx, 1.0 / my
    dt = 0.3 * dx
    for k in range(2):
        qbc = padded_with_bc(q, 2, (3, 3), (1, 1), True)
        auxbc = padded_with_bc(aux, 2, (3, 3), (1, 1), False) if aux is not None else None
        ref, cref = oracle_classic_step(qbc, auxbc, case, dt, 2, trans, dx=dx, dy=dy)
        got, cgot = gpu_classic_step(qbc, auxbc, case, dt, 2, trans, bc_lower=(3, 3), bc_upper=(1, 1),
                                     abc_lower=(3, 3), abc_upper=(1, 1), dx=dx, dy=dy)
        assert np.array_equal(got, ref) and cgot == cref, "step %d max %g" % (k, np.abs(got - ref).max())
        q = ref



@pytest.mark.parametrize("name,prec", [("euler", "f64"), ("euler", "f64fma"), ("shallow", "f64")])
def test_fill_consecutive_steps_equal_oracle(name, prec):
    """Six consecutive steps on one context through the paired fill
    decomposition (shape 0,2,1, 2000x512, walls low, extrapolation high, MC,
    transverse correction waves): every step equals the oracle bit for bit
    (the fused Euler build: to its rounding) and equals the step with 18-row
    bands on another context bit for bit (the fused Euler build's answers do
    not depend on even band starts)."""
    import torch
    from pyclaw_amd import _lib
    case = RP_CASES[name]
    mx, my = 2000, 512
    rng = np.random.default_rng(77)
    q = case['make'](rng, mx, my)
    dx, dy = 1.0 / mx, 1.0 / my
    dt = (0.1 if name == 'euler' else 0.3) * min(dx, dy)
    P = _lib.F64_FMA if prec == 'f64fma' else _lib.F64
    m = _lib.ClawMethod()
    m.order, m.transverse, m.split, m.num_waves = 2, 2, 0, case['nw']
    for k in range(case['nw']):
        m.mthlim[k] = 4
    grid = _dev_grid(mx, my, dx, dy)
    rp, bc = _rp_struct(case), _bc_struct((3, 3), (1, 1))
    ctx, ctx18 = _lib.Context(0), _lib.Context(0)
    _lib.check(_lib.lib().claw2d_classic_set_launch_ex(ctx.handle, grid, m, rp, P, 0, 2, 1), 'set_launch_ex')
    _lib.check(_lib.lib().claw2d_classic_set_launch_ex(ctx18.handle, grid, m, rp, P, 18, 2, 1), 'set_launch_ex')
    pad = np.zeros((q.shape[0], mx + 2 * NG, my + 2 * NG), order='F')
    dtd = torch.tensor(dt, dtype=torch.float64, device='cuda')
    sh = torch.cuda.current_stream().cuda_stream
    for k in range(6):
        qbc = padded_with_bc(q, 2, (3, 3), (1, 1), True)
        ref, cref = oracle_classic_step(qbc, None, case, dt, 2, 2, dx=dx, dy=dy)
        pad[:] = 0.0
        pad[:, NG - 2:NG - 2 + mx + 4, NG - 2:NG - 2 + my + 4] = qbc
        qd = to_device_padded(pad)
        outs = []
        for c in (ctx, ctx18):
            qn = torch.zeros_like(qd)
            cfl = torch.zeros((), dtype=torch.float64, device='cuda')
            _lib.check(_lib.lib().claw2d_classic_step(c.handle, grid, m, rp, bc, P, qd.data_ptr(), qn.data_ptr(), None,
                                                      dtd.data_ptr(), cfl.data_ptr(), sh), 'claw2d_classic_step')
            torch.cuda.synchronize()
            outs.append((from_device_interior(qn, mx, my), float(cfl.item())))
        (got, cgot), (got18, cgot18) = outs
        print("fill, consecutive steps:", name, prec, "step", k, flush=True)
        if prec == 'f64':
            assert np.array_equal(got, ref) and cgot == cref, "step %d max %g" % (k, np.abs(got - ref).max())
        else:
            scale = np.abs(ref).reshape(ref.shape[0], -1).max(axis=1)[:, None, None] + 1e-300
            assert (np.abs(got - ref) / scale).max() <= 1e-12
        assert np.array_equal(got, got18) and cgot == cgot18, "step %d" % k
        q = ref

@pytest.mark.parametrize("name,trans", [("euler", 2), ("euler", 1), ("acoustics", 2), ("shallow", 1),
                                        ("vc_acoustics", 1)])
def test_paired_fill_decomposition_equals_oracle(name, trans):
    """Launch shape (0, 2, 1), the paired fill decomposition: one-wavefront
    blocks of the two-waves build, one per wavefront slot, the older wavefront
    of each SIMD taking two units of rows for the younger one's one, every
    piece within one strip, its boundaries on even rows.  On 1000x512 (17
    strips of 180-181 units; pieces of 2-8 rows) and 2000x301 (34 strips, an
    odd row count) it
    tiles the grid; on 3000x64 it cannot (a strip would need units under two
    rows) and the step runs the default bands.  Each step equals the oracle
    bit for bit; the fused build through the fill agrees with the oracle to
    the fused build's rounding, and for Euler and acoustics equals its own
    step with 18-row bands bit for bit (the fused build's light-solver march
    solves a band's first y-interface in a separately compiled copy of the
    row step, whose contractions can round differently, so for shallow water
    and vc acoustics the fused answers move in the last bits with the band
    starts) (walls low, extrapolation high, MC, transverse correction or
    increment waves)."""
    from pyclaw_amd import _lib
    case = RP_CASES[name]
    for mx, my in ((1000, 512), (2000, 301), (3000, 64)):
        print("paired fill:", name, "transverse", trans, "grid", mx, my, flush=True)
        rng = np.random.default_rng(mx + my)
        q = case['make'](rng, mx, my)
        aux = None
        if case['naux']:
            from helpers import vc_aux
            aux = vc_aux(rng, mx, my)
        dx, dy = 1.0 / mx, 1.0 / my
        dt = 0.3 * min(dx, dy) if name != 'euler' else 0.1 * min(dx, dy)
        ctx = _lib.Context(0)
        m = _lib.ClawMethod()
        m.order, m.transverse, m.split, m.num_waves = 2, trans, 0, case['nw']
        for k in range(case['nw']):
            m.mthlim[k] = 4
        grid = _dev_grid(mx, my, dx, dy)
        _lib.check(_lib.lib().claw2d_classic_set_launch_ex(ctx.handle, grid, m, _rp_struct(case), _lib.F64, 0, 2, 1),
                   'claw2d_classic_set_launch_ex')
        band, w, wpb = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.lib().claw2d_classic_get_launch(ctx.handle, grid, m, _rp_struct(case), _lib.F64,
                                                       ctypes.byref(band), ctypes.byref(w), ctypes.byref(wpb)),
                   'claw2d_classic_get_launch')
        assert (band.value, w.value, wpb.value) == (0, 2, 1)
        qbc = padded_with_bc(q, 2, (3, 3), (1, 1), True)
        auxbc = padded_with_bc(aux, 2, (3, 3), (1, 1), False) if aux is not None else None
        ref, cref = oracle_classic_step(qbc, auxbc, case, dt, 2, trans, dx=dx, dy=dy)
        kw = dict(bc_lower=(3, 3), bc_upper=(1, 1), abc_lower=(3, 3), abc_upper=(1, 1), dx=dx, dy=dy)
        got, cgot = gpu_classic_step(qbc, auxbc, case, dt, 2, trans, shape=(0, 2, 1), **kw)
        assert np.array_equal(got, ref) and cgot == cref, "max %g" % np.abs(got - ref).max()
        fma, cfma = gpu_classic_step(qbc, auxbc, case, dt, 2, trans, shape=(0, 2, 1), precision='f64fma', **kw)
        scale = np.abs(ref).reshape(ref.shape[0], -1).max(axis=1)[:, None, None] + 1e-300
        assert (np.abs(fma - ref) / scale).max() <= 1e-12 and abs(cfma - cref) <= 1e-12 * cref
        if name in ('euler', 'acoustics'):
            fma18, cfma18 = gpu_classic_step(qbc, auxbc, case, dt, 2, trans, shape=(18, 2, 1), precision='f64fma',
                                             **kw)
            assert np.array_equal(fma, fma18) and cfma == cfma18
