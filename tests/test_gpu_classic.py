"""GPU parity: the fused classic kernel against the CPU oracle, one step from
identical seeded inputs.  FP64 is required to be bit-exact (the kernel keeps
the reference's operation order and is built without contraction); the CFL
number must be identical.  FP32 is checked against FP64 with a stated bound."""
import numpy as np
import pytest

import oracle as O
from helpers import RP_CASES, gpu_classic_step, oracle_classic_step, padded_with_bc, vc_aux

pytestmark = pytest.mark.gpu

BC_SETS = [
    ((1, 1), (1, 1)),      # extrap everywhere
    ((3, 3), (1, 1)),      # walls low, extrap high (config 1 / vc-acoustics app)
    ((2, 2), (2, 2)),      # periodic
    ((0, 3), (1, 0)),      # stored ("custom") x-low and y-high
    ((3, 1), (3, 3)),
]


def _run(name, mx, my, seed, order=2, trans=1, split=False, mthlim=None, bcs=BC_SETS[0]):
    case = RP_CASES[name]
    rng = np.random.default_rng(seed)
    q = case['make'](rng, mx, my)
    aux = vc_aux(rng, mx, my) if case['naux'] else None
    lo, hi = bcs
    qbc = padded_with_bc(q, 2, lo, hi, True, ghost_rng=rng)
    abc = padded_with_bc(aux, 2, (1, 3), (3, 1), False) if aux is not None else None
    ref, cref = oracle_classic_step(qbc, abc, case, case['dt'], order, trans, split, mthlim)
    got, cgot = gpu_classic_step(qbc, abc, case, case['dt'], order, trans, split, mthlim, lo, hi,
                                 (1, 3), (3, 1))
    return ref, cref, got, cgot


@pytest.mark.parametrize("name", list(RP_CASES))
@pytest.mark.parametrize("trans", [0, 1, 2])
def test_unsplit_bitexact(name, trans):
    print("case", name, "trans", trans, flush=True)
    ref, cref, got, cgot = _run(name, 70, 45, 1, trans=trans)
    assert np.array_equal(ref, got), np.abs(ref - got).max()
    assert cref == cgot


@pytest.mark.parametrize("bcs", BC_SETS)
@pytest.mark.parametrize("name", ['euler', 'vc_acoustics'])
def test_unsplit_bcs(name, bcs):
    print("case", name, bcs, flush=True)
    ref, cref, got, cgot = _run(name, 67, 38, 2, trans=2, bcs=bcs)
    assert np.array_equal(ref, got), np.abs(ref - got).max()
    assert cref == cgot


@pytest.mark.parametrize("lim", [0, 1, 2, 3, 4, 5])
def test_unsplit_limiters(lim):
    ref, cref, got, cgot = _run('euler', 130, 33, 3, trans=2, mthlim=[lim] * 5)
    assert np.array_equal(ref, got)
    assert cref == cgot


def test_unsplit_order1():
    ref, cref, got, cgot = _run('euler', 64, 64, 4, order=1, trans=1)
    assert np.array_equal(ref, got)
    assert cref == cgot


@pytest.mark.parametrize("name", list(RP_CASES))
def test_split_bitexact(name):
    ref, cref, got, cgot = _run(name, 75, 41, 5, split=True, mthlim=None, bcs=BC_SETS[1])
    assert np.array_equal(ref, got), np.abs(ref - got).max()
    assert cref == cgot


def test_split_shockbubble_limiters():
    ref, cref, got, cgot = _run('euler', 160, 40, 6, split=True, mthlim=[4, 4, 4, 4, 2], bcs=BC_SETS[3])
    assert np.array_equal(ref, got)
    assert cref == cgot


@pytest.mark.parametrize("shape", [(2, 2), (3, 200), (200, 3), (61, 61), (62, 1), (244, 123)])
def test_unsplit_ragged_shapes(shape):
    mx, my = shape
    if my < 2:
        # grids need at least 2 cells per side: the ABI says so before any launch
        from pyclaw_amd import _lib
        import ctypes
        ctx = _lib.Context(0)
        g = _lib.ClawGrid(mx, my, 3, 0, mx + 6, (mx + 6) * (my + 6), 0.1, 0.1)
        m = _lib.ClawMethod(); m.order = 2; m.transverse = 2; m.num_waves = 5
        for k in range(5):
            m.mthlim[k] = 4
        r = _lib.ClawRiemann(); r.kind = _lib.RP_EULER
        bc = _lib.ClawBC()
        for s in range(2):
            bc.lower[s], bc.upper[s] = BC_SETS[1][0][s], BC_SETS[1][1][s]
        rc = _lib.lib().claw2d_classic_step(ctx.handle, g, m, r, bc, _lib.F64, ctypes.c_void_p(16),
                                            ctypes.c_void_p(32), None, ctypes.c_void_p(48), ctypes.c_void_p(64),
                                            None)
        assert rc == _lib.CLAW_EINVAL
        return
    ref, cref, got, cgot = _run('euler', mx, my, 7, trans=2, bcs=BC_SETS[1])
    assert np.array_equal(ref, got)
    assert cref == cgot


@pytest.mark.parametrize("precision", ['f64', 'f64fma', 'f32'])
@pytest.mark.parametrize("name", list(RP_CASES))
def test_wide_offset_kernels_give_identical_results(name, precision):
    """The unsplit kernels with 64-bit per-lane offsets (used for component
    planes of 4 GiB and more) against the 32-bit-offset kernels every smaller
    grid runs, bit for bit, and (FP64) against the oracle."""
    print("case", name, precision, flush=True)
    case = RP_CASES[name]
    rng = np.random.default_rng(12)
    q = case['make'](rng, 131, 47)
    aux = vc_aux(rng, 131, 47) if case['naux'] else None
    lo, hi = BC_SETS[4]
    qbc = padded_with_bc(q, 2, lo, hi, True, ghost_rng=rng)
    abc = padded_with_bc(aux, 2, (1, 3), (3, 1), False) if aux is not None else None
    args = (qbc, abc, case, case['dt'], 2, 2, False, None, lo, hi, (1, 3), (3, 1))
    narrow, c1 = gpu_classic_step(*args, precision=precision)
    wide, c2 = gpu_classic_step(*args, precision=precision, wide_offsets=True)
    assert np.array_equal(narrow, wide) and c1 == c2
    if precision == 'f64':
        ref, cref = oracle_classic_step(qbc, abc, case, case['dt'], 2, 2)
        assert np.array_equal(ref, wide) and cref == c2


@pytest.mark.parametrize("trans", [1, 2])
def test_euler_without_entropy_fix(trans):
    """efix = 0 (rpn2_euler_5wave.f's plain Roe splitting) against the oracle
    with the same switch; with transverse 2 and MC this is the method-
    specialised kernel's fallback (its top level assumes the fix on)."""
    import copy
    case = copy.deepcopy(RP_CASES['euler'])
    case['pd']['efix'] = 0
    rng = np.random.default_rng(21)
    q = case['make'](rng, 83, 41)
    qbc = padded_with_bc(q, 2, (1, 1), (1, 1), True)
    ref, cref = oracle_classic_step(qbc, None, case, case['dt'], 2, trans)
    got, cgot = gpu_classic_step(qbc, None, case, case['dt'], 2, trans, efix=0)
    assert np.array_equal(ref, got) and cref == cgot
    with_fix, _ = gpu_classic_step(qbc, None, RP_CASES['euler'], case['dt'], 2, trans)
    assert not np.array_equal(with_fix, got)   # the switch reaches the kernel


def test_f32_error_vs_f64():
    """FP32 variant: relative L-inf error of one Euler step against the FP64 oracle."""
    case = RP_CASES['euler']
    rng = np.random.default_rng(8)
    q = case['make'](rng, 96, 64)
    qbc = padded_with_bc(q, 2, (1, 1), (1, 1), True)
    ref, cref = oracle_classic_step(qbc, None, case, case['dt'], 2, 2)
    got, cgot = gpu_classic_step(qbc, None, case, case['dt'], 2, 2, precision='f32')
    rel = np.abs(got - ref).max() / np.abs(ref).max()
    print("f32 rel Linf", rel)
    assert rel < 1e-5
    assert abs(cgot - cref) / cref < 1e-5


def test_autotuned_band_heights_give_identical_results():
    """The launch shape (row-band height, one or two wavefronts per SIMD) only
    changes scheduling: after autotuning and for shapes set explicitly the step
    equals the default-shape step and the oracle bit for bit, and the Courant
    number is the same (the last shape, the paired fill decomposition, cannot
    tile a grid this small and runs the default bands)."""
    import ctypes
    import torch
    from pyclaw_amd import _lib
    from pyclaw_amd.problems import shockbubble, euler_max_speed
    mx, my = 130, 70
    q0, dx, dy = shockbubble(mx, my)
    dt = 0.8 * min(dx, dy) / euler_max_speed(q0)
    dev = torch.device('cuda', 0)
    ng = 3
    buf = torch.zeros((5, my + 2 * ng, mx + 2 * ng), dtype=torch.float64)
    buf[:, ng:ng + my, ng:ng + mx] = torch.from_numpy(np.ascontiguousarray(q0.transpose(0, 2, 1)))
    qa = buf.to(dev)
    grid = _lib.ClawGrid(mx, my, ng, 0, mx + 2 * ng, (mx + 2 * ng) * (my + 2 * ng), dx, dy)
    m = _lib.ClawMethod()
    m.order, m.transverse, m.split, m.fwave, m.num_waves = 2, 2, 0, 0, 5
    for k in range(5):
        m.mthlim[k] = 4
    r = _lib.ClawRiemann()
    r.kind, r.efix = _lib.RP_EULER, 1
    r.p[0], r.p[1] = 1.4, 0.4
    bc = _lib.ClawBC()
    bc.lower[0], bc.lower[1], bc.upper[0], bc.upper[1] = 1, 3, 1, 1
    L = _lib.lib()
    dtd = torch.tensor(dt, dtype=torch.float64, device=dev)
    sh = torch.cuda.current_stream().cuda_stream
    outs, cfls = [], []
    for tuned in (False, True, (5, 1), (5, 2), (70, 2), (36, 2, 1), (7, 1, 2), (3, 2, 1), (1, 1, 1), (2, 2, 4),
                  (0, 2, 1)):
        ctx = _lib.Context(0)
        cfl = torch.zeros((), dtype=torch.float64, device=dev)
        qb = torch.zeros_like(qa)
        if isinstance(tuned, tuple) and len(tuned) == 3:
            _lib.check(L.claw2d_classic_set_launch_ex(ctx.handle, grid, m, r, _lib.F64, *tuned), "set_launch_ex")
            sb, sw, sp = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
            _lib.check(L.claw2d_classic_get_launch(ctx.handle, grid, m, r, _lib.F64, ctypes.byref(sb),
                                                   ctypes.byref(sw), ctypes.byref(sp)), "get_launch")
            assert (sb.value, sw.value, sp.value) == tuned
        elif isinstance(tuned, tuple):
            _lib.check(L.claw2d_classic_set_launch(ctx.handle, grid, m, r, _lib.F64, tuned[0], tuned[1]), "set_launch")
        elif tuned:
            band = ctypes.c_int(0)
            _lib.check(L.claw2d_classic_autotune(ctx.handle, grid, m, r, bc, _lib.F64, qa.data_ptr(), qb.data_ptr(),
                                                 None, dtd.data_ptr(), cfl.data_ptr(), sh, ctypes.byref(band)),
                       "autotune")
            assert 1 <= abs(band.value) <= my
            qb.zero_()
        _lib.check(L.claw2d_classic_step(ctx.handle, grid, m, r, bc, _lib.F64, qa.data_ptr(), qb.data_ptr(), None,
                                         dtd.data_ptr(), cfl.data_ptr(), sh), "step")
        torch.cuda.synchronize()
        outs.append(qb[:, ng:ng + my, ng:ng + mx].cpu().numpy())
        cfls.append(float(cfl))
    for k in range(1, len(outs)):
        assert np.array_equal(outs[0], outs[k]) and cfls[0] == cfls[k]
    O.set_params(O.RP_EULER, dict(gamma=1.4, gamma1=0.4))
    qbc = np.zeros((5, mx + 4, my + 4), order='F')
    qbc[:, 2:-2, 2:-2] = q0
    O.fill_bc(qbc, 2, [1, 3], [1, 1], True)
    ref, rcfl = O.step2(O.RP_EULER, qbc, None, dx, dy, dt, O.method_array(2, 2, False), O.mthlim_array(4, 5), 5)
    assert np.array_equal(outs[1], np.ascontiguousarray(ref[:, 2:-2, 2:-2].transpose(0, 2, 1)))
    assert cfls[1] == rcfl
