"""GPU parity of the fused SharpClaw kernels (WENO5 + Riemann + RK stage)
against the CPU oracle: one dq evaluation and one whole SSP104 step from the
same seeded data.  FP64 must be bit-exact (same operation order, no
contraction), and the Courant numbers identical."""
import numpy as np
import pytest

import oracle as O
from oracle import driver
from helpers import RP_CASES, gpu_sharp_dq, gpu_sharp_ssp104, padded_with_bc, vc_aux

pytestmark = pytest.mark.gpu

BCS = [((1, 1), (1, 1)), ((3, 3), (1, 1)), ((2, 2), (2, 2)), ((0, 3), (1, 0))]


@pytest.mark.parametrize("name", list(RP_CASES))
@pytest.mark.parametrize("bcs", BCS)
def test_sharp_dq_bitexact(name, bcs):
    print("case", name, bcs, flush=True)
    case = RP_CASES[name]
    rng = np.random.default_rng(21)
    mx, my = 67, 41
    q = case['make'](rng, mx, my)
    aux = vc_aux(rng, mx, my) if case['naux'] else None
    lo, hi = bcs
    qbc = padded_with_bc(q, 3, lo, hi, True, ghost_rng=rng)
    abc = padded_with_bc(aux, 3, (3, 1), (1, 3), False) if aux is not None else None
    O.set_params(case['rp'], case['pd'])
    ref, cref = O.sharp_dq(case['rp'], qbc, abc, 0.1, 0.07, case['dt'], case['nw'])
    got, cgot = gpu_sharp_dq(qbc, abc, case, case['dt'], lo, hi, (3, 1), (1, 3))
    assert np.array_equal(ref, got), np.abs(ref - got).max()
    assert cref == cgot



@pytest.mark.parametrize("name", ['shallow', 'euler', 'vc_acoustics', 'acoustics'])
def test_sharp_paired_fill_equals_oracle(name):
    """SharpClaw dq through the paired fill decomposition: launch shape
    (0, 1, 1), one-wavefront blocks, the older wavefront of each SIMD taking
    two units of rows for the younger one's one, and (0, 3, 1), equal pieces
    in the three-per-SIMD build; each piece one band within a strip.  On
    1000x401 (18 strips, pieces of 2-5 rows) it equals the oracle bit for bit,
    Courant number included (walls and extrapolation)."""
    case = RP_CASES[name]
    rng = np.random.default_rng(31)
    mx, my = 1000, 401
    q = case['make'](rng, mx, my)
    aux = vc_aux(rng, mx, my) if case['naux'] else None
    qbc = padded_with_bc(q, 3, (3, 1), (1, 3), True)
    abc = padded_with_bc(aux, 3, (3, 1), (1, 3), False) if aux is not None else None
    O.set_params(case['rp'], case['pd'])
    ref, cref = O.sharp_dq(case['rp'], qbc, abc, 0.1, 0.07, case['dt'], case['nw'])
    for shape in ((0, 1, 1), (0, 3, 1)):
        print("sharp paired fill:", name, shape, flush=True)
        got, cgot = gpu_sharp_dq(qbc, abc, case, case['dt'], (3, 1), (1, 3), (3, 1), (1, 3), shape=shape)
        assert np.array_equal(ref, got), (shape, np.abs(ref - got).max())
        assert cref == cgot

@pytest.mark.parametrize("name", ['shallow', 'euler', 'vc_acoustics'])
def test_sharp_ssp104_step_bitexact(name):
    case = RP_CASES[name]
    rng = np.random.default_rng(22)
    mx, my = 75, 52
    q = case['make'](rng, mx, my)
    aux = vc_aux(rng, mx, my) if case['naux'] else None
    P = driver.Problem(case['rp'], case['nw'], case['pd'], 0.1, 0.07, (2, 3), (2, 1), aux=aux,
                       aux_bc_lower=(2, 3), aux_bc_upper=(2, 1), kind='sharpclaw', cfl_max=1e9)
    ref, cref = driver.sharp_step(P, q, case['dt'], 0.0)
    got, cgot = gpu_sharp_ssp104(q, aux, case, case['dt'], (2, 3), (2, 1), (2, 3), (2, 1))
    assert np.array_equal(ref, got), np.abs(ref - got).max()
    assert cgot[-1] == cref


@pytest.mark.parametrize("shape", [(3, 3), (5, 130), (130, 5), (58, 58), (117, 64)])
def test_sharp_dq_ragged(shape):
    case = RP_CASES['euler']
    rng = np.random.default_rng(23)
    mx, my = shape
    q = case['make'](rng, mx, my)
    qbc = padded_with_bc(q, 3, (3, 1), (1, 3), True)
    O.set_params(case['rp'], case['pd'])
    ref, cref = O.sharp_dq(case['rp'], qbc, None, 0.1, 0.07, case['dt'], case['nw'])
    got, cgot = gpu_sharp_dq(qbc, None, case, case['dt'], (3, 1), (1, 3))
    assert np.array_equal(ref, got)
    assert cref == cgot


def test_sharp_autotuned_band_gives_identical_results():
    """SharpClaw dq after claw2d_sharpclaw_autotune and for shapes set
    explicitly (both builds) equals the default-shape dq bit for bit, Courant
    number included."""
    import ctypes
    import torch
    from pyclaw_amd import _lib
    from pyclaw_amd.problems import shallow_hump
    mx, my = 150, 90
    q0 = shallow_hump(mx, my, (0.4, 0.6))
    dev = torch.device('cuda', 0)
    ng = 3
    buf = torch.zeros((3, my + 2 * ng, mx + 2 * ng), dtype=torch.float64)
    buf[:, ng:ng + my, ng:ng + mx] = torch.from_numpy(np.ascontiguousarray(q0.transpose(0, 2, 1)))
    q = buf.to(dev)
    g = _lib.ClawGrid(mx, my, ng, 0, mx + 2 * ng, (mx + 2 * ng) * (my + 2 * ng), 1.0 / mx, 1.0 / my)
    r = _lib.ClawRiemann()
    r.kind, r.efix = _lib.RP_SHALLOW, 1
    r.p[0] = 1.0
    bc = _lib.ClawBC()
    bc.lower[0], bc.lower[1], bc.upper[0], bc.upper[1] = 2, 3, 2, 1
    L = _lib.lib()
    dtd = torch.tensor(0.3 / mx, dtype=torch.float64, device=dev)
    sh = torch.cuda.current_stream().cuda_stream
    outs, cfls = [], []
    for tuned in (False, True, (5, 1), (5, 3), (90, 3), (7, 3, 1), (6, 1, 2), (0, 1, 1), (0, 3, 1), (0, 1, 4)):
        ctx = _lib.Context(0)
        cfl = torch.zeros((), dtype=torch.float64, device=dev)
        dq = torch.zeros_like(q)
        if isinstance(tuned, tuple) and len(tuned) == 3:
            _lib.check(L.claw2d_sharpclaw_set_launch_ex(ctx.handle, g, r, _lib.F64, *tuned), "set_launch_ex")
            sb, sw, sp = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
            _lib.check(L.claw2d_sharpclaw_get_launch(ctx.handle, g, r, _lib.F64, ctypes.byref(sb), ctypes.byref(sw),
                                                     ctypes.byref(sp)), "get_launch")
            assert (sb.value, sw.value, sp.value) == tuned
        elif isinstance(tuned, tuple):
            _lib.check(L.claw2d_sharpclaw_set_launch(ctx.handle, g, r, _lib.F64, tuned[0], tuned[1]), "set_launch")
        elif tuned:
            band = ctypes.c_int(0)
            _lib.check(L.claw2d_sharpclaw_autotune(ctx.handle, g, r, bc, _lib.F64, q.data_ptr(), dq.data_ptr(), None,
                                                   dtd.data_ptr(), cfl.data_ptr(), sh, ctypes.byref(band)), "tune")
            assert 0 <= abs(band.value) <= my   # 0: the fill decomposition
            dq.zero_()
        _lib.check(L.claw2d_sharpclaw_dq(ctx.handle, g, r, bc, _lib.F64, q.data_ptr(), dq.data_ptr(), None,
                                         dtd.data_ptr(), cfl.data_ptr(), sh), "dq")
        torch.cuda.synchronize()
        outs.append(dq.cpu().numpy())
        cfls.append(float(cfl))
    for k in range(1, len(outs)):
        assert np.array_equal(outs[0], outs[k]) and cfls[0] == cfls[k]
