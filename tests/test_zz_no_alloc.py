"""claw2d_classic_step_noalloc (the step entry point that allocates no device
memory: the split step's intermediate state lives in the library's fixed
device array) against claw2d_classic_step, through the C ABI: bit for bit in
every precision, on two calls in a row, the second on a smaller grid, for
split steps (which use the array) and unsplit steps (which use none); sizes
beyond the array are rejected before anything is launched."""
import ctypes

import numpy as np
import pytest

from helpers import NG, RP_CASES, _bc_struct, _dev_grid, _rp_struct, vc_aux

pytestmark = pytest.mark.gpu


def _method(case, split, trans=2):
    from pyclaw_amd import _lib
    m = _lib.ClawMethod()
    m.order, m.transverse, m.split, m.fwave, m.num_waves = 2, trans, int(split), 0, case['nw']
    for k in range(case['nw']):
        m.mthlim[k] = 4
    return m


def _step(entry, ctx, case, q0, aux, split, precision, bc=((3, 1), (1, 3))):
    """One step of `entry` on device copies of q0 (padded device layout);
    returns (qnew on the host, cfl, rc)."""
    import torch
    from pyclaw_amd import _lib
    ne, rows, ldq = q0.shape
    mx, my = ldq - 2 * NG, rows - 2 * NG
    dtype = torch.float32 if precision == _lib.F32 else torch.float64
    qa = torch.from_numpy(q0).to('cuda', dtype)
    qb = torch.full_like(qa, 7.0)
    ad = torch.from_numpy(aux).to('cuda', dtype) if aux is not None else None
    dt = torch.tensor(case['dt'], dtype=torch.float64, device='cuda')
    cfl = torch.full((), -1.0, dtype=torch.float64, device='cuda')
    rc = entry(ctx.handle, _dev_grid(mx, my, 0.1, 0.07), _method(case, split), _rp_struct(case), _bc_struct(*bc),
               precision, qa.data_ptr(), qb.data_ptr(), ad.data_ptr() if ad is not None else None, dt.data_ptr(),
               cfl.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return qb.cpu().numpy(), float(cfl.item()), rc


def _padded_state(case, rng, mx, my):
    q = case['make'](rng, mx, my)
    pad = np.zeros((q.shape[0], my + 2 * NG, mx + 2 * NG))
    pad[:, NG:-NG, NG:-NG] = q.transpose(0, 2, 1)
    aux = None
    if case['naux']:
        a = vc_aux(rng, mx, my)
        aux = np.ones((2, my + 2 * NG, mx + 2 * NG))
        aux[:, NG:-NG, NG:-NG] = a.transpose(0, 2, 1)
    return pad, aux


@pytest.mark.parametrize("name", list(RP_CASES))
@pytest.mark.parametrize("precision", ['f64', 'f64fma', 'f32'])
def test_noalloc_step_equals_classic_step(name, precision):
    from pyclaw_amd import _lib
    prec = {'f64': _lib.F64, 'f64fma': _lib.F64_FMA, 'f32': _lib.F32}[precision]
    print("noalloc step:", name, precision, flush=True)
    L = _lib.lib()
    case = RP_CASES[name]
    rng = np.random.default_rng(90)
    ctx_a, ctx_b = _lib.Context(0), _lib.Context(0)
    # two calls in a row, the second on a smaller grid; split and unsplit
    for (mx, my) in [(301, 203), (77, 45)]:
        q0, aux = _padded_state(case, rng, mx, my)
        for split in (True, False):
            ref, cref, rc1 = _step(L.claw2d_classic_step, ctx_a, case, q0, aux, split, prec)
            got, cgot, rc2 = _step(L.claw2d_classic_step_noalloc, ctx_b, case, q0, aux, split, prec)
            assert rc1 == _lib.CLAW_OK and rc2 == _lib.CLAW_OK
            assert np.array_equal(ref, got), (mx, my, split, np.nanmax(np.abs(ref - got)))
            assert cref == cgot and cgot >= 0.0


def test_noalloc_rejects_split_steps_beyond_its_array():
    """A split Euler step whose state exceeds the fixed array returns
    CLAW_ETOOLARGE, and leaves the Courant slot untouched (nothing launched);
    the same grid unsplit (no scratch) runs and equals claw2d_classic_step."""
    from pyclaw_amd import _lib
    L = _lib.lib()
    case = RP_CASES['euler']
    mx, my = 1200, 800
    assert 5 * (mx + 2 * NG) * (my + 2 * NG) * 8 > _lib.NOALLOC_SCRATCH_BYTES
    rng = np.random.default_rng(91)
    q0, _ = _padded_state(case, rng, mx, my)
    ctx = _lib.Context(0)
    _, cfl, rc = _step(L.claw2d_classic_step_noalloc, ctx, case, q0, None, True, _lib.F64)
    assert rc == _lib.CLAW_ETOOLARGE and cfl == -1.0
    got, cgot, rc = _step(L.claw2d_classic_step_noalloc, ctx, case, q0, None, False, _lib.F64)
    ref, cref, rc0 = _step(L.claw2d_classic_step, _lib.Context(0), case, q0, None, False, _lib.F64)
    assert rc == rc0 == _lib.CLAW_OK and np.array_equal(got, ref) and cgot == cref
