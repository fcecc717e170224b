"""The FP64 classic and SharpClaw kernels built with fused multiply-add and
shared reciprocals (precision CLAW_F64_FMA): one step (or one dq evaluation)
from identical seeded inputs must agree with the CPU oracle (the reference's
arithmetic) to 1e-13 relative L-inf per field, the Courant number to 1e-13
relative, and a periodic step must conserve the conserved fields to 1e-13
relative."""
import numpy as np
import pytest

import oracle as O
from oracle import driver
from helpers import (RP_CASES, gpu_classic_step, gpu_sharp_dq, gpu_sharp_ssp104, oracle_classic_step,
                     padded_with_bc, vc_aux)

pytestmark = pytest.mark.gpu

TOL = 1e-13


def _rel_fields(got, ref):
    return [np.abs(got[m] - ref[m]).max() / np.abs(ref[m]).max() for m in range(ref.shape[0])]


def _case(name, mx, my, seed, trans, split, bcs):
    case = RP_CASES[name]
    rng = np.random.default_rng(seed)
    q = case['make'](rng, mx, my)
    aux = vc_aux(rng, mx, my) if case['naux'] else None
    lo, hi = bcs
    qbc = padded_with_bc(q, 2, lo, hi, True, ghost_rng=rng)
    abc = padded_with_bc(aux, 2, (1, 3), (3, 1), False) if aux is not None else None
    ref, cref = oracle_classic_step(qbc, abc, case, case['dt'], 2, trans, split)
    got, cgot = gpu_classic_step(qbc, abc, case, case['dt'], 2, trans, split, None, lo, hi, (1, 3), (3, 1),
                                 precision='f64fma')
    exact, cex = gpu_classic_step(qbc, abc, case, case['dt'], 2, trans, split, None, lo, hi, (1, 3), (3, 1))
    return q, ref, cref, got, cgot, exact, cex


@pytest.mark.parametrize("name", list(RP_CASES))
@pytest.mark.parametrize("trans,split", [(0, False), (1, False), (2, False), (0, True)])
def test_fma_step_within_tolerance(name, trans, split):
    print("case", name, "trans", trans, "split", split, flush=True)
    q, ref, cref, got, cgot, exact, cex = _case(name, 70, 45, 11, trans, split, ((3, 3), (1, 1)))
    assert np.array_equal(exact, ref)
    rel = _rel_fields(got, ref)
    print("fma rel Linf per field", rel, flush=True)
    assert max(rel) <= TOL, rel
    assert abs(cgot - cref) <= TOL * cref


@pytest.mark.parametrize("bcs", [((1, 1), (1, 1)), ((0, 3), (1, 0)), ((3, 1), (3, 3))])
def test_fma_euler_bcs(bcs):
    print("case euler", bcs, flush=True)
    q, ref, cref, got, cgot, exact, cex = _case('euler', 131, 67, 12, 2, False, bcs)
    assert max(_rel_fields(got, ref)) <= TOL


@pytest.mark.parametrize("name", ['euler', 'shallow', 'acoustics'])
def test_fma_periodic_conservation(name):
    """Periodic in both directions: the totals of the conserved fields change by
    rounding only (the Euler tracer is advected non-conservatively)."""
    print("case", name, "periodic", flush=True)
    q, ref, cref, got, cgot, exact, cex = _case(name, 96, 80, 13, 2, False, ((2, 2), (2, 2)))
    ncons = 4 if name == 'euler' else 3
    for m in range(ncons):
        t0 = np.sum(q[m], dtype=np.float64)
        scale = np.abs(q[m]).sum()
        assert abs(np.sum(got[m], dtype=np.float64) - t0) <= TOL * scale, m


def test_fma_shockbubble_steps_track_exact():
    """Shock-bubble 4096x1024-like set-up at 512x128: each of 20 steps, taken
    from the exact kernel's state, stays within tolerance of the exact step."""
    import torch
    from pyclaw_amd import _lib
    from pyclaw_amd.problems import shockbubble, euler_max_speed
    mx, my = 512, 128
    q0, dx, dy = shockbubble(mx, my)
    dt = 0.8 * min(dx, dy) / euler_max_speed(q0)
    dev = torch.device('cuda', 0)
    ng = 3
    buf = torch.zeros((5, my + 2 * ng, mx + 2 * ng), dtype=torch.float64)
    buf[:, ng:ng + my, ng:ng + mx] = torch.from_numpy(np.ascontiguousarray(q0.transpose(0, 2, 1)))
    qa = buf.to(dev)
    qe, qf = torch.zeros_like(qa), torch.zeros_like(qa)
    grid = _lib.ClawGrid(mx, my, ng, 0, mx + 2 * ng, (mx + 2 * ng) * (my + 2 * ng), dx, dy)
    m = _lib.ClawMethod()
    m.order, m.transverse, m.split, m.fwave, m.num_waves = 2, 1, 0, 0, 5
    for k in range(5):
        m.mthlim[k] = 4
    r = _lib.ClawRiemann()
    r.kind, r.efix = _lib.RP_EULER, 1
    r.p[0], r.p[1] = 1.4, 0.4
    bc = _lib.ClawBC()
    bc.lower[0], bc.lower[1], bc.upper[0], bc.upper[1] = 1, 3, 1, 1
    L = _lib.lib()
    ctx = _lib.Context(0)
    dtd = torch.tensor(dt, dtype=torch.float64, device=dev)
    cfl = torch.zeros((), dtype=torch.float64, device=dev)
    sh = torch.cuda.current_stream().cuda_stream
    worst = 0.0
    for n in range(20):
        _lib.check(L.claw2d_classic_step(ctx.handle, grid, m, r, bc, _lib.F64_FMA, qa.data_ptr(), qf.data_ptr(),
                                         None, dtd.data_ptr(), cfl.data_ptr(), sh), "fma step")
        _lib.check(L.claw2d_classic_step(ctx.handle, grid, m, r, bc, _lib.F64, qa.data_ptr(), qe.data_ptr(),
                                         None, dtd.data_ptr(), cfl.data_ptr(), sh), "exact step")
        d = (qf - qe)[:, ng:ng + my, ng:ng + mx].abs().amax(dim=(1, 2))
        s = qe[:, ng:ng + my, ng:ng + mx].abs().amax(dim=(1, 2))
        # both momenta relative to the momentum scale: rho v is exactly zero
        # until the shock reaches the bubble
        s[1] = s[2] = torch.maximum(s[1], s[2])
        worst = max(worst, float((d / s).max()))
        qa, qe = qe, qa
    print("worst per-step rel Linf over 20 steps", worst)
    assert worst <= TOL


@pytest.mark.parametrize("name", list(RP_CASES))
@pytest.mark.parametrize("bcs", [((1, 1), (1, 1)), ((3, 3), (1, 1)), ((2, 2), (2, 2)), ((0, 3), (1, 0))])
def test_fma_sharp_dq_within_tolerance(name, bcs):
    print("sharp dq", name, bcs, flush=True)
    case = RP_CASES[name]
    rng = np.random.default_rng(31)
    mx, my = 67, 41
    q = case['make'](rng, mx, my)
    aux = vc_aux(rng, mx, my) if case['naux'] else None
    lo, hi = bcs
    qbc = padded_with_bc(q, 3, lo, hi, True, ghost_rng=rng)
    abc = padded_with_bc(aux, 3, (3, 1), (1, 3), False) if aux is not None else None
    O.set_params(case['rp'], case['pd'])
    ref, cref = O.sharp_dq(case['rp'], qbc, abc, 0.1, 0.07, case['dt'], case['nw'])
    got, cgot = gpu_sharp_dq(qbc, abc, case, case['dt'], lo, hi, (3, 1), (1, 3), precision='f64fma')
    rel = _rel_fields(got, ref)
    print("sharp dq fma rel Linf per field", rel, flush=True)
    assert max(rel) <= TOL, rel
    assert abs(cgot - cref) <= TOL * cref


@pytest.mark.parametrize("name", ['shallow', 'euler', 'vc_acoustics', 'acoustics'])
def test_fma_sharp_ssp104_step(name):
    case = RP_CASES[name]
    rng = np.random.default_rng(32)
    mx, my = 75, 52
    q = case['make'](rng, mx, my)
    aux = vc_aux(rng, mx, my) if case['naux'] else None
    P = driver.Problem(case['rp'], case['nw'], case['pd'], 0.1, 0.07, (2, 3), (2, 1), aux=aux,
                       aux_bc_lower=(2, 3), aux_bc_upper=(2, 1), kind='sharpclaw', cfl_max=1e9)
    ref, cref = driver.sharp_step(P, q, case['dt'], 0.0)
    got, cgot = gpu_sharp_ssp104(q, aux, case, case['dt'], (2, 3), (2, 1), (2, 3), (2, 1), precision='f64fma')
    rel = _rel_fields(got, ref)
    print("sharp ssp104 fma rel Linf per field", rel, flush=True)
    assert max(rel) <= TOL, rel
    assert abs(cgot[-1] - cref) <= TOL * cref


def test_fma_sharp_periodic_conservation():
    """Shallow water, periodic: one fused SSP104 step keeps the totals of h, hu
    and hv to rounding."""
    case = RP_CASES['shallow']
    rng = np.random.default_rng(33)
    q = case['make'](rng, 96, 80)
    got, _ = gpu_sharp_ssp104(q, None, case, case['dt'], (2, 2), (2, 2), precision='f64fma')
    for m in range(3):
        scale = np.abs(q[m]).sum()
        assert abs(np.sum(got[m]) - np.sum(q[m])) <= TOL * scale, m


def test_fma_sharp_smooth_hump_steps():
    """Smooth data is where WENO5's smoothness indicators cancel: there the
    reference arithmetic itself moves by ~1e-9 relative in one SSP104 step when
    its input moves by one ulp, so no build that is not bit-identical can stay
    within 1e-13.  A shallow-water Gaussian hump (periodic, 192x160) is advanced
    12 steps with the exact build; at every step the fused build's deviation
    from the exact step must stay within 50x the exact build's own response to
    a random one-ulp perturbation of the input (and below 1e-6)."""
    import torch
    from pyclaw_amd.problems import shallow_hump
    case = RP_CASES['shallow']
    mx, my = 192, 160
    q = shallow_hump(mx, my, (0.43, 0.57))
    dx, dy = 1.0 / mx, 1.0 / my
    dt = 0.5 * min(dx, dy) / float(np.sqrt(q[0]).max() * 1.6)
    rng = np.random.default_rng(34)

    def rel(a, b):
        d = np.abs(a - b).max(axis=(1, 2))
        s = np.abs(b).max(axis=(1, 2))
        s[1] = s[2] = max(s[1], s[2])
        return float((d / s).max())
    worst_dev, worst_ratio = 0.0, 0.0
    for n in range(12):
        exact, _ = gpu_sharp_ssp104(q, None, case, dt, (2, 2), (2, 2), dx=dx, dy=dy)
        fused, _ = gpu_sharp_ssp104(q, None, case, dt, (2, 2), (2, 2), dx=dx, dy=dy, precision='f64fma')
        qp = q * (1 + rng.choice([-1.0, 1.0], size=q.shape) * 2.0 ** -52)
        noisy, _ = gpu_sharp_ssp104(qp, None, case, dt, (2, 2), (2, 2), dx=dx, dy=dy)
        dev, sens = rel(fused, exact), rel(noisy, exact)
        worst_dev = max(worst_dev, dev)
        worst_ratio = max(worst_ratio, dev / max(sens, 1e-15))
        q = exact
    print("smooth hump: fused deviation %.3g, worst deviation / one-ulp sensitivity %.3g" % (worst_dev, worst_ratio),
          flush=True)
    assert worst_dev < 1e-6
    assert worst_ratio <= 50
