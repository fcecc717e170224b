"""Shared test helpers: seeded inputs, and calling the C ABI with torch buffers."""
import ctypes

import numpy as np

import oracle as O

NG = 3  # ghost layers in the device layout used by the tests


def euler_state(rng, nx, ny, amp=0.5):
    rho = 1 + amp * rng.random((nx, ny))
    u = rng.standard_normal((nx, ny)) * amp
    v = rng.standard_normal((nx, ny)) * amp
    p = 1 + rng.random((nx, ny))
    q = np.zeros((5, nx, ny))
    q[0] = rho; q[1] = rho * u; q[2] = rho * v
    q[3] = p / 0.4 + 0.5 * rho * (u * u + v * v)
    q[4] = rng.random((nx, ny))
    return q


def shallow_state(rng, nx, ny):
    h = 1 + 0.5 * rng.random((nx, ny))
    return np.stack([h, 0.3 * rng.standard_normal((nx, ny)), 0.3 * rng.standard_normal((nx, ny))])


def acoustic_state(rng, nx, ny):
    return rng.standard_normal((3, nx, ny))


def vc_aux(rng, nx, ny):
    return np.stack([1 + 3 * rng.random((nx, ny)), 0.5 + 1.5 * rng.random((nx, ny))])


RP_CASES = {
    'acoustics': dict(rp=O.RP_ACOUSTICS, nw=2, naux=0, pd=dict(rho=1.0, bulk=4.0, cc=2.0, zz=2.0),
                      make=acoustic_state, dt=0.02),
    'vc_acoustics': dict(rp=O.RP_VC_ACOUSTICS, nw=2, naux=2, pd={}, make=acoustic_state, dt=0.02),
    'euler': dict(rp=O.RP_EULER, nw=5, naux=0, pd=dict(gamma=1.4, gamma1=0.4), make=euler_state, dt=0.01),
    'shallow': dict(rp=O.RP_SHALLOW, nw=3, naux=0, pd=dict(grav=1.0), make=shallow_state, dt=0.02),
}


def padded_with_bc(interior, ng, bc_lower, bc_upper, negate, ghost_rng=None):
    """Reference padded array (ncomp, mx+2ng, my+2ng) F-order; standard codes are
    filled like the reference; code 0 ('custom') gets arbitrary ghost values."""
    ne, mx, my = interior.shape
    qbc = np.zeros((ne, mx + 2 * ng, my + 2 * ng), order='F')
    qbc[:, ng:-ng, ng:-ng] = interior
    if ghost_rng is not None:
        # arbitrary but physically valid ghost states ("custom" boundary data):
        # cell states drawn from the interior, shuffled
        mask = np.ones(qbc.shape[1:], bool)
        mask[ng:-ng, ng:-ng] = False
        cells = interior.reshape(ne, -1)
        pick = ghost_rng.integers(0, cells.shape[1], mask.sum())
        qbc[:, mask] = cells[:, pick]
    # custom sides: the stored values stay; other sides are filled in reference order
    O.fill_bc(qbc, ng, [c if c != 0 else 99 for c in bc_lower], [c if c != 0 else 99 for c in bc_upper],
              negate)
    return qbc


def gpu_classic_step(qbc_ref, auxbc_ref, case, dt, order=2, trans=1, split=False, mthlim=None,
                     bc_lower=(1, 1), bc_upper=(1, 1), abc_lower=(1, 1), abc_upper=(1, 1),
                     dx=0.1, dy=0.07, precision='f64', wide_offsets=False, efix=1, shape=None, ctx=None):
    """Run claw2d_classic_step on the GPU.  qbc_ref/auxbc_ref are reference-layout
    padded arrays (ghost width 2 for the classic step); they are copied into the
    device layout with NG ghost layers (the outermost layer is padding).
    wide_offsets: the unsplit kernels with 64-bit per-lane offsets
    (claw_ctx_set_wide_offsets).  shape: (rows, waves per SIMD, waves per
    block) set on the context first (claw2d_classic_set_launch_ex; rows 0 is
    the paired fill decomposition).  ctx: a context to run on (a new one by
    default)."""
    import torch
    from pyclaw_amd import _lib
    ne, nxp, nyp = qbc_ref.shape
    g = (nxp - (nxp - 4)) // 2
    mx, my = nxp - 2 * g, nyp - 2 * g
    dtype = torch.float32 if precision == 'f32' else torch.float64
    dev = torch.device('cuda', 0)

    def to_dev(a):
        buf = torch.zeros((a.shape[0], my + 2 * NG, mx + 2 * NG), dtype=dtype)
        off = NG - g
        buf[:, off:off + nyp, off:off + nxp] = torch.from_numpy(np.ascontiguousarray(a.transpose(0, 2, 1)))
        return buf.to(dev)

    qd = to_dev(qbc_ref)
    qn = torch.zeros_like(qd)
    ad = to_dev(auxbc_ref) if auxbc_ref is not None else None
    grid = _lib.ClawGrid(mx, my, NG, 0, mx + 2 * NG, (mx + 2 * NG) * (my + 2 * NG), dx, dy)
    m = _lib.ClawMethod()
    m.order, m.transverse, m.split, m.fwave, m.num_waves = order, trans, int(split), 0, case['nw']
    lim = mthlim if mthlim is not None else [4] * case['nw']
    for k, v in enumerate(lim):
        m.mthlim[k] = v
    r = _lib.ClawRiemann()
    r.kind = case['rp']
    r.efix = efix
    vals = {O.RP_ACOUSTICS: ['rho', 'bulk', 'cc', 'zz'], O.RP_EULER: ['gamma', 'gamma1'],
            O.RP_SHALLOW: ['grav'], O.RP_VC_ACOUSTICS: []}[case['rp']]
    for k, key in enumerate(vals):
        r.p[k] = case['pd'][key]
    bc = _lib.ClawBC()
    for s in range(2):
        bc.lower[s], bc.upper[s] = bc_lower[s], bc_upper[s]
        bc.aux_lower[s], bc.aux_upper[s] = abc_lower[s], abc_upper[s]
    dtd = torch.tensor(dt, dtype=torch.float64, device=dev)
    cfl = torch.zeros((), dtype=torch.float64, device=dev)
    ctx = _lib.Context(0) if ctx is None else ctx
    if wide_offsets:
        _lib.check(_lib.lib().claw_ctx_set_wide_offsets(ctx.handle, 1), 'claw_ctx_set_wide_offsets')
    prec = {'f64': _lib.F64, 'f32': _lib.F32, 'f64fma': _lib.F64_FMA}[precision]
    if shape is not None:
        _lib.check(_lib.lib().claw2d_classic_set_launch_ex(ctx.handle, grid, m, r, prec, *shape),
                   'claw2d_classic_set_launch_ex')
    rc = _lib.lib().claw2d_classic_step(ctx.handle, grid, m, r, bc, prec,
                                        qd.data_ptr(), qn.data_ptr(), ad.data_ptr() if ad is not None else None,
                                        dtd.data_ptr(), cfl.data_ptr(),
                                        torch.cuda.current_stream().cuda_stream)
    _lib.check(rc, 'claw2d_classic_step')
    torch.cuda.synchronize()
    out = qn[:, NG:NG + my, NG:NG + mx].permute(0, 2, 1).to('cpu', torch.float64).numpy()
    return out, float(cfl.item())


def oracle_classic_step(qbc, auxbc, case, dt, order=2, trans=1, split=False, mthlim=None, dx=0.1, dy=0.07):
    O.set_params(case['rp'], case['pd'])
    meth = O.method_array(order, trans, split)
    lim = O.mthlim_array(mthlim if mthlim is not None else 4, case['nw'])
    if not split:
        qn, cfl = O.step2(case['rp'], qbc, auxbc, dx, dy, dt, meth, lim, case['nw'])
        return qn[:, 2:-2, 2:-2], cfl
    qold = np.asfortranarray(qbc)
    qn = qold.copy(order='F')
    c1 = O.step2ds(case['rp'], qold, qn, auxbc, dx, dy, dt, meth, lim, case['nw'], 1)
    c2 = O.step2ds(case['rp'], qn, qn, auxbc, dx, dy, dt, meth, lim, case['nw'], 2)
    return qn[:, 2:-2, 2:-2], max(c1, c2)


def _dev_grid(mx, my, dx, dy):
    from pyclaw_amd import _lib
    return _lib.ClawGrid(mx, my, NG, 0, mx + 2 * NG, (mx + 2 * NG) * (my + 2 * NG), dx, dy)


def _rp_struct(case):
    from pyclaw_amd import _lib
    r = _lib.ClawRiemann()
    r.kind = case['rp']
    r.efix = int(case['pd'].get('efix', 1))
    vals = {O.RP_ACOUSTICS: ['rho', 'bulk', 'cc', 'zz'], O.RP_EULER: ['gamma', 'gamma1'],
            O.RP_SHALLOW: ['grav'], O.RP_VC_ACOUSTICS: []}[case['rp']]
    for k, key in enumerate(vals):
        r.p[k] = case['pd'][key]
    return r


def _bc_struct(lo, hi, alo=(1, 1), ahi=(1, 1)):
    from pyclaw_amd import _lib
    bc = _lib.ClawBC()
    for s in range(2):
        bc.lower[s], bc.upper[s] = lo[s], hi[s]
        bc.aux_lower[s], bc.aux_upper[s] = alo[s], ahi[s]
    return bc


def to_device_padded(a, dtype=None):
    """Reference padded array (ncomp, mx+2*NG, my+2*NG) -> device planar layout."""
    import torch
    dtype = dtype or torch.float64
    return torch.from_numpy(np.ascontiguousarray(a.transpose(0, 2, 1))).to('cuda', dtype)


def from_device_interior(t, mx, my):
    import torch
    return t[:, NG:NG + my, NG:NG + mx].permute(0, 2, 1).to('cpu', torch.float64).numpy()


def gpu_sharp_dq(qbc3, auxbc3, case, dt, bc_lower, bc_upper, abc_lower=(1, 1), abc_upper=(1, 1),
                 dx=0.1, dy=0.07, precision='f64', shape=None, ctx=None):
    """claw2d_sharpclaw_dq on reference-layout padded arrays with 3 ghost layers.
    shape: (rows, waves per SIMD, waves per block) set on a context first
    (claw2d_sharpclaw_set_launch_ex; rows 0 is the paired fill decomposition).
    ctx: a context to run on (by default none, or a new one for a shape)."""
    import torch
    from pyclaw_amd import _lib
    ne, nxp, nyp = qbc3.shape
    mx, my = nxp - 2 * NG, nyp - 2 * NG
    q = to_device_padded(qbc3)
    dq = torch.zeros_like(q)
    aux = to_device_padded(auxbc3) if auxbc3 is not None else None
    dtd = torch.tensor(dt, dtype=torch.float64, device='cuda')
    cfl = torch.zeros((), dtype=torch.float64, device='cuda')
    prec = _lib.F64_FMA if precision == 'f64fma' else _lib.F64
    if shape is not None:
        ctx = _lib.Context(0) if ctx is None else ctx
        _lib.check(_lib.lib().claw2d_sharpclaw_set_launch_ex(ctx.handle, _dev_grid(mx, my, dx, dy), _rp_struct(case),
                                                             prec, *shape), 'claw2d_sharpclaw_set_launch_ex')
    rc = _lib.lib().claw2d_sharpclaw_dq(ctx.handle if ctx is not None else None, _dev_grid(mx, my, dx, dy),
                                        _rp_struct(case), _bc_struct(bc_lower, bc_upper, abc_lower, abc_upper),
                                        prec,
                                        q.data_ptr(), dq.data_ptr(), aux.data_ptr() if aux is not None else None,
                                        dtd.data_ptr(), cfl.data_ptr(), torch.cuda.current_stream().cuda_stream)
    _lib.check(rc, 'claw2d_sharpclaw_dq')
    torch.cuda.synchronize()
    return from_device_interior(dq, mx, my), float(cfl.item())


def gpu_sharp_ssp104(q0, aux, case, dt, bc_lower, bc_upper, abc_lower=(1, 1), abc_upper=(1, 1),
                     dx=0.1, dy=0.07, precision='f64'):
    """One fused SSP104 step; q0/aux interior arrays (BCs synthesised in the kernel)."""
    import torch
    from pyclaw_amd import _lib
    ne, mx, my = q0.shape
    pad = np.zeros((ne, mx + 2 * NG, my + 2 * NG))
    pad[:, NG:-NG, NG:-NG] = q0
    q = to_device_padded(pad)
    qn, s1a, s1b, s2 = (torch.zeros_like(q) for _ in range(4))
    auxd = None
    if aux is not None:
        apad = np.zeros((aux.shape[0], mx + 2 * NG, my + 2 * NG))
        apad[:, NG:-NG, NG:-NG] = aux
        auxd = to_device_padded(apad)
    dtd = torch.tensor(dt, dtype=torch.float64, device='cuda')
    cfl = torch.zeros(10, dtype=torch.float64, device='cuda')
    rc = _lib.lib().claw2d_sharpclaw_ssp104_step(None, _dev_grid(mx, my, dx, dy), _rp_struct(case),
                                                 _bc_struct(bc_lower, bc_upper, abc_lower, abc_upper),
                                                 _lib.F64_FMA if precision == 'f64fma' else _lib.F64,
                                                 q.data_ptr(), qn.data_ptr(), s1a.data_ptr(), s1b.data_ptr(),
                                                 s2.data_ptr(), auxd.data_ptr() if auxd is not None else None,
                                                 dtd.data_ptr(), cfl.data_ptr(),
                                                 torch.cuda.current_stream().cuda_stream)
    _lib.check(rc, 'claw2d_sharpclaw_ssp104_step')
    torch.cuda.synchronize()
    return from_device_interior(qn, mx, my), cfl.cpu().numpy()
