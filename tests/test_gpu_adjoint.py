"""Hand-written adjoint kernels of the vc-acoustics unsplit step on the GPU
(claw2d_vc_adjoint_step), checked against
  * the torch-autograd gradient of the CPU restatement (oracle/adjoint_ref.py,
    itself pinned to orc_step2 and to finite differences in
    test_adjoint_oracle.py), one step, every limiter and boundary code;
  * the exact transpose of the dense Jacobian built from orc_step2 in the
    linear case (unlimited Lax-Wendroff);
  * the same autograd oracle through a whole checkpointed misfit run;
  * central finite differences of the GPU forward misfit (tolerance 1e-6).
"""
import numpy as np
import pytest
import torch

import oracle as O
from oracle import adjoint_ref as A

pytestmark = pytest.mark.gpu

CONFIGS = [  # order, transverse, limiters, bc_lower, bc_upper
    (2, 1, (0, 0), (1, 1), (1, 1)),
    (2, 1, (4, 4), (1, 1), (1, 1)),
    (2, 1, (4, 4), (3, 1), (1, 3)),
    (2, 1, (4, 4), (2, 2), (2, 2)),
    (2, 2, (1, 2), (3, 3), (3, 3)),
    (2, 2, (3, 5), (1, 2), (1, 2)),
    (1, 1, (0, 0), (1, 1), (1, 1)),
    (2, 0, (6, 1), (3, 2), (3, 2)),
]


def _problem(mx, my, cfg, nsteps=1, receivers=((0, 0),), every=45, dx=0.1, dy=0.12, dt=0.03):
    from pyclaw_amd.adjoint import MisfitProblem
    order, trans, lim, bcl, bcu = cfg
    return MisfitProblem(mx, my, dx, dy, dt, nsteps, receivers, order=order, transverse=trans, limiters=list(lim),
                         bc_lower=bcl, bc_upper=bcu, aux_bc_lower=bcl, aux_bc_upper=bcu, every=every)


def _gpu_vjp(mx, my, cfg, q, aux, lam, dt=0.03):
    P = _problem(mx, my, cfg, dt=dt)
    qp, ap, lp = P.padded(q), P.padded(aux), P.padded(lam)
    lo = torch.zeros_like(lp)
    gaux = torch.zeros_like(ap)
    P._adjoint_step(qp, ap, lp, lo, gaux)
    torch.cuda.synchronize()
    return P.interior(lo).cpu().numpy(), P.interior(gaux).cpu().numpy()


def _torch_vjp(cfg, q, aux, lam, dx=0.1, dy=0.12, dt=0.03):
    order, trans, lim, bcl, bcu = cfg
    qt = torch.from_numpy(q).requires_grad_(True)
    at = torch.from_numpy(aux).requires_grad_(True)
    out = A.step(qt, at, dx, dy, dt, order=order, trans=trans, mthlim=lim, bc_lower=bcl, bc_upper=bcu,
                 aux_bc_lower=bcl, aux_bc_upper=bcu)
    gq, ga = torch.autograd.grad(out, (qt, at), torch.from_numpy(lam))
    return gq.numpy(), ga.numpy()


def _data(mx, my, seed):
    rng = np.random.default_rng(seed)
    q = rng.standard_normal((3, mx, my))
    aux = np.stack([1 + rng.random((mx, my)), 0.5 + rng.random((mx, my))])
    lam = rng.standard_normal((3, mx, my))
    return q, aux, lam


@pytest.mark.parametrize("cfg", CONFIGS)
@pytest.mark.parametrize("shape", [(9, 7), (70, 37)])
def test_adjoint_step_matches_autograd(cfg, shape):
    mx, my = shape
    print("adjoint step", cfg, shape, flush=True)
    q, aux, lam = _data(mx, my, mx * 100 + my)
    gq, ga = _gpu_vjp(mx, my, cfg, q, aux, lam)
    rq, ra = _torch_vjp(cfg, q, aux, lam)
    assert np.abs(gq - rq).max() <= 1e-13 * np.abs(rq).max()
    assert np.abs(ga - ra).max() <= 1e-13 * np.abs(ra).max()


def test_adjoint_step_is_the_transpose_of_the_dense_linear_step():
    """Unlimited Lax-Wendroff: the step is q -> M q with M built from orc_step2."""
    mx, my = 8, 6
    cfg = (2, 1, (0, 0), (3, 1), (1, 3))
    q, aux, lam = _data(mx, my, 4)
    n = 3 * mx * my
    M = np.zeros((n, n))
    O.set_params(O.RP_VC_ACOUSTICS, {})
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        qbc = np.zeros((3, mx + 4, my + 4), order='F')
        qbc[:, 2:-2, 2:-2] = e.reshape(3, mx, my)
        O.fill_bc(qbc, 2, cfg[3], cfg[4], True)
        abc = np.zeros((2, mx + 4, my + 4), order='F')
        abc[:, 2:-2, 2:-2] = aux
        O.fill_bc(abc, 2, cfg[3], cfg[4], False)
        qn, _ = O.step2(O.RP_VC_ACOUSTICS, qbc, abc, 0.1, 0.12, 0.03, O.method_array(2, 1, False),
                        O.mthlim_array([0, 0], 2), 2)
        M[:, k] = qn[:, 2:-2, 2:-2].ravel()
    gq, _ = _gpu_vjp(mx, my, cfg, q, aux, lam)
    ref = (M.T @ lam.ravel()).reshape(3, mx, my)
    assert np.abs(gq - ref).max() <= 1e-13 * np.abs(ref).max()


def test_checkpointed_misfit_gradient_matches_autograd():
    from pyclaw_amd.problems import adjoint_medium
    mx, my, N, every = 24, 20, 17, 5
    cfg = (2, 1, (4, 4), (3, 1), (1, 1))
    rng = np.random.default_rng(1)
    q0, K, rho, K_true = adjoint_medium(rng, mx, my, width=0.15)
    dx, dy = 1.0 / mx, 1.0 / my
    dt = 0.8 * min(dx, dy) / np.sqrt(K / rho).max()
    rec = [(i, my - 3) for i in range(2, mx - 2, 3)]
    ix = torch.tensor([r[0] for r in rec])
    iy = torch.tensor([r[1] for r in rec])
    kw = dict(order=2, trans=1, mthlim=(4, 4), bc_lower=cfg[3], bc_upper=cfg[4], aux_bc_lower=cfg[3],
              aux_bc_upper=cfg[4])
    with torch.no_grad():
        _, obs = A.misfit(torch.from_numpy(q0), torch.from_numpy(rho), torch.from_numpy(np.sqrt(K_true / rho)), N,
                          dx, dy, dt, (ix, iy), torch.zeros(N, len(rec), dtype=torch.float64), **kw)
    q0t = torch.from_numpy(q0).requires_grad_(True)
    Kt = torch.from_numpy(K).requires_grad_(True)
    rt = torch.from_numpy(rho).requires_grad_(True)
    Jref, _ = A.misfit(q0t, rt, torch.sqrt(Kt / rt), N, dx, dy, dt, (ix, iy), obs, **kw)
    gq_ref, gK_ref, gr_ref = torch.autograd.grad(Jref, (q0t, Kt, rt))
    from pyclaw_amd.adjoint import MisfitProblem
    P = MisfitProblem(mx, my, dx, dy, dt, N, rec, p_obs=obs.numpy(), order=2, transverse=1, limiters=4,
                      bc_lower=cfg[3], bc_upper=cfg[4], aux_bc_lower=cfg[3], aux_bc_upper=cfg[4], every=every)
    J, gq, gK, gr = P.gradient(q0, K, rho)
    assert abs(float(J) - float(Jref)) <= 1e-12 * abs(float(Jref))
    for got, ref in ((gq, gq_ref), (gK, gK_ref), (gr, gr_ref)):
        got = got.cpu().numpy()
        ref = ref.numpy()
        assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()


def test_misfit_gradient_matches_central_differences():
    """Unlimited Lax-Wendroff (smooth in q and aux): directional derivatives of
    the GPU misfit along smooth random directions agree with the adjoint
    gradient to 1e-6 relative."""
    from pyclaw_amd.adjoint import MisfitProblem, VCAcousticsMisfit, receivers_on_row
    from pyclaw_amd.problems import adjoint_medium, smooth_random_field
    mx = my = 96
    N = 120
    rng = np.random.default_rng(5)
    q0, K, rho, K_true = adjoint_medium(rng, mx, my, width=0.06)
    dx = dy = 1.0 / mx
    dt = 0.8 * dx / np.sqrt(K / rho).max()
    rec = receivers_on_row(mx, my, 0.95, 16)
    P = MisfitProblem(mx, my, dx, dy, dt, N, rec, order=2, transverse=1, limiters=0, every=11)
    obs, _, _ = P.forward(P.padded(q0), P.aux_from(K_true, rho), keep_checkpoints=False)
    P.p_obs = obs
    J, gq, gK, gr = P.gradient(q0, K, rho)
    vK = smooth_random_field(rng, mx, my, -1.0, 1.0) * K
    vr = smooth_random_field(rng, mx, my, -1.0, 1.0) * rho
    vq = np.zeros_like(q0)
    vq[0] = smooth_random_field(rng, mx, my, -1.0, 1.0) * q0[0]

    def Jat(q, Kk, rr):
        tr, _, _ = P.forward(P.padded(q), P.aux_from(Kk, rr), keep_checkpoints=False)
        return float(P.objective(tr))

    for (name, v, g) in (("K", vK, gK), ("rho", vr, gr), ("p0", vq[0], gq[0])):
        h = 1e-4
        if name == "K":
            jp, jm = Jat(q0, K + h * vK, rho), Jat(q0, K - h * vK, rho)
        elif name == "rho":
            jp, jm = Jat(q0, K, rho + h * vr), Jat(q0, K, rho - h * vr)
        else:
            jp, jm = Jat(q0 + h * vq, K, rho), Jat(q0 - h * vq, K, rho)
        fd = (jp - jm) / (2 * h)
        ad = float((g.cpu().numpy() * v).sum())
        print("FD check", name, fd, ad, abs(fd - ad) / abs(ad), flush=True)
        assert abs(fd - ad) <= 1e-6 * abs(ad)
    # the autograd.Function form gives the same gradients
    q0t = torch.tensor(q0, device='cuda', requires_grad=True)
    Kt = torch.tensor(K, device='cuda', requires_grad=True)
    rt = torch.tensor(rho, device='cuda', requires_grad=True)
    Jf = VCAcousticsMisfit.apply(P, q0t, Kt, rt)
    Jf.backward()
    assert float(Jf) == float(J)
    assert torch.equal(Kt.grad, gK) and torch.equal(rt.grad, gr) and torch.equal(q0t.grad, gq)


@pytest.mark.parametrize("scale", [1e-155, 1e-160])
def test_adjoint_step_is_finite_and_scale_invariant_at_subnormal_wave_norms(scale):
    """The limited step is positively homogeneous in q (theta and phi depend on
    ratios of waves), so lam_n does not change when q is scaled and the aux
    gradient scales with q.  At 1e-155 |W|^2 is just below the normal range
    (~44 significant bits left) and the invariance still holds; at 1e-160
    |W|^2 ~ 1e-320 keeps only a few bits, so the forward limiter itself is no
    longer scale-invariant and only finiteness is required (the limiter
    derivative must never form 1/|W|^2)."""
    mx, my = 21, 13
    cfg = (2, 1, (4, 4), (1, 1), (1, 1))
    q, aux, lam = _data(mx, my, 77)
    g1, a1 = _gpu_vjp(mx, my, cfg, q, aux, lam)
    g2, a2 = _gpu_vjp(mx, my, cfg, scale * q, aux, lam)
    assert np.isfinite(g2).all() and np.isfinite(a2).all()
    if scale >= 1e-155:
        assert np.abs(g2 - g1).max() <= 1e-10 * np.abs(g1).max()
        assert np.abs(a2 / scale - a1).max() <= 1e-10 * np.abs(a1).max()


@pytest.mark.parametrize("bc", [((1, 1), (1, 1)), ((3, 3), (3, 3)), ((3, 1), (1, 3)), ((2, 2), (2, 2))])
@pytest.mark.parametrize("shape", [(63, 37), (9, 6)])
def test_misfit_gradient_with_receivers_on_the_boundary_ring(bc, shape):
    """Receivers in the corners and in the two cells next to every side, where
    the ghosts' contributions are folded back into lam: inside the tile for
    extrapolation and wall sides (63 x 37 also moves the last tile of both
    axes back to hold the high ghosts), through the fold launch for periodic
    ones."""
    from pyclaw_amd.adjoint import MisfitProblem
    from pyclaw_amd.problems import adjoint_medium
    mx, my = shape
    N, every = 6, 4
    bcl, bcu = bc
    print("ring receivers", bc, shape, flush=True)
    rng = np.random.default_rng(11)
    q0, K, rho, K_true = adjoint_medium(rng, mx, my, width=0.2)
    dx, dy = 1.0 / mx, 1.0 / my
    dt = 0.8 * min(dx, dy) / np.sqrt(K / rho).max()
    rec = sorted({(i, j) for i in (0, 1, mx // 2, mx - 2, mx - 1) for j in (0, 1, my // 2, my - 2, my - 1)})
    ix = torch.tensor([r[0] for r in rec])
    iy = torch.tensor([r[1] for r in rec])
    kw = dict(order=2, trans=2, mthlim=(4, 4), bc_lower=bcl, bc_upper=bcu, aux_bc_lower=bcl, aux_bc_upper=bcu)
    with torch.no_grad():
        _, obs = A.misfit(torch.from_numpy(q0), torch.from_numpy(rho), torch.from_numpy(np.sqrt(K_true / rho)), N,
                          dx, dy, dt, (ix, iy), torch.zeros(N, len(rec), dtype=torch.float64), **kw)
    q0t = torch.from_numpy(q0).requires_grad_(True)
    Kt = torch.from_numpy(K).requires_grad_(True)
    rt = torch.from_numpy(rho).requires_grad_(True)
    Jref, _ = A.misfit(q0t, rt, torch.sqrt(Kt / rt), N, dx, dy, dt, (ix, iy), obs, **kw)
    refs = torch.autograd.grad(Jref, (q0t, Kt, rt))
    P = MisfitProblem(mx, my, dx, dy, dt, N, rec, p_obs=obs.numpy(), order=2, transverse=2, limiters=4,
                      bc_lower=bcl, bc_upper=bcu, aux_bc_lower=bcl, aux_bc_upper=bcu, every=every)
    J, gq, gK, gr = P.gradient(q0, K, rho)
    assert abs(float(J) - float(Jref)) <= 1e-12 * abs(float(Jref))
    for got, ref in zip((gq, gK, gr), refs):
        got = got.cpu().numpy()
        ref = ref.numpy()
        assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("every", [1, 4, 17, 40])
def test_checkpoint_interval_does_not_change_the_gradient(every):
    """Every checkpoint interval (1: all states kept, no recomputation; larger:
    segment recomputation, including a last partial segment and one segment
    longer than the run) gives bit-identical gradients."""
    from pyclaw_amd.adjoint import MisfitProblem
    from pyclaw_amd.problems import adjoint_medium
    mx, my, N = 30, 22, 17
    rng = np.random.default_rng(3)
    q0, K, rho, K_true = adjoint_medium(rng, mx, my, width=0.15)
    dx, dy = 1.0 / mx, 1.0 / my
    dt = 0.8 * min(dx, dy) / np.sqrt(K / rho).max()
    rec = [(i, my - 4) for i in range(1, mx - 1, 4)]
    ref = MisfitProblem(mx, my, dx, dy, dt, N, rec, limiters=4, every=3)
    obs, _, _ = ref.forward(ref.padded(q0), ref.aux_from(K_true, rho), keep_checkpoints=False)
    ref.p_obs = obs
    J0, g0, k0, r0 = ref.gradient(q0, K, rho)
    P = MisfitProblem(mx, my, dx, dy, dt, N, rec, p_obs=obs, limiters=4, every=every)
    J1, g1, k1, r1 = P.gradient(q0, K, rho)
    assert float(J0) == float(J1)
    assert torch.equal(g0, g1) and torch.equal(k0, k1) and torch.equal(r0, r1)
