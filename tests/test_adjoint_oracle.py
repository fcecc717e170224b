"""The gradient oracle of the adjoint (oracle/adjoint_ref.py) pinned on the CPU:
  * its forward step equals the C restatement of step2 (orc_step2) for
    vc acoustics, every limiter, trans_inc/trans_cor and every boundary code;
  * its autograd gradient of one step equals the exact transpose of the
    dense Jacobian built column by column from orc_step2 in the linear case
    (unlimited Lax-Wendroff), and central differences of orc_step2 for the
    aux gradient and for limited runs on smooth data.
"""
import numpy as np
import pytest
import torch

import oracle as O
from oracle import adjoint_ref as A

CONFIGS = [  # order, transverse, limiters, bc_lower, bc_upper
    (2, 1, (4, 4), (1, 1), (1, 1)),
    (2, 0, (0, 0), (3, 1), (1, 3)),
    (2, 2, (1, 2), (2, 3), (2, 3)),
    (1, 1, (0, 0), (3, 3), (3, 3)),
    (2, 1, (3, 5), (1, 2), (1, 2)),
]


def _data(mx, my, seed, smooth=False):
    rng = np.random.default_rng(seed)
    if smooth:
        from pyclaw_amd.problems import smooth_random_field
        q = np.stack([smooth_random_field(rng, mx, my, -1.0, 1.0) for _ in range(3)])
    else:
        q = rng.standard_normal((3, mx, my))
    rho = 1 + rng.random((mx, my))
    c = 0.5 + rng.random((mx, my))
    return q, np.stack([rho, c])


def _c_step(q, aux, dx, dy, dt, order, trans, lim, bcl, bcu):
    mx, my = q.shape[1:]
    qbc = np.zeros((3, mx + 4, my + 4), order='F')
    qbc[:, 2:-2, 2:-2] = q
    O.fill_bc(qbc, 2, bcl, bcu, True)
    abc = np.zeros((2, mx + 4, my + 4), order='F')
    abc[:, 2:-2, 2:-2] = aux
    O.fill_bc(abc, 2, bcl, bcu, False)
    O.set_params(O.RP_VC_ACOUSTICS, {})
    qn, _ = O.step2(O.RP_VC_ACOUSTICS, qbc, abc, dx, dy, dt, O.method_array(order, trans, False),
                    O.mthlim_array(list(lim), 2), 2)
    return np.array(qn[:, 2:-2, 2:-2])


def _t_step(q, aux, dx, dy, dt, order, trans, lim, bcl, bcu):
    return A.step(q, aux, dx, dy, dt, order=order, trans=trans, mthlim=lim, bc_lower=bcl, bc_upper=bcu,
                  aux_bc_lower=bcl, aux_bc_upper=bcu)


@pytest.mark.parametrize("cfg", CONFIGS)
def test_torch_step_equals_c_oracle(cfg):
    q, aux = _data(11, 9, 3)
    ref = _c_step(q, aux, 0.1, 0.12, 0.03, *cfg)
    got = _t_step(torch.from_numpy(q), torch.from_numpy(aux), 0.1, 0.12, 0.03, *cfg).numpy()
    assert np.abs(got - ref).max() <= 4e-15 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("trans", [0, 1, 2])
def test_autograd_equals_dense_transpose_linear_case(trans):
    """Unlimited Lax-Wendroff is linear in q: M from unit vectors, then M^T lam."""
    mx, my = 7, 6
    q, aux = _data(mx, my, 5)
    cfg = (2, trans, (0, 0), (3, 1), (1, 3))
    n = 3 * mx * my
    M = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        M[:, k] = _c_step(e.reshape(3, mx, my), aux, 0.1, 0.12, 0.03, *cfg).ravel()
    lam = np.random.default_rng(9).standard_normal((3, mx, my))
    qt = torch.from_numpy(q).requires_grad_(True)
    out = _t_step(qt, torch.from_numpy(aux), 0.1, 0.12, 0.03, *cfg)
    (g,) = torch.autograd.grad(out, qt, torch.from_numpy(lam))
    ref = (M.T @ lam.ravel()).reshape(3, mx, my)
    assert np.abs(g.numpy() - ref).max() <= 1e-14 * np.abs(ref).max()


@pytest.mark.parametrize("lim", [(0, 0), (4, 4), (1, 3)])
def test_autograd_aux_and_limited_gradients_match_central_differences(lim):
    mx, my = 9, 8
    q, aux = _data(mx, my, 7, smooth=True)
    cfg = (2, 1, lim, (1, 1), (1, 1))
    rng = np.random.default_rng(11)
    lam = rng.standard_normal((3, mx, my))
    vq = rng.standard_normal((3, mx, my))
    va = rng.standard_normal((2, mx, my)) * 0.1
    qt = torch.from_numpy(q).requires_grad_(True)
    at = torch.from_numpy(aux).requires_grad_(True)
    out = _t_step(qt, at, 0.1, 0.12, 0.03, *cfg)
    gq, ga = torch.autograd.grad(out, (qt, at), torch.from_numpy(lam))
    h = 1e-6
    fq = lambda s: float((_c_step(q + s * vq, aux, 0.1, 0.12, 0.03, *cfg) * lam).sum())
    fa = lambda s: float((_c_step(q, aux + s * va, 0.1, 0.12, 0.03, *cfg) * lam).sum())
    dq = (fq(h) - fq(-h)) / (2 * h)
    da = (fa(h) - fa(-h)) / (2 * h)
    assert abs(dq - float((gq.numpy() * vq).sum())) <= 1e-7 * abs(dq) + 1e-9
    assert abs(da - float((ga.numpy() * va).sum())) <= 1e-6 * abs(da) + 1e-9
