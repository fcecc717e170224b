"""Differentiable CPU restatement of the unsplit classic step for
variable-coefficient acoustics, used as the gradient oracle of the adjoint.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The step is the reference's step2 (classic/step2.f90:92-248) with flux2
(classic/flux2.f90:100-206), the wave limiter (classic/limiter.f90:31-61,
philim.f90:18-55) and the vc-acoustics Riemann solvers (call site
apps/acoustics_2d_variable/acoustics.py:25-26; aux = (rho, c) per cell),
written with torch float64 tensor operations on the CPU so that
torch.autograd gives the exact gradient of the discrete map.  Its forward
values are checked against the C restatement (orc_step2) to 1e-14; the
gradient of the discrete map is then what the hand-written adjoint kernels
must reproduce.

Boundary fill follows solver.py:338-434 (x sides over all rows, then y sides
over the padded width; wall negates the normal velocity) and is written with
differentiable indexing, so gradients flow from ghost cells to their sources.

Layout: padded arrays (meqn, mx+2ng, my+2ng), the reference's qbc[m, i+ng, j+ng].
"""
import torch

NG = 2


def _fill_dim(a, ng, lo, hi, dim, negate_comp):
    """Ghost fill along one dimension (1: x, 2: y) of a padded tensor; returns a new tensor."""
    n = a.shape[dim] - 2 * ng
    idx = torch.arange(a.shape[dim])
    src = idx.clone()
    flip = torch.zeros(a.shape[dim], dtype=torch.bool)
    for g in range(ng):
        k = g - ng                     # interior-relative ghost index -ng..-1
        if lo == 1:
            src[g] = ng
        elif lo == 2:
            src[g] = k + n + ng
        elif lo == 3:
            src[g] = (-1 - k) + ng
            flip[g] = True
        kk = n + g                     # upper ghosts n..n+ng-1
        if hi == 1:
            src[ng + n + g] = ng + n - 1
        elif hi == 2:
            src[ng + n + g] = kk - n + ng
        elif hi == 3:
            src[ng + n + g] = (2 * n - 1 - kk) + ng
            flip[ng + n + g] = True
    out = torch.index_select(a, dim, src)
    if negate_comp is not None and bool(flip.any()):
        shape = [1] * a.dim()
        shape[dim] = a.shape[dim]
        sign = torch.ones(a.shape, dtype=a.dtype)
        f = flip.view(shape).expand(a.shape[0:1] + a.shape[1:]).clone()
        f[torch.arange(a.shape[0]) != negate_comp] = False
        sign = torch.where(f, -torch.ones_like(sign), sign)
        out = out * sign
    return out


def fill_bc(a, bc_lower, bc_upper, is_q=True, ng=NG):
    """solver.py:338-434 order: x lower/upper over all rows, then y over the padded width."""
    a = _fill_dim(a, ng, bc_lower[0], bc_upper[0], 1, 1 if is_q else None)
    a = _fill_dim(a, ng, bc_lower[1], bc_upper[1], 2, 2 if is_q else None)
    return a


def _max2(a, b):
    """a > b ? a : b -- the selection the forward kernels make; at a tie the
    derivative follows the selected operand (torch.maximum would split it)."""
    return torch.where(a > b, a, b)


def _min2(a, b):
    return torch.where(a < b, a, b)


def philim(a, b, meth):
    """philim.f90:18-55 (r = b/a); ids outside 1..5 act as minmod."""
    r = b / a
    zero = torch.zeros_like(r)
    one = torch.ones_like(r)
    two = 2.0 * one
    if meth == 2:
        return _max2(_max2(zero, _min2(one, 2.0 * r)), _min2(two, r))
    if meth == 3:
        ar = torch.abs(r)
        return (r + ar) / (1.0 + ar)
    if meth == 4:
        c = (1.0 + r) / 2.0
        return _max2(zero, _min2(_min2(c, two), 2.0 * r))
    if meth == 5:
        return r
    return _max2(zero, _min2(one, r))


def _sweep(qb, ab, dtdn, dtdt, order, trans, mthlim, dirn):
    """Contributions of every interface normal to dimension `dirn` (0: x, 1: y)
    to the padded update, as a tensor of the padded shape."""
    mu, mv = (1, 2) if dirn == 0 else (2, 1)
    # put the normal direction on axis 1
    q = qb if dirn == 0 else qb.transpose(1, 2)
    a = ab if dirn == 0 else ab.transpose(1, 2)
    nn = q.shape[1]
    lo, hi = q[:, :-1, :], q[:, 1:, :]           # interface k between cells k and k+1 (padded)
    alo, ahi = a[:, :-1, :], a[:, 1:, :]
    d0 = hi[0] - lo[0]
    d1 = hi[mu] - lo[mu]
    zi = ahi[0] * ahi[1]
    zim = alo[0] * alo[1]
    a1 = (-d0 + zi * d1) / (zim + zi)
    a2 = (d0 + zim * d1) / (zim + zi)
    # waves: only components 0 and mu are non-zero
    W = [torch.stack([-a1 * zim, a1]), torch.stack([a2 * zi, a2])]
    S = [-alo[1], ahi[1]]
    amdq = S[0] * W[0]
    apdq = S[1] * W[1]
    nf = nn - 1                                   # interfaces 0..nf-1
    contrib = torch.zeros_like(q)
    cq = torch.zeros_like(amdq)
    if order == 2:
        Wl = []
        for p in range(2):
            w = W[p]
            wn2 = (w * w).sum(0)
            dr = (w[:, :-1] * w[:, 1:]).sum(0)     # W(k) . W(k+1), for k = 0..nf-2
            dotr = torch.cat([dr, torch.zeros_like(wn2[-1:])], 0)
            dotl = torch.cat([torch.zeros_like(wn2[:1]), dr], 0)
            wt = w
            if mthlim[p] != 0:
                b = torch.where(S[p] > 0, dotl, dotr)
                safe = torch.where(wn2 != 0, wn2, torch.ones_like(wn2))
                phi = philim(safe, b, mthlim[p])
                # the reference limits interfaces 1..mx+1 (1-based, limiter.f90:31-34), i.e.
                # padded interfaces 1..nf-2 (padded interface k lies between padded cells k, k+1)
                kk = torch.arange(nf).view(nf, 1)
                lim_rows = (kk >= 1) & (kk <= nf - 2) & (wn2 != 0)
                wt = torch.where(lim_rows.unsqueeze(0), phi.unsqueeze(0) * w, w)
            Wl.append(wt)
        for p in range(2):
            s = torch.abs(S[p])
            cq = cq + (s * (1.0 - s * (0.5 * (dtdn + dtdn)))) * Wl[p]
    fadd = 0.5 * cq
    # cell k gets -dtdn*amdq(k) ... interface k is between cells k and k+1
    up = torch.zeros_like(q)
    # amdq of interface k -> cell k; apdq -> cell k+1
    up[0, :-1, :] = up[0, :-1, :] - dtdn * amdq[0] - dtdn * fadd[0]
    up[mu, :-1, :] = up[mu, :-1, :] - dtdn * amdq[1] - dtdn * fadd[1]
    up[0, 1:, :] = up[0, 1:, :] - dtdn * apdq[0] + dtdn * fadd[0]
    up[mu, 1:, :] = up[mu, 1:, :] - dtdn * apdq[1] + dtdn * fadd[1]
    contrib = contrib + up
    if trans > 0:
        tA = [amdq[0], amdq[1]]
        tP = [apdq[0], apdq[1]]
        if order == 2 and trans == 2:
            tA = [tA[0] + cq[0], tA[1] + cq[1]]
            tP = [tP[0] - cq[0], tP[1] - cq[1]]
        gam = dtdt * (0.5 * dtdn)
        tr = torch.zeros_like(q)
        for (t, col) in ((tA, 'lo'), (tP, 'hi')):
            aa = alo if col == 'lo' else ahi
            # transverse neighbours along axis 2: rows t-1, t, t+1 for t = 1..nt-2
            am, a0, ap = aa[:, :, :-2], aa[:, :, 1:-1], aa[:, :, 2:]
            t0 = t[0][:, 1:-1]
            tmv = torch.zeros_like(t0)             # the normal solve leaves the transverse velocity at 0
            cm, cp = am[1], ap[1]
            zm = am[0] * am[1]
            zz = a0[0] * a0[1]
            zp = ap[0] * ap[1]
            b1 = (-t0 + tmv * zz) / (zm + zz)
            b2 = (t0 + tmv * zz) / (zz + zp)
            bm0, bmv = cm * b1 * zm, -cm * b1
            bp0, bpv = cp * b2 * zp, cp * b2
            sl = slice(0, nf) if col == 'lo' else slice(1, nf + 1)
            for (comp, bm, bp) in ((0, bm0, bp0), (mv, bmv, bpv)):
                # (c, T-1): +gam*bm ; (c, T): gam*(bp - bm) ; (c, T+1): -gam*bp
                tr[comp, sl, 0:-2] = tr[comp, sl, 0:-2] + gam * bm
                tr[comp, sl, 1:-1] = tr[comp, sl, 1:-1] + gam * (bp - bm)
                tr[comp, sl, 2:] = tr[comp, sl, 2:] - gam * bp
        contrib = contrib + tr
    return contrib if dirn == 0 else contrib.transpose(1, 2)


def step(q, aux, dx, dy, dt, order=2, trans=1, mthlim=(4, 4), bc_lower=(1, 1), bc_upper=(1, 1),
         aux_bc_lower=(1, 1), aux_bc_upper=(1, 1)):
    """One unsplit step of interior q (3, mx, my) with interior aux (2, mx, my); returns interior qnew."""
    ne, mx, my = q.shape
    qb = torch.zeros((ne, mx + 2 * NG, my + 2 * NG), dtype=q.dtype)
    qb = qb.index_put((torch.arange(ne).view(-1, 1, 1), torch.arange(NG, NG + mx).view(1, -1, 1),
                       torch.arange(NG, NG + my).view(1, 1, -1)), q)
    ab = torch.zeros((2, mx + 2 * NG, my + 2 * NG), dtype=q.dtype)
    ab = ab.index_put((torch.arange(2).view(-1, 1, 1), torch.arange(NG, NG + mx).view(1, -1, 1),
                       torch.arange(NG, NG + my).view(1, 1, -1)), aux)
    qb = fill_bc(qb, bc_lower, bc_upper, True)
    ab = fill_bc(ab, aux_bc_lower, aux_bc_upper, False)
    dtdx, dtdy = dt / dx, dt / dy
    cx = _sweep(qb, ab, dtdx, dtdy, order, trans, mthlim, 0)
    cy = _sweep(qb, ab, dtdy, dtdx, order, trans, mthlim, 1)
    return q + cx[:, NG:NG + mx, NG:NG + my] + cy[:, NG:NG + mx, NG:NG + my]


def misfit(q0, rho, c, nsteps, dx, dy, dt, receivers, p_obs, **kw):
    """J = 0.5 * sum_{n=1..nsteps} sum_r (p_n(r) - p_obs[n-1, r])^2 with receivers
    given as (ix, iy) index tensors; returns (J, traces)."""
    aux = torch.stack([rho, c])
    q = q0
    J = torch.zeros((), dtype=q0.dtype)
    traces = []
    ix, iy = receivers
    for n in range(nsteps):
        q = step(q, aux, dx, dy, dt, **kw)
        p = q[0, ix, iy]
        traces.append(p.detach().clone())
        J = J + 0.5 * ((p - p_obs[n]) ** 2).sum()
    return J, torch.stack(traces)
