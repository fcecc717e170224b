"""Gradients of a receiver-misfit objective through many classic steps of 2D
variable-coefficient acoustics, by the discrete adjoint on the GPU.

The reference has no adjoint; the forward steps it differentiates are the
reference's unsplit step2 (classic/step2.f90:92-248) for rp2_vc_acoustics
with aux = (rho, c) per cell (apps/acoustics_2d_variable/acoustics.py:25-26,
46-61), at a fixed dt.

    J(q0, K, rho) = 0.5 * sum_{n=1..N} sum_r (p_n(r) - p_obs[n-1, r])^2,
    c = sqrt(K / rho)

Forward: N steps through claw2d_classic_step, keeping every `every`-th state
as a checkpoint and the pressure at the receiver cells after each step
(claw2d_gather).  Reverse: for each checkpoint segment, last first, the
segment's states are recomputed from its checkpoint and the adjoint steps
(claw2d_vc_adjoint_step, hand-written HIP kernels) run backwards through it;
each receiver residual enters lam at its cell (claw2d_misfit_source).
The forward pass also keeps the last segment's states, which the reverse
sweep needs first, so it starts without waiting for a recomputation.
Memory: ceil(N/every) checkpoints plus three segments of `every` states (the
last segment, and two buffers that the reverse sweep recomputes into); with
every=1 all N+1 states stay in HBM and nothing is recomputed.

VCAcousticsMisfit.apply(q0, K, rho) is the torch.autograd.Function form.
"""
import ctypes

import torch

from . import _lib
from .state import NG_STORE


class Checkpoints(object):
    """States kept by a checkpointed forward pass: states[k] = q_{k*every}, and
    tail[n - s0] = q_n for the last segment's steps n = s0 .. N-1
    (s0 = (ceil(N/every) - 1) * every).  Indexing reads the checkpoints."""

    def __init__(self, states, tail, first):
        self.states, self.tail, self.first = states, tail, first

    def __getitem__(self, k):
        return self.states[k]

    def __len__(self):
        return self.states.shape[0]


class _Launches(object):
    """The launches of one pass with every constant argument bound once: states
    and traces are plain device addresses (base + index * stride), so a launch
    is one ctypes call with no tensor views (a pass issues thousands)."""

    def __init__(self, P, aux, stream, cfl):
        self.L = P.L
        self.ctx = P._ctx.handle
        self.grid, self.meth = ctypes.byref(P.grid), ctypes.byref(P.method)
        self.rp, self.bc = ctypes.byref(P.rp), ctypes.byref(P.bc)
        self.auxp, self.dtp, self.cflp = aux.data_ptr(), P._dt.data_ptr(), cfl.data_ptr()
        self.dt = P.dt
        self.wsp, self.wsb = P._ws.data_ptr(), P._ws.numel() * 8
        self.offp, self.nrec = P.offsets.data_ptr(), P.nrec
        self.st = stream.cuda_stream

    def step(self, q, qn):
        rc = self.L.claw2d_classic_step_ex(self.ctx, self.grid, self.meth, self.rp, self.bc, _lib.F64, q, qn,
                                           self.auxp, self.dtp, self.cflp, _lib.STEP_ACCUMULATE_CFL, self.st)
        if rc:
            _lib.check(rc, "claw2d_classic_step_ex")

    def record(self, q, out):
        rc = self.L.claw2d_gather(q, self.offp, self.nrec, out, self.st)
        if rc:
            _lib.check(rc, "claw2d_gather")

    def source(self, lam, tr, obs):
        rc = self.L.claw2d_misfit_source(lam, self.offp, self.nrec, tr, obs, self.st)
        if rc:
            _lib.check(rc, "claw2d_misfit_source")

    def adjoint_misfit(self, q, lam_in, lam_out, gaux, tr, obs):
        """Adjoint step, then the residual of the previous step's traces enters lam_out."""
        rc = self.L.claw2d_vc_adjoint_step_misfit(self.grid, self.meth, self.bc, q, self.auxp, lam_in, lam_out, gaux,
                                                  self.dt, self.wsp, self.wsb, self.offp, self.nrec, tr, obs, self.st)
        if rc:
            _lib.check(rc, "claw2d_vc_adjoint_step_misfit")

    def adjoint(self, q, lam_in, lam_out, gaux):
        rc = self.L.claw2d_vc_adjoint_step(self.grid, self.meth, self.bc, q, self.auxp, lam_in, lam_out, gaux,
                                           self.dt, self.wsp, self.wsb, self.st)
        if rc:
            _lib.check(rc, "claw2d_vc_adjoint_step")


class MisfitProblem(object):
    """Grid, method, boundary codes, receivers and observed traces of one
    misfit objective.  receivers: sequence of (i, j) interior cells.
    order/transverse/limiters as in ClawSolver2D (order 2 with limiters 0 is
    unlimited Lax-Wendroff, the exactly linear case).  bc: q and aux boundary
    codes per side, extrap (1), periodic (2) or wall (3)."""

    def __init__(self, mx, my, dx, dy, dt, nsteps, receivers, p_obs=None, order=2, transverse=1, limiters=4,
                 bc_lower=(1, 1), bc_upper=(1, 1), aux_bc_lower=(1, 1), aux_bc_upper=(1, 1), every=45,
                 device=None):
        if not torch.cuda.is_available():
            raise _lib.ClawError("the adjoint runs on the GPU; no device is available")
        self.L = _lib.lib()
        for name in ("claw2d_vc_adjoint_step", "claw2d_gather", "claw2d_misfit_source"):
            if not hasattr(self.L, name):
                raise _lib.ClawError("libclawhip.so lacks %s: rebuild it" % name)
        self.dev = torch.device(device) if device is not None else torch.device('cuda', 0)
        self.mx, self.my, self.dx, self.dy, self.dt = int(mx), int(my), float(dx), float(dy), float(dt)
        self.nsteps, self.every = int(nsteps), max(1, int(every))
        ng = NG_STORE
        self.ng, self.ldq = ng, self.mx + 2 * ng
        self.rows = self.my + 2 * ng
        self.plane = self.ldq * self.rows
        self.grid = _lib.ClawGrid(self.mx, self.my, ng, 0, self.ldq, self.plane, self.dx, self.dy)
        m = _lib.ClawMethod()
        m.order, m.transverse, m.split, m.fwave, m.num_waves = int(order), int(transverse), 0, 0, 2
        lim = limiters if isinstance(limiters, (list, tuple)) else [limiters] * 2
        for k, v in enumerate(lim):
            m.mthlim[k] = int(v)
        self.method = m
        self.rp = _lib.ClawRiemann()
        self.rp.kind, self.rp.efix = _lib.RP_VC_ACOUSTICS, 1
        bc = _lib.ClawBC()
        for s in range(2):
            bc.lower[s], bc.upper[s] = bc_lower[s], bc_upper[s]
            bc.aux_lower[s], bc.aux_upper[s] = aux_bc_lower[s], aux_bc_upper[s]
        self.bc = bc
        rec = [(int(i), int(j)) for (i, j) in receivers]
        for (i, j) in rec:
            if not (0 <= i < self.mx and 0 <= j < self.my):
                raise ValueError("receiver (%d, %d) outside the grid" % (i, j))
        if len(set(rec)) != len(rec):
            raise ValueError("receivers must be distinct cells")
        self.nrec = len(rec)
        self.offsets = torch.tensor([(j + ng) * self.ldq + (i + ng) for (i, j) in rec], dtype=torch.int64,
                                    device=self.dev)
        self.p_obs = None if p_obs is None else torch.as_tensor(p_obs, dtype=torch.float64, device=self.dev)
        self._ctx = _lib.Context(self.dev.index or 0)
        f64 = dict(dtype=torch.float64, device=self.dev)
        self._dt = torch.tensor([self.dt], **f64)
        self._cfl = torch.zeros(1, **f64)
        self._ws = torch.empty(int(self.L.claw2d_vc_adjoint_workspace_bytes(self.grid)) // 8, **f64)
        self._tuned = False

    # ---- layout helpers ----------------------------------------------------
    def padded(self, a):
        """(ncomp, mx, my) interior -> (ncomp, my+2ng, ldq) padded device tensor."""
        a = torch.as_tensor(a, dtype=torch.float64, device=self.dev)
        out = torch.zeros((a.shape[0], self.rows, self.ldq), dtype=torch.float64, device=self.dev)
        out[:, self.ng:self.ng + self.my, self.ng:self.ng + self.mx] = a.permute(0, 2, 1)
        return out

    def interior(self, buf):
        ng = self.ng
        return buf[:, ng:ng + self.my, ng:ng + self.mx].permute(0, 2, 1)

    def aux_from(self, K, rho):
        K = torch.as_tensor(K, dtype=torch.float64, device=self.dev)
        rho = torch.as_tensor(rho, dtype=torch.float64, device=self.dev)
        return self.padded(torch.stack([rho, torch.sqrt(K / rho)]))

    # ---- device calls ------------------------------------------------------
    def _stream(self):
        return torch.cuda.current_stream(self.dev).cuda_stream

    def _step(self, qold, qnew, aux, stream=None, cfl=None):
        rc = self.L.claw2d_classic_step_ex(self._ctx.handle, self.grid, self.method, self.rp, self.bc, _lib.F64,
                                           qold.data_ptr(), qnew.data_ptr(), aux.data_ptr(), self._dt.data_ptr(),
                                           (self._cfl if cfl is None else cfl).data_ptr(),
                                           _lib.STEP_ACCUMULATE_CFL,
                                           self._stream() if stream is None else stream.cuda_stream)
        _lib.check(rc, "claw2d_classic_step_ex")

    def _adjoint_step(self, q, aux, lam_in, lam_out, gaux):
        rc = self.L.claw2d_vc_adjoint_step(self.grid, self.method, self.bc, q.data_ptr(), aux.data_ptr(),
                                           lam_in.data_ptr(), lam_out.data_ptr(), gaux.data_ptr(), self.dt,
                                           self._ws.data_ptr(), self._ws.numel() * 8, self._stream())
        _lib.check(rc, "claw2d_vc_adjoint_step")

    def _side_stream(self):
        if getattr(self, "_side", None) is None:
            self._side = torch.cuda.Stream(self.dev)     # default priority
        return self._side

    # ---- passes ------------------------------------------------------------
    def tune(self, q0p, aux):
        """Pick the forward kernel's row-band height for this grid (claw2d_classic_autotune)."""
        b = torch.empty_like(q0p)
        band = ctypes.c_int(0)
        _lib.check(self.L.claw2d_classic_autotune(self._ctx.handle, self.grid, self.method, self.rp, self.bc,
                                                  _lib.F64, q0p.data_ptr(), b.data_ptr(), aux.data_ptr(),
                                                  self._dt.data_ptr(), self._cfl.data_ptr(), self._stream(),
                                                  ctypes.byref(band)), "claw2d_classic_autotune")
        self._tuned = True
        return band.value

    def forward(self, q0p, aux, keep_checkpoints=True):
        """N steps from the padded state q0p; returns (traces (N, nrec),
        checkpoints, final state).  With every == 1 the checkpoints are one
        tensor holding all N+1 states, written by the steps directly; otherwise
        a Checkpoints object: states 0, every, 2 every, ... and the last
        segment's states."""
        N, E = self.nsteps, self.every
        if not self._tuned:
            self.tune(q0p, aux)
        traces = torch.empty((N, self.nrec), dtype=torch.float64, device=self.dev)
        self._cfl.zero_()
        shape = tuple(q0p.shape)
        go = _Launches(self, aux, torch.cuda.current_stream(self.dev), self._cfl)
        sb, trp, rb = q0p.numel() * 8, traces.data_ptr(), self.nrec * 8
        if keep_checkpoints and E == 1:
            states = torch.empty((N + 1,) + shape, dtype=torch.float64, device=self.dev)
            states[0].copy_(q0p)
            base = states.data_ptr()
            for n in range(N):
                go.step(base + n * sb, base + (n + 1) * sb)
                go.record(base + (n + 1) * sb, trp + n * rb)
            return traces, states, states[N]
        cps, s0 = None, N
        if keep_checkpoints:
            nseg = (N + E - 1) // E
            s0 = (nseg - 1) * E
            cps = Checkpoints(torch.empty((nseg,) + shape, dtype=torch.float64, device=self.dev),
                              torch.empty((N - s0,) + shape, dtype=torch.float64, device=self.dev), s0)
        ab = [q0p.clone(), torch.empty_like(q0p)]   # ab[cur] holds the current state
        abp = [t.data_ptr() for t in ab]
        cur = 0
        tail = cps.tail.data_ptr() if cps is not None else 0
        for n in range(N):
            if cps is not None and n % E == 0:
                cps.states[n // E].copy_(ab[cur])
            if n == s0:
                cps.tail[0].copy_(ab[cur])
            if s0 <= n < N - 1:
                # the last segment's steps write into the kept states directly
                go.step(tail + (n - s0) * sb, tail + (n - s0 + 1) * sb)
                go.record(tail + (n - s0 + 1) * sb, trp + n * rb)
                continue
            go.step(tail + (n - s0) * sb if n > s0 else abp[cur], abp[1 - cur])
            go.record(abp[1 - cur], trp + n * rb)
            cur = 1 - cur
        return traces, cps, ab[cur]

    def backward(self, traces, cps, aux, p_obs=None):
        """Adjoint sweep; returns (dJ/dq0 padded, dJ/d(rho, c) padded)."""
        N, E = self.nsteps, self.every
        obs = self.p_obs if p_obs is None else torch.as_tensor(p_obs, dtype=torch.float64, device=self.dev)
        if obs is None:
            raise ValueError("observed traces are needed")
        obs = obs.to(device=self.dev, dtype=torch.float64).contiguous()
        traces = traces.contiguous()
        if tuple(obs.shape) != (N, self.nrec) or tuple(traces.shape) != (N, self.nrec):
            raise ValueError("traces and observed traces must be (nsteps, nreceivers)")
        lam = torch.zeros((3, self.rows, self.ldq), dtype=torch.float64, device=self.dev)
        lam2 = torch.zeros_like(lam)
        gaux = torch.zeros((2, self.rows, self.ldq), dtype=torch.float64, device=self.dev)
        main = torch.cuda.current_stream(self.dev)
        go = _Launches(self, aux, main, self._cfl)
        sb, rb = lam.numel() * 8, self.nrec * 8
        trp, obp = traces.data_ptr(), obs.data_ptr()
        lp = [lam.data_ptr(), lam2.data_ptr()]
        gp = gaux.data_ptr()
        go.source(lp[0], trp + (N - 1) * rb, obp + (N - 1) * rb)

        def sweep(base, s0, s1, cur):
            # adjoint steps n = s1-1 .. s0 with q_n at base + (n - s0) * sb; lam alternates between lp[0], lp[1]
            for n in range(s1 - 1, s0 - 1, -1):
                q = base + (n - s0) * sb
                if n >= 1:
                    go.adjoint_misfit(q, lp[cur], lp[1 - cur], gp, trp + (n - 1) * rb, obp + (n - 1) * rb)
                else:
                    go.adjoint(q, lp[cur], lp[1 - cur], gp)
                cur = 1 - cur
            return cur

        if E == 1:
            # every state is stored: no recomputation
            cur = sweep(cps.data_ptr(), 0, N, 0)
            return (lam if cur == 0 else lam2), gaux
        # Segments are recomputed on a second stream into the other of two
        # buffers while the adjoint steps of the previous segment run: the
        # recomputation of segment k-1 does not depend on segment k's adjoint.
        side = self._side_stream()
        seg = [torch.empty((E,) + tuple(lam.shape), dtype=torch.float64, device=self.dev) for _ in range(2)]
        segp = [t.data_ptr() for t in seg]
        ready = [torch.cuda.Event() for _ in range(2)]
        free = [torch.cuda.Event() for _ in range(2)]
        cfl_side = torch.zeros(1, dtype=torch.float64, device=self.dev)
        go_side = _Launches(self, aux, side, cfl_side)
        nseg = (N + E - 1) // E

        def recompute(k, b):
            s0, s1 = k * E, min((k + 1) * E, N)
            with torch.cuda.stream(side):
                side.wait_event(free[b])
                seg[b][0].copy_(cps[k])
                for n in range(s0, s1 - 1):
                    go_side.step(segp[b] + (n - s0) * sb, segp[b] + (n - s0 + 1) * sb)
                ready[b].record(side)

        # the last segment's states come from the forward pass when it kept them
        tail = getattr(cps, 'tail', None)
        side.wait_stream(main)           # checkpoints and aux are complete
        for b in range(2):
            free[b].record(main)
        if tail is None:
            recompute(nseg - 1, (nseg - 1) % 2)
        cur = 0
        for k in range(nseg - 1, -1, -1):
            b = k % 2
            if k >= 1:
                recompute(k - 1, (k - 1) % 2)
            if k == nseg - 1 and tail is not None:
                base = tail.data_ptr()
            else:
                main.wait_event(ready[b])
                base = segp[b]
            cur = sweep(base, k * E, min((k + 1) * E, N), cur)
            free[b].record(main)
        main.wait_stream(side)
        return (lam if cur == 0 else lam2), gaux

    def objective(self, traces, p_obs=None):
        obs = self.p_obs if p_obs is None else torch.as_tensor(p_obs, dtype=torch.float64, device=self.dev)
        return 0.5 * ((traces - obs) ** 2).sum()

    def gradient(self, q0, K, rho, p_obs=None):
        """(J, dJ/dq0 (3, mx, my), dJ/dK (mx, my), dJ/drho (mx, my)) at the given state and medium."""
        aux = self.aux_from(K, rho)
        traces, cps, _ = self.forward(self.padded(q0), aux)
        J = self.objective(traces, p_obs)
        lam, gaux = self.backward(traces, cps, aux, p_obs)
        gq0 = self.interior(lam)
        grho_a, gc = self.interior(gaux)[0], self.interior(gaux)[1]
        K = torch.as_tensor(K, dtype=torch.float64, device=self.dev)
        rho = torch.as_tensor(rho, dtype=torch.float64, device=self.dev)
        c = torch.sqrt(K / rho)
        # c = sqrt(K / rho): dc/dK = c / (2K), dc/drho = -c / (2 rho)
        gK = gc * (c / (2.0 * K))
        grho = grho_a - gc * (c / (2.0 * rho))
        return J, gq0, gK, grho


class VCAcousticsMisfit(torch.autograd.Function):
    """J = VCAcousticsMisfit.apply(problem, q0, K, rho): the misfit of
    MisfitProblem `problem` (which holds p_obs), differentiable in q0, K, rho."""

    @staticmethod
    def forward(ctx, problem, q0, K, rho):
        aux = problem.aux_from(K.detach(), rho.detach())
        traces, cps, _ = problem.forward(problem.padded(q0.detach()), aux)
        ctx.problem = problem
        ctx.saved = (traces, cps, aux)
        ctx.save_for_backward(K, rho)
        return problem.objective(traces)

    @staticmethod
    def backward(ctx, gJ):
        problem = ctx.problem
        traces, cps, aux = ctx.saved
        K, rho = ctx.saved_tensors
        lam, gaux = problem.backward(traces, cps, aux)
        gq0 = problem.interior(lam)
        gaux_i = problem.interior(gaux)
        Kd = K.detach().to(problem.dev, torch.float64)
        rd = rho.detach().to(problem.dev, torch.float64)
        c = torch.sqrt(Kd / rd)
        gK = gaux_i[1] * (c / (2.0 * Kd))
        grho = gaux_i[0] - gaux_i[1] * (c / (2.0 * rd))
        ctx.saved = None
        return None, (gJ * gq0).to(K.dtype), (gJ * gK).to(K.dtype), (gJ * grho).to(rho.dtype)


def receivers_on_row(mx, my, y, nrec, lower=0.0, upper=1.0):
    """nrec receiver cells evenly spread in x on the row whose centre is nearest y."""
    dy = (upper - lower) / my
    j = int(round((y - lower) / dy - 0.5))
    j = min(max(j, 0), my - 1)
    return [((2 * k + 1) * mx // (2 * nrec), j) for k in range(nrec)]


def fixed_dt(dx, cmax, cfl=0.8):
    return cfl * dx / cmax
