"""Batched runs: many independent 2D grids advanced together on one GPU.

The reference runs one grid per process.  BatchedClawSolver2D advances a
list of instances that share the equations, method and boundary codes but
differ in size, coefficients and data; every step is one persistent
work-queue launch over all instances followed by the reference's
accept/reject and dt rule per instance (solver.py:651-687) on the device, and
`run(nsteps)` replays the captured hipGraph of the whole step sequence.
The per-instance results are bit-identical to stepping each grid alone.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .state import NG_STORE


class ClawBatchInstance(ctypes.Structure):
    _fields_ = [("mx", ctypes.c_int32), ("my", ctypes.c_int32), ("q_offset", ctypes.c_int64 * 2),
                ("aux_offset", ctypes.c_int64), ("ldq", ctypes.c_int64), ("plane", ctypes.c_int64),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double)]


def _bind():
    L = _lib.lib()
    vp = ctypes.c_void_p
    L.claw2d_batch_plan_create.argtypes = [vp, ctypes.POINTER(ClawBatchInstance), ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(_lib.ClawMethod), ctypes.POINTER(_lib.ClawRiemann),
                                           ctypes.POINTER(_lib.ClawBC), ctypes.c_int, ctypes.POINTER(vp)]
    L.claw2d_batch_plan_destroy.argtypes = [vp]
    L.claw2d_batch_work_items.argtypes = [vp]
    L.claw2d_batch_steps.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     vp, vp, vp, vp, vp, vp, ctypes.c_int, vp]
    L.claw2d_step_update.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     vp, vp, vp, vp, vp, vp, vp, vp]
    L.claw2d_batch_set_shape.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.claw2d_batch_autotune.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_int)]
    return L


class BatchedClawSolver2D(object):
    """Unsplit classic solver over a batch of grids.

    instances: list of dicts with q (num_eqn, mx, my), optional aux, dx, dy,
    dt_initial.  All grids use ng=NG_STORE stored ghost layers; the boundary
    codes are synthesised in the kernel (standard codes only)."""

    def __init__(self, riemann_solver, problem_data, instances, order=2, transverse=1, limiters=4,
                 bc_lower=(1, 1), bc_upper=(1, 1), aux_bc_lower=(1, 1), aux_bc_upper=(1, 1), cfl_max=1.0,
                 cfl_desired=0.9, dt_max=1e99, dtype=torch.float64, device=None):
        if not torch.cuda.is_available():
            raise _lib.ClawError("batched runs need the GPU")
        self.L = _bind()
        self.dev = torch.device(device) if device is not None else torch.device('cuda', 0)
        self.rp = riemann_solver
        self.cfl_max, self.cfl_desired, self.dt_max = float(cfl_max), float(cfl_desired), float(dt_max)
        ne, na = riemann_solver.num_eqn, riemann_solver.num_aux
        ng = NG_STORE
        self.n = len(instances)
        descs = (ClawBatchInstance * self.n)()
        qoff = 0
        aoff = 0
        self.layout = []
        for k, inst in enumerate(instances):
            mx, my = inst['q'].shape[1:]
            ldq, plane = mx + 2 * ng, (mx + 2 * ng) * (my + 2 * ng)
            d = descs[k]
            d.mx, d.my, d.ldq, d.plane = mx, my, ldq, plane
            d.q_offset[0] = qoff
            d.q_offset[1] = qoff + ne * plane
            d.aux_offset = aoff if na else -1
            d.dx, d.dy = float(inst['dx']), float(inst['dy'])
            self.layout.append((mx, my, ldq, plane, qoff, aoff))
            qoff += 2 * ne * plane
            aoff += na * plane
        self.qpool = torch.zeros(qoff, dtype=dtype, device=self.dev)
        self.auxpool = torch.zeros(max(aoff, 1), dtype=dtype, device=self.dev) if na else None
        for k, inst in enumerate(instances):
            self._put(k, inst['q'], inst.get('aux'))
        m = _lib.ClawMethod()
        m.order, m.transverse, m.split, m.fwave, m.num_waves = order, transverse, 0, 0, riemann_solver.num_waves
        lim = limiters if isinstance(limiters, (list, tuple)) else [limiters] * riemann_solver.num_waves
        for k, v in enumerate(lim):
            m.mthlim[k] = int(v)
        r = _lib.ClawRiemann()
        r.kind, r.efix = riemann_solver.kind, 1
        for k, v in enumerate(riemann_solver.params(problem_data)):
            r.p[k] = v
        bc = _lib.ClawBC()
        for s in range(2):
            bc.lower[s], bc.upper[s] = bc_lower[s], bc_upper[s]
            bc.aux_lower[s], bc.aux_upper[s] = aux_bc_lower[s], aux_bc_upper[s]
        self.method, self.rps, self.bc = m, r, bc
        self.precision = _lib.F32 if dtype == torch.float32 else _lib.F64
        self._ctx = _lib.Context(self.dev.index)
        self._plan = ctypes.c_void_p()
        _lib.check(self.L.claw2d_batch_plan_create(self._ctx.handle, descs, self.n, ng, m, r, bc, self.precision,
                                                   ctypes.byref(self._plan)), "claw2d_batch_plan_create")
        f64 = dict(dtype=torch.float64, device=self.dev)
        i32 = dict(dtype=torch.int32, device=self.dev)
        self.dt = torch.tensor([float(inst['dt_initial']) for inst in instances], **f64)
        self.t = torch.zeros(self.n, **f64)
        self.cflmax = torch.zeros(self.n, **f64)
        self.parity = torch.zeros(self.n, **i32)
        self.naccepted = torch.zeros(self.n, **i32)
        self.nrejected = torch.zeros(self.n, **i32)

    @property
    def work_items(self):
        return self.L.claw2d_batch_work_items(self._plan)

    def _view(self, k, which):
        mx, my, ldq, plane, qoff, _ = self.layout[k]
        ne = self.rp.num_eqn
        base = qoff + which * ne * plane
        buf = self.qpool[base:base + ne * plane].view(ne, my + 2 * NG_STORE, ldq)
        return buf

    def _put(self, k, q, aux):
        mx, my, ldq, plane, qoff, aoff = self.layout[k]
        ng = NG_STORE
        qt = torch.from_numpy(np.ascontiguousarray(np.asarray(q, dtype=np.float64).transpose(0, 2, 1)))
        for which in (0, 1):
            self._view(k, which)[:, ng:ng + my, ng:ng + mx] = qt.to(self.dev, self.qpool.dtype)
        if aux is not None and self.auxpool is not None:
            na = aux.shape[0]
            a = self.auxpool[aoff:aoff + na * plane].view(na, my + 2 * ng, ldq)
            at = torch.from_numpy(np.ascontiguousarray(np.asarray(aux, dtype=np.float64).transpose(0, 2, 1)))
            a[:, ng:ng + my, ng:ng + mx] = at.to(self.dev, self.auxpool.dtype)

    def set_shape(self, rows, waves_per_block):
        """Tile shape of the persistent launch: rows per band, wavefronts per block."""
        _lib.check(self.L.claw2d_batch_set_shape(self._plan, int(rows), int(waves_per_block)),
                   "claw2d_batch_set_shape")

    def autotune(self, stream=None):
        """Time the tile shapes on the current state (left unchanged) and keep
        the fastest; returns (rows, waves_per_block)."""
        st = stream if stream is not None else torch.cuda.current_stream()
        rows, wpb = ctypes.c_int(0), ctypes.c_int(0)
        _lib.check(self.L.claw2d_batch_autotune(self._plan, self.qpool.data_ptr(),
                                                self.auxpool.data_ptr() if self.auxpool is not None else None,
                                                self.dt.data_ptr(), self.parity.data_ptr(), st.cuda_stream,
                                                ctypes.byref(rows), ctypes.byref(wpb)), "claw2d_batch_autotune")
        return rows.value, wpb.value

    def run(self, nsteps, use_graph=True, stream=None):
        """Advance every instance nsteps steps (device-side dt and accept/reject)."""
        st = stream if stream is not None else torch.cuda.current_stream()
        rc = self.L.claw2d_batch_steps(self._plan, self.qpool.data_ptr(),
                                       self.auxpool.data_ptr() if self.auxpool is not None else None, int(nsteps),
                                       self.cfl_max, self.cfl_desired, self.dt_max, self.dt.data_ptr(),
                                       self.t.data_ptr(), self.cflmax.data_ptr(), self.parity.data_ptr(),
                                       self.naccepted.data_ptr(), self.nrejected.data_ptr(), 1 if use_graph else 0,
                                       st.cuda_stream)
        _lib.check(rc, "claw2d_batch_steps")

    def q(self, k):
        """Current state of instance k as a host array (num_eqn, mx, my)."""
        mx, my = self.layout[k][:2]
        par = int(self.parity[k].item())
        b = self._view(k, par)
        ng = NG_STORE
        return b[:, ng:ng + my, ng:ng + mx].permute(0, 2, 1).to('cpu', torch.float64).numpy().copy()

    def __del__(self):
        try:
            if self._plan:
                self.L.claw2d_batch_plan_destroy(self._plan)
                self._plan = ctypes.c_void_p()
        except Exception:
            pass


class OneByOne(object):
    """The comparison path: each instance stepped alone with the plain
    single-grid entry point (claw2d_classic_step) and the same device-side dt
    rule (claw2d_step_update), one instance after another.  The constructor
    uploads every instance (device buffers, per-instance dt/cfl/t slots);
    run(nsteps) only queues the steps, so the two phases can be timed apart.
    Buffers ping-pong on the host side, which is valid while no step is
    rejected (checked by results())."""

    def __init__(self, riemann_solver, problem_data, instances, order=2, transverse=1, limiters=4,
                 bc_lower=(1, 1), bc_upper=(1, 1), aux_bc_lower=(1, 1), aux_bc_upper=(1, 1), cfl_max=1.0,
                 cfl_desired=0.9, dt_max=1e99, dtype=torch.float64):
        self.L = _bind()
        dev = torch.device('cuda', 0)
        ng = NG_STORE
        self.ctx = _lib.Context(0)
        m = _lib.ClawMethod()
        m.order, m.transverse, m.split, m.fwave, m.num_waves = order, transverse, 0, 0, riemann_solver.num_waves
        lim = limiters if isinstance(limiters, (list, tuple)) else [limiters] * riemann_solver.num_waves
        for k, v in enumerate(lim):
            m.mthlim[k] = int(v)
        r = _lib.ClawRiemann()
        r.kind, r.efix = riemann_solver.kind, 1
        for k, v in enumerate(riemann_solver.params(problem_data)):
            r.p[k] = v
        bc = _lib.ClawBC()
        for s in range(2):
            bc.lower[s], bc.upper[s] = bc_lower[s], bc_upper[s]
            bc.aux_lower[s], bc.aux_upper[s] = aux_bc_lower[s], aux_bc_upper[s]
        self.m, self.r, self.bc = m, r, bc
        self.rule = (cfl_max, cfl_desired, dt_max)
        self.prec = _lib.F32 if dtype == torch.float32 else _lib.F64
        f64 = dict(dtype=torch.float64, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        self.items = []
        for inst in instances:
            q = np.asarray(inst['q'], dtype=np.float64)
            ne, mx, my = q.shape
            b0 = torch.zeros((ne, my + 2 * ng, mx + 2 * ng), dtype=dtype)
            b0[:, ng:ng + my, ng:ng + mx] = torch.from_numpy(np.ascontiguousarray(q.transpose(0, 2, 1)))
            b0 = b0.to(dev)
            b1 = b0.clone()
            aux = inst.get('aux')
            ad = None
            if aux is not None:
                ad = torch.zeros((aux.shape[0], my + 2 * ng, mx + 2 * ng), dtype=dtype)
                ad[:, ng:ng + my, ng:ng + mx] = torch.from_numpy(
                    np.ascontiguousarray(np.asarray(aux).transpose(0, 2, 1)))
                ad = ad.to(dev)
            g = _lib.ClawGrid(mx, my, ng, 0, mx + 2 * ng, (mx + 2 * ng) * (my + 2 * ng), float(inst['dx']),
                              float(inst['dy']))
            slots = dict(dt=torch.tensor([float(inst['dt_initial'])], **f64), cfl=torch.zeros(1, **f64),
                         t=torch.zeros(1, **f64), cmax=torch.zeros(1, **f64), par=torch.zeros(1, **i32),
                         nacc=torch.zeros(1, **i32), nrej=torch.zeros(1, **i32))
            self.items.append(dict(pair=(b0, b1), aux=ad, grid=g, mx=mx, my=my, **slots))
        self.nsteps = 0
        torch.cuda.synchronize()

    def run(self, nsteps):
        """Queue nsteps steps of every instance, one instance after another."""
        L, m, r, bc, prec = self.L, self.m, self.r, self.bc, self.prec
        cfl_max, cfl_desired, dt_max = self.rule
        st = torch.cuda.current_stream().cuda_stream
        for it in self.items:
            pair, ad, g = it['pair'], it['aux'], it['grid']
            aptr = ad.data_ptr() if ad is not None else None
            dtp, cflp = it['dt'].data_ptr(), it['cfl'].data_ptr()
            for s in range(self.nsteps, self.nsteps + nsteps):
                rc = L.claw2d_classic_step(self.ctx.handle, g, m, r, bc, prec, pair[s % 2].data_ptr(),
                                           pair[(s + 1) % 2].data_ptr(), aptr, dtp, cflp, st)
                _lib.check(rc, "claw2d_classic_step")
                rc = L.claw2d_step_update(1, cfl_max, cfl_desired, dt_max, dtp, cflp, it['t'].data_ptr(),
                                          it['cmax'].data_ptr(), it['par'].data_ptr(), it['nacc'].data_ptr(),
                                          it['nrej'].data_ptr(), st)
                _lib.check(rc, "claw2d_step_update")
        self.nsteps += nsteps

    def results(self):
        """Each instance's state (num_eqn, mx, my) on the host, after checking
        that no step was rejected."""
        torch.cuda.synchronize()
        ng = NG_STORE
        outs = []
        for it in self.items:
            if int(it['nrej'].item()) != 0:
                raise RuntimeError("a step was rejected; the one-by-one comparison path assumes none")
            buf, mx, my = it['pair'][self.nsteps % 2], it['mx'], it['my']
            outs.append(buf[:, ng:ng + my, ng:ng + mx].permute(0, 2, 1).to('cpu', torch.float64).numpy().copy())
        return outs


def run_one_by_one(riemann_solver, problem_data, instances, nsteps, keep=True, **kw):
    """OneByOne set up and run for nsteps; the instances' final states (keep)
    or None."""
    ob = OneByOne(riemann_solver, problem_data, instances, **kw)
    ob.run(nsteps)
    out = ob.results()
    return out if keep else None
