"""State: the solution on one patch, stored on the GPU.

API of the reference's State (src/pyclaw/state.py:12-301): State(geom, num_eqn,
num_aux), .q (num_eqn, mx, my), .aux, .t, .problem_data, .index_capa,
.keep_gauges, .gauge_data, get_q_global(), get_aux_global().

Storage: the authoritative copy of q lives on the device in the planar padded
layout the kernels use, qbuf[m, j+ng, i+ng] (x fastest), with NG_STORE ghost
layers.  `state.q` returns a host numpy mirror with the reference's shape and
indexing q[m, i, j]; it is refreshed from the device when read after a step,
and uploaded again before the next step if it may have been modified (any
read hands out a mutable array), so problem set-up and source-term code
written with numpy runs unchanged.  Code that wants to stay on the device
uses `state.q_device` (a torch view with the same (m, i, j) indexing).
"""
import copy

import numpy as np
import torch

from .geometry import Domain, Patch

NG_STORE = 3   # ghost layers kept in device storage (classic needs 2, SharpClaw 3)


def default_device():
    return torch.device('cuda', 0) if torch.cuda.is_available() else torch.device('cpu')


class _Field(object):
    """One cell-centred field (q or aux): device storage + host mirror."""

    def __init__(self, ncomp, mx, my, dtype, device):
        self.ncomp, self.mx, self.my = ncomp, mx, my
        ng = NG_STORE
        self.buf = torch.zeros((ncomp, my + 2 * ng, mx + 2 * ng), dtype=dtype, device=device)
        self.host = np.zeros((ncomp, mx, my), order='F')
        self.host_valid = True
        self.host_dirty = False

    def interior(self, buf=None):
        ng = NG_STORE
        b = self.buf if buf is None else buf
        return b[:, ng:ng + self.my, ng:ng + self.mx].permute(0, 2, 1)

    def padded(self, buf=None):
        """(ncomp, mx+2ng, my+2ng) view indexed [m, i, j] like the reference's qbc."""
        b = self.buf if buf is None else buf
        return b.permute(0, 2, 1)

    def download(self):
        if not self.host_valid:
            self.host = np.asfortranarray(self.interior().detach().to('cpu', torch.float64).numpy())
            self.host_valid = True

    def upload(self):
        if self.host_dirty:
            src = torch.from_numpy(np.ascontiguousarray(self.host, dtype=np.float64))
            self.interior().copy_(src.to(self.buf.device, self.buf.dtype))
            self.host_dirty = False

    def to(self, device=None, dtype=None):
        self.upload()
        self.buf = self.buf.to(device=device if device is not None else self.buf.device,
                               dtype=dtype if dtype is not None else self.buf.dtype)


class State(object):
    """Solution state on a patch (state.py:93-126)."""

    def __init__(self, geom, num_eqn, num_aux=0, dtype=torch.float64, device=None):
        if isinstance(geom, Patch):
            self.patch = geom
        elif isinstance(geom, Domain):
            self.patch = geom.patches[0]
        else:
            raise Exception("A PyClaw State object must be initialized with a PyClaw Patch object.")
        if self.patch.num_dim != 2:
            raise NotImplementedError("pyclaw_amd implements the 2D time step only")
        mx, my = self.patch.num_cells_global
        dev = torch.device(device) if device is not None else default_device()
        self.p = None
        self.F = None
        self.problem_data = {}
        self.t = 0.
        self.index_capa = -1
        self.keep_gauges = False
        self.gauge_data = []
        self._q = _Field(num_eqn, mx, my, dtype, dev)
        self._aux = _Field(num_aux, mx, my, dtype, dev) if num_aux > 0 else None

    # ---- geometry -----------------------------------------------------------
    @property
    def grid(self):
        return self.patch.grid

    @property
    def num_eqn(self):
        return self._q.ncomp

    @property
    def num_aux(self):
        return self._aux.ncomp if self._aux is not None else 0

    @property
    def dtype(self):
        return self._q.buf.dtype

    @property
    def device(self):
        return self._q.buf.device

    # ---- q / aux: host mirrors ------------------------------------------------
    @property
    def q(self):
        self._q.download()
        self._q.host_dirty = True
        return self._q.host

    @q.setter
    def q(self, value):
        value = np.asarray(value, dtype=np.float64)
        if value.shape != (self._q.ncomp, self._q.mx, self._q.my):
            raise ValueError("q must have shape %s" % ((self._q.ncomp, self._q.mx, self._q.my),))
        self._q.host = np.array(value, order='F', copy=True)
        self._q.host_valid = True
        self._q.host_dirty = True

    @property
    def aux(self):
        if self._aux is None:
            return None
        self._aux.download()
        self._aux.host_dirty = True
        return self._aux.host

    @aux.setter
    def aux(self, value):
        if self._aux is None:
            raise Exception("State was created with num_aux=0")
        value = np.asarray(value, dtype=np.float64)
        self._aux.host = np.array(value, order='F', copy=True)
        self._aux.host_valid = True
        self._aux.host_dirty = True

    # ---- device views -----------------------------------------------------------
    @property
    def q_device(self):
        """Torch view (num_eqn, mx, my) of the device copy; modifying it in
        place changes the state (the host mirror is refreshed on next read)."""
        self._q.upload()
        self._q.host_valid = False
        return self._q.interior()

    @property
    def aux_device(self):
        if self._aux is None:
            return None
        self._aux.upload()
        self._aux.host_valid = False
        return self._aux.interior()

    def sync_to_device(self):
        """Push host-side modifications (if any) to the device copy."""
        self._q.upload()
        if self._aux is not None:
            self._aux.upload()

    def _device_updated(self):
        self._q.host_valid = False
        self._q.host_dirty = False

    def to(self, device=None, dtype=None):
        self._q.to(device, dtype)
        if self._aux is not None:
            self._aux.to(device, dtype)
        return self

    # ---- reference helpers --------------------------------------------------------
    def is_valid(self):
        return True

    def set_cparam(self, fortran_module):
        """Kept for API compatibility; Riemann parameters are read from
        problem_data when the solver is set up (state.py:162-182)."""
        pass

    def set_num_ghost(self, num_ghost):
        if num_ghost > NG_STORE:
            raise ValueError("num_ghost %d exceeds the stored ghost layers (%d)" % (num_ghost, NG_STORE))

    def get_q_global(self):
        return self.q.copy()

    def get_aux_global(self):
        a = self.aux
        return None if a is None else a.copy()

    def sum_F(self, i):
        return np.sum(np.abs(self.F[i, ...]))

    def new_array(self, dof):
        if dof == 0:
            return None
        return np.empty([dof] + list(self.grid.num_cells), order='F')

    def __deepcopy__(self, memo):
        self.sync_to_device()
        result = State.__new__(State)
        for k, v in self.__dict__.items():
            if k in ('_q', '_aux', 'patch'):   # the patch (and its grid's open gauge files) is shared
                continue
            setattr(result, k, copy.deepcopy(v, memo))
        result.patch = self.patch
        for name in ('_q', '_aux'):
            f = getattr(self, name)
            if f is None:
                setattr(result, name, None)
                continue
            g = _Field.__new__(_Field)
            g.ncomp, g.mx, g.my = f.ncomp, f.mx, f.my
            g.buf = f.buf.clone()
            g.host = f.host.copy(order='F')
            g.host_valid, g.host_dirty = f.host_valid, False
            setattr(result, name, g)
        return result

    def __str__(self):
        out = "PyClaw State object\n"
        out += "Patch dimensions: %s\n" % str(self.patch.num_cells_global)
        out += "Time  t=%s\n" % self.t
        out += "Number of conserved quantities: %s\n" % self.num_eqn
        if self._aux is not None:
            out += "Number of auxiliary fields: %s\n" % self.num_aux
        if self.problem_data:
            out += "problem_data: " + str(self.problem_data)
        return out
