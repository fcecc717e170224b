"""CPU oracle for the 2D classic and SharpClaw time step.

TEST INFRASTRUCTURE ONLY.  tests/, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py may import this package, always as the checker; the product
(pyclaw_amd) never imports it.

Two libraries are wrapped:
  * liborc.so          -- this repository's C restatement (claw_oracle.c),
                          built by oracle/Makefile.
  * _ref/libref_*.so   -- the reference's own Fortran kernels compiled from
                          /root/reference by oracle/build_ref.sh (present only
                          where the reference was available at build time).

Array convention (the reference's): padded arrays qbc of shape
(meqn, mx+2*ng, my+2*ng) in Fortran order, i.e. qbc[m, i+ng, j+ng].

Riemann-solver ids: 0 acoustics, 1 vc_acoustics, 2 euler (5-wave, efix),
3 shallow-water Roe with entropy fix.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
RP_ACOUSTICS, RP_VC_ACOUSTICS, RP_EULER, RP_SHALLOW = 0, 1, 2, 3

_lib = None
_ref_classic = None
_ref_sharp = None
_threads = 1


def set_threads(n):
    """Threads for step2 / step2ds / sharp_dq (OpenMP over row and column bands;
    results are bit-identical to one thread).  Returns the previous value."""
    global _threads
    old, _threads = _threads, max(1, int(n))
    return old


def host_threads(cap=16):
    """The host's usable cores (the process's CPU affinity), at most cap."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(cap, n))


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def lib():
    """Load liborc.so (the C restatement)."""
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liborc.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle library not built: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(path)
        L.orc_set_cparam.argtypes = [ctypes.c_double] * 7 + [ctypes.c_int]
        L.orc_step2.restype = ctypes.c_double
        L.orc_step2.argtypes = [ctypes.c_int] * 8 + [_dp, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                                     ctypes.c_double, _ip, _ip, ctypes.c_int]
        L.orc_step2ds.restype = ctypes.c_double
        L.orc_step2ds.argtypes = [ctypes.c_int] * 8 + [_dp, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                                       ctypes.c_double, _ip, _ip, ctypes.c_int, ctypes.c_int]
        L.orc_step2_mt.restype = ctypes.c_double
        L.orc_step2_mt.argtypes = L.orc_step2.argtypes + [ctypes.c_int]
        L.orc_step2ds_mt.restype = ctypes.c_double
        L.orc_step2ds_mt.argtypes = L.orc_step2ds.argtypes + [ctypes.c_int]
        L.orc_sharp_dq_mt.restype = ctypes.c_double
        L.orc_sharp_dq_mt.argtypes = [ctypes.c_int] * 7 + [_dp, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                                           ctypes.c_double, ctypes.c_int]
        L.orc_fill_bc.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _ip,
                                  ctypes.c_int]
        L.orc_sharp_dq.restype = ctypes.c_double
        L.orc_sharp_dq.argtypes = [ctypes.c_int] * 7 + [_dp, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                                        ctypes.c_double]
        L.orc_philim.restype = ctypes.c_double
        L.orc_philim.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_rpn2.argtypes = [ctypes.c_int] * 7 + [_dp] * 8 + [ctypes.c_int]
        L.orc_rpt2.argtypes = [ctypes.c_int] * 7 + [_dp] * 5 + [ctypes.c_int] + [_dp] * 3 + [ctypes.c_int]
        L.orc_weno5.argtypes = [_dp, _dp, _dp, ctypes.c_int, ctypes.c_int]
        L.orc_rp_pointer.restype = ctypes.c_void_p
        L.orc_rp_pointer.argtypes = [ctypes.c_int, ctypes.c_int]
        _lib = L
    return _lib


def set_params(rp, problem_data):
    """Set the Riemann-solver parameters (the reference's /cparam/ block)."""
    pd = problem_data
    lib().orc_set_cparam(pd.get("rho", 1.0), pd.get("bulk", 1.0), pd.get("cc", 1.0), pd.get("zz", 1.0),
                         pd.get("gamma", 1.4), pd.get("gamma1", 0.4), pd.get("grav", 1.0),
                         int(pd.get("efix", 1)))


def _f64F(a):
    return np.asfortranarray(a, dtype=np.float64)


def method_array(order=2, transverse=1, dimensional_split=False, dt_variable=True, num_aux=0):
    """The reference's method(1:7) (classic/solver.py:178-198)."""
    m = np.zeros(7, dtype=np.int32)
    m[0] = 1 if dt_variable else 0
    m[1] = order
    m[2] = -1 if dimensional_split else transverse
    m[5] = 0
    m[6] = num_aux
    return m


def mthlim_array(limiters, num_waves):
    """classic/solver.py:167-176 (set_mthlim)."""
    lim = limiters if isinstance(limiters, (list, tuple)) else [limiters]
    if len(lim) == 1:
        lim = list(lim) * num_waves
    if len(lim) != num_waves:
        raise ValueError("Length of solver.limiters is not equal to 1 or to solver.num_waves")
    return np.asarray(lim, dtype=np.int32)


def fill_bc(qbc, ng, bc_lower, bc_upper, negate=True):
    """In-place ghost fill of a padded array (extrap=1, periodic=2, wall=3)."""
    assert qbc.flags["F_CONTIGUOUS"] and qbc.dtype == np.float64
    ne, nxp, nyp = qbc.shape
    lo = np.asarray(bc_lower, dtype=np.int32)
    hi = np.asarray(bc_upper, dtype=np.int32)
    lib().orc_fill_bc(_ptr(qbc), ne, nxp - 2 * ng, nyp - 2 * ng, ng, lo.ctypes.data_as(_ip),
                      hi.ctypes.data_as(_ip), 1 if negate else 0)


def _aux_or_dummy(auxbc, nxp, nyp):
    if auxbc is None or auxbc.shape[0] == 0:
        return np.zeros((1, nxp, nyp), order="F"), 0
    return _f64F(auxbc), auxbc.shape[0]


def step2(rp, qbc, auxbc, dx, dy, dt, method, mthlim, num_waves, use_fwave=False):
    """One unsplit classic step (step2.f90); returns (qnew, cfl)."""
    qold = _f64F(qbc)
    qnew = qold.copy(order="F")
    ne, nxp, nyp = qold.shape
    ng = 2
    mx, my = nxp - 2 * ng, nyp - 2 * ng
    aux, na = _aux_or_dummy(auxbc, nxp, nyp)
    meth = np.asarray(method, dtype=np.int32).copy()
    meth[6] = na
    lim = np.asarray(mthlim, dtype=np.int32)
    cfl = lib().orc_step2_mt(rp, max(mx, my), ne, num_waves, na, ng, mx, my, _ptr(qold), _ptr(qnew),
                             _ptr(aux), dx, dy, dt, meth.ctypes.data_as(_ip), lim.ctypes.data_as(_ip),
                             1 if use_fwave else 0, _threads)
    return qnew, cfl


def step2ds(rp, qold, qnew, auxbc, dx, dy, dt, method, mthlim, num_waves, ids, use_fwave=False):
    """One step2ds call (ids 1: x sweeps, 2: y sweeps); qnew updated in place; returns cfl."""
    ne, nxp, nyp = qold.shape
    ng = 2
    mx, my = nxp - 2 * ng, nyp - 2 * ng
    aux, na = _aux_or_dummy(auxbc, nxp, nyp)
    meth = np.asarray(method, dtype=np.int32).copy()
    meth[6] = na
    lim = np.asarray(mthlim, dtype=np.int32)
    return lib().orc_step2ds_mt(rp, max(mx, my), ne, num_waves, na, ng, mx, my, _ptr(qold), _ptr(qnew),
                                _ptr(aux), dx, dy, dt, meth.ctypes.data_as(_ip), lim.ctypes.data_as(_ip),
                                ids, 1 if use_fwave else 0, _threads)


def sharp_dq(rp, qbc, auxbc, dx, dy, dt, num_waves):
    """SharpClaw 2D dq over the padded patch (ng=3); returns (dq interior (meqn,mx,my), cfl)."""
    q = _f64F(qbc)
    ne, nxp, nyp = q.shape
    ng = 3
    mx, my = nxp - 2 * ng, nyp - 2 * ng
    aux, na = _aux_or_dummy(auxbc, nxp, nyp)
    dq = np.zeros_like(q, order="F")
    cfl = lib().orc_sharp_dq_mt(rp, ne, num_waves, na, ng, mx, my, _ptr(q), _ptr(aux), _ptr(dq), dx, dy, dt,
                                _threads)
    return dq[:, ng:-ng, ng:-ng].copy(order="F"), cfl


def philim(a, b, meth):
    return lib().orc_philim(a, b, meth)


def weno5(q):
    """weno.f90:5-83 on a slice q (meqn, n): returns (ql, qr) with left/right
    edge values for cells 2..n-3 (0-based); other entries are zero."""
    q = np.asfortranarray(q, dtype=np.float64)
    ne, n = q.shape
    ql = np.zeros_like(q, order='F')
    qr = np.zeros_like(q, order='F')
    lib().orc_weno5(_ptr(q), _ptr(ql), _ptr(qr), ne, n)
    return ql, qr


def rpn2(rp, ixy, qlo, qhi, auxlo=None, auxhi=None, num_waves=None):
    """Riemann solve at single interfaces: qlo/qhi (meqn, k) are the states left
    and right of k interfaces.  Returns (wave (meqn, mw, k), s (mw, k), amdq, apdq)."""
    qlo = np.asarray(qlo, dtype=np.float64); qhi = np.asarray(qhi, dtype=np.float64)
    ne, k = qlo.shape
    nw = num_waves
    mbc = 1
    # slice of length 2k: cells (2j, 2j+1) hold (lo, hi) of interface j; rpn2 computes
    # interfaces between consecutive cells, the odd ones are the wanted ones.
    n = 2 * k
    q = np.zeros((ne, n + 2 * mbc), order='F')
    q[:, mbc:mbc + n:2] = qlo
    q[:, mbc + 1:mbc + n:2] = qhi
    na = 2
    a = np.ones((na, n + 2 * mbc), order='F')
    if auxlo is not None:
        a[:, mbc:mbc + n:2] = auxlo
        a[:, mbc + 1:mbc + n:2] = auxhi
    wave = np.zeros((ne, nw, n + 2 * mbc), order='F')
    s = np.zeros((nw, n + 2 * mbc), order='F')
    amdq = np.zeros((ne, n + 2 * mbc), order='F')
    apdq = np.zeros((ne, n + 2 * mbc), order='F')
    lib().orc_rpn2(rp, ixy, n, ne, nw, mbc, n, _ptr(q), _ptr(q), _ptr(a), _ptr(a), _ptr(wave), _ptr(s),
                   _ptr(amdq), _ptr(apdq), na)
    # interface between slice cells c-1 and c lives at Fortran index c (array index c-1+mbc)
    idx = np.arange(k) * 2 + 1 + mbc
    return wave[:, :, idx], s[:, idx], amdq[:, idx], apdq[:, idx]


def rpt2(rp, ixy, qlo, qhi, asdq, imp=1, aux1=None, aux2=None, aux3=None, num_waves=None):
    """Transverse solve at single interfaces, called as flux2.f90:184-185 calls
    it: qlo/qhi (meqn, k) are the states left and right of k interfaces,
    asdq (meqn, k) the fluctuation being split, imp 1 (amdq, goes to the left
    cell) or 2 (apdq, right cell).  aux1/aux2/aux3 are (naux, 2, k): for each
    interface the aux of its (left, right) cell in the row below, the row
    itself and the row above.  Returns (bmasdq, bpasdq), each (meqn, k)."""
    qlo = np.asarray(qlo, dtype=np.float64); qhi = np.asarray(qhi, dtype=np.float64)
    ne, k = qlo.shape
    nw = num_waves
    mbc = 1
    n = 2 * k
    q = np.zeros((ne, n + 2 * mbc), order='F')
    q[:, mbc:mbc + n:2] = qlo
    q[:, mbc + 1:mbc + n:2] = qhi
    idx = np.arange(k) * 2 + 1 + mbc
    a = np.zeros((ne, n + 2 * mbc), order='F')
    a[:, idx] = asdq
    auxs = []
    for ax in (aux1, aux2, aux3):
        if ax is None:
            arr = np.ones((2, n + 2 * mbc), order='F')
        else:
            ax = np.asarray(ax, dtype=np.float64)
            arr = np.ones((ax.shape[0], n + 2 * mbc), order='F')
            arr[:, mbc:mbc + n:2] = ax[:, 0]
            arr[:, mbc + 1:mbc + n:2] = ax[:, 1]
        auxs.append(arr)
    na = auxs[0].shape[0]
    bm = np.zeros((ne, n + 2 * mbc), order='F')
    bp = np.zeros((ne, n + 2 * mbc), order='F')
    lib().orc_rpt2(rp, ixy, n, ne, nw, mbc, n, _ptr(q), _ptr(q), _ptr(auxs[0]), _ptr(auxs[1]), _ptr(auxs[2]),
                   imp, _ptr(a), _ptr(bm), _ptr(bp), na)
    return bm[:, idx], bp[:, idx]


# ---------------------------------------------------------------------------
# The reference itself, compiled from /root/reference (oracle/build_ref.sh)
# ---------------------------------------------------------------------------
def ref_available():
    return (os.path.exists(os.path.join(HERE, "_ref", "libref_classic.so"))
            and os.path.exists(os.path.join(HERE, "_ref", "libref_sharpclaw.so")))


def _ref_libs():
    global _ref_classic, _ref_sharp
    if _ref_classic is None:
        _ref_classic = ctypes.CDLL(os.path.join(HERE, "_ref", "libref_classic.so"), mode=ctypes.RTLD_LOCAL)
        _ref_sharp = ctypes.CDLL(os.path.join(HERE, "_ref", "libref_sharpclaw.so"), mode=ctypes.RTLD_LOCAL)
    return _ref_classic, _ref_sharp


def _byref_int(v):
    return ctypes.byref(ctypes.c_int(int(v)))


def _byref_dbl(v):
    return ctypes.byref(ctypes.c_double(float(v)))


def _ref_rp(rp, use_ref_euler):
    """rpn2/rpt2 procedure pointers for the reference sweeps."""
    if rp == RP_EULER and use_ref_euler:
        C, _ = _ref_libs()
        return ctypes.cast(C.rpn2_, ctypes.c_void_p).value, ctypes.cast(C.rpt2_, ctypes.c_void_p).value
    L = lib()
    return L.orc_rp_pointer(rp, 0), L.orc_rp_pointer(rp, 1)


def ref_set_euler_params(gamma, gamma1):
    C, _ = _ref_libs()
    cp = (ctypes.c_double * 2).in_dll(C, "cparam_")
    cp[0] = gamma
    cp[1] = gamma1


def ref_step2(rp, qbc, auxbc, dx, dy, dt, method, mthlim, num_waves, use_fwave=False,
              ids=0, qnew=None, use_ref_euler=True):
    """Reference step2 (ids=0) or step2ds (ids=1,2) compiled from its own sources.

    The Euler solver is the reference's own; the other Riemann solvers (external
    to the reference) are this repository's C restatement, passed in as the
    rpn2/rpt2 procedure arguments exactly as clawpack.riemann's would be.
    Returns (qnew, cfl)."""
    C, _ = _ref_libs()
    qold = _f64F(qbc)
    if qnew is None:
        qnew = qold.copy(order="F")
    ne, nxp, nyp = qold.shape
    ng = 2
    mx, my = nxp - 2 * ng, nyp - 2 * ng
    maxm = max(mx, my)
    aux, na = _aux_or_dummy(auxbc, nxp, nyp)
    meth = np.asarray(method, dtype=np.int32).copy()
    meth[6] = na
    lim = np.asarray(mthlim, dtype=np.int32)
    nm = maxm + 2 * ng
    naw = max(na, 1)
    qadd = np.zeros((ne, nm), order="F"); fadd = np.zeros((ne, nm), order="F")
    gadd = np.zeros((ne, 2, nm), order="F"); q1d = np.zeros((ne, nm), order="F")
    dtdx1d = np.zeros(nm); dtdy1d = np.zeros(nm)
    aux1 = np.zeros((naw, nm), order="F"); aux2 = np.zeros((naw, nm), order="F")
    aux3 = np.zeros((naw, nm), order="F")
    mwork = nm * (5 * ne + num_waves + ne * num_waves)
    work = np.zeros(mwork)
    cfl = ctypes.c_double(0.0)
    fw = ctypes.c_int(1 if use_fwave else 0)
    rpn, rpt = _ref_rp(rp, use_ref_euler)
    args = [_byref_int(maxm), _byref_int(ne), _byref_int(num_waves), _byref_int(na), _byref_int(ng),
            _byref_int(mx), _byref_int(my), _ptr(qold), _ptr(qnew), _ptr(aux), _byref_dbl(dx),
            _byref_dbl(dy), _byref_dbl(dt), meth.ctypes.data_as(_ip), lim.ctypes.data_as(_ip),
            ctypes.byref(cfl), _ptr(qadd), _ptr(fadd), _ptr(gadd), _ptr(q1d), _ptr(dtdx1d), _ptr(dtdy1d),
            _ptr(aux1), _ptr(aux2), _ptr(aux3), _ptr(work), _byref_int(mwork)]
    if ids == 0:
        fn = C.step2_
        args += [ctypes.byref(fw), ctypes.c_void_p(rpn), ctypes.c_void_p(rpt)]
    else:
        fn = C.step2ds_
        args += [_byref_int(ids), ctypes.byref(fw), ctypes.c_void_p(rpn), ctypes.c_void_p(rpt)]
    fn.restype = None
    fn(*args)
    return qnew, cfl.value


def ref_sharp_dq(rp, qbc, auxbc, dx, dy, dt, num_waves, use_ref_euler=True):
    """Reference SharpClaw flux2 (dq, cfl) compiled from its own sources."""
    C, S = _ref_libs()
    q = _f64F(qbc)
    ne, nxp, nyp = q.shape
    ng = 3
    mx, my = nxp - 2 * ng, nyp - 2 * ng
    maxnx = max(mx, my)
    aux, na = _aux_or_dummy(auxbc, nxp, nyp)
    S.ref_sharp_setup.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double] * 2
    S.ref_sharp_setup(ne, num_waves, ng, maxnx, dx, dy)
    dq = np.zeros_like(q, order="F")
    nm = maxnx + 2 * ng
    q1d = np.zeros((ne, nm), order="F")
    dq1d = np.zeros((ne, nm), order="F")
    cfl = ctypes.c_double(0.0)
    rpn, _ = _ref_rp(rp, use_ref_euler)
    S.flux2_.restype = None
    S.flux2_(_ptr(q), _ptr(dq), _ptr(q1d), _ptr(dq1d), _ptr(aux), _byref_dbl(dt), ctypes.byref(cfl),
             _byref_dbl(0.0), _byref_int(na), _byref_int(ne), _byref_int(ng), _byref_int(maxnx),
             _byref_int(mx), _byref_int(my), ctypes.c_void_p(rpn))
    return dq[:, ng:-ng, ng:-ng].copy(order="F"), cfl.value
