"""ctypes binding of libclawhip.so (the C ABI declared in include/clawhip.h).

Loads the in-tree library built by pyclaw_amd.build.  There is no fallback:
every entry point that needs the HIP library raises if it is missing.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libclawhip.so")

CLAW_OK, CLAW_EINVAL, CLAW_EDEVICE, CLAW_ENOMEM, CLAW_ETOOLARGE = 0, 1, 2, 3, 4
NOALLOC_SCRATCH_BYTES = 32 << 20   # CLAW_NOALLOC_SCRATCH_BYTES (claw2d_classic_step_noalloc)
RP_ACOUSTICS, RP_VC_ACOUSTICS, RP_EULER, RP_SHALLOW = 0, 1, 2, 3
BC_STORED, BC_EXTRAP, BC_PERIODIC, BC_WALL = 0, 1, 2, 3
F64, F32, F64_FMA = 0, 1, 2
STEP_ACCUMULATE_CFL = 1
SM_DQ, SM_EULER, SM_SSP104_A, SM_SSP104_MID, SM_SSP104_END, SM_SSP33_2, SM_SSP33_3 = range(7)

_ERRORS = {
    CLAW_EINVAL: "invalid argument or unsupported option",
    CLAW_EDEVICE: "HIP runtime error",
    CLAW_ENOMEM: "device allocation failed",
    CLAW_ETOOLARGE: "size beyond the documented maximum",
}


class ClawGrid(ctypes.Structure):
    _fields_ = [("mx", ctypes.c_int32), ("my", ctypes.c_int32), ("ng", ctypes.c_int32),
                ("pad_", ctypes.c_int32), ("ldq", ctypes.c_int64), ("plane", ctypes.c_int64),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double)]


class ClawMethod(ctypes.Structure):
    _fields_ = [("order", ctypes.c_int32), ("transverse", ctypes.c_int32), ("split", ctypes.c_int32),
                ("fwave", ctypes.c_int32), ("num_waves", ctypes.c_int32), ("mthlim", ctypes.c_int32 * 8)]


class ClawRiemann(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("efix", ctypes.c_int32), ("p", ctypes.c_double * 8)]


class ClawBC(ctypes.Structure):
    _fields_ = [("lower", ctypes.c_int32 * 2), ("upper", ctypes.c_int32 * 2),
                ("aux_lower", ctypes.c_int32 * 2), ("aux_upper", ctypes.c_int32 * 2)]


class ClawLoopRule(ctypes.Structure):
    _fields_ = [("cfl_max", ctypes.c_double), ("cfl_desired", ctypes.c_double), ("dt_max", ctypes.c_double),
                ("tend", ctypes.c_double), ("tstart", ctypes.c_double), ("max_steps", ctypes.c_int32),
                ("ncfl", ctypes.c_int32), ("dt_variable", ctypes.c_int32), ("pad_", ctypes.c_int32)]


class ClawLoopState(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("t", ctypes.c_double), ("cflmax", ctypes.c_double),
                ("dtmin", ctypes.c_double), ("dtmax", ctypes.c_double), ("cfl_last", ctypes.c_double),
                ("naccepted", ctypes.c_int32), ("niter", ctypes.c_int32), ("done", ctypes.c_int32),
                ("failed", ctypes.c_int32)]


class ClawLoopGauges(ctypes.Structure):
    _fields_ = [("offsets", ctypes.c_void_p), ("vals", ctypes.c_void_p), ("times", ctypes.c_void_p),
                ("n", ctypes.c_int32), ("base", ctypes.c_int32), ("cap", ctypes.c_int32), ("pad_", ctypes.c_int32)]


class ClawError(RuntimeError):
    pass


_lib = None


def available():
    return os.path.exists(LIB_PATH)


def lib():
    """The loaded HIP library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ClawError("libclawhip.so is not built (run pyclaw_amd.build.build() or "
                            "__graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.claw_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        L.claw_ctx_destroy.argtypes = [vp]
        L.claw2d_classic_step.argtypes = [vp, ctypes.POINTER(ClawGrid), ctypes.POINTER(ClawMethod),
                                          ctypes.POINTER(ClawRiemann), ctypes.POINTER(ClawBC), ctypes.c_int,
                                          vp, vp, vp, vp, vp, vp]
        L.claw2d_classic_step_ex.argtypes = L.claw2d_classic_step.argtypes[:11] + [ctypes.c_int, vp]
        L.claw2d_classic_step_noalloc.argtypes = L.claw2d_classic_step.argtypes
        L.claw2d_classic_launches.argtypes = [ctypes.POINTER(ClawMethod)]
        for name, argtypes in _optional_symbols().items():
            fn = getattr(L, name, None)
            if fn is not None:
                fn.argtypes = argtypes
        if hasattr(L, "claw2d_vc_adjoint_workspace_bytes"):
            L.claw2d_vc_adjoint_workspace_bytes.restype = ctypes.c_longlong
        _lib = L
    return _lib


def _optional_symbols():
    vp = ctypes.c_void_p
    return {
        "claw_ctx_set_wide_offsets": [vp, ctypes.c_int],
        "claw2d_loop_advance": [ctypes.POINTER(ClawLoopRule), vp, vp, vp, vp, vp, vp, ctypes.c_longlong, vp],
        "claw2d_sharpclaw_loop": [vp, ctypes.POINTER(ClawGrid), ctypes.POINTER(ClawRiemann), ctypes.POINTER(ClawBC),
                                  ctypes.c_int, vp, vp, vp, vp, vp, vp, ctypes.POINTER(ClawLoopRule), vp, vp,
                                  ctypes.c_longlong, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ClawLoopGauges), vp],
        "claw2d_classic_loop": [vp, ctypes.POINTER(ClawGrid), ctypes.POINTER(ClawMethod), ctypes.POINTER(ClawRiemann),
                                ctypes.POINTER(ClawBC), ctypes.c_int, vp, vp, vp, ctypes.POINTER(ClawLoopRule), vp, vp,
                                ctypes.c_longlong, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ClawLoopGauges), vp],
        "claw2d_euler_radial_step": [ctypes.POINTER(ClawGrid), ctypes.c_int, ctypes.c_int, vp, vp, ctypes.c_double,
                                     ctypes.c_double, vp],
        "claw2d_euler_radial_dq": [ctypes.POINTER(ClawGrid), ctypes.c_int, ctypes.c_int, vp, vp, vp,
                                   ctypes.c_double, ctypes.c_double, vp],
        "claw2d_sharpclaw_dq": [vp, ctypes.POINTER(ClawGrid), ctypes.POINTER(ClawRiemann), ctypes.POINTER(ClawBC),
                                ctypes.c_int, vp, vp, vp, vp, vp, vp],
        "claw2d_sharpclaw_stage": [vp, ctypes.POINTER(ClawGrid), ctypes.POINTER(ClawRiemann),
                                   ctypes.POINTER(ClawBC), ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, vp, vp,
                                   vp, vp],
        "claw2d_sharpclaw_ssp104_step": [vp, ctypes.POINTER(ClawGrid), ctypes.POINTER(ClawRiemann),
                                         ctypes.POINTER(ClawBC), ctypes.c_int, vp, vp, vp, vp, vp, vp, vp, vp,
                                         vp],
        "claw2d_classic_autotune": [vp, ctypes.POINTER(ClawGrid), ctypes.POINTER(ClawMethod),
                                    ctypes.POINTER(ClawRiemann), ctypes.POINTER(ClawBC), ctypes.c_int, vp, vp, vp, vp,
                                    vp, vp, ctypes.POINTER(ctypes.c_int)],
        "claw2d_sharpclaw_autotune": [vp, ctypes.POINTER(ClawGrid), ctypes.POINTER(ClawRiemann),
                                      ctypes.POINTER(ClawBC), ctypes.c_int, vp, vp, vp, vp, vp, vp,
                                      ctypes.POINTER(ctypes.c_int)],
        "claw2d_sharpclaw_set_launch": [vp, ctypes.POINTER(ClawGrid), ctypes.POINTER(ClawRiemann), ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int],
        "claw2d_sharpclaw_set_launch_ex": [vp, ctypes.POINTER(ClawGrid), ctypes.POINTER(ClawRiemann), ctypes.c_int,
                                           ctypes.c_int, ctypes.c_int, ctypes.c_int],
        "claw2d_sharpclaw_get_launch": [vp, ctypes.POINTER(ClawGrid), ctypes.POINTER(ClawRiemann), ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_int)],
        "claw2d_classic_set_launch_ex": [vp, ctypes.POINTER(ClawGrid), ctypes.POINTER(ClawMethod),
                                         ctypes.POINTER(ClawRiemann), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int],
        "claw2d_classic_get_launch": [vp, ctypes.POINTER(ClawGrid), ctypes.POINTER(ClawMethod),
                                      ctypes.POINTER(ClawRiemann), ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                      ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
        "claw2d_classic_set_launch": [vp, ctypes.POINTER(ClawGrid), ctypes.POINTER(ClawMethod),
                                      ctypes.POINTER(ClawRiemann), ctypes.c_int, ctypes.c_int, ctypes.c_int],
        "claw2d_vc_adjoint_workspace_bytes": [ctypes.POINTER(ClawGrid)],
        "claw2d_vc_adjoint_step": [ctypes.POINTER(ClawGrid), ctypes.POINTER(ClawMethod), ctypes.POINTER(ClawBC),
                                   vp, vp, vp, vp, vp, ctypes.c_double, vp, ctypes.c_longlong, vp],
        "claw2d_vc_adjoint_step_misfit": [ctypes.POINTER(ClawGrid), ctypes.POINTER(ClawMethod),
                                          ctypes.POINTER(ClawBC), vp, vp, vp, vp, vp, ctypes.c_double, vp,
                                          ctypes.c_longlong, vp, ctypes.c_int, vp, vp, vp],
        "claw2d_gather": [vp, vp, ctypes.c_int, vp, vp],
        "claw2d_misfit_source": [vp, vp, ctypes.c_int, vp, vp, vp],
    }


def check(rc, what):
    if rc != CLAW_OK:
        raise ClawError("%s failed: %s (code %d)" % (what, _ERRORS.get(rc, "unknown error"), rc))


class Context(object):
    """Owns a libclawhip context (scratch buffers) for one device."""

    def __init__(self, device=0):
        self._ctx = ctypes.c_void_p()
        check(lib().claw_ctx_create(int(device), ctypes.byref(self._ctx)), "claw_ctx_create")
        self.device = device

    @property
    def handle(self):
        return self._ctx

    def __del__(self):
        try:
            if self._ctx and _lib is not None:
                _lib.claw_ctx_destroy(self._ctx)
                self._ctx = ctypes.c_void_p()
        except Exception:
            pass
