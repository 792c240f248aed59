"""ctypes binding of ``libmdsf.so`` (declarations in ``include/mdsf.h``).

The library is the only compute path: importing this module fails loudly when
it has not been built (``python -m mdsf.build``); there is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MDSF_LIB", os.path.join(HERE, "libmdsf.so"))

OK = 0
ERR_ARG = 1
ERR_HIP = 2
ERR_RANGE = 3
ERR_TYPE = 4
ERR_UNSUPPORTED = 5
F32 = 0
F64 = 1
FFT_FP64 = 1
AUTOTUNE = 2

c_int32 = ctypes.c_int32
c_int64 = ctypes.c_int64
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p
c_char_p = ctypes.c_char_p


class TypeDesc(ctypes.Structure):
    """``mdsf_type_desc``."""
    _fields_ = [("amp", c_double), ("two_sig2", c_double), ("half", c_int32 * 3),
                ("count", c_int32)]


class GridDesc(ctypes.Structure):
    """``mdsf_grid_desc``."""
    _fields_ = [("n", c_int32 * 3), ("nborder", c_int32), ("dr", c_double * 3),
                ("ucell", c_double * 9), ("ntypes", c_int32),
                ("types", ctypes.POINTER(TypeDesc)), ("natoms", c_int32), ("flags", c_int32)]


# name -> (restype, argtypes); every symbol include/mdsf.h declares
SIGNATURES = {
    "mdsf_plan_create": (c_int32, [ctypes.POINTER(GridDesc), ctypes.POINTER(c_void_p)]),
    "mdsf_plan_destroy": (c_int32, [c_void_p]),
    "mdsf_last_error": (c_char_p, []),
    "mdsf_version": (c_int32, []),
    "mdsf_sf_reset": (c_int32, [c_void_p, c_void_p]),
    "mdsf_check_errors": (c_int32, [c_void_p, c_void_p]),
    "mdsf_accumulate_frames": (c_int32, [c_void_p, c_void_p, c_int32, c_int64, c_void_p, c_void_p,
                                         ctypes.POINTER(c_double), c_int32, c_int64, c_int64,
                                         c_void_p, c_void_p, c_void_p]),
    "mdsf_rescale_wrap": (c_int32, [c_void_p, c_void_p, c_int32, c_int64, c_void_p,
                                    ctypes.POINTER(c_double), c_int32, c_int64, c_int64,
                                    c_void_p]),
    "mdsf_rescale": (c_int32, [c_void_p, c_int32, c_int64, c_int64, c_void_p, c_int32,
                               c_void_p]),
    "mdsf_fold_padded": (c_int32, [c_void_p, ctypes.POINTER(c_int32), ctypes.POINTER(c_int32),
                                   c_void_p, c_void_p]),
    "mdsf_deposit_fold": (c_int32, [c_void_p, c_void_p, c_int32, c_void_p, c_void_p,
                                    ctypes.POINTER(c_double), c_int32, c_int64, c_int64,
                                    c_void_p, c_void_p, c_void_p]),
    "mdsf_prepare_frame": (c_int32, [c_void_p, c_void_p, c_int32, c_void_p, c_void_p,
                                     ctypes.POINTER(c_double), c_int32, c_int64, c_int64,
                                     c_void_p, c_void_p, c_void_p]),
    "mdsf_deposit": (c_int32, [c_void_p, c_void_p, c_void_p]),
    "mdsf_set_frame_buffer": (c_int32, [c_void_p, c_int32]),
    "mdsf_set_stage_timing": (c_int32, [c_void_p, c_int32]),
    "mdsf_stage_times": (c_int32, [c_void_p, ctypes.POINTER(c_double)]),
    "mdsf_fft_pass": (c_int32, [c_void_p, c_int32, c_void_p, c_void_p]),
    "mdsf_accumulate_density": (c_int32, [c_void_p, c_void_p, c_void_p]),
    "mdsf_rfft3": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "mdsf_rfft3_noalloc": (c_int32, [ctypes.POINTER(c_int32), c_int32, c_void_p, c_void_p,
                                     c_void_p]),
    "mdsf_sf_read": (c_int32, [c_void_p, c_void_p, c_void_p]),
    "mdsf_outputs": (c_int32, [c_void_p, c_void_p, ctypes.POINTER(c_double), c_int32, c_void_p,
                               c_void_p, c_void_p, c_void_p]),
    "mdsf_ewald_slices": (c_int32, [c_void_p, c_int32, c_int32, c_int32, c_double, c_void_p,
                                    c_void_p, c_void_p, c_void_p]),
    "mdsf_rz_average": (c_int32, [c_void_p, c_int32, c_int32, c_int32, c_void_p, c_void_p,
                                  c_int32, c_void_p, c_int32, ctypes.POINTER(c_double), c_void_p,
                                  c_void_p]),
    "mdsf_interp_points": (c_int32, [c_void_p, c_int32, c_int32, c_int32, c_void_p, c_int64,
                                     c_void_p, c_void_p]),
    "mdsf_adjoint_frames": (c_int32, [c_void_p, c_void_p, c_int32, c_int64, c_void_p, c_void_p,
                                      ctypes.POINTER(c_double), c_int32, c_int64, c_int64,
                                      c_void_p, c_void_p, c_void_p]),
    "mdsf_density_adjoint": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "mdsf_deposit_adjoint": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "mdsf_batch_create": (c_int32, [ctypes.POINTER(GridDesc), c_int32,
                                    ctypes.POINTER(c_void_p)]),
    "mdsf_batch_destroy": (c_int32, [c_void_p]),
    "mdsf_batch_plan": (c_void_p, [c_void_p, c_int32]),
    "mdsf_batch_set_frames": (c_int32, [c_void_p, c_int32, c_void_p, c_int32, c_int64, c_void_p,
                                        c_void_p, ctypes.POINTER(c_double), c_int32, c_int64,
                                        c_int64]),
    "mdsf_batch_run": (c_int32, [c_void_p, c_void_p]),
    "mdsf_batch_check_errors": (c_int32, [c_void_p, c_void_p]),
}


class MdsfError(RuntimeError):
    """A HIP runtime failure inside libmdsf."""


_lib = None


def load():
    """Load ``libmdsf.so`` once and attach the signatures; raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libmdsf.so not found at {LIB_PATH}: build it with `python -m mdsf.build` "
            "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    """Map an ``mdsf_*`` return code to the exception the reference would raise."""
    if rc == OK:
        return
    msg = load().mdsf_last_error().decode(errors="replace")
    if rc in (ERR_ARG, ERR_RANGE):
        raise ValueError(msg)
    if rc == ERR_TYPE:
        raise KeyError(msg)
    if rc == ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise MdsfError(msg)


def doubles(vals):
    arr = (c_double * len(vals))(*[float(v) for v in vals])
    return arr
