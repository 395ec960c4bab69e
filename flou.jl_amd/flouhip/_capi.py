"""ctypes binding of libflouhip.so (include/flouhip.h).

The library is loaded from the package directory (flou.jl_amd/libflouhip.so).  There is
no fallback: if it is missing or cannot be loaded, importing flouhip raises ImportError.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# $FLOUHIP_LIB selects another build of the library (same ABI), e.g. for A/B timing
LIB_PATH = os.environ.get("FLOUHIP_LIB") or os.path.join(os.path.dirname(_HERE), "libflouhip.so")

OK, EINVAL, EUNSUPPORTED, EDEVICE, ENOMEM, EDOMAIN = 0, -1, -2, -3, -4, -5
FP64, FP32 = 0, 1
ADVECTION, EULER = 0, 1
STRONG, SPLIT, HYBRID = 0, 1, 2
TP_STDAVG, TP_CHANDRASEKHAR, TP_KENNEDY_GRUBER = 0, 1, 2
AVG_STD, AVG_CHANDRASEKHAR = 0, 1
DISS_NONE, DISS_LXF, DISS_MATRIX, DISS_SCALAR = 0, 1, 2, 3
LIMITER_NONE, LIMITER_ZHANG_SHU = 0, 1
MONITOR_KINETIC_ENERGY, MONITOR_ENTROPY = 0, 1
BC_DIRICHLET, BC_INFLOW, BC_OUTFLOW, BC_SLIP = 0, 1, 2, 3
SSPRK33, CK2N54, ORK256 = 0, 1, 2
GEOM_AFFINE, GEOM_CURVED = 0, 1

_d = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.POINTER(ctypes.c_int64)
_i8 = ctypes.POINTER(ctypes.c_int8)
_u8 = ctypes.POINTER(ctypes.c_uint8)
_i32 = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p


class Desc(ctypes.Structure):
    """flou_hip_desc."""
    _fields_ = [
        ("nd", ctypes.c_int32), ("n", ctypes.c_int32), ("nvars", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("nelements", ctypes.c_int64), ("nfaces", ctypes.c_int64),
        ("nboundaries", ctypes.c_int32),
        ("equation", ctypes.c_int32), ("gamma", ctypes.c_double),
        ("velocity", ctypes.c_double * 3),
        ("volume_op", ctypes.c_int32), ("tp_flux", ctypes.c_int32),
        ("surface_avg", ctypes.c_int32), ("surface_diss", ctypes.c_int32),
        ("intensity", ctypes.c_double),
        ("D", _d), ("Ds", _d), ("Dsharp", _d), ("l_minus", _d), ("l_plus", _d),
        ("dg_minus", _d), ("dg_plus", _d), ("omega", _d),
        ("elem_faceinds", _i64), ("elem_facepos", _i8), ("face_eleminds", _i64),
        ("face_elempos", _i8), ("face_orientation", _u8), ("face_boundary", _i32),
        ("geometry", ctypes.c_int32),
        ("elem_jac", _d), ("elem_metric", _d), ("elem_volume", _d),
        ("face_frames", _d), ("face_jac", _d),
        ("blend", ctypes.c_double),
    ]


SYMBOLS = {
    "flou_hip_version": (ctypes.c_char_p, []),
    "flou_hip_create": (ctypes.c_int, [ctypes.POINTER(Desc), ctypes.POINTER(_vp)]),
    "flou_hip_destroy": (ctypes.c_int, [_vp]),
    "flou_hip_last_error": (ctypes.c_char_p, [_vp]),
    "flou_hip_create_error": (ctypes.c_char_p, []),
    "flou_hip_set_option": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32]),
    "flou_hip_set_stream": (ctypes.c_int, [_vp, _vp]),
    "flou_hip_set_bc": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, _d]),
    "flou_hip_ndofs": (ctypes.c_int64, [_vp]),
    "flou_hip_state_bytes": (ctypes.c_size_t, [_vp]),
    "flou_hip_alloc": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.c_size_t]),
    "flou_hip_free": (ctypes.c_int, [_vp, _vp]),
    "flou_hip_upload": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "flou_hip_download": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "flou_hip_synchronize": (ctypes.c_int, [_vp]),
    "flou_hip_rhs": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_double]),
    "flou_hip_rk_stage": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                         ctypes.c_double, _vp, _vp, _vp, _vp]),
    "flou_hip_timeintegrate": (ctypes.c_int, [_vp, ctypes.c_int32, _vp, ctypes.c_double,
                                              ctypes.c_int64, ctypes.c_double, _d]),
    "flou_hip_max_dt": (ctypes.c_int, [_vp, _vp, ctypes.c_double, _d]),
    "flou_hip_integrate": (ctypes.c_int, [_vp, _vp, ctypes.c_int32, _d]),
    "flou_hip_check": (ctypes.c_int, [_vp]),
    "flou_hip_set_limiter": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_double]),
    "flou_hip_limit": (ctypes.c_int, [_vp, _vp, ctypes.c_double]),
    "flou_hip_monitor": (ctypes.c_int, [_vp, _vp, ctypes.c_int32, _d]),
    "flou_hip_equispaced": (ctypes.c_int, [_vp, ctypes.c_int32, _d, _vp, ctypes.c_int32, _vp]),
    "flou_hip_rhs_vjp": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "flou_hip_rk_step_record": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                               ctypes.POINTER(_vp), _vp]),
    "flou_hip_rk_step_vjp": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                            ctypes.POINTER(_vp), _vp]),
    "flou_hip_rk_step_vjp_noalloc": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                            ctypes.POINTER(_vp), _vp]),
}


def _share_torch_hip_runtime():
    """Load torch's HIP runtime first.

    torch ships its own libamdhip64.so (same soname, libamdhip64.so.7, as /opt/rocm's).  A
    process must hold a single HIP/HSA runtime, so when torch provides the device buffers
    libflouhip must bind to torch's copy: loading it globally before libflouhip makes the
    dynamic linker resolve libflouhip's DT_NEEDED entry to it."""
    try:
        import torch  # noqa: F401
    except ImportError:
        return
    libdir = os.path.join(os.path.dirname(torch.__file__), "lib")
    hip = os.path.join(libdir, "libamdhip64.so")
    if os.path.exists(hip):
        ctypes.CDLL(hip, mode=ctypes.RTLD_GLOBAL)


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libflouhip.so not found at {LIB_PATH}; build it with "
            "`python flou.jl_amd/build_lib.py` (there is no CPU fallback)")
    _share_torch_hip_runtime()
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SYMBOLS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


class DomainError(ArithmeticError):
    """A non-physical state was met (Julia's DomainError from sqrt/log of a negative)."""


def check(rc, handle=None, what=""):
    if rc == OK:
        return
    msg = (lib.flou_hip_last_error(handle) if handle else lib.flou_hip_create_error()) or b""
    msg = f"{what}: {msg.decode()}" if what else msg.decode()
    if rc == EINVAL:
        raise ValueError(msg)
    if rc == EUNSUPPORTED:
        raise NotImplementedError(msg)
    if rc == EDOMAIN:
        raise DomainError(msg)
    if rc == ENOMEM:
        raise MemoryError(msg)
    raise RuntimeError(f"flouhip error {rc}: {msg}")
