"""ctypes binding of the C ABI in ``include/amg.h`` (``libanelastic_mg.so``).

The library is loaded from this package directory; there is no fallback:
a missing library raises ``ImportError`` with the build command.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libanelastic_mg.so")

AMG_OK = 0
AMG_EINVAL = -1
AMG_EBADSIZE = -2
AMG_EBADBC = -3
AMG_ENOMEM = -4
AMG_ENOTCONV = -5
AMG_ENAN = -6
AMG_EHIP = -7
AMG_EUNSUPPORTED = -8
AMG_EPARSE = -9
AMG_ENOGPU = -10
AMG_EBUSY = -11

BC_PERIODIC, BC_NEUMANN, BC_DIRICHLET = 0, 1, 2
SMOOTHER_RBPGS, SMOOTHER_LJ = 0, 2
FP64, MIXED = 0, 1
STOP_RESIDUAL, STOP_RELMAX = 0, 1
MAX_LEVELS = 16
AMG_FP64, AMG_MIXED = 0, 1
AMG_STOP_RESIDUAL, AMG_STOP_RELMAX = 0, 1

_dp = C.POINTER(C.c_double)


class amg_config(C.Structure):
    _fields_ = [
        ("nlevels", C.c_int * 3),
        ("smoother", C.c_int),
        ("maxcy", C.c_int),
        ("tol", C.c_double),
        ("stop_rule", C.c_int),
        ("pre_relax", C.c_int * MAX_LEVELS),
        ("post_relax", C.c_int * MAX_LEVELS),
        ("tbc", C.c_int * 6),
        ("vbc", C.c_double * 6),
        ("precision", C.c_int),
        ("report_relmax", C.c_int),
        ("verbose", C.c_int),
        ("strict", C.c_int),
        ("use_graph", C.c_int),
        ("coarse_direct", C.c_int),
        ("pair_cycles", C.c_int),
    ]


class amg_grid(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("xe", _dp), ("xc", _dp), ("dxe", _dp), ("dxc", _dp),
        ("ye", _dp), ("yc", _dp), ("dye", _dp), ("dyc", _dp),
        ("ze", _dp), ("zc", _dp), ("dze", _dp), ("dzc", _dp),
        ("rho", _dp),
    ]


class amg_stats(C.Structure):
    _fields_ = [
        ("iters", C.c_int),
        ("converged", C.c_int),
        ("relres", C.c_double),
        ("relmax", C.c_double),
        ("res0", C.c_double),
        ("bnorm", C.c_double),
        ("ms_total", C.c_double),
        ("ms_cycles", C.c_double),
        ("bytes_cycles", C.c_double),
        ("nlevels", C.c_int),
        ("level_n", (C.c_int * 3) * MAX_LEVELS),
    ]


class amg_flowbc(C.Structure):
    """TB/VB(1:3,1:2,1:12) of the reference's module boundC, Fortran order."""
    _fields_ = [
        ("tb", C.c_int * 72),
        ("vb", C.c_double * 72),
    ]


_PROTOS = {
    "amg_config_default": (None, [C.POINTER(amg_config)]),
    "amg_read_params": (C.c_int, [C.c_char_p, C.POINTER(amg_config)]),
    "amg_create": (C.c_int, [C.POINTER(amg_config), C.POINTER(amg_grid), C.POINTER(C.c_void_p)]),
    "amg_destroy": (C.c_int, [C.c_void_p]),
    "amg_solve_host": (C.c_int, [C.c_void_p, _dp, _dp, C.POINTER(C.c_int), _dp,
                                 C.POINTER(amg_stats)]),
    "amg_solve_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int), _dp,
                                   C.POINTER(amg_stats)]),
    "amg_stream": (C.c_void_p, [C.c_void_p]),
    "amg_levels": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int * 3)]),
    "amg_direct_level": (C.c_int, [C.c_void_p]),
    "amg_psource_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_double]),
    "amg_pgrad_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_double]),
    "amg_divergence_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p, _dp, _dp, C.POINTER(C.c_int)]),
    "amg_divergence_device_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                                C.c_void_p, _dp, _dp, C.POINTER(C.c_int)]),
    "amg_ghost_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_char_p, C.POINTER(amg_flowbc)]),
    "amg_pgrad_ghost_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_double, C.POINTER(amg_flowbc)]),
    "amg_cn_advance_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.c_int, C.c_void_p,
                                        C.POINTER(amg_flowbc), C.c_double, C.c_double, C.c_double]),
    "amg_batch_create": (C.c_int, [C.c_int, C.POINTER(amg_config), C.POINTER(amg_grid),
                                   C.POINTER(C.c_void_p)]),
    "amg_batch_destroy": (C.c_int, [C.c_void_p]),
    "amg_batch_solve_device": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                         C.POINTER(C.c_int), _dp, C.POINTER(amg_stats)]),
    "amg_server_create": (C.c_int, [C.POINTER(amg_config), C.c_double, C.c_int, C.c_int,
                                    C.POINTER(C.c_void_p)]),
    "amg_server_destroy": (C.c_int, [C.c_void_p]),
    "amg_server_add_member": (C.c_int, [C.c_void_p, C.POINTER(amg_grid)]),
    "amg_server_submit": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "amg_server_poll": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), _dp, _dp]),
    "amg_server_drain": (C.c_int, [C.c_void_p]),
    "amg_run_kernel": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _dp]),
    "amg_level_field": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp]),
    "amg_strerror": (C.c_char_p, [C.c_int]),
    "amg_last_error": (C.c_char_p, []),
}

EXPORTED = sorted(_PROTOS)

_lib = None


def load():
    """Load ``libanelastic_mg.so`` and set every prototype (no compute)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m anelastic_solve_amd.build` "
            "(there is no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class AMGError(RuntimeError):
    def __init__(self, code, where=""):
        lib = load()
        msg = lib.amg_strerror(code).decode()
        detail = lib.amg_last_error().decode()
        super().__init__(f"{where}: {msg} ({code})" + (f": {detail}" if detail else ""))
        self.code = code


def check(code, where):
    if code != AMG_OK:
        raise AMGError(code, where)
    return code
