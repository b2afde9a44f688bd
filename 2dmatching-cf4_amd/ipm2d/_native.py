"""ctypes binding of the gfx950 hot-path library (include/ipm_hip.h).

The shared library is built in-tree by ``build.py`` into ``ipm2d/_lib``.
There is no fallback: if the library is missing or cannot be loaded, every
entry point raises, so a run can never silently compute somewhere else.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libipm_hip.so")

IPM_OK = 0
IPM_EINVAL = -1
IPM_EHIP = -2
IPM_ECAPACITY = -3
IPM_ENODEV = -4
LAYOUT_AOS = 0
LAYOUT_SOA = 1
NS_MAX = 9
NG_MAX = 3
NR_MAX = 128
CAPACITY_MAX = 536870880        # IPM_CAPACITY_MAX: 32-bit byte offsets per field


class IPMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[ipm {code}] {msg}")
        self.code = code


class CapacityError(IPMError):
    pass


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i64p = C.POINTER(C.c_int64)
_vp = C.c_void_p

# name -> (restype, argtypes); every function returns int status
SIGNATURES = {
    "ipm_create": [C.POINTER(_vp), C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int],
    "ipm_destroy": [_vp],
    "ipm_set_species": [_vp, C.c_int, C.c_double, C.c_double, C.c_int64],
    "ipm_set_geometry": [_vp, _dp, _ip, _ip, _ip, C.c_int, C.c_int, _ip],
    "ipm_upload_species": [_vp, C.c_int, C.c_int64, _dp, C.c_int],
    "ipm_upload_species_device": [_vp, C.c_int, C.c_int64, C.c_int64, _vp, C.c_int64],
    "ipm_download_species": [_vp, C.c_int, _i64p, _dp, C.c_int64, C.c_int],
    "ipm_moments_noalloc": [_vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp],
    "ipm_count": [_vp, C.c_int, _i64p],
    "ipm_set_fields": [_vp, _dp, _dp],
    "ipm_move_es": [_vp, C.c_int],
    "ipm_absorb": [_vp, C.c_int],
    "ipm_move_absorb": [_vp, C.c_int],
    "ipm_take_tallies": [_vp, _dp, _dp],
    "ipm_p2c": [_vp, C.c_int, _dp, _dp],
    "ipm_field_sum": [_vp, _dp, _dp],
    "ipm_set_mcc": [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                    C.c_double, C.c_double, _ip, _ip, _dp, _dp, C.c_double, C.c_double, C.c_double],
    "ipm_set_rng": [_vp, C.c_uint64, C.c_uint32],
    "ipm_mcc": [_vp],
    "ipm_set_event_counting": [_vp, C.c_int],
    "ipm_set_null_screen": [_vp, C.c_int],
    "ipm_take_mcc_counts": [_vp, _i64p],
    "ipm_step": [_vp],
    "ipm_step_graph": [_vp],
    "ipm_set_timing": [_vp, C.c_int],
    "ipm_take_timing": [_vp, _dp, _i64p],
    "ipm_sort": [_vp, C.c_int],
    "ipm_set_step_rezone": [_vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double],
    "ipm_set_sort_interval": [_vp, C.c_int],
    "ipm_set_launch": [_vp, C.c_int, C.c_int],
    "ipm_set_deposit_window": [_vp, C.c_int],
    "ipm_set_sort_tile": [_vp, C.c_int],
    "ipm_set_collide_grid": [_vp, C.c_int],
    "ipm_synchronize": [_vp],
    "ipm_species_device_ptrs": [_vp, C.c_int, C.POINTER(_dp)],
    "ipm_set_potential": [_vp, _dp],
    "ipm_set_solver_geometry": [_vp, _dp, _ip],
    "ipm_solver_system": [_vp, C.c_int, _dp, _dp, _ip, _dp, _dp],
    "ipm_moments": [_vp, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp],
    "ipm_recombine": [_vp, C.c_int, _ip, _dp, C.c_double, _i64p],
    "ipm_split_merge": [_vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _i64p],
    "ipm_window_create": [_vp, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp],
    "ipm_window_forward": [_vp, _dp, C.c_int, _dp],
    "ipm_window_backward": [_vp, _dp, _dp, _dp],
    "ipm_window_info": [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int), C.POINTER(C.c_float),
                        C.POINTER(C.c_float)],
    "ipm_window_download": [_vp, C.c_int, _dp, _ip],
    "ipm_window_destroy": [_vp],
}

_lib = None


def load():
    """Load the library (once).  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise IPMError(IPM_ENODEV, f"native library not built: {LIB_PATH} (run build.py)")
    lib = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = C.c_int
        fn.argtypes = args
    lib.ipm_last_error.restype = C.c_char_p
    lib.ipm_last_error.argtypes = []
    _lib = lib
    return lib


def check(rc):
    if rc != IPM_OK:
        msg = load().ipm_last_error().decode(errors="replace")
        if rc == IPM_ECAPACITY:
            raise CapacityError(rc, msg)
        raise IPMError(rc, msg)
    return rc


def call(name, *args):
    return check(getattr(load(), name)(*args))
