"""ctypes binding of libamcafe.so (include/amcafe.h).  No fallback path exists."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("AMCAFE_LIB_PATH") or os.path.join(_HERE, "_lib", "libamcafe.so")


class LibraryMissing(RuntimeError):
    pass


class AmcafeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"amcafe error {code}: {msg}")
        self.code = code


class Stats(ctypes.Structure):
    _fields_ = [("n_active", ctypes.c_int), ("n_mushy0", ctypes.c_int), ("n_events", ctypes.c_int),
                ("n_captures", ctypes.c_int), ("n_nucleations", ctypes.c_int), ("n_growing", ctypes.c_int),
                ("nGrain", ctypes.c_int), ("draws", ctypes.c_longlong), ("tinc", ctypes.c_double),
                ("slot_visits", ctypes.c_longlong), ("cand_evals", ctypes.c_longlong),
                ("interface_events", ctypes.c_longlong), ("loop_cycles", ctypes.c_longlong)]


class Tuning(ctypes.Structure):
    _fields_ = [("list_blocks", ctypes.c_int), ("tunings", ctypes.c_int), ("growing_cells", ctypes.c_int),
                ("candidates", ctypes.c_int * 5), ("ms", ctypes.c_float * 5)]


class Info(ctypes.Structure):
    _fields_ = [("nX", ctypes.c_int * 3), ("nGrain", ctypes.c_int), ("tInd", ctypes.c_int), ("isp", ctypes.c_int),
                ("indlayer", ctypes.c_int), ("ilaserLoc", ctypes.c_int), ("nlayerTot", ctypes.c_int),
                ("done", ctypes.c_int), ("inewscanflg", ctypes.c_int), ("inewlayerflg", ctypes.c_int),
                ("time", ctypes.c_double), ("gth", ctypes.c_double)]


class DomainDesc(ctypes.Structure):
    _fields_ = [("nX", ctypes.c_int * 3), ("dX", ctypes.c_double), ("tL", ctypes.c_double), ("tS", ctypes.c_double),
                ("T0", ctypes.c_double), ("meltparam", ctypes.c_double * 4), ("bmV", ctypes.c_double),
                ("bmDelT", ctypes.c_double), ("rNmax", ctypes.c_double), ("Avel", ctypes.c_double),
                ("nvel", ctypes.c_double)]


class Laser(ctypes.Structure):
    _fields_ = [("x", ctypes.c_double), ("y", ctypes.c_double), ("z", ctypes.c_double), ("dirp", ctypes.c_double),
                ("gth", ctypes.c_double), ("ilaserLoc", ctypes.c_int)]


# every symbol include/amcafe.h declares, with its signature
_P, _I, _D, _C = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_char_p
SIGNATURES = {
    "amcafe_last_error": (_C, []),
    "amcafe_device_name": (_I, [_C, _I]),
    "amcafe_domain_create": (_I, [ctypes.POINTER(DomainDesc), ctypes.POINTER(_P)]),
    "amcafe_domain_destroy": (None, [_P]),
    "amcafe_domain_upload": (_I, [_P, _P, _P, _P, _P, _P]),
    "amcafe_domain_download": (_I, [_P, _P, _P, _P, _P, _P, _P]),
    "amcafe_domain_set_grains": (_I, [_P, _P, _I]),
    "amcafe_domain_get_grains": (_I, [_P, _P, ctypes.POINTER(_I)]),
    "amcafe_analytic_temp_curr": (_I, [_P, ctypes.POINTER(Laser), ctypes.POINTER(_I)]),
    "amcafe_easm_temp_curr": (_I, [_P, ctypes.POINTER(Laser), _D, _D, ctypes.POINTER(_I),
                                   ctypes.POINTER(ctypes.c_longlong)]),
    "amcafe_update_voxels": (_I, [_P, _I, _I, ctypes.POINTER(Stats)]),
    "amcafe_sim_create": (_I, [ctypes.POINTER(_C), _I, ctypes.POINTER(_P)]),
    "amcafe_sim_destroy": (None, [_P]),
    "amcafe_sim_step": (_I, [_P, ctypes.POINTER(_I)]),
    "amcafe_sim_info": (_I, [_P, _I, ctypes.POINTER(Info)]),
    "amcafe_sim_stats": (_I, [_P, _I, ctypes.POINTER(Stats)]),
    "amcafe_sim_tuning": (_I, [_P, ctypes.POINTER(Tuning)]),
    "amcafe_sim_get_fields": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P]),
    "amcafe_sim_get_grains": (_I, [_P, _I, _P, ctypes.POINTER(_I)]),
    "amcafe_sim_write_hdf5": (_I, [_P, _I, _C]),
    "amcafe_sim_set_output": (_I, [_P, _I, _C]),
    "amcafe_write_hdf5_fields": (_I, [_C, _P, _P, _P, _P, _P, _P, _P, _I]),
    "amcafe_ipf_colours": (_I, [_P, _I, _P]),
    "amcafe_heat_loss_f32": (_I, [_P, _P, _P, _I, _I, _P, _P]),
    "amcafe_heat_loss_grad_f32": (_I, [_P, _P, _P, _I, _I, _P, _P, _P]),
    "amcafe_heat_loss_f64": (_I, [_P, _P, _P, _I, _I, _P, _P]),
    "amcafe_heat_loss_grad_batch_f32": (_I, [_P, _I, _P, _P, _I, _I, _P, _P, _P]),
    "amcafe_heat_last_evaluations": (ctypes.c_longlong, []),
    "amcafe_heat_loss_grad_batch_f32_noalloc": (_I, [_P, _I, _P, _P, _I, _I, _P, _P]),
    "amcafe_heat_noalloc_bytes": (ctypes.c_longlong, [_P, _I]),
    "amcafe_heat_noalloc_max_bytes": (ctypes.c_longlong, []),
    "amcafe_sim_set_timing": (_I, [_P, _I]),
    "amcafe_sim_timings": (_I, [_P, _P, _I]),
}

_lib = None


def lib():
    """Load libamcafe.so once; raise LibraryMissing when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(f"{LIB_PATH} not found; build it with "
                             "`python -c 'import __graft_entry__ as g; g.build()'`")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    # known-answer helpers of the device library (host-callable)
    L.amcd_uniform_stream.argtypes = [ctypes.c_ulonglong, _I, _P]
    L.amcd_uniform_jump.argtypes = [ctypes.c_ulonglong, _I, _P]
    L.amcd_host_shape_factor.argtypes = [_P, _I, _P]
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise AmcafeError(rc, lib().amcafe_last_error().decode())
    return rc
