"""ctypes binding of libfcfc_gpu.so (the C ABI in include/fcfc_gpu.h).

The library is loaded from fcfc_amd/lib/.  There is no fallback: if the
library is missing or cannot be loaded, importing fcfc_amd raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libfcfc_gpu.so")

# error codes (include/fcfc_gpu.h, same values as FCFC's define_comm.h)
FCFC_OK = 0
FCFC_ERR_MEMORY = -1
FCFC_ERR_ARG = -2
FCFC_ERR_FILE = -3
FCFC_ERR_CFG = -4
FCFC_ERR_GPU = -60

FCFC_BIN_ISO, FCFC_BIN_SMU, FCFC_BIN_SPI = 0, 1, 2
FCFC_GEOM_BOX, FCFC_GEOM_SURVEY = 0, 1

FCFC_GPU_AUTO = 0x01
FCFC_GPU_WEIGHTED = 0x02
FCFC_GPU_PREC_FP64 = 0x00
FCFC_GPU_PREC_FP32 = 0x10
FCFC_GPU_CNT_DEVICE = 0x20
FCFC_GPU_CENSUS = 0x40
FCFC_GPU_GRADIENT = 0x100
FCFC_GPU_PREC_PLAIN = 0x80

_dp = C.POINTER(C.c_double)


class fcfc_bins(C.Structure):
    _fields_ = [
        ("geometry", C.c_int), ("bintype", C.c_int),
        ("ns", C.c_int), ("nmu", C.c_int), ("np", C.c_int), ("ntot", C.c_int),
        ("rescale", C.c_double),
        ("sbin_raw", _dp), ("pbin_raw", _dp), ("sbin", _dp), ("s2bin", _dp),
        ("pbin", _dp),
        ("tabtype", C.c_int), ("stab_len", C.c_int), ("ptab_len", C.c_int),
        ("stab", C.POINTER(C.c_uint16)), ("ptab", C.POINTER(C.c_uint16)),
        ("soff", C.c_int), ("poff", C.c_int),
    ]


class fcfc_cat(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p),
        ("s", C.c_void_p), ("w", C.c_void_p),
        ("on_device", C.c_int),
    ]


class fcfc_gpu_stats(C.Structure):
    _fields_ = [
        ("evals", C.c_double), ("rechecked", C.c_double), ("census", C.c_double),
        ("kernel_ms", C.c_double), ("total_ms", C.c_double),
        ("cell_div", C.c_int), ("tile", C.c_int),
    ]


class fcfc_batch_job(C.Structure):
    _fields_ = [
        ("a", fcfc_cat), ("b", fcfc_cat),
        ("bsize", C.c_double * 3),
        ("flags", C.c_uint32),
    ]


class fcfc_cosmology(C.Structure):
    _fields_ = [("omega_m", C.c_double), ("omega_l", C.c_double), ("omega_k", C.c_double),
                ("w", C.c_double), ("err", C.c_double)]


# every symbol the header declares, with its ctypes signature
SIGNATURES = {
    "fcfc_bins_linear": (C.c_int, [C.POINTER(fcfc_bins), C.c_int, C.c_int, C.c_double,
                                   C.c_double, C.c_double, C.c_int, C.c_double,
                                   C.c_double, C.c_double, _dp]),
    "fcfc_bins_edges": (C.c_int, [C.POINTER(fcfc_bins), C.c_int, C.c_int, _dp, C.c_int,
                                  C.c_int, _dp, C.c_int, _dp]),
    "fcfc_bins_free": (None, [C.POINTER(fcfc_bins)]),
    "fcfc_prepare_catalog": (C.c_int, [C.POINTER(fcfc_bins), C.c_int64, _dp, _dp, _dp,
                                       _dp, _dp, _dp]),
    "fcfc_gpu_init": (C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    "fcfc_gpu_destroy": (C.c_int, [C.c_void_p]),
    "fcfc_gpu_last_error": (C.c_char_p, [C.c_void_p]),
    "fcfc_gpu_count": (C.c_int, [C.c_void_p, C.POINTER(fcfc_cat), C.POINTER(fcfc_cat),
                                 C.POINTER(fcfc_bins), _dp, C.c_uint32, C.c_void_p]),
    "fcfc_gpu_count_noalloc": (C.c_int, [C.c_void_p, C.POINTER(fcfc_cat), C.POINTER(fcfc_cat),
                                         C.POINTER(fcfc_bins), _dp, C.c_uint32, C.c_void_p]),
    "fcfc_gpu_last_stats": (C.c_int, [C.c_void_p, C.POINTER(fcfc_gpu_stats)]),
    "fcfc_gpu_count_batch": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(fcfc_batch_job),
                                       C.POINTER(fcfc_bins), C.c_uint32, C.c_void_p]),
    "fcfc_gpu_plan_create": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(fcfc_batch_job),
                                       C.POINTER(fcfc_bins), C.c_uint32,
                                       C.POINTER(C.c_void_p)]),
    "fcfc_gpu_plan_run": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fcfc_gpu_plan_destroy": (C.c_int, [C.c_void_p]),
    "fcfc_gpu_count_backward": (C.c_int, [C.c_void_p, C.POINTER(fcfc_cat),
                                          C.POINTER(fcfc_cat), C.POINTER(fcfc_bins), _dp,
                                          C.c_uint32, _dp, _dp, _dp]),
    "fcfc_gpu_set_tuning": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "fcfc_gpu_autotune": (C.c_int, [C.c_void_p, C.POINTER(fcfc_cat), C.POINTER(fcfc_cat),
                                    C.POINTER(fcfc_bins), _dp, C.c_uint32, C.POINTER(C.c_int)]),
    "fcfc_save_pair_count": (C.c_int, [C.c_char_p, C.POINTER(fcfc_bins), C.c_int, _dp,
                                       C.c_int, C.c_int, C.c_uint64, C.c_uint64,
                                       C.c_double, C.c_double, C.c_void_p]),
    "fcfc_cosmology_check": (C.c_int, [C.POINTER(fcfc_cosmology)]),
    "fcfc_cmvdst_order": (C.c_int, [C.POINTER(fcfc_cosmology), C.c_int64, _dp,
                                    C.POINTER(C.c_int)]),
    "fcfc_convert_coords": (C.c_int, [C.POINTER(fcfc_cosmology), C.c_int64, _dp, _dp, _dp,
                                      C.POINTER(C.c_int)]),
    "fcfc_convert_coords_table": (C.c_int, [_dp, _dp, C.c_int64, C.c_int64, _dp, _dp, _dp]),
    "fcfc_gen_uniform": (C.c_int, [C.c_uint64, C.c_int64, C.c_double, _dp, _dp, _dp]),
    "fcfc_gen_clustered": (C.c_int, [C.c_uint64, C.c_int64, C.c_int64, C.c_int64,
                                     C.c_double, C.c_double, C.c_double, _dp, _dp, _dp,
                                     _dp]),
    "fcfc_gen_survey": (C.c_int, [C.c_uint64, C.c_int64, C.c_double, C.c_double,
                                  C.c_double, C.c_double, C.c_double, C.c_double, _dp,
                                  _dp, _dp, _dp]),
    "fcfc_gen_sizes": (C.c_int, [C.c_uint64, C.c_int, C.c_double, C.c_double,
                                 C.POINTER(C.c_int64)]),
}

_lib = None


def load():
    """Load libfcfc_gpu.so (raises OSError if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} is not built; run `python -m fcfc_amd.build`")
    # Share one HIP runtime with PyTorch: torch bundles its own libamdhip64
    # (same SONAME, different file name in its DT_NEEDED).  Loaded first, it
    # satisfies this library's dependency; loaded second, a process would hold
    # two HIP runtimes and torch would find no GPU.
    try:
        import torch  # noqa: F401
    except Exception:
        pass
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib
