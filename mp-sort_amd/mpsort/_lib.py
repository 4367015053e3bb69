"""ctypes view of libmpsort_gpu.so (include/mpsort_gpu.h).

The library is the only compute path: importing this module loads it and
raises ImportError when it is missing (build it with ``make -C mp-sort_amd``).
There is no CPU fallback.
"""
import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.dirname(_PKG)
LIBPATH = os.environ.get("MPSORT_GPU_LIB", os.path.join(LIBDIR, "libmpsort_gpu.so"))
GENPATH = os.path.join(LIBDIR, "libmpsort_gen.so")

MPSORT_OK = 0
MPSORT_ERR_ARG = -1
MPSORT_ERR_SIZE_MISMATCH = -2
MPSORT_ERR_LAYOUT = -3
MPSORT_ERR_CHECKSUM = -4
MPSORT_ERR_HIP = -5
MPSORT_ERR_NOMEM = -6
MPSORT_ERR_UNSUPPORTED = -7
MPSORT_ERR_NODEVICE = -8
MPSORT_ERR_TIMEOUT = -9

DEVICE_PTRS = 1 << 0
NO_CHECKSUM = 1 << 1
TIMING = 1 << 2
LEARNED = 1 << 3
FORCE_GENERAL = 1 << 4
LEARNED_EXCHANGE = 1 << 5
EXCHANGE = LEARNED_EXCHANGE  # explicit exchange before the SecondSort (either path)
EXCHANGE_DENSE = 1 << 6
EXCHANGE_SPARSE = 1 << 7
FIRST_RADIX = 1 << 8
FIRST_BUCKETS = 1 << 9
FAULT_SPLITTERS = 1 << 0
FAULT_BUCKET_CAP = 1 << 1
FAULT_PLAIN = 1 << 2
FAULT_CORRUPT = 1 << 3

# mpsort.h:17-20
MPSORT_DISABLE_SPARSE_ALLTOALLV = 1 << 1
MPSORT_DISABLE_GATHER_SORT = 1 << 3
MPSORT_REQUIRE_GATHER_SORT = 1 << 4
MPSORT_REQUIRE_SPARSE_ALLTOALLV = 1 << 6

# every symbol include/mpsort_gpu.h and include/mpsort.h declare
EXPORTS = [
    "mpsort_gpu_newarray", "mpsort_gpu_newarray_radix", "mpsort_gpu_last_stats",
    "mpsort_gpu_report_last_run", "mpsort_gpu_last_counts", "mpsort_gpu_set_allocator",
    "mpsort_gpu_release", "mpsort_gpu_set_vranks", "mpsort_gpu_get_vranks",
    "mpsort_gpu_last_error", "mpsort_gpu_version", "mpsort_gpu_dmalloc", "mpsort_gpu_dfree",
    "mpsort_gpu_memcpy_h2d", "mpsort_gpu_memcpy_d2h", "mpsort_gpu_synchronize", "mpsort_gpu_dcopy",
    "mpsort_gpu_device_count", "mpsort_mpi_set_options", "mpsort_mpi_has_options",
    "mpsort_mpi_unset_options", "mpsort_mpi_report_last_run", "radix_sort", "mpsort_omp",
    "mpsort_gpu_model_train", "mpsort_gpu_model_free", "mpsort_gpu_model_get_info",
    "mpsort_gpu_model_eval", "mpsort_gpu_model_params", "mpsort_gpu_mlp_grad",
    "mpsort_gpu_newarray_learned", "mpsort_gpu_segplan_create", "mpsort_gpu_segplan_add",
    "mpsort_gpu_segplan_run", "mpsort_gpu_segplan_info", "mpsort_gpu_segplan_free",
    "mpsort_gpu_sort_segments",
    "mpiu_malloc", "mpiu_free", "mpiu_set_malloc", "mpsort_gpu_newarray_noalloc", "mpsort_gpu_set_faults",
    "mpsort_gpu_segplan_shape",
]

SEGPLAN_GRAPH = 1 << 0
SEGPLAN_TUNE = 1 << 1
SEGPLAN_SMALL_512x9 = 1 << 2
SEGPLAN_SMALL_512x6 = 1 << 3
SEGPLAN_EACH = 1 << 4

MODEL_LINEAR_ROOT = 1 << 8
MLP_PARAMS = 1153


class KeyDesc(ctypes.Structure):
    _fields_ = [("offset", ctypes.c_int64), ("width", ctypes.c_int32),
                ("is_signed", ctypes.c_int32), ("nwords", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [
        ("t_first_sort_ms", ctypes.c_double), ("t_findp_ms", ctypes.c_double),
        ("t_laysolve_ms", ctypes.c_double), ("t_exchange_ms", ctypes.c_double),
        ("t_second_sort_ms", ctypes.c_double), ("t_total_ms", ctypes.c_double),
        ("t_h2d_ms", ctypes.c_double), ("t_d2h_ms", ctypes.c_double),
        ("nvranks", ctypes.c_int32), ("record_bytes", ctypes.c_int32),
        ("first_passes", ctypes.c_int32), ("second_passes", ctypes.c_int32),
        ("bisect_rounds", ctypes.c_int32), ("general_path", ctypes.c_int32),
        ("nrecords", ctypes.c_int64), ("checksum_in", ctypes.c_int64),
        ("checksum_out", ctypes.c_int64), ("npass_ms", ctypes.c_int32),
        ("learned", ctypes.c_int32), ("pass_ms", ctypes.c_double * 64),
        ("splitters_exact", ctypes.c_int32), ("splitters_repaired", ctypes.c_int32),
        ("buckets", ctypes.c_int32), ("buckets_overflow", ctypes.c_int32),
        ("repair_rounds", ctypes.c_int32), ("plain_sides", ctypes.c_int32),
        ("t_model_ms", ctypes.c_double), ("t_bucket_ms", ctypes.c_double),
        ("buckets_lds_radix", ctypes.c_int32), ("buckets_overflow_first", ctypes.c_int32),
        ("buckets_listed", ctypes.c_int32), ("exchange_sparse", ctypes.c_int32),
        ("exchange_runs", ctypes.c_int32), ("second_merged", ctypes.c_int32),
    ]

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            if name == "pass_ms":
                v = [self.pass_ms[i] for i in range(self.npass_ms)]
            d[name] = v
        return d


class ModelInfo(ctypes.Structure):
    _fields_ = [("nsample", ctypes.c_int64), ("leaves", ctypes.c_int32), ("linear_root", ctypes.c_int32),
                ("steps", ctypes.c_int32), ("nheavy", ctypes.c_int32), ("max_leaf_samples", ctypes.c_int32),
                ("log_input", ctypes.c_int32), ("train_ms", ctypes.c_double), ("loss", ctypes.c_float),
                ("err_max", ctypes.c_float), ("kmin", ctypes.c_uint64), ("kmax", ctypes.c_uint64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


RADIX_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)


def _load():
    if not os.path.exists(LIBPATH):
        raise ImportError("libmpsort_gpu.so not found at %s; build it with "
                          "`make -C mp-sort_amd` (hipcc, gfx950)" % LIBPATH)
    lib = ctypes.CDLL(LIBPATH)
    P = ctypes.POINTER
    i64p = P(ctypes.c_int64)
    lib.mpsort_gpu_newarray.argtypes = [ctypes.c_void_p, i64p, ctypes.c_int, ctypes.c_void_p, i64p,
                                        ctypes.c_size_t, P(KeyDesc), ctypes.c_uint]
    lib.mpsort_gpu_newarray.restype = ctypes.c_int
    lib.mpsort_gpu_newarray_noalloc.argtypes = [ctypes.c_void_p, i64p, ctypes.c_int, ctypes.c_void_p, i64p,
                                                ctypes.c_size_t, P(KeyDesc), ctypes.c_uint]
    lib.mpsort_gpu_newarray_noalloc.restype = ctypes.c_int
    lib.mpsort_gpu_newarray_radix.argtypes = [ctypes.c_void_p, i64p, ctypes.c_int, ctypes.c_void_p,
                                              i64p, ctypes.c_size_t, RADIX_FN, ctypes.c_size_t,
                                              ctypes.c_void_p, ctypes.c_uint]
    lib.mpsort_gpu_newarray_radix.restype = ctypes.c_int
    lib.mpsort_gpu_last_stats.argtypes = [P(Stats)]
    lib.mpsort_gpu_set_faults.argtypes = [ctypes.c_uint, ctypes.c_int64]
    lib.mpsort_gpu_last_stats.restype = ctypes.c_int
    lib.mpsort_gpu_last_counts.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                           i64p, i64p, i64p, i64p, i64p, i64p]
    lib.mpsort_gpu_last_counts.restype = ctypes.c_int
    lib.mpsort_gpu_last_error.restype = ctypes.c_char_p
    lib.mpsort_gpu_version.restype = ctypes.c_char_p
    lib.mpsort_gpu_dmalloc.argtypes = [ctypes.c_size_t]
    lib.mpsort_gpu_dmalloc.restype = ctypes.c_void_p
    lib.mpsort_gpu_dfree.argtypes = [ctypes.c_void_p]
    lib.mpsort_gpu_memcpy_h2d.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    lib.mpsort_gpu_memcpy_d2h.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    lib.mpsort_gpu_dcopy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    vp = ctypes.c_void_p
    lib.mpsort_gpu_model_train.argtypes = [P(vp), vp, ctypes.c_int64, ctypes.c_size_t, P(KeyDesc),
                                           ctypes.c_int64, ctypes.c_uint64, ctypes.c_uint]
    lib.mpsort_gpu_model_free.argtypes = [vp]
    lib.mpsort_gpu_model_get_info.argtypes = [vp, P(ModelInfo)]
    lib.mpsort_gpu_model_eval.argtypes = [vp, vp, ctypes.c_int64, vp]
    lib.mpsort_gpu_model_params.argtypes = [vp, vp, vp]
    lib.mpsort_gpu_mlp_grad.argtypes = [vp, vp, vp, ctypes.c_int64, vp, P(ctypes.c_float)]
    lib.mpsort_gpu_newarray_learned.argtypes = [vp, i64p, ctypes.c_int, vp, i64p, ctypes.c_size_t,
                                                P(KeyDesc), vp, ctypes.c_uint]
    lib.mpsort_gpu_segplan_create.argtypes = [P(ctypes.c_void_p), ctypes.c_uint]
    lib.mpsort_gpu_segplan_add.argtypes = [vp, vp, vp, i64p, ctypes.c_int, ctypes.c_size_t, P(KeyDesc)]
    lib.mpsort_gpu_segplan_run.argtypes = [vp, P(Stats)]
    lib.mpsort_gpu_segplan_info.argtypes = [vp, ctypes.c_int] + [P(ctypes.c_int)] * 5
    lib.mpsort_gpu_segplan_free.argtypes = [vp]
    lib.mpsort_gpu_sort_segments.argtypes = [vp, vp, i64p, ctypes.c_int, ctypes.c_size_t, P(KeyDesc),
                                             ctypes.c_uint]
    lib.mpsort_gpu_set_vranks.argtypes = [ctypes.c_int]
    lib.mpsort_gpu_get_vranks.restype = ctypes.c_int
    lib.mpsort_mpi_set_options.argtypes = [ctypes.c_int]
    lib.mpsort_mpi_has_options.argtypes = [ctypes.c_int]
    lib.mpsort_mpi_has_options.restype = ctypes.c_int
    lib.mpsort_mpi_unset_options.argtypes = [ctypes.c_int]
    return lib


lib = _load()


class MPSortError(RuntimeError):
    def __init__(self, code, msg):
        RuntimeError.__init__(self, "mpsort_gpu error %d: %s" % (code, msg))
        self.code = code


def check(rc):
    """Turn a C return code into the exception the reference binding raises
    for the same condition (binding.pyx:157-189), or MPSortError."""
    if rc == MPSORT_OK:
        return
    msg = lib.mpsort_gpu_last_error().decode(errors="replace")
    if rc == MPSORT_ERR_SIZE_MISMATCH or rc == MPSORT_ERR_ARG:
        raise ValueError(msg)
    if rc == MPSORT_ERR_UNSUPPORTED:
        raise TypeError(msg)
    raise MPSortError(rc, msg)


def last_stats():
    s = Stats()
    check(lib.mpsort_gpu_last_stats(ctypes.byref(s)))
    return s


def device_count():
    return int(lib.mpsort_gpu_device_count())
