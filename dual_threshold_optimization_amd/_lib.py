"""ctypes binding of libdto_amd.so (include/dto.h).

The library is built in-tree by ``__graft_entry__.build()`` (or ``make -C
dual_threshold_optimization_amd/csrc``).  Loading fails loudly when it is
missing: there is no CPU fallback for any compute entry point.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdto_amd.so")
CLI_PATH = os.path.join(_HERE, "dto_amd")

DTO_OK = 0
DTO_ERR_INVALID_ARGUMENT = 1
DTO_ERR_POPULATION = 2
DTO_ERR_HYPERGEOMETRIC = 3
DTO_ERR_EMPTY_GRID = 4
DTO_ERR_NO_DEVICE = 5
DTO_ERR_DEVICE = 6
DTO_ERR_BAD_PERMUTATION = 7
DTO_ERR_TOO_LARGE = 8
DTO_ERR_IO = 9
DTO_ERR_PARSE = 10
DTO_ERR_BUSY = 11

DTO_DROP_OUTSIDE_BACKGROUND = 0x1
DTO_DEFER_UPLOAD = 0x2
DTO_NOALLOC_MAX_CELLS = 1 << 22
DTO_MODE_SCREENED = 0
DTO_MODE_PLAIN = 1
MODES = {"screened": DTO_MODE_SCREENED, "plain": DTO_MODE_PLAIN}


class DtoError(RuntimeError):
    def __init__(self, code, message):
        super().__init__("dto error %d: %s" % (code, message))
        self.code = code
        self.message = message


class Record(ctypes.Structure):
    """dto_record (include/dto.h)."""
    _fields_ = [
        ("rank1", ctypes.c_uint64),
        ("rank2", ctypes.c_uint64),
        ("set1_len", ctypes.c_uint64),
        ("set2_len", ctypes.c_uint64),
        ("population_size", ctypes.c_uint64),
        ("intersection_size", ctypes.c_uint64),
        ("pvalue", ctypes.c_double),
        ("permuted", ctypes.c_int32),
        ("idx1", ctypes.c_uint32),
        ("idx2", ctypes.c_uint32),
        ("ties", ctypes.c_uint32),
    ]


RECORD_DTYPE = np.dtype([
    ("rank1", "<u8"), ("rank2", "<u8"), ("set1_len", "<u8"), ("set2_len", "<u8"),
    ("population_size", "<u8"), ("intersection_size", "<u8"), ("pvalue", "<f8"),
    ("permuted", "<i4"), ("idx1", "<u4"), ("idx2", "<u4"), ("ties", "<u4"),
])
assert RECORD_DTYPE.itemsize == ctypes.sizeof(Record) == 72


class Result(ctypes.Structure):
    """dto_result (include/dto.h)."""
    _fields_ = [
        ("unpermuted", Record),
        ("empirical_pvalue", ctypes.c_double),
        ("fdr", ctypes.c_double),
        ("n_permutations", ctypes.c_uint64),
        ("n_le", ctypes.c_uint64),
    ]


class JobInfo(ctypes.Structure):
    _fields_ = [
        ("population_size", ctypes.c_uint64),
        ("len1", ctypes.c_uint64),
        ("len2", ctypes.c_uint64),
        ("n_thresholds1", ctypes.c_uint32),
        ("n_thresholds2", ctypes.c_uint32),
        ("dropped1", ctypes.c_uint64),
        ("dropped2", ctypes.c_uint64),
        ("device_bytes", ctypes.c_uint64),
    ]


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("hbm_budget_bytes", ctypes.c_uint64),
        ("max_active_jobs", ctypes.c_uint32),
        ("perms_per_slice", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
    ]


# every entry point include/dto.h declares, with (restype, argtypes)
_u64, _u32, _i32, _dbl, _int, _vp, _cp = (ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int32, ctypes.c_double,
                                          ctypes.c_int, ctypes.c_void_p, ctypes.c_char_p)
_P = ctypes.POINTER
SIGNATURES = {
    "dto_last_error": (_cp, []),
    "dto_device_count": (_int, []),
    "dto_hypergeometric_pvalue": (_int, [_P(_u64), _P(_u64), _P(_u64), _P(_u64), _u64, _P(_dbl)]),
    "dto_fdr": (_dbl, [_u64, _u64, _u64, _u64, _dbl]),
    "dto_fdr_checked": (_int, [_u64, _u64, _u64, _u64, _dbl, _P(_dbl)]),
    "dto_generate_thresholds": (_u32, [_u32, _P(_u32), _u32]),
    "dto_ln_factorial_table": (_int, [_u64, _P(_dbl)]),
    "dto_job_create": (_int, [_P(_cp), _P(_u32), _u64, _P(_cp), _P(_u32), _u64, _P(_cp), _u64, _u32, _P(_vp)]),
    "dto_job_create_with_population": (_int, [_P(_cp), _P(_u32), _u64, _P(_cp), _P(_u32), _u64, _u64, _P(_vp)]),
    "dto_job_create_from_files": (_int, [_cp, _cp, _cp, _u32, _P(_vp)]),
    "dto_job_set_thresholds": (_int, [_vp, _P(_u32), _u32, _P(_u32), _u32]),
    "dto_job_info_get": (_int, [_vp, _P(JobInfo)]),
    "dto_job_thresholds": (_int, [_vp, _int, _P(_u32), _u32]),
    "dto_job_destroy": (None, [_vp]),
    "dto_optimize": (_int, [_vp, _P(_u32), _P(_u32), _u32, _P(Record)]),
    "dto_process_threshold_pairs": (_int, [_vp, _P(_u32), _P(_u32), _vp]),
    "dto_process_threshold_pairs_noalloc": (_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "dto_run": (_int, [_vp, _u64, _u64, _P(_u32), _P(_u32), _u32, _P(Result), _vp]),
    "dto_run_device": (_int, [_vp, _u64, _u64, _u64, _vp, _vp, _u32, _vp, _vp]),
    "dto_philox_permutation": (_int, [_u64, _u32, _u32, _u32, _u32, _P(_u32)]),
    "dto_empirical_pvalue": (_int, [_P(Record), _vp, _u64, _P(Result)]),
    "dto_result_json": (_int, [_P(Result), ctypes.c_char_p, ctypes.c_size_t]),
    "dto_batch_create": (_int, [_P(_vp), _P(_u64), _u64, _u64, _P(_vp)]),
    "dto_batch_launch": (_int, [_vp, _vp]),
    "dto_batch_results": (_int, [_vp, _P(Result)]),
    "dto_batch_destroy": (None, [_vp]),
    "dto_batch_autotune": (_int, [_vp, _P(_int)]),
    "dto_server_create": (_int, [_P(ServerConfig), _P(_vp)]),
    "dto_server_submit": (_int, [_vp, _vp, _u64, _P(_u64)]),
    "dto_server_poll": (_int, [_vp, _u32, _P(_u64), _P(Result), _P(_dbl), _u32, _P(_u32)]),
    "dto_server_pending": (_int, [_vp, _P(_u64), _P(_u64)]),
    "dto_server_destroy": (None, [_vp]),
}
# test hooks exported by the library (not part of include/dto.h)
INTERNAL = {
    "dto_internal_exp": (_int, [_P(_dbl), _u64, _P(_dbl)]),
    "dto_internal_exp_host": (_int, [_P(_dbl), _u64, _P(_dbl)]),
    "dto_internal_shuffle_device": (_int, [_u64, _u32, _u32, _u32, _u32, _P(_u32)]),
    "dto_internal_stats": (_int, [_vp, _P(_u64), _int]),
    "dto_internal_philox": (_int, [_P(_u32), _P(_u32), _P(_u32)]),
    "dto_internal_read_ranked_csv": (_int, [_cp, _vp, _u64, _P(_u64), _P(_u32), _u64, _P(_u64)]),
    "dto_internal_read_feature_lines": (_int, [_cp, _vp, _u64, _P(_u64), _P(_u64)]),
}

_lib = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libdto_amd.so is not built (%s): run __graft_entry__.build() or "
                               "make -C dual_threshold_optimization_amd/csrc" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in list(SIGNATURES.items()) + list(INTERNAL.items()):
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != DTO_OK:
        raise DtoError(rc, lib().dto_last_error().decode(errors="replace"))
    return rc


def read_ranked_csv_native(path):
    """The library's CSV reader alone (test hook): (ids as bytes, ranks)."""
    L, n, used = lib(), ctypes.c_uint64(), ctypes.c_uint64()
    check(L.dto_internal_read_ranked_csv(os.fsencode(path), None, 0, ctypes.byref(used), None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(max(used.value, 1))
    ranks = np.zeros(max(n.value, 1), dtype=np.uint32)
    check(L.dto_internal_read_ranked_csv(os.fsencode(path), buf, used.value, ctypes.byref(used), u32p(ranks), n.value,
                                         ctypes.byref(n)))
    ids = buf.raw[:used.value].split(b"\0")[:-1] if used.value else []
    return ids, ranks[:n.value]


def read_feature_lines_native(path):
    """The library's background reader alone (test hook): ids as bytes."""
    L, n, used = lib(), ctypes.c_uint64(), ctypes.c_uint64()
    check(L.dto_internal_read_feature_lines(os.fsencode(path), None, 0, ctypes.byref(used), ctypes.byref(n)))
    buf = ctypes.create_string_buffer(max(used.value, 1))
    check(L.dto_internal_read_feature_lines(os.fsencode(path), buf, used.value, ctypes.byref(used), ctypes.byref(n)))
    return buf.raw[:used.value].split(b"\0")[:-1] if used.value else []


def device_count():
    return lib().dto_device_count()


def require_device():
    if device_count() <= 0:
        raise DtoError(DTO_ERR_NO_DEVICE, "no HIP device: the DTO kernels run on an MI355X (gfx950) only")


def u32p(a):
    return None if a is None else a.ctypes.data_as(_P(_u32))


def u64p(a):
    return None if a is None else a.ctypes.data_as(_P(_u64))


def dblp(a):
    return None if a is None else a.ctypes.data_as(_P(_dbl))


def cstr_array(ids):
    arr = (ctypes.c_char_p * max(len(ids), 1))()
    for i, s in enumerate(ids):
        arr[i] = s.encode() if isinstance(s, str) else s
    return arr


def idx_array(idx, length, n_perm=None):
    if idx is None:
        return None
    a = np.ascontiguousarray(idx, dtype=np.uint32)
    if n_perm is None:
        if a.shape != (length,):
            raise ValueError("permutation must have shape (%d,)" % length)
    elif a.shape != (n_perm, length):
        raise ValueError("permutations must have shape (%d, %d)" % (n_perm, length))
    return a
