"""ctypes binding of libharmonia_amd.so (the C ABI declared in include/harmonia.h).

The library is loaded from this package directory only; a missing library raises
immediately -- there is no CPU fallback for any operation.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libharmonia_amd.so")

HM_OK = 0
HM_EINVAL = -1
HM_ENOMEM = -2
HM_EBUDGET = -3
HM_EDEVICE = -4
HM_KEY_MAX = 0xFFFFFFFFFFFFFFFF
HM_FLAG_KEYS_SORTED = 1
HM_OUT_DEFAULT = 0
HM_OUT_UNPERMUTE = 1
HM_OUT_SCATTER = 2
HM_OUT_PSA_ORDER = 3

u8p = ctypes.POINTER(ctypes.c_uint8)
u32p = ctypes.POINTER(ctypes.c_uint32)
u64p = ctypes.POINTER(ctypes.c_uint64)
vp = ctypes.c_void_p


class HmParams(ctypes.Structure):
    _fields_ = [
        ("fanout", ctypes.c_uint32),
        ("leaf_fill", ctypes.c_uint32),
        ("lds_levels", ctypes.c_int32),
        ("group_size", ctypes.c_uint32),
        ("psa_bits", ctypes.c_uint32),
        ("reserved0", ctypes.c_uint32),
        ("sort_threshold", ctypes.c_uint64),
        ("mem_budget", ctypes.c_uint64),
        ("flags", ctypes.c_uint32),
        ("reserved1", ctypes.c_uint32),
    ]


class HmStats(ctypes.Structure):
    _fields_ = [
        ("n_keys", ctypes.c_uint64),
        ("n_leaves", ctypes.c_uint64),
        ("n_inner", ctypes.c_uint64),
        ("height", ctypes.c_uint32),
        ("fanout", ctypes.c_uint32),
        ("node_slots", ctypes.c_uint32),
        ("lds_levels", ctypes.c_uint32),
        ("level_start", ctypes.c_uint64 * 17),
        ("bytes_inner", ctypes.c_uint64),
        ("bytes_prefix", ctypes.c_uint64),
        ("bytes_leaves", ctypes.c_uint64),
        ("bytes_device", ctypes.c_uint64),
        ("epoch", ctypes.c_uint64),
    ]


class HmSearchOpts(ctypes.Structure):
    _fields_ = [
        ("group_size", ctypes.c_uint32),
        ("lds_levels", ctypes.c_int32),
        ("psa_bits", ctypes.c_int32),
        ("blocks_per_cu", ctypes.c_uint32),
        ("output_mode", ctypes.c_uint32),
        ("queries_per_group", ctypes.c_uint32),
        ("out_pos", vp),
    ]


# (name, restype, argtypes) for every entry point of include/harmonia.h
SIGNATURES = [
    ("hm_params_init", ctypes.c_int, [ctypes.POINTER(HmParams)]),
    ("hm_tree_build", ctypes.c_int, [vp, vp, ctypes.c_uint64, ctypes.POINTER(HmParams), ctypes.POINTER(vp)]),
    ("hm_tree_destroy", ctypes.c_int, [vp]),
    ("hm_search", ctypes.c_int, [vp, vp, ctypes.c_uint64, vp, vp, vp]),
    ("hm_search_ex", ctypes.c_int, [vp, vp, ctypes.c_uint64, vp, vp, ctypes.POINTER(HmSearchOpts), vp]),
    ("hm_range", ctypes.c_int, [vp, vp, vp, ctypes.c_uint64, vp, vp, vp, vp, ctypes.c_uint64, vp]),
    ("hm_range_between", ctypes.c_int, [vp, vp, vp, ctypes.c_uint64, ctypes.c_uint32, vp, vp, vp, vp]),
    ("hm_update", ctypes.c_int, [vp, vp, vp, ctypes.c_uint64, vp]),
    ("hm_epoch_publish", ctypes.c_int, [vp, vp, ctypes.POINTER(ctypes.c_int)]),
    ("hm_tree_stats", ctypes.c_int, [vp, ctypes.POINTER(HmStats)]),
    ("hm_tree_export", ctypes.c_int, [vp, vp, vp, vp]),
    ("hm_psa_sort", ctypes.c_int, [vp, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, vp, vp, vp]),
    ("hm_profile_enable", ctypes.c_int, [ctypes.c_int]),
    ("hm_profile_reset", ctypes.c_int, []),
    ("hm_profile_get", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), u64p]),
    ("hm_device_synchronize", ctypes.c_int, []),
    ("hm_last_error", ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint64]),
]

_lib = None


class HarmoniaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"harmonia error {code}: {msg}")
        self.code = code


def load():
    """Load libharmonia_amd.so from the package directory (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `make -C harmonia_amd` (or __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error():
    buf = ctypes.create_string_buffer(512)
    load().hm_last_error(buf, 512)
    return buf.value.decode(errors="replace")


def check(rc):
    if rc != HM_OK:
        raise HarmoniaError(rc, last_error())
    return rc
