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
HM_BASELINE_POINTER_CHILD = 1
HM_BASELINE_POINTER_LINEAR = 2
HM_BASELINE_PREFIX_PLAIN = 3
HM_BATCH_DEVICE = 1
HM_SORTED_AUTO = 0
HM_SORTED_QUERY = 1
HM_SORTED_WINDOW = 2

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
        ("relayouts", ctypes.c_uint64),
        ("respreads", ctypes.c_uint64),
        ("key_min", ctypes.c_uint64),
        ("key_max", ctypes.c_uint64),
        ("psa_window_top", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
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
        ("sorted_kernel", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class HmTuneResult(ctypes.Structure):
    _fields_ = [
        ("fanout", ctypes.c_uint32),
        ("direct_group", ctypes.c_uint32),
        ("direct_queries_per_group", ctypes.c_uint32),
        ("direct_blocks_per_cu", ctypes.c_uint32),
        ("lds_levels", ctypes.c_int32),
        ("psa_group", ctypes.c_uint32),
        ("psa_queries_per_group", ctypes.c_uint32),
        ("psa_blocks_per_cu", ctypes.c_uint32),
        ("psa_bits", ctypes.c_uint32),
        ("output_mode", ctypes.c_uint32),
        ("sort_threshold", ctypes.c_uint64),
        ("direct_ms", ctypes.c_double),
        ("psa_ms", ctypes.c_double),
        ("best_ms", ctypes.c_double),
        ("trials", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


HM_TUNE_QUICK = 1


class HmServerParams(ctypes.Structure):
    _fields_ = [
        ("mem_budget", ctypes.c_uint64),
        ("max_batch", ctypes.c_uint32),
        ("max_inflight", ctypes.c_uint32),
        ("coalesce", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class HmServerStats(ctypes.Structure):
    _fields_ = [
        ("completed", ctypes.c_uint64),
        ("lookups", ctypes.c_uint64),
        ("ranges", ctypes.c_uint64),
        ("updates", ctypes.c_uint64),
        ("queries", ctypes.c_uint64),
        ("rejected", ctypes.c_uint64),
        ("launches", ctypes.c_uint64),
        ("epochs", ctypes.c_uint64),
        ("p50_us", ctypes.c_double),
        ("p99_us", ctypes.c_double),
        ("mean_us", ctypes.c_double),
        ("max_us", ctypes.c_double),
        ("span_s", ctypes.c_double),
        ("max_read_batch_us", ctypes.c_double),
        ("max_update_us", ctypes.c_double),
        ("max_prepare_us", ctypes.c_double),
        ("max_commit_wait_us", ctypes.c_double),
        ("p99_lookup_us", ctypes.c_double),
        ("p99_range_us", ctypes.c_double),
        ("p99_update_us", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}

# (name, restype, argtypes) for every entry point of include/harmonia.h
SIGNATURES = [
    ("hm_params_init", ctypes.c_int, [ctypes.POINTER(HmParams)]),
    ("hm_tree_build", ctypes.c_int, [vp, vp, ctypes.c_uint64, ctypes.POINTER(HmParams), ctypes.POINTER(vp)]),
    ("hm_tree_destroy", ctypes.c_int, [vp]),
    ("hm_search", ctypes.c_int, [vp, vp, ctypes.c_uint64, vp, vp, vp]),
    ("hm_search_ex", ctypes.c_int, [vp, vp, ctypes.c_uint64, vp, vp, ctypes.POINTER(HmSearchOpts), vp]),
    ("hm_search_noalloc", ctypes.c_int, [vp, vp, ctypes.c_uint64, vp, vp, vp]),
    ("hm_search_noalloc_max", ctypes.c_uint64, []),
    ("hm_baseline_search", ctypes.c_int, [vp, ctypes.c_int, ctypes.c_uint32, vp, ctypes.c_uint64, vp, vp, vp]),
    ("hm_range", ctypes.c_int, [vp, vp, vp, ctypes.c_uint64, vp, vp, vp, vp, ctypes.c_uint64, vp]),
    ("hm_range_between", ctypes.c_int, [vp, vp, vp, ctypes.c_uint64, ctypes.c_uint32, vp, vp, vp, vp]),
    ("hm_update", ctypes.c_int, [vp, vp, vp, ctypes.c_uint64, vp]),
    ("hm_epoch_publish", ctypes.c_int, [vp, vp, ctypes.POINTER(ctypes.c_int)]),
    ("hm_batcher_create", ctypes.c_int, [vp, ctypes.c_uint64, ctypes.POINTER(HmSearchOpts), vp, ctypes.POINTER(vp)]),
    ("hm_batcher_add", ctypes.c_int, [vp, vp, ctypes.c_uint64, u64p]),
    ("hm_batcher_add_many", ctypes.c_int, [vp, vp, vp, ctypes.c_uint64, ctypes.c_uint32, vp, u64p]),
    ("hm_batcher_flush", ctypes.c_int, [vp]),
    ("hm_batcher_results", ctypes.c_int, [vp, ctypes.POINTER(vp), ctypes.POINTER(vp)]),
    ("hm_batcher_copy_out", ctypes.c_int, [vp, ctypes.c_uint64, ctypes.c_uint64, vp, vp]),
    ("hm_batcher_sync", ctypes.c_int, [vp]),
    ("hm_batcher_destroy", ctypes.c_int, [vp]),
    ("hm_server_create", ctypes.c_int, [vp, ctypes.POINTER(HmServerParams), ctypes.POINTER(vp)]),
    ("hm_server_lookup", ctypes.c_int, [vp, vp, ctypes.c_uint32, vp, vp, u64p]),
    ("hm_server_range", ctypes.c_int, [vp, vp, vp, ctypes.c_uint32, vp, vp, vp, ctypes.c_uint64, u64p]),
    ("hm_server_update", ctypes.c_int, [vp, vp, vp, ctypes.c_uint32, u64p]),
    ("hm_server_wait", ctypes.c_int, [vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int), u64p]),
    ("hm_server_drain", ctypes.c_int, [vp]),
    ("hm_server_get_stats", ctypes.c_int, [vp, ctypes.POINTER(HmServerStats)]),
    ("hm_server_reset_stats", ctypes.c_int, [vp]),
    ("hm_server_destroy", ctypes.c_int, [vp]),
    ("hm_autotune", ctypes.c_int, [vp, vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(HmTuneResult)]),
    ("hm_tree_build_tuned", ctypes.c_int, [vp, vp, ctypes.c_uint64, ctypes.POINTER(HmParams), vp, ctypes.c_uint64,
                                           ctypes.c_uint32, ctypes.POINTER(vp), ctypes.POINTER(HmTuneResult)]),
    ("hm_tree_stats", ctypes.c_int, [vp, ctypes.POINTER(HmStats)]),
    ("hm_tree_export", ctypes.c_int, [vp, vp, vp, vp]),
    ("hm_read_keys_text", ctypes.c_int, [ctypes.c_char_p, vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
    ("hm_read_ranges_text", ctypes.c_int,
     [ctypes.c_char_p, vp, vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
    ("hm_tree_save", ctypes.c_int, [vp, ctypes.c_char_p]),
    ("hm_tree_load", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(HmParams), ctypes.POINTER(vp)]),
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
