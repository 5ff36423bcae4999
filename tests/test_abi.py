"""The C ABI library loads and exports every entry point include/harmonia.h declares; host-only
argument checks return HM_EINVAL without touching a device.  No compute call runs here."""
import ctypes
import os
import re

import pytest

import harmonia_amd
from harmonia_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "harmonia.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|uint64_t)\s+(hm_\w+)\s*\(", src, flags=re.M)))


def test_library_is_built():
    assert os.path.exists(_lib.LIB_PATH), "run make -C harmonia_amd"


def test_every_declared_symbol_is_exported():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name


def test_binding_covers_header():
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert set(declared_functions()) == bound


def test_struct_sizes_match_header_layout():
    assert ctypes.sizeof(_lib.HmParams) == 48
    assert ctypes.sizeof(_lib.HmSearchOpts) == 40
    assert ctypes.sizeof(_lib.HmStats) == 8 * 3 + 4 * 4 + 8 * 17 + 8 * 7 + 8 * 2 + 4 * 2
    assert ctypes.sizeof(_lib.HmServerStats) == 8 * 8 + 8 * 5 + 8 * 4 + 8 * 3


def test_params_defaults():
    lib = _lib.load()
    p = _lib.HmParams()
    assert lib.hm_params_init(ctypes.byref(p)) == 0
    assert (p.fanout, p.leaf_fill, p.lds_levels, p.psa_bits) == (64, 8, -1, 16)


@pytest.mark.parametrize("field,value", [("fanout", 48), ("leaf_fill", 0), ("leaf_fill", 9), ("lds_levels", 5),
                                         ("group_size", 3), ("psa_bits", 40)])
def test_bad_params_rejected_before_device(field, value):
    lib = _lib.load()
    p = _lib.HmParams()
    lib.hm_params_init(ctypes.byref(p))
    setattr(p, field, value)
    h = ctypes.c_void_p()
    assert lib.hm_tree_build(None, None, 0, ctypes.byref(p), ctypes.byref(h)) == _lib.HM_EINVAL
    assert h.value is None
    assert _lib.last_error()


def test_null_tree_rejected():
    lib = _lib.load()
    assert lib.hm_search(None, None, 1, None, None, None) == _lib.HM_EINVAL
    assert lib.hm_tree_destroy(None) == _lib.HM_EINVAL
    assert lib.hm_update(None, None, None, 1, None) == _lib.HM_EINVAL
    assert lib.hm_epoch_publish(None, None, None) == _lib.HM_EINVAL
    assert "NULL" in _lib.last_error()


def test_package_exports():
    assert harmonia_amd.HM_KEY_MAX == 0xFFFFFFFFFFFFFFFF
    assert callable(harmonia_amd.BPlusTree) and callable(harmonia_amd.psa_sort)
