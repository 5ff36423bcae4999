"""harmonia_amd -- MI355X-native Harmonia GPU B+tree (HIP kernels behind a C ABI).

The compute path is libharmonia_amd.so in this directory (built by `make -C harmonia_amd`);
importing the package does not load it, `BPlusTree(...)` does and fails loudly if it is missing.
"""
from ._lib import HarmoniaError, HM_KEY_MAX, LIB_PATH, load  # noqa: F401
from .tree import Batcher, BPlusTree, Server, psa_sort, read_keys_text, read_ranges_text  # noqa: F401
from . import workloads  # noqa: F401

__all__ = ["BPlusTree", "Batcher", "Server", "psa_sort", "read_keys_text", "read_ranges_text", "HarmoniaError",
           "HM_KEY_MAX", "load", "workloads"]
