"""CPU oracle for the Harmonia B+tree path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this package,
and only as the checker (or as the timed CPU reference path).  The product never uses it.

liboracle.so (oracle.c, plain C) restates the reference algorithms; see oracle.c's header for the
reference file:line each function follows.  oracle/_ref/ holds the reference's own CPU sources
compiled in place (regular_bpt and the Harmonia BPlusTree) where the reference tree is present;
those libraries pin the restatement.

Parity pinning (what checks each function; details in tests/test_oracle.py):
  build_bfs      structural invariants (check_layout), brute-force BFS of tiny trees, and the
                 device build compared slot for slot
  search_cfb     regular_bpt find + Harmonia BPlusTree::search (oracle/_ref), golden vectors,
                 closed form search_sorted, brute force
  search_harmonia the same checks as search_cfb (reference BPlusTree::search and regular find via
                 oracle/_ref, golden vectors, closed form), the AVX2 and scalar paths against each
                 other, and a planted index-step mistake that those checks catch
  search_sorted  brute force (python set/dict) on tiny inputs
  psa_sort       python's stable sorted() on tiny inputs, permutation + stability invariants
  range_walk     regular find_range + BPlusTree::range_search (oracle/_ref), golden vectors,
                 closed form range_sorted
  upsert_sorted  dict semantics on tiny inputs, regular insert (oracle/_ref) for new keys
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_DIR = os.path.join(_HERE, "_ref")

_u64 = np.uint64
_vp = ctypes.c_void_p
_lib = None


def build_lib():
    import subprocess
    subprocess.check_call(["make", "-s", "-B", "-C", _HERE, "all"])


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build_lib()
    lib = ctypes.CDLL(LIB_PATH)
    sig = {
        "hmo_splitmix64_mix": (ctypes.c_uint64, [ctypes.c_uint64]),
        "hmo_plan": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp]),
        "hmo_build_bfs": (ctypes.c_int, [_vp, _vp, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp,
                                         _vp, _vp, _vp]),
        "hmo_search_cfb": (None, [_vp, _vp, _vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _vp,
                                  ctypes.c_uint64, _vp, _vp, ctypes.c_int]),
        "hmo_search_harmonia": (None, [_vp, _vp, _vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _vp,
                                       ctypes.c_uint64, _vp, _vp, ctypes.c_int, ctypes.c_int]),
        "hmo_have_avx2": (ctypes.c_int, []),
        "hmo_search_sorted": (None, [_vp, _vp, ctypes.c_uint64, _vp, ctypes.c_uint64, _vp, _vp]),
        "hmo_psa_sort": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, _vp, _vp]),
        "hmo_range_walk": (None, [_vp, _vp, _vp, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int,
                                  _vp, _vp, _vp, ctypes.c_uint32, ctypes.c_uint64, _vp, _vp, _vp, _vp]),
        "hmo_range_sorted": (None, [_vp, _vp, ctypes.c_uint64, _vp, _vp, _vp, ctypes.c_uint32, ctypes.c_uint64,
                                    _vp, _vp, _vp, _vp]),
        "hmo_upsert_sorted": (ctypes.c_uint64, [_vp, _vp, ctypes.c_uint64, _vp, _vp, ctypes.c_uint64, _vp, _vp]),
        "hmo_check_layout": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32, _vp,
                                            ctypes.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


class Layout:
    """Host copy of a tree layout (inner, prefix, leaves) with its shape."""

    def __init__(self, inner, prefix, leaves, height, fanout, level_start):
        self.inner = _c(inner, _u64)
        self.prefix = _c(prefix, np.uint32)
        self.leaves = _c(leaves, _u64)
        self.height = int(height)
        self.fanout = int(fanout)
        self.level_start = _c(level_start, np.uint32)

    @property
    def n_inner(self):
        return int(self.prefix.shape[0])

    @property
    def n_leaves(self):
        return int(self.leaves.shape[0])


def node_slots(F):
    return max(8, F // 8) + F


def plan(n, F, fill):
    lib = load()
    ls = np.zeros(17, dtype=np.uint32)
    ni = np.zeros(1, dtype=np.uint32)
    nl = np.zeros(1, dtype=_u64)
    h = lib.hmo_plan(n, F, fill, _p(ls), _p(ni), _p(nl))
    if h < 0:
        raise ValueError("bad plan parameters")
    return h, int(ni[0]), int(nl[0]), ls


def build_bfs(sorted_keys, vals, F=64, fill=8):
    lib = load()
    k = _c(sorted_keys, _u64)
    v = _c(vals, _u64)
    h, ni, nl, _ = plan(len(k), F, fill)
    inner = np.empty((ni, node_slots(F)), dtype=_u64)
    prefix = np.empty(ni, dtype=np.uint32)
    leaves = np.empty((nl, 16), dtype=_u64)
    ls = np.zeros(17, dtype=np.uint32)
    oh = np.zeros(1, dtype=np.uint32)
    oni = np.zeros(1, dtype=np.uint32)
    onl = np.zeros(1, dtype=_u64)
    rc = lib.hmo_build_bfs(_p(k), _p(v), len(k), F, fill, _p(inner), _p(prefix), _p(leaves), _p(ls), _p(oh),
                           _p(oni), _p(onl))
    if rc:
        raise ValueError("hmo_build_bfs failed")
    return Layout(inner, prefix, leaves, oh[0], F, ls)


def search_cfb(layout, q, nthreads=1):
    lib = load()
    q = _c(q, _u64)
    vals = np.empty(len(q), dtype=_u64)
    found = np.empty(len(q), dtype=np.uint8)
    lib.hmo_search_cfb(_p(layout.inner), _p(layout.prefix), _p(layout.leaves), layout.n_inner, layout.height,
                       layout.fanout, _p(q), len(q), _p(vals), _p(found), nthreads)
    return vals, found


def search_harmonia(layout, q, nthreads=1, avx2=True):
    """Harmonia's BPlusTree::search steps (index-then-key, 4 + 4 compares; backward leaf scan) over
    the layout, OpenMP over queries as search_cpu_test does."""
    lib = load()
    q = _c(q, _u64)
    vals = np.empty(len(q), dtype=_u64)
    found = np.empty(len(q), dtype=np.uint8)
    lib.hmo_search_harmonia(_p(layout.inner), _p(layout.prefix), _p(layout.leaves), layout.n_inner, layout.height,
                            layout.fanout, _p(q), len(q), _p(vals), _p(found), nthreads, 1 if avx2 else 0)
    return vals, found


def have_avx2():
    return bool(load().hmo_have_avx2())


def search_sorted(sorted_keys, vals, q):
    lib = load()
    k, v, q = _c(sorted_keys, _u64), _c(vals, _u64), _c(q, _u64)
    ov = np.empty(len(q), dtype=_u64)
    of = np.empty(len(q), dtype=np.uint8)
    lib.hmo_search_sorted(_p(k), _p(v), len(k), _p(q), len(q), _p(ov), _p(of))
    return ov, of


def psa_sort(keys, begin_bit=48, end_bit=64):
    lib = load()
    k = _c(keys, _u64)
    ko = np.empty(len(k), dtype=_u64)
    po = np.empty(len(k), dtype=np.uint32)
    if lib.hmo_psa_sort(_p(k), len(k), begin_bit, end_bit, _p(ko), _p(po)):
        raise ValueError("bad bit range")
    return ko, po


def range_walk(layout, starts, lens=None, his=None, cap=0):
    """Leaf-chain range scan over a layout; returns (offsets, keys, vals, counts)."""
    lib = load()
    s = _c(starts, _u64)
    m = len(s)
    hi = None if his is None else _c(his, _u64)
    ln = None if lens is None else _c(lens, np.uint32)
    if ln is not None:
        offsets = np.zeros(m + 1, dtype=_u64)
        offsets[1:] = np.cumsum(ln.astype(_u64))
        total = int(offsets[-1])
    else:
        offsets = (np.arange(m + 1, dtype=_u64) * _u64(cap))
        total = m * cap
    ok = np.zeros(max(total, 1), dtype=_u64)
    ov = np.zeros(max(total, 1), dtype=_u64)
    oc = np.zeros(m, dtype=np.uint32)
    lib.hmo_range_walk(_p(layout.inner), _p(layout.prefix), _p(layout.leaves), layout.n_inner, layout.n_leaves,
                       layout.height, layout.fanout, _p(s), _p(hi), _p(ln), cap, m,
                       _p(offsets) if ln is not None else None, _p(ok), _p(ov), _p(oc))
    return offsets, ok, ov, oc


def range_sorted(sorted_keys, vals, starts, lens=None, his=None, cap=0):
    lib = load()
    k, v, s = _c(sorted_keys, _u64), _c(vals, _u64), _c(starts, _u64)
    m = len(s)
    hi = None if his is None else _c(his, _u64)
    ln = None if lens is None else _c(lens, np.uint32)
    if ln is not None:
        offsets = np.zeros(m + 1, dtype=_u64)
        offsets[1:] = np.cumsum(ln.astype(_u64))
        total = int(offsets[-1])
    else:
        offsets = (np.arange(m + 1, dtype=_u64) * _u64(cap))
        total = m * cap
    ok = np.zeros(max(total, 1), dtype=_u64)
    ov = np.zeros(max(total, 1), dtype=_u64)
    oc = np.zeros(m, dtype=np.uint32)
    lib.hmo_range_sorted(_p(k), _p(v), len(k), _p(s), _p(hi), _p(ln), cap, m,
                         _p(offsets) if ln is not None else None, _p(ok), _p(ov), _p(oc))
    return offsets, ok, ov, oc


def upsert_sorted(sorted_keys, vals, bkeys, bvals):
    lib = load()
    k, v, bk, bv = _c(sorted_keys, _u64), _c(vals, _u64), _c(bkeys, _u64), _c(bvals, _u64)
    ok = np.empty(len(k) + len(bk) + 1, dtype=_u64)
    ov = np.empty(len(k) + len(bk) + 1, dtype=_u64)
    n = lib.hmo_upsert_sorted(_p(k), _p(v), len(k), _p(bk), _p(bv), len(bk), _p(ok), _p(ov))
    return ok[:n].copy(), ov[:n].copy()


def check_layout(layout):
    lib = load()
    return lib.hmo_check_layout(_p(layout.inner), _p(layout.prefix), _p(layout.leaves), layout.n_inner,
                                layout.n_leaves, layout.height, _p(layout.level_start), layout.fanout)


def splitmix64_mix(x):
    return load().hmo_splitmix64_mix(int(x))


# ---- the reference's own CPU code (oracle/_ref) ----------------------------------------------
class _RefLib:
    def __init__(self, path, kind):
        self.lib = ctypes.CDLL(path)
        self.kind = kind
        L = self.lib
        if kind == "rbpt":
            L.ref_rbpt_build.restype = _vp
            L.ref_rbpt_build.argtypes = [_vp, _vp, ctypes.c_uint64]
            L.ref_rbpt_insert.restype = _vp
            L.ref_rbpt_insert.argtypes = [_vp, _vp, _vp, ctypes.c_uint64]
            L.ref_rbpt_find.restype = None
            L.ref_rbpt_find.argtypes = [_vp, _vp, ctypes.c_uint64, _vp, _vp]
            L.ref_rbpt_find_omp.restype = None
            L.ref_rbpt_find_omp.argtypes = [_vp, _vp, ctypes.c_uint64, _vp, _vp, ctypes.c_int]
            L.ref_rbpt_find_range.restype = ctypes.c_uint64
            L.ref_rbpt_find_range.argtypes = [_vp, ctypes.c_int64, ctypes.c_int64, _vp, _vp, ctypes.c_uint64,
                                              ctypes.c_uint64]
            L.ref_rbpt_order.restype = ctypes.c_int
            L.ref_rbpt_height.restype = ctypes.c_int
            L.ref_rbpt_height.argtypes = [_vp]
        else:
            L.ref_h_build.restype = _vp
            L.ref_h_build.argtypes = [_vp, _vp, ctypes.c_uint64]
            L.ref_h_search.restype = None
            L.ref_h_search.argtypes = [_vp, _vp, ctypes.c_uint64, _vp]
            L.ref_h_search_omp.restype = None
            L.ref_h_search_omp.argtypes = [_vp, _vp, ctypes.c_uint64, _vp, ctypes.c_int]
            L.ref_h_range.restype = ctypes.c_uint64
            L.ref_h_range.argtypes = [_vp, ctypes.c_int64, ctypes.c_int64, _vp, _vp, ctypes.c_uint64]
            L.ref_h_height.restype = ctypes.c_int
            L.ref_h_height.argtypes = [_vp]


def ref_available():
    return all(os.path.exists(os.path.join(REF_DIR, f))
               for f in ("libref_rbpt64.so", "libref_rbpt128.so", "libref_harmonia.so"))


_refs = {}


def ref_lib(name):
    if name not in _refs:
        _refs[name] = _RefLib(os.path.join(REF_DIR, f"libref_{name}.so"), "harmonia" if name == "harmonia" else "rbpt")
    return _refs[name]


class RefRegularTree:
    """The reference's regular B+tree (order 64 or 128) built by repeated insert."""

    def __init__(self, keys, vals, order=64):
        self.r = ref_lib(f"rbpt{order}")
        k = _c(keys, _u64).view(np.int64)
        v = _c(vals, _u64).view(np.int64)
        self.n = len(k)
        self.root = self.r.lib.ref_rbpt_build(_p(k), _p(v), len(k))

    def insert(self, keys, vals):
        k = _c(keys, _u64).view(np.int64)
        v = _c(vals, _u64).view(np.int64)
        self.root = self.r.lib.ref_rbpt_insert(self.root, _p(k), _p(v), len(k))
        self.n += len(k)

    def find(self, q, nthreads=0):
        """find() per query; nthreads > 0 runs search_cpu_multithread's OpenMP loop."""
        q = _c(q, _u64).view(np.int64)
        ov = np.empty(len(q), dtype=np.int64)
        of = np.empty(len(q), dtype=np.uint8)
        if nthreads > 0:
            self.r.lib.ref_rbpt_find_omp(self.root, _p(q), len(q), _p(ov), _p(of), nthreads)
        else:
            self.r.lib.ref_rbpt_find(self.root, _p(q), len(q), _p(ov), _p(of))
        return ov.view(_u64), of

    def find_range(self, lo, hi, cap):
        ok = np.empty(max(cap, 1), dtype=np.int64)
        ov = np.empty(max(cap, 1), dtype=np.int64)
        c = self.r.lib.ref_rbpt_find_range(self.root, int(np.int64(np.uint64(lo).view(np.int64))),
                                           int(np.int64(np.uint64(hi).view(np.int64))), _p(ok), _p(ov), cap,
                                           self.n)
        c = min(int(c), cap)
        return ok[:c].view(_u64).copy(), ov[:c].view(_u64).copy()


class RefHarmoniaTree:
    """The reference's Harmonia tree: inserts, then BPlusTree::sort into the prefix layout."""

    def __init__(self, keys, vals):
        self.r = ref_lib("harmonia")
        k = _c(keys, _u64).view(np.int64)
        v = _c(vals, _u64).view(np.int64)
        self.t = self.r.lib.ref_h_build(_p(k), _p(v), len(k))

    def search(self, q, nthreads=0):
        """Values as int64 (-1 = not found, the reference's sentinel); nthreads > 0 runs
        search_cpu_test's OpenMP loop."""
        q = _c(q, _u64).view(np.int64)
        ov = np.empty(len(q), dtype=np.int64)
        if nthreads > 0:
            self.r.lib.ref_h_search_omp(self.t, _p(q), len(q), _p(ov), nthreads)
        else:
            self.r.lib.ref_h_search(self.t, _p(q), len(q), _p(ov))
        return ov

    def range_search(self, lo, hi, cap):
        ok = np.empty(max(cap, 1), dtype=np.int64)
        ov = np.empty(max(cap, 1), dtype=np.int64)
        c = self.r.lib.ref_h_range(self.t, int(lo), int(hi), _p(ok), _p(ov), cap)
        c = min(int(c), cap)
        return ok[:c].view(_u64).copy(), ov[:c].view(_u64).copy()
