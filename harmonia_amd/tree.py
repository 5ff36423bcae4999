"""Host-side mirror of the reference's tree interface over the C ABI.

`BPlusTree` keeps the reference's method names (``insert``, ``search``, ``range_search``,
``insert_parallel``, ``getHeight``, ``getPrefixArray``, ``getLevelIdx``; H/src/bpt.h:456-561 and the
regular tree's ``find`` / ``find_range``, H/regular_bpt/regular_bpt.h:218-246) and adds the batched
forms the GPU path is built for.  Every operation runs on the GPU through libharmonia_amd; nothing
here computes a result on the CPU.

Arrays may be numpy arrays (host memory, staged by the library) or torch tensors on the GPU
(device pointers, used in place).
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import HmParams, HmSearchOpts, HmStats, check

KEY_MAX = _lib.HM_KEY_MAX


def _ptr(x):
    """Raw address of a numpy array or torch tensor (None -> NULL)."""
    if x is None:
        return None
    if isinstance(x, np.ndarray):
        if not x.flags["C_CONTIGUOUS"]:
            raise ValueError("arrays must be C-contiguous")
        return x.ctypes.data
    if hasattr(x, "data_ptr"):
        if hasattr(x, "is_contiguous") and not x.is_contiguous():
            raise ValueError("tensors must be contiguous")
        return x.data_ptr()
    raise TypeError(f"unsupported buffer type {type(x)!r}")


def _u64(a):
    if hasattr(a, "data_ptr"):
        return a
    return np.ascontiguousarray(a, dtype=np.uint64)


def _len(a):
    return int(a.numel()) if hasattr(a, "numel") else int(len(a))


def _stream(stream):
    if stream is None:
        return None
    if hasattr(stream, "cuda_stream"):
        return stream.cuda_stream
    return int(stream)


class BPlusTree:
    """A device-resident Harmonia B+tree (fanout F in {32, 64, 128}, 8-key leaves)."""

    def __init__(self, fanout=64, leaf_fill=8, lds_levels=-1, group_size=0, psa_bits=16,
                 sort_threshold=None, mem_budget=0):
        self._lib = _lib.load()
        self.params = HmParams()
        check(self._lib.hm_params_init(ctypes.byref(self.params)))
        self.params.fanout = fanout
        self.params.leaf_fill = leaf_fill
        self.params.lds_levels = lds_levels
        self.params.group_size = group_size
        self.params.psa_bits = psa_bits
        if sort_threshold is not None:  # else the library default
            self.params.sort_threshold = sort_threshold
        self.params.mem_budget = mem_budget
        self._h = None
        self._pending_keys = []
        self._pending_vals = []

    # ---- construction -------------------------------------------------------------------
    @classmethod
    def from_keys(cls, keys, vals, keys_sorted=False, **params):
        t = cls(**params)
        t.build(keys, vals, keys_sorted=keys_sorted)
        return t

    @classmethod
    def load(cls, path, **params):
        """Reload a snapshot written by save() (fanout and leaf fill come from the file)."""
        t = cls(**params)
        p = HmParams.from_buffer_copy(t.params)
        h = ctypes.c_void_p()
        check(t._lib.hm_tree_load(os.fsencode(path), ctypes.byref(p), ctypes.byref(h)))
        t._h = h
        return t

    def save(self, path):
        """Binary snapshot of the published tree."""
        check(self._lib.hm_tree_save(self._handle(), os.fsencode(path)))

    def build(self, keys=None, vals=None, keys_sorted=False):
        """Bulk build (replaces the reference's insert loop + BPlusTree::sort)."""
        if keys is None:
            keys = np.asarray(self._pending_keys, dtype=np.uint64)
            vals = np.asarray(self._pending_vals, dtype=np.uint64)
            self._pending_keys, self._pending_vals = [], []
        keys, vals = _u64(keys), _u64(vals)
        n = _len(keys)
        if _len(vals) != n:
            raise ValueError("keys and vals differ in length")
        self.close()
        p = HmParams.from_buffer_copy(self.params)
        if keys_sorted:
            p.flags |= _lib.HM_FLAG_KEYS_SORTED
        h = ctypes.c_void_p()
        check(self._lib.hm_tree_build(_ptr(keys), _ptr(vals), n, ctypes.byref(p), ctypes.byref(h)))
        self._h = h
        return self

    def insert(self, key, val):
        """Reference-style single insert: staged before build, an upsert after it."""
        if self._h is None:
            self._pending_keys.append(key)
            self._pending_vals.append(val)
            return True
        self.insert_parallel(np.array([key], dtype=np.uint64), np.array([val], dtype=np.uint64))
        return True

    def close(self):
        if self._h is not None:
            self._lib.hm_tree_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _handle(self):
        if self._h is None:
            raise RuntimeError("tree is not built")
        return self._h

    # ---- point search --------------------------------------------------------------------
    def search_batch(self, q, out_vals=None, out_found=None, stream=None, group_size=0, lds_levels=-1,
                     psa_bits=-1, blocks_per_cu=0, queries_per_group=0, output_mode=0, sorted_output=False,
                     out_pos=None, sorted_kernel=0):
        """Batched search; returns (values, found) in input order (or PSA order with out_pos).

        output_mode: 0 tree default, 1 input order via the replayed sort passes, 2 input order via direct
        scatter, 3 (or sorted_output=True) PSA order with the permutation in out_pos.
        sorted_kernel (PSA batches): 0 automatic, 1 one descent per query, 2 leaf window."""
        if sorted_output:
            output_mode = _lib.HM_OUT_PSA_ORDER
        q = _u64(q)
        n = _len(q)
        if out_vals is None:
            out_vals = np.empty(n, dtype=np.uint64)
        if out_found is None:
            out_found = np.empty(n, dtype=np.uint8)
        o = HmSearchOpts(group_size, lds_levels, psa_bits, blocks_per_cu, output_mode, queries_per_group,
                         _ptr(out_pos), sorted_kernel)
        check(self._lib.hm_search_ex(self._handle(), _ptr(q), n, _ptr(out_vals), _ptr(out_found),
                                     ctypes.byref(o), _stream(stream)))
        return out_vals, out_found

    def search_baseline(self, q, out_vals, out_found, kind="pointer", group_size=1, stream=None):
        """Comparator search over a pointer-child tree (same keys and leaves; device tensors only).

        kind "pointer": index granule, key granule, then child[] (HB+-tree, H/src/gbpt.cu:34-113);
        kind "linear": granule-by-granule node scan, then child[] (H/regular_bpt/noprefix-nosort.cu:89-237);
        kind "prefix": the same plain kernel as "pointer" over the prefix-sum array (the control).
        Results in input order, identical to search_batch."""
        k = {"pointer": _lib.HM_BASELINE_POINTER_CHILD, "linear": _lib.HM_BASELINE_POINTER_LINEAR,
             "prefix": _lib.HM_BASELINE_PREFIX_PLAIN}[kind]
        check(self._lib.hm_baseline_search(self._handle(), k, group_size, _ptr(q), _len(q), _ptr(out_vals),
                                           _ptr(out_found), _stream(stream)))
        return out_vals, out_found

    def search(self, key):
        """Single search: the value, or None when absent (the reference returns -1)."""
        v, f = self.search_batch(np.array([key], dtype=np.uint64), psa_bits=0)
        return int(v[0]) if f[0] else None

    find = search

    # ---- range ----------------------------------------------------------------------------
    def range_batch(self, starts, lens, stream=None):
        """First len[i] keys >= start[i]; returns (offsets, keys, vals, counts)."""
        starts = _u64(starts)
        lens = np.ascontiguousarray(lens, dtype=np.uint32) if not hasattr(lens, "data_ptr") else lens
        n = _len(starts)
        cap = int(np.asarray(lens, dtype=np.uint64).sum()) if isinstance(lens, np.ndarray) else int(lens.sum())
        offsets = np.empty(n + 1, dtype=np.uint64)
        keys = np.empty(max(cap, 1), dtype=np.uint64)
        vals = np.empty(max(cap, 1), dtype=np.uint64)
        counts = np.empty(n, dtype=np.uint32)
        check(self._lib.hm_range(self._handle(), _ptr(starts), _ptr(lens), n, _ptr(offsets), _ptr(keys),
                                 _ptr(vals), _ptr(counts), cap, _stream(stream)))
        return offsets, keys, vals, counts

    def range_between(self, lo, hi, cap, stream=None):
        """Keys in [lo[i], hi[i]], at most cap per range; returns (keys, vals, counts)."""
        lo, hi = _u64(lo), _u64(hi)
        n = _len(lo)
        keys = np.empty(max(n * cap, 1), dtype=np.uint64)
        vals = np.empty(max(n * cap, 1), dtype=np.uint64)
        counts = np.empty(n, dtype=np.uint32)
        check(self._lib.hm_range_between(self._handle(), _ptr(lo), _ptr(hi), n, cap, _ptr(keys), _ptr(vals),
                                         _ptr(counts), _stream(stream)))
        return keys, vals, counts

    def range_search(self, start_key, end_key, cap=4096):
        """All (key, value) with start <= key <= end, ascending (BPlusTree::range_search)."""
        while True:
            keys, vals, counts = self.range_between([start_key], [end_key], cap)
            c = int(counts[0])
            if c < cap:
                return list(zip(keys[:c].tolist(), vals[:c].tolist()))
            cap *= 4

    # ---- update ---------------------------------------------------------------------------
    def update(self, keys, vals, stream=None):
        """Stage an upsert batch; visible after publish()."""
        keys, vals = _u64(keys), _u64(vals)
        n = _len(keys)
        if _len(vals) != n:
            raise ValueError("keys and vals differ in length")
        check(self._lib.hm_update(self._handle(), _ptr(keys), _ptr(vals), n, _stream(stream)))

    def publish(self, stream=None):
        """Apply staged updates after every earlier search; returns True if leaves were re-laid out."""
        r = ctypes.c_int(0)
        check(self._lib.hm_epoch_publish(self._handle(), _stream(stream), ctypes.byref(r)))
        return bool(r.value)

    def insert_parallel(self, keys, vals):
        """Batched upsert applied immediately (BPlusTree::insert_parallel)."""
        self.update(keys, vals)
        return self.publish()

    # ---- tuning ---------------------------------------------------------------------------
    def autotune(self, sample, quick=False):
        """Time launch shapes on a sample batch and make the winners the tree's defaults."""
        sample = _u64(sample)
        r = _lib.HmTuneResult()
        check(self._lib.hm_autotune(self._handle(), _ptr(sample), _len(sample),
                                    _lib.HM_TUNE_QUICK if quick else 0, ctypes.byref(r)))
        return r.as_dict()

    @classmethod
    def from_keys_tuned(cls, keys, vals, sample, quick=False, **params):
        """Build with fanout 32, 64 and 128, autotune each on the sample, keep the fastest."""
        t = cls(**params)
        keys, vals, sample = _u64(keys), _u64(vals), _u64(sample)
        h = ctypes.c_void_p()
        r = _lib.HmTuneResult()
        check(t._lib.hm_tree_build_tuned(_ptr(keys), _ptr(vals), _len(keys), ctypes.byref(t.params), _ptr(sample),
                                         _len(sample), _lib.HM_TUNE_QUICK if quick else 0, ctypes.byref(h),
                                         ctypes.byref(r)))
        t._h = h
        t.params.fanout = r.fanout
        return t, r.as_dict()

    # ---- introspection --------------------------------------------------------------------
    def stats(self):
        s = HmStats()
        check(self._lib.hm_tree_stats(self._handle(), ctypes.byref(s)))
        return s

    def getHeight(self):
        return int(self.stats().height)

    def getLevelIdx(self, level):
        return int(self.stats().level_start[level])

    def export(self):
        """Host copies of the layout: inner (n_inner, node_slots), prefix, leaves (n_leaves, 16)."""
        s = self.stats()
        inner = np.empty((s.n_inner, s.node_slots), dtype=np.uint64)
        prefix = np.empty(s.n_inner, dtype=np.uint32)
        leaves = np.empty((s.n_leaves, 16), dtype=np.uint64)
        check(self._lib.hm_tree_export(self._handle(), _ptr(inner), _ptr(prefix), _ptr(leaves)))
        return {"inner": inner, "prefix": prefix, "leaves": leaves, "stats": s}

    def getPrefixArray(self):
        return self.export()["prefix"]


def read_keys_text(path):
    """Keys of a text dataset in the reference's format (one decimal key per line)."""
    lib = _lib.load()
    n = ctypes.c_uint64(0)
    check(lib.hm_read_keys_text(os.fsencode(path), None, 0, ctypes.byref(n)))
    out = np.empty(max(n.value, 1), dtype=np.uint64)
    check(lib.hm_read_keys_text(os.fsencode(path), _ptr(out), n.value, ctypes.byref(n)))
    return out[:n.value]


def read_ranges_text(path):
    """(starts, ends) of a range file: a start line and an end line per range."""
    lib = _lib.load()
    n = ctypes.c_uint64(0)
    check(lib.hm_read_ranges_text(os.fsencode(path), None, None, 0, ctypes.byref(n)))
    a = np.empty(max(n.value, 1), dtype=np.uint64)
    b = np.empty(max(n.value, 1), dtype=np.uint64)
    check(lib.hm_read_ranges_text(os.fsencode(path), _ptr(a), _ptr(b), n.value, ctypes.byref(n)))
    return a[:n.value], b[:n.value]


def psa_sort(keys, begin_bit=48, end_bit=64, stream=None):
    """Stable radix sort of (key, position) over key bits [begin_bit, end_bit) on the GPU."""
    lib = _lib.load()
    keys = _u64(keys)
    n = _len(keys)
    ko = np.empty(n, dtype=np.uint64)
    po = np.empty(n, dtype=np.uint32)
    check(lib.hm_psa_sort(_ptr(keys), n, begin_bit, end_bit, _ptr(ko), _ptr(po), _stream(stream)))
    return ko, po


class Batcher:
    """Many uneven query batches served by one hipGraph replay (hm_batcher_*).

    add() packs a batch into the batcher's device buffer and returns its offset; flush() launches
    PSA sort -> one persistent search launch -> un-sort over everything added; copy_out() reads
    one batch's results (or results() gives the device pointers)."""

    def __init__(self, tree, capacity, stream=None, group_size=0, lds_levels=-1, psa_bits=-1, blocks_per_cu=0,
                 queries_per_group=0, output_mode=0):
        self._lib = _lib.load()
        self.tree = tree
        self.capacity = int(capacity)
        o = HmSearchOpts(group_size, lds_levels, psa_bits, blocks_per_cu, output_mode, queries_per_group, None)
        h = ctypes.c_void_p()
        check(self._lib.hm_batcher_create(tree._handle(), self.capacity, ctypes.byref(o), _stream(stream),
                                          ctypes.byref(h)))
        self._h = h
        self._keep = []  # device batches are read by the flush: keep them alive until sync()

    def add(self, q):
        q = _u64(q)
        off = ctypes.c_uint64(0)
        check(self._lib.hm_batcher_add(self._h, _ptr(q), _len(q), ctypes.byref(off)))
        if hasattr(q, "data_ptr"):
            self._keep.append(q)
        return int(off.value)

    def add_many(self, ptrs, sizes, device=True, offsets=None):
        """Add batches given as raw addresses (uint64 array) and sizes until the buffer is full; returns
        how many were added.  With device=True every address must be device memory, kept valid until
        the next flush has run (sync())."""
        ptrs = np.ascontiguousarray(ptrs, dtype=np.uint64)
        sizes = np.ascontiguousarray(sizes, dtype=np.uint64)
        if len(ptrs) != len(sizes):
            raise ValueError("ptrs and sizes differ in length")
        added = ctypes.c_uint64(0)
        check(self._lib.hm_batcher_add_many(self._h, _ptr(ptrs), _ptr(sizes), len(ptrs),
                                            _lib.HM_BATCH_DEVICE if device else 0, _ptr(offsets),
                                            ctypes.byref(added)))
        return int(added.value)

    def flush(self):
        check(self._lib.hm_batcher_flush(self._h))

    def results(self):
        v, f = ctypes.c_void_p(), ctypes.c_void_p()
        check(self._lib.hm_batcher_results(self._h, ctypes.byref(v), ctypes.byref(f)))
        return v.value, f.value

    def copy_out(self, offset, n, out_vals=None, out_found=None):
        if out_vals is None:
            out_vals = np.empty(n, dtype=np.uint64)
        if out_found is None:
            out_found = np.empty(n, dtype=np.uint8)
        check(self._lib.hm_batcher_copy_out(self._h, offset, n, _ptr(out_vals), _ptr(out_found)))
        return out_vals, out_found

    def sync(self):
        check(self._lib.hm_batcher_sync(self._h))
        self._keep.clear()

    def close(self):
        if self._h is not None:
            self._lib.hm_batcher_destroy(self._h)  # waits for the batcher's stream
            self._h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The serving loop (hm_server_*): requests are admitted as they arrive, coalesced into batches
    while other batches run, and updates are applied between read epochs."""

    def __init__(self, tree, mem_budget=0, max_batch=0, max_inflight=0, coalesce=True):
        self._lib = _lib.load()
        self.tree = tree
        p = _lib.HmServerParams(mem_budget, max_batch, max_inflight, 1 if coalesce else 0, 0)
        h = ctypes.c_void_p()
        check(self._lib.hm_server_create(tree._handle(), ctypes.byref(p), ctypes.byref(h)))
        self._h = h
        self._outs = {}
        self._inputs = {}  # the library reads inputs when a request joins a batch: keep them alive

    def lookup(self, keys):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        v = np.empty(len(keys), dtype=np.uint64)
        f = np.empty(len(keys), dtype=np.uint8)
        t = ctypes.c_uint64(0)
        check(self._lib.hm_server_lookup(self._h, _ptr(keys), len(keys), _ptr(v), _ptr(f), ctypes.byref(t)))
        self._outs[t.value] = (v, f)
        self._inputs[t.value] = keys
        return t.value

    def range(self, starts, lens):
        starts = np.ascontiguousarray(starts, dtype=np.uint64)
        lens = np.ascontiguousarray(lens, dtype=np.uint32)
        cap = int(lens.astype(np.uint64).sum())
        k = np.empty(max(cap, 1), dtype=np.uint64)
        v = np.empty(max(cap, 1), dtype=np.uint64)
        c = np.empty(len(starts), dtype=np.uint32)
        t = ctypes.c_uint64(0)
        check(self._lib.hm_server_range(self._h, _ptr(starts), _ptr(lens), len(starts), _ptr(k), _ptr(v), _ptr(c),
                                         cap, ctypes.byref(t)))
        off = np.zeros(len(lens) + 1, dtype=np.uint64)
        off[1:] = np.cumsum(lens.astype(np.uint64))
        self._outs[t.value] = (off, k, v, c)
        self._inputs[t.value] = (starts, lens)
        return t.value

    def update(self, keys, vals):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        vals = np.ascontiguousarray(vals, dtype=np.uint64)
        t = ctypes.c_uint64(0)
        check(self._lib.hm_server_update(self._h, _ptr(keys), _ptr(vals), len(keys), ctypes.byref(t)))
        self._outs[t.value] = ()
        self._inputs[t.value] = (keys, vals)
        return t.value

    def wait(self, ticket):
        """Block until the request is done; returns (status, epoch, outputs)."""
        st = ctypes.c_int(0)
        ep = ctypes.c_uint64(0)
        check(self._lib.hm_server_wait(self._h, ticket, ctypes.byref(st), ctypes.byref(ep)))
        self._inputs.pop(ticket, None)
        return st.value, ep.value, self._outs.pop(ticket)

    def drain(self):
        check(self._lib.hm_server_drain(self._h))
        self._inputs.clear()  # every request has completed

    def stats(self):
        s = _lib.HmServerStats()
        check(self._lib.hm_server_get_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    def reset_stats(self):
        check(self._lib.hm_server_reset_stats(self._h))

    def close(self):
        if self._h is not None:
            self._lib.hm_server_destroy(self._h)  # finishes every pending request first
            self._h = None
            self._inputs.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
