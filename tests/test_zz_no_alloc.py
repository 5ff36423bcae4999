"""hm_search_noalloc (fixed device scratch, no allocation) against hm_search through the C ABI:
bit-identical results on two calls in a row, the second smaller, plus the oracle on a sample and
the rejections of oversized batches and host buffers."""
import ctypes

import numpy as np
import pytest

import oracle
from harmonia_amd import BPlusTree, _lib
from harmonia_amd import workloads as W

pytestmark = pytest.mark.gpu


def _noalloc(lib, tree, dq, ov, of, stream):
    return lib.hm_search_noalloc(tree._handle(), dq.data_ptr(), dq.numel(), ov.data_ptr(), of.data_ptr(),
                                 stream.cuda_stream)


@pytest.mark.parametrize("psa_threshold", [1024, 1 << 40])  # PSA pipeline (scratch in use) / direct search
def test_noalloc_equals_search(gpu, psa_threshold):
    import torch

    lib = _lib.load()
    print(f"noalloc: 2^20-key tree, sort threshold {psa_threshold}", flush=True)
    keys, vals = W.gen_keys(1 << 20, 71)
    sk = np.sort(keys)
    sv = W.splitmix64_mix(sk)
    tree = BPlusTree.from_keys(keys, vals, fanout=64, sort_threshold=psa_threshold)
    stream = torch.cuda.Stream()
    nmax = int(lib.hm_search_noalloc_max())
    assert nmax == 1 << 20
    for i, m in enumerate((nmax, 100003)):  # two calls in a row, the second smaller
        q = W.point_queries(sk, m, 72 + i)
        q[:4] = [0, (1 << 63) - 1, 0xFFFFFFFFFFFFFFFF, sk[-1] + 1]
        dq = torch.from_numpy(q.view(np.int64)).to(gpu)
        ov = torch.full((m,), 7, dtype=torch.int64, device=gpu)
        of = torch.full((m,), 7, dtype=torch.uint8, device=gpu)
        rv = torch.empty(m, dtype=torch.int64, device=gpu)
        rf = torch.empty(m, dtype=torch.uint8, device=gpu)
        torch.cuda.synchronize()
        assert _noalloc(lib, tree, dq, ov, of, stream) == 0, _lib.last_error()
        tree.search_batch(dq, out_vals=rv, out_found=rf, stream=stream)
        stream.synchronize()
        got_v, got_f = ov.cpu().numpy().view(np.uint64), of.cpu().numpy()
        assert np.array_equal(got_f, rf.cpu().numpy())
        assert np.array_equal(got_v, rv.cpu().numpy().view(np.uint64))
        hit = got_f.astype(bool)
        assert np.array_equal(hit, np.isin(q, sk))
        assert np.array_equal(got_v[hit], W.splitmix64_mix(q[hit]))
        assert not got_v[~hit].any()
        sample = slice(0, 1 << 14)
        ev, ef = oracle.search_sorted(sk, sv, q[sample])
        assert np.array_equal(got_f[sample], ef) and np.array_equal(got_v[sample], ev)
    tree.close()


def test_noalloc_rejections(gpu):
    import torch

    lib = _lib.load()
    keys, vals = W.gen_keys(5000, 73)
    tree = BPlusTree.from_keys(keys, vals)
    stream = torch.cuda.Stream()
    m = int(lib.hm_search_noalloc_max()) + 1
    dq = torch.zeros(m, dtype=torch.int64, device=gpu)
    ov = torch.empty(m, dtype=torch.int64, device=gpu)
    of = torch.empty(m, dtype=torch.uint8, device=gpu)
    assert _noalloc(lib, tree, dq, ov, of, stream) == _lib.HM_EINVAL
    assert "exceed" in _lib.last_error()
    hq = np.zeros(16, dtype=np.uint64)
    hv = np.zeros(16, dtype=np.uint64)
    hf = np.zeros(16, dtype=np.uint8)
    rc = lib.hm_search_noalloc(tree._handle(), hq.ctypes.data, 16, hv.ctypes.data, hf.ctypes.data,
                               stream.cuda_stream)
    assert rc == _lib.HM_EINVAL and "device buffers" in _lib.last_error()
    assert lib.hm_search_noalloc(tree._handle(), None, 0, None, None, None) == 0  # empty batch
    stream.synchronize()
    tree.close()


def test_noalloc_sees_published_epochs(gpu):
    """An upsert published between two calls is visible to the second (the call waits for the publish,
    and a publish waits for the call's reads)."""
    import torch

    lib = _lib.load()
    keys, vals = W.gen_keys(200000, 74)
    tree = BPlusTree.from_keys(keys, vals, sort_threshold=1024)
    stream = torch.cuda.Stream()
    q = np.sort(keys)[:50000].copy()
    dq = torch.from_numpy(q.view(np.int64)).to(gpu)
    ov = torch.empty(len(q), dtype=torch.int64, device=gpu)
    of = torch.empty(len(q), dtype=torch.uint8, device=gpu)
    assert _noalloc(lib, tree, dq, ov, of, stream) == 0, _lib.last_error()
    newv = np.arange(len(q), dtype=np.uint64) + np.uint64(12345)
    tree.update(torch.from_numpy(q.view(np.int64)).to(gpu), torch.from_numpy(newv.view(np.int64)).to(gpu),
                stream=stream)
    tree.publish(stream=stream)
    assert _noalloc(lib, tree, dq, ov, of, stream) == 0, _lib.last_error()
    stream.synchronize()
    assert of.cpu().numpy().all()
    assert np.array_equal(ov.cpu().numpy().view(np.uint64), newv)
    tree.close()
