"""Full-size checks on the 2^26-key, fanout-64 tree with 2^24 queries: size-independent properties
over every result (found <=> the query is in the key set; value = splitmix64_mix(query)), plus a
bit-exact sample against the oracle, for both query orders and PSA on and off."""
import numpy as np
import pytest

import oracle
from harmonia_amd import BPlusTree
from harmonia_amd import workloads as W

pytestmark = pytest.mark.gpu

N_KEYS = 1 << 26
N_Q = 1 << 24


@pytest.fixture(scope="module")
def big(gpu):
    import torch

    keys, vals = W.gen_keys(N_KEYS, W.config_seed(1))
    dk = torch.from_numpy(keys.view(np.int64)).to(gpu)
    dv = torch.from_numpy(vals.view(np.int64)).to(gpu)
    t = BPlusTree.from_keys(dk, dv, fanout=64, leaf_fill=8)
    del dk, dv
    sk = np.sort(keys)
    return t, sk


@pytest.mark.parametrize("order", ["unsorted", "runs4096"])
def test_fullsize_search_properties(big, gpu, order):
    import torch

    t, sk = big
    q = W.point_queries(sk, N_Q, W.config_seed(1) + 7, order=order)
    hit = np.isin(q, sk, assume_unique=False)
    assert hit.sum() == round(0.9 * N_Q)
    dq = torch.from_numpy(q.view(np.int64)).to(gpu)
    dv = torch.empty(N_Q, dtype=torch.int64, device=gpu)
    df = torch.empty(N_Q, dtype=torch.uint8, device=gpu)
    sample = W.rng(3, 3).integers(0, N_Q, 1 << 16)
    ev, ef = oracle.search_sorted(sk, W.splitmix64_mix(sk), q[sample])
    for psa in (0, 16):
        print(f"fullsize order={order} psa={psa}", flush=True)
        t.search_batch(dq, out_vals=dv, out_found=df, psa_bits=psa)
        torch.cuda.synchronize()
        f = df.cpu().numpy()
        v = dv.cpu().numpy().view(np.uint64)
        assert np.array_equal(f.astype(bool), hit)
        assert np.array_equal(v[hit], W.splitmix64_mix(q[hit]))
        assert np.all(v[~hit] == 0)
        assert np.array_equal(f[sample], ef) and np.array_equal(v[sample], ev)


def test_fullsize_layout_shape(big):
    t, sk = big
    s = t.stats()
    h, ni, nl, ls = oracle.plan(N_KEYS, 64, 8)
    assert (s.height, s.n_inner, s.n_leaves) == (h, ni, nl)
    assert s.n_leaves == N_KEYS // 8
    e = t.export()
    lk = e["leaves"][:, :8].ravel()
    assert np.array_equal(lk, sk)  # full leaves: the leaf chain is exactly the sorted key set


def test_fullsize_fanout128_zipf_and_ranges(gpu):
    """Fanout-128 tree of 2^26 keys: 2^24 Zipf(0.99) queries over the sorted key order (all hits),
    and 2^20 ranges (start uniform in [0, 2^63), length 16..256) checked against the sorted key
    array: the i-th result of a range is key[lower_bound(start) + i], truncated at the end."""
    import torch

    seed = W.config_seed(2)
    keys, vals = W.gen_keys(N_KEYS, seed)
    dk = torch.from_numpy(keys.view(np.int64)).to(gpu)
    dv = torch.from_numpy(vals.view(np.int64)).to(gpu)
    t = BPlusTree.from_keys(dk, dv, fanout=128, leaf_fill=8)
    del dk, dv, vals
    sk = np.sort(keys)
    del keys
    q = W.zipf_queries(sk, N_Q, seed + 11)
    dq = torch.from_numpy(q.view(np.int64)).to(gpu)
    ov = torch.empty(N_Q, dtype=torch.int64, device=gpu)
    of = torch.empty(N_Q, dtype=torch.uint8, device=gpu)
    for psa in (0, 16):
        print(f"fullsize f128 zipf psa={psa}", flush=True)
        t.search_batch(dq, out_vals=ov, out_found=of, psa_bits=psa)
        torch.cuda.synchronize()
        assert bool(of.all())
        assert np.array_equal(ov.cpu().numpy().view(np.uint64), W.splitmix64_mix(q))
    del dq, ov, of
    print("fullsize f128 ranges", flush=True)
    starts, lens = W.range_queries(1 << 20, seed + 13)
    off, rk, rv, cnt = t.range_batch(starts, lens)
    lb = np.searchsorted(sk, starts, side="left").astype(np.int64)
    exp_cnt = np.minimum(lens.astype(np.int64), N_KEYS - lb)
    assert np.array_equal(cnt.astype(np.int64), exp_cnt)
    assert off[0] == 0 and np.array_equal(np.diff(off.astype(np.int64)), exp_cnt)
    total = int(exp_cnt.sum())
    idx = np.repeat(lb - off[:-1].astype(np.int64), exp_cnt) + np.arange(total, dtype=np.int64)
    assert np.array_equal(rk[:total], sk[idx])
    assert np.array_equal(rv[:total], W.splitmix64_mix(sk[idx]))
    t.close()
